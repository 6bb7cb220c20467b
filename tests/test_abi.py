"""The C-ABI libraries load without a GPU and export every entry point include/omm_hip.h declares;
the Fortran drop-in exports omm and its C facade.  CPU only (no compute calls)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(omx_[A-Za-z0-9_]+)\s*\(", text)))


def test_libommhip_exports_header():
    lib = ctypes.CDLL(os.path.join(ROOT, "omm-bundle_amd", "libommhip.so"))
    names = declared("omm_hip.h")
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    import ommhip

    names = declared("omm_hip.h")
    assert sorted(ommhip.SIGNATURES) == names


def test_libommf_exports_drop_in():
    lib = ctypes.CDLL(os.path.join(ROOT, "omm-bundle_amd", "libommf.so"))
    for n in ("omm_", "omm_c_setup", "omm_c_solve", "omm_c_mat_alloc", "omm_c_mat_from_csc", "omm_c_stats"):
        assert hasattr(lib, n), n


def test_no_device_without_gpu_is_reported():
    import ommhip

    if ommhip.device_count() == 0:
        h = ctypes.c_void_p()
        assert ommhip.lib().omx_init(0, ctypes.byref(h)) != 0
        assert b"no HIP device" in ommhip.lib().omx_last_error()
