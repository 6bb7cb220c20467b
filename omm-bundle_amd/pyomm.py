"""ctypes access to the Fortran drop-in (libommf.so): MatrixSwitch matrices by handle and omm().

This mirrors how a Fortran SCF code drives libOMM (libOMM/examples/example_gamma.F90:134-245):
allocate matrices with a storage label, fill them, call omm() once per SCF step with the
reference argument list, read e_min and the density matrix back.  Storage labels starting with
'g' keep the matrices in HBM and run the HIP kernels of libommhip.so.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libommf.so")

_i, _d, _vp, _i64 = C.c_int, C.c_double, C.c_void_p, C.c_int64
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} not found: build it with omm-bundle_amd/build.sh")
        L = C.CDLL(LIB_PATH)
        L.omm_c_setup.argtypes = [_i]
        L.omm_c_mat_alloc.argtypes = [_i, _i, C.c_char_p]
        L.omm_c_mat_from_csc.argtypes = [_i, _vp, _vp, _vp, C.c_char_p]
        L.omm_c_mat_set_walk.argtypes = [_i, _i, _vp, _i]
        L.omm_c_mat_set_walk.restype = _i
        L.omm_c_mat_from_bsr.argtypes = [_i, _vp, _vp, _vp]
        L.omm_c_mat_register_pcsc.argtypes = [_i, _i, _vp, _vp, _vp]
        L.omm_c_mat_convert.argtypes = [_i, C.c_char_p, _d]
        L.omm_c_mat_bsr.argtypes = [_i]
        L.omm_c_mat_bsr.restype = _vp
        L.omm_c_mat_upload.argtypes = [_i, _vp]
        L.omm_c_mat_download.argtypes = [_i, _vp]
        L.omm_c_mat_free.argtypes = [_i]
        L.omm_c_mat_dptr.argtypes = [_i, C.POINTER(_i64)]
        L.omm_c_mat_dptr.restype = _vp
        L.omm_c_solve.argtypes = [_i, _i, _i, _i, _i, C.POINTER(_d), _i, _i, _d, _i, _i, _i, _i, _i, _d, _i, _i,
                                  C.c_char_p, C.c_char_p]
        L.omm_c_stats.argtypes = [C.POINTER(_i), C.POINTER(_i), C.POINTER(_i), C.POINTER(_i), C.POINTER(_d)]
        L.omm_c_elog.argtypes = [_vp, _i]
        L.omm_c_refine_stats.argtypes = [C.POINTER(_d), C.POINTER(_d)]
        L.omm_c_set_option.argtypes = [C.c_char_p, _d]
        L.omm_c_tlog.argtypes = [_vp, _i]
        L.omm_c_ctx.restype = _vp
        L.omm_c_param_grad.argtypes = [_i, _i, _vp, _vp, _vp, _vp]
        L.omm_c_mat_from_bsr_z.argtypes = [_i, _vp, _vp, _vp]
        L.omm_c_mat_from_dense_z.argtypes = [_i, _vp, _d, C.c_char_p]
        L.omm_c_mat_upload_z.argtypes = [_i, _vp]
        L.omm_c_mat_download_z.argtypes = [_i, _vp]
        L.omm_c_get_element_z.argtypes = [_i, _i, _i, C.POINTER(_d), C.POINTER(_d)]
        L.omm_c_get_element.argtypes = [_i, _i, _i, C.POINTER(_d)]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class Mat:
    """A MatrixSwitch matrix held by the Fortran library (handle)."""

    def __init__(self, rows, cols, label="gdden", data=None):
        self.rows, self.cols, self.label = rows, cols, label
        self.h = lib().omm_c_mat_alloc(rows, cols, label.encode())
        if self.h <= 0:
            raise RuntimeError("omm_c_mat_alloc failed")
        if data is not None:
            self.upload(data)

    @classmethod
    def from_csc(cls, nb, colptr, rowidx, val, label="gdcsc"):
        self = cls.__new__(cls)
        self.rows = self.cols = nb
        self.label = label
        cp = np.ascontiguousarray(colptr, np.int32)
        ri = np.ascontiguousarray(rowidx, np.int32)
        v = np.ascontiguousarray(val, np.float64)
        self.h = lib().omm_c_mat_from_csc(nb, _p(cp), _p(ri), _p(v), label.encode())
        if self.h <= 0:
            raise RuntimeError("omm_c_mat_from_csc failed")
        return self

    @classmethod
    def register_pcsc(cls, nb, idx1, idx2, val):
        """Device operator registered through m_register_pcsc from a CSC in pspBLAS's convention:
        1-based row indices idx1 and column pointers idx2 (idx2[0] == 1)."""
        self = cls.__new__(cls)
        self.rows = self.cols = nb
        self.label = "gdcsc"
        i1 = np.ascontiguousarray(idx1, np.int32)
        i2 = np.ascontiguousarray(idx2, np.int32)
        v = np.ascontiguousarray(val, np.float64)
        self.h = lib().omm_c_mat_register_pcsc(nb, len(v), _p(i1), _p(i2), _p(v))
        if self.h <= 0:
            raise RuntimeError("omm_c_mat_register_pcsc failed")
        return self

    def bsr_values(self):
        """Block values (nnzb x 16, column-major blocks) of a real device sparse matrix."""
        import ommhip as oh

        b = lib().omm_c_mat_bsr(self.h)
        if not b:
            raise RuntimeError("not a device sparse matrix")
        nnzb = C.c_int64(0)
        oh.check(oh.lib().omx_bsr_info(b, None, C.byref(nnzb), None), "bsr_info")
        out = np.zeros((nnzb.value, 16))
        oh.check(oh.lib().omx_bsr_get_values(b, _p(out)), "bsr_get_values")
        return out

    def convert(self, label, threshold=None):
        """m_convert to another 5-character label (optionally with a threshold)."""
        rc = lib().omm_c_mat_convert(self.h, label.encode(), -1.0 if threshold is None else float(threshold))
        if rc != 0:
            raise RuntimeError("omm_c_mat_convert failed (bad handle)")
        self.label = label

    @classmethod
    def from_bsr(cls, nb, bcolptr, browidx, bval):
        """Device operator ('gdcsc') from 0-based BSR-4 block columns."""
        self = cls.__new__(cls)
        self.rows = self.cols = nb
        self.label = "gdcsc"
        cp = np.ascontiguousarray(bcolptr, np.int32)
        ri = np.ascontiguousarray(browidx, np.int32)
        v = np.ascontiguousarray(bval, np.float64)
        self.h = lib().omm_c_mat_from_bsr(nb, _p(cp), _p(ri), _p(v))
        if self.h <= 0:
            raise RuntimeError("omm_c_mat_from_bsr failed")
        return self

    def set_walk(self, order, group=0):
        """Block-column visiting order of a real device operator, LDS-staged groups of `group`."""
        o = np.ascontiguousarray(order, np.int32)
        if lib().omm_c_mat_set_walk(self.h, len(o), _p(o), int(group)) != 0:
            raise RuntimeError("omm_c_mat_set_walk failed")

    @property
    def is_complex(self):
        return self.label[1] in "zc"

    @classmethod
    def from_bsr_z(cls, nb, bcolptr, browidx, bval):
        """Complex device operator ('gzcsc') from BSR-4 arrays with complex128 values (nnzb, 16)."""
        self = cls.__new__(cls)
        self.rows = self.cols = nb
        self.label = "gzcsc"
        cp = np.ascontiguousarray(bcolptr, np.int32)
        ri = np.ascontiguousarray(browidx, np.int32)
        v = np.ascontiguousarray(bval, np.complex128)
        self.h = lib().omm_c_mat_from_bsr_z(nb, _p(cp), _p(ri), _p(v))
        if self.h <= 0:
            raise RuntimeError("omm_c_mat_from_bsr_z failed")
        return self

    @classmethod
    def from_dense_z(cls, a, thr=0.0, label="gzcsc"):
        """Complex operator from a dense complex host matrix (m_copy, entries |a| > thr)."""
        a = np.asfortranarray(a, np.complex128)
        self = cls.__new__(cls)
        self.rows = self.cols = a.shape[0]
        self.label = label
        self.h = lib().omm_c_mat_from_dense_z(a.shape[0], _p(a), thr, label.encode())
        if self.h <= 0:
            raise RuntimeError("omm_c_mat_from_dense_z failed")
        return self

    def upload(self, a):
        if self.is_complex:
            a = np.asfortranarray(a, np.complex128)
            assert a.shape == (self.rows, self.cols)
            if lib().omm_c_mat_upload_z(self.h, _p(a)) != 0:
                raise RuntimeError("upload failed")
            return
        a = np.asfortranarray(a, np.float64)
        assert a.shape == (self.rows, self.cols)
        if lib().omm_c_mat_upload(self.h, _p(a)) != 0:
            raise RuntimeError("upload failed")

    def download(self):
        if self.is_complex:
            out = np.zeros((self.rows, self.cols), np.complex128, order="F")
            if lib().omm_c_mat_download_z(self.h, _p(out)) != 0:
                raise RuntimeError("download failed")
            return out
        out = np.zeros((self.rows, self.cols), order="F")
        if lib().omm_c_mat_download(self.h, _p(out)) != 0:
            raise RuntimeError("download failed")
        return out

    def dptr(self):
        ld = C.c_int64(0)
        p = lib().omm_c_mat_dptr(self.h, C.byref(ld))
        return p, ld.value

    def get_element_z(self, i, j):
        """Complex element (i, j), 1-based (m_get_element with a complex result)."""
        re, im = C.c_double(0), C.c_double(0)
        if lib().omm_c_get_element_z(self.h, i, j, C.byref(re), C.byref(im)) != 0:
            raise RuntimeError("get_element failed")
        return complex(re.value, im.value)

    def get_element(self, i, j):
        """Element (i, j), 1-based (m_get_element; sparse device matrices included)."""
        v = C.c_double(0)
        if lib().omm_c_get_element(self.h, i, j, C.byref(v)) != 0:
            raise RuntimeError("get_element failed")
        return v.value

    def free(self):
        if self.h > 0:
            lib().omm_c_mat_free(self.h)
            self.h = 0


def setup(device=0):
    lib().omm_c_setup(device)


def sync():
    lib().omm_c_sync()


def set_option(name: str, value: float):
    lib().omm_c_set_option(name.encode(), float(value))


def omm(m, n, H: Mat, S: Mat | None, C_min: Mat, *, new_S=True, D_min: Mat | None = None, calc_ED=False, eta=0.0,
        init_C=True, flavour=0, np_=1, ip=1, cg_tol=-1.0, long_out=False, dealloc=False, storage="gdden",
        operation="hip"):
    """Call the Fortran omm() (same arguments as libOMM/src/omm.F90:5-6).  Returns a dict."""
    e = C.c_double(0)
    rc = lib().omm_c_solve(m, n, H.h, S.h if S is not None else 0, int(new_S), C.byref(e),
                           D_min.h if D_min is not None else 0, int(calc_ED), eta, C_min.h, int(init_C), flavour, np_,
                           ip, cg_tol, int(long_out), int(dealloc), storage.encode(), operation.encode())
    if rc != 0:
        raise RuntimeError("omm_c_solve failed (bad handle)")
    icg, conv, nfail, icg_lowp = C.c_int(0), C.c_int(0), C.c_int(0), C.c_int(0)
    trqs = C.c_double(0)
    lib().omm_c_stats(C.byref(icg), C.byref(conv), C.byref(nfail), C.byref(icg_lowp), C.byref(trqs))
    buf = np.zeros(100000)
    k = lib().omm_c_elog(_p(buf), len(buf))
    tbuf = np.zeros(100000)
    kt = lib().omm_c_tlog(_p(tbuf), len(tbuf))
    g0, gend = C.c_double(0), C.c_double(0)
    lib().omm_c_refine_stats(C.byref(g0), C.byref(gend))
    return {"e_min": e.value, "icg": icg.value, "converged": bool(conv.value), "nfail": nfail.value,
            "icg_lowp": icg_lowp.value, "TrQS": trqs.value, "e_log": buf[:k].copy(), "t0": tbuf[0],
            "t_log": tbuf[1:kt].copy(), "refine_g0": g0.value, "refine_gend": gend.value}


def param_grad(D: Mat, nparam: int, seg_p, ent_p, coef_p) -> np.ndarray:
    """dE/dtheta (omm_param_grad) from a sparse device D_min and device Jacobian arrays."""
    out = np.zeros(nparam)
    if lib().omm_c_param_grad(D.h, nparam, seg_p, ent_p, coef_p, _p(out)) != 0:
        raise RuntimeError("param_grad failed (bad handle)")
    return out


def ctx_handle():
    """The libommhip context (void*) behind the 'g' storage, for omx_prof_* event timing."""
    return lib().omm_c_ctx()
