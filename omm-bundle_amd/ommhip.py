"""ctypes binding of libommhip.so (the C-ABI declared in include/omm_hip.h).

Used by the tests and the benchmark to call the HIP path directly; the Fortran host reaches the
same entry points through ISO_C_BINDING.  Loading fails loudly when the library has not been
built: there is no CPU fallback behind these functions.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libommhip.so")

OMX_F64, OMX_F32, OMX_BF16, OMX_C128, OMX_C64 = 0, 1, 2, 3, 4
PREC_F64, PREC_F32, PREC_BF16, PREC_BF16X3, PREC_BF16X6 = 0, 1, 2, 3, 4
K_SPMM, K_GEMM, K_COEFF, K_ELEM, K_DOT, K_GEMM_PAIR = 0, 1, 2, 3, 4, 5

_NP = {OMX_F64: np.float64, OMX_F32: np.float32, OMX_C128: np.complex128, OMX_C64: np.complex64}
_ESIZE = {OMX_F64: 8, OMX_F32: 4, OMX_BF16: 2, OMX_C128: 16, OMX_C64: 8}

_vp, _i, _i64, _d, _c = C.c_void_p, C.c_int, C.c_int64, C.c_double, C.c_char
_pd, _pi = C.POINTER(C.c_double), C.POINTER(C.c_int)

# name -> (restype, argtypes)
SIGNATURES = {
    "omx_device_count": (_i, [_pi]),
    "omx_init": (_i, [_i, C.POINTER(_vp)]),
    "omx_finalize": (_i, [_vp]),
    "omx_sync": (_i, [_vp]),
    "omx_last_error": (C.c_char_p, []),
    "omx_stream": (_vp, [_vp]),
    "omx_malloc": (_i, [_vp, C.c_size_t, C.POINTER(_vp)]),
    "omx_free": (_i, [_vp, _vp]),
    "omx_zero": (_i, [_vp, _vp, C.c_size_t]),
    "omx_h2d": (_i, [_vp, _vp, _vp, C.c_size_t]),
    "omx_d2h": (_i, [_vp, _vp, _vp, C.c_size_t]),
    "omx_d2d": (_i, [_vp, _vp, _vp, C.c_size_t]),
    "omx_h2d_2d": (_i, [_vp, _vp, _i64, _vp, _i64, _i, _i, _i]),
    "omx_d2h_2d": (_i, [_vp, _vp, _i64, _vp, _i64, _i, _i, _i]),
    "omx_convert": (_i, [_vp, _i, _i, _vp, _i, _i64, _vp, _i, _i64]),
    "omx_bsr_create": (_i, [_vp, _i, _vp, _vp, _vp, C.POINTER(_vp)]),
    "omx_bsr_from_csc": (_i, [_vp, _i, _vp, _vp, _vp, _i, _i, C.POINTER(_vp)]),
    "omx_csc_to_bsr": (_i, [_i, _vp, _vp, _vp, _i, _i, _i, C.POINTER(_i64), _vp, _vp, _vp]),
    "omx_bsr_destroy": (_i, [_vp]),
    "omx_bsr_info": (_i, [_vp, _pi, C.POINTER(_i64), _pi]),
    "omx_bsr_update_values": (_i, [_vp, _vp]),
    "omx_bsr_get_values": (_i, [_vp, _vp]),
    "omx_bsr_get_pattern": (_i, [_vp, _vp, _vp]),
    "omx_sddmm_bsr": (_i, [_vp, _i, _i, _i, _d, _vp, _i64, _vp, _i64, _d, _vp]),
    "omx_bsr_param_grad": (_i, [_vp, _vp, _i, _vp, _vp, _vp, _vp]),
    "omx_bsr_create_z": (_i, [_vp, _i, _vp, _vp, _vp, C.POINTER(_vp)]),
    "omx_potrf": (_i, [_vp, _i, _i, _vp, _i64, C.POINTER(_i)]),
    "omx_zbatch_create": (_i, [_vp, _i, _i, _i, _vp, _vp, _vp, _vp, C.POINTER(_vp)]),
    "omx_zbatch_destroy": (_i, [_vp]),
    "omx_zbatch_create_dt": (_i, [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _vp, C.POINTER(_vp)]),
    "omx_zbatch_set_c": (_i, [_vp, _vp]),
    "omx_zbatch_get_c": (_i, [_vp, _vp]),
    "omx_zbatch_solve": (_i, [_vp, _d, _d, _i, _i, _i, _vp, _vp, _vp, C.POINTER(_i)]),
    "omx_trsm": (_i, [_vp, _i, C.c_char, C.c_char, _i, _i, _d, _vp, _i64, _vp, _i64]),
    "omx_trmm": (_i, [_vp, _i, C.c_char, C.c_char, _i, _i, _d, _vp, _i64, _vp, _i64]),
    "omx_sygst": (_i, [_vp, _i, _i, _vp, _i64, _vp, _i64]),
    "omx_inverse_spd": (_i, [_vp, _i, _i, _vp, _i64, C.POINTER(_i)]),
    "omx_bsr_from_csc_z": (_i, [_vp, _i, _vp, _vp, _vp, _i, _i, C.POINTER(_vp)]),
    "omx_set_element_z": (_i, [_vp, _i, _vp, _i64, _i, _i, _d, _d, _d]),
    "omx_lattice_bsr_k": (_i, [_i, _i, _i, C.c_uint64, _i, _vp, C.POINTER(_i64), _vp, _vp, _vp]),
    "omx_bsr_get_element": (_i, [_vp, _i, _i, C.POINTER(_d)]),
    "omx_spmm": (_i, [_vp, _i, _i, _i, _d, _vp, _i64, _vp, _d, _vp, _i64]),
    "omx_gemm": (_i, [_vp, _i, _i, _c, _c, _i, _i, _i, _d, _vp, _i64, _vp, _i64, _d, _vp, _i64]),
    "omx_gemm_sym": (_i, [_vp, _i, _i, _i, _i, _d, _vp, _i64, _vp, _i64, _vp, _i64]),
    "omx_gemm_set_tile": (_i, [_vp, _i]),
    "omx_spmm_set_tile": (_i, [_vp, _i, _i, _i, _i]),
    "omx_bsr_set_walk": (_i, [_vp, _vp]),
    "omx_bsr_set_walk_groups": (_i, [_vp, _vp, _i]),
    "omx_lattice_walk3": (_i, [_i, _i, _i, _i, _i, _i, _vp]),
    "omx_lattice_walk": (_i, [_i, _i, _i, _i, _i, _vp]),
    "omx_autotune": (_i, [_vp, _i]),
    "omx_group_create": (_i, [_vp, _i, _vp, _vp, _vp, C.POINTER(_vp)]),
    "omx_group_destroy": (_i, [_vp]),
    "omx_group_set_c": (_i, [_vp, _i, _vp]),
    "omx_group_get_c": (_i, [_vp, _i, _vp]),
    "omx_group_solve": (_i, [_vp, _vp, _d, _i, _i, _i, _vp, _vp, _vp, C.POINTER(_i)]),
    "omx_autotune_report": (_i, [_vp, _vp, _i64]),
    "omx_add": (_i, [_vp, _i, _c, _i, _i, _d, _vp, _i64, _d, _vp, _i64]),
    "omx_scale": (_i, [_vp, _i, _i, _i, _d, _vp, _i64]),
    "omx_set": (_i, [_vp, _i, _c, _i, _i, _d, _d, _vp, _i64]),
    "omx_trace": (_i, [_vp, _i, _i, _vp, _i64, _pd]),
    "omx_mm_trace": (_i, [_vp, _i, _i, _i, _vp, _i64, _vp, _i64, _pd]),
    "omx_set_element": (_i, [_vp, _i, _vp, _i64, _i, _i, _d, _d]),
    "omx_get_element": (_i, [_vp, _i, _vp, _i64, _i, _i, _pd]),
    "omx_omm_coeff": (_i, [_vp, _i, _i, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "omx_omm_calc_G": (_i, [_vp, _i, _i, _i, _i, _vp, _vp, _i64, _vp, _vp, _vp, _i64]),
    "omx_omm_pr_dots": (_i, [_vp, _i, _i, _i, _vp, _vp, _i64, _vp]),
    "omx_spmm_stencil": (_i, [_vp, _i]),
    "omx_bsr_lattice": (_i, [_vp, _vp]),
    "omx_omm_pr_dots_noalloc": (_i, [_vp, _i, _i, _i, _vp, _vp, _i64, _vp]),
    "omx_omm_cg_direction": (_i, [_vp, _i, _i, _i, _d, _vp, _vp, _vp, _i64]),
    "omx_omm_update_w": (_i, [_vp, _i, _i, _i64, _d, _vp, _vp, _vp]),
    "omx_prof_enable": (_i, [_vp, _i]),
    "omx_prof_read": (_i, [_vp, _i, _pd, C.POINTER(_i64)]),
    "omx_prof_reset": (_i, [_vp]),
    "omx_prof_work": (_i, [_vp, _i, _pd, _pd]),
    "omx_lattice_bsr": (_i, [_i, _i, _i, C.c_uint64, _i, C.POINTER(_i64), _vp, _vp, _vp]),
    "omx_lattice_bsr_theta": (_i, [_i, _i, _i, _vp, C.POINTER(_i64), _vp, _vp, _vp]),
    "omx_lattice_theta0": (_i, [C.c_uint64, _vp]),
    "omx_lattice_theta_jacobian": (_i, [_i, _i, _i, C.POINTER(_i64), _vp, _vp, _vp]),
    "omx_random_c0": (_i, [_i, _i, C.c_uint64, _vp, _i64]),
    "omx_u01": (_d, [C.c_uint64, C.c_uint64, C.c_uint64]),
}

_lib = None


class OmxError(RuntimeError):
    pass


def lib():
    """The loaded libommhip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OmxError(f"{LIB_PATH} not found: build it with omm-bundle_amd/build.sh")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int, what: str = ""):
    if rc != 0:
        msg = lib().omx_last_error().decode(errors="replace")
        raise OmxError(f"{what} failed ({rc}): {msg}")


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def device_count() -> int:
    c = C.c_int(0)
    check(lib().omx_device_count(C.byref(c)), "omx_device_count")
    return c.value


# ---------------------------------------------------------------------------------------------
# synthetic lattice inputs (host)
# ---------------------------------------------------------------------------------------------

DEFAULT_SEED = 0x0EEDC0DE
THETA_COUNT = 20


def csc_to_bsr(nb, colptr, rowidx, val, ptrbase=0, rowbase=0, is_complex=False):
    """BSR-4 block columns (bcolptr, browidx, bval) of a CSC matrix, on the host (omx_csc_to_bsr)."""
    colptr = np.ascontiguousarray(colptr, np.int32)
    rowidx = np.ascontiguousarray(rowidx, np.int32)
    val = np.ascontiguousarray(val, np.complex128 if is_complex else np.float64)
    nnzb = C.c_int64(0)
    check(lib().omx_csc_to_bsr(nb, _ptr(colptr), _ptr(rowidx), _ptr(val), ptrbase, rowbase, int(is_complex),
                               C.byref(nnzb), None, None, None), "omx_csc_to_bsr")
    cp = np.zeros((nb + 3) // 4 + 1, np.int32)
    ri = np.zeros(nnzb.value, np.int32)
    bv = np.zeros((nnzb.value, 32 if is_complex else 16), np.float64)
    check(lib().omx_csc_to_bsr(nb, _ptr(colptr), _ptr(rowidx), _ptr(val), ptrbase, rowbase, int(is_complex),
                               C.byref(nnzb), _ptr(cp), _ptr(ri), _ptr(bv)), "omx_csc_to_bsr")
    return cp, ri, bv


def lattice_bsr(L, which: int, seed: int = DEFAULT_SEED):
    """BSR-4 arrays (bcolptr, browidx, bval[nnzb,16]) of H (which=0) or S (which=1)."""
    Lx, Ly, Lz = (L, L, L) if np.isscalar(L) else L
    nnzb = C.c_int64(0)
    check(lib().omx_lattice_bsr(Lx, Ly, Lz, seed, which, C.byref(nnzb), None, None, None), "omx_lattice_bsr")
    ns = Lx * Ly * Lz
    cp = np.zeros(ns + 1, np.int32)
    ri = np.zeros(nnzb.value, np.int32)
    bv = np.zeros((nnzb.value, 16), np.float64)
    check(lib().omx_lattice_bsr(Lx, Ly, Lz, seed, which, C.byref(nnzb), _ptr(cp), _ptr(ri), _ptr(bv)),
          "omx_lattice_bsr")
    return cp, ri, bv


def lattice_bsr_k(L, which: int, phi, seed: int = DEFAULT_SEED):
    """Complex Bloch operator H(k) / S(k): (bcolptr, browidx, bval[nnzb,16] complex128)."""
    Lx, Ly, Lz = (L, L, L) if np.isscalar(L) else L
    phi = np.ascontiguousarray(phi, np.float64)
    nnzb = C.c_int64(0)
    check(lib().omx_lattice_bsr_k(Lx, Ly, Lz, seed, which, _ptr(phi), C.byref(nnzb), None, None, None), "lattice k")
    cp = np.zeros(Lx * Ly * Lz + 1, np.int32)
    ri = np.zeros(nnzb.value, np.int32)
    bv = np.zeros((nnzb.value, 16), np.complex128)
    check(lib().omx_lattice_bsr_k(Lx, Ly, Lz, seed, which, _ptr(phi), C.byref(nnzb), _ptr(cp), _ptr(ri), _ptr(bv)),
          "lattice k")
    return cp, ri, bv


def lattice_bsr_theta(L, theta):
    Lx, Ly, Lz = (L, L, L) if np.isscalar(L) else L
    theta = np.ascontiguousarray(theta, np.float64)
    nnzb = C.c_int64(0)
    check(lib().omx_lattice_bsr_theta(Lx, Ly, Lz, _ptr(theta), C.byref(nnzb), None, None, None), "lattice theta")
    ns = Lx * Ly * Lz
    cp = np.zeros(ns + 1, np.int32)
    ri = np.zeros(nnzb.value, np.int32)
    bv = np.zeros((nnzb.value, 16), np.float64)
    check(lib().omx_lattice_bsr_theta(Lx, Ly, Lz, _ptr(theta), C.byref(nnzb), _ptr(cp), _ptr(ri), _ptr(bv)),
          "lattice theta")
    return cp, ri, bv


def lattice_theta_jacobian(L):
    """Jacobian of H(theta) on the lattice block pattern: (seg[21], ent, coef); parameter p owns
    entries seg[p]..seg[p+1]-1, ent = index 16*l + r + 4*c into the block values."""
    Lx, Ly, Lz = (L, L, L) if np.isscalar(L) else L
    nent = C.c_int64(0)
    check(lib().omx_lattice_theta_jacobian(Lx, Ly, Lz, C.byref(nent), None, None, None), "theta jacobian")
    seg = np.zeros(THETA_COUNT + 1, np.int64)
    ent = np.zeros(max(nent.value, 1), np.int64)
    coef = np.zeros(max(nent.value, 1), np.float64)
    check(lib().omx_lattice_theta_jacobian(Lx, Ly, Lz, C.byref(nent), _ptr(seg), _ptr(ent), _ptr(coef)),
          "theta jacobian")
    return seg, ent[:nent.value], coef[:nent.value]


def lattice_walk(L, tx: int, ty: int) -> np.ndarray:
    """Tiled visiting order of the lattice sites (x-y tiles of tx x ty swept through z)."""
    Lx, Ly, Lz = (L, L, L) if np.isscalar(L) else L
    order = np.zeros(Lx * Ly * Lz, np.int32)
    check(lib().omx_lattice_walk(Lx, Ly, Lz, tx, ty, _ptr(order)), "lattice walk")
    return order


def lattice_walk3(L, tx: int, ty: int, tz: int) -> np.ndarray:
    """Patch visiting order of the lattice sites: tx x ty x tz patches listed one after another."""
    Lx, Ly, Lz = (L, L, L) if np.isscalar(L) else L
    order = np.zeros(Lx * Ly * Lz, np.int32)
    check(lib().omx_lattice_walk3(Lx, Ly, Lz, tx, ty, tz, _ptr(order)), "lattice walk3")
    return order


def lattice_theta0(seed: int = DEFAULT_SEED):
    th = np.zeros(20, np.float64)
    check(lib().omx_lattice_theta0(seed, _ptr(th)), "theta0")
    return th


def random_c0(no: int, nb: int, seed: int = DEFAULT_SEED):
    c = np.zeros((no, nb), np.float64, order="F")
    check(lib().omx_random_c0(no, nb, seed, _ptr(c), no), "omx_random_c0")
    return c


def u01(seed, stream, idx) -> float:
    return lib().omx_u01(seed, stream, idx)


def param_grad(ctx, Dp: "Bsr", seg_d: "DMat", ent_d: "DMat", coef_d: "DMat", nparam: int) -> np.ndarray:
    """dE/dtheta from dE/dH on the pattern (Dp) and a device Jacobian (omx_bsr_param_grad)."""
    out = np.zeros(nparam)
    check(lib().omx_bsr_param_grad(ctx.h, Dp.h, nparam, seg_d.p, ent_d.p, coef_d.p, _ptr(out)), "param_grad")
    return out


def bsr_get_element(Dp: "Bsr", i: int, j: int) -> float:
    out = C.c_double(0)
    check(lib().omx_bsr_get_element(Dp.h, i, j, C.byref(out)), "bsr_get_element")
    return out.value


def bsr_to_dense(nb: int, cp, ri, bv) -> np.ndarray:
    """Dense nb x nb matrix of a BSR-4 operator (host, for checks)."""
    nbp = 4 * (len(cp) - 1)
    A = np.zeros((nbp, nbp), dtype=np.result_type(np.asarray(bv).dtype, np.float64))
    for J in range(len(cp) - 1):
        for l in range(cp[J], cp[J + 1]):
            I = ri[l]
            A[4 * I:4 * I + 4, 4 * J:4 * J + 4] = bv[l].reshape(4, 4, order="F")
    return A[:nb, :nb]


def bsr_to_csc(nb: int, cp, ri, bv):
    """0-based CSC (colptr, rowidx, val) with all stored block entries, rows ascending."""
    colptr = [0]
    rows, vals = [], []
    for j in range(nb):
        J, c = divmod(j, 4)
        for l in range(cp[J], cp[J + 1]):
            I = ri[l]
            for r in range(4):
                i = 4 * I + r
                if i < nb:
                    rows.append(i)
                    vals.append(bv[l][r + 4 * c])
        colptr.append(len(rows))
    return np.array(colptr, np.int32), np.array(rows, np.int32), np.array(vals, np.float64)


# ---------------------------------------------------------------------------------------------
# device context, matrices, operators
# ---------------------------------------------------------------------------------------------


class Ctx:
    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(lib().omx_init(device, C.byref(h)), "omx_init")
        self.h = h

    def close(self):
        if self.h:
            lib().omx_finalize(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def sync(self):
        check(lib().omx_sync(self.h), "omx_sync")

    # -- tuning --
    def autotune(self, enable=True):
        check(lib().omx_autotune(self.h, int(enable)), "autotune")

    def autotune_report(self) -> str:
        buf = C.create_string_buffer(1 << 16)
        lib().omx_autotune_report(self.h, buf, len(buf))
        return buf.value.decode()

    def spmm_stencil(self, enable=True):
        """Use the 2.5-D stencil SpMM for recognised lattice operators (default off)."""
        check(lib().omx_spmm_stencil(self.h, int(enable)), "spmm_stencil")

    def spmm_shape(self, batch=0, wpg=4, jper=1, mode=0):
        """Force the real SpMM launch shape (batch 0: back to the default or tuned shape)."""
        check(lib().omx_spmm_set_tile(self.h, batch, wpg, jper, mode), "spmm_set_tile")

    # -- profiling --
    def prof(self, enable=True):
        check(lib().omx_prof_enable(self.h, int(enable)), "prof_enable")

    def prof_reset(self):
        check(lib().omx_prof_reset(self.h), "prof_reset")

    def prof_work(self, kclass):
        fl, by = C.c_double(0), C.c_double(0)
        check(lib().omx_prof_work(self.h, kclass, C.byref(fl), C.byref(by)), "prof_work")
        return fl.value, by.value

    def prof_read(self, kclass):
        ms, n = C.c_double(0), C.c_int64(0)
        check(lib().omx_prof_read(self.h, kclass, C.byref(ms), C.byref(n)), "prof_read")
        return ms.value, n.value


class DMat:
    """A device matrix (column-major, ld = rows), zero-initialised."""

    def __init__(self, ctx: Ctx, rows: int, cols: int, dt: int = OMX_F64, ld: int | None = None):
        self.ctx, self.rows, self.cols, self.dt = ctx, rows, cols, dt
        self.ld = ld if ld is not None else max(rows, 1)
        p = C.c_void_p()
        check(lib().omx_malloc(ctx.h, self.ld * max(cols, 1) * _ESIZE[dt], C.byref(p)), "omx_malloc")
        self.p = p

    @classmethod
    def from_host(cls, ctx, a: np.ndarray, dt: int = OMX_F64, ld: int | None = None):
        m = cls(ctx, a.shape[0], a.shape[1], dt, ld)
        m.upload(a)
        return m

    def upload(self, a: np.ndarray):
        a = np.asfortranarray(a, dtype=_NP[self.dt])
        assert a.shape == (self.rows, self.cols)
        check(lib().omx_h2d_2d(self.ctx.h, self.p, self.ld, _ptr(a), max(self.rows, 1), self.rows, self.cols,
                               _ESIZE[self.dt]), "h2d")

    def download(self) -> np.ndarray:
        out = np.zeros((self.rows, self.cols), _NP[self.dt], order="F")
        check(lib().omx_d2h_2d(self.ctx.h, _ptr(out), max(self.rows, 1), self.p, self.ld, self.rows, self.cols,
                               _ESIZE[self.dt]), "d2h")
        return out

    def free(self):
        if self.p:
            lib().omx_free(self.ctx.h, self.p)
            self.p = None

    def __del__(self):
        try:
            if self.p and self.ctx.h:
                self.free()
        except Exception:
            pass


class Bsr:
    """A BSR-4 operator in HBM."""

    def __init__(self, ctx: Ctx, nb: int, cp, ri, bv):
        self.ctx, self.nb = ctx, nb
        cp = np.ascontiguousarray(cp, np.int32)
        ri = np.ascontiguousarray(ri, np.int32)
        bv = np.ascontiguousarray(bv, np.float64)
        h = C.c_void_p()
        check(lib().omx_bsr_create(ctx.h, nb, _ptr(cp), _ptr(ri), _ptr(bv), C.byref(h)), "omx_bsr_create")
        self.h = h
        self.nnzb = len(ri)

    @classmethod
    def from_csc(cls, ctx, nb, colptr, rowidx, val, ptrbase=0, rowbase=0):
        self = cls.__new__(cls)
        self.ctx, self.nb = ctx, nb
        colptr = np.ascontiguousarray(colptr, np.int32)
        rowidx = np.ascontiguousarray(rowidx, np.int32)
        val = np.ascontiguousarray(val, np.float64)
        h = C.c_void_p()
        check(lib().omx_bsr_from_csc(ctx.h, nb, _ptr(colptr), _ptr(rowidx), _ptr(val), ptrbase, rowbase,
                                     C.byref(h)), "omx_bsr_from_csc")
        self.h = h
        nnzb = C.c_int64(0)
        check(lib().omx_bsr_info(h, None, C.byref(nnzb), None), "bsr_info")
        self.nnzb = nnzb.value
        return self

    @classmethod
    def complex(cls, ctx, nb, cp, ri, bv):
        """Complex operator from BSR-4 arrays with complex128 values (nnzb, 16)."""
        self = cls.__new__(cls)
        self.ctx, self.nb = ctx, nb
        cp = np.ascontiguousarray(cp, np.int32)
        ri = np.ascontiguousarray(ri, np.int32)
        bv = np.ascontiguousarray(bv, np.complex128)
        h = C.c_void_p()
        check(lib().omx_bsr_create_z(ctx.h, nb, _ptr(cp), _ptr(ri), _ptr(bv), C.byref(h)), "omx_bsr_create_z")
        self.h = h
        self.nnzb = len(ri)
        return self

    def values(self):
        out = np.zeros((self.nnzb, 16))
        check(lib().omx_bsr_get_values(self.h, _ptr(out)), "bsr_get_values")
        return out

    def pattern(self):
        """(block colptr, block row indices) of the stored blocks, 0-based."""
        nbc = (self.nb + 3) // 4
        cp = np.zeros(nbc + 1, np.int32)
        ri = np.zeros(max(self.nnzb, 1), np.int32)
        check(lib().omx_bsr_get_pattern(self.h, _ptr(cp), _ptr(ri)), "bsr_get_pattern")
        return cp, ri[: self.nnzb]

    def set_walk(self, order, group: int = 0):
        """Block-column visiting order of later SpMMs (None clears it), with LDS-staged groups of
        `group` block columns when group > 0; results do not change."""
        if order is None:
            check(lib().omx_bsr_set_walk(self.h, None), "bsr_set_walk")
            return
        o = np.ascontiguousarray(order, np.int32)
        check(lib().omx_bsr_set_walk_groups(self.h, _ptr(o), int(group)), "bsr_set_walk")

    def free(self):
        if self.h:
            lib().omx_bsr_destroy(self.h)
            self.h = None

    def __del__(self):
        # only while the context lives: destroying an operator synchronises its context's stream,
        # which is gone once the context is finalised (the memory is then released with the process)
        try:
            if self.h and self.ctx.h:
                self.free()
        except Exception:
            pass


# ---------------------------------------------------------------------------------------------
# thin op wrappers (raise on error)
# ---------------------------------------------------------------------------------------------


def spmm(ctx, X: DMat, A: Bsr, Y: DMat, alpha=1.0, beta=0.0):
    check(lib().omx_spmm(ctx.h, X.dt, X.rows, X.cols, alpha, X.p, X.ld, A.h, beta, Y.p, Y.ld), "omx_spmm")


def sddmm(ctx, A: DMat, B: DMat, Cp: Bsr, alpha=1.0, beta=0.0):
    """Cp := alpha*(A^T B) on Cp's blocks + beta*Cp."""
    check(lib().omx_sddmm_bsr(ctx.h, A.dt, A.rows, A.cols, alpha, A.p, A.ld, B.p, B.ld, beta, Cp.h), "omx_sddmm_bsr")


def gemm(ctx, opA, opB, A: DMat, B: DMat, Cm: DMat, alpha=1.0, beta=0.0, prec=PREC_F64):
    m = A.cols if opA in "tTcC" else A.rows
    k = A.rows if opA in "tTcC" else A.cols
    n = B.rows if opB in "tTcC" else B.cols
    check(lib().omx_gemm(ctx.h, prec, A.dt, opA.encode(), opB.encode(), m, n, k, alpha, A.p, A.ld, B.p, B.ld, beta,
                         Cm.p, Cm.ld), "omx_gemm")


def gemm_sym(ctx, A: DMat, B: DMat, Cm: DMat, alpha=1.0, prec=PREC_F64):
    """Cm := alpha*A*B^T for a product known to be symmetric (lower-triangle tiles, mirrored)."""
    check(lib().omx_gemm_sym(ctx.h, prec, A.dt, A.rows, A.cols, alpha, A.p, A.ld, B.p, B.ld, Cm.p, Cm.ld),
          "omx_gemm_sym")


def add(ctx, opA, A: DMat, Cm: DMat, alpha, beta):
    check(lib().omx_add(ctx.h, Cm.dt, opA.encode(), Cm.rows, Cm.cols, alpha, A.p, A.ld, beta, Cm.p, Cm.ld),
          "omx_add")


def scale(ctx, Cm: DMat, beta):
    check(lib().omx_scale(ctx.h, Cm.dt, Cm.rows, Cm.cols, beta, Cm.p, Cm.ld), "omx_scale")


def mset(ctx, seC, Cm: DMat, alpha, beta):
    check(lib().omx_set(ctx.h, Cm.dt, seC.encode(), Cm.rows, Cm.cols, alpha, beta, Cm.p, Cm.ld), "omx_set")


def trace(ctx, A: DMat):
    """tr(A); complex matrices give a complex value (the C call writes (re, im))."""
    out = np.zeros(2)
    check(lib().omx_trace(ctx.h, A.dt, A.rows, A.p, A.ld, out.ctypes.data_as(_pd)), "omx_trace")
    return complex(out[0], out[1]) if A.dt in (OMX_C128, OMX_C64) else float(out[0])


def mm_trace(ctx, A: DMat, B: DMat):
    """sum_ij A_ij B_ij (complex: sum conj(A_ij) B_ij, a complex value)."""
    out = np.zeros(2)
    check(lib().omx_mm_trace(ctx.h, A.dt, A.rows, A.cols, A.p, A.ld, B.p, B.ld, out.ctypes.data_as(_pd)),
          "omx_mm_trace")
    return complex(out[0], out[1]) if A.dt in (OMX_C128, OMX_C64) else float(out[0])


def omm_coeff(ctx, HW, SW, HWd, SWd, HWdd, SWdd) -> np.ndarray:
    out = np.zeros(5)
    check(lib().omx_omm_coeff(ctx.h, HW.dt, HW.rows, HW.ld, HW.p, SW.p, HWd.p, SWd.p, HWdd.p, SWdd.p, _ptr(out)),
          "omx_omm_coeff")
    return out


def omm_calc_G(ctx, HW, SW, HC, SC, G, prec=PREC_F64):
    check(lib().omx_omm_calc_G(ctx.h, prec, HC.dt, HC.rows, HC.cols, HW.p, SW.p, HW.ld, HC.p, SC.p, G.p, G.ld),
          "omx_omm_calc_G")


def omm_pr_dots(ctx, G, Gp):
    out = np.zeros(2)
    check(lib().omx_omm_pr_dots(ctx.h, G.dt, G.rows, G.cols, G.p, Gp.p, G.ld, _ptr(out)), "omx_omm_pr_dots")
    return out


def omm_cg_direction(ctx, lam, G, D, Gp):
    check(lib().omx_omm_cg_direction(ctx.h, G.dt, G.rows, G.cols, lam, G.p, D.p, Gp.p, G.ld), "cg_direction")


def omm_update_w(ctx, x, Wd, Wdd, W):
    check(lib().omx_omm_update_w(ctx.h, W.dt, W.rows, W.ld, x, Wd.p, Wdd.p, W.p), "update_w")


def set_element(ctx, M: DMat, i, j, alpha, beta):
    check(lib().omx_set_element(ctx.h, M.dt, M.p, M.ld, i, j, alpha, beta), "set_element")


def get_element(ctx, M: DMat, i, j) -> float:
    out = C.c_double(0)
    check(lib().omx_get_element(ctx.h, M.dt, M.p, M.ld, i, j, C.byref(out)), "get_element")
    return out.value


class ZBatch:
    """Batched complex128 k-point solver (omx_zbatch_*): nsys systems sharing one block pattern."""

    def __init__(self, ctx: Ctx, no: int, nb: int, cp, ri, hvals, svals, dt: int = OMX_C128):
        """hvals, svals: lists (one per system) of complex (nnzb, 16) block values, H already shifted;
        dt: element type of the state (OMX_C128 or OMX_C64)."""
        self.ctx, self.no, self.nb, self.nsys, self.dt = ctx, no, nb, len(hvals), dt
        cp = np.ascontiguousarray(cp, np.int32)
        ri = np.ascontiguousarray(ri, np.int32)
        hv = np.ascontiguousarray(np.stack(hvals), np.complex128)
        sv = np.ascontiguousarray(np.stack(svals), np.complex128)
        h = C.c_void_p()
        check(lib().omx_zbatch_create_dt(ctx.h, dt, self.nsys, no, nb, _ptr(cp), _ptr(ri), _ptr(hv), _ptr(sv),
                                         C.byref(h)), "omx_zbatch_create")
        self.h = h

    def set_c(self, Cs):
        t = _NP[self.dt]
        a = np.ascontiguousarray(np.stack([np.asfortranarray(c, t).ravel(order="F") for c in Cs]))
        check(lib().omx_zbatch_set_c(self.h, _ptr(a)), "zbatch_set_c")

    def get_c(self):
        a = np.zeros((self.nsys, self.no * self.nb), _NP[self.dt])
        check(lib().omx_zbatch_get_c(self.h, _ptr(a)), "zbatch_get_c")
        return [a[s].reshape((self.no, self.nb), order="F") for s in range(self.nsys)]

    def solve(self, eta=0.0, cg_tol=1e-9, max_steps=10000, check_every=10, use_graph=True):
        e = np.zeros(self.nsys)
        icg = np.zeros(self.nsys, np.int32)
        conv = np.zeros(self.nsys, np.int32)
        steps = C.c_int(0)
        check(lib().omx_zbatch_solve(self.h, eta, cg_tol, max_steps, check_every, int(use_graph), _ptr(e), _ptr(icg),
                                     _ptr(conv), C.byref(steps)), "zbatch_solve")
        return {"e_min": e, "icg": icg, "converged": conv.astype(bool), "steps": steps.value}

    def free(self):
        if self.h:
            lib().omx_zbatch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            if self.h and self.ctx.h:
                self.free()
        except Exception:
            pass


class Group:
    """Uneven-size group of real systems (omx_group_*): system s has no[s] states and its own
    operators H[s] (already shifted by -eta_s S) and S[s] (Bsr objects on ctx, kept alive here)."""

    def __init__(self, ctx: Ctx, no, H, S):
        self.ctx, self.nsys = ctx, len(no)
        self.no = [int(x) for x in no]
        self.nb = [h.nb for h in H]
        self._ops = list(H) + list(S)
        nos = np.ascontiguousarray(self.no, np.int32)
        hp = (C.c_void_p * self.nsys)(*[h.h.value if isinstance(h.h, C.c_void_p) else h.h for h in H])
        sp = (C.c_void_p * self.nsys)(*[x.h.value if isinstance(x.h, C.c_void_p) else x.h for x in S])
        g = C.c_void_p()
        check(lib().omx_group_create(ctx.h, self.nsys, _ptr(nos), C.cast(hp, C.c_void_p), C.cast(sp, C.c_void_p),
                                     C.byref(g)), "omx_group_create")
        self.h = g

    def set_c(self, s, Cm):
        a = np.asfortranarray(Cm, np.float64)
        assert a.shape == (self.no[s], self.nb[s])
        check(lib().omx_group_set_c(self.h, s, _ptr(a)), "group_set_c")

    def get_c(self, s):
        a = np.zeros((self.no[s], self.nb[s]), np.float64, order="F")
        check(lib().omx_group_get_c(self.h, s, _ptr(a)), "group_get_c")
        return a

    def solve(self, eta, cg_tol=1e-9, max_steps=10000, check_every=10, use_graph=True):
        eta = np.ascontiguousarray(np.broadcast_to(np.asarray(eta, np.float64), (self.nsys,)))
        e = np.zeros(self.nsys)
        icg = np.zeros(self.nsys, np.int32)
        conv = np.zeros(self.nsys, np.int32)
        steps = C.c_int(0)
        check(lib().omx_group_solve(self.h, _ptr(eta), cg_tol, max_steps, check_every, int(use_graph), _ptr(e),
                                    _ptr(icg), _ptr(conv), C.byref(steps)), "group_solve")
        return {"e_min": e, "icg": icg, "converged": conv.astype(bool), "steps": steps.value}

    def free(self):
        if self.h:
            lib().omx_group_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            if self.h and self.ctx.h:
                self.free()
        except Exception:
            pass
