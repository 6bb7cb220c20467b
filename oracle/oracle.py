"""ctypes wrapper of the CPU oracle (liboommoracle.so) -- TEST INFRASTRUCTURE ONLY.

The oracle restates libOMM's omm() flavour 0 and MatrixSwitch's 'ref' operations in C
(omm_oracle.c).  Only tests, the smoke check and the benchmark's CPU-baseline leg use it, as the
checker; the HIP product path never calls it.  Also drives the reference serial build in
oracle/_ref (ref_driver) when that has been built from the reference sources.
"""
from __future__ import annotations

import ctypes as C
import os
import struct
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboommoracle.so")
REF_DRIVER = os.path.join(HERE, "_ref", "ref_driver")
REF_MM = os.path.join(HERE, "_ref", "ref_mm")

_vp, _i, _d, _c = C.c_void_p, C.c_int, C.c_double, C.c_char
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} not built (make -C oracle)")
        L = C.CDLL(LIB_PATH)
        L.or_gemm.argtypes = [_c, _c, _i, _i, _i, _d, _vp, _i, _vp, _i, _d, _vp, _i]
        L.or_spmm_csc.argtypes = [_c, _i, _i, _d, _vp, _i, _vp, _vp, _vp, _d, _vp, _i]
        L.or_sddmm_csc.argtypes = [_c, _c, _i, _d, _vp, _i, _vp, _i, _i, _vp, _vp, _d, _vp]
        L.or_add.argtypes = [_c, _i, _i, _d, _vp, _i, _d, _vp, _i]
        L.or_scale.argtypes = [_i, _i, _d, _vp, _i]
        L.or_set.argtypes = [_c, _i, _i, _d, _d, _vp, _i]
        L.or_trace.argtypes = [_i, _vp, _i]
        L.or_trace.restype = _d
        L.or_mm_trace.argtypes = [_i, _i, _vp, _i, _vp, _i]
        L.or_mm_trace.restype = _d
        L.or_bsd_lcg.argtypes = [C.POINTER(C.c_int32)]
        L.or_bsd_lcg.restype = _d
        L.or_solve_quartic.argtypes = [_vp, C.POINTER(_d), C.POINTER(_i)]
        L.or_calc_coeff.argtypes = [_i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
        L.or_omm_create.argtypes = [_i]
        L.or_omm_create.restype = _vp
        L.or_omm_destroy.argtypes = [_vp]
        L.or_omm_flat.argtypes = [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _i, _vp, _vp, _vp, _vp, _i,
                                  C.POINTER(_d), _vp, _i, _d, _vp, _i, _i, _i, _d, _i, _i, _vp, _i, _vp, _vp]
        L.or_zgemm.argtypes = L.or_gemm.argtypes
        L.or_zspmm_csc.argtypes = L.or_spmm_csc.argtypes
        L.or_zadd.argtypes = L.or_add.argtypes
        L.or_zscale.argtypes = L.or_scale.argtypes
        L.or_zset.argtypes = L.or_set.argtypes
        L.or_ztrace.argtypes = [_i, _vp, _i, _vp]
        L.or_ztrace.restype = _d
        L.or_zmm_trace.argtypes = [_i, _i, _vp, _i, _vp, _i, _vp]
        L.or_zmm_trace.restype = _d
        L.or_zcalc_coeff.argtypes = L.or_calc_coeff.argtypes
        L.or_num_threads.restype = _i
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f(a):
    return np.asfortranarray(a, dtype=np.float64)


# ---- MatrixSwitch 'ref' ops on numpy (Fortran-ordered) arrays --------------------------------

def gemm(opA, opB, A, B, Cm, alpha=1.0, beta=0.0):
    assert Cm.flags.f_contiguous and Cm.dtype == np.float64
    A, B = _f(A), _f(B)
    m, n = Cm.shape
    k = A.shape[0] if opA in "tTcC" else A.shape[1]
    lib().or_gemm(opA.encode(), opB.encode(), m, n, k, alpha, _p(A), A.shape[0], _p(B), B.shape[0], beta, _p(Cm),
                  Cm.shape[0])
    return Cm


def spmm_csc(A, colptr, rowidx, val, Cm, alpha=1.0, beta=0.0, opA="n"):
    A = _f(A)
    colptr = np.ascontiguousarray(colptr, np.int32)
    rowidx = np.ascontiguousarray(rowidx, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    m, nb = Cm.shape
    lib().or_spmm_csc(opA.encode(), m, nb, alpha, _p(A), A.shape[0], _p(colptr), _p(rowidx), _p(val), beta, _p(Cm),
                      m)
    return Cm


def sddmm_csc(opA, opB, A, B, colptr, rowidx, cval, alpha=1.0, beta=0.0):
    """C := alpha*op(A)*op(B) + beta*C on C's CSC pattern (0-based colptr/rowidx); returns the
    updated values (MatrixSwitch_mm_multiply.F90:334-398)."""
    A, B = _f(A), _f(B)
    colptr = np.ascontiguousarray(colptr, np.int32)
    rowidx = np.ascontiguousarray(rowidx, np.int32)
    out = np.array(cval, np.float64, copy=True)
    n = len(colptr) - 1
    k = A.shape[0] if opA in "tTcC" else A.shape[1]
    lib().or_sddmm_csc(opA.encode(), opB.encode(), k, alpha, _p(A), A.shape[0], _p(B), B.shape[0], n, _p(colptr),
                       _p(rowidx), beta, _p(out))
    return out


def add(opA, A, Cm, alpha, beta):
    A = _f(A)
    m, n = Cm.shape
    lib().or_add(opA.encode(), m, n, alpha, _p(A), A.shape[0], beta, _p(Cm), m)
    return Cm


def trace(A):
    A = _f(A)
    return lib().or_trace(A.shape[0], _p(A), A.shape[0])


def mm_trace(A, B):
    A, B = _f(A), _f(B)
    return lib().or_mm_trace(A.shape[0], A.shape[1], _p(A), A.shape[0], _p(B), B.shape[0])


def mset(seC, Cm, alpha, beta):
    lib().or_set(seC.encode(), Cm.shape[0], Cm.shape[1], alpha, beta, _p(Cm), Cm.shape[0])
    return Cm


# ---- the same operations on complex128 ('szden' / 'szcsc') ------------------------------------

def _z(a):
    return np.asfortranarray(a, dtype=np.complex128)


def zgemm(opA, opB, A, B, Cm, alpha=1.0, beta=0.0):
    assert Cm.flags.f_contiguous and Cm.dtype == np.complex128
    A, B = _z(A), _z(B)
    m, n = Cm.shape
    k = A.shape[0] if opA in "tTcC" else A.shape[1]
    lib().or_zgemm(opA.encode(), opB.encode(), m, n, k, alpha, _p(A), A.shape[0], _p(B), B.shape[0], beta, _p(Cm),
                   Cm.shape[0])
    return Cm


def zspmm_csc(A, colptr, rowidx, val, Cm, alpha=1.0, beta=0.0, opA="n"):
    A = _z(A)
    colptr = np.ascontiguousarray(colptr, np.int32)
    rowidx = np.ascontiguousarray(rowidx, np.int32)
    val = np.ascontiguousarray(val, np.complex128)
    m, nb = Cm.shape
    lib().or_zspmm_csc(opA.encode(), m, nb, alpha, _p(A), A.shape[0], _p(colptr), _p(rowidx), _p(val), beta,
                       _p(Cm), m)
    return Cm


def zadd(opA, A, Cm, alpha, beta):
    A = _z(A)
    m, n = Cm.shape
    lib().or_zadd(opA.encode(), m, n, alpha, _p(A), A.shape[0], beta, _p(Cm), m)
    return Cm


def ztrace(A):
    A = _z(A)
    out = np.zeros(2)
    lib().or_ztrace(A.shape[0], _p(A), A.shape[0], _p(out))
    return complex(out[0], out[1])


def zmm_trace(A, B):
    A, B = _z(A), _z(B)
    out = np.zeros(2)
    lib().or_zmm_trace(A.shape[0], A.shape[1], _p(A), A.shape[0], _p(B), B.shape[0], _p(out))
    return complex(out[0], out[1])


def zcalc_coeff(HW, SW, HWd, SWd, HWdd, SWdd):
    n = HW.shape[0]
    mats = [_z(X) for X in (HW, SW, HWd, SWd, HWdd, SWdd)]
    work = np.zeros((n, n), np.complex128, order="F")
    coeff = np.zeros(5)
    lib().or_zcalc_coeff(n, *[_p(X) for X in mats], _p(work), _p(coeff))
    return coeff


def bsd_lcg_seq(seed: int, count: int):
    x = C.c_int32(seed)
    out = []
    for _ in range(count):
        r = lib().or_bsd_lcg(C.byref(x))
        out.append((x.value, r))
    return out


def solve_quartic(c):
    c = np.ascontiguousarray(c, np.float64)
    x, fail = C.c_double(0), C.c_int(0)
    lib().or_solve_quartic(_p(c), C.byref(x), C.byref(fail))
    return x.value, bool(fail.value)


def calc_coeff(HW, SW, HWd, SWd, HWdd, SWdd):
    n = HW.shape[0]
    mats = [_f(M) for M in (HW, SW, HWd, SWd, HWdd, SWdd)]
    work = np.zeros((n, n), order="F")
    out = np.zeros(5)
    lib().or_calc_coeff(n, *[_p(M) for M in mats], _p(work), _p(out))
    return out


def num_threads():
    return lib().or_num_threads()


# ---- omm() -------------------------------------------------------------------------------------

class Omm:
    """omm() flavour 0 with per-ip state, mirroring the reference call sequence."""

    def __init__(self, np_: int = 1):
        self.h = lib().or_omm_create(np_)

    def close(self):
        if self.h:
            lib().or_omm_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __call__(self, m, n, H, S, C_min, *, new_S=True, calc_ED=False, eta=0.0, init_C=True, ip=1, cg_tol=-1.0,
                 dealloc=False, max_iter=0, want_D=True, log_cap=100000, flavour=0):
        """H, S: dense (m x m array) or a CSC tuple (colptr, rowidx, val), 0-based; S may be None.
        C_min: Fortran-ordered n x m array, updated in place; float64 runs the real ('sdden'/'sdcsc')
        operations, complex128 the complex ('szden'/'szcsc') ones, and H, S are taken in the same
        type.  Returns a dict."""
        assert C_min.flags.f_contiguous and C_min.shape == (n, m)
        cplx = C_min.dtype == np.complex128
        assert cplx or C_min.dtype == np.float64
        vt = np.complex128 if cplx else np.float64
        keep = []

        def op(A):
            if A is None:
                return 0, None, None, None, None
            if isinstance(A, tuple):
                cp, ri, v = (np.ascontiguousarray(A[0], np.int32), np.ascontiguousarray(A[1], np.int32),
                             np.ascontiguousarray(A[2], vt))
                keep.extend([cp, ri, v])
                return 2, None, _p(cp), _p(ri), _p(v)
            Ad = np.asfortranarray(A, dtype=vt)
            keep.append(Ad)
            return 1, _p(Ad), None, None, None

        hk, hd, hc, hr, hv = op(H)
        sk, sd, sc, sr, sv = op(S)
        D = np.zeros((m, m), vt, order="F") if want_D else None
        e = C.c_double(0)
        elog = np.zeros(log_cap)
        info = np.zeros(4, np.int32)
        out = np.zeros(2)
        rc = lib().or_omm_flat(self.h, int(cplx), m, n, hk, hd, hc, hr, hv, sk, sd, sc, sr, sv, int(new_S), C.byref(e), _p(D),
                               int(calc_ED), eta, _p(C_min), int(init_C), flavour, ip, cg_tol, int(dealloc), max_iter,
                               _p(elog), log_cap, _p(info), _p(out))
        if rc != 0:
            raise RuntimeError(f"or_omm failed with {rc}")
        return {"e_min": e.value, "D": D, "icg": int(info[0]), "converged": bool(info[1]), "nfail": int(info[2]),
                "e_log": elog[:int(info[3])].copy(), "TrQS": out[0]}


# ---- the reference serial build (oracle/_ref) --------------------------------------------------

def have_ref() -> bool:
    return os.path.exists(REF_DRIVER)


def run_ref(m, n, calls, s_present=True, sparse=False, workdir=None):
    """Replay omm() calls through the reference build.  Each call is a dict with keys H (m x m),
    S (m x m, if s_present), optional C (n x m, given to the solver), new_S, init_C, calc_ED,
    dealloc, eta, cg_tol.  Complex H (complex128) runs the whole replay on 'szden' storage (dense
    only).  Returns a list of dicts (e_min, C, D) and the libOMM.log text."""
    cplx = np.iscomplexobj(calls[0]["H"])
    assert not (cplx and sparse), "the reference's complex CSC conversion is not usable"
    vt = np.complex128 if cplx else np.float64
    own = workdir is None
    wd = tempfile.mkdtemp(prefix="omm_ref_") if own else workdir
    fin, fout = os.path.join(wd, "in.bin"), os.path.join(wd, "out.bin")
    with open(fin, "wb") as f:
        f.write(struct.pack("=6i", m, n, len(calls), int(s_present), int(sparse), int(cplx)))
        for c in calls:
            give_C = int(c.get("C") is not None)
            f.write(struct.pack("=6i", int(c.get("new_S", True)), int(c.get("init_C", False)),
                                int(c.get("calc_ED", False)), int(c.get("dealloc", False)), give_C,
                                int(c.get("flavour", 0))))
            f.write(struct.pack("=2d", float(c.get("eta", 0.0)), float(c.get("cg_tol", -1.0))))
            f.write(np.asfortranarray(c["H"], vt).tobytes(order="F"))
            if s_present:
                f.write(np.asfortranarray(c["S"], vt).tobytes(order="F"))
            if give_C:
                f.write(np.asfortranarray(c["C"], vt).tobytes(order="F"))
    subprocess.run([REF_DRIVER, fin, fout], cwd=wd, check=True, capture_output=True)
    raw = open(fout, "rb").read()
    es = 16 if cplx else 8
    res, off = [], 0
    for _ in calls:
        e = struct.unpack_from("=d", raw, off)[0]
        off += 8
        Cm = np.frombuffer(raw, vt, n * m, off).reshape((n, m), order="F")
        off += n * m * es
        D = np.frombuffer(raw, vt, m * m, off).reshape((m, m), order="F")
        off += m * m * es
        res.append({"e_min": float(e), "C": Cm.copy(), "D": D.copy()})
    log = open(os.path.join(wd, "libOMM.log")).read() if os.path.exists(os.path.join(wd, "libOMM.log")) else ""
    return res, log


def run_ref_sddmm(opA, opB, A, B, C0, alpha, beta):
    """One mm_multiply(A, opA, B, opB, C, alpha, beta, 'ref') of the reference build with C stored
    as 'sdcsc' on the nonzeros of C0 (oracle/ref_mm.F90).  Returns (colptr 0-based, rowidx 0-based,
    values)."""
    m, n = C0.shape
    k = A.shape[0] if opA in "tT" else A.shape[1]
    wd = tempfile.mkdtemp(prefix="omm_refmm_")
    fin, fout = os.path.join(wd, "in.bin"), os.path.join(wd, "out.bin")
    with open(fin, "wb") as f:
        f.write(struct.pack("=5i", m, n, k, int(opA in "tT"), int(opB in "tT")))
        f.write(struct.pack("=2d", alpha, beta))
        for M in (A, B, C0):
            f.write(np.asfortranarray(M, np.float64).tobytes(order="F"))
    subprocess.run([REF_MM, fin, fout], cwd=wd, check=True, capture_output=True)
    raw = open(fout, "rb").read()
    nnz = struct.unpack_from("=i", raw, 0)[0]
    off = 4
    colptr = np.frombuffer(raw, np.int32, n + 1, off)
    off += 4 * (n + 1)
    rowidx = np.frombuffer(raw, np.int32, nnz, off) - 1
    off += 4 * nnz
    val = np.frombuffer(raw, np.float64, nnz, off)
    return colptr.copy(), rowidx.copy(), val.copy()


def icg_from_log(log: str):
    """The 'minim: icg' counts printed by the reference (omm.F90:588), in call order."""
    out = []
    for line in log.splitlines():
        if "minim: icg" in line:
            out.append(int(line.split("=")[1].split("|")[0]))
    return out


def e_log_from_log(log: str):
    """Per-step energies printed with long_out (omm.F90:504), grouped per call."""
    calls, cur = [], None
    for line in log.splitlines():
        s = line.strip()
        if s.startswith("| libOMM"):
            cur = []
            calls.append(cur)
        elif cur is not None and s.startswith("|") and len(s.split()) == 6:
            parts = s.strip("|").split()
            try:
                cur.append(float(parts[2]))
            except ValueError:
                pass
    return calls


def read_example_gamma(path):
    """example_gamma.dat: header 'm n num', then 'i j H_ij S_ij' (1-based)."""
    with open(path) as f:
        m, n, num = (int(x) for x in f.readline().split())
        H = np.zeros((m, m))
        S = np.zeros((m, m))
        for _ in range(num):
            i, j, h, s = f.readline().split()
            H[int(i) - 1, int(j) - 1] = float(h)
            S[int(i) - 1, int(j) - 1] = float(s)
    return m, n, H, S


def read_example_kpoints(path_gamma, path_kpoints):
    """The two k-points of example_kpoints.F90:116-131, 173-199: k-point 1 is example_gamma.dat
    (real H, S taken as complex), k-point 2 example_kpoints.dat ('i j Re H Im H Re S Im S', 1-based,
    its first line a header).  Returns m, n, [(H1, S1), (H2, S2)] as complex128 arrays."""
    m, n, H1, S1 = read_example_gamma(path_gamma)
    H2 = np.zeros((m, m), np.complex128)
    S2 = np.zeros((m, m), np.complex128)
    with open(path_kpoints) as f:
        f.readline()
        for line in f:
            p = line.split()
            if len(p) < 6:
                continue
            i, j = int(p[0]) - 1, int(p[1]) - 1
            H2[i, j] = complex(float(p[2]), float(p[3]))
            S2[i, j] = complex(float(p[4]), float(p[5]))
    return m, n, [(H1.astype(np.complex128), S1.astype(np.complex128)), (H2, S2)]
