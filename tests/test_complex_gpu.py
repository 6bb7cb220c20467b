"""Complex (k-point) path on the device: the complex MFMA products, the complex BSR SpMM, the
complex element-wise and reduction kernels against numpy complex128 references; complex omm()
trajectories against the oracle's complex ('szden') table, itself bit-identical to the reference
build on the same inputs (the reference's example_kpoints data, a Bloch lattice with S and eta, the
no-S branch); and complex omm() solves against the closed form (sum of the occupied eigenvalues of
the Hermitian generalised problem H(k) x = lambda S(k) x) and against the real solve at k = 0.

Tolerances: complex128 paths 1e-12 relative to the operand scale (FMA/summation order); complex64
paths 1e-4; against the oracle (as for the real solves) converged energies 1e-10 relative, per-step
energies 1e-9 relative, density matrices 1e-8 absolute, CG step counts within 2; converged energies
1e-10 relative to the real solve at k = 0 and to the closed form."""
import os

import numpy as np
import pytest
import scipy.linalg as sla

import ommhip as oh
import oracle as O
import pyomm as F

pytestmark = pytest.mark.gpu
rng = np.random.default_rng(21)


def _note(*a):
    print("CASE", *a, flush=True)


def crandn(*shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _op(X, op):
    return X if op == "n" else (X.T if op == "t" else X.conj().T)


@pytest.mark.parametrize("dt,tol", [(oh.OMX_C128, 1e-12), (oh.OMX_C64, 2e-5)])
@pytest.mark.parametrize("opA,opB", [("n", "n"), ("n", "c"), ("c", "n"), ("t", "n"), ("n", "t"), ("c", "c")])
@pytest.mark.parametrize("shape", [(70, 90, 40), (129, 64, 300)])
def test_zgemm(ctx, dt, tol, opA, opB, shape):
    m, n, k = shape
    _note("zgemm", dt, opA, opB, shape)
    A = crandn(*((k, m) if opA != "n" else (m, k)))
    B = crandn(*((n, k) if opB != "n" else (k, n)))
    C0 = crandn(m, n)
    dA, dB, dC = (oh.DMat.from_host(ctx, M, dt) for M in (A, B, C0))
    oh.gemm(ctx, opA, opB, dA, dB, dC, alpha=-1.5, beta=0.5)
    if dt == oh.OMX_C64:
        A, B, C0 = (M.astype(np.complex64).astype(np.complex128) for M in (A, B, C0))
    ref = -1.5 * _op(A, opA) @ _op(B, opB) + 0.5 * C0
    assert np.max(np.abs(dC.download() - ref)) <= tol * np.sqrt(k) * 8


@pytest.mark.parametrize("dt", [oh.OMX_C128, oh.OMX_C64])
@pytest.mark.parametrize("n,k", [(70, 40), (2048 + 70, 100), (2048, 64)])
def test_zgemm_herm(ctx, dt, n, k):
    """omx_gemm_sym on complex storage, AW = (C H) C^H with H Hermitian: the strict upper triangle is
    exactly the conjugate mirror of the lower one.  The lower triangle is bitwise the full 'n','c'
    product's where both launches take the same K split (the small shape); the large shapes take
    the lower-triangle tiles (306 of 578 and 272 of 512 tiles of 64 x 128, one with partial edge
    tiles), fewer than two per CU, so they are split along K where the full product is not:
    rounding-level differences, 1e-12 (complex128) and 1e-5 (complex64) of the product's scale."""
    _note("zgemm_herm", dt, n, k)
    C = crandn(n, k)
    Hm = crandn(k, k)
    Hm = Hm + Hm.conj().T
    AC = C @ Hm
    dAC, dC = (oh.DMat.from_host(ctx, M, dt) for M in (AC, C))
    full, herm = oh.DMat(ctx, n, n, dt), oh.DMat(ctx, n, n, dt)
    oh.gemm(ctx, "n", "c", dAC, dC, full)
    oh.gemm_sym(ctx, dAC, dC, herm)
    F_, S_ = full.download(), herm.download()
    il, iu = np.tril_indices(n), np.triu_indices(n, 1)
    if n < 128:
        assert np.array_equal(S_[il], F_[il])
    else:
        scale = np.max(np.abs(F_))
        assert np.max(np.abs(S_[il] - F_[il])) <= (1e-12 if dt == oh.OMX_C128 else 1e-5) * scale
    assert np.array_equal(S_[iu], S_.T[iu].conj())
    tol = 1e-12 if dt == oh.OMX_C128 else 2e-5
    assert np.max(np.abs(S_ - AC @ C.conj().T)) <= tol * np.sqrt(k) * 8 * np.max(np.abs(AC))


@pytest.mark.parametrize("dt,tol", [(oh.OMX_C128, 1e-12), (oh.OMX_C64, 2e-5)])
def test_zcalc_G(ctx, dt, tol):
    n, m = 48, 100
    _note("zcalc_G", dt)
    HW, SW, HC, SC = crandn(n, n), crandn(n, n), crandn(n, m), crandn(n, m)
    d = [oh.DMat.from_host(ctx, M, dt) for M in (HW, SW, HC, SC)]
    G = oh.DMat(ctx, n, m, dt)
    oh.omm_calc_G(ctx, d[0], d[1], d[2], d[3], G)
    ref = 4 * HC - 2 * SW @ HC - 2 * HW @ SC
    assert np.max(np.abs(G.download() - ref)) < tol * np.sqrt(n) * 20


@pytest.mark.parametrize("dt,tol", [(oh.OMX_C128, 1e-12), (oh.OMX_C64, 1e-4)])
@pytest.mark.parametrize("L,no", [(3, 54), (4, 127)])
def test_zspmm_bloch(ctx, dt, tol, L, no):
    _note("zspmm", dt, L, no)
    phi = [0.25, 0.5, 0.125]
    cp, ri, bv = oh.lattice_bsr_k(L, 0, phi)
    nb = 4 * L ** 3
    Hd = oh.bsr_to_dense(nb, cp, ri, bv)
    assert np.allclose(Hd, Hd.conj().T, atol=1e-15)
    A = oh.Bsr.complex(ctx, nb, cp, ri, bv)
    X, Y0 = crandn(no, nb), crandn(no, nb)
    dX, dY = oh.DMat.from_host(ctx, X, dt), oh.DMat.from_host(ctx, Y0, dt)
    oh.spmm(ctx, dX, A, dY, alpha=1.5, beta=-0.5)
    ref = 1.5 * X @ Hd - 0.5 * Y0
    assert np.max(np.abs(dY.download() - ref)) < tol * 40


@pytest.mark.parametrize("dt", [oh.OMX_C128, oh.OMX_C64])
def test_zelementwise(ctx, dt):
    n, m = 45, 70
    _note("zelementwise", dt)
    tol = 1e-12 if dt == oh.OMX_C128 else 1e-5
    A, C0 = crandn(m, n), crandn(n, m)
    for op in ("n", "t", "c"):
        src = A if op != "n" else C0
        dA, dC = oh.DMat.from_host(ctx, src, dt), oh.DMat.from_host(ctx, C0, dt)
        oh.add(ctx, op, dA, dC, 0.7, -1.25)
        ref = 0.7 * _op(src, op) - 1.25 * C0
        assert np.max(np.abs(dC.download() - ref)) < tol * 10
    dC = oh.DMat.from_host(ctx, C0, dt)
    oh.scale(ctx, dC, -0.5)
    assert np.max(np.abs(dC.download() + 0.5 * C0)) < tol
    W = crandn(n, n)
    dW = oh.DMat.from_host(ctx, W, dt)
    tr = np.zeros(2)
    oh.check(oh.lib().omx_trace(ctx.h, dt, n, dW.p, dW.ld, tr.ctypes.data_as(oh._pd)), "trace")
    assert abs(complex(*tr) - np.trace(W)) < tol * 100
    B = crandn(n, n)
    dB = oh.DMat.from_host(ctx, B, dt)
    oh.check(oh.lib().omx_mm_trace(ctx.h, dt, n, n, dW.p, dW.ld, dB.p, dB.ld, tr.ctypes.data_as(oh._pd)),
             "mm_trace")
    assert abs(complex(*tr) - np.sum(W.conj() * B)) < tol * 1000
    oh.mset(ctx, "a", dW, 0.25, 2.0)
    got = dW.download()
    assert np.allclose(got, np.where(np.eye(n, dtype=bool), 2.0, 0.25), atol=0)
    # CG updates: D = G + lam*D, Gp = G; W += x (Wd^H + Wd) + x^2 Wdd; PR dots (real parts)
    G, D = crandn(n, m), crandn(n, m)
    dG, dD, dGp = oh.DMat.from_host(ctx, G, dt), oh.DMat.from_host(ctx, D, dt), oh.DMat(ctx, n, m, dt)
    oh.omm_cg_direction(ctx, 0.3, dG, dD, dGp)
    assert np.max(np.abs(dD.download() - (G + 0.3 * D))) < tol * 10
    G2 = crandn(n, m)
    dG.upload(G2)
    num, den = oh.omm_pr_dots(ctx, dG, dGp)
    assert abs(num - np.real(np.sum(G2.conj() * (G2 - G)))) < tol * 1e4
    assert abs(den - np.real(np.sum(G.conj() * G))) < tol * 1e4
    W, Wd, Wdd = crandn(n, n), crandn(n, n), crandn(n, n)
    dW, dWd, dWdd = (oh.DMat.from_host(ctx, M, dt) for M in (W, Wd, Wdd))
    oh.omm_update_w(ctx, -0.37, dWd, dWdd, dW)
    ref = W - 0.37 * (Wd.conj().T + Wd) + 0.37 ** 2 * Wdd
    assert np.max(np.abs(dW.download() - ref)) < tol * 20


def _coeff_np_z(HW, SW, HWd, SWd, HWdd, SWdd):
    mt = lambda A, B: np.real(np.sum(A.conj() * B))
    TrH, TrHS = np.real(np.trace(HW)), mt(HW, SW)
    TrHd, TrHdS, TrHSd = np.real(np.trace(HWd)), mt(SW, HWd), mt(HW, SWd)
    TrHdd, TrHddS, TrHSdd = np.real(np.trace(HWdd)), mt(HWdd, SW), mt(HW, SWdd)
    TrHdSd, TrHdSdH = mt(SWd.conj().T, HWd), mt(SWd, HWd)
    TrHddSd, TrHdSdd, TrHddSdd = mt(HWdd, SWd), mt(SWdd, HWd), mt(HWdd, SWdd)
    return np.array([2 * TrH - TrHS, 2 * (2 * TrHd - TrHdS - TrHSd),
                     2 * (TrHdd - TrHdSd - TrHdSdH) - TrHddS - TrHSdd, -2 * (TrHddSd + TrHdSdd), -TrHddSdd])


@pytest.mark.parametrize("n", [31, 64])
def test_zcoeff(ctx, n):
    _note("zcoeff", n)
    mats = [crandn(n, n) for _ in range(6)]
    d = [oh.DMat.from_host(ctx, M, oh.OMX_C128) for M in mats]
    got = oh.omm_coeff(ctx, *d)
    assert np.allclose(got, _coeff_np_z(*mats), rtol=1e-12, atol=1e-10 * n)


@pytest.fixture(scope="module")
def fsetup():
    F.setup(0)
    yield


GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _same_trajectory(r, o, steps=20):
    assert abs(r["e_min"] - o["e_min"]) <= 1e-10 * abs(o["e_min"]), (r["e_min"], o["e_min"])
    assert abs(r["icg"] - o["icg"]) <= 2, (r["icg"], o["icg"])
    k = min(len(r["e_log"]), len(o["e_log"]), steps)
    assert k > 0 and np.allclose(r["e_log"][:k], o["e_log"][:k], rtol=1e-9, atol=0)


def test_example_kpoints_device_matches_oracle(fsetup):
    """example_kpoints.F90:154-305 (flavour 0) on 'gzden': two k-points with their own state
    (np = 2), the LCG start, the perturbed second SCF step from the previous C, then the
    energy-weighted density matrices; every call against the oracle's complex table."""
    _note("example_kpoints device vs oracle")
    m, n, ks = O.read_example_kpoints(os.path.join(GOLD, "example_gamma.dat"), os.path.join(GOLD, "example_kpoints.dat"))
    H2s = []
    for H, _ in ks:
        H2 = H.copy()
        H2[np.diag_indices(m)] += -0.001
        H2s.append(H2)
    om = O.Omm(2)
    Co = [np.zeros((n, m), np.complex128, order="F") for _ in ks]
    Hm = [F.Mat(m, m, "gzden", data=H) for H, _ in ks]
    Sm = [F.Mat(m, m, "gzden", data=S) for _, S in ks]
    Cm = [F.Mat(n, m, "gzden") for _ in ks]
    Dm = [F.Mat(m, m, "gzden") for _ in ks]
    for step in range(2):
        for k, (H, S) in enumerate(ks):
            Hh = H if step == 0 else H2s[k]
            if step == 1:
                Hm[k].upload(Hh)
            o = om(m, n, Hh, S, Co[k], new_S=(step == 0), init_C=(step == 1), ip=k + 1)
            r = F.omm(m, n, Hm[k], Sm[k], Cm[k], new_S=(step == 0), D_min=Dm[k], init_C=(step == 1), np_=2,
                      ip=k + 1, storage="gzden")
            _note("step", step, "k", k, "device", r["e_min"], r["icg"], "oracle", o["e_min"], o["icg"])
            _same_trajectory(r, o)
            assert np.max(np.abs(Dm[k].download() - o["D"])) < 1e-8
            assert np.max(np.abs(Cm[k].download() - Co[k])) < 1e-7
    for k, (H, S) in enumerate(ks):
        o = om(m, n, H2s[k], S, Co[k], new_S=False, init_C=True, calc_ED=True, ip=k + 1, dealloc=(k == 1))
        F.omm(m, n, Hm[k], Sm[k], Cm[k], new_S=False, D_min=Dm[k], calc_ED=True, init_C=True, np_=2, ip=k + 1,
              dealloc=(k == 1), storage="gzden")
        assert np.max(np.abs(Dm[k].download() - o["D"])) < 1e-8
    for M in Hm + Sm + Cm + Dm:
        M.free()


@pytest.mark.parametrize("L,phi,eta", [(3, (0.25, 0.5, 0.75), 0.2), (4, (0.125, 0.0, 0.375), 0.0)])
def test_complex_lattice_device_matches_oracle(fsetup, L, phi, eta):
    """Bloch operators H(k), S(k) of the L^3 cell on the device's complex BSR path ('gzcsc') from a
    complex start, against the oracle's complex trajectory on the same dense operators."""
    nb, no = 4 * L ** 3, 2 * L ** 3
    _note("complex lattice device vs oracle", L, phi, eta)
    Hb, Sb = oh.lattice_bsr_k(L, 0, phi), oh.lattice_bsr_k(L, 1, phi)
    Hs = (Hb[0], Hb[1], Hb[2] - eta * Sb[2])
    r0 = np.random.default_rng(L)
    C0 = np.asfortranarray((r0.uniform(-0.5, 0.5, (no, nb)) + 1j * r0.uniform(-0.5, 0.5, (no, nb))) / np.sqrt(nb))
    Co = C0.copy(order="F")
    o = O.Omm()(nb, no, oh.bsr_to_dense(nb, *Hs), oh.bsr_to_dense(nb, *Sb), Co, init_C=True, eta=eta, cg_tol=1e-9,
                max_iter=300)
    Hm, Sm = F.Mat.from_bsr_z(nb, *Hs), F.Mat.from_bsr_z(nb, *Sb)
    Cm, Dm = F.Mat(no, nb, "gzden", data=C0), F.Mat(nb, nb, "gzden")
    F.set_option("max_iter", 300)
    try:
        r = F.omm(nb, no, Hm, Sm, Cm, D_min=Dm, init_C=True, eta=eta, cg_tol=1e-9, dealloc=True, storage="gzden")
    finally:
        F.set_option("max_iter", 0)
    _note("device", r["e_min"], r["icg"], "oracle", o["e_min"], o["icg"])
    _same_trajectory(r, o, 30)
    assert np.max(np.abs(Dm.download() - o["D"])) < 1e-7
    assert abs(r["TrQS"] - o["TrQS"]) < 1e-7
    for M in (Hm, Sm, Cm, Dm):
        M.free()


def test_complex_no_S_device_matches_oracle(fsetup):
    L, no = 3, 40
    nb = 4 * L ** 3
    _note("complex no_S device vs oracle")
    H = oh.bsr_to_dense(nb, *oh.lattice_bsr_k(L, 0, (0.25, 0.5, 0.0))) - 0.5 * np.eye(nb)
    r0 = np.random.default_rng(5)
    C0 = np.asfortranarray((r0.uniform(-0.5, 0.5, (no, nb)) + 1j * r0.uniform(-0.5, 0.5, (no, nb))) / np.sqrt(nb))
    Co = C0.copy(order="F")
    o = O.Omm()(nb, no, H, None, Co, init_C=True)
    Hm, Cm = F.Mat.from_dense_z(H), F.Mat(no, nb, "gzden", data=C0)
    r = F.omm(nb, no, Hm, None, Cm, init_C=True, dealloc=True, storage="gzden")
    _note("device", r["e_min"], r["icg"], "oracle", o["e_min"], o["icg"])
    _same_trajectory(r, o)
    assert np.max(np.abs(Cm.download() - Co)) < 1e-7
    Hm.free()
    Cm.free()


def _exact(Hd, Sd, no, eta):
    w = sla.eigh(Hd, Sd, eigvals_only=True)
    return float(np.sum(w[:no][w[:no] < eta]))


@pytest.mark.parametrize("phi", [(0.25, 0.5, 0.75), (0.125, 0.0, 0.375)])
def test_complex_omm_matches_closed_form(fsetup, phi):
    L = 3
    nb, no = 4 * L ** 3, 2 * L ** 3
    _note("complex omm", phi)
    Hb, Sb = oh.lattice_bsr_k(L, 0, phi), oh.lattice_bsr_k(L, 1, phi)
    Hd, Sd = oh.bsr_to_dense(nb, *Hb), oh.bsr_to_dense(nb, *Sb)
    w = sla.eigh(Hd, Sd, eigvals_only=True)
    eta = 0.5 * (w[no - 1] + w[no])
    exact = float(np.sum(w[:no]))
    Hm = F.Mat.from_bsr_z(nb, Hb[0], Hb[1], Hb[2] - eta * Sb[2])
    Sm = F.Mat.from_bsr_z(nb, *Sb)
    Cm = F.Mat(no, nb, "gzden", data=oh.random_c0(no, nb).astype(np.complex128))
    r = F.omm(nb, no, Hm, Sm, Cm, init_C=True, eta=eta, cg_tol=1e-12, dealloc=True, storage="gzden")
    print("CASE complex omm icg", r["icg"], "e", r["e_min"], "exact", exact, flush=True)
    assert r["converged"]
    assert abs(r["e_min"] - exact) < 1e-8 * abs(exact)
    for M in (Hm, Sm, Cm):
        M.free()


def test_complex_omm_at_k0_equals_real(fsetup):
    L = 3
    nb, no = 4 * L ** 3, 2 * L ** 3
    _note("complex omm k=0")
    eta = 0.2
    Hb, Sb = oh.lattice_bsr(L, 0), oh.lattice_bsr(L, 1)
    Hs = (Hb[0], Hb[1], Hb[2] - eta * Sb[2])
    C0 = oh.random_c0(no, nb)
    Hr, Sr = F.Mat.from_bsr(nb, *Hs), F.Mat.from_bsr(nb, *Sb)
    Cr = F.Mat(no, nb, data=C0)
    rr = F.omm(nb, no, Hr, Sr, Cr, init_C=True, eta=eta, cg_tol=1e-10, dealloc=True)
    Hz = F.Mat.from_bsr_z(nb, Hs[0], Hs[1], Hs[2].astype(np.complex128))
    Sz = F.Mat.from_bsr_z(nb, Sb[0], Sb[1], Sb[2].astype(np.complex128))
    Cz = F.Mat(no, nb, "gzden", data=C0.astype(np.complex128))
    rz = F.omm(nb, no, Hz, Sz, Cz, init_C=True, eta=eta, cg_tol=1e-10, dealloc=True, storage="gzden")
    # same arithmetic on the real parts (imaginary parts stay exactly zero): same trajectory
    assert abs(rz["e_min"] - rr["e_min"]) < 1e-10 * abs(rr["e_min"])
    assert abs(rz["icg"] - rr["icg"]) <= 2
    Cd = Cz.download()
    assert np.max(np.abs(Cd.imag)) == 0.0
    for M in (Hr, Sr, Cr, Hz, Sz, Cz):
        M.free()


def test_complex64_omm(fsetup):
    L = 3
    nb, no = 4 * L ** 3, 2 * L ** 3
    phi = (0.25, 0.5, 0.75)
    _note("complex64 omm")
    Hb, Sb = oh.lattice_bsr_k(L, 0, phi), oh.lattice_bsr_k(L, 1, phi)
    Hd, Sd = oh.bsr_to_dense(nb, *Hb), oh.bsr_to_dense(nb, *Sb)
    w = sla.eigh(Hd, Sd, eigvals_only=True)
    eta = 0.5 * (w[no - 1] + w[no])
    exact = float(np.sum(w[:no]))
    Hm = F.Mat.from_bsr_z(nb, Hb[0], Hb[1], Hb[2] - eta * Sb[2])
    Sm = F.Mat.from_bsr_z(nb, *Sb)
    Cm = F.Mat(no, nb, "gcden", data=oh.random_c0(no, nb).astype(np.complex128))
    r = F.omm(nb, no, Hm, Sm, Cm, init_C=True, eta=eta, cg_tol=1e-7, dealloc=True, storage="gcden")
    print("CASE complex64 omm icg", r["icg"], "e", r["e_min"], "exact", exact, flush=True)
    assert abs(r["e_min"] - exact) < 1e-5 * abs(exact)
    for M in (Hm, Sm, Cm):
        M.free()
