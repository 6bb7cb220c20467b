"""The Fortran omm() drop-in on device storage ('gdden' + 'gdcsc', HIP kernels) against the CPU
oracle (oracle/omm_oracle.c, itself bit-identical to the reference serial build).

Tolerances (FP64): converged energies agree to 1e-10 relative, per-step
energies of the trajectory to 1e-9 relative, density matrices to 1e-8 absolute, CG step counts
within 2 (the e_diff test can flip at the threshold when the last bits differ).
"""
import os

import numpy as np
import pytest

import ommhip as oh
import oracle as O
import pyomm as F

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _note(*a):
    print("CASE", *a, flush=True)


@pytest.fixture(scope="module", autouse=True)
def _setup():
    F.setup(0)
    yield


def csc_of(A):
    nb = A.shape[0]
    cols = [np.nonzero(A[:, j])[0] for j in range(nb)]
    return (np.cumsum([0] + [len(c) for c in cols]).astype(np.int32), np.concatenate(cols).astype(np.int32),
            np.concatenate([A[c, j] for j, c in enumerate(cols)]))


def lattice(L, seed=oh.DEFAULT_SEED):
    nb = 4 * L ** 3
    Hb, Sb = oh.lattice_bsr(L, 0, seed), oh.lattice_bsr(L, 1, seed)
    return nb, Hb, Sb


def test_example_gamma_device_matches_oracle():
    _note("example_gamma device")
    m, n, H, S = O.read_example_gamma(os.path.join(GOLD, "example_gamma.dat"))
    H2 = H.copy()
    H2[np.diag_indices(m)] += -0.001
    om = O.Omm()
    Cc = np.zeros((n, m), order="F")
    o1 = om(m, n, H, S, Cc, new_S=True, init_C=False)
    o2 = om(m, n, H2, S, Cc, new_S=False, init_C=True)
    o3 = om(m, n, H2, S, Cc, new_S=False, init_C=True, calc_ED=True, dealloc=True)

    Hm, Sm = F.Mat.from_csc(m, *csc_of(H)), F.Mat.from_csc(m, *csc_of(S))
    Cm, Dm, EDm = F.Mat(n, m), F.Mat(m, m), F.Mat(m, m)
    r1 = F.omm(m, n, Hm, Sm, Cm, new_S=True, D_min=Dm, init_C=False)
    D1 = Dm.download()
    Hm.free()
    Hm = F.Mat.from_csc(m, *csc_of(H2))
    r2 = F.omm(m, n, Hm, Sm, Cm, new_S=False, D_min=Dm, init_C=True)
    D2 = Dm.download()
    F.omm(m, n, Hm, Sm, Cm, new_S=False, D_min=EDm, calc_ED=True, init_C=True, dealloc=True)
    for r, o in ((r1, o1), (r2, o2)):
        assert abs(r["e_min"] - o["e_min"]) <= 1e-10 * abs(o["e_min"])
        assert abs(r["icg"] - o["icg"]) <= 2
        k = min(len(r["e_log"]), len(o["e_log"]), 20)
        assert np.allclose(r["e_log"][:k], o["e_log"][:k], rtol=1e-9, atol=0)
    assert np.max(np.abs(D1 - o1["D"])) < 1e-8 and np.max(np.abs(D2 - o2["D"])) < 1e-8
    assert np.max(np.abs(EDm.download() - o3["D"])) < 1e-8
    for M in (Hm, Sm, Cm, Dm, EDm):
        M.free()


@pytest.mark.parametrize("L,eta", [(3, 0.0), (4, 0.2)])
def test_lattice_device_matches_oracle(L, eta):
    _note("lattice device", L, eta)
    nb, Hb, Sb = lattice(L)
    no = nb // 2
    H = oh.bsr_to_dense(nb, *Hb)
    S = oh.bsr_to_dense(nb, *Sb)
    Hs = H - eta * S
    C0 = oh.random_c0(no, nb)
    Cc = C0.copy(order="F")
    o = O.Omm()(nb, no, csc_of(Hs), csc_of(S), Cc, init_C=True, eta=eta, cg_tol=1e-9, max_iter=300)
    Hm, Sm = F.Mat.from_csc(nb, *csc_of(Hs)), F.Mat.from_csc(nb, *csc_of(S))
    Cm, Dm = F.Mat(no, nb, data=C0), F.Mat(nb, nb)
    F.set_option("max_iter", 300)
    try:
        r = F.omm(nb, no, Hm, Sm, Cm, D_min=Dm, init_C=True, eta=eta, cg_tol=1e-9, dealloc=True)
    finally:
        F.set_option("max_iter", 0)
    assert abs(r["e_min"] - o["e_min"]) <= 1e-10 * abs(o["e_min"])
    assert abs(r["icg"] - o["icg"]) <= 2
    k = min(len(r["e_log"]), len(o["e_log"]), 30)
    assert np.allclose(r["e_log"][:k], o["e_log"][:k], rtol=1e-9, atol=0)
    assert np.max(np.abs(Dm.download() - o["D"])) < 1e-7
    assert abs(r["TrQS"] - o["TrQS"]) < 1e-7
    for M in (Hm, Sm, Cm, Dm):
        M.free()


def test_no_S_device_matches_oracle():
    _note("no_S device")
    nb, Hb, _ = lattice(3)
    no = 40
    H = oh.bsr_to_dense(nb, *Hb) - 0.5 * np.eye(nb)
    C0 = oh.random_c0(no, nb)
    Cc = C0.copy(order="F")
    o = O.Omm()(nb, no, csc_of(H), None, Cc, init_C=True)
    Hm, Cm = F.Mat.from_csc(nb, *csc_of(H)), F.Mat(no, nb, data=C0)
    r = F.omm(nb, no, Hm, None, Cm, init_C=True, dealloc=True)
    assert abs(r["e_min"] - o["e_min"]) <= 1e-10 * abs(o["e_min"])
    assert np.max(np.abs(Cm.download() - Cc)) < 1e-7
    Hm.free()
    Cm.free()


def _exact_density(H, S, no):
    """Closed form: the S-orthonormal projector onto the no lowest generalised eigenvectors."""
    import scipy.linalg as sla

    w, X = sla.eigh(H, S)
    return w, X[:, :no] @ X[:, :no].T


@pytest.mark.parametrize("lowp", ["hbf", "h32"])
def test_mixed_precision_reaches_fp64_energy(lowp):
    """'hmx': a low-precision phase (BF16 or FP32 line-search Grams, FP32 gradient and state, the
    state recomputed every 5 steps) until the predicted decrease is below 0.3*cg_tol, then FP64
    refinement to e_diff <= cg_tol (once the gradient has dropped tenfold).  On the gapped 4x4x4
    lattice (onsite levels -6, -6, 6, 6, eta mid-gap) the mixed energy agrees with the FP64 solve's
    and with the exact minimum (sum of the occupied generalised eigenvalues) to 1e-10 relative, and
    its density matrix is at least as close to the closed form as the FP64 solve's."""
    _note("mixed", lowp)
    L = 4
    nb, no = 4 * L ** 3, 2 * L ** 3
    cp, ri, h = oh.lattice_bsr_theta(L, oh.lattice_theta0())
    Sb = oh.lattice_bsr(L, 1)
    H, S = oh.bsr_to_dense(nb, cp, ri, h), oh.bsr_to_dense(nb, *Sb)
    w, Dx = _exact_density(H, S, no)
    eta = 0.5 * (w[no - 1] + w[no])
    exact = float(np.sum(w[:no]))
    Hs = H - eta * S
    C0 = oh.random_c0(no, nb)
    Hm, Sm = F.Mat.from_csc(nb, *csc_of(Hs)), F.Mat.from_csc(nb, *csc_of(S))
    Cm, D64, Dmx = F.Mat(no, nb, data=C0), F.Mat(nb, nb), F.Mat(nb, nb)
    ref = F.omm(nb, no, Hm, Sm, Cm, D_min=D64, init_C=True, eta=eta, cg_tol=1e-10, dealloc=True)
    Cm.upload(C0)
    F.set_option("lowp_bf16", 1 if lowp == "hbf" else 0)
    F.set_option("lowp_grad_bf16x3", 0)  # the FP32-gradient form
    try:
        r = F.omm(nb, no, Hm, Sm, Cm, D_min=Dmx, init_C=True, eta=eta, cg_tol=1e-10, dealloc=True, operation="hmx")
    finally:
        F.set_option("lowp_bf16", 1)
        F.set_option("lowp_grad_bf16x3", 1)
    d64, dmx = D64.download(), Dmx.download()
    print("CASE mixed", lowp, "icg", r["icg"], "lowp", r["icg_lowp"], "fp64 icg", ref["icg"], "err",
          abs(r["e_min"] - exact) / abs(exact), "fp64 err", abs(ref["e_min"] - exact) / abs(exact), "D err",
          np.max(np.abs(dmx - Dx)), "fp64 D err", np.max(np.abs(d64 - Dx)), "grad", r["refine_g0"], r["refine_gend"],
          flush=True)
    assert r["converged"] and r["icg_lowp"] > 0
    assert abs(r["e_min"] - ref["e_min"]) <= 1e-10 * abs(ref["e_min"])
    assert abs(r["e_min"] - exact) <= 1e-10 * abs(exact)
    # the FP64 solve's own density is ~1e-6 from the closed form at this cg_tol; the mixed one is
    # at least as close
    assert np.max(np.abs(dmx - Dx)) <= max(np.max(np.abs(d64 - Dx)), 1e-7)
    assert np.max(np.abs(dmx - d64)) < 1e-5
    for M in (Hm, Sm, Cm, D64, Dmx):
        M.free()


@pytest.mark.parametrize("state", ["fp32", "bf16x6"])
def test_mixed_precision_bf16x3_gradient(state):
    """'hmx' with the BF16x3 gradient (hi*hi + hi*lo + lo*hi of the FP32 operands on BF16 MFMA) at
    the 8x8x8 lattice, where the Gram and gradient shapes run on the k-contiguous BF16 panels.  The
    gapped parameterised lattice (onsite levels -6, -6, 6, 6) is used, so that both solves can
    reach the exact minimum: the same bars as the FP32-gradient form.  With state = "bf16x6" the
    periodic state recomputation runs as BF16x6 panel products (FP32-grade) instead of FP32."""
    _note("mixed bf16x3", state)
    L = 8
    nb, no = 4 * L ** 3, 2 * L ** 3
    cp, ri, h = oh.lattice_bsr_theta(L, oh.lattice_theta0())
    Sb = oh.lattice_bsr(L, 1)
    H, S = oh.bsr_to_dense(nb, cp, ri, h), oh.bsr_to_dense(nb, *Sb)
    w = np.linalg.eigvalsh(np.linalg.solve(np.linalg.cholesky(S), np.linalg.solve(np.linalg.cholesky(S), H).T))
    eta = 0.5 * (w[no - 1] + w[no])
    exact = float(np.sum(w[:no]))
    Hs = H - eta * S
    C0 = oh.random_c0(no, nb)
    Hm, Sm = F.Mat.from_csc(nb, *csc_of(Hs)), F.Mat.from_csc(nb, *csc_of(S))
    Cm = F.Mat(no, nb, data=C0)
    ref = F.omm(nb, no, Hm, Sm, Cm, init_C=True, eta=eta, cg_tol=1e-10, dealloc=True)
    Cm.upload(C0)
    F.set_option("lowp_grad_bf16x3", 1)
    F.set_option("lowp_max_iter", 3000)
    F.set_option("lowp_state_bf16x6", 1 if state == "bf16x6" else 0)
    try:
        r = F.omm(nb, no, Hm, Sm, Cm, init_C=True, eta=eta, cg_tol=1e-10, dealloc=True, operation="hmx")
    finally:
        F.set_option("lowp_grad_bf16x3", 1)
        F.set_option("lowp_max_iter", 0)
        F.set_option("lowp_state_bf16x6", 0)
    print("CASE mixed bf16x3", state, "icg", r["icg"], "lowp", r["icg_lowp"], "fp64 icg", ref["icg"], "err",
          abs(r["e_min"] - exact) / abs(exact), "fp64 err", abs(ref["e_min"] - exact) / abs(exact), flush=True)
    assert r["converged"] and r["icg_lowp"] > 0
    assert abs(r["e_min"] - ref["e_min"]) <= 1e-10 * abs(ref["e_min"])
    assert abs(r["e_min"] - exact) <= 1e-10 * abs(exact)
    for M in (Hm, Sm, Cm):
        M.free()


def test_pdcsc_registered_operators_match_oracle():
    """H and S registered through m_register_pcsc in pspBLAS's convention (1-based row indices and
    column pointers, MatrixSwitch_m_register.F90:394-443, pspMat.F90:87-140), then omm() on device
    storage against the oracle on the same CSC: the caller of the 1-rank 'pdcsc' path, drop-in."""
    _note("pdcsc registration")
    L, eta = 3, 0.2
    nb, Hb, Sb = lattice(L)
    no = nb // 2
    H = oh.bsr_to_dense(nb, *Hb)
    S = oh.bsr_to_dense(nb, *Sb)
    Hs = H - eta * S
    C0 = oh.random_c0(no, nb)
    Cc = C0.copy(order="F")
    o = O.Omm()(nb, no, csc_of(Hs), csc_of(S), Cc, init_C=True, eta=eta, cg_tol=1e-9, max_iter=300)
    hp, sp = csc_of(Hs), csc_of(S)
    Hm = F.Mat.register_pcsc(nb, hp[1] + 1, hp[0] + 1, hp[2])
    Sm = F.Mat.register_pcsc(nb, sp[1] + 1, sp[0] + 1, sp[2])
    for i, j in ((1, 1), (1, 5), (7, 2), (nb, nb)):
        assert Hm.get_element(i, j) == Hs[i - 1, j - 1]
    Cm = F.Mat(no, nb, data=C0)
    F.set_option("max_iter", 300)
    try:
        r = F.omm(nb, no, Hm, Sm, Cm, init_C=True, eta=eta, cg_tol=1e-9, dealloc=True)
    finally:
        F.set_option("max_iter", 0)
    assert abs(r["e_min"] - o["e_min"]) <= 1e-10 * abs(o["e_min"])
    assert abs(r["icg"] - o["icg"]) <= 2
    for M in (Hm, Sm, Cm):
        M.free()


def test_m_convert_device_storages():
    """m_convert (MatrixSwitch.F90:987-1041) between host and device storage: 'sdden' -> 'gdden'
    (values unchanged), 'gdden' -> 'gsden' (FP32), and 'gdden' -> 'gdcsc' with a threshold used as
    the SpMM operator of mm_multiply against the oracle's CSC product."""
    _note("m_convert")
    nb, no = 24, 8
    r = np.random.default_rng(4)
    A = r.standard_normal((nb, nb)) * (r.random((nb, nb)) < 0.4)
    M = F.Mat(nb, nb, "sdden", data=A)
    M.convert("gdden")
    assert np.array_equal(M.download(), A)
    M.convert("gsden")
    assert np.array_equal(M.download(), A.astype(np.float32).astype(np.float64))
    M.free()
    M = F.Mat(nb, nb, "sdden", data=A)
    M.convert("gdden")
    M.convert("gdcsc", threshold=0.5)
    At = np.where(np.abs(A) > 0.5, A, 0.0)
    for i in range(1, nb + 1, 5):
        for j in range(1, nb + 1, 3):
            assert M.get_element(i, j) == At[i - 1, j - 1]
    M.free()


@pytest.mark.parametrize("flavour", [1, 2, 3])
def test_flavours_reach_the_closed_form(flavour):
    """The Cholesky flavours 1 (S factored by omm()) and 2 (S given as its upper Cholesky factor U,
    omm.F90:133-136, 305-309) and the preconditioned flavour 3 (P = S^-1, omm.F90:379-403) on dense
    device H and S: the gapped 4x4x4 lattice (onsite levels -6, -6, 6, 6, eta mid-gap) solved to
    cg_tol 1e-11; the energy agrees with the sum of the occupied generalised eigenvalues to 1e-10
    relative and D_min with the S-orthonormal projector to 1e-6, as the flavour-0 solve does."""
    import scipy.linalg as sla

    _note("flavour", flavour)
    L = 4
    nb, no = 4 * L ** 3, 2 * L ** 3
    cp, ri, h = oh.lattice_bsr_theta(L, oh.lattice_theta0())
    Sb = oh.lattice_bsr(L, 1)
    H, S = oh.bsr_to_dense(nb, cp, ri, h), oh.bsr_to_dense(nb, *Sb)
    w, Dx = _exact_density(H, S, no)
    eta = 0.5 * (w[no - 1] + w[no])
    exact = float(np.sum(w[:no]))
    Sgiven = sla.cholesky(S, lower=False) if flavour == 2 else S
    Hm, Sm = F.Mat(nb, nb, data=H - eta * S), F.Mat(nb, nb, data=Sgiven)
    C0 = oh.random_c0(no, nb)
    Cm, Dm = F.Mat(no, nb, data=C0), F.Mat(nb, nb)
    r = F.omm(nb, no, Hm, Sm, Cm, D_min=Dm, init_C=True, eta=eta, cg_tol=1e-11, flavour=flavour, dealloc=True)
    rel = abs(r["e_min"] - exact) / abs(exact)
    derr = np.max(np.abs(Dm.download() - Dx))
    _note("flavour", flavour, "icg", r["icg"], "energy err", rel, "D err", derr)
    assert r["converged"] and rel <= 1e-10 and derr < 1e-6
    if flavour == 1:
        # S now holds U in its upper triangle; its strict lower triangle is the caller's S
        Sout = Sm.download()
        assert np.max(np.abs(np.triu(Sout) - sla.cholesky(S, lower=False))) < 1e-12
        assert np.array_equal(np.tril(Sout, -1), np.tril(S, -1))
    for M in (Hm, Sm, Cm, Dm):
        M.free()
