"""Pins of the CPU oracle (oracle/omm_oracle.c) -- CPU only.

The oracle is the checker of every GPU parity test, so each of its functions is pinned here by
something other than a restatement of its own formula:
  * the reference serial build (oracle/_ref: libOMM + MatrixSwitch compiled from the reference
    sources) on the same inputs -- bitwise for omm() trajectories, D_min, C_min and icg;
  * the reference's published known answers (example_gamma, example_sdden);
  * closed forms: generalised eigenproblem (converged energy, density matrix, TrQS), brute-force
    quartic minimisation, the polynomial identity E(C + xD) = sum c_i x^i, central finite
    differences of E for the gradient, and the LCG integer recurrence.
"""
import json
import os

import numpy as np
import pytest
import scipy.linalg as sla

import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
KA = json.load(open(os.path.join(GOLD, "known_answers.json")))
rng = np.random.default_rng(11)


def energy(H, S, C):
    """Kim functional E = tr[(2I - SW) HW] with HW = C H C^T, SW = C S C^T (omm_ops.F90:213)."""
    HW, SW = C @ H @ C.T, C @ S @ C.T
    return np.trace((2 * np.eye(C.shape[0]) - SW) @ HW)


def small_lattice(L=2, seed=5):
    """A dense symmetric block-sparse H and SPD S with the lattice structure (numpy only)."""
    import ommhip as oh  # generator lives in the product library (host code, no GPU needed)

    nb = 4 * L ** 3
    H = oh.bsr_to_dense(nb, *oh.lattice_bsr(L, 0, seed))
    S = oh.bsr_to_dense(nb, *oh.lattice_bsr(L, 1, seed))
    return nb, H, S


# ---------------------------------------------------------------------------------------------
# reference build (bitwise)
# ---------------------------------------------------------------------------------------------

needs_ref = pytest.mark.skipif(not O.have_ref(), reason="reference build oracle/_ref not present (make -C oracle ref)")


@needs_ref
def test_example_gamma_matches_reference_build_bitwise():
    m, n, H, S = O.read_example_gamma(os.path.join(GOLD, "example_gamma.dat"))
    H2 = H.copy()
    H2[np.diag_indices(m)] += -0.001  # example_gamma.F90:175-182
    calls = [dict(H=H, S=S, new_S=True, init_C=False), dict(H=H2, S=S, new_S=False, init_C=True),
             dict(H=H2, S=S, new_S=False, init_C=True, calc_ED=True, dealloc=True)]
    ref, log = O.run_ref(m, n, calls)
    om = O.Omm()
    Cm = np.zeros((n, m), order="F")
    r1 = om(m, n, H, S, Cm, new_S=True, init_C=False)
    r2 = om(m, n, H2, S, Cm, new_S=False, init_C=True)
    r3 = om(m, n, H2, S, Cm, new_S=False, init_C=True, calc_ED=True, dealloc=True)
    assert r1["e_min"] == ref[0]["e_min"] and r2["e_min"] == ref[1]["e_min"]
    assert np.array_equal(r1["D"], ref[0]["D"]) and np.array_equal(r2["D"], ref[1]["D"])
    assert np.array_equal(r3["D"], ref[2]["D"])  # energy-weighted density matrix
    assert np.array_equal(Cm, ref[2]["C"])
    assert O.icg_from_log(log) == [r1["icg"], r2["icg"]]


@needs_ref
@pytest.mark.parametrize("sparse", [False, True])
@pytest.mark.parametrize("eta", [0.0, 0.3])
def test_lattice_matches_reference_build(sparse, eta):
    nb, H, S = small_lattice(2)
    no = nb // 2
    C0 = np.asfortranarray(rng.uniform(-0.5, 0.5, (no, nb)) / np.sqrt(nb))
    Hs = H - eta * S
    calls = [dict(H=Hs, S=S, C=C0, new_S=True, init_C=True, eta=eta, cg_tol=1e-9)]
    ref, log = O.run_ref(nb, no, calls, sparse=sparse)
    Cm = C0.copy(order="F")
    if sparse:
        def csc(A):
            cols = [np.nonzero(A[:, j])[0] for j in range(nb)]
            return (np.cumsum([0] + [len(c) for c in cols]), np.concatenate(cols),
                    np.concatenate([A[c, j] for j, c in enumerate(cols)]))
        Hop, Sop = csc(Hs), csc(S)
    else:
        Hop, Sop = Hs, S
    r = O.Omm()(nb, no, Hop, Sop, Cm, new_S=True, init_C=True, eta=eta, cg_tol=1e-9)
    assert r["e_min"] == ref[0]["e_min"]
    assert np.array_equal(r["D"], ref[0]["D"])
    assert np.array_equal(Cm, ref[0]["C"])
    assert O.icg_from_log(log) == [r["icg"]]


def _kpoint_calls(H, S):
    """The call sequence of example_kpoints.F90:176-305 for one k-point (flavour 0): LCG start,
    then the perturbed H (diagonal -0.001) from the previous C, then the energy-weighted density."""
    H2 = H.copy()
    H2[np.diag_indices(H.shape[0])] += -0.001
    return H2, [dict(H=H, S=S, new_S=True, init_C=False), dict(H=H2, S=S, new_S=False, init_C=True),
                dict(H=H2, S=S, new_S=False, init_C=True, calc_ED=True, dealloc=True)]


@needs_ref
@pytest.mark.parametrize("k", [0, 1])
def test_example_kpoints_matches_reference_build_bitwise(k):
    """Complex storage ('szden'): the reference build's mm_multiply_szdenref, m_add_szdenref and
    mm_ztrace (MatrixSwitch_mm_multiply.F90:85-185, MatrixSwitch_m_add.F90:135-175,
    MatrixSwitch.F90:2382-2491) against the oracle's complex table, on the reference's own k-point
    data.  k-point 2 is genuinely complex (Hermitian with nonzero imaginary parts)."""
    m, n, ks = O.read_example_kpoints(os.path.join(GOLD, "example_gamma.dat"), os.path.join(GOLD, "example_kpoints.dat"))
    H, S = ks[k]
    assert k == 0 or np.abs(H.imag).max() > 1e-3
    H2, calls = _kpoint_calls(H, S)
    ref, log = O.run_ref(m, n, calls)
    om = O.Omm()
    Cm = np.zeros((n, m), np.complex128, order="F")
    r1 = om(m, n, H, S, Cm, new_S=True, init_C=False)
    r2 = om(m, n, H2, S, Cm, new_S=False, init_C=True)
    r3 = om(m, n, H2, S, Cm, new_S=False, init_C=True, calc_ED=True, dealloc=True)
    assert r1["e_min"] == ref[0]["e_min"] and r2["e_min"] == ref[1]["e_min"]
    assert np.array_equal(r1["D"], ref[0]["D"]) and np.array_equal(r2["D"], ref[1]["D"])
    assert np.array_equal(r3["D"], ref[2]["D"])
    assert np.array_equal(Cm, ref[2]["C"])
    assert O.icg_from_log(log) == [r1["icg"], r2["icg"]]
    # the reference's check constants (example_kpoints.F90:66-70, 256-259)
    ka = KA["example_kpoints"]
    want = ka["e_min_check"]["k%d" % (k + 1)]
    assert abs(r1["e_min"] - want[0]) < ka["tolerance"] and abs(r2["e_min"] - want[1]) < ka["tolerance"]


def _bloch(L, phi, seed=5):
    import ommhip as oh

    nb = 4 * L ** 3
    H = oh.bsr_to_dense(nb, *oh.lattice_bsr_k(L, 0, phi, seed))
    S = oh.bsr_to_dense(nb, *oh.lattice_bsr_k(L, 1, phi, seed))
    return nb, H, S


@needs_ref
@pytest.mark.parametrize("s_present", [True, False])
def test_complex_lattice_matches_reference_build_bitwise(s_present):
    """A Bloch lattice operator (k = (1/4, 1/2, 0) on the 2x2x2 cell) from a complex start C, with
    the eta shift (with S) or the no-S branch: trajectory, C, D and icg bitwise."""
    nb, H, S = _bloch(2, (0.25, 0.5, 0.0))
    no = nb // 2 if s_present else 12
    eta = 0.3 if s_present else 0.0
    C0 = np.asfortranarray((rng.uniform(-0.5, 0.5, (no, nb)) + 1j * rng.uniform(-0.5, 0.5, (no, nb))) / np.sqrt(nb))
    Hs = H - eta * S if s_present else H - 0.5 * np.eye(nb)
    call = dict(H=Hs, C=C0, new_S=True, init_C=True, eta=eta, cg_tol=1e-9)
    if s_present:
        call["S"] = S
    ref, log = O.run_ref(nb, no, [call], s_present=s_present)
    Cm = C0.copy(order="F")
    r = O.Omm()(nb, no, Hs, S if s_present else None, Cm, new_S=True, init_C=True, eta=eta, cg_tol=1e-9)
    assert r["e_min"] == ref[0]["e_min"]
    assert np.array_equal(r["D"], ref[0]["D"]) and np.array_equal(Cm, ref[0]["C"])
    assert O.icg_from_log(log) == [r["icg"]] and r["icg"] > 5


@pytest.mark.skipif(not os.path.exists(O.REF_MM), reason="reference build oracle/_ref/ref_mm not present")
@pytest.mark.parametrize("opA,opB", [("n", "n"), ("t", "n"), ("n", "t"), ("t", "t")])
def test_sddmm_csc_matches_reference_build_bitwise(opA, opB):
    """or_sddmm_csc against mm_multiply_sddensddensdcscref (MatrixSwitch_mm_multiply.F90:334-398),
    dense x dense -> 'sdcsc', with beta applied to the stored values first."""
    r = np.random.default_rng(11)
    m, n, k = 37, 29, 23
    A = r.standard_normal((k, m) if opA == "t" else (m, k))
    B = r.standard_normal((n, k) if opB == "t" else (k, n))
    C0 = r.standard_normal((m, n)) * (r.random((m, n)) < 0.3)
    C0[:, 5] = 0.0  # an empty column
    cp, ri, ref_val = O.run_ref_sddmm(opA, opB, A, B, C0, 0.7, -1.3)
    cval0 = np.concatenate([C0[ri[cp[j]:cp[j + 1]], j] for j in range(n)])
    got = O.sddmm_csc(opA, opB, A, B, cp, ri, cval0, 0.7, -1.3)
    assert np.array_equal(got, ref_val)
    # the same numbers from dense algebra (summation order aside)
    P = 0.7 * ((A.T if opA == "t" else A) @ (B.T if opB == "t" else B)) - 1.3 * C0
    dense = np.concatenate([P[ri[cp[j]:cp[j + 1]], j] for j in range(n)])
    assert np.allclose(got, dense, rtol=1e-13, atol=1e-13)
    assert cp[6] == cp[5]


@needs_ref
def test_no_S_matches_reference_build():
    nb, H, _ = small_lattice(2, seed=9)
    no = 24
    C0 = np.asfortranarray(rng.uniform(-0.5, 0.5, (no, nb)) / np.sqrt(nb))
    ref, log = O.run_ref(nb, no, [dict(H=H - 0.5 * np.eye(nb), C=C0, new_S=True, init_C=True)], s_present=False)
    Cm = C0.copy(order="F")
    r = O.Omm()(nb, no, H - 0.5 * np.eye(nb), None, Cm, init_C=True)
    assert r["e_min"] == ref[0]["e_min"] and np.array_equal(Cm, ref[0]["C"])
    assert O.icg_from_log(log) == [r["icg"]]


# ---------------------------------------------------------------------------------------------
# published known answers
# ---------------------------------------------------------------------------------------------

def test_example_gamma_known_answers():
    ka = KA["example_gamma"]
    m, n, H, S = O.read_example_gamma(os.path.join(GOLD, "example_gamma.dat"))
    H2 = H.copy()
    H2[np.diag_indices(m)] += -0.001
    om = O.Omm()
    Cm = np.zeros((n, m), order="F")
    r1 = om(m, n, H, S, Cm, new_S=True, init_C=False)
    r2 = om(m, n, H2, S, Cm, new_S=False, init_C=True)
    r3 = om(m, n, H2, S, Cm, new_S=False, init_C=True, calc_ED=True, dealloc=True)
    tol = ka["tolerance"]
    assert abs(r1["e_min"] - ka["e_min_check"][0]) < tol and abs(r2["e_min"] - ka["e_min_check"][1]) < tol
    assert abs(r1["D"][0, 0] - ka["D11_check"][0]) < tol and abs(r2["D"][0, 0] - ka["D11_check"][1]) < tol
    assert abs(r3["D"][0, 0] - ka["ED11_check"]) < tol
    # the full-precision sample output agrees far below the check tolerance
    assert abs(r1["e_min"] - ka["sample_flavour0"]["e_min"][0]) < 1e-6


def test_example_sdden_known_answers():
    ka = KA["example_sdden"]
    N, M = 15, 8
    A, B = np.zeros((N, M), order="F"), np.zeros((N, M), order="F")
    rn = 0.1
    for i in range(N):  # example_sdden.F90:78-86
        for j in range(M):
            rn = np.fmod(4.2 * rn, 1.0)
            A[i, j] = rn
            rn = np.fmod(4.2 * rn, 1.0)
            B[i, j] = rn
    res1 = O.mm_trace(A, B)
    D = O.gemm("t", "n", A, B, np.zeros((M, M), order="F"))
    E = O.gemm("n", "t", B, A, np.zeros((N, N), order="F"))
    Cm = O.add("t", A, np.zeros((M, N), order="F"), 1.0, 0.0)
    D2 = O.gemm("n", "n", Cm, B, np.zeros((M, M), order="F"))
    E2 = O.gemm("n", "n", B, Cm, np.zeros((N, N), order="F"))
    tol = ka["tolerance"]
    for r in (res1, O.trace(D), O.trace(E), O.trace(D2), O.trace(E2)):
        assert abs(r - ka["res_check"]) < tol
    assert abs(E2[0, 0] - ka["el_check"]) < tol


def test_lcg_matches_golden():
    g = json.load(open(os.path.join(GOLD, "lcg_123456.json")))
    seq = O.bsd_lcg_seq(123456, len(g["states"]))
    assert [s for s, _ in seq] == g["states"]
    assert [r for _, r in seq] == g["values"]


# ---------------------------------------------------------------------------------------------
# closed forms
# ---------------------------------------------------------------------------------------------

def _quartic_brute(c, lo=-50, hi=50):
    xs = np.linspace(lo, hi, 2000001)
    p = np.polyval(c[::-1], xs)
    return xs[np.argmin(p)]


@pytest.mark.parametrize("seed", range(40))
def test_quartic_returns_a_stationary_minimum(seed):
    """The solver returns a local minimiser of the quartic (a real root of its derivative with
    positive curvature, located independently with numpy.roots).  With two wells and the
    separating maximum at t >= 0 it is the left-hand well (omm_quartic.F90:120-132)."""
    r = np.random.default_rng(seed)
    c = np.array([r.normal(), r.normal(), r.normal(), r.normal(), abs(r.normal()) + 0.1])
    x, fail = O.solve_quartic(c)
    assert not fail
    d1 = np.polyder(c[::-1])
    d2 = np.polyder(d1)
    roots = np.roots(d1)
    real = np.sort(roots[np.abs(roots.imag) < 1e-9].real)
    assert np.min(np.abs(real - x)) < 1e-7 * max(1.0, abs(x))
    assert np.polyval(d2, x) > 0
    if len(real) == 3 and real[1] >= 0:
        assert abs(x - real[0]) < 1e-7 * max(1.0, abs(x))
    if len(real) == 1:
        assert abs(x - _quartic_brute(c)) < 1e-3


def test_quartic_special_branches():
    # quadratic branch: |c4| <= 1e-11 -> x = -c1/(2 c2)  (omm_quartic.F90:105-109)
    x, fail = O.solve_quartic(np.array([1.0, -3.0, 2.0, 0.0, 0.0]))
    assert x == 0.75 and not fail
    # one real root of the cubic with c4 < 0 -> fail (omm_quartic.F90:142-150)
    x, fail = O.solve_quartic(np.array([0.0, 1.0, 1.0, 1.0, -1.0]))
    assert fail
    # three real roots, c4 > 0: the minimum among them (double well x^4 - 2x^2 + 0.1x)
    c = np.array([0.0, 0.1, -2.0, 0.0, 1.0])
    x, fail = O.solve_quartic(c)
    assert not fail and abs(x - _quartic_brute(c, -3, 3)) < 1e-5


def test_calc_coeff_is_the_line_polynomial():
    nb, H, S = small_lattice(2)
    no = 12
    Cm = rng.standard_normal((no, nb)) * 0.1
    D = rng.standard_normal((no, nb)) * 0.1
    HW, SW = Cm @ H @ Cm.T, Cm @ S @ Cm.T
    HWd, SWd = (Cm @ H) @ D.T, (Cm @ S) @ D.T
    HWdd, SWdd = (D @ H) @ D.T, (D @ S) @ D.T
    c = O.calc_coeff(HW, SW, HWd, SWd, HWdd, SWdd)
    for x in (-1.3, -0.4, 0.0, 0.25, 0.9, 2.0):
        assert abs(np.polyval(c[::-1], x) - energy(H, S, Cm + x * D)) < 1e-10 * (1 + abs(x) ** 4)


def test_gradient_is_dE_dC():
    """G = 4HC - 2 SW HC - 2 HW SC (omm_ops.F90:17-39) vs central differences of E."""
    nb, H, S = small_lattice(2)
    no = 6
    Cm = rng.standard_normal((no, nb)) * 0.2
    HC, SC = Cm @ H, Cm @ S
    HW, SW = HC @ Cm.T, SC @ Cm.T
    G = O.add("n", HC, np.zeros((no, nb), order="F"), 4.0, 0.0)
    O.gemm("n", "n", SW, HC, G, -2.0, 1.0)
    O.gemm("n", "n", HW, SC, G, -2.0, 1.0)
    h = 1e-6
    for (i, k) in [(0, 0), (3, 17), (5, nb - 1), (2, 9)]:
        E1, E2 = Cm.copy(), Cm.copy()
        E1[i, k] += h
        E2[i, k] -= h
        fd = (energy(H, S, E1) - energy(H, S, E2)) / (2 * h)
        assert abs(fd - G[i, k]) < 1e-7 * max(1, abs(G[i, k]))


@pytest.mark.parametrize("eta", [0.0, None])
def test_converged_energy_is_sum_of_occupied_eigenvalues(eta):
    """At the fixed point e_min = sum of the occupied generalised eigenvalues of (H - eta S, S)
    shifted back by eta*TrQS, D_min = the S-orthonormal projector, TrQS = occupied count."""
    nb, H, S = small_lattice(3, seed=21)
    no = nb // 2
    w, V = sla.eigh(H, S)
    if eta is None:
        eta = 0.5 * (w[no - 1] + w[no])  # mid-gap shift: exactly `no` states occupied
    ws = w - eta
    occ = np.nonzero(ws < 0)[0][:no]
    e_exact = np.sum(w[occ])
    Dex = V[:, occ] @ V[:, occ].T
    C0 = np.asfortranarray(rng.uniform(-0.5, 0.5, (no, nb)) / np.sqrt(nb))
    r = O.Omm()(nb, no, H - eta * S, S, C0, init_C=True, eta=eta, cg_tol=1e-14)
    assert abs(r["e_min"] - e_exact) < 1e-7 * abs(e_exact)
    assert abs(r["TrQS"] - len(occ)) < 1e-4
    assert np.max(np.abs(r["D"] - Dex)) < 1e-3


# ---------------------------------------------------------------------------------------------
# complex storage ('szden' / 'szcsc')
# ---------------------------------------------------------------------------------------------

def _crandn(*shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _zop(X, op):
    return X if op == "n" else (X.T if op == "t" else X.conj().T)


@pytest.mark.parametrize("opA,opB", [("n", "n"), ("n", "c"), ("c", "n"), ("t", "n"), ("n", "t"), ("c", "c")])
def test_complex_products_match_numpy(opA, opB):
    m, n, k = 13, 11, 17
    A = _crandn(*((k, m) if opA != "n" else (m, k)))
    B = _crandn(*((n, k) if opB != "n" else (k, n)))
    C0 = np.asfortranarray(_crandn(m, n))
    got = O.zgemm(opA, opB, A, B, C0.copy(order="F"), alpha=-1.5, beta=0.5)
    ref = -1.5 * _zop(A, opA) @ _zop(B, opB) + 0.5 * C0
    assert np.max(np.abs(got - ref)) < 1e-13 * np.sqrt(k) * 8


def test_complex_csc_product_equals_dense_product_bitwise():
    """Per output element both consume the contraction in ascending order, and a skipped zero adds
    only a signed zero, so the CSC kernel (MatrixSwitch_mm_multiply.F90:734-778) returns the dense
    kernel's numbers (MatrixSwitch_mm_multiply.F90:85-185) on the same operator."""
    nb, H, _ = _bloch(2, (0.25, 0.5, 0.125))
    cols = [np.nonzero(H[:, j])[0] for j in range(nb)]
    cp = np.cumsum([0] + [len(c) for c in cols])
    ri = np.concatenate(cols)
    val = np.concatenate([H[c, j] for j, c in enumerate(cols)])
    A = np.asfortranarray(_crandn(9, nb))
    C0 = np.asfortranarray(_crandn(9, nb))
    sp = O.zspmm_csc(A, cp, ri, val, C0.copy(order="F"), alpha=0.7, beta=-1.1)
    de = O.zgemm("n", "n", A, H, C0.copy(order="F"), alpha=0.7, beta=-1.1)
    assert np.array_equal(sp, de)
    assert np.max(np.abs(sp - (0.7 * A @ H - 1.1 * C0))) < 1e-12


def test_complex_add_and_traces_match_numpy():
    A, B = _crandn(7, 7), _crandn(7, 7)
    for op in ("n", "t", "c"):
        C0 = np.asfortranarray(_crandn(7, 7))
        got = O.zadd(op, A, C0.copy(order="F"), 0.3, -2.0)
        assert np.max(np.abs(got - (0.3 * _zop(A, op) - 2.0 * C0))) < 1e-14
    assert abs(O.ztrace(A) - np.trace(A)) < 1e-13
    assert abs(O.zmm_trace(A, B) - np.sum(A.conj() * B)) < 1e-12


def test_complex_calc_coeff_is_the_line_polynomial():
    """E(C + xD) = Re tr[(2I - SW) HW] with HW = C H C^H, SW = C S C^H is the quartic sum c_i x^i
    for Hermitian H, S and complex C, D."""
    nb, H, S = _bloch(2, (0.25, 0.5, 0.0))
    no = 10
    Cm, D = _crandn(no, nb) * 0.1, _crandn(no, nb) * 0.1

    def e(Cx):
        HW, SW = Cx @ H @ Cx.conj().T, Cx @ S @ Cx.conj().T
        return np.trace((2 * np.eye(no) - SW) @ HW).real

    HW, SW = Cm @ H @ Cm.conj().T, Cm @ S @ Cm.conj().T
    HWd, SWd = (Cm @ H) @ D.conj().T, (Cm @ S) @ D.conj().T
    HWdd, SWdd = (D @ H) @ D.conj().T, (D @ S) @ D.conj().T
    c = O.zcalc_coeff(HW, SW, HWd, SWd, HWdd, SWdd)
    for x in (-1.3, -0.4, 0.0, 0.25, 0.9, 2.0):
        assert abs(np.polyval(c[::-1], x) - e(Cm + x * D)) < 1e-10 * (1 + abs(x) ** 4)


def test_complex_oracle_on_real_data_equals_real_oracle_bitwise():
    """(a + 0i)(b + 0i) = ab + 0i exactly, so on real operators and a real start the complex table
    reproduces the real one to the last bit: same trajectory, C, D and icg, zero imaginary parts."""
    nb, H, S = small_lattice(2, seed=3)
    no, eta = nb // 2, 0.2
    C0 = np.asfortranarray(rng.uniform(-0.5, 0.5, (no, nb)) / np.sqrt(nb))
    Cr, Cz = C0.copy(order="F"), C0.astype(np.complex128, order="F")
    rr = O.Omm()(nb, no, H - eta * S, S, Cr, init_C=True, eta=eta, cg_tol=1e-10)
    rz = O.Omm()(nb, no, (H - eta * S).astype(np.complex128), S.astype(np.complex128), Cz, init_C=True, eta=eta,
                 cg_tol=1e-10)
    assert rr["e_min"] == rz["e_min"] and rr["icg"] == rz["icg"] and np.array_equal(rr["e_log"], rz["e_log"])
    assert np.array_equal(Cz.real, Cr) and not np.any(Cz.imag)
    assert np.array_equal(rz["D"].real, rr["D"]) and not np.any(rz["D"].imag)


def test_complex_converged_energy_is_sum_of_occupied_eigenvalues():
    """At a general k-point of the 3x3x3 cell with the mid-gap shift: e_min = sum of the occupied
    eigenvalues of the Hermitian generalised problem, TrQS = no, D_min = X X^H over the occupied
    S-orthonormal eigenvectors X."""
    nb, H, S = _bloch(3, (0.25, 0.5, 0.75), seed=21)
    no = nb // 2
    w, V = sla.eigh(H, S)
    eta = 0.5 * (w[no - 1] + w[no])
    C0 = np.asfortranarray(rng.uniform(-0.5, 0.5, (no, nb)) / np.sqrt(nb)).astype(np.complex128)
    r = O.Omm()(nb, no, H - eta * S, S, C0, init_C=True, eta=eta, cg_tol=1e-14)
    exact = np.sum(w[:no])
    assert abs(r["e_min"] - exact) < 1e-9 * abs(exact)
    assert abs(r["TrQS"] - no) < 1e-4
    Dex = V[:, :no] @ V[:, :no].conj().T
    assert np.max(np.abs(r["D"] - Dex)) < 1e-3
