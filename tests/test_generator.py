"""The synthetic lattice inputs (csrc/lattice.cpp) against an independent numpy restatement
(tests/golden/lattice_gen.py) and structural invariants.  CPU only."""
import os
import sys

import numpy as np
import pytest

import ommhip as oh

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
import lattice_gen as G  # noqa: E402


def test_u01_matches_numpy_restatement():
    r = np.random.default_rng(1)
    for _ in range(200):
        seed, stream, idx = int(r.integers(0, 2 ** 63)), int(r.integers(1, 6)), int(r.integers(0, 2 ** 40))
        assert oh.u01(seed, stream, idx) == G.u01(seed, stream, idx)


@pytest.mark.parametrize("dims", [(3, 3, 3), (2, 3, 4), (2, 2, 2), (4, 3, 3)])
@pytest.mark.parametrize("which", [0, 1])
def test_lattice_matches_numpy_restatement(dims, which):
    nb = 4 * dims[0] * dims[1] * dims[2]
    A = oh.bsr_to_dense(nb, *oh.lattice_bsr(dims, which, oh.DEFAULT_SEED))
    B = G.dense_lattice(*dims, oh.DEFAULT_SEED, which)
    assert np.max(np.abs(A - B)) < 1e-15


@pytest.mark.parametrize("L", [3, 5])
def test_lattice_structure(L):
    nb = 4 * L ** 3
    cp, ri, bv = oh.lattice_bsr(L, 0)
    assert np.all(np.diff(cp) == 7)  # onsite + 6 neighbours per block column
    H = oh.bsr_to_dense(nb, cp, ri, bv)
    S = oh.bsr_to_dense(nb, *oh.lattice_bsr(L, 1))
    assert np.array_equal(H, H.T) and np.array_equal(S, S.T)
    assert np.count_nonzero(H) == 28 * nb
    w = np.linalg.eigvalsh(S)
    assert w[0] > 0.25  # Gershgorin: 1 - 24*0.03 = 0.28
    assert np.allclose(np.diag(H)[::4], -5.0, atol=0.1) and np.allclose(np.diag(H)[1::4], 1.0, atol=0.1)


def test_random_c0_statistics_and_determinism():
    c1, c2 = oh.random_c0(64, 512), oh.random_c0(64, 512)
    assert np.array_equal(c1, c2)
    assert np.max(np.abs(c1)) <= 0.5 / np.sqrt(512)
    assert c1[3, 7] == (G.u01(oh.DEFAULT_SEED, 4, 3 + 64 * 7) - 0.5) * (1.0 / np.sqrt(512))


def test_theta_generator_reproduces_shared_hopping():
    th = oh.lattice_theta0(oh.DEFAULT_SEED)
    L = 3
    nb = 4 * L ** 3
    H = oh.bsr_to_dense(nb, *oh.lattice_bsr_theta(L, th))
    T = th[:16].reshape(4, 4, order="F")
    assert np.allclose(H[0:4, 4:8], T) and np.allclose(H[4:8, 0:4], T.T)
    assert np.allclose(H[0:4, 0:4], np.diag(th[16:20]))
    assert np.array_equal(H, H.T)


def test_bloch_generator_phases():
    # k = 0 gives the real operator; the Bloch operator is Hermitian and its spectrum over the
    # phases of a 2-cell supercell is the spectrum of the doubled cell
    L = 3
    nb = 4 * L ** 3
    cp, ri, bv = oh.lattice_bsr(L, 0)
    cpk, rik, bvk = oh.lattice_bsr_k(L, 0, [0.0, 0.0, 0.0])
    assert np.array_equal(cp, cpk) and np.array_equal(ri, rik)
    assert np.array_equal(bvk.real, bv) and np.all(bvk.imag == 0)
    Hk = oh.bsr_to_dense(nb, *oh.lattice_bsr_k(L, 0, [0.3, 0.0, 0.7]))
    assert np.allclose(Hk, Hk.conj().T, atol=1e-15)


def test_exact_eigenvalue_fixture_reproduces():
    # the tabulated exact eigenvalue data (used to report the benchmark's energy error against the
    # exact answer) regenerate from the numpy generator at L = 4
    import json
    import scipy.linalg as sl
    tab = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "lattice_eigs.json")))
    e = tab["L4"]
    H = G.dense_lattice(4, 4, 4, e["seed"], 0)
    S = G.dense_lattice(4, 4, 4, e["seed"], 1)
    w = sl.eigh(H, S, eigvals_only=True)
    m = e["m"]
    assert abs(np.sum(w[:m]) - e["sum_lowest_m"]) < 1e-9
    assert abs((w[m] - w[m - 1]) - e["gap_at_m"]) < 1e-12
    assert set(tab) >= {"L4", "L8", "L16"}
    for v in tab.values():
        assert v["gap_at_m"] > 0 and v["m"] == v["n"] // 2
