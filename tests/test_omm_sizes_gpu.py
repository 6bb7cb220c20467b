"""omm() on device storage at the benchmark sizes against oracle fixtures (tests/golden/
make_cfg_fixtures.py: the CPU oracle, itself bit-identical to the reference serial build on the
small cases, run once in the build container).

  * 8x8x8 lattice (n=2048, m=1024): the whole solve, cg_tol 1e-9 with the 100-step cap.  Every
    step's energy agrees to 1e-9 relative, the final energy to 1e-10, the step count exactly,
    D_min to 1e-8 (its diagonal, its first 8 columns, and 8 +-1 probes scaled by 1/sqrt(n)).
  * 16x16x16 lattice (n=16384, m=8192): the first three CG steps (the initial products, the
    gradient, two rounds of the four Gram products and two line searches) agree to 1e-11 relative.
  * One k-point of the 10x10x10 cell's 4x4x4 grid, k = (1/4, 1/2, 3/4) (complex128, n=4000,
    m=2000, eta 0.05): the first three CG steps on the complex path ('gzden' state, complex BSR
    operators, Hermitian Grams) against the oracle's complex table, to 1e-11 relative.

The inputs are rebuilt from the seeded generator; their sums are checked against the fixture's so
a generator change cannot pass silently.
"""
import os

import numpy as np
import pytest

import ommhip as oh
import pyomm as F

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _note(*a):
    print("CASE", *a, flush=True)


@pytest.fixture(scope="module", autouse=True)
def _setup():
    F.setup(0)
    yield
    F.set_option("max_iter", 0)


def _inputs(fx):
    L = int(fx["L"])
    nb, no = 4 * L ** 3, 2 * L ** 3
    eta = float(fx["eta"])
    Hb, Sb = oh.lattice_bsr(L, 0), oh.lattice_bsr(L, 1)
    hv = Hb[2] - eta * Sb[2]
    assert float(hv.sum()) == float(fx["h_sum"])
    C0 = oh.random_c0(no, nb)
    assert float(C0.sum()) == float(fx["c0_sum"])
    H = F.Mat.from_bsr(nb, Hb[0], Hb[1], hv)
    S = F.Mat.from_bsr(nb, *Sb)
    return nb, no, eta, H, S, C0


def test_cfg0_8cubed_full_solve_matches_oracle():
    fx = np.load(os.path.join(GOLD, "omm_cfg0_L8.npz"))
    _note("cfg0 8x8x8 n=2048 full solve, cap", int(fx["max_iter"]))
    nb, no, eta, H, S, C0 = _inputs(fx)
    Cm, Dm = F.Mat(no, nb, data=C0), F.Mat(nb, nb)
    F.set_option("max_iter", int(fx["max_iter"]))
    try:
        r = F.omm(nb, no, H, S, Cm, D_min=Dm, init_C=True, eta=eta, cg_tol=float(fx["cg_tol"]), dealloc=True)
    finally:
        F.set_option("max_iter", 0)
    ref_log = fx["e_log"]
    assert r["icg"] == int(fx["icg"])
    assert len(r["e_log"]) == len(ref_log)
    assert np.max(np.abs(r["e_log"] - ref_log) / np.abs(ref_log)) <= 1e-9
    assert abs(r["e_min"] - float(fx["e_min"])) <= 1e-10 * abs(float(fx["e_min"]))
    assert abs(r["TrQS"] - float(fx["TrQS"])) <= 1e-8
    D = Dm.download()
    P = np.where(np.random.default_rng(7).random((nb, 8)) < 0.5, -1.0, 1.0)
    assert np.max(np.abs(np.diag(D) - fx["D_diag"])) < 1e-8
    assert np.max(np.abs(D[:, :8] - fx["D_cols"])) < 1e-8
    assert np.max(np.abs(D @ P - fx["D_probe"])) / np.sqrt(nb) < 1e-8
    for M in (H, S, Cm, Dm):
        M.free()


def test_cfg1_16cubed_first_steps_match_oracle():
    path = os.path.join(GOLD, "omm_cfg1_L16_3steps.npz")
    fx = np.load(path)
    _note("cfg1 16x16x16 n=16384 first", int(fx["max_iter"]), "steps")
    nb, no, eta, H, S, C0 = _inputs(fx)
    Cm = F.Mat(no, nb, data=C0)
    F.set_option("max_iter", int(fx["max_iter"]))
    try:
        r = F.omm(nb, no, H, S, Cm, init_C=True, eta=eta, cg_tol=float(fx["cg_tol"]), dealloc=True)
    finally:
        F.set_option("max_iter", 0)
    ref_log = fx["e_log"]
    assert r["icg"] == int(fx["icg"]) == len(ref_log)
    assert np.max(np.abs(r["e_log"] - ref_log) / np.abs(ref_log)) <= 1e-11
    assert abs(r["e_min"] - float(fx["e_min"])) <= 1e-11 * abs(float(fx["e_min"]))
    assert abs(r["TrQS"] - float(fx["TrQS"])) <= 1e-7
    for M in (H, S, Cm):
        M.free()


def test_cfg3_kpoint_of_the_10cubed_cell_first_steps_match_oracle():
    fx = np.load(os.path.join(GOLD, "omm_cfg3_L10k_3steps.npz"))
    L = int(fx["L"])
    nb, no = 4 * L ** 3, 2 * L ** 3
    phi = tuple(float(x) for x in fx["phi"])
    eta = float(fx["eta"])
    _note("cfg3 k-point", phi, "of the 10x10x10 cell, n =", nb, "first", int(fx["max_iter"]), "steps")
    Hb, Sb = oh.lattice_bsr_k(L, 0, phi), oh.lattice_bsr_k(L, 1, phi)
    hv = Hb[2] - eta * Sb[2]
    assert complex(hv.sum()) == complex(fx["h_sum"])
    C0 = oh.random_c0(no, nb).astype(np.complex128)
    assert complex(C0.sum()) == complex(fx["c0_sum"])
    H, S = F.Mat.from_bsr_z(nb, Hb[0], Hb[1], hv), F.Mat.from_bsr_z(nb, *Sb)
    Cm = F.Mat(no, nb, "gzden", data=C0)
    F.set_option("max_iter", int(fx["max_iter"]))
    try:
        r = F.omm(nb, no, H, S, Cm, init_C=True, eta=eta, cg_tol=float(fx["cg_tol"]), dealloc=True, storage="gzden")
    finally:
        F.set_option("max_iter", 0)
    ref_log = fx["e_log"]
    _note("cfg3 device", r["e_log"].tolist(), "oracle", ref_log.tolist())
    assert r["icg"] == int(fx["icg"]) == len(ref_log)
    assert np.max(np.abs(r["e_log"] - ref_log) / np.abs(ref_log)) <= 1e-11
    assert abs(r["e_min"] - float(fx["e_min"])) <= 1e-11 * abs(float(fx["e_min"]))
    assert abs(r["TrQS"] - float(fx["TrQS"])) <= 1e-7
    P = np.where(np.random.default_rng(7).random((nb, 8)) < 0.5, -1.0, 1.0)
    Cp = Cm.download() @ P
    assert np.max(np.abs(Cp - fx["C_probe"])) / np.sqrt(nb) < 1e-9
    for M in (H, S, Cm):
        M.free()
