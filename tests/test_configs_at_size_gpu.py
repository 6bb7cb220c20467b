"""The batched, grouped and large-system solvers at the benchmark workloads' sizes.

  * 64 k-points of the 10x10x10 cell (n = 4000, m = 2000, complex128, Bloch phases on the 4x4x4
    grid): one batched solve (graph-captured steps, device-side line search) against one omm()
    call per k-point (np = 64) after 1, 2 and 3 CG steps, every k-point's energy to 1e-10
    relative, and each energy below the previous step's.
  * 24 real systems with edges L drawn from U{6..20} (n from 864 to 32000): one group solve against
    one omm() call per system after 3 CG steps, energies to 1e-10 relative.
  * The 32x32x16 lattice (n = 65536, m = 32768, 215 GB of FP64 state): the mixed solve's low-
    precision phase for 4 steps and one FP64 refinement step, with the density matrix on H's
    pattern.  The energies decrease step by step, and the density matrix obeys the size-independent
    identity sum_ij D_ij S_ij = Tr[(2I - SW) SW] (both are tr[(2I - SW) C S C^T]).

The small-size checks of the same solvers against the closed form and the oracle are in
test_batch_gpu.py, test_group_gpu.py and test_omm_gpu.py.
"""
import numpy as np
import pytest

import ommhip as oh
import pyomm as F

pytestmark = pytest.mark.gpu


def _note(*a):
    print("CASE", *a, flush=True)


@pytest.fixture(scope="module", autouse=True)
def _setup():
    F.setup(0)
    yield
    F.set_option("max_iter", 0)
    F.set_option("lowp_max_iter", 0)


def test_64_kpoints_of_the_10cubed_cell_batched_matches_sequential(ctx):
    L, nk1, eta = 10, 4, 0.05
    nb, no = 4 * L ** 3, 2 * L ** 3
    g = [i / nk1 for i in range(nk1)]
    phis = [(a, b, c) for a in g for b in g for c in g]
    _note("k-point batch", len(phis), "x n", nb)
    hv, sv = [], []
    for phi in phis:
        cp, ri, h = oh.lattice_bsr_k(L, 0, phi)
        cps, ris, s = oh.lattice_bsr_k(L, 1, phi)
        assert np.array_equal(cp, cps) and np.array_equal(ri, ris)
        hv.append(h - eta * s)
        sv.append(s)
    C0 = oh.random_c0(no, nb).astype(np.complex128)
    steps = (1, 2, 3)
    B = oh.ZBatch(ctx, no, nb, cp, ri, hv, sv)
    batched = {}
    for K in steps:
        B.set_c([C0] * len(phis))
        r = B.solve(eta=eta, cg_tol=1e-14, max_steps=K, check_every=K, use_graph=True)
        assert (r["icg"] == K).all()
        batched[K] = r["e_min"].copy()
    B.free()
    Hm = [F.Mat.from_bsr_z(nb, cp, ri, h) for h in hv]
    Sm = [F.Mat.from_bsr_z(nb, cp, ri, s) for s in sv]
    Cm = F.Mat(no, nb, "gzden")
    worst = 0.0
    try:
        for K in steps:
            F.set_option("max_iter", K)
            for k in range(len(phis)):
                Cm.upload(C0)
                rs = F.omm(nb, no, Hm[k], Sm[k], Cm, init_C=True, eta=eta, cg_tol=1e-14, storage="gzden",
                           np_=len(phis), ip=k + 1, dealloc=(k == len(phis) - 1))
                assert rs["icg"] == K
                rel = abs(batched[K][k] - rs["e_min"]) / abs(rs["e_min"])
                worst = max(worst, rel)
                assert rel <= 1e-10, (K, k, batched[K][k], rs["e_min"])
    finally:
        F.set_option("max_iter", 0)
    _note("k-point batch worst relative difference", worst)
    assert (batched[2] < batched[1]).all() and (batched[3] < batched[2]).all()
    for M in Hm + Sm + [Cm]:
        M.free()


def test_24_uneven_systems_group_matches_sequential(ctx):
    count, stream, eta, K = 24, 6, 0.0517, 3
    seed = oh.DEFAULT_SEED
    Ls = [6 + int(15 * oh.u01(seed, stream, s)) for s in range(count)]
    _note("group", count, "systems, L =", Ls)
    ops, C0 = [], []
    for L in Ls:
        cp, ri, h = oh.lattice_bsr(L, 0)
        _, _, s = oh.lattice_bsr(L, 1)
        nb = 4 * L ** 3
        ops.append((nb, cp, ri, h - eta * s, s))
        C0.append(oh.random_c0(nb // 2, nb))
    H = [oh.Bsr(ctx, nb, cp, ri, h) for nb, cp, ri, h, _ in ops]
    S = [oh.Bsr(ctx, nb, cp, ri, s) for nb, cp, ri, _, s in ops]
    grp = oh.Group(ctx, [nb // 2 for nb, *_ in ops], H, S)
    for k, c in enumerate(C0):
        grp.set_c(k, c)
    r = grp.solve([eta] * count, cg_tol=1e-14, max_steps=K, check_every=K, use_graph=True)
    grp.free()
    for M in H + S:
        M.free()
    assert (r["icg"] == K).all()
    worst = 0.0
    F.set_option("max_iter", K)
    try:
        for k, (nb, cp, ri, h, s) in enumerate(ops):
            Hm, Sm = F.Mat.from_bsr(nb, cp, ri, h), F.Mat.from_bsr(nb, cp, ri, s)
            Cm = F.Mat(nb // 2, nb, data=C0[k])
            rs = F.omm(nb, nb // 2, Hm, Sm, Cm, init_C=True, eta=eta, cg_tol=1e-14, dealloc=True)
            for M in (Hm, Sm, Cm):
                M.free()
            rel = abs(r["e_min"][k] - rs["e_min"]) / abs(rs["e_min"])
            worst = max(worst, rel)
            assert rs["icg"] == K and rel <= 1e-10, (k, Ls[k], r["e_min"][k], rs["e_min"])
    finally:
        F.set_option("max_iter", 0)
    _note("group worst relative difference", worst)


def test_32x32x16_mixed_steps_and_density_identity():
    dims, eta, K = (32, 32, 16), 0.0517, 4
    nb = 4 * dims[0] * dims[1] * dims[2]
    no = nb // 2
    _note("32x32x16 mixed", K, "low-precision steps + 1 FP64 step, n =", nb)
    cp, ri, h = oh.lattice_bsr(dims, 0)
    _, _, s = oh.lattice_bsr(dims, 1)
    Hm, Sm = F.Mat.from_bsr(nb, cp, ri, h - eta * s), F.Mat.from_bsr(nb, cp, ri, s)
    Dsp = F.Mat.from_bsr(nb, cp, ri, np.zeros_like(h))
    Cm = F.Mat(no, nb, data=oh.random_c0(no, nb))
    F.set_option("lowp_max_iter", K)
    F.set_option("max_iter", 1)
    try:
        r = F.omm(nb, no, Hm, Sm, Cm, D_min=Dsp, init_C=True, eta=eta, cg_tol=1e-300, operation="hmx",
                  dealloc=True)
    finally:
        F.set_option("lowp_max_iter", 0)
        F.set_option("max_iter", 0)
    e = r["e_log"]
    _note("32x32x16 energies", e.tolist(), "TrQS", r["TrQS"])
    assert r["icg_lowp"] == K and r["icg"] == K + 1 and len(e) == K + 1
    assert np.all(np.isfinite(e)) and np.all(np.diff(e) < 0)
    dv = Dsp.bsr_values()
    assert np.all(np.isfinite(dv))
    tr_ds = float(np.sum(dv * s))  # sum over the stored entries of D_ij S_ij (S lives on H's pattern)
    _note("32x32x16 sum D.S", tr_ds, "TrQS", r["TrQS"])
    assert abs(tr_ds - r["TrQS"]) <= 1e-9 * abs(r["TrQS"])
    for M in (Hm, Sm, Dsp, Cm):
        M.free()
