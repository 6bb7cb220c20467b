"""Batched k-point solver (omx_zbatch_*): the k-points of one cell advanced in lockstep by omm()'s
flavour-0 CG loop with every decision on the device, one step captured in a hipGraph.

Checked against the same systems solved one omm() call at a time through the Fortran drop-in
(np = nk per-k state; converged energies to 1e-10 relative, step counts within 3: the device's
quartic solver rounds its transcendental functions differently from the host's), against the
closed form (sum of the occupied generalised eigenvalues below eta, 1e-8 relative), and graph
replay against eager launches (bitwise)."""
import numpy as np
import pytest
import scipy.linalg as sla

import ommhip as oh
import pyomm as F

pytestmark = pytest.mark.gpu


def _note(*a):
    print("CASE", *a, flush=True)


def _kgrid(nk1):
    g = [i / nk1 for i in range(nk1)]
    return [(a, b, c) for a in g for b in g for c in g]


def _systems(L, phis, eta):
    hv, sv = [], []
    for phi in phis:
        cp, ri, h = oh.lattice_bsr_k(L, 0, phi)
        cps, ris, s = oh.lattice_bsr_k(L, 1, phi)
        assert np.array_equal(cp, cps) and np.array_equal(ri, ris)
        hv.append(h - eta * s)
        sv.append(s)
    return cp, ri, hv, sv


@pytest.fixture(scope="module")
def fsetup():
    F.setup(0)
    yield


def test_batch_matches_sequential_and_closed_form(ctx, fsetup):
    L, eta = 3, 0.2
    nb, no = 4 * L ** 3, 2 * L ** 3
    phis = _kgrid(2)
    _note("zbatch", L, len(phis))
    cp, ri, hv, sv = _systems(L, phis, eta)
    C0 = oh.random_c0(no, nb).astype(np.complex128)
    B = oh.ZBatch(ctx, no, nb, cp, ri, hv, sv)
    B.set_c([C0] * len(phis))
    r = B.solve(eta=eta, cg_tol=1e-10, max_steps=20000, check_every=16, use_graph=True)
    _note("zbatch steps", r["steps"], "icg", r["icg"].tolist())
    assert r["converged"].all()
    for k, phi in enumerate(phis):
        Hd = oh.bsr_to_dense(nb, cp, ri, hv[k] + eta * sv[k])
        Sd = oh.bsr_to_dense(nb, cp, ri, sv[k])
        w = sla.eigh(Hd, Sd, eigvals_only=True)
        exact = float(np.sum(w[:no][w[:no] < eta]))
        Hm, Sm = F.Mat.from_bsr_z(nb, cp, ri, hv[k]), F.Mat.from_bsr_z(nb, cp, ri, sv[k])
        Cm = F.Mat(no, nb, "gzden", data=C0)
        rs = F.omm(nb, no, Hm, Sm, Cm, init_C=True, eta=eta, cg_tol=1e-10, dealloc=True, storage="gzden")
        _note("k", k, "batch", r["e_min"][k], r["icg"][k], "sequential", rs["e_min"], rs["icg"], "exact", exact)
        assert abs(r["e_min"][k] - rs["e_min"]) < 1e-10 * abs(rs["e_min"])
        assert abs(int(r["icg"][k]) - rs["icg"]) <= 3
        assert abs(r["e_min"][k] - exact) < 1e-8 * abs(exact)
        for M in (Hm, Sm, Cm):
            M.free()
    B.free()


def test_batch_graph_replay_equals_eager(ctx):
    L, eta = 3, 0.2
    nb, no = 4 * L ** 3, 2 * L ** 3
    phis = [(0.0, 0.0, 0.0), (0.25, 0.5, 0.0), (0.5, 0.5, 0.5)]
    _note("zbatch graph vs eager")
    cp, ri, hv, sv = _systems(L, phis, eta)
    C0 = oh.random_c0(no, nb).astype(np.complex128)
    out = []
    for use_graph in (False, True):
        B = oh.ZBatch(ctx, no, nb, cp, ri, hv, sv)
        B.set_c([C0] * len(phis))
        r = B.solve(eta=eta, cg_tol=1e-10, max_steps=60, check_every=7, use_graph=use_graph)
        out.append((r, B.get_c()))
        B.free()
    (r0, c0), (r1, c1) = out
    assert np.array_equal(r0["e_min"], r1["e_min"]) and np.array_equal(r0["icg"], r1["icg"])
    for a, b in zip(c0, c1):
        assert np.array_equal(a, b)


def test_batch_complex64_matches_closed_form(ctx):
    # the k-point batch with complex64 state (FP32 MFMA products, FP64 traces and line search):
    # every system converged to cg_tol = 1e-7 and within 1e-5 relative of the closed form, as a
    # single complex64 omm() solve
    L = 3
    nb, no = 4 * L ** 3, 2 * L ** 3
    phis = [(0.0, 0.0, 0.0), (0.25, 0.5, 0.75), (0.0, 0.5, 0.25)]  # a common gap (0.051, 0.134)
    _note("zbatch complex64")
    hv, sv, ws = [], [], []
    for phi in phis:
        cp, ri, h = oh.lattice_bsr_k(L, 0, phi)
        _, _, s = oh.lattice_bsr_k(L, 1, phi)
        ws.append(sla.eigh(oh.bsr_to_dense(nb, cp, ri, h), oh.bsr_to_dense(nb, cp, ri, s), eigvals_only=True))
        hv.append(h)
        sv.append(s)
    # one shift for the batch, inside every system's gap
    lo, hi = max(w[no - 1] for w in ws), min(w[no] for w in ws)
    assert lo < hi
    eta = 0.5 * (lo + hi)
    exact = [float(np.sum(w[:no])) for w in ws]
    hv = [h - eta * s for h, s in zip(hv, sv)]
    C0 = oh.random_c0(no, nb).astype(np.complex64)
    B = oh.ZBatch(ctx, no, nb, cp, ri, hv, sv, dt=oh.OMX_C64)
    B.set_c([C0] * len(phis))
    r = B.solve(eta=eta, cg_tol=1e-7, max_steps=20000, check_every=16, use_graph=True)
    _note("zbatch c64 icg", r["icg"].tolist(), "e", r["e_min"].tolist(), "exact", exact)
    assert r["converged"].all()
    for k in range(len(phis)):
        assert abs(r["e_min"][k] - exact[k]) < 1e-5 * abs(exact[k])
    assert B.get_c()[0].dtype == np.complex64
    B.free()


def test_batch_second_solve_uses_its_own_tol_and_cap(ctx):
    # cg_tol and max_steps of each solve reach the captured steps through device memory: a batch
    # solved at 1e-6, then again at 1e-10 and with a step cap, gives bitwise the result of a fresh
    # eager solve with those values
    L, eta = 3, 0.2
    nb, no = 4 * L ** 3, 2 * L ** 3
    phis = [(0.0, 0.0, 0.0), (0.25, 0.5, 0.0)]
    _note("zbatch re-solve tol/cap")
    cp, ri, hv, sv = _systems(L, phis, eta)
    C0 = oh.random_c0(no, nb).astype(np.complex128)
    cases = [(1e-6, 20000), (1e-10, 20000), (1e-10, 9)]
    B = oh.ZBatch(ctx, no, nb, cp, ri, hv, sv)
    got = []
    for tol, cap in cases:
        B.set_c([C0] * len(phis))
        got.append((B.solve(eta=eta, cg_tol=tol, max_steps=cap, check_every=8, use_graph=True), B.get_c()))
    B.free()
    for (tol, cap), (r, cs) in zip(cases, got):
        E = oh.ZBatch(ctx, no, nb, cp, ri, hv, sv)
        E.set_c([C0] * len(phis))
        re_ = E.solve(eta=eta, cg_tol=tol, max_steps=cap, check_every=8, use_graph=False)
        ce = E.get_c()
        E.free()
        _note("tol", tol, "cap", cap, "graph icg", r["icg"].tolist(), "eager icg", re_["icg"].tolist())
        assert np.array_equal(r["icg"], re_["icg"]) and np.array_equal(r["e_min"], re_["e_min"])
        assert np.array_equal(r["converged"], re_["converged"])
        for a, b in zip(cs, ce):
            assert np.array_equal(a, b)
    (r6, _), (r10, _), (rc, _) = got
    assert r6["converged"].all() and r10["converged"].all()
    assert (r6["icg"] < r10["icg"]).all()
    assert (rc["icg"] == 9).all() and not rc["converged"].any()
