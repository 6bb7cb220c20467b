"""Uneven-size group solver (omx_group_*): real systems of different sizes, each with its own H and
S, advanced by omm()'s flavour-0 CG loop with every decision on the device and one step of all
systems captured as one graph (a parallel branch per system).

Checked against the same systems solved one omm() call at a time through the Fortran drop-in
(converged energies to 1e-10 relative, step counts within 3: the device quartic rounds its
transcendental functions differently from the host's), against the closed form (sum of the
occupied generalised eigenvalues, 1e-8 relative), and graph replay against eager launches
(bitwise)."""
import numpy as np
import pytest
import scipy.linalg as sla

import ommhip as oh
import pyomm as F

pytestmark = pytest.mark.gpu

DIMS = [(3, 3, 3), (4, 4, 4), (2, 3, 4), (5, 4, 3)]


def _note(*a):
    print("CASE", *a, flush=True)


def _system(dims):
    """Gapped lattice (theta0 onsite levels -6, -6, 6, 6): H - eta S, S, eta mid-gap, exact sum."""
    th = oh.lattice_theta0()
    cp, ri, h = oh.lattice_bsr_theta(dims, th)
    cps, ris, s = oh.lattice_bsr(dims, 1)
    assert np.array_equal(cp, cps) and np.array_equal(ri, ris)
    nb = 4 * int(np.prod(dims))
    no = nb // 2
    w = sla.eigh(oh.bsr_to_dense(nb, cp, ri, h), oh.bsr_to_dense(nb, cp, ri, s), eigvals_only=True)
    eta = 0.5 * (w[no - 1] + w[no])
    return {"nb": nb, "no": no, "cp": cp, "ri": ri, "h": h - eta * s, "s": s, "eta": eta,
            "exact": float(np.sum(w[:no]))}


@pytest.fixture(scope="module")
def systems():
    return [_system(d) for d in DIMS]


@pytest.fixture(scope="module")
def fsetup():
    F.setup(0)
    yield


def _group(ctx, systems):
    H = [oh.Bsr(ctx, x["nb"], x["cp"], x["ri"], x["h"]) for x in systems]
    S = [oh.Bsr(ctx, x["nb"], x["cp"], x["ri"], x["s"]) for x in systems]
    return oh.Group(ctx, [x["no"] for x in systems], H, S), H + S


def test_group_matches_sequential_and_closed_form(ctx, systems, fsetup):
    _note("group", DIMS)
    g, ops = _group(ctx, systems)
    C0 = [oh.random_c0(x["no"], x["nb"]) for x in systems]
    for k, c in enumerate(C0):
        g.set_c(k, c)
    etas = [x["eta"] for x in systems]
    r = g.solve(etas, cg_tol=1e-10, max_steps=20000, check_every=16, use_graph=True)
    _note("group steps", r["steps"], "icg", r["icg"].tolist())
    assert r["converged"].all()
    for k, x in enumerate(systems):
        nb, no = x["nb"], x["no"]
        Hm, Sm = F.Mat.from_bsr(nb, x["cp"], x["ri"], x["h"]), F.Mat.from_bsr(nb, x["cp"], x["ri"], x["s"])
        Cm = F.Mat(no, nb, "gdden", data=C0[k])
        rs = F.omm(nb, no, Hm, Sm, Cm, init_C=True, eta=x["eta"], cg_tol=1e-10, dealloc=True)
        _note("sys", k, DIMS[k], "group", r["e_min"][k], r["icg"][k], "sequential", rs["e_min"], rs["icg"],
              "exact", x["exact"])
        assert abs(r["e_min"][k] - rs["e_min"]) < 1e-10 * abs(rs["e_min"])
        assert abs(int(r["icg"][k]) - rs["icg"]) <= 3
        assert abs(r["e_min"][k] - x["exact"]) < 1e-8 * abs(x["exact"])
        # the converged coefficients span the same occupied subspace: equal density matrices
        Cg = g.get_c(k)
        Cs = Cm.download()
        Sd = oh.bsr_to_dense(nb, x["cp"], x["ri"], x["s"])

        def dens(Cc):
            W = Cc @ Sd @ Cc.T
            return Cc.T @ (2 * np.eye(no) - W) @ Cc

        assert np.max(np.abs(dens(Cg) - dens(Cs))) < 1e-6
        for M in (Hm, Sm, Cm):
            M.free()
    g.free()
    for o in ops:
        o.free()


def test_group_graph_replay_equals_eager(ctx, systems):
    _note("group graph vs eager")
    res = []
    for use_graph in (True, False):
        g, ops = _group(ctx, systems)
        for k, x in enumerate(systems):
            g.set_c(k, oh.random_c0(x["no"], x["nb"]))
        r = g.solve([x["eta"] for x in systems], cg_tol=1e-10, max_steps=40, check_every=8, use_graph=use_graph)
        res.append((r, [g.get_c(k) for k in range(len(systems))]))
        g.free()
        for o in ops:
            o.free()
    (rg, cg), (re_, ce) = res
    assert np.array_equal(rg["icg"], re_["icg"]) and np.array_equal(rg["e_min"], re_["e_min"])
    for a, b in zip(cg, ce):
        assert np.array_equal(a, b)


def test_group_autotuned_tiles_agree():
    # with tuning on, each system's shapes are timed in the eager first pass; the chosen tiles
    # change only rounding
    sys2 = [_system(d) for d in [(4, 4, 4), (5, 5, 4)]]
    _note("group autotune")
    with oh.Ctx(0) as c:
        out = []
        for tune in (False, True):
            c.autotune(tune)
            g, ops = _group(c, sys2)
            for k, x in enumerate(sys2):
                g.set_c(k, oh.random_c0(x["no"], x["nb"]))
            out.append(g.solve([x["eta"] for x in sys2], cg_tol=1e-10, max_steps=20000, check_every=16))
            g.free()
            for o in ops:
                o.free()
        assert out[0]["converged"].all() and out[1]["converged"].all()
        assert np.max(np.abs(out[0]["e_min"] - out[1]["e_min"]) / np.abs(out[0]["e_min"])) < 1e-10
        assert "gemm" in c.autotune_report()


def test_group_rejects_mismatched_operators(ctx, systems):
    x = systems[0]
    H = oh.Bsr(ctx, x["nb"], x["cp"], x["ri"], x["h"])
    y = systems[1]
    S = oh.Bsr(ctx, y["nb"], y["cp"], y["ri"], y["s"])
    with pytest.raises(oh.OmxError, match="bad sizes"):
        oh.Group(ctx, [x["no"]], [H], [S])
    H.free()
    S.free()


def test_group_second_solve_uses_its_own_tol_and_cap(ctx, systems):
    # as for the k-point batch: cg_tol and max_steps of each solve reach the captured graph
    _note("group re-solve tol/cap")
    sys2 = systems[:2]
    etas = [x["eta"] for x in sys2]
    C0 = [oh.random_c0(x["no"], x["nb"]) for x in sys2]
    cases = [(1e-6, 20000), (1e-10, 20000), (1e-10, 11)]
    g, ops = _group(ctx, sys2)
    got = []
    for tol, cap in cases:
        for k, c in enumerate(C0):
            g.set_c(k, c)
        got.append(g.solve(etas, cg_tol=tol, max_steps=cap, check_every=8, use_graph=True))
    g.free()
    for o in ops:
        o.free()
    for (tol, cap), r in zip(cases, got):
        e, ops = _group(ctx, sys2)
        for k, c in enumerate(C0):
            e.set_c(k, c)
        re_ = e.solve(etas, cg_tol=tol, max_steps=cap, check_every=8, use_graph=False)
        e.free()
        for o in ops:
            o.free()
        _note("tol", tol, "cap", cap, "graph icg", r["icg"].tolist(), "eager icg", re_["icg"].tolist())
        assert np.array_equal(r["icg"], re_["icg"]) and np.array_equal(r["e_min"], re_["e_min"])
    assert got[0]["converged"].all() and got[1]["converged"].all()
    assert (got[0]["icg"] < got[1]["icg"]).all()
    assert (got[2]["icg"] == 11).all() and not got[2]["converged"].any()
