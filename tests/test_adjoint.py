"""Adjoint of the converged OMM energy on the CPU: the lattice generator's Jacobian, and the
Hellmann-Feynman formula dE/dtheta_p = <D, dH/dtheta_p> (D = C^T (2I - SW) C, the density matrix of
the converged solve) against central finite differences of oracle solves and against Euler's
identity sum_p theta_p dE/dtheta_p = e_min (the energy is homogeneous of degree one in theta
while no level crosses eta)."""
import numpy as np
import pytest

import ommhip as oh
import oracle as O


def _apply_jac(seg, ent, coef, theta, size):
    v = np.zeros(size)
    for p in range(oh.THETA_COUNT):
        np.add.at(v, ent[seg[p]:seg[p + 1]], theta[p] * coef[seg[p]:seg[p + 1]])
    return v


@pytest.mark.parametrize("L", [2, 3, (3, 4, 2)])
def test_jacobian_reproduces_generator(L):
    rng = np.random.default_rng(3)
    th = oh.lattice_theta0() + 0.1 * rng.standard_normal(oh.THETA_COUNT)
    cp, ri, bv = oh.lattice_bsr_theta(L, th)
    seg, ent, coef = oh.lattice_theta_jacobian(L)
    assert seg[0] == 0 and np.all(np.diff(seg) > 0) and seg[-1] == len(ent)
    v = _apply_jac(seg, ent, coef, th, bv.size)
    dims = (L, L, L) if np.isscalar(L) else L
    if min(dims) < 3:
        # the +d and -d bonds of an edge-2 direction coincide and their blocks are merged
        assert np.allclose(v, bv.ravel(), rtol=0, atol=1e-14)
    else:
        assert np.array_equal(v, bv.ravel())
        ns = int(np.prod(dims))
        # every hopping entry sits on 3 bonds per site, in both directions; onsite levels once per site
        assert np.all(np.diff(seg)[:16] == 6 * ns) and np.all(np.diff(seg)[16:] == ns)


def _dense_of_entries(nb, cp, ri, vals):
    """Map block-value index e = 16*l + r + 4*c -> dense (row, col)."""
    nbc = len(cp) - 1
    J = np.repeat(np.arange(nbc), np.diff(cp))
    e = np.arange(16 * len(ri))
    l, q = e // 16, e % 16
    return 4 * ri[l] + q % 4, 4 * J[l] + q // 4


def test_hellmann_feynman_against_finite_differences():
    L = 2
    nb, no = 4 * L ** 3, 2 * L ** 3
    th = oh.lattice_theta0()
    cp, ri, bv = oh.lattice_bsr_theta(L, th)
    S = oh.bsr_to_dense(nb, *oh.lattice_bsr(L, 1))
    import scipy.linalg as sla
    w = sla.eigh(oh.bsr_to_dense(nb, cp, ri, bv), S, eigvals_only=True)
    eta = 0.5 * (w[no - 1] + w[no])
    exact = float(np.sum(w[:no]))

    def solve(theta, C):
        H = oh.bsr_to_dense(nb, *oh.lattice_bsr_theta(L, theta)) - eta * S
        return O.Omm()(nb, no, H, S, C, init_C=True, eta=eta, cg_tol=1e-15, dealloc=True, max_iter=5000)

    C = np.asfortranarray(oh.random_c0(no, nb))
    r = solve(th, C)
    assert abs(r["e_min"] - exact) < 1e-10 * abs(exact)
    rows, cols = _dense_of_entries(nb, cp, ri, bv)
    dvals = r["D"][rows, cols]
    seg, ent, coef = oh.lattice_theta_jacobian(L)
    grad = np.array([np.dot(coef[seg[p]:seg[p + 1]], dvals[ent[seg[p]:seg[p + 1]]]) for p in range(oh.THETA_COUNT)])
    # Euler: e_min(lambda*theta) = lambda*e_min(theta) locally
    assert abs(np.dot(th, grad) - r["e_min"]) < 1e-9 * abs(r["e_min"])
    h = 1e-5
    fd = np.zeros(oh.THETA_COUNT)
    for p in range(oh.THETA_COUNT):
        e = []
        for sgn in (1, -1):
            t2 = th.copy()
            t2[p] += sgn * h
            e.append(solve(t2, C.copy(order="F"))["e_min"])
        fd[p] = (e[0] - e[1]) / (2 * h)
    # D enters to first order in the error of C, which FP64 CG drives only to about sqrt(eps) of
    # the energy; the energy-based finite differences are second order.  Observed agreement
    # 4e-6 of the largest derivative.
    assert np.max(np.abs(fd - grad)) < 2e-5 * np.max(np.abs(grad))
