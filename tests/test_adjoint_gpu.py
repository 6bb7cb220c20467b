"""Adjoint of the converged OMM energy on the device: the dE/dtheta contraction kernel, the density
matrix on H's pattern ('gdcsc' D_min, SDDMM) against the dense one, and dE/dtheta against central
finite differences (h = 1e-5, warm-started FP64 solves) and Euler's identity
sum_p theta_p dE/dtheta_p = e_min, on a 3x3x3 lattice and on the 12x12x12 system (n = 6912).

Tolerances: the contraction 1e-13 relative (summation order); sparse vs dense density 1e-12;
finite differences 2e-5 (3x3x3) and 1e-4 (12x12x12) of the largest derivative: D, and so the
adjoint, carries the error of C to first order, and FP64 CG on the energy drives C only to about
sqrt(eps) of the energy; the finite differences of the (variational) energy are second order."""
import numpy as np
import pytest

import ommhip as oh
import pyomm as F

pytestmark = pytest.mark.gpu

# mid-gap eta = (lambda_m + lambda_{m+1})/2 of H(theta0) x = lambda S x for the default seed
# (scipy generalised eigensolver on the host), and the gap
ETA_THETA0 = {3: (0.5273898630346915, 4.144775489523899), 4: (0.6716930386128855, 3.845393787313842),
              12: (0.6723397363468351, 3.8575065331396274)}


def _note(*a):
    print("CASE", *a, flush=True)


@pytest.fixture(scope="module", autouse=True)
def _setup():
    F.setup(0)
    yield


class DevJac:
    """The lattice Jacobian in HBM (int64 seg/ent, FP64 coef) for omx_bsr_param_grad."""

    def __init__(self, ctx, L):
        self.seg, self.ent, self.coef = oh.lattice_theta_jacobian(L)
        lib = oh.lib()
        self.bufs = []
        for a in (self.seg, self.ent, self.coef):
            p = oh.C.c_void_p()
            oh.check(lib.omx_malloc(ctx.h, a.nbytes, oh.C.byref(p)), "malloc")
            oh.check(lib.omx_h2d(ctx.h, p, a.ctypes.data_as(oh.C.c_void_p), a.nbytes), "h2d")
            self.bufs.append(p)
        self.ctx = ctx

    def ptrs(self):
        return [b.value for b in self.bufs]

    def free(self):
        for b in self.bufs:
            oh.lib().omx_free(self.ctx.h, b)
        self.bufs = []


def test_param_grad_kernel(ctx):
    L = 4
    _note("param_grad kernel", L)
    cp, ri, bv = oh.lattice_bsr_theta(L, oh.lattice_theta0())
    rng = np.random.default_rng(11)
    vals = rng.standard_normal(bv.shape)
    Dp = oh.Bsr(ctx, 4 * L ** 3, cp, ri, vals)
    J = DevJac(ctx, L)
    dseg, dent, dcoef = J.ptrs()
    got = np.zeros(oh.THETA_COUNT)
    oh.check(oh.lib().omx_bsr_param_grad(ctx.h, Dp.h, oh.THETA_COUNT, dseg, dent, dcoef,
                                         got.ctypes.data_as(oh.C.c_void_p)), "param_grad")
    flat = vals.ravel()
    ref = np.array([np.dot(J.coef[J.seg[p]:J.seg[p + 1]], flat[J.ent[J.seg[p]:J.seg[p + 1]]])
                    for p in range(oh.THETA_COUNT)])
    assert np.allclose(got, ref, rtol=1e-13, atol=1e-12)
    # element access: (2, 5) is entry (1, 0) of block (0, 1)
    l = cp[1] + int(np.flatnonzero(ri[cp[1]:cp[2]] == 0)[0])
    assert oh.bsr_get_element(Dp, 2, 5) == vals[l, 1]
    assert oh.bsr_get_element(Dp, 1, 4 * L ** 3 // 2 + 1) == 0.0
    J.free()


def _system(L):
    nb, no = 4 * L ** 3, 2 * L ** 3
    th = oh.lattice_theta0()
    eta = ETA_THETA0[L][0]
    Sb = oh.lattice_bsr(L, 1)
    return nb, no, th, eta, Sb


def _H_mat(L, th, eta, Sb):
    cp, ri, bv = oh.lattice_bsr_theta(L, th)
    assert np.array_equal(cp, Sb[0]) and np.array_equal(ri, Sb[1])
    return F.Mat.from_bsr(4 * L ** 3, cp, ri, bv - eta * Sb[2]), (cp, ri, bv)


def test_sparse_density_matches_dense():
    L = 3
    _note("sparse D", L)
    nb, no, th, eta, Sb = _system(L)
    Hm, (cp, ri, bv) = _H_mat(L, th, eta, Sb)
    Sm = F.Mat.from_bsr(nb, *Sb)
    C0 = oh.random_c0(no, nb)
    Cm = F.Mat(no, nb, data=C0)
    Dd = F.Mat(nb, nb)
    r1 = F.omm(nb, no, Hm, Sm, Cm, D_min=Dd, init_C=True, eta=eta, cg_tol=1e-10, dealloc=True)
    Cm.upload(C0)
    Dsp = F.Mat.from_bsr(nb, cp, ri, np.zeros_like(bv))
    r2 = F.omm(nb, no, Hm, Sm, Cm, D_min=Dsp, init_C=True, eta=eta, cg_tol=1e-10, dealloc=True)
    assert r1["icg"] == r2["icg"] and r1["e_min"] == r2["e_min"]
    Ddense = Dd.download()
    for (i, j) in [(1, 1), (2, 1), (4, 4), (5, 1), (nb, nb), (1, 13), (13, 1)]:
        dv = Dsp.get_element(i, j)
        blockI, blockJ = (i - 1) // 4, (j - 1) // 4
        stored = blockI in ri[cp[blockJ]:cp[blockJ + 1]]
        assert (abs(dv - Ddense[i - 1, j - 1]) < 1e-12) if stored else dv == 0.0
    for M in (Hm, Sm, Cm, Dd, Dsp):
        M.free()


def _adjoint_fd(L, params, h=1e-5, tol=1e-14, fd_tol=1e-14, base_max_iter=5000, fd_max_iter=0):
    nb, no, th, eta, Sb = _system(L)
    ctx_h = F.ctx_handle()
    Sm = F.Mat.from_bsr(nb, *Sb)
    Hm, (cp, ri, bv) = _H_mat(L, th, eta, Sb)
    Cm = F.Mat(no, nb, data=oh.random_c0(no, nb))
    Dsp = F.Mat.from_bsr(nb, cp, ri, np.zeros_like(bv))
    F.set_option("max_iter", base_max_iter)
    try:
        r = F.omm(nb, no, Hm, Sm, Cm, D_min=Dsp, init_C=True, eta=eta, cg_tol=tol, dealloc=True)
    finally:
        F.set_option("max_iter", 0)
    _note("adjoint base solve", L, "icg", r["icg"], "e", r["e_min"])
    Cstar = Cm.download()

    class _Ctx:
        h = ctx_h
    J = DevJac(_Ctx, L)
    grad = F.param_grad(Dsp, oh.THETA_COUNT, *J.ptrs())
    J.free()
    euler = float(np.dot(th, grad))
    fd = {}
    for p in params:
        e = []
        for sgn in (1, -1):
            t2 = th.copy()
            t2[p] += sgn * h
            H2, _ = _H_mat(L, t2, eta, Sb)
            Cm.upload(Cstar)
            F.set_option("max_iter", fd_max_iter)
            try:
                e.append(F.omm(nb, no, H2, Sm, Cm, init_C=True, eta=eta, cg_tol=fd_tol, dealloc=True)["e_min"])
            finally:
                F.set_option("max_iter", 0)
            H2.free()
        fd[p] = (e[0] - e[1]) / (2 * h)
        _note("adjoint fd", L, p, grad[p], fd[p])
    for M in (Hm, Sm, Cm, Dsp):
        M.free()
    return r, grad, euler, fd


def test_adjoint_small_lattice():
    r, grad, euler, fd = _adjoint_fd(3, range(oh.THETA_COUNT))
    assert abs(euler - r["e_min"]) < 1e-9 * abs(r["e_min"])
    scale = np.max(np.abs(grad))
    assert max(abs(fd[p] - grad[p]) for p in fd) < 2e-5 * scale


def test_adjoint_12_cubed():
    """The 12x12x12 system (n = 6912, m = 3456, FP64): all 20 parameters; warm-started finite
    difference solves capped at 150 steps, the base solve at 3000."""
    r, grad, euler, fd = _adjoint_fd(12, range(oh.THETA_COUNT), tol=1e-13, fd_tol=1e-13, base_max_iter=3000,
                                     fd_max_iter=150)
    assert abs(euler - r["e_min"]) < 1e-9 * abs(r["e_min"])
    scale = np.max(np.abs(grad))
    err = max(abs(fd[p] - grad[p]) for p in fd)
    print("CASE adjoint 12^3 max |fd - adjoint| / max |adjoint| =", err / scale, flush=True)
    assert err < 1e-4 * scale
