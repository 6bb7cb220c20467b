"""Dense factorisations of the Cholesky and preconditioned flavours (omx_potrf, omx_trsm, omx_trmm,
omx_sygst, omx_inverse_spd) against numpy/scipy FP64, on orders that exercise partial 64- and
128-wide diagonal blocks and more than one 512-wide panel.  Tolerance 1e-11 relative to the
operand scale (blocked order and block inverses against LAPACK's substitution)."""
import ctypes as C

import numpy as np
import pytest
import scipy.linalg as sla

import ommhip as oh

pytestmark = pytest.mark.gpu
rng = np.random.default_rng(5)


def _spd(n):
    A = rng.standard_normal((n, n))
    return A @ A.T / n + np.eye(n)


@pytest.mark.parametrize("n", [1, 31, 64, 100, 257, 600, 1100, 1600])
def test_potrf(ctx, n):
    print("CASE potrf", n, flush=True)
    S = _spd(n)
    dS = oh.DMat.from_host(ctx, S)
    info = C.c_int(-1)
    oh.check(oh.lib().omx_potrf(ctx.h, oh.OMX_F64, n, dS.p, dS.ld, C.byref(info)), "potrf")
    assert info.value == 0
    U = np.triu(dS.download())
    assert np.max(np.abs(U - sla.cholesky(S, lower=False))) < 1e-11 * n
    # lower triangle of the input is never read: a non-SPD lower part changes nothing
    S2 = np.triu(S) + np.tril(rng.standard_normal((n, n)), -1)
    dS2 = oh.DMat.from_host(ctx, S2)
    oh.check(oh.lib().omx_potrf(ctx.h, oh.OMX_F64, n, dS2.p, dS2.ld, C.byref(info)), "potrf")
    assert np.array_equal(np.triu(dS2.download()), U)
    # and, as with dpotrf 'u', it is left as it was
    assert np.array_equal(np.tril(dS2.download(), -1), np.tril(S2, -1))


@pytest.mark.parametrize("n,bad", [(40, 35), (600, 400), (700, 600), (1600, 1300)])
def test_potrf_not_spd(ctx, n, bad):
    """A negative diagonal entry makes that pivot the first non-positive one (earlier pivots do not
    depend on it): info names it, 1-based, also past the first 64-wide diagonal blocks and past the
    first 512-wide panel."""
    print("CASE potrf not spd", n, bad, flush=True)
    S = _spd(n)
    S[bad, bad] = -1.0
    dS = oh.DMat.from_host(ctx, S)
    info = C.c_int(-1)
    oh.check(oh.lib().omx_potrf(ctx.h, oh.OMX_F64, n, dS.p, dS.ld, C.byref(info)), "potrf")
    assert info.value == bad + 1


@pytest.mark.parametrize("side", ["l", "r"])
@pytest.mark.parametrize("trans", ["n", "t"])
@pytest.mark.parametrize("nu", [75, 600, 1100])
def test_trsm_trmm(ctx, side, trans, nu):
    other = 50
    print("CASE trsm", side, trans, nu, flush=True)
    U = sla.cholesky(_spd(nu), lower=False)
    Ug = U + np.tril(rng.standard_normal((nu, nu)), -1)  # garbage below the diagonal is ignored
    B = rng.standard_normal((nu, other) if side == "l" else (other, nu))
    m, n = B.shape
    opU = U.T if trans == "t" else U
    dU, dB = oh.DMat.from_host(ctx, Ug), oh.DMat.from_host(ctx, B)
    oh.check(oh.lib().omx_trsm(ctx.h, oh.OMX_F64, side.encode(), trans.encode(), m, n, 0.5, dU.p, dU.ld, dB.p,
                               dB.ld), "trsm")
    ref = 0.5 * (np.linalg.solve(opU, B) if side == "l" else np.linalg.solve(opU.T, B.T).T)
    assert np.max(np.abs(dB.download() - ref)) < 1e-11 * np.max(np.abs(ref)) * nu
    dB = oh.DMat.from_host(ctx, B)
    oh.check(oh.lib().omx_trmm(ctx.h, oh.OMX_F64, side.encode(), trans.encode(), m, n, -2.0, dU.p, dU.ld, dB.p,
                               dB.ld), "trmm")
    ref = -2.0 * (opU @ B if side == "l" else B @ opU)
    assert np.max(np.abs(dB.download() - ref)) < 1e-12 * np.max(np.abs(ref)) * nu


@pytest.mark.parametrize("n", [90, 600, 1100])
def test_sygst_and_inverse(ctx, n):
    print("CASE sygst/inverse", n, flush=True)
    S = _spd(n)
    H = rng.standard_normal((n, n))
    H = H + H.T
    U = sla.cholesky(S, lower=False)
    dH, dU = oh.DMat.from_host(ctx, H), oh.DMat.from_host(ctx, U)
    oh.check(oh.lib().omx_sygst(ctx.h, oh.OMX_F64, n, dH.p, dH.ld, dU.p, dU.ld), "sygst")
    Ui = np.linalg.inv(U)
    ref = Ui.T @ H @ Ui
    got = dH.download()
    assert np.array_equal(got, got.T)
    assert np.max(np.abs(got - ref)) < 1e-11 * np.max(np.abs(ref)) * n
    # eigenvalues of the standard problem are those of the generalized one
    assert np.allclose(np.linalg.eigvalsh(got), sla.eigh(H, S, eigvals_only=True), atol=1e-10)
    dS = oh.DMat.from_host(ctx, S)
    info = C.c_int(-1)
    oh.check(oh.lib().omx_inverse_spd(ctx.h, oh.OMX_F64, n, dS.p, dS.ld, C.byref(info)), "inverse")
    assert info.value == 0
    assert np.max(np.abs(dS.download() - np.linalg.inv(S))) < 1e-11 * n
