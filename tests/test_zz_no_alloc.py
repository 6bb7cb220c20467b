"""omx_omm_pr_dots_noalloc (the allocation-free Polak-Ribiere dots, static device scratch) through
the C-ABI against omx_omm_pr_dots on the same inputs: the same launch, so the same bits.  Two calls
in a row with the second input smaller (the scratch keeps the first call's contents), real and
complex, and a numpy FP64 check of the values (1e-12 relative)."""
import ctypes

import numpy as np
import pytest

import ommhip as oh

pytestmark = pytest.mark.gpu


def _dots(fn, ctx, G, Gp):
    out = np.zeros(2)
    oh.check(fn(ctx.h, G.dt, G.rows, G.cols, G.p, Gp.p, G.ld, out.ctypes.data_as(ctypes.c_void_p)), "pr_dots")
    return out


@pytest.mark.parametrize("dt", [oh.OMX_F64, oh.OMX_C128])
def test_pr_dots_noalloc_matches_allocating_entry(ctx, dt):
    r = np.random.default_rng(9)
    for n, m in ((640, 2048), (100, 37)):  # the second call is smaller
        print("CASE pr_dots_noalloc", dt, n, m, flush=True)
        shape = (n, m)
        if dt == oh.OMX_C128:
            g = r.standard_normal(shape) + 1j * r.standard_normal(shape)
            gp = r.standard_normal(shape) + 1j * r.standard_normal(shape)
        else:
            g, gp = r.standard_normal(shape), r.standard_normal(shape)
        G, Gp = oh.DMat(ctx, n, m, dt), oh.DMat(ctx, n, m, dt)
        G.upload(g)
        Gp.upload(gp)
        a = _dots(oh.lib().omx_omm_pr_dots, ctx, G, Gp)
        b = _dots(oh.lib().omx_omm_pr_dots_noalloc, ctx, G, Gp)
        assert np.array_equal(a, b), (a, b)
        want = np.array([np.real(np.sum(np.conj(g) * (g - gp))), np.real(np.sum(np.conj(gp) * gp))])
        assert np.allclose(b, want, rtol=1e-12, atol=0)
        G.free()
        Gp.free()


def test_pr_dots_noalloc_rejects_bad_arguments(ctx):
    G = oh.DMat(ctx, 8, 4)
    out = np.zeros(2)
    rc = oh.lib().omx_omm_pr_dots_noalloc(ctx.h, G.dt, 8, 4, G.p, G.p, 4, out.ctypes.data_as(ctypes.c_void_p))
    assert rc != 0 and b"bad arguments" in oh.lib().omx_last_error()
    G.free()
