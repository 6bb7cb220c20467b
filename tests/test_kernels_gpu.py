"""HIP kernels of libommhip.so against numpy FP64 (and rounded-operand) references.

Each test prints the case it runs before the GPU call so that a failing run names it.
Tolerances: FP64 paths 1e-12 relative to the operand scale (FMA and summation order differ from
numpy); FP32 paths 1e-5; BF16 paths 2e-2 (8-bit mantissa operands, FP32 accumulation).
"""
import ctypes as C

import numpy as np
import pytest

import ommhip as oh

pytestmark = pytest.mark.gpu

rng = np.random.default_rng(7)


def _note(*a):
    print("CASE", *a, flush=True)


def _bf16(x):
    """Round-to-nearest-even of float64 values to bfloat16 in one step (the kernels convert the
    stored FP64 value directly; rounding through FP32 first differs near ties), as float64."""
    u = np.ascontiguousarray(x, np.float64).view(np.uint64)
    u = (u + np.uint64((1 << 44) - 1) + ((u >> np.uint64(45)) & np.uint64(1))) & ~np.uint64((1 << 45) - 1)
    return u.view(np.float64)


OPS = [("n", "n"), ("n", "t"), ("t", "n"), ("t", "t")]


@pytest.mark.parametrize("opA,opB", OPS)
@pytest.mark.parametrize("shape", [(64, 48, 40), (200, 130, 300), (256, 256, 1024), (1, 5, 3)])
def test_gemm_f64(ctx, opA, opB, shape):
    m, n, k = shape
    _note("gemm_f64", opA, opB, shape)
    A = rng.standard_normal((k, m) if opA == "t" else (m, k))
    B = rng.standard_normal((n, k) if opB == "t" else (k, n))
    C0 = rng.standard_normal((m, n))
    dA, dB, dC = oh.DMat.from_host(ctx, A), oh.DMat.from_host(ctx, B), oh.DMat.from_host(ctx, C0)
    oh.gemm(ctx, opA, opB, dA, dB, dC, alpha=-2.0, beta=0.5)
    ref = -2.0 * ((A.T if opA == "t" else A) @ (B.T if opB == "t" else B)) + 0.5 * C0
    got = dC.download()
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.sqrt(k) * 4


@pytest.mark.parametrize("prec", [oh.PREC_F32, oh.PREC_BF16])
@pytest.mark.parametrize("opA,opB", [("n", "t"), ("n", "n"), ("t", "n")])
def test_gemm_low_precision(ctx, prec, opA, opB):
    m, n, k = 192, 160, 512
    _note("gemm_lowp", prec, opA, opB)
    A = rng.standard_normal((k, m) if opA == "t" else (m, k))
    B = rng.standard_normal((n, k) if opB == "t" else (k, n))
    dA, dB, dC = oh.DMat.from_host(ctx, A), oh.DMat.from_host(ctx, B), oh.DMat(ctx, m, n)
    oh.gemm(ctx, opA, opB, dA, dB, dC, prec=prec)
    rnd = (lambda x: x.astype(np.float32).astype(np.float64)) if prec == oh.PREC_F32 else _bf16
    ref = (rnd(A).T if opA == "t" else rnd(A)) @ (rnd(B).T if opB == "t" else rnd(B))
    got = dC.download()
    tol = 1e-5 if prec == oh.PREC_F32 else 1e-4
    assert np.max(np.abs(got - ref)) <= tol * np.sqrt(k)


def test_gemm_f32_storage(ctx):
    m, n, k = 130, 70, 256
    _note("gemm_f32_storage")
    A = rng.standard_normal((m, k)).astype(np.float32)
    B = rng.standard_normal((n, k)).astype(np.float32)
    dA = oh.DMat.from_host(ctx, A, oh.OMX_F32)
    dB = oh.DMat.from_host(ctx, B, oh.OMX_F32)
    dC = oh.DMat(ctx, m, n, oh.OMX_F32)
    oh.gemm(ctx, "n", "t", dA, dB, dC, prec=oh.PREC_F32)
    ref = A.astype(np.float64) @ B.astype(np.float64).T
    assert np.max(np.abs(dC.download() - ref)) <= 1e-4 * np.sqrt(k)


def _bf16_of_f32(x):
    """Round-to-nearest-even of float32 values to bfloat16, as float64."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + np.uint64(0x7FFF) + ((u >> np.uint64(16)) & np.uint64(1))) & np.uint64(0xFFFF0000)
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _split3(x):
    """The BF16x3 operand terms of float32 values: hi = bf16(x), lo = bf16(x - hi) (in FP32)."""
    hi = _bf16_of_f32(x)
    lo = _bf16_of_f32((np.asarray(x, np.float32) - hi.astype(np.float32)).astype(np.float32))
    return hi, lo


def _terms3(x):
    """The three BF16 terms of values in their storage precision (float32: exact split; float64:
    each difference exact in float64): hi = bf16(x), mid = bf16(x - hi), lo = bf16(x - hi - mid)."""
    if x.dtype == np.float32:
        hi = _bf16_of_f32(x)
        r1 = (x - hi.astype(np.float32)).astype(np.float32)
        mid = _bf16_of_f32(r1)
        lo = _bf16_of_f32((r1 - mid.astype(np.float32)).astype(np.float32))
    else:
        hi = _bf16(x)
        r1 = x - hi
        mid = _bf16(r1)
        lo = _bf16(r1 - mid)
    return hi, mid, lo


@pytest.mark.parametrize("opA,opB", OPS)
@pytest.mark.parametrize("prec", [oh.PREC_BF16, oh.PREC_BF16X3, oh.PREC_BF16X6])
@pytest.mark.parametrize("dt", [oh.OMX_F32, oh.OMX_F64])
@pytest.mark.parametrize("shape", [(300, 260, 512), (301, 259, 576)])
def test_gemm_bf16_panels(ctx, opA, opB, prec, dt, shape):
    # shapes the k-contiguous BF16 panels cover (m, n >= 256, k % 64 == 0), FP32 or FP64 storage:
    # BF16 against the product of the operands rounded once from the stored value; BF16x3 against
    # hi*hi + hi*lo + lo*hi and, much tighter than BF16, against the exact product.  The odd shape
    # gives odd leading dimensions, so the panel conversion takes its element-wise paths.
    m, n, k = shape
    _note("gemm_bf16_panels", dt, prec, opA, opB, shape)
    npdt = np.float32 if dt == oh.OMX_F32 else np.float64
    A = rng.standard_normal((k, m) if opA == "t" else (m, k)).astype(npdt)
    B = rng.standard_normal((n, k) if opB == "t" else (k, n)).astype(npdt)
    C0 = rng.standard_normal((m, n)).astype(npdt)
    dA, dB = oh.DMat.from_host(ctx, A, dt), oh.DMat.from_host(ctx, B, dt)
    dC = oh.DMat.from_host(ctx, C0, dt)
    oh.gemm(ctx, opA, opB, dA, dB, dC, alpha=0.5, beta=-1.0, prec=prec)
    got = dC.download().astype(np.float64)
    op = lambda X, o: X.T if o == "t" else X  # noqa: E731
    rnd = _bf16_of_f32 if dt == oh.OMX_F32 else _bf16
    exact = 0.5 * op(A.astype(np.float64), opA) @ op(B.astype(np.float64), opB) - C0
    if prec == oh.PREC_BF16:
        ref = 0.5 * op(rnd(A), opA) @ op(rnd(B), opB) - C0
        assert np.max(np.abs(got - ref)) <= 1e-5 * np.sqrt(k)
    elif prec == oh.PREC_BF16X6:
        # the six term products of order >= 2^-16; FP32-grade against the exact product
        (A1, A2, A3), (B1, B2, B3) = _terms3(A), _terms3(B)
        pairs = [(A1, B3), (A3, B1), (A2, B2), (A1, B2), (A2, B1), (A1, B1)]
        ref = 0.5 * sum(op(a, opA) @ op(b, opB) for a, b in pairs) - C0
        assert np.max(np.abs(got - ref)) <= 1e-5 * np.sqrt(k)
        assert np.max(np.abs(got - exact)) <= 1e-5 * np.sqrt(k)
    else:
        if dt == oh.OMX_F32:
            Ah, Al = _split3(A)
            Bh, Bl = _split3(B)
        else:
            Ah, Bh = _bf16(A), _bf16(B)
            Al, Bl = _bf16(A - Ah), _bf16(B - Bh)
        ref = 0.5 * (op(Ah, opA) @ op(Bh, opB) + op(Ah, opA) @ op(Bl, opB) + op(Al, opA) @ op(Bh, opB)) - C0
        assert np.max(np.abs(got - ref)) <= 1e-5 * np.sqrt(k)
        bf = 0.5 * op(rnd(A), opA) @ op(rnd(B), opB) - C0
        assert np.max(np.abs(got - exact)) < np.max(np.abs(bf - exact)) / 50


@pytest.mark.parametrize("prec", [oh.PREC_BF16, oh.PREC_BF16X3, oh.PREC_BF16X6])
@pytest.mark.parametrize("n,m", [(256, 512), (200, 300)])
def test_calc_G_bf16_panels(ctx, prec, n, m):
    # the gradient pair on FP32 storage: n % 64 == 0 runs on the panels, n = 200 falls back (BF16
    # staging path, or FP32 for BF16x3)
    _note("calc_G_bf16_panels", prec, n, m)
    HW, SW = (rng.standard_normal((n, n)).astype(np.float32) for _ in range(2))
    HC, SC = (rng.standard_normal((n, m)).astype(np.float32) for _ in range(2))
    d = [oh.DMat.from_host(ctx, M, oh.OMX_F32) for M in (HW, SW, HC, SC)]
    G = oh.DMat(ctx, n, m, oh.OMX_F32)
    oh.omm_calc_G(ctx, d[0], d[1], d[2], d[3], G, prec=prec)
    f = lambda X: X.astype(np.float64)  # noqa: E731
    exact = 4 * f(HC) - 2 * f(SW) @ f(HC) - 2 * f(HW) @ f(SC)
    got = G.download().astype(np.float64)
    if prec == oh.PREC_BF16:
        b = _bf16_of_f32
        ref = 4 * f(HC) - 2 * b(SW) @ b(HC) - 2 * b(HW) @ b(SC)
        assert np.max(np.abs(got - ref)) <= 1e-5 * np.sqrt(n) * 8
    elif prec == oh.PREC_BF16X6:
        assert np.max(np.abs(got - exact)) <= 2e-5 * np.sqrt(n)  # FP32-grade
    else:
        assert np.max(np.abs(got - exact)) <= 2e-4 * np.sqrt(n)


@pytest.mark.parametrize("L,no", [(3, 54), (4, 128), (4, 127), (5, 250)])
def test_spmm_lattice(ctx, L, no):
    _note("spmm", L, no)
    cp, ri, bv = oh.lattice_bsr(L, 0)
    nb = 4 * L ** 3
    H = oh.bsr_to_dense(nb, cp, ri, bv)
    assert np.allclose(H, H.T)
    X = rng.standard_normal((no, nb))
    Y0 = rng.standard_normal((no, nb))
    A = oh.Bsr(ctx, nb, cp, ri, bv)
    dX, dY = oh.DMat.from_host(ctx, X), oh.DMat.from_host(ctx, Y0)
    oh.spmm(ctx, dX, A, dY, alpha=1.5, beta=-0.5)
    ref = 1.5 * X @ H - 0.5 * Y0
    assert np.max(np.abs(dY.download() - ref)) < 1e-12 * 30


def test_spmm_from_csc_general(ctx):
    # non-multiple-of-4 basis and an irregular pattern, 1-based CSC as MatrixSwitch 'sdcsc'
    nb, no = 37, 20
    _note("spmm_csc_general", nb, no)
    A = rng.standard_normal((nb, nb)) * (rng.random((nb, nb)) < 0.2)
    cols = [np.nonzero(A[:, j])[0] for j in range(nb)]
    colptr = np.cumsum([0] + [len(c) for c in cols]).astype(np.int32)
    rows = np.concatenate(cols).astype(np.int32) + 1
    vals = np.concatenate([A[c, j] for j, c in enumerate(cols)])
    B = oh.Bsr.from_csc(ctx, nb, colptr, rows, vals, ptrbase=0, rowbase=1)
    X = rng.standard_normal((no, nb))
    dX, dY = oh.DMat.from_host(ctx, X), oh.DMat(ctx, no, nb)
    oh.spmm(ctx, dX, B, dY)
    assert np.max(np.abs(dY.download() - X @ A)) < 1e-12 * 30


def test_spmm_f32(ctx):
    L, no = 4, 256
    _note("spmm_f32")
    cp, ri, bv = oh.lattice_bsr(L, 1)
    nb = 4 * L ** 3
    S = oh.bsr_to_dense(nb, cp, ri, bv)
    X = rng.standard_normal((no, nb)).astype(np.float32)
    A = oh.Bsr(ctx, nb, cp, ri, bv)
    dX = oh.DMat.from_host(ctx, X, oh.OMX_F32)
    dY = oh.DMat(ctx, no, nb, oh.OMX_F32)
    oh.spmm(ctx, dX, A, dY)
    assert np.max(np.abs(dY.download() - X.astype(np.float64) @ S)) < 1e-5


@pytest.mark.parametrize("dt", [oh.OMX_F64, oh.OMX_F32])
@pytest.mark.parametrize("no", [1030, 2048])
def test_spmm_shapes_and_walks(ctx, dt, no):
    # every launch shape and block-column walk computes each output element with the same
    # ascending (block, r) sum, so all of them agree bit for bit
    L = 6
    _note("spmm_shapes", dt, no)
    cp, ri, bv = oh.lattice_bsr(L, 0)
    nb = 4 * L ** 3
    npdt = np.float64 if dt == oh.OMX_F64 else np.float32
    X = rng.standard_normal((no, nb)).astype(npdt)
    Y0 = rng.standard_normal((no, nb)).astype(npdt)
    A = oh.Bsr(ctx, nb, cp, ri, bv)
    dX = oh.DMat.from_host(ctx, X, dt)
    dY = oh.DMat.from_host(ctx, Y0, dt)
    ctx.spmm_shape(0)
    oh.spmm(ctx, dX, A, dY, alpha=0.75, beta=-1.5)
    ref = dY.download()
    H = oh.bsr_to_dense(nb, cp, ri, bv)
    tol = 1e-12 if dt == oh.OMX_F64 else 1e-4
    assert np.max(np.abs(ref - (0.75 * X.astype(np.float64) @ H - 1.5 * Y0))) < tol * 30
    walks = [None, oh.lattice_walk(L, 2, 3), rng.permutation(nb // 4).astype(np.int32)]
    try:
        for walk in walks:
            A.set_walk(walk)
            for mode in (0, 1):
                for wpg in (1, 2, 4):
                    for batch in (1, 4, 8):
                        for jper in (1, 3, 16):
                            ctx.spmm_shape(batch, wpg, jper, mode)
                            dY.upload(Y0)
                            oh.spmm(ctx, dX, A, dY, alpha=0.75, beta=-1.5)
                            got = dY.download()
                            assert np.array_equal(got, ref), (walk is None, mode, wpg, batch, jper)
    finally:
        ctx.spmm_shape(0)
        A.set_walk(None)


@pytest.mark.parametrize("dt", [oh.OMX_F64, oh.OMX_F32])
@pytest.mark.parametrize("no", [1030, 2048])
def test_spmm_lds_groups_bitwise(ctx, dt, no):
    # LDS-staged gathers over compact patches of block columns: each output element is the same
    # ascending (block, r) FMA sequence as the default kernel, so the results are bitwise equal
    L = 6
    _note("spmm_lds_groups", dt, no)
    cp, ri, bv = oh.lattice_bsr(L, 0)
    nb = 4 * L ** 3
    npdt = np.float64 if dt == oh.OMX_F64 else np.float32
    X = rng.standard_normal((no, nb)).astype(npdt)
    Y0 = rng.standard_normal((no, nb)).astype(npdt)
    A = oh.Bsr(ctx, nb, cp, ri, bv)
    dX, dY = oh.DMat.from_host(ctx, X, dt), oh.DMat.from_host(ctx, Y0, dt)
    oh.spmm(ctx, dX, A, dY, alpha=0.75, beta=-1.5)
    ref = dY.download()
    try:
        for (tx, ty, tz) in [(2, 2, 2), (3, 3, 1), (1, 1, 1), (2, 3, 1)]:
            A.set_walk(oh.lattice_walk3(L, tx, ty, tz), group=tx * ty * tz)
            dY.upload(Y0)
            oh.spmm(ctx, dX, A, dY, alpha=0.75, beta=-1.5)
            assert np.array_equal(dY.download(), ref), (tx, ty, tz)
        A.set_walk(None)  # plan cleared: the default kernel again
        dY.upload(Y0)
        oh.spmm(ctx, dX, A, dY, alpha=0.75, beta=-1.5)
        assert np.array_equal(dY.download(), ref)
    finally:
        A.set_walk(None)
    for M in (dX, dY, A):
        M.free()


def test_spmm_lds_groups_reject_large_unions(ctx):
    L = 6
    cp, ri, bv = oh.lattice_bsr(L, 0)
    A = oh.Bsr(ctx, 4 * L ** 3, cp, ri, bv)
    with pytest.raises(oh.OmxError, match="distinct blocks"):
        A.set_walk(np.arange(L ** 3, dtype=np.int32), group=36)  # 6x6 sites of one plane: 36 + 72 blocks
    A.free()


def test_spmm_walk_rejects_non_permutation(ctx):
    L = 3
    cp, ri, bv = oh.lattice_bsr(L, 0)
    A = oh.Bsr(ctx, 4 * L ** 3, cp, ri, bv)
    bad = np.arange(L ** 3, dtype=np.int32)
    bad[1] = 0
    with pytest.raises(oh.OmxError, match="permutation"):
        A.set_walk(bad)
    with pytest.raises(oh.OmxError, match="bad shape"):
        ctx.spmm_shape(3, 4, 1, 0)


def test_autotune_matches_default():
    # tuned shapes change only the schedule: SpMM results are bitwise those of the default shape,
    # GEMM results agree to rounding (tile configurations may split K differently)
    _note("autotune")
    L, no = 4, 256
    nb = 4 * L ** 3
    cp, ri, bv = oh.lattice_bsr(L, 0)
    X = rng.standard_normal((no, nb))
    with oh.Ctx(0) as c:
        A = oh.Bsr(c, nb, cp, ri, bv)
        dX, dY = oh.DMat.from_host(c, X), oh.DMat(c, no, nb)
        oh.spmm(c, dX, A, dY)
        ref = dY.download()
        G = oh.DMat(c, no, no)
        oh.gemm(c, "n", "t", dX, dX, G)
        gref = G.download()
        c.autotune(True)
        for _ in range(2):  # the first call tunes, the second reuses the choice
            oh.spmm(c, dX, A, dY)
            assert np.array_equal(dY.download(), ref)
            oh.gemm(c, "n", "t", dX, dX, G)
            assert np.max(np.abs(G.download() - gref)) < 1e-12 * np.max(np.abs(gref))
        rep = c.autotune_report()
        assert "spmm" in rep and "gemm" in rep, rep
        for M in (dX, dY, G, A):
            M.free()


@pytest.mark.parametrize("shape", [(4096, 2048, 272), (4096 + 64, 2048, 272), (4096, 2048, 270), (256, 256, 64)])
def test_gemm_f64_gram_direct_lds(shape):
    """FP64 Gram C = -2 A B^T + 0.5 C0 (op 'n','t'): the default configuration 7 loads operand rows
    straight into LDS where m, n % 128 == 0, k % 16 == 0 and the tiles fill the chip (the first
    shape); the others fall back to the staged kernel with the same tiles (configuration 5).
    Results are bitwise those of configuration 5 and within the FP64 bar of numpy."""
    m, n, k = shape
    _note("gemm_f64 gram direct-to-LDS", shape)
    A = rng.standard_normal((m, k))
    B = rng.standard_normal((n, k))
    C0 = rng.standard_normal((m, n))
    ref = -2.0 * (A @ B.T) + 0.5 * C0
    with oh.Ctx(0) as c:
        dA, dB = oh.DMat.from_host(c, A), oh.DMat.from_host(c, B)
        out = []
        for cfg in (-1, 5):
            oh.check(oh.lib().omx_gemm_set_tile(c.h, cfg), "gemm_set_tile")
            dC = oh.DMat.from_host(c, C0)
            oh.gemm(c, "n", "t", dA, dB, dC, alpha=-2.0, beta=0.5)
            out.append(dC.download())
            dC.free()
        oh.check(oh.lib().omx_gemm_set_tile(c.h, -1), "gemm_set_tile")
        dA.free()
        dB.free()
    assert np.array_equal(out[0], out[1])
    assert np.max(np.abs(out[0] - ref)) <= 1e-12 * np.sqrt(k) * 4


@pytest.mark.parametrize("shape", [(4096, 2048, 288), (4096 + 64, 2048, 288), (4096, 2048, 280), (256, 256, 64)])
def test_gemm_f32_gram_direct_lds(shape):
    """FP32 Gram C = -2 A B^T + 0.5 C0 (op 'n','t', FP32 products on FP32 storage): the default
    configuration 7 loads operand rows straight into LDS (two k-rows per wave load, chunks swizzled
    by k-row) where m, n % 128 == 0, k % 32 == 0 and the tiles fill the chip (the first shape); the
    others fall back to the staged kernel (configuration 0).  Results are bitwise those of
    configuration 0 and within the FP32 bar of numpy."""
    m, n, k = shape
    _note("gemm_f32 gram direct-to-LDS", shape)
    A = rng.standard_normal((m, k)).astype(np.float32)
    B = rng.standard_normal((n, k)).astype(np.float32)
    C0 = rng.standard_normal((m, n)).astype(np.float32)
    ref = -2.0 * (A.astype(np.float64) @ B.astype(np.float64).T) + 0.5 * C0
    with oh.Ctx(0) as c:
        dA, dB = oh.DMat.from_host(c, A, oh.OMX_F32), oh.DMat.from_host(c, B, oh.OMX_F32)
        out = []
        for cfg in (-1, 0):
            oh.check(oh.lib().omx_gemm_set_tile(c.h, cfg), "gemm_set_tile")
            dC = oh.DMat.from_host(c, C0, oh.OMX_F32)
            oh.gemm(c, "n", "t", dA, dB, dC, alpha=-2.0, beta=0.5, prec=oh.PREC_F32)
            out.append(dC.download())
            dC.free()
        oh.check(oh.lib().omx_gemm_set_tile(c.h, -1), "gemm_set_tile")
        dA.free()
        dB.free()
    assert np.array_equal(out[0], out[1])
    assert np.max(np.abs(out[0] - ref)) <= 1e-5 * np.sqrt(k) * 4


@pytest.mark.parametrize("dt,prec,no,nb", [(oh.OMX_F64, oh.PREC_F64, 4096, 1024),
                                           (oh.OMX_F64, oh.PREC_F64, 300, 512),
                                           (oh.OMX_F64, oh.PREC_F64, 4096 + 64, 1024),
                                           (oh.OMX_F32, oh.PREC_F32, 1024, 512),
                                           (oh.OMX_F32, oh.PREC_BF16, 1024, 512),
                                           (oh.OMX_F32, oh.PREC_BF16X3, 300, 512)])
def test_gemm_sym(dt, prec, no, nb):
    """omx_gemm_sym on AW = (C H) C^T with H symmetric: the upper triangle is the mirror of the lower
    one, which is bitwise the full product's (same kernel per element), except on the FP64
    direct-to-LDS path, whose tiles beyond the last whole round of workgroups are split along K
    (rounding-level differences, 1e-12 of the product's scale).  Shapes cover that path (4096: 528
    tiles, 16 of them split), the staged FP64 kernel, FP32 and the BF16 panels; the odd sizes take
    partial edge tiles."""
    _note("gemm_sym", dt, prec, no, nb)
    npdt = np.float64 if dt == oh.OMX_F64 else np.float32
    C = rng.standard_normal((no, nb)).astype(npdt)
    Hm = rng.standard_normal((nb, nb))
    Hm = (Hm + Hm.T).astype(npdt)
    AC = (C.astype(np.float64) @ Hm.astype(np.float64)).astype(npdt)
    with oh.Ctx(0) as c:
        dAC, dC = oh.DMat.from_host(c, AC, dt), oh.DMat.from_host(c, C, dt)
        full, sym = oh.DMat(c, no, no, dt), oh.DMat(c, no, no, dt)
        oh.gemm(c, "n", "t", dAC, dC, full, prec=prec)
        oh.gemm_sym(c, dAC, dC, sym, prec=prec)
        F_, S_ = full.download(), sym.download()
        for M in (dAC, dC, full, sym):
            M.free()
    il = np.tril_indices(no)
    if dt == oh.OMX_F64 and no % 128 == 0:
        assert np.max(np.abs(S_[il] - F_[il])) <= 1e-12 * np.max(np.abs(F_))
    else:
        assert np.array_equal(S_[il], F_[il])
    assert np.array_equal(S_, S_.T)


@pytest.mark.parametrize("no,nb", [(1024, 8192), (1024 + 64, 8192), (256, 512)])
def test_calc_G_f64_direct_lds(no, nb):
    """FP64 calc_G pair G = 4HC - 2 SW HC - 2 HW SC: configuration 7 (operands loaded straight into
    LDS, HC/SC through the swizzled column-major tile) where m, n % 128 == 0, k % 16 == 0 and the
    tiles fill the chip (the first shape), the staged kernel otherwise.  Bitwise the same as
    configuration 5; sampled columns within the FP64 bar of numpy."""
    _note("calc_G f64 direct-to-LDS", no, nb)
    HC, SC = rng.standard_normal((no, nb)), rng.standard_normal((no, nb))
    W1, W2 = rng.standard_normal((no, no)), rng.standard_normal((no, no))
    HW, SW = W1 + W1.T, W2 + W2.T
    cols = rng.choice(nb, 64, replace=False)
    ref = 4 * HC[:, cols] - 2 * SW @ HC[:, cols] - 2 * HW @ SC[:, cols]
    with oh.Ctx(0) as c:
        d = {k_: oh.DMat.from_host(c, v) for k_, v in (("HC", HC), ("SC", SC), ("HW", HW), ("SW", SW))}
        out = []
        for cfg in (-1, 5):
            oh.check(oh.lib().omx_gemm_set_tile(c.h, cfg), "gemm_set_tile")
            G = oh.DMat(c, no, nb)
            oh.omm_calc_G(c, d["HW"], d["SW"], d["HC"], d["SC"], G)
            out.append(G.download())
            G.free()
        oh.check(oh.lib().omx_gemm_set_tile(c.h, -1), "gemm_set_tile")
        for v in d.values():
            v.free()
    assert np.array_equal(out[0], out[1])
    assert np.max(np.abs(out[0][:, cols] - ref)) <= 1e-12 * np.sqrt(no) * 8 * np.max(np.abs(ref))


@pytest.mark.parametrize("op", ["n", "t"])
def test_add_scale_set(ctx, op):
    m, n = 70, 45
    _note("add", op)
    A = rng.standard_normal((n, m) if op == "t" else (m, n))
    C0 = rng.standard_normal((m, n))
    dA, dC = oh.DMat.from_host(ctx, A), oh.DMat.from_host(ctx, C0)
    oh.add(ctx, op, dA, dC, 0.7, -1.25)
    ref = -1.25 * C0 + 0.7 * (A.T if op == "t" else A)
    assert np.max(np.abs(dC.download() - ref)) < 1e-14 * 10
    oh.scale(ctx, dC, 0.25)
    assert np.max(np.abs(dC.download() - 0.25 * ref)) < 1e-14 * 10
    oh.mset(ctx, "a", dC, 0.0, 2.0)
    E = np.zeros((m, n))
    np.fill_diagonal(E, 2.0)
    assert np.array_equal(dC.download(), E)
    oh.mset(ctx, "u", dC, 3.0, 1.0)
    E2 = E.copy()
    for i in range(m):
        for j in range(i, n):
            E2[i, j] = 1.0 if i == j else 3.0
    assert np.array_equal(dC.download(), E2)


def test_traces_and_elements(ctx):
    n = 300
    _note("traces")
    A, B = rng.standard_normal((n, n)), rng.standard_normal((n, n))
    dA, dB = oh.DMat.from_host(ctx, A), oh.DMat.from_host(ctx, B)
    assert abs(oh.trace(ctx, dA) - np.trace(A)) < 1e-12
    assert abs(oh.mm_trace(ctx, dA, dB) - np.sum(A * B)) < 1e-10
    oh.set_element(ctx, dA, 3, 5, 2.0, 0.5)
    assert oh.get_element(ctx, dA, 3, 5) == 2.0 + 0.5 * A[2, 4]


def _coeff_np(HW, SW, HWd, SWd, HWdd, SWdd):
    TrH, TrHS = np.trace(HW), np.sum(HW * SW)
    TrHd, TrHdS, TrHSd = np.trace(HWd), np.sum(SW * HWd), np.sum(HW * SWd)
    TrHdd, TrHddS, TrHSdd = np.trace(HWdd), np.sum(HWdd * SW), np.sum(HW * SWdd)
    TrHdSd, TrHdSdH = np.sum(SWd.T * HWd), np.sum(SWd * HWd)
    TrHddSd, TrHdSdd, TrHddSdd = np.sum(HWdd * SWd), np.sum(SWdd * HWd), np.sum(HWdd * SWdd)
    return np.array([2 * TrH - TrHS, 2 * (2 * TrHd - TrHdS - TrHSd),
                     2 * (TrHdd - TrHdSd - TrHdSdH) - TrHddS - TrHSdd, -2 * (TrHddSd + TrHdSdd), -TrHddSdd])


@pytest.mark.parametrize("n", [31, 64, 257])
def test_omm_coeff(ctx, n):
    _note("coeff", n)
    mats = [rng.standard_normal((n, n)) for _ in range(6)]
    d = [oh.DMat.from_host(ctx, M) for M in mats]
    got = oh.omm_coeff(ctx, *d)
    ref = _coeff_np(*mats)
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-10 * n)


@pytest.mark.parametrize("n,m", [(64, 128), (96, 200), (33, 70)])
def test_calc_G(ctx, n, m):
    _note("calc_G", n, m)
    HW, SW = rng.standard_normal((n, n)), rng.standard_normal((n, n))
    HC, SC = rng.standard_normal((n, m)), rng.standard_normal((n, m))
    d = [oh.DMat.from_host(ctx, M) for M in (HW, SW, HC, SC)]
    G = oh.DMat(ctx, n, m)
    oh.omm_calc_G(ctx, d[0], d[1], d[2], d[3], G)
    ref = 4 * HC - 2 * SW @ HC - 2 * HW @ SC
    assert np.max(np.abs(G.download() - ref)) < 1e-12 * np.sqrt(n) * 10


def test_cg_updates(ctx):
    n, m = 50, 90
    _note("cg_updates")
    G, D = rng.standard_normal((n, m)), rng.standard_normal((n, m))
    dG, dD, dGp = oh.DMat.from_host(ctx, G), oh.DMat.from_host(ctx, D), oh.DMat(ctx, n, m)
    oh.omm_cg_direction(ctx, 0.3, dG, dD, dGp)
    assert np.allclose(dD.download(), G + 0.3 * D, rtol=0, atol=1e-15)
    assert np.array_equal(dGp.download(), G)
    G2 = rng.standard_normal((n, m))
    dG.upload(G2)
    num, den = oh.omm_pr_dots(ctx, dG, dGp)
    assert abs(num - np.sum(G2 * (G2 - G))) < 1e-10 and abs(den - np.sum(G * G)) < 1e-10
    W, Wd, Wdd = (rng.standard_normal((n, n)) for _ in range(3))
    dW, dWd, dWdd = (oh.DMat.from_host(ctx, M) for M in (W, Wd, Wdd))
    oh.omm_update_w(ctx, -0.37, dWd, dWdd, dW)
    ref = W - 0.37 * (Wd.T + Wd) + 0.37 ** 2 * Wdd
    assert np.max(np.abs(dW.download() - ref)) < 1e-14 * 20


def test_bsr_roundtrip(ctx):
    # values and pattern download return what was uploaded; update_values replaces the values
    L = 3
    _note("bsr_roundtrip", L)
    cp, ri, bv = oh.lattice_bsr(L, 0)
    A = oh.Bsr(ctx, 4 * L ** 3, cp, ri, bv)
    gcp, gri = A.pattern()
    assert np.array_equal(gcp, cp) and np.array_equal(gri, ri)
    assert np.array_equal(A.values().ravel(), np.asarray(bv).ravel())
    bv2 = rng.standard_normal(np.asarray(bv).shape)
    from ctypes import c_void_p
    oh.check(oh.lib().omx_bsr_update_values(A.h, bv2.ctypes.data_as(c_void_p)), "update_values")
    assert np.array_equal(A.values().ravel(), bv2.ravel())


def _pattern_mask(nb, cp, ri):
    M = np.zeros((nb, nb), bool)
    for jb in range(len(cp) - 1):
        for l in range(cp[jb], cp[jb + 1]):
            i0, j0 = 4 * ri[l], 4 * jb
            M[i0:i0 + 4, j0:j0 + 4] = True
    return M


@pytest.mark.parametrize("L,no,beta", [(3, 54, 0.0), (4, 127, 0.5), (5, 250, -1.0)])
def test_sddmm_lattice(ctx, L, no, beta):
    # Cp := alpha*(A^T B) + beta*Cp on the stored blocks of the lattice H pattern
    _note("sddmm", L, no, beta)
    cp, ri, bv = oh.lattice_bsr(L, 0)
    nb = 4 * L ** 3
    P = oh.Bsr(ctx, nb, cp, ri, bv)
    Hd = oh.bsr_to_dense(nb, cp, ri, bv)
    A, B = rng.standard_normal((no, nb)), rng.standard_normal((no, nb))
    dA, dB = oh.DMat.from_host(ctx, A), oh.DMat.from_host(ctx, B)
    oh.sddmm(ctx, dA, dB, P, alpha=2.0, beta=beta)
    mask = _pattern_mask(nb, cp, ri)
    ref = np.where(mask, 2.0 * (A.T @ B) + beta * Hd, 0.0)
    got = oh.bsr_to_dense(nb, cp, ri, P.values())
    assert np.max(np.abs(got - ref)) < 1e-12 * np.sqrt(no) * 10


def test_sddmm_general_pattern(ctx):
    # non-multiple-of-4 basis from 1-based CSC; padding rows/columns of edge blocks stay zero
    nb, no = 37, 21
    _note("sddmm_general", nb, no)
    Am = rng.standard_normal((nb, nb)) * (rng.random((nb, nb)) < 0.2)
    cols = [np.nonzero(Am[:, j])[0] for j in range(nb)]
    colptr = np.cumsum([0] + [len(c) for c in cols]).astype(np.int32)
    rows = np.concatenate(cols).astype(np.int32) + 1
    vals = np.concatenate([Am[c, j] for j, c in enumerate(cols)])
    P = oh.Bsr.from_csc(ctx, nb, colptr, rows, vals, ptrbase=0, rowbase=1)
    X, Y = rng.standard_normal((no, nb)), rng.standard_normal((no, nb))
    oh.sddmm(ctx, oh.DMat.from_host(ctx, X), oh.DMat.from_host(ctx, Y), P)
    gcp, gri = P.pattern()
    nbp = 4 * ((nb + 3) // 4)
    got = oh.bsr_to_dense(nbp, gcp, gri, P.values())
    mask = _pattern_mask(nbp, gcp, gri)[:nb, :nb]
    ref = np.where(mask, X.T @ Y, 0.0)
    assert np.max(np.abs(got[:nb, :nb] - ref)) < 1e-12 * 50
    assert np.all(got[nb:, :] == 0) and np.all(got[:, nb:] == 0)


def test_sddmm_f32(ctx):
    L, no = 4, 200
    _note("sddmm_f32", L, no)
    cp, ri, bv = oh.lattice_bsr(L, 1)
    nb = 4 * L ** 3
    P = oh.Bsr(ctx, nb, cp, ri, bv)
    A = rng.standard_normal((no, nb)).astype(np.float32)
    B = rng.standard_normal((no, nb)).astype(np.float32)
    oh.sddmm(ctx, oh.DMat.from_host(ctx, A, oh.OMX_F32), oh.DMat.from_host(ctx, B, oh.OMX_F32), P)
    ref = np.where(_pattern_mask(nb, cp, ri), A.astype(np.float64).T @ B.astype(np.float64), 0.0)
    got = oh.bsr_to_dense(nb, cp, ri, P.values())
    assert np.max(np.abs(got - ref)) < 1e-4


@pytest.mark.parametrize("dims,no,dt", [((16, 16, 16), 8192, oh.OMX_F64), ((8, 8, 8), 1024, oh.OMX_F64),
                                        ((4, 8, 3), 200, oh.OMX_F64), ((12, 4, 5), 37, oh.OMX_F64),
                                        ((5, 3, 7), 130, oh.OMX_F64), ((3, 9, 4), 64, oh.OMX_F64),
                                        ((16, 16, 16), 8192, oh.OMX_F32), ((8, 4, 6), 300, oh.OMX_F32),
                                        ((7, 5, 3), 258, oh.OMX_F32)])
def test_spmm_stencil_equals_general_kernel(ctx, dims, no, dt):
    """The 2.5-D kernel of lattice operators (z-planes in registers and LDS) against the general
    BSR kernel on the same operator and X: the same bits, including partial row chunks, Lz = 3,
    beta != 0 and edges that are not multiples of 4 (where the stencil kernel falls back to the
    general one); and the lattice edges recognised in the pattern."""
    _note("spmm stencil", dims, no, dt)
    cp, ri, bv = oh.lattice_bsr(dims, 0)
    nb = 4 * int(np.prod(dims))
    A = oh.Bsr(ctx, nb, cp, ri, bv)
    lat = np.zeros(3, np.int32)
    oh.check(oh.lib().omx_bsr_lattice(A.h, lat.ctypes.data_as(C.c_void_p)), "bsr_lattice")
    assert tuple(lat) == tuple(dims)
    X = rng.standard_normal((no, nb))
    Y0 = rng.standard_normal((no, nb))
    dX = oh.DMat.from_host(ctx, X, dt)
    out = {}
    try:
        for kern in (1, 0):
            ctx.spmm_stencil(kern)
            for beta in (-0.5, 0.0):
                dY = oh.DMat.from_host(ctx, Y0, dt)
                oh.spmm(ctx, dX, A, dY, alpha=1.5, beta=beta)
                out[kern, beta] = dY.download()
                dY.free()
    finally:
        ctx.spmm_stencil(0)
        dX.free()
        A.free()
    for beta in (-0.5, 0.0):
        assert np.array_equal(out[1, beta], out[0, beta])
    H = oh.bsr_to_dense(nb, cp, ri, bv)
    tol = 1e-12 if dt == oh.OMX_F64 else 1e-4
    assert np.max(np.abs(out[1, -0.5] - (1.5 * X @ H - 0.5 * Y0))) < tol * 30


def test_non_lattice_pattern_is_not_taken_for_a_lattice(ctx):
    # a 7-block pattern that is not the periodic stencil (one block row moved) stays on the general
    # kernel: no lattice recognised
    cp, ri, bv = oh.lattice_bsr((4, 4, 4), 0)
    ri = ri.copy()
    J = 5
    rows = ri[cp[J]:cp[J + 1]]
    free = [I for I in range(64) if I not in rows]
    rows[-1] = free[-1]
    ri[cp[J]:cp[J + 1]] = np.sort(rows)
    A = oh.Bsr(ctx, 256, cp, ri, bv)
    lat = np.zeros(3, np.int32)
    try:
        oh.check(oh.lib().omx_bsr_lattice(A.h, lat.ctypes.data_as(C.c_void_p)), "bsr_lattice")
    finally:
        A.free()
    assert tuple(lat) == (0, 0, 0)


@pytest.mark.parametrize("dt", [oh.OMX_F64, oh.OMX_F32])
def test_elementwise_past_one_grid_stride(ctx, dt):
    """The element-wise kernels walk (row, column) by their grid stride (up to 65536 x 256 threads,
    2048 x 256 for the Polak-Ribiere dots).  A 3000 x 6000 matrix with ld 3010 has 18M elements,
    more than one stride, and 3000 divides neither stride, so each thread's walk wraps columns at
    varying rows: add, scale, set (lower), dot, Polak-Ribiere dots and the CG direction against
    numpy."""
    m, n, ld = 3000, 6000, 3010
    _note("elementwise wrap", dt, m, n, ld)
    npdt = np.float64 if dt == oh.OMX_F64 else np.float32
    tol = 1e-13 if dt == oh.OMX_F64 else 1e-5
    A = rng.standard_normal((m, n)).astype(npdt)
    B = rng.standard_normal((m, n)).astype(npdt)
    Cm = rng.standard_normal((m, n)).astype(npdt)
    dA, dB = oh.DMat.from_host(ctx, A, dt, ld=ld), oh.DMat.from_host(ctx, B, dt, ld=ld)
    dC = oh.DMat.from_host(ctx, Cm, dt, ld=ld)
    try:
        oh.add(ctx, "n", dA, dC, 0.75, -1.25)
        ref = (-1.25 * Cm.astype(np.float64) + 0.75 * A.astype(np.float64))
        assert np.max(np.abs(dC.download() - ref)) <= tol * 4
        oh.scale(ctx, dC, -2.0)
        ref = -2.0 * ref.astype(npdt).astype(np.float64)
        assert np.max(np.abs(dC.download() - ref)) <= tol * 8
        before = dC.download()
        oh.mset(ctx, "l", dC, 0.5, 3.0)
        got = dC.download()
        il = np.tril_indices(m, 0, n)
        lower = np.zeros((m, n), bool)
        lower[il] = True
        assert np.all(got[lower & (np.arange(m)[:, None] != np.arange(n)[None, :])] == 0.5)
        assert np.all(np.diag(got) == 3.0)
        assert np.array_equal(got[~lower], before[~lower])
        dot = oh.mm_trace(ctx, dA, dB)
        assert abs(dot - np.sum(A.astype(np.float64) * B)) <= tol * 1e3 * np.sqrt(m * n)
        pr = oh.omm_pr_dots(ctx, dA, dB)
        W = (A - B).astype(npdt).astype(np.float64)
        assert abs(pr[0] - np.sum(A.astype(np.float64) * W)) <= tol * 1e3 * np.sqrt(m * n)
        assert abs(pr[1] - np.sum(B.astype(np.float64) ** 2)) <= tol * 1e3 * np.sqrt(m * n)
        dGp = oh.DMat(ctx, m, n, dt, ld=ld)
        oh.omm_cg_direction(ctx, 0.375, dA, dB, dGp)
        assert np.max(np.abs(dB.download() - (A.astype(np.float64) + 0.375 * B))) <= tol * 4
        assert np.array_equal(dGp.download(), A)
        dGp.free()
    finally:
        for M in (dA, dB, dC):
            M.free()
