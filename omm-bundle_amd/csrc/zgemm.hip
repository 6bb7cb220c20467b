// Complex dense products on the CDNA4 matrix cores: C := alpha*op(A)*op(B) + beta*C for
// complex128 (FP64 MFMA) and complex64 (FP32 MFMA), op = 'n', 't' or 'c' (conjugate transpose),
// real alpha and beta (the only scalars the OMM loop uses).  These are MatrixSwitch's
// mm_zmultiply products (MatrixSwitch/src/MatrixSwitch.F90:1395-1726, 'ref' semantics of
// MatrixSwitch_mm_multiply.F90:85-185) for the k-point solves.
//
// Storage is the Fortran complex layout: (re, im) interleaved, column-major.  A workgroup of 4
// wave64s owns a BM x BN tile of C; operand slices are staged through LDS as separate real and
// imaginary k-major planes (conjugation is applied while staging, by negating the imaginary
// plane), and each 16x16 output block keeps a real and an imaginary accumulator:
//     acc_re += a_re*b_re - a_im*b_im,    acc_im += a_re*b_im + a_im*b_re,
// four real MFMAs per block and k step.  K slices are double-buffered; split-K with an
// in-order reduction covers shapes with too few tiles; the pair mode (two contractions back to
// back over one K range) gives calc_G in one launch.  Tile ids are XCD-remapped as in the real
// kernel.
#include <algorithm>

#include "omx_internal.h"

namespace omx {

typedef double zd4v __attribute__((ext_vector_type(4)));
typedef float zf4v __attribute__((ext_vector_type(4)));

template <typename R>
struct ZPrec;
template <>
struct ZPrec<double> {
    typedef zd4v accv;
    typedef double2 cplx;
    __device__ static inline accv mfma(double a, double b, accv c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // result lane l, register v -> row (l/16) + 4v
    __device__ static inline int out_row(int lq, int v) { return lq + 4 * v; }
};
template <>
struct ZPrec<float> {
    typedef zf4v accv;
    typedef float2 cplx;
    __device__ static inline accv mfma(float a, float b, accv c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    // result lane l, register v -> row 4(l/16) + v
    __device__ static inline int out_row(int lq, int v) { return 4 * lq + v; }
};

constexpr int ZBM = 64, ZBN = 128, ZBK = 8, ZNT = 256, ZPAD = 4;

// Batched products: blockIdx.y selects the system; operand and result offsets are the system index
// times these strides (complex elements); systems with active[s] == 0 are skipped.
struct ZBatch {
    int64_t sA = 0, sB = 0, sC = 0;
    const int *active = nullptr;
};
constexpr int ZWM = ZBM / 2, ZWN = ZBN / 2, ZFBM = ZWM / 16, ZFBN = ZWN / 16;
static_assert(ZBN == 2 * ZBM, "the lower-triangle tile count assumes tiles twice as wide as tall");

// Tiles of an n x n product that hold lower-triangle elements (row >= column): tile row tm (ZBM
// rows) reaches column ZBM*tm + ZBM-1, so it holds tm/2 + 1 tiles of ZBN columns, and tile rows
// 2a and 2a+1 hold a+1 each.  Ids are dealt over the XCDs as tile_of does; within a band the
// tiles go by rows.
static inline int lower_tiles_z(int tiles_m) {
    int t = 0;
    for (int r = 0; r < tiles_m; ++r) t += r / 2 + 1;
    return t;
}
__device__ inline void tile_of_lower_z(int w, int T, int &tm, int &tn) {
    int t = w;
    if ((T & 7) == 0) t = (w & 7) * (T >> 3) + (w >> 3);
    int a = (int)((sqrtf(4.0f * (float)t + 1.0f) - 1.0f) * 0.5f);
    while (a > 0 && a * (a + 1) > t) --a;
    while ((a + 1) * (a + 2) <= t) ++a;
    const int r = t - a * (a + 1);
    tm = 2 * a + r / (a + 1);
    tn = r % (a + 1);
}

// One BK slice of op(A) (BM x BK, "rows" r = i) or op(B)^T (BN x BK, r = j) into registers.
// CONTIG_R: r is the contiguous index in memory (A 'n', B 't'/'c'); otherwise k is.
template <typename R, bool CONTIG_R, bool CONJ, int ROWS>
struct ZStage {
    typedef typename ZPrec<R>::cplx C2;
    static constexpr int NE = ROWS * ZBK / ZNT;
    R re[NE], im[NE];
    __device__ static inline void coord(int t, int q, int &r, int &k) {
        const int idx = t + ZNT * q;
        if constexpr (CONTIG_R) {
            r = idx % ROWS;
            k = idx / ROWS;
        } else {
            k = idx % ZBK;
            r = idx / ZBK;
        }
    }
    __device__ inline void load(const C2 *__restrict__ P, int64_t ld, int r0, int k0, int rlim, int klim, int t) {
#pragma unroll
        for (int q = 0; q < NE; ++q) {
            int r, k;
            coord(t, q, r, k);
            const int gr = r0 + r, gk = k0 + k;
            C2 v;
            if (gr < rlim && gk < klim) {
                v = CONTIG_R ? P[(int64_t)gr + (int64_t)gk * ld] : P[(int64_t)gk + (int64_t)gr * ld];
            } else {
                v.x = (R)0;
                v.y = (R)0;
            }
            re[q] = v.x;
            im[q] = CONJ ? -v.y : v.y;
        }
    }
    __device__ inline void store(R *lre, R *lim, int t) const {
#pragma unroll
        for (int q = 0; q < NE; ++q) {
            int r, k;
            coord(t, q, r, k);
            lre[k * (ROWS + ZPAD) + r] = re[q];
            lim[k * (ROWS + ZPAD) + r] = im[q];
        }
    }
};

template <typename R, bool TRA, bool CJA, bool TRB, bool CJB, int NPAIR>
__global__ __launch_bounds__(ZNT, 2) void k_zgemm(int m, int n, int k, double alpha,
                                               const typename ZPrec<R>::cplx *__restrict__ A1,
                                               const typename ZPrec<R>::cplx *__restrict__ A2, int64_t lda,
                                               const typename ZPrec<R>::cplx *__restrict__ B1,
                                               const typename ZPrec<R>::cplx *__restrict__ B2, int64_t ldb,
                                               double beta0, const typename ZPrec<R>::cplx *__restrict__ C0,
                                               double beta, typename ZPrec<R>::cplx *__restrict__ C, int64_t ldc,
                                               int k_per_split, R *__restrict__ partial, int sym_tiles,
                                               ZBatch bat) {
    typedef typename ZPrec<R>::accv AV;
    typedef typename ZPrec<R>::cplx C2;
    constexpr int LA = ZBK * (ZBM + ZPAD), LB = ZBK * (ZBN + ZPAD);
    constexpr int STAGE = 2 * (LA + LB);  // re and im planes of A and B
    __shared__ R smem[2 * STAGE];

    const int bsys = blockIdx.y;
    if (bat.active && !bat.active[bsys]) return;  // whole workgroup, before any barrier
    A1 += bsys * bat.sA;
    A2 += bsys * bat.sA;
    B1 += bsys * bat.sB;
    B2 += bsys * bat.sB;
    C += bsys * bat.sC;
    if (C0) C0 += bsys * bat.sC;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int tiles_m = (m + ZBM - 1) / ZBM, tiles_n = (n + ZBN - 1) / ZBN;
    int tm, tn;
    if (sym_tiles)
        tile_of_lower_z(blockIdx.x, sym_tiles, tm, tn);
    else
        tile_of(blockIdx.x, tiles_m, tiles_n, tm, tn);
    const int i0 = tm * ZBM, j0 = tn * ZBN;
    const int ktot = NPAIR * k;
    const int kb = blockIdx.z * k_per_split;
    const int ke = min(ktot, kb + k_per_split);

    AV are[ZFBM][ZFBN], aim[ZFBM][ZFBN];
#pragma unroll
    for (int a = 0; a < ZFBM; ++a)
#pragma unroll
        for (int b = 0; b < ZFBN; ++b) {
            are[a][b] = AV{0, 0, 0, 0};
            aim[a][b] = AV{0, 0, 0, 0};
        }

    ZStage<R, !TRA, CJA, ZBM> sa;
    ZStage<R, TRB, CJB, ZBN> sb;
    auto fetch = [&](int kk) {
        const C2 *Ap = A1, *Bp = B1;
        int kl = kk;
        if (NPAIR == 2 && kk >= k) {
            Ap = A2;
            Bp = B2;
            kl = kk - k;
        }
        const int klim = min(k, kl + (ke - kk));
        sa.load(Ap, lda, i0, kl, m, klim, t);
        sb.load(Bp, ldb, j0, kl, n, klim, t);
    };
    auto planes = [&](int buf, R *&Are, R *&Aim, R *&Bre, R *&Bim) {
        R *base = smem + buf * STAGE;
        Are = base;
        Aim = base + LA;
        Bre = base + 2 * LA;
        Bim = base + 2 * LA + LB;
    };

    const int li = lane & 15, lk = lane >> 4;
    const int nslices = (ke - kb + ZBK - 1) / ZBK;
    if (nslices > 0) {
        R *Are, *Aim, *Bre, *Bim;
        fetch(kb);
        planes(0, Are, Aim, Bre, Bim);
        sa.store(Are, Aim, t);
        sb.store(Bre, Bim, t);
        __syncthreads();
        for (int s = 0; s < nslices; ++s) {
            const bool more = (s + 1 < nslices);
            if (more) fetch(kb + (s + 1) * ZBK);
            planes(s & 1, Are, Aim, Bre, Bim);
#pragma unroll
            for (int ks = 0; ks < ZBK; ks += 4) {
                R xr[ZFBM], xi[ZFBM], nxi[ZFBM], yr[ZFBN], yi[ZFBN];
#pragma unroll
                for (int f = 0; f < ZFBM; ++f) {
                    const int o = (ks + lk) * (ZBM + ZPAD) + wm * ZWM + f * 16 + li;
                    xr[f] = Are[o];
                    xi[f] = Aim[o];
                    nxi[f] = -xi[f];
                }
#pragma unroll
                for (int f = 0; f < ZFBN; ++f) {
                    const int o = (ks + lk) * (ZBN + ZPAD) + wn * ZWN + f * 16 + li;
                    yr[f] = Bre[o];
                    yi[f] = Bim[o];
                }
#pragma unroll
                for (int fm = 0; fm < ZFBM; ++fm)
#pragma unroll
                    for (int fn = 0; fn < ZFBN; ++fn) {
                        are[fm][fn] = ZPrec<R>::mfma(xr[fm], yr[fn], are[fm][fn]);
                        are[fm][fn] = ZPrec<R>::mfma(nxi[fm], yi[fn], are[fm][fn]);
                        aim[fm][fn] = ZPrec<R>::mfma(xr[fm], yi[fn], aim[fm][fn]);
                        aim[fm][fn] = ZPrec<R>::mfma(xi[fm], yr[fn], aim[fm][fn]);
                    }
            }
            if (more) {
                planes((s + 1) & 1, Are, Aim, Bre, Bim);
                sa.store(Are, Aim, t);
                sb.store(Bre, Bim, t);
            }
            __syncthreads();
        }
    }

    const int lq = lane >> 4;
#pragma unroll
    for (int fm = 0; fm < ZFBM; ++fm)
#pragma unroll
        for (int fn = 0; fn < ZFBN; ++fn) {
            const int j = j0 + wn * ZWN + fn * 16 + li;
            if (j >= n) continue;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = i0 + wm * ZWM + fm * 16 + ZPrec<R>::out_row(lq, v);
                if (i >= m) continue;
                const R vr = are[fm][fn][v], vi = aim[fm][fn][v];
                if (partial) {
                    const int64_t off = 2 * ((int64_t)blockIdx.z * m * n + (int64_t)i + (int64_t)j * m);
                    partial[off] = vr;
                    partial[off + 1] = vi;
                } else {
                    const int64_t off = (int64_t)i + (int64_t)j * ldc;
                    double rr = alpha * (double)vr, ri = alpha * (double)vi;
                    if (C0) {
                        rr += beta0 * (double)C0[off].x;
                        ri += beta0 * (double)C0[off].y;
                    }
                    if (beta != 0.0) {
                        rr += beta * (double)C[off].x;
                        ri += beta * (double)C[off].y;
                    }
                    C2 o;
                    o.x = (R)rr;
                    o.y = (R)ri;
                    C[off] = o;
                }
            }
        }
}

template <typename R>
__global__ void k_zsplitk_reduce(int m, int n, int splits, const R *__restrict__ partial, double alpha, double beta0,
                                 const typename ZPrec<R>::cplx *__restrict__ C0, double beta,
                                 typename ZPrec<R>::cplx *__restrict__ C, int64_t ldc) {
    const int64_t total = (int64_t)m * n;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = q % m, j = q / m;
        R sr = partial[2 * q], si = partial[2 * q + 1];
        for (int z = 1; z < splits; ++z) {
            sr += partial[2 * ((int64_t)z * total + q)];
            si += partial[2 * ((int64_t)z * total + q) + 1];
        }
        const int64_t off = i + j * ldc;
        double rr = alpha * (double)sr, ri = alpha * (double)si;
        if (C0) {
            rr += beta0 * (double)C0[off].x;
            ri += beta0 * (double)C0[off].y;
        }
        if (beta != 0.0) {
            rr += beta * (double)C[off].x;
            ri += beta * (double)C[off].y;
        }
        typename ZPrec<R>::cplx o;
        o.x = (R)rr;
        o.y = (R)ri;
        C[off] = o;
    }
}

template <typename R, bool TRA, bool CJA, bool TRB, bool CJB, int NPAIR>
static int run_zgemm(omx_ctx ctx, int m, int n, int k, double alpha, const void *A1, const void *A2, int64_t lda,
                     const void *B1, const void *B2, int64_t ldb, double beta0, const void *C0, double beta, void *C,
                     int64_t ldc, int nbatch = 1, ZBatch bat = ZBatch(), bool allow_split = true) {
    typedef typename ZPrec<R>::cplx C2;
    const int full_tiles = ((m + ZBM - 1) / ZBM) * ((n + ZBN - 1) / ZBN);
    const int ktot = NPAIR * k;
    int splits = 1;
    const int target = 2 * ctx->num_cus;
    // products known to be Hermitian (A B^H, ctx->gemm_sym): lower-triangle tiles only.  Split along
    // K when the tiles fill less than two workgroups per CU; the reduction also writes the upper
    // elements no tile computed, which the caller's mirror then overwrites.
    const int sym_tiles =
        (ctx->gemm_sym && NPAIR == 1 && m == n && !TRA && TRB && CJB) ? lower_tiles_z((m + ZBM - 1) / ZBM) : 0;
    const int tiles = sym_tiles ? sym_tiles : full_tiles;
    // batched launches never split K: no workspace, so the launch can be graph-captured
    if (allow_split && nbatch == 1 && tiles < target) {
        const int max_splits = std::max(1, ktot / (4 * ZBK));
        splits = std::max(1, std::min(std::min(max_splits, (target + tiles - 1) / tiles), 64));
        if (sym_tiles) {
            // the fewest rounds of workgroups per unit of K, up to 4 splits (272 tiles: 3 splits,
            // two rounds of thirds, not 2 splits and a second round of 32 workgroups)
            auto cost = [&](int q) { return (double)((tiles * q + target - 1) / target) / q; };
            splits = 1;
            for (int q = 2; q <= std::min(max_splits, 4); ++q)
                if (cost(q) < cost(splits)) splits = q;
        }
    }
    int kps = (ktot + splits - 1) / splits;
    kps = (kps + ZBK - 1) / ZBK * ZBK;
    splits = (ktot + kps - 1) / kps;
    R *partial = nullptr;
    if (splits > 1) {
        int rc = ensure_ws(ctx, sizeof(R) * 2 * (size_t)splits * m * n);
        if (rc) return rc;
        partial = (R *)ctx->ws;
    }
    hipEvent_t ev;
    prof_begin(ctx, NPAIR == 2 ? OMX_K_GEMM_PAIR : OMX_K_GEMM, &ev);
    prof_work(ctx, NPAIR == 2 ? OMX_K_GEMM_PAIR : OMX_K_GEMM,
              nbatch * 8.0 * m * n * (double)ktot * ((double)tiles / full_tiles),
              nbatch * 2.0 * sizeof(R) * ((double)NPAIR * ((double)m * k + (double)k * n) +
                                          (double)m * n * (1 + (C0 ? 1 : 0) + (beta != 0.0 ? 1 : 0))));
    hipLaunchKernelGGL((k_zgemm<R, TRA, CJA, TRB, CJB, NPAIR>), dim3(tiles, nbatch, splits), dim3(ZNT), 0, ctx->stream,
                       m, n, k, alpha, (const C2 *)A1, (const C2 *)A2, lda, (const C2 *)B1, (const C2 *)B2, ldb, beta0,
                       (const C2 *)C0, beta, (C2 *)C, ldc, kps, partial, sym_tiles, bat);
    if (splits > 1) {
        const int64_t total = (int64_t)m * n;
        const int grid = (int)std::min<int64_t>((total + 255) / 256, 8192);
        hipLaunchKernelGGL((k_zsplitk_reduce<R>), dim3(grid), dim3(256), 0, ctx->stream, m, n, splits, partial, alpha,
                           beta0, (const C2 *)C0, beta, (C2 *)C, ldc);
    }
    prof_end(ctx, NPAIR == 2 ? OMX_K_GEMM_PAIR : OMX_K_GEMM, ev);
    OMX_CHECK_HIP(hipGetLastError());
    return OMX_OK;
}

// op -> (transposed, conjugated) template arguments.
template <typename R, bool TRB, bool CJB>
static int zdispatch_a(omx_ctx ctx, char opA, int m, int n, int k, double alpha, const void *A, int64_t lda,
                       const void *B, int64_t ldb, double beta, void *C, int64_t ldc) {
    if (opA == 'n' || opA == 'N')
        return run_zgemm<R, false, false, TRB, CJB, 1>(ctx, m, n, k, alpha, A, A, lda, B, B, ldb, 0.0, nullptr, beta,
                                                       C, ldc);
    if (opA == 't' || opA == 'T')
        return run_zgemm<R, true, false, TRB, CJB, 1>(ctx, m, n, k, alpha, A, A, lda, B, B, ldb, 0.0, nullptr, beta, C,
                                                      ldc);
    return run_zgemm<R, true, true, TRB, CJB, 1>(ctx, m, n, k, alpha, A, A, lda, B, B, ldb, 0.0, nullptr, beta, C,
                                                 ldc);
}

template <typename R>
static int zdispatch(omx_ctx ctx, char opA, char opB, int m, int n, int k, double alpha, const void *A, int64_t lda,
                     const void *B, int64_t ldb, double beta, void *C, int64_t ldc) {
    if (opB == 'n' || opB == 'N')
        return zdispatch_a<R, false, false>(ctx, opA, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    if (opB == 't' || opB == 'T')
        return zdispatch_a<R, true, false>(ctx, opA, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    return zdispatch_a<R, true, true>(ctx, opA, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

int launch_zgemm(omx_ctx ctx, int dt, char opA, char opB, int m, int n, int k, double alpha, const void *A,
                 int64_t lda, const void *B, int64_t ldb, double beta, void *C, int64_t ldc) {
    if (dt == OMX_C128) return zdispatch<double>(ctx, opA, opB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    if (dt == OMX_C64) return zdispatch<float>(ctx, opA, opB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    set_error("complex product: unsupported element type %d", dt);
    return OMX_ERR_UNSUPPORTED;
}

// Batched complex products (complex128 or complex64) over nbatch systems with element strides sA, sB,
// sC (result and C0).
template <typename R>
static int zgemm_batched_r(omx_ctx ctx, char opA, char opB, int m, int n, int k, double alpha, const void *A,
                           int64_t lda, const void *B, int64_t ldb, double beta, void *C, int64_t ldc, int nbatch,
                           const ZBatch &bat) {
    const bool ta = (opA != 'n' && opA != 'N');
    const bool tb = (opB != 'n' && opB != 'N'), cb = (opB == 'c' || opB == 'C');
    if (!ta && tb && cb)  // the Gram products X Y^H
        return run_zgemm<R, false, false, true, true, 1>(ctx, m, n, k, alpha, A, A, lda, B, B, ldb, 0.0, nullptr,
                                                         beta, C, ldc, nbatch, bat, false);
    if (!ta && !tb)
        return run_zgemm<R, false, false, false, false, 1>(ctx, m, n, k, alpha, A, A, lda, B, B, ldb, 0.0, nullptr,
                                                           beta, C, ldc, nbatch, bat, false);
    set_error("batched complex product: op (%c,%c) not available", opA, opB);
    return OMX_ERR_UNSUPPORTED;
}

int launch_zgemm_batched(omx_ctx ctx, int dt, char opA, char opB, int m, int n, int k, double alpha, const void *A,
                         int64_t lda, int64_t sA, const void *B, int64_t ldb, int64_t sB, double beta, void *C,
                         int64_t ldc, int64_t sC, int nbatch, const int *active) {
    ZBatch bat;
    bat.sA = sA;
    bat.sB = sB;
    bat.sC = sC;
    bat.active = active;
    if (dt == OMX_C128)
        return zgemm_batched_r<double>(ctx, opA, opB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, nbatch, bat);
    if (dt == OMX_C64)
        return zgemm_batched_r<float>(ctx, opA, opB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, nbatch, bat);
    set_error("batched complex product: unsupported element type %d", dt);
    return OMX_ERR_UNSUPPORTED;
}

// C(i, j) := conj(C(j, i)) for i < j in each system's n x n block: copies the lower triangle of a
// Hermitian product onto the upper one.  One 32 x 32 tile pair per workgroup (the diagonal tiles
// on themselves), through an LDS transpose so that reads and writes run along columns.
template <typename R>
__global__ __launch_bounds__(256) void k_herm_mirror(int n, typename ZPrec<R>::cplx *__restrict__ C, int64_t ldc,
                                                     ZBatch bat) {
    typedef typename ZPrec<R>::cplx C2;
    constexpr int T = 32;
    __shared__ C2 tile[T][T + 1];
    const int bsys = blockIdx.y;
    if (bat.active && !bat.active[bsys]) return;
    C += bsys * bat.sC;
    int a, b;  // lower tile (a, b), a >= b; it fills the upper tile (b, a)
    tile_of_lower(blockIdx.x, (n + T - 1) / T, a, b);
    const int r0 = a * T, c0 = b * T;
    for (int q = threadIdx.x; q < T * T; q += 256) {
        const int r = q % T, c = q / T;
        if (r0 + r < n && c0 + c < n) tile[c][r] = C[(int64_t)(r0 + r) + (int64_t)(c0 + c) * ldc];
    }
    __syncthreads();
    for (int q = threadIdx.x; q < T * T; q += 256) {
        const int i = q % T, j = q / T;  // upper element (c0 + i, r0 + j)
        const int gi = c0 + i, gj = r0 + j;
        if (gi < gj && gj < n) {
            C2 v = tile[i][j];
            v.y = -v.y;
            C[(int64_t)gi + (int64_t)gj * ldc] = v;
        }
    }
}

// C := alpha * A * B^H for a product known to be Hermitian (n x n; A, B n x k), as AW = (C*H)*C^H in
// omm()'s calc_AW (omm_ops.F90:44-64) on complex storage.  The kernel computes the lower-triangle
// tiles where it can, and the lower triangle is then mirrored, so C is exactly Hermitian.
template <typename R>
static int zgemm_herm_r(omx_ctx ctx, int n, int k, double alpha, const void *A, int64_t lda, const void *B,
                        int64_t ldb, void *C, int64_t ldc, int nbatch, const ZBatch &bat, bool allow_split) {
    ctx->gemm_sym = true;
    int rc = run_zgemm<R, false, false, true, true, 1>(ctx, n, n, k, alpha, A, A, lda, B, B, ldb, 0.0, nullptr, 0.0,
                                                       C, ldc, nbatch, bat, allow_split);
    ctx->gemm_sym = false;
    if (rc) return rc;
    const int nt = (n + 31) / 32;
    hipLaunchKernelGGL(k_herm_mirror<R>, dim3(nt * (nt + 1) / 2, nbatch), dim3(256), 0, ctx->stream, n,
                       (typename ZPrec<R>::cplx *)C, ldc, bat);
    OMX_CHECK_HIP(hipGetLastError());
    return OMX_OK;
}

int launch_zgemm_herm(omx_ctx ctx, int dt, int n, int k, double alpha, const void *A, int64_t lda, const void *B,
                      int64_t ldb, void *C, int64_t ldc) {
    if (dt == OMX_C128) return zgemm_herm_r<double>(ctx, n, k, alpha, A, lda, B, ldb, C, ldc, 1, ZBatch(), true);
    if (dt == OMX_C64) return zgemm_herm_r<float>(ctx, n, k, alpha, A, lda, B, ldb, C, ldc, 1, ZBatch(), true);
    set_error("Hermitian product: unsupported element type %d", dt);
    return OMX_ERR_UNSUPPORTED;
}

int launch_zgemm_herm_batched(omx_ctx ctx, int dt, int n, int k, double alpha, const void *A, int64_t lda,
                              int64_t sA, const void *B, int64_t ldb, int64_t sB, void *C, int64_t ldc, int64_t sC,
                              int nbatch, const int *active) {
    ZBatch bat;
    bat.sA = sA;
    bat.sB = sB;
    bat.sC = sC;
    bat.active = active;
    if (dt == OMX_C128) return zgemm_herm_r<double>(ctx, n, k, alpha, A, lda, B, ldb, C, ldc, nbatch, bat, false);
    if (dt == OMX_C64) return zgemm_herm_r<float>(ctx, n, k, alpha, A, lda, B, ldb, C, ldc, nbatch, bat, false);
    set_error("batched Hermitian product: unsupported element type %d", dt);
    return OMX_ERR_UNSUPPORTED;
}

int launch_zgemm_pair_batched(omx_ctx ctx, int dt, int m, int n, int k, double alpha, const void *A1, const void *A2,
                              int64_t lda, int64_t sA, const void *B1, const void *B2, int64_t ldb, int64_t sB,
                              double beta0, const void *C0, double beta, void *C, int64_t ldc, int64_t sC, int nbatch,
                              const int *active) {
    if (k % ZBK != 0) return OMX_ERR_UNSUPPORTED;
    ZBatch bat;
    bat.sA = sA;
    bat.sB = sB;
    bat.sC = sC;
    bat.active = active;
    if (dt == OMX_C64)
        return run_zgemm<float, false, false, false, false, 2>(ctx, m, n, k, alpha, A1, A2, lda, B1, B2, ldb, beta0,
                                                               C0, beta, C, ldc, nbatch, bat, false);
    return run_zgemm<double, false, false, false, false, 2>(ctx, m, n, k, alpha, A1, A2, lda, B1, B2, ldb, beta0, C0,
                                                            beta, C, ldc, nbatch, bat, false);
}

int launch_zgemm_pair(omx_ctx ctx, int dt, int m, int n, int k, double alpha, const void *A1, const void *A2,
                      int64_t lda, const void *B1, const void *B2, int64_t ldb, double beta0, const void *C0,
                      double beta, void *C, int64_t ldc) {
    if (k % ZBK != 0) return OMX_ERR_UNSUPPORTED;  // the pair boundary must fall on a slice boundary
    if (dt == OMX_C128)
        return run_zgemm<double, false, false, false, false, 2>(ctx, m, n, k, alpha, A1, A2, lda, B1, B2, ldb, beta0,
                                                                C0, beta, C, ldc);
    if (dt == OMX_C64)
        return run_zgemm<float, false, false, false, false, 2>(ctx, m, n, k, alpha, A1, A2, lda, B1, B2, ldb, beta0,
                                                               C0, beta, C, ldc);
    set_error("complex calc_G: unsupported element type %d", dt);
    return OMX_ERR_UNSUPPORTED;
}

}  // namespace omx
