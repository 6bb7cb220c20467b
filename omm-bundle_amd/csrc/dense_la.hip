// Dense factorisations for libOMM's Cholesky and preconditioned flavours (FP64, column-major):
//   omx_potrf      S = U^T U, upper (dpotrf 'u', libOMM/src/omm_ops.F90:301)
//   omx_trsm       B := alpha * op(U)^-1 * B  or  alpha * B * op(U)^-1, U upper (dtrsm)
//   omx_trmm       B := alpha * op(U) * B     or  alpha * B * op(U),     U upper (dtrmm)
//   omx_sygst      H := U^-T H U^-1, both triangles (dsygst itype 1 'u' + the mirror copy,
//                  omm_ops.F90:481-487)
//   omx_inverse_spd A := A^-1 for symmetric positive definite A (m_inverse, omm_ops.F90:1065-1075
//                  uses dsytrf/dsytri; on an SPD matrix the Cholesky route gives the same inverse)
//
// Everything is blocked by LB = 128.  A 128x128 diagonal block is factorised and inverted by one
// workgroup entirely in LDS (132 KB); every other operation is an MFMA GEMM of gemm.hip with K = 128:
//   * potrf runs the right-looking lower Cholesky of S^T (S's upper triangle, transposed into a
//     workspace padded with the identity to a multiple of 128): per block, L_kk and L_kk^-1 from the
//     diagonal kernel, the panel L21 = A21 L_kk^-T as a GEMM, and the trailing update
//     A22 -= L21 L21^T on the lower-triangle tiles only.  U = L^T is written back to the upper
//     triangle; the strict lower triangle of the input is never read nor written (as dpotrf 'u').
//   * the triangular solves apply the explicit inverses of U's diagonal blocks (all computed by one
//     launch, one workgroup per block), so each block step is two GEMMs.  Solves whose right-hand
//     side or result is known to be triangular restrict the GEMMs to the nonzero part.
//   * sygst = a full left solve, then a right solve that forms only the block-lower part of the
//     (symmetric) result, mirrored onto the upper triangle: 5/3 n^3 flops.
//   * the SPD inverse = potrf, X = U^-1 (upper-triangular right-hand side, n^3/3 flops), then
//     A^-1 = X X^T by row chunks that skip X's zero lower part (n^3/3 flops) and a mirror.
// Multiplying by the inverse of a diagonal block instead of substituting changes rounding at the
// level of cond(block) * eps; the diagonal blocks of an SPD factor are well conditioned when S is.
#include <algorithm>

#include "omx_internal.h"

namespace omx {

constexpr int LB = 128;       // diagonal block
constexpr int LB_OUT = 512;   // outer block: the updates beyond it run with K = 512
constexpr int LB_S = LB + 1;  // LDS column stride (odd: rows and columns both conflict-free)
constexpr int LB_NT = 1024;   // threads of the diagonal-block kernels (a 32 x 32 grid)
constexpr int LA_NT = 256;    // threads of the element-wise kernels
constexpr int TT = 64;        // transpose tile

// Register layout of the block inverse of the triangular solves (order DB = LB): thread (tx, ty) of a
// TG x TG grid owns the elements of rows tx + TG a and columns ty + TG b (a, b < DB / TG): four waves
// per SIMD, so the short dependent chains of one row step overlap across waves.  Row or column
// groups that a step cannot touch are skipped by uniform branches.
constexpr int TG = 32;
constexpr int DC = 64;  // potrf's diagonal block: one wave, one row per lane
static_assert(DC == 64, "k_chol_diag_wave maps one row to each lane of a 64-wide wave");

// X = L^-1 for the lower-triangular LB x LB matrix L in LDS (a, column-major, stride LB_S; rdiag[k] =
// 1 / L(k, k)), by right-looking forward substitution of L X = I in registers.  At row k the owners
// of row k finish it, X(k, j) = B(k, j) / L(k, k), and publish it to rowk[k & 1]; after one barrier
// every thread updates its rows below k, B(i, j) -= L(i, k) X(k, j).  A rowk buffer is rewritten two
// rows later, after every thread has passed the barrier that follows its last reads.  X is 0 above
// the diagonal.
template <int DB>
__device__ void reg_lower_inverse(const double *a, const double *rdiag, double (*rowk)[DB],
                                  double (&x)[DB / TG][DB / TG]) {
    constexpr int RG = DB / TG, DS = DB + 1;
    const int t = threadIdx.x, tx = t % TG, ty = t / TG;
#pragma unroll
    for (int ai = 0; ai < RG; ++ai)
#pragma unroll
        for (int bj = 0; bj < RG; ++bj) x[ai][bj] = (tx + TG * ai == ty + TG * bj) ? 1.0 : 0.0;
#pragma unroll 1
    for (int k = 0; k < DB; ++k) {
        const int kg = k / TG;
        double *rk = rowk[k & 1];
        if (tx == k % TG) {
            const double rl = rdiag[k];
#pragma unroll
            for (int ai = 0; ai < RG; ++ai) {
                if (ai != kg) continue;
#pragma unroll
                for (int bj = 0; bj < RG; ++bj) {
                    if (TG * bj > k) continue;
                    const int j = ty + TG * bj;
                    if (j <= k) {
                        x[ai][bj] *= rl;
                        rk[j] = x[ai][bj];
                    }
                }
            }
        }
        __syncthreads();
        // all reads first (LDS, always in bounds), then the updates with multipliers that are zero
        // outside the active range (x - 0*y = x exactly), so no per-element branches
        double lik[RG], xkj[RG];
#pragma unroll
        for (int q = 0; q < RG; ++q) {
            const int i = tx + TG * q, j = ty + TG * q;
            const double av = a[k * DS + i], rv = rk[j];
            lik[q] = (i > k) ? av : 0.0;
            xkj[q] = (j <= k) ? rv : 0.0;
        }
#pragma unroll
        for (int bj = 0; bj < RG; ++bj) {
            if (TG * bj > k) continue;  // column groups right of k are still zero (uniform)
#pragma unroll
            for (int ai = 0; ai < RG; ++ai) {
                if (TG * ai + TG - 1 <= k) continue;  // row groups at or above k are final (uniform)
                x[ai][bj] = fma(-lik[ai], xkj[bj], x[ai][bj]);
            }
        }
    }
}

// v broadcast from lane `lane` (a compile-time constant in the callers) through two scalar reads.
__device__ inline double bcast_lane(double v, int lane) {
    const long long b = __double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(b & 0xffffffffLL), lane);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// The DC x DC diagonal block of W at (k0, k0) by one wave, without barriers: lane i keeps row i of
// the block in registers.  Lower Cholesky, right-looking, with the pivot and L(j, k) broadcast from
// the lanes that hold them; then X = L^-1 by forward substitution of L X = I, row k of X broadcast
// from lane k, with L's columns read back from LDS.  L goes to W's lower triangle, X into LI (DC x DC,
// zeros above the diagonal); *info := global index + 1 of the first non-positive pivot below n (a
// non-positive pivot is replaced by 1 so the remaining blocks stay finite).  W is padded with the
// identity beyond n, so the block is always whole.
// Elements above the diagonal take harmless updates and are never written out.
__global__ __launch_bounds__(64) void k_chol_diag_wave(double *__restrict__ W, int64_t ldw, int k0, int n,
                                                       double *__restrict__ LI, int *__restrict__ info) {
    constexpr int DS = DC + 1;
    __shared__ double a[DC * DS];
    const int i = threadIdx.x;
    double *Wkk = W + k0 + (int64_t)k0 * ldw;
    double r[DC];
#pragma unroll
    for (int j = 0; j < DC; ++j) r[j] = (i >= j) ? Wkk[i + (int64_t)j * ldw] : 0.0;
    int bad = 0;
#pragma unroll
    for (int k = 0; k < DC; ++k) {
        const double d = bcast_lane(r[k], k);
        if (!(d > 0.0) && !bad && k0 + k < n) bad = k0 + k + 1;
        const double dk = sqrt(d > 0.0 ? d : 1.0), rdk = 1.0 / dk;
        const double li = r[k] * rdk;  // L(i, k) for i > k
        r[k] = (i == k) ? dk : (i > k ? li : r[k]);
#pragma unroll
        for (int j = k + 1; j < DC; ++j) r[j] = fma(-li, bcast_lane(li, j), r[j]);
    }
#pragma unroll
    for (int j = 0; j < DC; ++j) {
        if (i >= j) Wkk[i + (int64_t)j * ldw] = r[j];
        a[j * DS + i] = (i >= j) ? r[j] : 0.0;  // column j of L
    }
    __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): this wave's LDS writes are done
    __builtin_amdgcn_wave_barrier();
    double x[DC];
#pragma unroll
    for (int j = 0; j < DC; ++j) x[j] = (i == j) ? 1.0 : 0.0;
#pragma unroll
    for (int k = 0; k < DC; ++k) {
        const double lik = (i > k) ? a[k * DS + i] : 0.0;  // L(i, k) below the diagonal
        const double s = (i == k) ? 1.0 / a[k * DS + k] : 1.0;
#pragma unroll
        for (int j = 0; j <= k; ++j) {
            x[j] *= s;  // lane k finishes row k of X
            x[j] = fma(-lik, bcast_lane(x[j], k), x[j]);
        }
    }
#pragma unroll
    for (int j = 0; j < DC; ++j) LI[i + j * DC] = x[j];
    if (i == 0 && bad && *info == 0) *info = bad;
}

// Inverses of the diagonal blocks of the upper-triangular U (order nu): block b (rows and columns
// b*LB .. b*LB+nb) into UI + b*LB*LB as an LB x LB matrix, U_bb^-1 (upper, zeros below the diagonal)
// or, with LOWER_OUT, its transpose; a partial last block is padded with the identity.  One
// workgroup per block: U_bb^T is inverted as a lower matrix in LDS.
template <bool LOWER_OUT>
__global__ __launch_bounds__(LB_NT) void k_trinv_upper_blocks(int nu, const double *__restrict__ U, int64_t ldu,
                                                              double *__restrict__ UI) {
    __shared__ double a[LB * LB_S];
    __shared__ double rowk[2][LB];
    __shared__ double rdiag[LB];
    const int t = threadIdx.x, tx = t % TG, ty = t / TG, k0 = blockIdx.x * LB, nb = min(LB, nu - k0);
    const double *Ubb = U + k0 + (int64_t)k0 * ldu;
#pragma unroll 4
    for (int e = t; e < LB * LB; e += LB_NT) {
        const int i = e % LB, j = e / LB;  // a(i, j) = U(j, i) for i >= j
        double v = 0.0;
        if (i >= j) v = (i < nb) ? Ubb[j + (int64_t)i * ldu] : (i == j ? 1.0 : 0.0);
        a[j * LB_S + i] = v;
        if (i == j) rdiag[i] = 1.0 / v;
    }
    __syncthreads();
    constexpr int RG = LB / TG;
    double x[RG][RG];
    reg_lower_inverse<LB>(a, rdiag, rowk, x);
    double *out = UI + (int64_t)blockIdx.x * LB * LB;
#pragma unroll
    for (int ai = 0; ai < RG; ++ai)
#pragma unroll
        for (int bj = 0; bj < RG; ++bj) {
            const int i = tx + TG * ai, j = ty + TG * bj;  // x(i, j) = (U^-T)(i, j) = (U^-1)(j, i)
            out[LOWER_OUT ? i + j * LB : j + i * LB] = x[ai][bj];
        }
}

// Transposed copies of triangles through 64 x 64 LDS tiles (coalesced on both sides).
//   UP_TO_LOWER: dst = W (N x N, N >= n, ld ldd) from the upper triangle of src (n x n):
//     W(i, j) = src(j, i) for j <= i < n; the identity where i or j >= n; 0 above the diagonal.
//   else: the upper triangle of dst (n x n) from the lower triangle of src: dst(i, j) = src(j, i)
//     for i <= j < n; nothing else is written (N unused).
template <bool UP_TO_LOWER>
__global__ __launch_bounds__(LA_NT) void k_tri_transpose(int n, int N, const double *__restrict__ src, int64_t lds,
                                                         double *__restrict__ dst, int64_t ldd) {
    __shared__ double tile[TT][TT + 1];
    const int i0 = blockIdx.x * TT, j0 = blockIdx.y * TT;  // dst tile rows i0.., columns j0..
    const int tx = threadIdx.x % TT, ty = threadIdx.x / TT;  // 64 x 4
    const bool need = UP_TO_LOWER ? (i0 + TT - 1 >= j0) : (i0 <= j0 + TT - 1);
    if (!need) {
        if (UP_TO_LOWER)
            for (int c = ty; c < TT; c += 4)
                if (i0 + tx < N && j0 + c < N) dst[(i0 + tx) + (int64_t)(j0 + c) * ldd] = 0.0;
        return;
    }
    // src(j, i) for j in the dst column range, i in the dst row range: coalesced over j
    for (int c = ty; c < TT; c += 4) {
        const int i = i0 + c, j = j0 + tx;
        double v = 0.0;
        if (UP_TO_LOWER ? (j <= i && i < n) : (i <= j && j < n && i < n)) v = src[j + (int64_t)i * lds];
        tile[c][tx] = v;  // tile[i - i0][j - j0]
    }
    __syncthreads();
    for (int c = ty; c < TT; c += 4) {
        const int i = i0 + tx, j = j0 + c;
        const double v = tile[tx][c];
        if (UP_TO_LOWER) {
            double w = 0.0;
            if (i >= j) w = (i < n && j < n) ? v : (i == j ? 1.0 : 0.0);
            if (i < N && j < N) dst[i + (int64_t)j * ldd] = w;
        } else if (i <= j && j < n) {
            dst[i + (int64_t)j * ldd] = v;
        }
    }
}

// Lower triangle -> upper triangle (symmetrise), 64 x 64 tiles through LDS.
__global__ __launch_bounds__(LA_NT) void k_mirror_lower(int n, double *__restrict__ A, int64_t lda) {
    __shared__ double tile[TT][TT + 1];
    const int i0 = blockIdx.x * TT, j0 = blockIdx.y * TT;  // upper tile (i0 <= j0)
    if (i0 > j0) return;
    const int tx = threadIdx.x % TT, ty = threadIdx.x / TT;
    for (int c = ty; c < TT; c += 4) {  // read A(j, i) from the lower tile: rows j0.., columns i0..
        const int j = j0 + tx, i = i0 + c;
        tile[c][tx] = (j < n && i < n) ? A[j + (int64_t)i * lda] : 0.0;
    }
    __syncthreads();
    for (int c = ty; c < TT; c += 4) {
        const int i = i0 + tx, j = j0 + c;
        if (i < j && j < n) A[i + (int64_t)j * lda] = tile[tx][c];
    }
}

__global__ void k_scale_mat(int m, int n, double alpha, double *__restrict__ B, int64_t ldb) {
    const int64_t total = (int64_t)m * n;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += (int64_t)gridDim.x * blockDim.x)
        B[q % m + (q / m) * ldb] *= alpha;
}

// W := upper triangle of U (zeros below), n x n, ld n.
__global__ __launch_bounds__(LA_NT) void k_triu_copy(int n, const double *__restrict__ U, int64_t ldu,
                                                     double *__restrict__ W) {
    const int64_t total = (int64_t)n * n;
    for (int64_t q = (int64_t)blockIdx.x * LA_NT + threadIdx.x; q < total; q += (int64_t)gridDim.x * LA_NT) {
        const int64_t i = q % n, j = q / n;
        W[q] = (i <= j) ? U[i + j * ldu] : 0.0;
    }
}

__global__ void k_set_identity(int n, double *__restrict__ A, int64_t lda) {
    const int64_t total = (int64_t)n * n;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = q % n, j = q / n;
        A[i + j * lda] = (i == j) ? 1.0 : 0.0;
    }
}

static inline int la_grid(int64_t total) {
    return (int)std::max<int64_t>(1, std::min<int64_t>((total + LA_NT - 1) / LA_NT, 65536));
}

static int gemm64(omx_ctx ctx, bool trA, bool trB, int m, int n, int k, double alpha, const double *A, int64_t lda,
                  const double *B, int64_t ldb, double beta, double *C, int64_t ldc) {
    if (m <= 0 || n <= 0 || k <= 0) return OMX_OK;
    return launch_gemm(ctx, OMX_PREC_F64, OMX_F64, trA, trB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

// C (n x n) := alpha A B^T + beta C on the lower-triangle tiles where the kernel can restrict itself
// to them (the caller reads only the lower triangle), A, B n x k.
static int gemm64_lower(omx_ctx ctx, int n, int k, double alpha, const double *A, int64_t lda, const double *B,
                        int64_t ldb, double beta, double *C, int64_t ldc) {
    if (n <= 0 || k <= 0) return OMX_OK;
    ctx->gemm_sym = true;
    const int rc = launch_gemm(ctx, OMX_PREC_F64, OMX_F64, false, true, n, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    ctx->gemm_sym = false;
    return rc;
}

// Stream-ordered device buffers released when the scope ends.
struct DevBuf {
    omx_ctx ctx;
    double *p = nullptr;
    DevBuf(omx_ctx c) : ctx(c) {}
    int alloc(size_t count) {
        if (hipMallocAsync((void **)&p, sizeof(double) * std::max<size_t>(count, 1), ctx->stream) != hipSuccess) {
            p = nullptr;
            set_error("dense factorisation: device allocation of %zu doubles failed", count);
            return OMX_ERR_NOMEM;
        }
        return OMX_OK;
    }
    ~DevBuf() {
        if (p) (void)hipFreeAsync(p, ctx->stream);
    }
};

// Shape of the solution known in advance, so the GEMMs can skip what stays zero or is not needed:
//   TRI_NONE:   full solve;
//   TRI_COLS_GE (left 'n'):  B is upper triangular with the block structure of U (B = I at the
//               start): block row b of X lives in columns >= b*LB;
//   TRI_ROWS_GE (right 'n'): only the block-lower part of X is wanted: block column b for rows >= b*LB.
enum TriShape { TRI_NONE = 0, TRI_COLS_GE = 1, TRI_ROWS_GE = 2 };

static int trsm_impl(omx_ctx ctx, char side, char trans, int m, int n, double alpha, const double *U, int64_t ldu,
                     double *B, int64_t ldb, TriShape shape = TRI_NONE) {
    const bool left = (side == 'l' || side == 'L');
    const bool tr = is_trans(trans);
    const int nu = left ? m : n;  // order of U
    const int nblk = (nu + LB - 1) / LB;
    // U^T X = B runs on Lt = U^T (a transposed copy) and the transposed block inverses, so that all
    // its products are 'n','n' (the direct-to-LDS FP64 kernel)
    const bool lt = left && tr;
    DevBuf UI(ctx), T(ctx), Lt(ctx);
    int rc = UI.alloc((size_t)nblk * LB * LB);
    if (!rc) rc = T.alloc((size_t)LB * (left ? n : m));
    if (!rc && lt) rc = Lt.alloc((size_t)nu * nu);
    if (rc) return rc;
    if (lt) {
        const int nt = (nu + TT - 1) / TT;
        hipLaunchKernelGGL(k_tri_transpose<true>, dim3(nt, nt), dim3(LA_NT), 0, ctx->stream, nu, nu, U, ldu, Lt.p,
                           (int64_t)nu);
        hipLaunchKernelGGL(k_trinv_upper_blocks<true>, dim3(nblk), dim3(LB_NT), 0, ctx->stream, nu, U, ldu, UI.p);
    } else {
        hipLaunchKernelGGL(k_trinv_upper_blocks<false>, dim3(nblk), dim3(LB_NT), 0, ctx->stream, nu, U, ldu, UI.p);
    }
    if (alpha != 1.0)
        hipLaunchKernelGGL(k_scale_mat, dim3(la_grid((int64_t)m * n)), dim3(LA_NT), 0, ctx->stream, m, n, alpha, B,
                           ldb);
    auto copy_back = [&](double *dst, int rows, int cols, int64_t ldt) {
        return hipMemcpy2DAsync(dst, sizeof(double) * ldb, T.p, sizeof(double) * ldt, sizeof(double) * rows, cols,
                                hipMemcpyDeviceToDevice, ctx->stream) == hipSuccess
                   ? OMX_OK
                   : OMX_ERR_HIP;
    };
    // Two levels: outer blocks of LB_OUT, each solved block by block with updates confined to the
    // outer block (K = LB), then one update of everything beyond it with K = the outer width.
    const int nout = (nu + LB_OUT - 1) / LB_OUT;
    const bool fwd = left ? tr : !tr;  // forward solves take the blocks in order, backward in reverse
    for (int so = 0; so < nout; ++so) {
        const int o = fwd ? so : nout - 1 - so;
        const int o0 = o * LB_OUT, ob = std::min(LB_OUT, nu - o0), end = o0 + ob;
        const int nin = (ob + LB - 1) / LB;
        for (int si = 0; si < nin; ++si) {
            const int c = fwd ? si : nin - 1 - si;
            const int k0 = o0 + c * LB, nb = std::min(LB, nu - k0);
            const double *Ui = UI.p + (int64_t)(k0 / LB) * LB * LB;
            if (left) {
                const int c0 = shape == TRI_COLS_GE ? k0 : 0, nc = n - c0;
                double *Bb = B + k0 + (int64_t)c0 * ldb;
                // X_b = op(U_bb)^-1 B_b (UI holds the transposed inverses when tr), then the block rows
                // of this outer block that are still to come: B_r -= op(U_rb) X_b
                if ((rc = gemm64(ctx, false, false, nb, nc, nb, 1.0, Ui, LB, Bb, ldb, 0.0, T.p, nb))) return rc;
                if ((rc = copy_back(Bb, nb, nc, nb))) return rc;
                if (tr) {  // U^T X = B: rows k0+nb .. end, Lt(r, b) = U(b, r)^T
                    if ((rc = gemm64(ctx, false, false, end - k0 - nb, nc, nb, -1.0, Lt.p + (k0 + nb) + (int64_t)k0 * nu,
                                     nu, Bb, ldb, 1.0, Bb + nb, ldb)))
                        return rc;
                } else {  // U X = B: rows o0 .. k0
                    if ((rc = gemm64(ctx, false, false, k0 - o0, nc, nb, -1.0, U + o0 + (int64_t)k0 * ldu, ldu, Bb, ldb,
                                     1.0, B + o0 + (int64_t)c0 * ldb, ldb)))
                        return rc;
                }
            } else {
                const int r0 = shape == TRI_ROWS_GE ? k0 : 0, nr = m - r0;
                double *Bb = B + r0 + (int64_t)k0 * ldb;
                // X_b = B_b op(U_bb)^-1, then the block columns of this outer block still to come
                if ((rc = gemm64(ctx, false, tr, nr, nb, nb, 1.0, Bb, ldb, Ui, LB, 0.0, T.p, nr))) return rc;
                if ((rc = copy_back(Bb, nr, nb, nr))) return rc;
                if (!tr) {  // X U = B: columns k0+nb .. end (only their rows >= k0 + nb under TRI_ROWS_GE)
                    const int skip = shape == TRI_ROWS_GE ? nb : 0;
                    if ((rc = gemm64(ctx, false, false, nr - skip, end - k0 - nb, nb, -1.0, Bb + skip, ldb,
                                     U + k0 + (int64_t)(k0 + nb) * ldu, ldu, 1.0,
                                     B + r0 + skip + (int64_t)(k0 + nb) * ldb, ldb)))
                        return rc;
                } else {  // X U^T = B: columns o0 .. k0
                    if ((rc = gemm64(ctx, false, true, nr, k0 - o0, nb, -1.0, Bb, ldb, U + o0 + (int64_t)k0 * ldu, ldu,
                                     1.0, B + r0 + (int64_t)o0 * ldb, ldb)))
                        return rc;
                }
            }
        }
        // the outer block's update of everything beyond it, K = ob
        if (left) {
            const int c0 = shape == TRI_COLS_GE ? o0 : 0, nc = n - c0;
            const double *Xo = B + o0 + (int64_t)c0 * ldb;
            if (tr)
                rc = gemm64(ctx, false, false, nu - end, nc, ob, -1.0, Lt.p + end + (int64_t)o0 * nu, nu, Xo, ldb, 1.0,
                            B + end + (int64_t)c0 * ldb, ldb);
            else
                rc = gemm64(ctx, false, false, o0, nc, ob, -1.0, U + (int64_t)o0 * ldu, ldu, Xo, ldb, 1.0,
                            B + (int64_t)c0 * ldb, ldb);
        } else {
            if (!tr) {
                const int r0 = shape == TRI_ROWS_GE ? end : 0;
                rc = gemm64(ctx, false, false, m - r0, nu - end, ob, -1.0, B + r0 + (int64_t)o0 * ldb, ldb,
                            U + o0 + (int64_t)end * ldu, ldu, 1.0, B + r0 + (int64_t)end * ldb, ldb);
            } else {
                rc = gemm64(ctx, false, true, m, o0, ob, -1.0, B + (int64_t)o0 * ldb, ldb, U + (int64_t)o0 * ldu, ldu,
                            1.0, B, ldb);
            }
        }
        if (rc) return rc;
    }
    OMX_CHECK_HIP(hipGetLastError());
    return OMX_OK;
}

// Upper Cholesky of A (order n) through the lower Cholesky of its transposed upper triangle, in
// panels of LB_OUT columns: each panel is factorised block by block (the diagonal kernel, the panel
// solve below the block as a GEMM with its inverse, and the update of the panel's remaining columns),
// then the trailing matrix takes one lower-triangle update with K = the panel width.
static int potrf_impl(omx_ctx ctx, int n, double *A, int64_t lda, int *info) {
    const int N = (n + LB - 1) / LB * LB;
    int *d_info = (int *)(ctx->d_scalars + (kScalarSlots - 1));  // the GEMMs use the workspace
    OMX_CHECK_HIP(hipMemsetAsync(d_info, 0, sizeof(int), ctx->stream));
    DevBuf W(ctx), LI(ctx), T(ctx);
    int rc = W.alloc((size_t)N * N);
    if (!rc) rc = LI.alloc((size_t)LB * LB);
    if (!rc) rc = T.alloc((size_t)N * LB);
    if (rc) return rc;
    hipLaunchKernelGGL(k_tri_transpose<true>, dim3(N / TT, N / TT), dim3(LA_NT), 0, ctx->stream, n, N,
                       (const double *)A, lda, W.p, (int64_t)N);
    for (int o0 = 0; o0 < N; o0 += LB_OUT) {
        const int end = std::min(N, o0 + LB_OUT);
        for (int k0 = o0; k0 < end; k0 += DC) {
            hipLaunchKernelGGL(k_chol_diag_wave, dim3(1), dim3(64), 0, ctx->stream, W.p, (int64_t)N, k0, n, LI.p,
                               d_info);
            const int rest = N - k0 - DC;
            if (rest == 0) break;
            double *P = W.p + (k0 + DC) + (int64_t)k0 * N;  // A21, rest x DC
            // L21 = A21 L_kk^-T into T, back into W; the panel's remaining columns -= L21 L21(cols)^T
            if ((rc = gemm64(ctx, false, true, rest, DC, DC, 1.0, P, N, LI.p, DC, 0.0, T.p, N))) return rc;
            OMX_CHECK_HIP(hipMemcpy2DAsync(P, sizeof(double) * N, T.p, sizeof(double) * N, sizeof(double) * rest, DC,
                                           hipMemcpyDeviceToDevice, ctx->stream));
            if ((rc = gemm64(ctx, false, true, rest, end - k0 - DC, DC, -1.0, T.p, N, T.p, N, 1.0,
                             W.p + (k0 + DC) + (int64_t)(k0 + DC) * N, N)))
                return rc;
        }
        const int rest = N - end;
        if (rest > 0 &&
            (rc = gemm64_lower(ctx, rest, end - o0, -1.0, W.p + end + (int64_t)o0 * N, N, W.p + end + (int64_t)o0 * N, N,
                               1.0, W.p + end + (int64_t)end * N, N)))
            return rc;
    }
    const int nt = (n + TT - 1) / TT;
    hipLaunchKernelGGL(k_tri_transpose<false>, dim3(nt, nt), dim3(LA_NT), 0, ctx->stream, n, n, (const double *)W.p,
                       (int64_t)N, A, lda);
    OMX_CHECK_HIP(hipGetLastError());
    int h_info = 0;
    OMX_CHECK_HIP(hipMemcpyAsync(&h_info, d_info, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    OMX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    *info = h_info;
    return OMX_OK;
}

}  // namespace omx

using namespace omx;

extern "C" {

int omx_potrf(omx_ctx ctx, int dt, int n, void *A, int64_t lda, int *info) {
    OMX_REQUIRE(ctx && info && n >= 0 && lda >= std::max(1, n), "omx_potrf: bad arguments");
    OMX_REQUIRE(dt == OMX_F64, "omx_potrf: FP64 only");
    *info = 0;
    if (n == 0) return OMX_OK;
    OMX_REQUIRE(A, "omx_potrf: null operand");
    return potrf_impl(ctx, n, (double *)A, lda, info);
}

int omx_trsm(omx_ctx ctx, int dt, char side, char trans, int m, int n, double alpha, const void *U, int64_t ldu,
             void *B, int64_t ldb) {
    OMX_REQUIRE(ctx && m >= 0 && n >= 0 && valid_op(trans), "omx_trsm: bad arguments");
    OMX_REQUIRE(side == 'l' || side == 'L' || side == 'r' || side == 'R', "omx_trsm: side must be l or r");
    OMX_REQUIRE(dt == OMX_F64, "omx_trsm: FP64 only");
    const bool left = (side == 'l' || side == 'L');
    OMX_REQUIRE(ldu >= std::max(1, left ? m : n) && ldb >= std::max(1, m), "omx_trsm: bad leading dimension");
    if (!m || !n) return OMX_OK;
    OMX_REQUIRE(U && B, "omx_trsm: null operand");
    return trsm_impl(ctx, side, trans, m, n, alpha, (const double *)U, ldu, (double *)B, ldb);
}

int omx_trmm(omx_ctx ctx, int dt, char side, char trans, int m, int n, double alpha, const void *U, int64_t ldu,
             void *B, int64_t ldb) {
    OMX_REQUIRE(ctx && m >= 0 && n >= 0 && valid_op(trans), "omx_trmm: bad arguments");
    OMX_REQUIRE(side == 'l' || side == 'L' || side == 'r' || side == 'R', "omx_trmm: side must be l or r");
    OMX_REQUIRE(dt == OMX_F64, "omx_trmm: FP64 only");
    const bool left = (side == 'l' || side == 'L');
    const int nu = left ? m : n;
    OMX_REQUIRE(ldu >= std::max(1, nu) && ldb >= std::max(1, m), "omx_trmm: bad leading dimension");
    if (!m || !n) return OMX_OK;
    OMX_REQUIRE(U && B, "omx_trmm: null operand");
    // W = triu(U) (n_u^2) and T = copy of B (m*n) in stream-ordered buffers of their own (the GEMM
    // uses the context workspace), then B = alpha * op(W) * T or T * op(W)
    DevBuf W(ctx), T(ctx);
    int rc = W.alloc((size_t)nu * nu);
    if (!rc) rc = T.alloc((size_t)m * n);
    if (rc) return rc;
    hipLaunchKernelGGL(k_triu_copy, dim3(la_grid((int64_t)nu * nu)), dim3(LA_NT), 0, ctx->stream, nu,
                       (const double *)U, ldu, W.p);
    OMX_CHECK_HIP(hipMemcpy2DAsync(T.p, sizeof(double) * m, B, sizeof(double) * ldb, sizeof(double) * m, n,
                                   hipMemcpyDeviceToDevice, ctx->stream));
    const bool tr = is_trans(trans);
    if (left)
        rc = gemm64(ctx, tr, false, m, n, m, alpha, W.p, nu, T.p, m, 0.0, (double *)B, ldb);
    else
        rc = gemm64(ctx, false, tr, m, n, n, alpha, T.p, m, W.p, nu, 0.0, (double *)B, ldb);
    if (rc) return rc;
    OMX_CHECK_HIP(hipGetLastError());
    return OMX_OK;
}

int omx_sygst(omx_ctx ctx, int dt, int n, void *H, int64_t ldh, const void *U, int64_t ldu) {
    OMX_REQUIRE(ctx && n >= 0 && ldh >= std::max(1, n) && ldu >= std::max(1, n), "omx_sygst: bad arguments");
    OMX_REQUIRE(dt == OMX_F64, "omx_sygst: FP64 only");
    if (!n) return OMX_OK;
    OMX_REQUIRE(H && U, "omx_sygst: null operand");
    // Y := U^-T H (full), then the block-lower part of Y U^-1, mirrored: the result is symmetric, as
    // the reference's after dsygst and its mirror copy.
    int rc = trsm_impl(ctx, 'l', 't', n, n, 1.0, (const double *)U, ldu, (double *)H, ldh);
    if (rc) return rc;
    rc = trsm_impl(ctx, 'r', 'n', n, n, 1.0, (const double *)U, ldu, (double *)H, ldh, TRI_ROWS_GE);
    if (rc) return rc;
    const int nt = (n + TT - 1) / TT;
    hipLaunchKernelGGL(k_mirror_lower, dim3(nt, nt), dim3(LA_NT), 0, ctx->stream, n, (double *)H, ldh);
    OMX_CHECK_HIP(hipGetLastError());
    return OMX_OK;
}

int omx_inverse_spd(omx_ctx ctx, int dt, int n, void *A, int64_t lda, int *info) {
    OMX_REQUIRE(ctx && info && n >= 0 && lda >= std::max(1, n), "omx_inverse_spd: bad arguments");
    OMX_REQUIRE(dt == OMX_F64, "omx_inverse_spd: FP64 only");
    *info = 0;
    if (!n) return OMX_OK;
    OMX_REQUIRE(A, "omx_inverse_spd: null operand");
    // U = chol(A) in a workspace copy; X = U^-1 (upper triangular: the solve skips the zero block
    // columns); A^-1 = X X^T by row chunks of LB_OUT: rows [r0, r0 + w) against columns [0, r0 + w)
    // need only p >= r0 (X(i, p) = 0 for p < i), so each chunk is one product with K = n - r0
    // (n^3/3 flops in all), then the lower triangle is mirrored
    DevBuf U(ctx), X(ctx);
    int rc = U.alloc((size_t)n * n);
    if (!rc) rc = X.alloc((size_t)n * n);
    if (rc) return rc;
    OMX_CHECK_HIP(hipMemcpy2DAsync(U.p, sizeof(double) * n, A, sizeof(double) * lda, sizeof(double) * n, n,
                                   hipMemcpyDeviceToDevice, ctx->stream));
    rc = potrf_impl(ctx, n, U.p, n, info);
    if (rc || *info) return rc;
    hipLaunchKernelGGL(k_set_identity, dim3(la_grid((int64_t)n * n)), dim3(LA_NT), 0, ctx->stream, n, X.p,
                       (int64_t)n);
    if ((rc = trsm_impl(ctx, 'l', 'n', n, n, 1.0, U.p, n, X.p, n, TRI_COLS_GE))) return rc;
    double *Ad = (double *)A;
    for (int r0 = 0; r0 < n; r0 += LB_OUT) {
        const int w = std::min(LB_OUT, n - r0);
        if ((rc = gemm64(ctx, false, true, w, r0 + w, n - r0, 1.0, X.p + r0 + (int64_t)r0 * n, n, X.p + (int64_t)r0 * n,
                         n, 0.0, Ad + r0, lda)))
            return rc;
    }
    const int nt = (n + TT - 1) / TT;
    hipLaunchKernelGGL(k_mirror_lower, dim3(nt, nt), dim3(LA_NT), 0, ctx->stream, n, Ad, lda);
    OMX_CHECK_HIP(hipGetLastError());
    OMX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    return OMX_OK;
}

}  // extern "C"
