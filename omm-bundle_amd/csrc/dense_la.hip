// Dense factorisations for libOMM's Cholesky and preconditioned flavours (FP64, column-major):
//   omx_potrf      S = U^T U, upper (dpotrf 'u', libOMM/src/omm_ops.F90:301)
//   omx_trsm       B := alpha * op(U)^-1 * B  or  alpha * B * op(U)^-1, U upper (dtrsm)
//   omx_trmm       B := alpha * op(U) * B     or  alpha * B * op(U),     U upper (dtrmm)
//   omx_sygst      H := U^-T H U^-1, both triangles (dsygst itype 1 'u' + the mirror copy,
//                  omm_ops.F90:481-487)
//   omx_inverse_spd A := A^-1 for symmetric positive definite A (m_inverse, omm_ops.F90:1065-1075
//                  uses dsytrf/dsytri; on an SPD matrix the Cholesky route gives the same inverse)
// All are blocked with NB = 32: the diagonal-block work (the unblocked factorisation and the
// triangular solves against a 32x32 block) runs in one workgroup out of LDS, and everything else
// is a GEMM update on the MFMA kernel of gemm.hip.  Only the upper triangle of U is read.
#include <algorithm>

#include "omx_internal.h"

namespace omx {

constexpr int LA_NB = 32;
// Outer block of the factorisations and solves: each outer step's update is one GEMM with K = 256
// (compute-bound), and the 32-wide kernels run only inside the 256-wide diagonal blocks.
constexpr int LA_OUTER = 256;
constexpr int LA_NT = 256;

// Unblocked upper Cholesky of the nb x nb diagonal block at A (in place), one workgroup; sets
// *info = global index + 1 of the first non-positive pivot (0 on success).
__global__ __launch_bounds__(LA_NT) void k_potrf_diag(int nb, double *__restrict__ A, int64_t lda, int k0,
                                                      int *__restrict__ info) {
    __shared__ double a[LA_NB][LA_NB + 1];
    __shared__ int bad;
    for (int e = threadIdx.x; e < nb * nb; e += LA_NT) {
        const int i = e % nb, j = e / nb;
        a[i][j] = (i <= j) ? A[i + (int64_t)j * lda] : 0.0;
    }
    if (threadIdx.x == 0) bad = 0;
    __syncthreads();
    for (int k = 0; k < nb; ++k) {
        if (threadIdx.x == 0) {
            const double d = a[k][k];
            if (!(d > 0.0) && !bad) bad = k0 + k + 1;
            a[k][k] = sqrt(d > 0.0 ? d : 1.0);
        }
        __syncthreads();
        const double dk = a[k][k];
        for (int j = k + 1 + threadIdx.x; j < nb; j += LA_NT) a[k][j] /= dk;
        __syncthreads();
        // trailing update of the upper triangle: a[i][j] -= a[k][i] * a[k][j], k < i <= j
        for (int e = threadIdx.x; e < nb * nb; e += LA_NT) {
            const int i = e % nb, j = e / nb;
            if (i > k && i <= j) a[i][j] -= a[k][i] * a[k][j];
        }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < nb * nb; e += LA_NT) {
        const int i = e % nb, j = e / nb;
        if (i <= j) A[i + (int64_t)j * lda] = a[i][j];
    }
    if (threadIdx.x == 0 && bad && *info == 0) *info = bad;
}

// Left solves against the nb x nb upper block U (at lda): each thread owns one column of B
// (nb rows).  TRANS: U^T X = B (forward); else U X = B (backward).  X overwrites B; alpha scales.
template <bool TRANS>
__global__ __launch_bounds__(LA_NT) void k_trsm_left_block(int nb, int ncols, const double *__restrict__ U,
                                                           int64_t ldu, double alpha, double *__restrict__ B,
                                                           int64_t ldb) {
    __shared__ double u[LA_NB][LA_NB + 1];
    for (int e = threadIdx.x; e < nb * nb; e += LA_NT) {
        const int i = e % nb, j = e / nb;
        u[i][j] = (i <= j) ? U[i + (int64_t)j * ldu] : 0.0;
    }
    __syncthreads();
    const int c = blockIdx.x * LA_NT + threadIdx.x;
    if (c >= ncols) return;
    double x[LA_NB];
    double *col = B + (int64_t)c * ldb;
#pragma unroll
    for (int i = 0; i < LA_NB; ++i) x[i] = (i < nb) ? alpha * col[i] : 0.0;
    if constexpr (TRANS) {
#pragma unroll
        for (int i = 0; i < LA_NB; ++i) {
            if (i < nb) {
                double s = x[i];
#pragma unroll
                for (int p = 0; p < LA_NB; ++p)
                    if (p < i) s -= u[p][i] * x[p];
                x[i] = s / u[i][i];
            }
        }
    } else {
#pragma unroll
        for (int ii = LA_NB - 1; ii >= 0; --ii) {
            if (ii < nb) {
                double s = x[ii];
#pragma unroll
                for (int p = 0; p < LA_NB; ++p)
                    if (p > ii && p < nb) s -= u[ii][p] * x[p];
                x[ii] = s / u[ii][ii];
            }
        }
    }
#pragma unroll
    for (int i = 0; i < LA_NB; ++i)
        if (i < nb) col[i] = x[i];
}

// Right solves against the nb x nb upper block U: each thread owns one row of B (nb columns).
// TRANS: X U^T = B (backward over columns); else X U = B (forward).
template <bool TRANS>
__global__ __launch_bounds__(LA_NT) void k_trsm_right_block(int nb, int nrows, const double *__restrict__ U,
                                                            int64_t ldu, double alpha, double *__restrict__ B,
                                                            int64_t ldb) {
    __shared__ double u[LA_NB][LA_NB + 1];
    for (int e = threadIdx.x; e < nb * nb; e += LA_NT) {
        const int i = e % nb, j = e / nb;
        u[i][j] = (i <= j) ? U[i + (int64_t)j * ldu] : 0.0;
    }
    __syncthreads();
    const int r = blockIdx.x * LA_NT + threadIdx.x;
    if (r >= nrows) return;
    double x[LA_NB];
#pragma unroll
    for (int j = 0; j < LA_NB; ++j) x[j] = (j < nb) ? alpha * B[r + (int64_t)j * ldb] : 0.0;
    if constexpr (!TRANS) {
        // x_j = (b_j - sum_{p<j} x_p U_pj) / U_jj
#pragma unroll
        for (int j = 0; j < LA_NB; ++j) {
            if (j < nb) {
                double s = x[j];
#pragma unroll
                for (int p = 0; p < LA_NB; ++p)
                    if (p < j) s -= x[p] * u[p][j];
                x[j] = s / u[j][j];
            }
        }
    } else {
        // X U^T = B: x_j = (b_j - sum_{p>j} x_p U_jp) / U_jj
#pragma unroll
        for (int jj = LA_NB - 1; jj >= 0; --jj) {
            if (jj < nb) {
                double s = x[jj];
#pragma unroll
                for (int p = 0; p < LA_NB; ++p)
                    if (p > jj && p < nb) s -= x[p] * u[jj][p];
                x[jj] = s / u[jj][jj];
            }
        }
    }
#pragma unroll
    for (int j = 0; j < LA_NB; ++j)
        if (j < nb) B[r + (int64_t)j * ldb] = x[j];
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

// Upper triangle -> lower triangle (symmetrise).
__global__ __launch_bounds__(LA_NT) void k_mirror_upper(int n, double *__restrict__ A, int64_t lda) {
    const int64_t total = (int64_t)n * n;
    for (int64_t q = (int64_t)blockIdx.x * LA_NT + threadIdx.x; q < total; q += (int64_t)gridDim.x * LA_NT) {
        const int64_t i = q % n, j = q / n;
        if (i > j) A[i + j * lda] = A[j + i * lda];
    }
}

// Strict lower triangle of src (n x n) into dst, the upper triangle untouched.
__global__ __launch_bounds__(LA_NT) void k_copy_lower(int n, const double *__restrict__ src, int64_t lds,
                                                      double *__restrict__ dst, int64_t ldd) {
    const int64_t total = (int64_t)n * n;
    for (int64_t q = (int64_t)blockIdx.x * LA_NT + threadIdx.x; q < total; q += (int64_t)gridDim.x * LA_NT) {
        const int64_t i = q % n, j = q / n;
        if (i > j) dst[i + j * ldd] = src[i + j * lds];
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

static int trsm_impl(omx_ctx ctx, char side, char trans, int m, int n, double alpha, const double *U, int64_t ldu,
                     double *B, int64_t ldb, int bs = LA_OUTER) {
    const bool left = (side == 'l' || side == 'L');
    const bool tr = is_trans(trans);
    const int nu = left ? m : n;  // order of U
    const int nblk = (nu + bs - 1) / bs;
    double a = alpha;
    // diagonal blocks wider than the kernels' 32 are solved by the same routine in 32-wide blocks
    auto sub = [&](int k0, int nb, double aa) {
        return left ? trsm_impl(ctx, side, trans, nb, n, aa, U + k0 + (int64_t)k0 * ldu, ldu, B + k0, ldb, LA_NB)
                    : trsm_impl(ctx, side, trans, m, nb, aa, U + k0 + (int64_t)k0 * ldu, ldu, B + (int64_t)k0 * ldb,
                                ldb, LA_NB);
    };
    if (left) {
        if (tr) {  // U^T X = B: forward over block rows
            for (int b = 0; b < nblk; ++b) {
                const int k0 = b * bs, nb = std::min(bs, nu - k0);
                if (nb > LA_NB) {
                    if (int rc = sub(k0, nb, a)) return rc;
                } else {
                    hipLaunchKernelGGL(k_trsm_left_block<true>, dim3((n + LA_NT - 1) / LA_NT), dim3(LA_NT), 0,
                                       ctx->stream, nb, n, U + k0 + (int64_t)k0 * ldu, ldu, a, B + k0, ldb);
                }
                const int rest = nu - k0 - nb;
                if (rest > 0) {
                    // B[k0+nb:, :] = a*B[k0+nb:, :] - U[k0:k0+nb, k0+nb:]^T X_b
                    int rc = launch_gemm(ctx, OMX_PREC_F64, OMX_F64, true, false, rest, n, nb, -1.0,
                                         U + k0 + (int64_t)(k0 + nb) * ldu, ldu, B + k0, ldb, a, B + k0 + nb, ldb);
                    if (rc) return rc;
                }
                a = 1.0;  // alpha applied once to every row block
            }
        } else {  // U X = B: backward
            for (int b = nblk - 1; b >= 0; --b) {
                const int k0 = b * bs, nb = std::min(bs, nu - k0);
                if (nb > LA_NB) {
                    if (int rc = sub(k0, nb, a)) return rc;
                } else {
                    hipLaunchKernelGGL(k_trsm_left_block<false>, dim3((n + LA_NT - 1) / LA_NT), dim3(LA_NT), 0,
                                       ctx->stream, nb, n, U + k0 + (int64_t)k0 * ldu, ldu, a, B + k0, ldb);
                }
                if (k0 > 0) {
                    // B[:k0, :] = a*B[:k0, :] - U[:k0, k0:k0+nb] X_b
                    int rc = launch_gemm(ctx, OMX_PREC_F64, OMX_F64, false, false, k0, n, nb, -1.0,
                                         U + (int64_t)k0 * ldu, ldu, B + k0, ldb, a, B, ldb);
                    if (rc) return rc;
                }
                a = 1.0;
            }
        }
    } else {
        if (!tr) {  // X U = B: forward over block columns
            for (int b = 0; b < nblk; ++b) {
                const int k0 = b * bs, nb = std::min(bs, nu - k0);
                if (nb > LA_NB) {
                    if (int rc = sub(k0, nb, a)) return rc;
                } else {
                    hipLaunchKernelGGL(k_trsm_right_block<false>, dim3((m + LA_NT - 1) / LA_NT), dim3(LA_NT), 0,
                                       ctx->stream, nb, m, U + k0 + (int64_t)k0 * ldu, ldu, a,
                                       B + (int64_t)k0 * ldb, ldb);
                }
                const int rest = nu - k0 - nb;
                if (rest > 0) {
                    // B[:, k0+nb:] = a*B[:, k0+nb:] - X_b U[k0:k0+nb, k0+nb:]
                    int rc = launch_gemm(ctx, OMX_PREC_F64, OMX_F64, false, false, m, rest, nb, -1.0,
                                         B + (int64_t)k0 * ldb, ldb, U + k0 + (int64_t)(k0 + nb) * ldu, ldu, a,
                                         B + (int64_t)(k0 + nb) * ldb, ldb);
                    if (rc) return rc;
                }
                a = 1.0;
            }
        } else {  // X U^T = B: backward over block columns
            for (int b = nblk - 1; b >= 0; --b) {
                const int k0 = b * bs, nb = std::min(bs, nu - k0);
                if (nb > LA_NB) {
                    if (int rc = sub(k0, nb, a)) return rc;
                } else {
                    hipLaunchKernelGGL(k_trsm_right_block<true>, dim3((m + LA_NT - 1) / LA_NT), dim3(LA_NT), 0,
                                       ctx->stream, nb, m, U + k0 + (int64_t)k0 * ldu, ldu, a,
                                       B + (int64_t)k0 * ldb, ldb);
                }
                if (k0 > 0) {
                    // B[:, :k0] = a*B[:, :k0] - X_b U[:k0, k0:k0+nb]^T
                    int rc = launch_gemm(ctx, OMX_PREC_F64, OMX_F64, false, true, m, k0, nb, -1.0,
                                         B + (int64_t)k0 * ldb, ldb, U + (int64_t)k0 * ldu, ldu, a, B, ldb);
                    if (rc) return rc;
                }
                a = 1.0;
            }
        }
    }
    OMX_CHECK_HIP(hipGetLastError());
    return OMX_OK;
}

// The Cholesky of one diagonal block (order n <= LA_OUTER) in 32-wide steps; pivots are reported
// at their global index koff + local index.
static int potrf_block(omx_ctx ctx, int n, double *A, int64_t lda, int koff, int *d_info) {
    for (int k0 = 0; k0 < n; k0 += LA_NB) {
        const int nb = std::min(LA_NB, n - k0);
        hipLaunchKernelGGL(k_potrf_diag, dim3(1), dim3(LA_NT), 0, ctx->stream, nb, A + k0 + (int64_t)k0 * lda, lda,
                           koff + k0, d_info);
        const int rest = n - k0 - nb;
        if (rest > 0) {
            // panel: U[k0:k0+nb, k0+nb:] = U_kk^-T A[k0:k0+nb, k0+nb:]
            hipLaunchKernelGGL(k_trsm_left_block<true>, dim3((rest + LA_NT - 1) / LA_NT), dim3(LA_NT), 0,
                               ctx->stream, nb, rest, A + k0 + (int64_t)k0 * lda, lda, 1.0,
                               A + k0 + (int64_t)(k0 + nb) * lda, lda);
            // trailing: A[k0+nb:, k0+nb:] -= panel^T panel (both triangles; only the upper is used)
            int rc = launch_gemm(ctx, OMX_PREC_F64, OMX_F64, true, false, rest, rest, nb, -1.0,
                                 A + k0 + (int64_t)(k0 + nb) * lda, lda, A + k0 + (int64_t)(k0 + nb) * lda, lda, 1.0,
                                 A + k0 + nb + (int64_t)(k0 + nb) * lda, lda);
            if (rc) return rc;
        }
    }
    return OMX_OK;
}

static int potrf_impl(omx_ctx ctx, int n, double *A, int64_t lda, int *info) {
    // the pivot flag lives in the last scalar slot (the GEMM updates use the workspace)
    int *d_info = (int *)(ctx->d_scalars + (kScalarSlots - 1));
    OMX_CHECK_HIP(hipMemsetAsync(d_info, 0, sizeof(int), ctx->stream));
    for (int k0 = 0; k0 < n; k0 += LA_OUTER) {
        const int kb = std::min(LA_OUTER, n - k0);
        double *Akk = A + k0 + (int64_t)k0 * lda;
        int rc = potrf_block(ctx, kb, Akk, lda, k0, d_info);
        if (rc) return rc;
        const int rest = n - k0 - kb;
        if (rest > 0) {
            // panel rows k0..k0+kb: U_kk^-T A, then the trailing update with K = kb
            double *P = A + k0 + (int64_t)(k0 + kb) * lda;
            if ((rc = trsm_impl(ctx, 'l', 't', kb, rest, 1.0, Akk, lda, P, lda))) return rc;
            if ((rc = launch_gemm(ctx, OMX_PREC_F64, OMX_F64, true, false, rest, rest, kb, -1.0, P, lda, P, lda, 1.0,
                                  A + k0 + kb + (int64_t)(k0 + kb) * lda, lda)))
                return rc;
        }
    }
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
    // The trailing updates write both triangles; like dpotrf 'u' the call leaves the strict lower
    // triangle as it was, so it is saved here and put back afterwards.
    double *save = nullptr;
    if (n > 1) {
        if (hipMalloc(&save, sizeof(double) * (size_t)n * n) != hipSuccess) {
            set_error("omx_potrf: device allocation of %d x %d failed", n, n);
            return OMX_ERR_NOMEM;
        }
        hipLaunchKernelGGL(k_copy_lower, dim3(la_grid((int64_t)n * n)), dim3(LA_NT), 0, ctx->stream, n,
                           (const double *)A, lda, save, (int64_t)n);
    }
    int rc = potrf_impl(ctx, n, (double *)A, lda, info);
    if (save) {
        hipLaunchKernelGGL(k_copy_lower, dim3(la_grid((int64_t)n * n)), dim3(LA_NT), 0, ctx->stream, n,
                           (const double *)save, (int64_t)n, (double *)A, lda);
        const hipError_t e = hipStreamSynchronize(ctx->stream);
        (void)hipFree(save);
        if (!rc && e != hipSuccess) {
            set_error("omx_potrf: %s", hipGetErrorString(e));
            rc = OMX_ERR_HIP;
        }
    }
    return rc;
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
    const size_t wbytes = sizeof(double) * (size_t)nu * nu, tbytes = sizeof(double) * (size_t)m * n;
    double *W = nullptr;
    OMX_CHECK_HIP(hipMallocAsync((void **)&W, wbytes + tbytes, ctx->stream));
    double *T = W + (size_t)nu * nu;
    int rc;
    hipLaunchKernelGGL(k_triu_copy, dim3(la_grid((int64_t)nu * nu)), dim3(LA_NT), 0, ctx->stream, nu,
                       (const double *)U, ldu, W);
    OMX_CHECK_HIP(hipMemcpy2DAsync(T, sizeof(double) * m, B, sizeof(double) * ldb, sizeof(double) * m, n,
                                   hipMemcpyDeviceToDevice, ctx->stream));
    const bool tr = is_trans(trans);
    if (left)
        rc = launch_gemm(ctx, OMX_PREC_F64, OMX_F64, tr, false, m, n, m, alpha, W, nu, T, m, 0.0, B, ldb);
    else
        rc = launch_gemm(ctx, OMX_PREC_F64, OMX_F64, false, tr, m, n, n, alpha, T, m, W, nu, 0.0, B, ldb);
    (void)hipFreeAsync(W, ctx->stream);
    if (rc) return rc;
    OMX_CHECK_HIP(hipGetLastError());
    return OMX_OK;
}

int omx_sygst(omx_ctx ctx, int dt, int n, void *H, int64_t ldh, const void *U, int64_t ldu) {
    OMX_REQUIRE(ctx && n >= 0 && ldh >= std::max(1, n) && ldu >= std::max(1, n), "omx_sygst: bad arguments");
    OMX_REQUIRE(dt == OMX_F64, "omx_sygst: FP64 only");
    if (!n) return OMX_OK;
    OMX_REQUIRE(H && U, "omx_sygst: null operand");
    // H := U^-T H, then H := H U^-1; the result is symmetric up to rounding: keep the upper triangle
    // and mirror it, as the reference does after dsygst.
    int rc = trsm_impl(ctx, 'l', 't', n, n, 1.0, (const double *)U, ldu, (double *)H, ldh);
    if (rc) return rc;
    rc = trsm_impl(ctx, 'r', 'n', n, n, 1.0, (const double *)U, ldu, (double *)H, ldh);
    if (rc) return rc;
    hipLaunchKernelGGL(k_mirror_upper, dim3(la_grid((int64_t)n * n)), dim3(LA_NT), 0, ctx->stream, n, (double *)H,
                       ldh);
    OMX_CHECK_HIP(hipGetLastError());
    return OMX_OK;
}

int omx_inverse_spd(omx_ctx ctx, int dt, int n, void *A, int64_t lda, int *info) {
    OMX_REQUIRE(ctx && info && n >= 0 && lda >= std::max(1, n), "omx_inverse_spd: bad arguments");
    OMX_REQUIRE(dt == OMX_F64, "omx_inverse_spd: FP64 only");
    *info = 0;
    if (!n) return OMX_OK;
    OMX_REQUIRE(A, "omx_inverse_spd: null operand");
    // U = chol(A) in a workspace copy; X = I; X := U^-T X; X := U^-1 X  ->  X = A^-1
    const size_t bytes = sizeof(double) * (size_t)n * n;
    double *U = nullptr;
    OMX_CHECK_HIP(hipMallocAsync((void **)&U, bytes, ctx->stream));
    OMX_CHECK_HIP(hipMemcpy2DAsync(U, sizeof(double) * n, A, sizeof(double) * lda, sizeof(double) * n, n,
                                   hipMemcpyDeviceToDevice, ctx->stream));
    int rc = potrf_impl(ctx, n, U, n, info);
    if (rc == OMX_OK && *info == 0) {
        hipLaunchKernelGGL(k_set_identity, dim3(la_grid((int64_t)n * n)), dim3(LA_NT), 0, ctx->stream, n,
                           (double *)A, lda);
        rc = trsm_impl(ctx, 'l', 't', n, n, 1.0, U, n, (double *)A, lda);
        if (rc == OMX_OK) rc = trsm_impl(ctx, 'l', 'n', n, n, 1.0, U, n, (double *)A, lda);
    }
    (void)hipFreeAsync(U, ctx->stream);
    if (rc) return rc;
    OMX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    return OMX_OK;
}

}  // extern "C"
