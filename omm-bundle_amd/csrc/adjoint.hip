// Adjoint of the converged OMM energy.
//
// At a minimiser C of E(C; H) = tr[(2I - SW) HW], HW = C H C^T, SW = C S C^T, the gradient in C
// vanishes, so the total derivative of the minimum with respect to H is the partial one:
//     dE/dH(k,l) = sum_ij (2I - SW)(i,j) C(j,k) C(i,l) = [C^T (2I - SW) C](k,l),
// the density matrix (libOMM/src/omm_ops.F90:104-124 computes the same matrix as D_min).  On
// the stored blocks of H it is the SDDMM omx_sddmm_bsr(C, (2I - SW) C).  For a Hamiltonian that
// is linear in parameters theta, H = sum_p theta_p J_p, the chain rule gives
//     dE/dtheta_p = sum over stored entries e of J_p(e) * dE/dH(e),
// a sparse contraction done here by omx_bsr_param_grad: the Jacobian is given per parameter as a
// segment of (entry, coefficient) pairs; one workgroup reduces one parameter's segment in a fixed
// order (deterministic).
#include <algorithm>

#include "omx_internal.h"

namespace omx {

constexpr int PG_NT = 256;

__global__ __launch_bounds__(PG_NT) void k_param_grad(int nparam, const int64_t *__restrict__ seg,
                                                      const int64_t *__restrict__ ent,
                                                      const double *__restrict__ coef,
                                                      const double *__restrict__ dval, int64_t nval,
                                                      double *__restrict__ out) {
    const int p = blockIdx.x;
    if (p >= nparam) return;
    const int64_t b = seg[p], e = seg[p + 1];
    double s = 0.0;
    for (int64_t k = b + threadIdx.x; k < e; k += PG_NT) {
        const int64_t idx = ent[k];
        if (idx >= 0 && idx < nval) s = fma(coef[k], dval[idx], s);
    }
    __shared__ double red[PG_NT / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
#pragma unroll
        for (int w = 0; w < PG_NT / 64; ++w) t += red[w];
        out[p] = t;
    }
}

}  // namespace omx

using namespace omx;

extern "C" int omx_bsr_param_grad(omx_ctx ctx, omx_bsr Dp, int nparam, const int64_t *seg, const int64_t *ent,
                                  const double *coef, double *grad) {
    OMX_REQUIRE(ctx && Dp && grad && nparam >= 0, "omx_bsr_param_grad: bad arguments");
    OMX_REQUIRE(!Dp->is_complex, "omx_bsr_param_grad: real operator expected");
    OMX_REQUIRE(nparam <= omx::kScalarSlots, "omx_bsr_param_grad: at most %d parameters", omx::kScalarSlots);
    if (nparam == 0) return OMX_OK;
    OMX_REQUIRE(seg && ent && coef, "omx_bsr_param_grad: null Jacobian");
    hipEvent_t ev;
    prof_begin(ctx, OMX_K_DOT, &ev);
    hipLaunchKernelGGL(k_param_grad, dim3(nparam), dim3(PG_NT), 0, ctx->stream, nparam, seg, ent, coef, Dp->d_val,
                       16 * Dp->nnzb, ctx->d_scalars);
    prof_end(ctx, OMX_K_DOT, ev);
    OMX_CHECK_HIP(hipGetLastError());
    int rc = read_scalars(ctx, nparam);
    if (rc) return rc;
    for (int p = 0; p < nparam; ++p) grad[p] = ctx->h_scalars[p];
    return OMX_OK;
}

// Element (i, j) (1-based) of a BSR-4 operator; 0 when its block is not stored.
extern "C" int omx_bsr_get_element(omx_bsr a, int i, int j, double *out) {
    OMX_REQUIRE(a && out && i >= 1 && j >= 1 && i <= a->nb && j <= a->nb, "omx_bsr_get_element: bad arguments");
    OMX_REQUIRE(!a->is_complex, "omx_bsr_get_element: real operator expected");
    const int I = (i - 1) / 4, J = (j - 1) / 4, r = (i - 1) % 4, c = (j - 1) % 4;
    hipStream_t st = a->ctx->stream;
    int cp[2];
    OMX_CHECK_HIP(hipMemcpyAsync(cp, a->d_colptr + J, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
    OMX_CHECK_HIP(hipStreamSynchronize(st));
    *out = 0.0;
    const int n = cp[1] - cp[0];
    if (n <= 0) return OMX_OK;
    std::vector<int> rows(n);
    OMX_CHECK_HIP(hipMemcpyAsync(rows.data(), a->d_rowidx + cp[0], n * sizeof(int), hipMemcpyDeviceToHost, st));
    OMX_CHECK_HIP(hipStreamSynchronize(st));
    auto it = std::lower_bound(rows.begin(), rows.end(), I);
    if (it == rows.end() || *it != I) return OMX_OK;
    const int64_t l = cp[0] + (it - rows.begin());
    OMX_CHECK_HIP(hipMemcpyAsync(out, a->d_val + 16 * l + r + 4 * c, sizeof(double), hipMemcpyDeviceToHost, st));
    OMX_CHECK_HIP(hipStreamSynchronize(st));
    return OMX_OK;
}
