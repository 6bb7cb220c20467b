// Dense x block-sparse product Y := alpha*X*A + beta*Y on CDNA4 (the H*C and S*C products of
// the OMM loop; MatrixSwitch/src/MatrixSwitch_mm_multiply.F90:263-327 and
// pspBLAS/src/pspLevel3/pspMspm_nn.F90:58-212 compute the same product from CSC).
//
// X, Y: no x nb column-major; a basis column X(:,k) is a contiguous vector of `no` state
// coefficients.  A: BSR-4 by block columns.  Output column 4J+c is
//     Y(:,4J+c) = sum over blocks (I,J), r = 0..3 of X(:,4I+r) * A_IJ(r,c),
// summed in ascending (I, r), the order of the reference's CSC loop.
//
// Mapping: a workgroup owns a chunk of ROWS consecutive states (VEC per thread, loaded as one
// 16-byte vector per basis column) and a run of JPER block columns.  Each thread keeps 4*VEC
// accumulators; the 16 block values are wave-uniform and come through the scalar cache.  The
// workgroup id is remapped so that each XCD owns a contiguous band of state chunks and walks
// the block columns in order: the neighbouring basis columns a block column gathers were read
// a few block columns earlier by the same XCD and are still in its L2 (lattice neighbours in
// +-x, +-y) or in the shared memory-side cache (+-z), so HBM sees X about once.
#include <algorithm>

#include "omx_internal.h"

namespace omx {

template <typename T, int VEC>
struct VecT;
template <>
struct VecT<double, 2> {
    typedef double2 type;
};
template <>
struct VecT<double, 1> {
    typedef double type;
};
template <>
struct VecT<float, 4> {
    typedef float4 type;
};
template <>
struct VecT<float, 1> {
    typedef float type;
};

template <typename T, int VEC>
__device__ inline void vload(const T *p, T (&out)[VEC]) {
    typedef typename VecT<T, VEC>::type V;
    V v = *reinterpret_cast<const V *>(p);
    const T *s = reinterpret_cast<const T *>(&v);
#pragma unroll
    for (int q = 0; q < VEC; ++q) out[q] = s[q];
}
template <typename T, int VEC>
__device__ inline void vstore(T *p, const T (&in)[VEC]) {
    typedef typename VecT<T, VEC>::type V;
    V v;
    T *d = reinterpret_cast<T *>(&v);
#pragma unroll
    for (int q = 0; q < VEC; ++q) d[q] = in[q];
    *reinterpret_cast<V *>(p) = v;
}

constexpr int SPMM_NT = 256;

template <typename T, int VEC>
__global__ __launch_bounds__(SPMM_NT) void k_spmm_bsr4(int no, int nb, int nbc, double alpha,
                                                       const T *__restrict__ X, int64_t ldx,
                                                       const int *__restrict__ colptr, const int *__restrict__ rowidx,
                                                       const double *__restrict__ val, double beta,
                                                       T *__restrict__ Y, int64_t ldy, int rchunks, int jper) {
    constexpr int ROWS = SPMM_NT * VEC;
    const int njg = (nbc + jper - 1) / jper;
    const int total = rchunks * njg;
    int w = blockIdx.x;
    int rc, jg;
    if ((rchunks & 7) == 0 && (total & 7) == 0) {
        const int per = rchunks >> 3;          // state chunks per XCD
        const int xcd = w & 7, local = w >> 3;  // dispatch is round-robin over 8 XCDs
        rc = xcd * per + local % per;
        jg = local / per;
    } else {
        rc = w % rchunks;
        jg = w / rchunks;
    }
    const int i0 = rc * ROWS + threadIdx.x * VEC;
    const bool full = (i0 + VEC <= no);
    const bool any = (i0 < no);
    const int jbeg = jg * jper, jend = min(nbc, jbeg + jper);

    for (int J = jbeg; J < jend; ++J) {
        T acc[4][VEC];
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int q = 0; q < VEC; ++q) acc[c][q] = (T)0;
        const int lb = colptr[J], le = colptr[J + 1];
        for (int l = lb; l < le; ++l) {
            const int I = rowidx[l];
            const double *v = val + 16 * (int64_t)l;
            double bv[16];
#pragma unroll
            for (int e = 0; e < 16; ++e) bv[e] = v[e];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int kcol = 4 * I + r;
                if (kcol >= nb) break;  // partial last block (nb % 4 != 0)
                T x[VEC];
                const T *src = X + (int64_t)kcol * ldx + i0;
                if (full) {
                    vload<T, VEC>(src, x);
                } else {
#pragma unroll
                    for (int q = 0; q < VEC; ++q) x[q] = (any && i0 + q < no) ? src[q] : (T)0;
                }
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const T a = (T)bv[r + 4 * c];
#pragma unroll
                    for (int q = 0; q < VEC; ++q) acc[c][q] = fma(x[q], a, acc[c][q]);
                }
            }
        }
        if (!any) continue;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int jcol = 4 * J + c;
            if (jcol >= nb) break;
            T *dst = Y + (int64_t)jcol * ldy + i0;
            T out[VEC];
            if (full) {
                if (beta != 0.0) {
                    T old[VEC];
                    vload<T, VEC>(dst, old);
#pragma unroll
                    for (int q = 0; q < VEC; ++q) out[q] = (T)(alpha * (double)acc[c][q] + beta * (double)old[q]);
                } else {
#pragma unroll
                    for (int q = 0; q < VEC; ++q) out[q] = (T)(alpha * (double)acc[c][q]);
                }
                vstore<T, VEC>(dst, out);
            } else {
#pragma unroll
                for (int q = 0; q < VEC; ++q)
                    if (i0 + q < no) {
                        double o = alpha * (double)acc[c][q];
                        if (beta != 0.0) o += beta * (double)dst[q];
                        dst[q] = (T)o;
                    }
            }
        }
    }
}

template <typename T, int VEC>
static int run_spmm(omx_ctx ctx, int no, int nb, double alpha, const T *X, int64_t ldx, omx_bsr A, double beta, T *Y,
                    int64_t ldy) {
    constexpr int ROWS = SPMM_NT * VEC;
    int rchunks = (no + ROWS - 1) / ROWS;
    // Enough workgroups to cover the chip several times; each walks jper block columns.
    int64_t target = 16LL * ctx->num_cus;
    int jper = 1;
    while ((int64_t)rchunks * ((A->nbc + jper * 2 - 1) / (jper * 2)) >= target && jper < 64) jper *= 2;
    int njg = (A->nbc + jper - 1) / jper;
    int64_t total = (int64_t)rchunks * njg;
    OMX_REQUIRE(total < (1LL << 31), "omx_spmm: grid too large");
    hipEvent_t ev;
    prof_begin(ctx, OMX_K_SPMM, &ev);
    prof_work(ctx, OMX_K_SPMM, 2.0 * 16.0 * A->nnzb * no,
              sizeof(T) * 2.0 * (double)no * nb * (beta != 0.0 ? 1.5 : 1.0) + A->nnzb * (16.0 * 8 + 4) + 4.0 * (A->nbc + 1));
    hipLaunchKernelGGL((k_spmm_bsr4<T, VEC>), dim3((unsigned)total), dim3(SPMM_NT), 0, ctx->stream, no, nb, A->nbc,
                       alpha, X, ldx, A->d_colptr, A->d_rowidx, A->d_val, beta, Y, ldy, rchunks, jper);
    prof_end(ctx, OMX_K_SPMM, ev);
    OMX_CHECK_HIP(hipGetLastError());
    return OMX_OK;
}

static bool aligned16(const void *p) { return ((uintptr_t)p & 15) == 0; }

}  // namespace omx

using namespace omx;

extern "C" int omx_spmm(omx_ctx ctx, int dt, int no, int nb, double alpha, const void *X, int64_t ldx, omx_bsr A,
                        double beta, void *Y, int64_t ldy) {
    OMX_REQUIRE(ctx && A, "omx_spmm: null argument");
    OMX_REQUIRE(nb == A->nb, "omx_spmm: operator is %d x %d but X has %d columns", A->nb, A->nb, nb);
    OMX_REQUIRE(no >= 0 && ldx >= std::max(1, no) && ldy >= std::max(1, no), "omx_spmm: bad leading dimension");
    OMX_REQUIRE(!A->is_complex, "omx_spmm: complex operator needs a complex element type");
    if (no == 0) return OMX_OK;
    OMX_REQUIRE(X && Y, "omx_spmm: null operand");
    OMX_REQUIRE(X != Y, "omx_spmm: X and Y must not alias");
    if (dt == OMX_F64) {
        bool vec = (no % 2 == 0) && (ldx % 2 == 0) && (ldy % 2 == 0) && aligned16(X) && aligned16(Y);
        if (vec) return run_spmm<double, 2>(ctx, no, nb, alpha, (const double *)X, ldx, A, beta, (double *)Y, ldy);
        return run_spmm<double, 1>(ctx, no, nb, alpha, (const double *)X, ldx, A, beta, (double *)Y, ldy);
    }
    if (dt == OMX_F32) {
        bool vec = (no % 4 == 0) && (ldx % 4 == 0) && (ldy % 4 == 0) && aligned16(X) && aligned16(Y);
        if (vec) return run_spmm<float, 4>(ctx, no, nb, alpha, (const float *)X, ldx, A, beta, (float *)Y, ldy);
        return run_spmm<float, 1>(ctx, no, nb, alpha, (const float *)X, ldx, A, beta, (float *)Y, ldy);
    }
    set_error("omx_spmm: unsupported element type %d", dt);
    return OMX_ERR_UNSUPPORTED;
}
