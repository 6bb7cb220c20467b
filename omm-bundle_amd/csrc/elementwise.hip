// Memory-bound kernels of the OMM loop: MatrixSwitch m_add / m_scale / m_set, traces, the fused
// line-search coefficient reduction, and the fused CG vector updates.
//
// Reductions are two-stage and deterministic: every workgroup writes FP64 partial sums for its
// tile range into the context workspace, then one workgroup adds the partials in workgroup
// order and writes the result to the context's scalar slots (read back through pinned memory).
#include <algorithm>
#include <mutex>

#include "omx_internal.h"

namespace omx {

constexpr int EW_NT = 256;
constexpr int TT = 32;  // transpose / trace tile edge

template <typename T>
__device__ inline double ld(const T *p) { return (double)*p; }

// ---- block reduction helpers ----------------------------------------------------------------

template <int NV>
__device__ inline void block_reduce_store(double (&v)[NV], double *__restrict__ out) {
    __shared__ double red[NV][EW_NT / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        double s = v[q];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
        if (lane == 0) red[q][wave] = s;
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < EW_NT / 64; ++w) s += red[threadIdx.x][w];
        out[threadIdx.x] = s;
    }
}

// Final stage: out[q] = sum over blocks b (ascending) of partial[b*NV + q].
template <int NV>
__global__ __launch_bounds__(EW_NT) void k_reduce_partials(int nblocks, const double *__restrict__ partial,
                                                           double *__restrict__ out,
                                                           const int *__restrict__ mask = nullptr) {
    if (mask && !mask[blockIdx.x]) return;          // system masked off (whole workgroup)
    partial += (int64_t)blockIdx.x * nblocks * NV;  // one workgroup per system in batched use
    out += blockIdx.x * NV;
    double v[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += EW_NT)
#pragma unroll
        for (int q = 0; q < NV; ++q) v[q] += partial[(int64_t)b * NV + q];
    block_reduce_store<NV>(v, out);
}

// ---- m_add / m_scale / m_set -----------------------------------------------------------------

// Row and column (i, j) of flat index q in an m-row column-major walk, advanced by a fixed grid
// stride with one division per thread instead of a 64-bit division per element.
struct ColWalk {
    int64_t i, j, si, sj, m;
    __device__ ColWalk(int64_t q, int64_t stride, int64_t rows)
        : i(q % rows), j(q / rows), si(stride % rows), sj(stride / rows), m(rows) {}
    __device__ void next() {
        i += si;
        j += sj;
        if (i >= m) {
            i -= m;
            ++j;
        }
    }
};

template <typename T>
__global__ __launch_bounds__(EW_NT) void k_add_n(int m, int n, double alpha, const T *__restrict__ A, int64_t lda,
                                                 double beta, T *__restrict__ C, int64_t ldc) {
    typedef El<T> E;
    const int64_t total = (int64_t)m * n, stride = (int64_t)gridDim.x * EW_NT;
    ColWalk w((int64_t)blockIdx.x * EW_NT + threadIdx.x, stride, m);
    for (int64_t q = (int64_t)blockIdx.x * EW_NT + threadIdx.x; q < total; q += stride, w.next()) {
        const int64_t i = w.i, j = w.j;
        const T a = A[i + j * lda];
        T *p = C + i + j * ldc;
        double ar = alpha * E::re(a), ai = 0.0;
        if constexpr (E::cplx) ai = alpha * E::im(a);
        if (beta == 0.0) {
            *p = E::make(ar, ai);
        } else {
            const T c = *p;
            double cr = beta * E::re(c) + ar, ci = 0.0;
            if constexpr (E::cplx) ci = beta * E::im(c) + ai;
            *p = E::make(cr, ci);
        }
    }
}

// C(i,j) := beta*C(i,j) + alpha*op(A)(i,j), op(A)(i,j) = A(j,i) or conj(A(j,i)): 32x32 tiles
// through LDS.
template <typename T>
__global__ __launch_bounds__(EW_NT) void k_add_t(int m, int n, double alpha, const T *__restrict__ A, int64_t lda,
                                                 double beta, T *__restrict__ C, int64_t ldc, int conj) {
    typedef El<T> E;
    __shared__ T tile[TT][TT + 1];
    const int tx = threadIdx.x % TT, ty = threadIdx.x / TT;  // 32 x 8
    const int ti = blockIdx.x * TT, tj = blockIdx.y * TT;     // C tile origin
    // A is n x m; read A(tj + r, ti + c) for the C tile (ti.., tj..).
    for (int c = ty; c < TT; c += EW_NT / TT) {
        int ar = tj + tx, ac = ti + c;
        tile[c][tx] = (ar < n && ac < m) ? A[(int64_t)ar + (int64_t)ac * lda] : E::make(0.0, 0.0);
    }
    __syncthreads();
    for (int c = ty; c < TT; c += EW_NT / TT) {
        int i = ti + tx, j = tj + c;
        if (i < m && j < n) {
            const T a = tile[tx][c];
            double vr = alpha * E::re(a), vi = 0.0;
            if constexpr (E::cplx) vi = alpha * (conj ? -E::im(a) : E::im(a));
            T *p = C + (int64_t)i + (int64_t)j * ldc;
            if (beta != 0.0) {
                const T o = *p;
                vr = beta * E::re(o) + vr;
                if constexpr (E::cplx) vi = beta * E::im(o) + vi;
            }
            *p = E::make(vr, vi);
        }
    }
}

template <typename T>
__global__ __launch_bounds__(EW_NT) void k_scale(int m, int n, double beta, T *__restrict__ C, int64_t ldc) {
    typedef El<T> E;
    const int64_t total = (int64_t)m * n, stride = (int64_t)gridDim.x * EW_NT;
    ColWalk w((int64_t)blockIdx.x * EW_NT + threadIdx.x, stride, m);
    for (int64_t q = (int64_t)blockIdx.x * EW_NT + threadIdx.x; q < total; q += stride, w.next()) {
        const int64_t i = w.i, j = w.j;
        const T c = C[i + j * ldc];
        C[i + j * ldc] = E::make(beta * E::re(c), beta * E::im(c));
    }
}

template <typename T>
__global__ __launch_bounds__(EW_NT) void k_set(int lu, int m, int n, double alpha, double beta, T *__restrict__ C,
                                               int64_t ldc) {
    const int64_t total = (int64_t)m * n, stride = (int64_t)gridDim.x * EW_NT;
    ColWalk w((int64_t)blockIdx.x * EW_NT + threadIdx.x, stride, m);
    for (int64_t q = (int64_t)blockIdx.x * EW_NT + threadIdx.x; q < total; q += stride, w.next()) {
        const int64_t i = w.i, j = w.j;
        if (lu == 1 && j < i) continue;
        if (lu == 2 && j > i) continue;
        C[i + j * ldc] = El<T>::make(i != j ? alpha : beta, 0.0);
    }
}

// ---- traces ----------------------------------------------------------------------------------

// (re, im) of tr(A)
template <typename T>
__global__ __launch_bounds__(EW_NT) void k_trace_partial(int n, const T *__restrict__ A, int64_t lda,
                                                         double *__restrict__ partial) {
    double v[2] = {0.0, 0.0};
    for (int i = blockIdx.x * EW_NT + threadIdx.x; i < n; i += gridDim.x * EW_NT) {
        const T a = A[(int64_t)i * (lda + 1)];
        v[0] += El<T>::re(a);
        if constexpr (El<T>::cplx) v[1] += El<T>::im(a);
    }
    block_reduce_store<2>(v, partial + 2 * (int64_t)blockIdx.x);
}

// (re, im) of sum_ij conj(A_ij)*B_ij
template <typename T>
__global__ __launch_bounds__(EW_NT) void k_dot_partial(int m, int n, const T *__restrict__ A, int64_t lda,
                                                       const T *__restrict__ B, int64_t ldb,
                                                       double *__restrict__ partial) {
    typedef El<T> E;
    double v[2] = {0.0, 0.0};
    const int64_t total = (int64_t)m * n, stride = (int64_t)gridDim.x * EW_NT;
    ColWalk w((int64_t)blockIdx.x * EW_NT + threadIdx.x, stride, m);
    for (int64_t q = (int64_t)blockIdx.x * EW_NT + threadIdx.x; q < total; q += stride, w.next()) {
        const int64_t i = w.i, j = w.j;
        const T a = A[i + j * lda], b = B[i + j * ldb];
        if constexpr (E::cplx) {
            v[0] += E::re(a) * E::re(b) + E::im(a) * E::im(b);
            v[1] += E::re(a) * E::im(b) - E::im(a) * E::re(b);
        } else {
            v[0] += E::re(a) * E::re(b);
        }
    }
    block_reduce_store<2>(v, partial + 2 * (int64_t)blockIdx.x);
}

// Re <G, G - G_p> and <G_p, G_p> (Polak-Ribiere numerator and denominator; G - G_p rounded to
// the storage type first, as the reference's work2 copy).
template <typename T>
__global__ __launch_bounds__(EW_NT) void k_pr_partial(int64_t total_m, int m, const T *__restrict__ G,
                                                      const T *__restrict__ Gp, int64_t ld,
                                                      double *__restrict__ partial, int64_t sG = 0,
                                                      const int *__restrict__ mask = nullptr) {
    typedef El<T> E;
    if (mask && !mask[blockIdx.y]) return;
    // batched use: blockIdx.y selects the system
    G += blockIdx.y * sG;
    Gp += blockIdx.y * sG;
    partial += 2 * (int64_t)blockIdx.y * gridDim.x;
    double v[2] = {0.0, 0.0};
    const int64_t stride = (int64_t)gridDim.x * EW_NT;
    ColWalk w((int64_t)blockIdx.x * EW_NT + threadIdx.x, stride, m);
    for (int64_t q = (int64_t)blockIdx.x * EW_NT + threadIdx.x; q < total_m; q += stride, w.next()) {
        const int64_t i = w.i, j = w.j;
        const T gt = G[i + j * ld], pt = Gp[i + j * ld];
        const double g = E::re(gt), p = E::re(pt);
        const T wt = E::make(g - p, E::im(gt) - E::im(pt));
        v[0] += g * E::re(wt);
        v[1] += p * p;
        if constexpr (E::cplx) {
            const double gi = E::im(gt), pi = E::im(pt);
            v[0] += gi * E::im(wt);
            v[1] += pi * pi;
        }
    }
    block_reduce_store<2>(v, partial + 2 * (int64_t)blockIdx.x);
}

// ---- fused line-search coefficients (calc_coeff, libOMM/src/omm_ops.F90:156-219) --------------
// Thirteen sums over n x n matrices (real parts; mm_trace(A,B) = Re sum conj(A_ij) B_ij); tile
// (I,J) also reads the SWd tile (J,I) for the transposed term TrHdSd = Re sum_ij SWd(j,i)*HWd(i,j).
// Partial layout per block: [TrH, TrHS, TrHd, TrHdS, TrHSd, TrHdd, TrHddS, TrHSdd, TrHdSd,
//                            TrHdSdH, TrHddSd, TrHdSdd, TrHddSdd]
template <typename T>
__device__ inline double cdot(const T &a, const T &b) {
    if constexpr (El<T>::cplx)
        return El<T>::re(a) * El<T>::re(b) + El<T>::im(a) * El<T>::im(b);
    else
        return El<T>::re(a) * El<T>::re(b);
}

template <typename T>
__global__ __launch_bounds__(EW_NT) void k_coeff_partial(int n, int64_t ldm, const T *__restrict__ HW,
                                                         const T *__restrict__ SW, const T *__restrict__ HWd,
                                                         const T *__restrict__ SWd, const T *__restrict__ HWdd,
                                                         const T *__restrict__ SWdd, double *__restrict__ partial,
                                                         int64_t sW = 0, const int *__restrict__ mask = nullptr) {
    typedef El<T> E;
    if (mask && !mask[blockIdx.y]) return;  // whole workgroup, before the barrier
    // batched use: blockIdx.y selects the system (all six matrices offset by sW elements)
    {
        const int64_t o = blockIdx.y * sW;
        HW += o;
        SW += o;
        HWd += o;
        SWd += o;
        HWdd += o;
        SWdd += o;
        partial += 13 * (int64_t)blockIdx.y * gridDim.x;
    }
    __shared__ double swdT[TT][TT + 1];
    __shared__ double swdTi[El<T>::cplx ? TT : 1][TT + 1];
    const int tx = threadIdx.x % TT, ty = threadIdx.x / TT;
    const int tiles = (n + TT - 1) / TT;
    const int I = blockIdx.x % tiles, J = blockIdx.x / tiles;
    const int i0 = I * TT, j0 = J * TT;
    // SWd tile (J,I): element (j0 + a, i0 + b) -> swdT[a][b]
    for (int b = ty; b < TT; b += EW_NT / TT) {
        int r = j0 + tx, c = i0 + b;
        const bool in = (r < n && c < n);
        const T w = in ? SWd[(int64_t)r + (int64_t)c * ldm] : E::make(0.0, 0.0);
        swdT[tx][b] = E::re(w);
        if constexpr (E::cplx) swdTi[tx][b] = E::im(w);
    }
    __syncthreads();
    double v[13];
#pragma unroll
    for (int q = 0; q < 13; ++q) v[q] = 0.0;
    for (int c = ty; c < TT; c += EW_NT / TT) {
        const int i = i0 + tx, j = j0 + c;
        if (i >= n || j >= n) continue;
        const int64_t o = (int64_t)i + (int64_t)j * ldm;
        const T hw = HW[o], sw = SW[o], hwd = HWd[o], swd = SWd[o], hwdd = HWdd[o], swdd = SWdd[o];
        if (i == j) {
            v[0] += E::re(hw);
            v[2] += E::re(hwd);
            v[5] += E::re(hwdd);
        }
        v[1] += cdot(hw, sw);
        v[3] += cdot(sw, hwd);
        v[4] += cdot(hw, swd);
        v[6] += cdot(hwdd, sw);
        v[7] += cdot(hw, swdd);
        if constexpr (E::cplx)
            v[8] += swdT[c][tx] * E::re(hwd) - swdTi[c][tx] * E::im(hwd);  // Re SWd(j,i) * HWd(i,j)
        else
            v[8] += swdT[c][tx] * E::re(hwd);
        v[9] += cdot(swd, hwd);
        v[10] += cdot(hwdd, swd);
        v[11] += cdot(swdd, hwd);
        v[12] += cdot(hwdd, swdd);
    }
    block_reduce_store<13>(v, partial + 13 * (int64_t)blockIdx.x);
}

__global__ void k_coeff_final(int nblocks, const double *__restrict__ partial, double *__restrict__ out,
                              const int *__restrict__ mask = nullptr) {
    // One workgroup (per system in batched use): deterministic per-term sums, then the closed-form
    // coefficients.
    if (mask && !mask[blockIdx.x]) return;
    partial += 13 * (int64_t)blockIdx.x * nblocks;
    out += 18 * blockIdx.x;
    double v[13];
#pragma unroll
    for (int q = 0; q < 13; ++q) v[q] = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += EW_NT)
#pragma unroll
        for (int q = 0; q < 13; ++q) v[q] += partial[(int64_t)b * 13 + q];
    __shared__ double tr[13];
    block_reduce_store<13>(v, tr);
    __syncthreads();
    if (threadIdx.x == 0) {
        const double TrH = tr[0], TrHS = tr[1], TrHd = tr[2], TrHdS = tr[3], TrHSd = tr[4], TrHdd = tr[5],
                     TrHddS = tr[6], TrHSdd = tr[7], TrHdSd = tr[8], TrHdSdH = tr[9], TrHddSd = tr[10],
                     TrHdSdd = tr[11], TrHddSdd = tr[12];
        out[0] = 2.0 * TrH - TrHS;
        out[1] = 2.0 * (2.0 * TrHd - TrHdS - TrHSd);
        out[2] = 2.0 * (TrHdd - TrHdSd - TrHdSdH) - TrHddS - TrHSdd;
        out[3] = -2.0 * (TrHddSd + TrHdSdd);
        out[4] = -TrHddSdd;
        for (int q = 0; q < 13; ++q) out[5 + q] = tr[q];
    }
}

// ---- fused CG updates ------------------------------------------------------------------------

// D := G + lambda*D; G_p := G  (omm.F90:440-445)
template <typename T>
__global__ __launch_bounds__(EW_NT) void k_cg_direction(int m, int n, double lambda, const T *__restrict__ G,
                                                        T *__restrict__ D, T *__restrict__ Gp, int64_t ld) {
    typedef El<T> E;
    const int64_t total = (int64_t)m * n, stride = (int64_t)gridDim.x * EW_NT;
    ColWalk w((int64_t)blockIdx.x * EW_NT + threadIdx.x, stride, m);
    for (int64_t q = (int64_t)blockIdx.x * EW_NT + threadIdx.x; q < total; q += stride, w.next()) {
        const int64_t o = w.i + w.j * ld;
        const T gt = G[o];
        const double g = E::re(gt), gi = E::im(gt);
        double d, di = 0.0;
        if (lambda == 0.0) {
            d = g;
            di = gi;
        } else {
            const T dt = D[o];
            d = lambda * E::re(dt) + g;
            if constexpr (E::cplx) di = lambda * E::im(dt) + gi;
        }
        D[o] = E::make(d, di);
        Gp[o] = gt;
    }
}

// W := W + x*Wd^H + x*Wd + x^2*Wdd, accumulated in that order (omm.F90:499-501, 514-516).
template <typename T>
__global__ __launch_bounds__(EW_NT) void k_update_w(int n, int64_t ldm, double x, const T *__restrict__ Wd,
                                                    const T *__restrict__ Wdd, T *__restrict__ W) {
    typedef El<T> E;
    __shared__ double tileT[TT][TT + 1];
    __shared__ double tileTi[El<T>::cplx ? TT : 1][TT + 1];
    const int tx = threadIdx.x % TT, ty = threadIdx.x / TT;
    const int i0 = blockIdx.x * TT, j0 = blockIdx.y * TT;
    for (int b = ty; b < TT; b += EW_NT / TT) {
        int r = j0 + tx, c = i0 + b;  // Wd(j0+tx, i0+b)
        const bool in = (r < n && c < n);
        const T w = in ? Wd[(int64_t)r + (int64_t)c * ldm] : E::make(0.0, 0.0);
        tileT[tx][b] = E::re(w);
        if constexpr (E::cplx) tileTi[tx][b] = -E::im(w);  // conjugate transpose
    }
    __syncthreads();
    const double x2 = x * x;
    for (int c = ty; c < TT; c += EW_NT / TT) {
        int i = i0 + tx, j = j0 + c;
        if (i >= n || j >= n) continue;
        int64_t o = (int64_t)i + (int64_t)j * ldm;
        const T wt = W[o], wdt = Wd[o], wddt = Wdd[o];
        double w = E::re(wt);
        w = w + x * tileT[c][tx];
        w = w + x * E::re(wdt);
        w = w + x2 * E::re(wddt);
        double wi = 0.0;
        if constexpr (E::cplx) {
            wi = E::im(wt);
            wi = wi + x * tileTi[c][tx];
            wi = wi + x * E::im(wdt);
            wi = wi + x2 * E::im(wddt);
        }
        W[o] = E::make(w, wi);
    }
}

// Calls f(T{}) with T the element type of dt; false for an unknown dt.
template <typename F>
inline bool with_type(int dt, F &&f) {
    switch (dt) {
        case OMX_F64: f(double{}); return true;
        case OMX_F32: f(float{}); return true;
        case OMX_C128: f(Cx<double>{}); return true;
        case OMX_C64: f(Cx<float>{}); return true;
        default: return false;
    }
}

// Batched complex128 reductions over nbatch systems (used by the batched solver): coefficients to
// out[18*s ...] (5 coefficients then the 13 traces) and Polak-Ribiere dots to out2[2*s ...], on
// the device, no readback.
int launch_coeff_batched(omx_ctx ctx, int dt, int n, int64_t ldm, int64_t sW, const void *HW, const void *SW,
                         const void *HWd, const void *SWd, const void *HWdd, const void *SWdd, int nbatch,
                         double *partial, double *out, const int *mask);
int launch_pr_batched(omx_ctx ctx, int dt, int n, int m, int64_t ld, int64_t sG, const void *G, const void *Gp,
                      int nbatch, double *partial, double *out2, const int *mask);
size_t coeff_partial_doubles(int n, int nbatch);
size_t pr_partial_doubles(int n, int m, int nbatch);

inline int grid_for(int64_t total) { return (int)std::max<int64_t>(1, std::min<int64_t>((total + EW_NT - 1) / EW_NT, 65536)); }

// Two-stage sum: partial kernel already launched with `blocks` workgroups and NV values each.
template <int NV>
static int finish_reduce(omx_ctx ctx, int blocks, double *partial) {
    hipLaunchKernelGGL((k_reduce_partials<NV>), dim3(1), dim3(EW_NT), 0, ctx->stream, blocks, partial, ctx->d_scalars);
    OMX_CHECK_HIP(hipGetLastError());
    return read_scalars(ctx, NV);
}

// ---- scratch of the allocation-free Polak-Ribiere dots (omx_omm_pr_dots_noalloc) ---------------
// The partial sums of at most kPrBlocks workgroups (2 each), then the 2 final sums: 32 KB per GPU,
// held until the process ends.  In host code `scratch` only names the symbol; the device address
// of each GPU's copy comes from hipGetSymbolAddress (noalloc_setup) and is what kernels receive.
constexpr int kPrBlocks = 2048;
constexpr int kScratchDoubles = 2 * kPrBlocks + 2;
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[kMaxDevices] = {};

// Once per device, with that device current (called by omx_init): look up this GPU's copy.
int noalloc_setup(int device) {
    OMX_REQUIRE(device >= 0 && device < kMaxDevices, "omx_init: device %d beyond %d", device, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return OMX_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        set_error("omx_init: hipGetSymbolAddress(scratch) failed on device %d: %s", device, hipGetErrorString(e));
        return OMX_ERR_HIP;
    }
    g_scratch_ptr[device] = (double *)p;
    return OMX_OK;
}

static double *noalloc_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch_ptr[device];
}

}  // namespace omx

using namespace omx;

extern "C" {

int omx_add(omx_ctx ctx, int dt, char opA, int m, int n, double alpha, const void *A, int64_t lda, double beta,
            void *C, int64_t ldc) {
    OMX_REQUIRE(ctx && valid_op(opA) && m >= 0 && n >= 0, "omx_add: bad arguments");
    bool tr = is_trans(opA);
    const int conj = (opA == 'c' || opA == 'C');
    OMX_REQUIRE(ldc >= std::max(1, m) && lda >= std::max(1, tr ? n : m), "omx_add: bad leading dimension");
    if (!m || !n) return OMX_OK;
    OMX_REQUIRE(A && C, "omx_add: null operand");
    OMX_REQUIRE(!(tr && A == C), "omx_add: in-place transpose is not supported");
    hipEvent_t ev;
    prof_begin(ctx, OMX_K_ELEM, &ev);
    const bool ok = with_type(dt, [&](auto tag) {
        typedef decltype(tag) T;
        if (!tr) {
            hipLaunchKernelGGL(k_add_n<T>, dim3(grid_for((int64_t)m * n)), dim3(EW_NT), 0, ctx->stream, m, n, alpha,
                               (const T *)A, lda, beta, (T *)C, ldc);
        } else {
            dim3 grid((m + TT - 1) / TT, (n + TT - 1) / TT);
            hipLaunchKernelGGL(k_add_t<T>, grid, dim3(EW_NT), 0, ctx->stream, m, n, alpha, (const T *)A, lda, beta,
                               (T *)C, ldc, conj);
        }
    });
    prof_end(ctx, OMX_K_ELEM, ev);
    if (!ok) {
        set_error("omx_add: unsupported element type %d", dt);
        return OMX_ERR_UNSUPPORTED;
    }
    OMX_CHECK_HIP(hipGetLastError());
    return OMX_OK;
}

int omx_scale(omx_ctx ctx, int dt, int m, int n, double beta, void *C, int64_t ldc) {
    OMX_REQUIRE(ctx && m >= 0 && n >= 0 && ldc >= std::max(1, m), "omx_scale: bad arguments");
    if (!m || !n) return OMX_OK;
    OMX_REQUIRE(C, "omx_scale: null operand");
    const bool ok = with_type(dt, [&](auto tag) {
        typedef decltype(tag) T;
        hipLaunchKernelGGL(k_scale<T>, dim3(grid_for((int64_t)m * n)), dim3(EW_NT), 0, ctx->stream, m, n, beta,
                           (T *)C, ldc);
    });
    if (!ok) {
        set_error("omx_scale: unsupported element type %d", dt);
        return OMX_ERR_UNSUPPORTED;
    }
    OMX_CHECK_HIP(hipGetLastError());
    return OMX_OK;
}

int omx_set(omx_ctx ctx, int dt, char seC, int m, int n, double alpha, double beta, void *C, int64_t ldc) {
    OMX_REQUIRE(ctx && m >= 0 && n >= 0 && ldc >= std::max(1, m), "omx_set: bad arguments");
    if (!m || !n) return OMX_OK;
    OMX_REQUIRE(C, "omx_set: null operand");
    int lu = (seC == 'l' || seC == 'L') ? 2 : ((seC == 'u' || seC == 'U') ? 1 : 0);
    const bool ok = with_type(dt, [&](auto tag) {
        typedef decltype(tag) T;
        hipLaunchKernelGGL(k_set<T>, dim3(grid_for((int64_t)m * n)), dim3(EW_NT), 0, ctx->stream, lu, m, n, alpha,
                           beta, (T *)C, ldc);
    });
    if (!ok) {
        set_error("omx_set: unsupported element type %d", dt);
        return OMX_ERR_UNSUPPORTED;
    }
    OMX_CHECK_HIP(hipGetLastError());
    return OMX_OK;
}

static inline bool is_complex_dt(int dt) { return dt == OMX_C128 || dt == OMX_C64; }

int omx_trace(omx_ctx ctx, int dt, int n, const void *A, int64_t lda, double *out) {
    OMX_REQUIRE(ctx && out && n >= 0 && lda >= std::max(1, n), "omx_trace: bad arguments");
    if (!n) {
        out[0] = 0.0;
        if (is_complex_dt(dt)) out[1] = 0.0;
        return OMX_OK;
    }
    OMX_REQUIRE(A, "omx_trace: null operand");
    int blocks = std::min(256, (n + EW_NT - 1) / EW_NT);
    int rc = ensure_ws(ctx, sizeof(double) * 2 * blocks);
    if (rc) return rc;
    double *partial = (double *)ctx->ws;
    const bool ok = with_type(dt, [&](auto tag) {
        typedef decltype(tag) T;
        hipLaunchKernelGGL(k_trace_partial<T>, dim3(blocks), dim3(EW_NT), 0, ctx->stream, n, (const T *)A, lda,
                           partial);
    });
    if (!ok) {
        set_error("omx_trace: unsupported element type %d", dt);
        return OMX_ERR_UNSUPPORTED;
    }
    OMX_CHECK_HIP(hipGetLastError());
    rc = finish_reduce<2>(ctx, blocks, partial);
    if (rc) return rc;
    out[0] = ctx->h_scalars[0];
    if (is_complex_dt(dt)) out[1] = ctx->h_scalars[1];
    return OMX_OK;
}

int omx_mm_trace(omx_ctx ctx, int dt, int m, int n, const void *A, int64_t lda, const void *B, int64_t ldb,
                 double *out) {
    OMX_REQUIRE(ctx && out && m >= 0 && n >= 0 && lda >= std::max(1, m) && ldb >= std::max(1, m),
                "omx_mm_trace: bad arguments");
    if (!m || !n) {
        out[0] = 0.0;
        if (is_complex_dt(dt)) out[1] = 0.0;
        return OMX_OK;
    }
    OMX_REQUIRE(A && B, "omx_mm_trace: null operand");
    int blocks = std::min(2048, grid_for((int64_t)m * n));
    int rc = ensure_ws(ctx, sizeof(double) * 2 * blocks);
    if (rc) return rc;
    double *partial = (double *)ctx->ws;
    hipEvent_t ev;
    prof_begin(ctx, OMX_K_DOT, &ev);
    const bool ok = with_type(dt, [&](auto tag) {
        typedef decltype(tag) T;
        hipLaunchKernelGGL(k_dot_partial<T>, dim3(blocks), dim3(EW_NT), 0, ctx->stream, m, n, (const T *)A, lda,
                           (const T *)B, ldb, partial);
    });
    prof_end(ctx, OMX_K_DOT, ev);
    if (!ok) {
        set_error("omx_mm_trace: unsupported element type %d", dt);
        return OMX_ERR_UNSUPPORTED;
    }
    OMX_CHECK_HIP(hipGetLastError());
    rc = finish_reduce<2>(ctx, blocks, partial);
    if (rc) return rc;
    out[0] = ctx->h_scalars[0];
    if (is_complex_dt(dt)) out[1] = ctx->h_scalars[1];
    return OMX_OK;
}

int omx_omm_coeff(omx_ctx ctx, int dt, int n, int64_t ldm, const void *HW, const void *SW, const void *HWd,
                  const void *SWd, const void *HWdd, const void *SWdd, double *coeff) {
    OMX_REQUIRE(ctx && coeff && n > 0 && ldm >= n, "omx_omm_coeff: bad arguments");
    OMX_REQUIRE(HW && SW && HWd && SWd && HWdd && SWdd, "omx_omm_coeff: null operand");
    int tiles = (n + TT - 1) / TT;
    int64_t blocks = (int64_t)tiles * tiles;
    OMX_REQUIRE(blocks < (1LL << 31), "omx_omm_coeff: n too large");
    int rc = ensure_ws(ctx, sizeof(double) * 13 * blocks);
    if (rc) return rc;
    double *partial = (double *)ctx->ws;
    hipEvent_t ev;
    prof_begin(ctx, OMX_K_COEFF, &ev);
    prof_work(ctx, OMX_K_COEFF, (is_complex_dt(dt) ? 4.0 : 1.0) * 26.0 * n * (double)n,
              dt_size(dt) * 6.0 * n * (double)n);
    const bool ok = with_type(dt, [&](auto tag) {
        typedef decltype(tag) T;
        hipLaunchKernelGGL(k_coeff_partial<T>, dim3((unsigned)blocks), dim3(EW_NT), 0, ctx->stream, n, ldm,
                           (const T *)HW, (const T *)SW, (const T *)HWd, (const T *)SWd, (const T *)HWdd,
                           (const T *)SWdd, partial);
    });
    if (!ok) {
        prof_end(ctx, OMX_K_COEFF, ev);
        set_error("omx_omm_coeff: unsupported element type %d", dt);
        return OMX_ERR_UNSUPPORTED;
    }
    hipLaunchKernelGGL(k_coeff_final, dim3(1), dim3(EW_NT), 0, ctx->stream, (int)blocks, partial, ctx->d_scalars);
    prof_end(ctx, OMX_K_COEFF, ev);
    OMX_CHECK_HIP(hipGetLastError());
    rc = read_scalars(ctx, 18);
    if (rc) return rc;
    for (int q = 0; q < 5; ++q) coeff[q] = ctx->h_scalars[q];
    return OMX_OK;
}

int omx_omm_pr_dots(omx_ctx ctx, int dt, int n, int m, const void *G, const void *G_p, int64_t ld, double *out) {
    OMX_REQUIRE(ctx && out && n > 0 && m > 0 && ld >= n && G && G_p, "omx_omm_pr_dots: bad arguments");
    int64_t total = (int64_t)n * m;
    int blocks = std::min(2048, grid_for(total));
    int rc = ensure_ws(ctx, sizeof(double) * 2 * blocks);
    if (rc) return rc;
    double *partial = (double *)ctx->ws;
    hipEvent_t ev;
    prof_begin(ctx, OMX_K_DOT, &ev);
    const bool ok = with_type(dt, [&](auto tag) {
        typedef decltype(tag) T;
        hipLaunchKernelGGL(k_pr_partial<T>, dim3(blocks), dim3(EW_NT), 0, ctx->stream, total, n, (const T *)G,
                           (const T *)G_p, ld, partial);
    });
    prof_end(ctx, OMX_K_DOT, ev);
    if (!ok) {
        set_error("omx_omm_pr_dots: unsupported element type %d", dt);
        return OMX_ERR_UNSUPPORTED;
    }
    OMX_CHECK_HIP(hipGetLastError());
    rc = finish_reduce<2>(ctx, blocks, partial);
    if (rc) return rc;
    out[0] = ctx->h_scalars[0];
    out[1] = ctx->h_scalars[1];
    return OMX_OK;
}

int omx_omm_pr_dots_noalloc(omx_ctx ctx, int dt, int n, int m, const void *G, const void *G_p, int64_t ld,
                            double *out) {
    OMX_REQUIRE(ctx && out && n > 0 && m > 0 && ld >= n && G && G_p, "omx_omm_pr_dots_noalloc: bad arguments");
    double *sp = noalloc_scratch(ctx->device);
    OMX_REQUIRE(sp, "omx_omm_pr_dots_noalloc: no scratch on device %d (context not initialised)", ctx->device);
    const int64_t total = (int64_t)n * m;
    // the same workgroup count as omx_omm_pr_dots (so the same sums, bit for bit); it never exceeds
    // the scratch's kPrBlocks partial slots, whatever the size
    const int blocks = std::min(2048, grid_for(total));
    OMX_REQUIRE(blocks >= 1 && blocks <= kPrBlocks,
                "omx_omm_pr_dots_noalloc: %d workgroups exceed the scratch's %d partial slots", blocks, kPrBlocks);
    const bool ok = with_type(dt, [&](auto tag) {
        typedef decltype(tag) T;
        hipLaunchKernelGGL(k_pr_partial<T>, dim3(blocks), dim3(EW_NT), 0, ctx->stream, total, n, (const T *)G,
                           (const T *)G_p, ld, sp);
    });
    if (!ok) {
        set_error("omx_omm_pr_dots_noalloc: unsupported element type %d", dt);
        return OMX_ERR_UNSUPPORTED;
    }
    hipLaunchKernelGGL((k_reduce_partials<2>), dim3(1), dim3(EW_NT), 0, ctx->stream, blocks, sp, sp + 2 * kPrBlocks);
    OMX_CHECK_HIP(hipGetLastError());
    OMX_CHECK_HIP(hipMemcpyAsync(out, sp + 2 * kPrBlocks, sizeof(double) * 2, hipMemcpyDeviceToHost, ctx->stream));
    OMX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    return OMX_OK;
}

int omx_omm_cg_direction(omx_ctx ctx, int dt, int n, int m, double lambda, const void *G, void *D, void *G_p,
                         int64_t ld) {
    OMX_REQUIRE(ctx && n > 0 && m > 0 && ld >= n && G && D && G_p, "omx_omm_cg_direction: bad arguments");
    int grid = grid_for((int64_t)n * m);
    hipEvent_t ev;
    prof_begin(ctx, OMX_K_ELEM, &ev);
    const bool ok = with_type(dt, [&](auto tag) {
        typedef decltype(tag) T;
        hipLaunchKernelGGL(k_cg_direction<T>, dim3(grid), dim3(EW_NT), 0, ctx->stream, n, m, lambda, (const T *)G,
                           (T *)D, (T *)G_p, ld);
    });
    prof_end(ctx, OMX_K_ELEM, ev);
    if (!ok) {
        set_error("omx_omm_cg_direction: unsupported element type %d", dt);
        return OMX_ERR_UNSUPPORTED;
    }
    OMX_CHECK_HIP(hipGetLastError());
    return OMX_OK;
}

int omx_omm_update_w(omx_ctx ctx, int dt, int n, int64_t ld, double x, const void *Wd, const void *Wdd, void *W) {
    OMX_REQUIRE(ctx && n > 0 && ld >= n && Wd && Wdd && W, "omx_omm_update_w: bad arguments");
    OMX_REQUIRE(W != Wd && W != Wdd, "omx_omm_update_w: W must not alias Wd or Wdd");
    dim3 grid((n + TT - 1) / TT, (n + TT - 1) / TT);
    hipEvent_t ev;
    prof_begin(ctx, OMX_K_ELEM, &ev);
    const bool ok = with_type(dt, [&](auto tag) {
        typedef decltype(tag) T;
        hipLaunchKernelGGL(k_update_w<T>, grid, dim3(EW_NT), 0, ctx->stream, n, ld, x, (const T *)Wd, (const T *)Wdd,
                           (T *)W);
    });
    prof_end(ctx, OMX_K_ELEM, ev);
    if (!ok) {
        set_error("omx_omm_update_w: unsupported element type %d", dt);
        return OMX_ERR_UNSUPPORTED;
    }
    OMX_CHECK_HIP(hipGetLastError());
    return OMX_OK;
}

}  // extern "C"

namespace omx {

size_t coeff_partial_doubles(int n, int nbatch) {
    const int64_t tiles = (n + TT - 1) / TT;
    return (size_t)13 * tiles * tiles * nbatch;
}
size_t pr_partial_doubles(int n, int m, int nbatch) {
    return (size_t)2 * std::min(2048, grid_for((int64_t)n * m)) * nbatch;
}

int launch_coeff_batched(omx_ctx ctx, int dt, int n, int64_t ldm, int64_t sW, const void *HW, const void *SW,
                         const void *HWd, const void *SWd, const void *HWdd, const void *SWdd, int nbatch,
                         double *partial, double *out, const int *mask) {
    const int tiles = (n + TT - 1) / TT;
    const int blocks = tiles * tiles;
    const bool ok = with_type(dt, [&](auto tag) {
        typedef decltype(tag) T;
        hipLaunchKernelGGL(k_coeff_partial<T>, dim3((unsigned)blocks, nbatch), dim3(EW_NT), 0, ctx->stream, n, ldm,
                           (const T *)HW, (const T *)SW, (const T *)HWd, (const T *)SWd, (const T *)HWdd,
                           (const T *)SWdd, partial, sW, mask);
    });
    OMX_REQUIRE(ok, "batched coefficients: unsupported element type %d", dt);
    hipLaunchKernelGGL(k_coeff_final, dim3(nbatch), dim3(EW_NT), 0, ctx->stream, blocks, partial, out, mask);
    OMX_CHECK_HIP(hipGetLastError());
    return OMX_OK;
}

int launch_pr_batched(omx_ctx ctx, int dt, int n, int m, int64_t ld, int64_t sG, const void *G, const void *Gp,
                      int nbatch, double *partial, double *out2, const int *mask) {
    const int64_t total = (int64_t)n * m;
    const int blocks = std::min(2048, grid_for(total));
    const bool ok = with_type(dt, [&](auto tag) {
        typedef decltype(tag) T;
        hipLaunchKernelGGL(k_pr_partial<T>, dim3(blocks, nbatch), dim3(EW_NT), 0, ctx->stream, total, n,
                           (const T *)G, (const T *)Gp, ld, partial, sG, mask);
    });
    OMX_REQUIRE(ok, "batched dots: unsupported element type %d", dt);
    hipLaunchKernelGGL((k_reduce_partials<2>), dim3(nbatch), dim3(EW_NT), 0, ctx->stream, blocks, partial, out2, mask);
    OMX_CHECK_HIP(hipGetLastError());
    return OMX_OK;
}

}  // namespace omx
