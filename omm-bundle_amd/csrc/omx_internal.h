// Internal definitions shared by the libommhip.so translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "omm_hip.h"

namespace omx {

// Per-thread last error message (omx_last_error()).
void set_error(const char *fmt, ...);
void clear_error();

#define OMX_CHECK_HIP(expr)                                                                    \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess) {                                                                \
            omx::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__,    \
                           __LINE__);                                                          \
            return OMX_ERR_HIP;                                                                \
        }                                                                                      \
    } while (0)

#define OMX_REQUIRE(cond, ...)                                                                 \
    do {                                                                                       \
        if (!(cond)) {                                                                         \
            omx::set_error(__VA_ARGS__);                                                       \
            return OMX_ERR_ARG;                                                                \
        }                                                                                      \
    } while (0)

// Event pair recorded around one launch of a timed kernel class.
struct ProfEvent {
    hipEvent_t start, stop;
    int kclass;
};

}  // namespace omx

// Launch shape of a real SpMM: blocks whose loads are issued together (1, 4, 8), waves stacked in
// rows per workgroup (1, 2, 4), block columns walked per workgroup, XCD map (0: each XCD owns
// contiguous rows; 1: each XCD walks one row chunk at a time).
struct omx_spmm_shape_s {
    int batch = 4;
    int wpg = 4;
    int jper = 1;
    int mode = 1;
};

struct omx_ctx_s {
    int device = 0;
    hipStream_t stream = nullptr;
    // Device workspace for split-K partial products and reduction partials (grown on demand).
    void *ws = nullptr;
    size_t ws_bytes = 0;
    // Small device buffer for reduction results and pinned host mirror for scalar readback.
    double *d_scalars = nullptr;
    double *h_scalars = nullptr;
    // Event timing.
    bool prof_on = false;
    std::vector<omx::ProfEvent> prof_pending;
    std::vector<hipEvent_t> prof_pool;
    double prof_ms[OMX_K_COUNT] = {0};
    int64_t prof_count[OMX_K_COUNT] = {0};
    double prof_flops[OMX_K_COUNT] = {0};  // algorithmic flops of the timed launches
    double prof_bytes[OMX_K_COUNT] = {0};  // algorithmic (minimum) HBM bytes of the timed launches
    int num_cus = 256;
    int gemm_cfg = -1;    // forced GEMM tile configuration (-1: chosen by shape)
    bool gemm_sym = false;  // set by omx_gemm_sym around its product: lower-triangle tiles only
    omx_spmm_shape_s spmm_shape;  // forced SpMM launch shape (when spmm_forced)
    bool spmm_forced = false;
    bool spmm_stencil = false;  // lattice operators use the 2.5-D stencil SpMM (omx_spmm_stencil)
    // Device flag for the grouped solver: while set, GEMM, SpMM and reduction launches do nothing
    // when *launch_mask == 0, and GEMMs do not split K (no workspace, so launches can be captured).
    const int *launch_mask = nullptr;
    // autotuner: per-shape choices (configuration or SpMM launch shape, best time in ms)
    bool autotune = false;
    std::map<std::string, std::pair<int, double>> tune_gemm;
    std::map<std::string, std::pair<omx_spmm_shape_s, double>> tune_spmm;
};

struct omx_bsr_s {
    omx_ctx ctx = nullptr;
    int nb = 0;        // basis size
    int nbc = 0;       // block columns = ceil(nb/4)
    int64_t nnzb = 0;  // stored 4x4 blocks
    int is_complex = 0;
    int max_blocks_per_col = 0;
    int *d_colptr = nullptr;  // nbc+1
    int *d_rowidx = nullptr;  // nnzb
    double *d_val = nullptr;  // 16*nnzb (x2 if complex)
    int *d_walk = nullptr;    // optional order in which SpMMs visit the block columns (nbc)
    // Lx, Ly, Lz when the pattern is the periodic 7-point stencil of that lattice of 4x4 blocks
    // (recognised at upload; 0 otherwise)
    int lat[3] = {0, 0, 0};
    // LDS-staged gathers: block columns taken `group` at a time in walk order; group g gathers the
    // distinct block rows uidx[gptr[g] .. gptr[g+1]) and block l reads slot[l] of them
    int group = 0, ngroups = 0;
    int *d_gptr = nullptr, *d_uidx = nullptr;
    unsigned char *d_slot = nullptr;
};

namespace omx {

constexpr int kScalarSlots = 256;

// ---- element types -----------------------------------------------------------------------------
// Real (double, float) and complex (Cx<double>, Cx<float>: the Fortran (re, im) layout).  The
// element-wise kernels are written once over El<T>; for real types the imaginary terms are
// compiled out, so the real arithmetic (and its rounding) is exactly the real-only one.

template <typename R_>
struct Cx {
    R_ x, y;
};
template <typename T>
struct El {
    static constexpr bool cplx = false;
    __device__ static inline double re(const T &v) { return (double)v; }
    __device__ static inline double im(const T &) { return 0.0; }
    __device__ static inline T make(double r, double) { return (T)r; }
};
template <typename R_>
struct El<Cx<R_>> {
    static constexpr bool cplx = true;
    __device__ static inline double re(const Cx<R_> &v) { return (double)v.x; }
    __device__ static inline double im(const Cx<R_> &v) { return (double)v.y; }
    __device__ static inline Cx<R_> make(double r, double i) { return Cx<R_>{(R_)r, (R_)i}; }
};
template <>
struct El<double2> {
    static constexpr bool cplx = true;
    __device__ static inline double re(const double2 &v) { return v.x; }
    __device__ static inline double im(const double2 &v) { return v.y; }
    __device__ static inline double2 make(double r, double i) { return make_double2(r, i); }
};

// Ensure ctx->ws holds at least `bytes`.  Returns OMX_OK or an error.
int ensure_ws(omx_ctx ctx, size_t bytes);

// Look up this device's copy of the static scratch of omx_omm_pr_dots_noalloc (once per device,
// with the device current).  Returns OMX_OK or an error.
int noalloc_setup(int device);

// Begin / end a timed launch of kernel class `kclass` (no-ops unless profiling is on).
void prof_begin(omx_ctx ctx, int kclass, hipEvent_t *ev_start);
void prof_end(omx_ctx ctx, int kclass, hipEvent_t ev_start);
// Add the algorithmic work of one launch to the class totals (only while profiling is on).
inline void prof_work(omx_ctx ctx, int kclass, double flops, double bytes) {
    if (!ctx->prof_on) return;
    ctx->prof_flops[kclass] += flops;
    ctx->prof_bytes[kclass] += bytes;
}

inline bool valid_real_dt(int dt) { return dt == OMX_F64 || dt == OMX_F32; }
inline int dt_size(int dt) {
    switch (dt) {
        case OMX_F64: return 8;
        case OMX_F32: return 4;
        case OMX_BF16: return 2;
        case OMX_C128: return 16;
        case OMX_C64: return 8;
        default: return 0;
    }
}
inline bool is_trans(char op) { return op == 't' || op == 'T' || op == 'c' || op == 'C'; }
inline bool valid_op(char op) { return op == 'n' || op == 'N' || is_trans(op); }

// Scalar readback: copy `n` doubles from ctx->d_scalars to ctx->h_scalars and wait.
int read_scalars(omx_ctx ctx, int n);

// Tile index for workgroup id w of a tiles_m x tiles_n product: workgroup ids are dealt
// round-robin over the 8 XCDs, so each XCD is given a contiguous band of tiles; inside a band
// the tiles go in groups of 8 tile rows, column by column, so operand panels are shared in one
// XCD's L2.
__device__ inline void tile_of(int w, int tiles_m, int tiles_n, int &tm, int &tn) {
    int T = tiles_m * tiles_n;
    int t = w;
    if ((T & 7) == 0) t = (w & 7) * (T >> 3) + (w >> 3);
    constexpr int GM = 8;
    int group = t / (GM * tiles_n);
    int first_m = group * GM;
    int gm = min(tiles_m - first_m, GM);
    int in = t % (GM * tiles_n);
    tm = first_m + in % gm;
    tn = in / gm;
}

// Lower-triangle tiles (tm >= tn) of an nt x nt tile grid, nt(nt+1)/2 of them, dealt over the
// XCDs as tile_of does and enumerated by tile rows (products known to be symmetric compute only
// these and mirror them).
__device__ inline void tile_of_lower(int w, int nt, int &tm, int &tn) {
    const int T = nt * (nt + 1) / 2;
    int t = w;
    if ((T & 7) == 0) t = (w & 7) * (T >> 3) + (w >> 3);
    int r = (int)((sqrtf(8.0f * (float)t + 1.0f) - 1.0f) * 0.5f);
    while (r > 0 && r * (r + 1) / 2 > t) --r;
    while ((r + 1) * (r + 2) / 2 <= t) ++r;
    tm = r;
    tn = t - r * (r + 1) / 2;
}

// Launch helpers implemented in the .hip files.
int launch_gemm(omx_ctx ctx, int prec, int dt, bool trA, bool trB, int m, int n, int k, double alpha,
                const void *A, int64_t lda, const void *B, int64_t ldb, double beta, void *C, int64_t ldc);
// C := alpha*A*B^T known to be symmetric (n x n, 'n','t' operands of k columns): the lower-triangle
// tiles where the kernels allow it, then a mirror of the lower triangle onto the upper (both honour
// ctx->launch_mask).
int launch_gemm_sym(omx_ctx ctx, int prec, int dt, int n, int k, double alpha, const void *A, int64_t lda,
                    const void *B, int64_t ldb, void *C, int64_t ldc);
int launch_gemm_pair(omx_ctx ctx, int prec, int dt, int m, int n, int k, double alpha, const void *A1,
                     const void *A2, int64_t lda, const void *B1, const void *B2, int64_t ldb, double beta0,
                     const void *C0, double beta, void *C, int64_t ldc);
// Complex products (OMX_C128 on FP64 MFMA, OMX_C64 on FP32 MFMA); op 'n', 't' or 'c' (conjugate
// transpose); real alpha, beta.
int launch_zgemm(omx_ctx ctx, int dt, char opA, char opB, int m, int n, int k, double alpha, const void *A,
                 int64_t lda, const void *B, int64_t ldb, double beta, void *C, int64_t ldc);
int launch_zgemm_pair(omx_ctx ctx, int dt, int m, int n, int k, double alpha, const void *A1, const void *A2,
                      int64_t lda, const void *B1, const void *B2, int64_t ldb, double beta0, const void *C0,
                      double beta, void *C, int64_t ldc);
// Batched complex128 products (blockIdx.y = system; strides in complex elements; systems with
// active[s] == 0 skipped; no split-K).  Ops: ('n','c') and ('n','n').
int launch_zgemm_batched(omx_ctx ctx, int dt, char opA, char opB, int m, int n, int k, double alpha, const void *A,
                         int64_t lda, int64_t sA, const void *B, int64_t ldb, int64_t sB, double beta, void *C,
                         int64_t ldc, int64_t sC, int nbatch, const int *active);
// C := alpha * A * B^H for products known to be Hermitian (n x n): lower-triangle tiles, then the
// conjugated mirror.  The batched form takes the strides and mask of launch_zgemm_batched.
int launch_zgemm_herm(omx_ctx ctx, int dt, int n, int k, double alpha, const void *A, int64_t lda, const void *B,
                      int64_t ldb, void *C, int64_t ldc);
int launch_zgemm_herm_batched(omx_ctx ctx, int dt, int n, int k, double alpha, const void *A, int64_t lda,
                              int64_t sA, const void *B, int64_t ldb, int64_t sB, void *C, int64_t ldc, int64_t sC,
                              int nbatch, const int *active);
int launch_zgemm_pair_batched(omx_ctx ctx, int dt, int m, int n, int k, double alpha, const void *A1, const void *A2,
                              int64_t lda, int64_t sA, const void *B1, const void *B2, int64_t ldb, int64_t sB,
                              double beta0, const void *C0, double beta, void *C, int64_t ldc, int64_t sC, int nbatch,
                              const int *active);

// Batched pieces of the k-point solver (complex128 or complex64, systems sharing one block pattern).
int launch_spmm_z_batched(omx_ctx ctx, int dt, int no, int nb, int nbc, const int *colptr, const int *rowidx,
                          const double *val, int64_t sV, const void *X, int64_t ldx, int64_t sX, void *Y, int64_t ldy,
                          int64_t sY, int nbatch, const int *active);
// Line-search coefficients (to out[18*s ...]: the 5 coefficients, then the 13 traces) and
// Polak-Ribiere dots (to out2[2*s ...]) of nbatch systems of element type dt, on the device; a
// system s with mask && !mask[s] is skipped.
int launch_coeff_batched(omx_ctx ctx, int dt, int n, int64_t ldm, int64_t sW, const void *HW, const void *SW,
                         const void *HWd, const void *SWd, const void *HWdd, const void *SWdd, int nbatch,
                         double *partial, double *out, const int *mask);
int launch_pr_batched(omx_ctx ctx, int dt, int n, int m, int64_t ld, int64_t sG, const void *G, const void *Gp,
                      int nbatch, double *partial, double *out2, const int *mask);
size_t coeff_partial_doubles(int n, int nbatch);
size_t pr_partial_doubles(int n, int m, int nbatch);

}  // namespace omx
