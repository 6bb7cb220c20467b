// Context, memory, sparse-operator upload, event timing and element access of libommhip.so.
#include <algorithm>
#include <cstdarg>
#include <cstring>
#include <map>
#include <utility>

#include "omx_internal.h"

namespace omx {

static thread_local char g_err[1024] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}
void clear_error() { g_err[0] = 0; }

int ensure_ws(omx_ctx ctx, size_t bytes) {
    if (ctx->ws_bytes >= bytes) return OMX_OK;
    if (ctx->ws) {
        OMX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
        OMX_CHECK_HIP(hipFree(ctx->ws));
        ctx->ws = nullptr;
        ctx->ws_bytes = 0;
    }
    size_t want = std::max(bytes, (size_t)1 << 20);
    if (hipMalloc(&ctx->ws, want) != hipSuccess) {
        set_error("workspace allocation of %zu bytes failed", want);
        return OMX_ERR_NOMEM;
    }
    ctx->ws_bytes = want;
    return OMX_OK;
}

static hipEvent_t pool_get(omx_ctx ctx) {
    if (!ctx->prof_pool.empty()) {
        hipEvent_t e = ctx->prof_pool.back();
        ctx->prof_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

void prof_begin(omx_ctx ctx, int kclass, hipEvent_t *ev_start) {
    *ev_start = nullptr;
    if (!ctx->prof_on) return;
    hipEvent_t e = pool_get(ctx);
    if (!e) return;
    (void)hipEventRecord(e, ctx->stream);
    *ev_start = e;
}

static void prof_drain(omx_ctx ctx) {
    for (auto &p : ctx->prof_pending) {
        (void)hipEventSynchronize(p.stop);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, p.start, p.stop);
        ctx->prof_ms[p.kclass] += ms;
        ctx->prof_count[p.kclass] += 1;
        ctx->prof_pool.push_back(p.start);
        ctx->prof_pool.push_back(p.stop);
    }
    ctx->prof_pending.clear();
}

void prof_end(omx_ctx ctx, int kclass, hipEvent_t ev_start) {
    if (!ctx->prof_on || !ev_start) return;
    hipEvent_t e = pool_get(ctx);
    if (!e) return;
    (void)hipEventRecord(e, ctx->stream);
    ctx->prof_pending.push_back({ev_start, e, kclass});
    if (ctx->prof_pending.size() > 4096) prof_drain(ctx);
}

int read_scalars(omx_ctx ctx, int n) {
    OMX_CHECK_HIP(hipMemcpyAsync(ctx->h_scalars, ctx->d_scalars, sizeof(double) * n, hipMemcpyDeviceToHost,
                                 ctx->stream));
    OMX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    return OMX_OK;
}

// Grid-stride zero fill (used instead of hipMemsetAsync so that the call is capture-safe).
__global__ void k_zero_bytes(uint8_t *p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    // 16-byte stores only when the base is 16-byte aligned; bytes otherwise.
    size_t n16 = (((uintptr_t)p & 15) == 0) ? n / 16 : 0;
    uint4 *q = reinterpret_cast<uint4 *>(p);
    for (size_t t = i; t < n16; t += stride) q[t] = make_uint4(0, 0, 0, 0);
    for (size_t t = n16 * 16 + i; t < n; t += stride) p[t] = 0;
}

template <typename TS, typename TD>
__device__ inline TD cvt(TS v) {
    return (TD)(float)v;
}
template <>
__device__ inline double cvt<double, double>(double v) { return v; }
template <>
__device__ inline float cvt<double, float>(double v) { return (float)v; }
template <>
__device__ inline double cvt<float, double>(float v) { return (double)v; }
template <>
__device__ inline float cvt<float, float>(float v) { return v; }
template <>
__device__ inline __hip_bfloat16 cvt<double, __hip_bfloat16>(double v) { return __float2bfloat16((float)v); }
template <>
__device__ inline __hip_bfloat16 cvt<float, __hip_bfloat16>(float v) { return __float2bfloat16(v); }
template <>
__device__ inline double cvt<__hip_bfloat16, double>(__hip_bfloat16 v) { return (double)__bfloat162float(v); }
template <>
__device__ inline float cvt<__hip_bfloat16, float>(__hip_bfloat16 v) { return __bfloat162float(v); }
template <>
__device__ inline __hip_bfloat16 cvt<__hip_bfloat16, __hip_bfloat16>(__hip_bfloat16 v) { return v; }

template <typename TS, typename TD>
__global__ void k_convert(int m, int n, const TS *__restrict__ src, int64_t lds, TD *__restrict__ dst, int64_t ldd) {
    int64_t total = (int64_t)m * n;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        int64_t i = t % m, j = t / m;
        dst[i + j * ldd] = cvt<TS, TD>(src[i + j * lds]);
    }
}

template <typename TS>
static int convert_from(omx_ctx ctx, int m, int n, const TS *src, int64_t lds, void *dst, int dt_dst, int64_t ldd) {
    int64_t total = (int64_t)m * n;
    int grid = (int)std::min<int64_t>((total + 255) / 256, 65536);
    if (grid < 1) return OMX_OK;
    switch (dt_dst) {
        case OMX_F64:
            hipLaunchKernelGGL((k_convert<TS, double>), dim3(grid), dim3(256), 0, ctx->stream, m, n, src, lds,
                               (double *)dst, ldd);
            break;
        case OMX_F32:
            hipLaunchKernelGGL((k_convert<TS, float>), dim3(grid), dim3(256), 0, ctx->stream, m, n, src, lds,
                               (float *)dst, ldd);
            break;
        case OMX_BF16:
            hipLaunchKernelGGL((k_convert<TS, __hip_bfloat16>), dim3(grid), dim3(256), 0, ctx->stream, m, n, src,
                               lds, (__hip_bfloat16 *)dst, ldd);
            break;
        default:
            set_error("omx_convert: unsupported destination type %d", dt_dst);
            return OMX_ERR_UNSUPPORTED;
    }
    OMX_CHECK_HIP(hipGetLastError());
    return OMX_OK;
}

// Element read/write kernels (single thread; used by m_set_element / m_get_element).
template <typename T>
__global__ void k_set_element(T *C, int64_t off, double alpha, double beta) {
    C[off] = (T)(alpha + beta * (double)C[off]);
}
template <typename T>
__global__ void k_get_element(const T *C, int64_t off, double *out) {
    out[0] = (double)C[off];
}
// Complex element (re, im) at element offset off: C := (alpha_re + beta*re, alpha_im + beta*im).
template <typename R>
__global__ void k_set_element_z(R *C, int64_t off, double are, double aim, double beta) {
    C[2 * off] = (R)(are + beta * (double)C[2 * off]);
    C[2 * off + 1] = (R)(aim + beta * (double)C[2 * off + 1]);
}
template <typename R>
__global__ void k_get_element_z(const R *C, int64_t off, double *out) {
    out[0] = (double)C[2 * off];
    out[1] = (double)C[2 * off + 1];
}

}  // namespace omx

using namespace omx;

extern "C" {

const char *omx_last_error(void) { return g_err; }

int omx_device_count(int *count) {
    if (!count) return OMX_ERR_ARG;
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *count = c;
    return OMX_OK;
}

int omx_init(int device, omx_ctx *out) {
    clear_error();
    OMX_REQUIRE(out != nullptr, "omx_init: null output");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        set_error("omx_init: no HIP device available");
        return OMX_ERR_HIP;
    }
    OMX_REQUIRE(device >= 0 && device < count, "omx_init: device %d out of range (%d devices)", device, count);
    OMX_CHECK_HIP(hipSetDevice(device));
    omx_ctx ctx = new omx_ctx_s();
    ctx->device = device;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        set_error("omx_init: stream creation failed");
        return OMX_ERR_HIP;
    }
    if (hipMalloc(&ctx->d_scalars, sizeof(double) * kScalarSlots) != hipSuccess ||
        hipHostMalloc(&ctx->h_scalars, sizeof(double) * kScalarSlots, hipHostMallocDefault) != hipSuccess) {
        set_error("omx_init: scalar buffer allocation failed");
        return OMX_ERR_NOMEM;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) ctx->num_cus = prop.multiProcessorCount;
    const int rc = noalloc_setup(device);
    if (rc) {
        (void)omx_finalize(ctx);
        return rc;
    }
    *out = ctx;
    return OMX_OK;
}

int omx_finalize(omx_ctx ctx) {
    if (!ctx) return OMX_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    prof_drain(ctx);
    for (hipEvent_t e : ctx->prof_pool) (void)hipEventDestroy(e);
    if (ctx->ws) (void)hipFree(ctx->ws);
    if (ctx->d_scalars) (void)hipFree(ctx->d_scalars);
    if (ctx->h_scalars) (void)hipHostFree(ctx->h_scalars);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return OMX_OK;
}

int omx_sync(omx_ctx ctx) {
    OMX_REQUIRE(ctx, "omx_sync: null context");
    OMX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    return OMX_OK;
}

void *omx_stream(omx_ctx ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int omx_zero(omx_ctx ctx, void *dptr, size_t bytes) {
    OMX_REQUIRE(ctx, "omx_zero: null context");
    if (bytes == 0) return OMX_OK;
    OMX_REQUIRE(dptr, "omx_zero: null pointer");
    size_t blocks = std::min<size_t>((bytes / 16 + 255) / 256 + 1, 16384);
    hipLaunchKernelGGL(k_zero_bytes, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, (uint8_t *)dptr, bytes);
    OMX_CHECK_HIP(hipGetLastError());
    return OMX_OK;
}

int omx_malloc(omx_ctx ctx, size_t bytes, void **dptr) {
    OMX_REQUIRE(ctx && dptr, "omx_malloc: null argument");
    *dptr = nullptr;
    if (bytes == 0) bytes = 16;
    if (hipMalloc(dptr, bytes) != hipSuccess) {
        set_error("omx_malloc: allocation of %zu bytes failed", bytes);
        *dptr = nullptr;
        return OMX_ERR_NOMEM;
    }
    return omx_zero(ctx, *dptr, bytes);
}

int omx_free(omx_ctx ctx, void *dptr) {
    OMX_REQUIRE(ctx, "omx_free: null context");
    if (!dptr) return OMX_OK;
    OMX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    OMX_CHECK_HIP(hipFree(dptr));
    return OMX_OK;
}

int omx_h2d(omx_ctx ctx, void *dst, const void *src, size_t bytes) {
    OMX_REQUIRE(ctx, "omx_h2d: null context");
    if (!bytes) return OMX_OK;
    OMX_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    OMX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    return OMX_OK;
}

int omx_d2h(omx_ctx ctx, void *dst, const void *src, size_t bytes) {
    OMX_REQUIRE(ctx, "omx_d2h: null context");
    if (!bytes) return OMX_OK;
    OMX_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    OMX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    return OMX_OK;
}

int omx_d2d(omx_ctx ctx, void *dst, const void *src, size_t bytes) {
    OMX_REQUIRE(ctx, "omx_d2d: null context");
    if (!bytes) return OMX_OK;
    OMX_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    return OMX_OK;
}

int omx_h2d_2d(omx_ctx ctx, void *dst, int64_t ldd, const void *src, int64_t lds, int m, int n, int esize) {
    OMX_REQUIRE(ctx && m >= 0 && n >= 0 && esize > 0 && ldd >= m && lds >= m, "omx_h2d_2d: bad arguments");
    if (!m || !n) return OMX_OK;
    OMX_CHECK_HIP(hipMemcpy2DAsync(dst, ldd * esize, src, lds * esize, (size_t)m * esize, n, hipMemcpyHostToDevice,
                                   ctx->stream));
    OMX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    return OMX_OK;
}

int omx_d2h_2d(omx_ctx ctx, void *dst, int64_t ldd, const void *src, int64_t lds, int m, int n, int esize) {
    OMX_REQUIRE(ctx && m >= 0 && n >= 0 && esize > 0 && ldd >= m && lds >= m, "omx_d2h_2d: bad arguments");
    if (!m || !n) return OMX_OK;
    OMX_CHECK_HIP(hipMemcpy2DAsync(dst, ldd * esize, src, lds * esize, (size_t)m * esize, n, hipMemcpyDeviceToHost,
                                   ctx->stream));
    OMX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    return OMX_OK;
}

int omx_convert(omx_ctx ctx, int m, int n, const void *src, int dt_src, int64_t lds, void *dst, int dt_dst,
                int64_t ldd) {
    OMX_REQUIRE(ctx && src && dst && m >= 0 && n >= 0 && lds >= m && ldd >= m, "omx_convert: bad arguments");
    switch (dt_src) {
        case OMX_F64: return convert_from(ctx, m, n, (const double *)src, lds, dst, dt_dst, ldd);
        case OMX_F32: return convert_from(ctx, m, n, (const float *)src, lds, dst, dt_dst, ldd);
        case OMX_BF16: return convert_from(ctx, m, n, (const __hip_bfloat16 *)src, lds, dst, dt_dst, ldd);
        default: set_error("omx_convert: unsupported source type %d", dt_src); return OMX_ERR_UNSUPPORTED;
    }
}

// ---- sparse operators -----------------------------------------------------------------------

// Lx, Ly, Lz when block column J = x + Lx*(y + Ly*z) of an nbc-site operator holds exactly the
// blocks of J and of its six periodic neighbours x+-1, y+-1, z+-1 (every edge >= 3), else false.
// Site 0's sorted block rows are 0, 1, Lx-1, Lx, Lx*(Ly-1), Lx*Ly, Lx*Ly*(Lz-1), which give the
// edges; every column is then checked against them.
static bool detect_lattice(int nbc, const std::vector<int> &colptr, const std::vector<int> &rowidx, int lat[3]) {
    lat[0] = lat[1] = lat[2] = 0;
    if (nbc < 27 || colptr[1] - colptr[0] != 7) return false;
    const int *r = rowidx.data();
    const int Lx = r[3];
    if (Lx < 3 || r[0] != 0 || r[1] != 1 || r[2] != Lx - 1) return false;
    const int LxLy = r[5];
    if (LxLy % Lx != 0) return false;
    const int Ly = LxLy / Lx;
    if (Ly < 3 || r[4] != Lx * (Ly - 1) || nbc % LxLy != 0) return false;
    const int Lz = nbc / LxLy;
    if (Lz < 3 || r[6] != LxLy * (Lz - 1)) return false;
    for (int J = 0; J < nbc; ++J) {
        if (colptr[J + 1] - colptr[J] != 7) return false;
        const int x = J % Lx, y = (J / Lx) % Ly, z = J / LxLy;
        int want[7] = {J,
                       (x + 1) % Lx + Lx * (y + Ly * z),
                       (x + Lx - 1) % Lx + Lx * (y + Ly * z),
                       x + Lx * ((y + 1) % Ly + Ly * z),
                       x + Lx * ((y + Ly - 1) % Ly + Ly * z),
                       x + Lx * (y + Ly * ((z + 1) % Lz)),
                       x + Lx * (y + Ly * ((z + Lz - 1) % Lz))};
        std::sort(want, want + 7);
        for (int q = 0; q < 7; ++q)
            if (rowidx[colptr[J] + q] != want[q]) return false;
    }
    lat[0] = Lx;
    lat[1] = Ly;
    lat[2] = Lz;
    return true;
}

static int bsr_upload(omx_ctx ctx, int nb, const std::vector<int> &colptr, const std::vector<int> &rowidx,
                      const double *val, int is_complex, omx_bsr *out) {
    omx_bsr a = new omx_bsr_s();
    a->ctx = ctx;
    a->nb = nb;
    a->nbc = (nb + 3) / 4;
    a->nnzb = (int64_t)rowidx.size();
    a->is_complex = is_complex;
    int mx = 0;
    for (int j = 0; j < a->nbc; ++j) mx = std::max(mx, colptr[j + 1] - colptr[j]);
    a->max_blocks_per_col = mx;
    if (nb == 4 * a->nbc) (void)detect_lattice(a->nbc, colptr, rowidx, a->lat);
    size_t vbytes = sizeof(double) * 16 * (size_t)a->nnzb * (is_complex ? 2 : 1);
    if (hipMalloc(&a->d_colptr, sizeof(int) * (a->nbc + 1)) != hipSuccess ||
        hipMalloc(&a->d_rowidx, sizeof(int) * std::max<int64_t>(a->nnzb, 1)) != hipSuccess ||
        hipMalloc(&a->d_val, std::max<size_t>(vbytes, 16)) != hipSuccess) {
        omx_bsr_destroy(a);
        set_error("omx_bsr: device allocation failed");
        return OMX_ERR_NOMEM;
    }
    (void)hipMemcpyAsync(a->d_colptr, colptr.data(), sizeof(int) * (a->nbc + 1), hipMemcpyHostToDevice, ctx->stream);
    if (a->nnzb) {
        (void)hipMemcpyAsync(a->d_rowidx, rowidx.data(), sizeof(int) * a->nnzb, hipMemcpyHostToDevice, ctx->stream);
        (void)hipMemcpyAsync(a->d_val, val, vbytes, hipMemcpyHostToDevice, ctx->stream);
    }
    OMX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    *out = a;
    return OMX_OK;
}

static int bsr_create_common(omx_ctx ctx, int nb, const int *bcolptr, const int *browidx, const double *bval,
                             int is_complex, omx_bsr *out) {
    OMX_REQUIRE(ctx && out && bcolptr && nb > 0, "omx_bsr_create: bad arguments");
    int nbc = (nb + 3) / 4;
    OMX_REQUIRE(bcolptr[0] == 0, "omx_bsr_create: bcolptr must be 0-based");
    for (int j = 0; j < nbc; ++j)
        OMX_REQUIRE(bcolptr[j + 1] >= bcolptr[j], "omx_bsr_create: bcolptr not monotone at %d", j);
    int64_t nnzb = bcolptr[nbc];
    OMX_REQUIRE(nnzb == 0 || (browidx && bval), "omx_bsr_create: null block arrays");
    for (int j = 0; j < nbc; ++j)
        for (int l = bcolptr[j]; l < bcolptr[j + 1]; ++l) {
            OMX_REQUIRE(browidx[l] >= 0 && browidx[l] < nbc, "omx_bsr_create: block row %d out of range", browidx[l]);
            OMX_REQUIRE(l == bcolptr[j] || browidx[l] > browidx[l - 1],
                        "omx_bsr_create: block rows of column %d not strictly ascending", j);
        }
    std::vector<int> cp(bcolptr, bcolptr + nbc + 1), ri(browidx, browidx + nnzb);
    return bsr_upload(ctx, nb, cp, ri, bval, is_complex, out);
}

int omx_bsr_create_z(omx_ctx ctx, int nb, const int *bcolptr, const int *browidx, const double *bval,
                     omx_bsr *out) {
    return bsr_create_common(ctx, nb, bcolptr, browidx, bval, 1, out);
}

// CSC (colptr with base ptrbase, row indices with base rowbase; values (re, im) interleaved when
// cplx) to BSR-4 block columns: block rows ascending, duplicates added, empty blocks not stored.
// The column pointers are checked before any index or value is read: colptr[0] == ptrbase and
// non-decreasing, so every entry offset lies in [0, nnz).
static int csc_to_bsr(int nb, const int *colptr, const int *rowidx, const double *val, int ptrbase, int rowbase,
                      bool cplx, std::vector<int> &cp, std::vector<int> &ri, std::vector<double> &bv,
                      const char *who) {
    OMX_REQUIRE(colptr && nb > 0, "%s: bad arguments", who);
    OMX_REQUIRE((ptrbase == 0 || ptrbase == 1) && (rowbase == 0 || rowbase == 1), "%s: bad base", who);
    OMX_REQUIRE(colptr[0] == ptrbase, "%s: colptr[0] is %d, expected %d", who, colptr[0], ptrbase);
    for (int j = 0; j < nb; ++j)
        OMX_REQUIRE(colptr[j + 1] >= colptr[j], "%s: colptr decreases at column %d", who, j);
    const int64_t nnz = (int64_t)colptr[nb] - ptrbase;
    OMX_REQUIRE(nnz == 0 || (rowidx && val), "%s: null index or value array", who);
    const int nbc = (nb + 3) / 4;
    const int bs = cplx ? 32 : 16;
    cp.assign(nbc + 1, 0);
    ri.clear();
    bv.clear();
    for (int J = 0; J < nbc; ++J) {
        std::map<int, std::vector<std::pair<int, int64_t>>> blocks;  // I -> (r + 4c, entry)
        for (int c = 0; c < 4; ++c) {
            const int j = 4 * J + c;
            if (j >= nb) break;
            for (int64_t l = colptr[j] - ptrbase; l < colptr[j + 1] - ptrbase; ++l) {
                const int i = rowidx[l] - rowbase;
                OMX_REQUIRE(i >= 0 && i < nb, "%s: row index %d out of range", who, i + rowbase);
                blocks[i / 4].push_back({(i % 4) + 4 * c, l});
            }
        }
        for (auto &kv : blocks) {
            ri.push_back(kv.first);
            const size_t base = bv.size();
            bv.resize(base + bs, 0.0);
            for (auto &e : kv.second) {
                if (cplx) {
                    bv[base + 2 * e.first] += val[2 * e.second];
                    bv[base + 2 * e.first + 1] += val[2 * e.second + 1];
                } else {
                    bv[base + e.first] += val[e.second];
                }
            }
        }
        cp[J + 1] = (int)ri.size();
    }
    return OMX_OK;
}

int omx_bsr_from_csc_z(omx_ctx ctx, int nb, const int *colptr, const int *rowidx, const double *val, int ptrbase,
                       int rowbase, omx_bsr *out) {
    OMX_REQUIRE(ctx && out, "omx_bsr_from_csc_z: bad arguments");
    std::vector<int> cp, ri;
    std::vector<double> bv;
    int rc = csc_to_bsr(nb, colptr, rowidx, val, ptrbase, rowbase, 1, cp, ri, bv, "omx_bsr_from_csc_z");
    if (rc) return rc;
    return bsr_upload(ctx, nb, cp, ri, bv.data(), 1, out);
}

int omx_bsr_create(omx_ctx ctx, int nb, const int *bcolptr, const int *browidx, const double *bval, omx_bsr *out) {
    OMX_REQUIRE(ctx && out && bcolptr && nb > 0, "omx_bsr_create: bad arguments");
    int nbc = (nb + 3) / 4;
    OMX_REQUIRE(bcolptr[0] == 0, "omx_bsr_create: bcolptr must be 0-based");
    for (int j = 0; j < nbc; ++j)
        OMX_REQUIRE(bcolptr[j + 1] >= bcolptr[j], "omx_bsr_create: bcolptr not monotone at %d", j);
    int64_t nnzb = bcolptr[nbc];
    OMX_REQUIRE(nnzb == 0 || (browidx && bval), "omx_bsr_create: null block arrays");
    for (int j = 0; j < nbc; ++j)
        for (int l = bcolptr[j]; l < bcolptr[j + 1]; ++l) {
            OMX_REQUIRE(browidx[l] >= 0 && browidx[l] < nbc, "omx_bsr_create: block row %d out of range", browidx[l]);
            OMX_REQUIRE(l == bcolptr[j] || browidx[l] > browidx[l - 1],
                        "omx_bsr_create: block rows of column %d not strictly ascending", j);
        }
    std::vector<int> cp(bcolptr, bcolptr + nbc + 1), ri(browidx, browidx + nnzb);
    return bsr_upload(ctx, nb, cp, ri, bval, 0, out);
}

int omx_bsr_from_csc(omx_ctx ctx, int nb, const int *colptr, const int *rowidx, const double *val, int ptrbase,
                     int rowbase, omx_bsr *out) {
    OMX_REQUIRE(ctx && out, "omx_bsr_from_csc: bad arguments");
    std::vector<int> cp, ri;
    std::vector<double> bv;
    int rc = csc_to_bsr(nb, colptr, rowidx, val, ptrbase, rowbase, 0, cp, ri, bv, "omx_bsr_from_csc");
    if (rc) return rc;
    return bsr_upload(ctx, nb, cp, ri, bv.data(), 0, out);
}

int omx_csc_to_bsr(int nb, const int *colptr, const int *rowidx, const double *val, int ptrbase, int rowbase,
                   int is_complex, int64_t *nnzb, int *bcolptr, int *browidx, double *bval) {
    OMX_REQUIRE(nnzb, "omx_csc_to_bsr: bad arguments");
    std::vector<int> cp, ri;
    std::vector<double> bv;
    int rc = csc_to_bsr(nb, colptr, rowidx, val, ptrbase, rowbase, is_complex != 0, cp, ri, bv, "omx_csc_to_bsr");
    if (rc) return rc;
    *nnzb = (int64_t)ri.size();
    if (bcolptr) std::copy(cp.begin(), cp.end(), bcolptr);
    if (browidx) std::copy(ri.begin(), ri.end(), browidx);
    if (bval) std::copy(bv.begin(), bv.end(), bval);
    return OMX_OK;
}

int omx_bsr_lattice(omx_bsr a, int *dims) {
    OMX_REQUIRE(a && dims, "omx_bsr_lattice: bad arguments");
    for (int q = 0; q < 3; ++q) dims[q] = a->lat[q];
    return OMX_OK;
}

int omx_bsr_destroy(omx_bsr a) {
    if (!a) return OMX_OK;
    if (a->ctx) (void)hipStreamSynchronize(a->ctx->stream);
    if (a->d_colptr) (void)hipFree(a->d_colptr);
    if (a->d_rowidx) (void)hipFree(a->d_rowidx);
    if (a->d_val) (void)hipFree(a->d_val);
    if (a->d_walk) (void)hipFree(a->d_walk);
    if (a->d_gptr) (void)hipFree(a->d_gptr);
    if (a->d_uidx) (void)hipFree(a->d_uidx);
    if (a->d_slot) (void)hipFree(a->d_slot);
    delete a;
    return OMX_OK;
}

int omx_bsr_info(omx_bsr a, int *nb, int64_t *nnzb, int *is_complex) {
    OMX_REQUIRE(a, "omx_bsr_info: null operator");
    if (nb) *nb = a->nb;
    if (nnzb) *nnzb = a->nnzb;
    if (is_complex) *is_complex = a->is_complex;
    return OMX_OK;
}

int omx_bsr_update_values(omx_bsr a, const double *bval) {
    OMX_REQUIRE(a && (bval || a->nnzb == 0), "omx_bsr_update_values: bad arguments");
    size_t vbytes = sizeof(double) * 16 * (size_t)a->nnzb * (a->is_complex ? 2 : 1);
    if (!vbytes) return OMX_OK;
    OMX_CHECK_HIP(hipMemcpyAsync(a->d_val, bval, vbytes, hipMemcpyHostToDevice, a->ctx->stream));
    OMX_CHECK_HIP(hipStreamSynchronize(a->ctx->stream));
    return OMX_OK;
}

int omx_bsr_get_values(omx_bsr a, double *bval) {
    OMX_REQUIRE(a && (bval || a->nnzb == 0), "omx_bsr_get_values: bad arguments");
    size_t vbytes = sizeof(double) * 16 * (size_t)a->nnzb * (a->is_complex ? 2 : 1);
    if (!vbytes) return OMX_OK;
    OMX_CHECK_HIP(hipMemcpyAsync(bval, a->d_val, vbytes, hipMemcpyDeviceToHost, a->ctx->stream));
    OMX_CHECK_HIP(hipStreamSynchronize(a->ctx->stream));
    return OMX_OK;
}

int omx_bsr_get_pattern(omx_bsr a, int *bcolptr, int *browidx) {
    OMX_REQUIRE(a && bcolptr && (browidx || a->nnzb == 0), "omx_bsr_get_pattern: bad arguments");
    OMX_CHECK_HIP(hipMemcpyAsync(bcolptr, a->d_colptr, sizeof(int) * (a->nbc + 1), hipMemcpyDeviceToHost,
                                 a->ctx->stream));
    if (a->nnzb)
        OMX_CHECK_HIP(hipMemcpyAsync(browidx, a->d_rowidx, sizeof(int) * a->nnzb, hipMemcpyDeviceToHost,
                                     a->ctx->stream));
    OMX_CHECK_HIP(hipStreamSynchronize(a->ctx->stream));
    return OMX_OK;
}

// ---- autotuner -------------------------------------------------------------------------------

int omx_autotune(omx_ctx ctx, int enable) {
    OMX_REQUIRE(ctx, "omx_autotune: null context");
    ctx->autotune = (enable != 0);
    return OMX_OK;
}

int omx_autotune_report(omx_ctx ctx, char *buf, int64_t len) {
    OMX_REQUIRE(ctx && buf && len > 0, "omx_autotune_report: bad arguments");
    std::string out;
    char line[256];
    for (auto &kv : ctx->tune_gemm) {
        snprintf(line, sizeof line, "%s -> cfg %d (%.4f ms)\n", kv.first.c_str(), kv.second.first, kv.second.second);
        out += line;
    }
    for (auto &kv : ctx->tune_spmm) {
        const omx_spmm_shape_s &sh = kv.second.first;
        snprintf(line, sizeof line, "%s -> batch %d, waves %d, block columns %d, XCD map %d (%.4f ms)\n",
                 kv.first.c_str(), sh.batch, sh.wpg, sh.jper, sh.mode, kv.second.second);
        out += line;
    }
    const size_t n = std::min<size_t>(out.size(), (size_t)len - 1);
    memcpy(buf, out.data(), n);
    buf[n] = 0;
    return (int)std::min<size_t>(out.size(), (size_t)INT32_MAX);
}

// ---- element access --------------------------------------------------------------------------

int omx_set_element(omx_ctx ctx, int dt, void *C, int64_t ldc, int i, int j, double alpha, double beta) {
    OMX_REQUIRE(ctx && C && i >= 1 && j >= 1 && ldc >= i, "omx_set_element: bad arguments");
    int64_t off = (int64_t)(i - 1) + (int64_t)(j - 1) * ldc;
    if (dt == OMX_F64)
        hipLaunchKernelGGL(k_set_element<double>, dim3(1), dim3(1), 0, ctx->stream, (double *)C, off, alpha, beta);
    else if (dt == OMX_F32)
        hipLaunchKernelGGL(k_set_element<float>, dim3(1), dim3(1), 0, ctx->stream, (float *)C, off, alpha, beta);
    else if (dt == OMX_C128)
        hipLaunchKernelGGL(k_set_element_z<double>, dim3(1), dim3(1), 0, ctx->stream, (double *)C, off, alpha, 0.0, beta);
    else if (dt == OMX_C64)
        hipLaunchKernelGGL(k_set_element_z<float>, dim3(1), dim3(1), 0, ctx->stream, (float *)C, off, alpha, 0.0, beta);
    else {
        set_error("omx_set_element: unsupported type %d", dt);
        return OMX_ERR_UNSUPPORTED;
    }
    OMX_CHECK_HIP(hipGetLastError());
    return OMX_OK;
}

int omx_get_element(omx_ctx ctx, int dt, const void *C, int64_t ldc, int i, int j, double *out) {
    OMX_REQUIRE(ctx && C && out && i >= 1 && j >= 1 && ldc >= i, "omx_get_element: bad arguments");
    int64_t off = (int64_t)(i - 1) + (int64_t)(j - 1) * ldc;
    if (dt == OMX_F64)
        hipLaunchKernelGGL(k_get_element<double>, dim3(1), dim3(1), 0, ctx->stream, (const double *)C, off,
                           ctx->d_scalars);
    else if (dt == OMX_F32)
        hipLaunchKernelGGL(k_get_element<float>, dim3(1), dim3(1), 0, ctx->stream, (const float *)C, off,
                           ctx->d_scalars);
    else if (dt == OMX_C128)
        hipLaunchKernelGGL(k_get_element_z<double>, dim3(1), dim3(1), 0, ctx->stream, (const double *)C, off,
                           ctx->d_scalars);
    else if (dt == OMX_C64)
        hipLaunchKernelGGL(k_get_element_z<float>, dim3(1), dim3(1), 0, ctx->stream, (const float *)C, off,
                           ctx->d_scalars);
    else {
        set_error("omx_get_element: unsupported type %d", dt);
        return OMX_ERR_UNSUPPORTED;
    }
    OMX_CHECK_HIP(hipGetLastError());
    const bool z = (dt == OMX_C128 || dt == OMX_C64);
    int rc = read_scalars(ctx, z ? 2 : 1);
    if (rc) return rc;
    out[0] = ctx->h_scalars[0];
    if (z) out[1] = ctx->h_scalars[1];
    return OMX_OK;
}

int omx_set_element_z(omx_ctx ctx, int dt, void *C, int64_t ldc, int i, int j, double alpha_re, double alpha_im,
                      double beta) {
    OMX_REQUIRE(ctx && C && i >= 1 && j >= 1 && ldc >= i, "omx_set_element_z: bad arguments");
    int64_t off = (int64_t)(i - 1) + (int64_t)(j - 1) * ldc;
    if (dt == OMX_C128)
        hipLaunchKernelGGL(k_set_element_z<double>, dim3(1), dim3(1), 0, ctx->stream, (double *)C, off, alpha_re,
                           alpha_im, beta);
    else if (dt == OMX_C64)
        hipLaunchKernelGGL(k_set_element_z<float>, dim3(1), dim3(1), 0, ctx->stream, (float *)C, off, alpha_re,
                           alpha_im, beta);
    else {
        set_error("omx_set_element_z: complex element type expected, got %d", dt);
        return OMX_ERR_UNSUPPORTED;
    }
    OMX_CHECK_HIP(hipGetLastError());
    return OMX_OK;
}

// ---- event timing ----------------------------------------------------------------------------

int omx_prof_enable(omx_ctx ctx, int enable) {
    OMX_REQUIRE(ctx, "omx_prof_enable: null context");
    ctx->prof_on = enable != 0;
    return OMX_OK;
}

int omx_prof_read(omx_ctx ctx, int kclass, double *ms, int64_t *launches) {
    OMX_REQUIRE(ctx && kclass >= 0 && kclass < OMX_K_COUNT, "omx_prof_read: bad arguments");
    OMX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    prof_drain(ctx);
    if (ms) *ms = ctx->prof_ms[kclass];
    if (launches) *launches = ctx->prof_count[kclass];
    return OMX_OK;
}

int omx_prof_reset(omx_ctx ctx) {
    OMX_REQUIRE(ctx, "omx_prof_reset: null context");
    OMX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    prof_drain(ctx);
    for (int k = 0; k < OMX_K_COUNT; ++k) {
        ctx->prof_ms[k] = 0;
        ctx->prof_count[k] = 0;
        ctx->prof_flops[k] = 0;
        ctx->prof_bytes[k] = 0;
    }
    return OMX_OK;
}

int omx_prof_work(omx_ctx ctx, int kclass, double *flops, double *bytes) {
    OMX_REQUIRE(ctx && kclass >= 0 && kclass < OMX_K_COUNT, "omx_prof_work: bad arguments");
    if (flops) *flops = ctx->prof_flops[kclass];
    if (bytes) *bytes = ctx->prof_bytes[kclass];
    return OMX_OK;
}

}  // extern "C"
