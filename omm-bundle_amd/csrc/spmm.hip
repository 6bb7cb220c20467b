// Dense x block-sparse product Y := alpha*X*A + beta*Y on CDNA4 (the H*C and S*C products of
// the OMM loop; MatrixSwitch/src/MatrixSwitch_mm_multiply.F90:263-327 and
// pspBLAS/src/pspLevel3/pspMspm_nn.F90:58-212 compute the same product from CSC).
//
// X, Y: no x nb column-major; a basis column X(:,k) is a contiguous vector of `no` state
// coefficients.  A: BSR-4 by block columns.  Output column 4J+c is
//     Y(:,4J+c) = sum over blocks (I,J), r = 0..3 of X(:,4I+r) * A_IJ(r,c),
// summed in ascending (I, r), the order of the reference's CSC loop.
//
// Mapping: a workgroup owns a chunk of RC = 64*VEC*WPG consecutive states (its WPG waves are
// stacked in rows, each lane loading one 8- or 16-byte vector per basis column) and a run of jper
// block columns, taken in walk order (the identity, or an order attached to the operator that
// keeps the basis columns a block column gathers close in time to their other uses, e.g. the
// lattice in x-y tiles).  Each lane keeps 4*VEC accumulators and issues the loads of a whole
// batch of NB blocks before using them (memory-level parallelism); the 16 block values are
// wave-uniform and come through the scalar cache.  Workgroup ids are dealt round-robin over the 8
// XCDs, and the id -> (row chunk, block-column run) map is XCD-aware:
//   mode 0: XCD x owns a contiguous eighth of the row chunks and cycles through them for one
//           block-column run before the next, so the chip reads whole basis columns at once;
//   mode 1: XCD x owns chunks x, x+8, ... and walks every block column of one chunk before the
//           next, so an XCD's L2 holds the neighbourhood of the current block columns of a
//           single chunk.
// A zero entry stored inside a block contributes an exact zero, so the per-element sum equals
// the reference's CSC sum over its stored entries.
#include <algorithm>
#include <utility>

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
template <>
struct VecT<float, 2> {
    typedef float2 type;
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

// y := alpha*acc + beta*y_old in FP64 with one explicit rounding sequence (product, then a fused
// multiply-add), so that every SpMM kernel rounds the epilogue identically whatever the compiler's
// contraction choices.
template <typename T>
__device__ inline T spmm_out(double alpha, T acc, double beta, T old) {
    const double p = alpha * (double)acc;
    return (T)(beta != 0.0 ? __builtin_fma(beta, (double)old, p) : p);
}

template <typename T, int VEC, int NB, int WPG>
__global__ __launch_bounds__(64 * WPG) void k_spmm_bsr4(int no, int nb, int nbc, double alpha,
                                                      const T *__restrict__ X, int64_t ldx,
                                                      const int *__restrict__ colptr, const int *__restrict__ rowidx,
                                                      const double *__restrict__ val, double beta,
                                                      T *__restrict__ Y, int64_t ldy, int nrc, int njg, int jper,
                                                      int mode, const int *__restrict__ walk,
                                                      const int *__restrict__ mask) {
    constexpr int RC = 64 * VEC * WPG;
    if (mask && !*mask) return;
    const int w = blockIdx.x;
    int rc, jg;
    if ((nrc & 7) == 0) {
        const int xcd = w & 7, local = w >> 3;  // dispatch is round-robin over 8 XCDs
        if (mode == 0) {
            const int per = nrc >> 3;
            rc = xcd * per + local % per;
            jg = local / per;
        } else {
            rc = (local / njg) * 8 + xcd;
            jg = local % njg;
        }
    } else {
        rc = w % nrc;
        jg = w / nrc;
    }
    const int i0 = rc * RC + threadIdx.x * VEC;
    const bool full = (i0 + VEC <= no);
    const bool any = (i0 < no);
    const int jbeg = jg * jper, jend = min(nbc, jbeg + jper);

    for (int jj = jbeg; jj < jend; ++jj) {
        const int J = walk ? walk[jj] : jj;
        T acc[4][VEC];
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int q = 0; q < VEC; ++q) acc[c][q] = (T)0;
        const int lb = colptr[J], le = colptr[J + 1];
        // Blocks are taken in batches of up to NB: all their basis columns are requested before
        // any is used, so each lane keeps 4*NB vector loads in flight; the accumulation then runs
        // in ascending (block, r) order.
        for (int l0 = lb; l0 < le; l0 += NB) {
            const int nblk = min(NB, le - l0);
            T x[NB][4][VEC];
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                if (b < nblk) {
                    const int I = rowidx[l0 + b];
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int kcol = 4 * I + r;
                        const T *src = X + (int64_t)kcol * ldx + i0;
                        if (kcol < nb && full) {
                            vload<T, VEC>(src, x[b][r]);
                        } else {
#pragma unroll
                            for (int q = 0; q < VEC; ++q)
                                x[b][r][q] = (kcol < nb && any && i0 + q < no) ? src[q] : (T)0;
                        }
                    }
                }
            }
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                if (b < nblk) {
                    const double *v = val + 16 * (int64_t)(l0 + b);
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
#pragma unroll
                        for (int c = 0; c < 4; ++c) {
                            const T a = (T)v[r + 4 * c];
#pragma unroll
                            for (int q = 0; q < VEC; ++q) acc[c][q] = fma(x[b][r][q], a, acc[c][q]);
                        }
                    }
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
                    for (int q = 0; q < VEC; ++q) out[q] = spmm_out(alpha, acc[c][q], beta, old[q]);
                } else {
#pragma unroll
                    for (int q = 0; q < VEC; ++q) out[q] = spmm_out(alpha, acc[c][q], 0.0, (T)0);
                }
                vstore<T, VEC>(dst, out);
            } else {
#pragma unroll
                for (int q = 0; q < VEC; ++q)
                    if (i0 + q < no) dst[q] = spmm_out(alpha, acc[c][q], beta, beta != 0.0 ? dst[q] : (T)0);
            }
        }
    }
}

// ---- 2.5-D kernel for lattice operators ------------------------------------------------------
// An operator whose pattern is the periodic 7-point stencil of an Lx x Ly x Lz lattice of 4x4
// blocks (block column J = site x + Lx*(y + Ly*z) holds exactly the blocks of J, x+-1, y+-1,
// z+-1; recognised at upload, omx_bsr_s::lat) gathers each basis column 7 times.  Here a
// workgroup owns one row chunk (RW = 64*VEC states) of a 4x4 tile of sites and marches through
// the tile's z-planes: wave w owns site w of the tile and keeps its own basis columns of planes
// z-1, z, z+1 in registers (z+2 in flight); the plane-z columns of the tile and of its 16-site halo
// (x-1, x+4, y-1, y+4 edges, one site per wave) are in LDS, double-buffered by plane parity, where
// the in-plane neighbours are read.  Each basis column then crosses from L2 about twice (own +
// halo) instead of 7 times.  The 7 blocks of a column are taken in stored order with the FMA
// sequence of k_spmm_bsr4, so the results are the same bits.
// Measured at n = 16384 (FP64): 1.2x the algorithmic HBM bytes (2.6 GB per launch against 4.4 GB
// for k_spmm_bsr4), but 0.89 ms against 0.83 ms in the same run (FP32: 0.57 against 0.47 ms): one
// workgroup per CU (128 KB of LDS) and a barrier per plane.  The block values come through the
// scalar cache (broadcast with v_readlane they cost 2.6x the VALU instructions of k_spmm_bsr4 and
// the kernel 1.08 ms).  It is therefore opt-in (omx_spmm_stencil).
constexpr int ST_TX = 4, ST_TY = 4, ST_WAVES = ST_TX * ST_TY, ST_SLOTS = ST_WAVES + 2 * (ST_TX + ST_TY);
constexpr int ST_VBYTES = 8;  // bytes per lane and basis column: 64 FP64 or 128 FP32 states per wave

// Workgroup barrier for LDS traffic only: this wave's LDS operations complete (lgkmcnt(0)), then
// s_barrier.  Global loads in flight (the prefetch of later planes into registers) are not waited
// for; the compiler waits for them where their registers are used.  The signal fences keep the
// compiler from moving LDS accesses across the barrier.
__device__ inline void lds_barrier() {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0xc07f);  // gfx9 encoding: vmcnt 63, expcnt 7, lgkmcnt 0
    __builtin_amdgcn_s_barrier();
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// Sorts the 7 (key, tag) pairs by key: a 16-comparator network, checked on all 0/1 inputs.
__device__ inline void sort7(int (&k)[7], int (&d)[7]) {
    constexpr int P[16][2] = {{0, 6}, {2, 3}, {4, 5}, {0, 2}, {1, 4}, {3, 6}, {0, 1}, {2, 5},
                              {3, 4}, {1, 2}, {4, 6}, {2, 3}, {4, 5}, {1, 2}, {3, 4}, {5, 6}};
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int a = P[q][0], b = P[q][1];
        const bool sw = k[a] > k[b];
        const int ka = sw ? k[b] : k[a], kb = sw ? k[a] : k[b];
        const int da = sw ? d[b] : d[a], db = sw ? d[a] : d[b];
        k[a] = ka;
        k[b] = kb;
        d[a] = da;
        d[b] = db;
    }
}

// f(z0 + k, integral_constant<k>) for k = 0.. while z0 + k < Lz: the march unrolled so that every
// register-ring slot index is a compile-time constant.
template <typename F, int... Ks>
__device__ inline void unroll_steps(F &f, int z0, int Lz, std::integer_sequence<int, Ks...>) {
    ((z0 + Ks < Lz ? f(z0 + Ks, std::integral_constant<int, Ks>{}) : void()), ...);
}

template <typename T, int MINB>
__global__ __launch_bounds__(64 * ST_WAVES, MINB) void k_spmm_stencil(int no, int Lx, int Ly, int Lz, double alpha,
                                                                 const T *__restrict__ X, int64_t ldx,
                                                                 const int *__restrict__ colptr,
                                                                 const int *__restrict__ rowidx,
                                                                 const double *__restrict__ val, double beta,
                                                                 T *__restrict__ Y, int64_t ldy, int nrc, int ntx,
                                                                 int ntiles, const int *__restrict__ mask) {
    constexpr int VEC = ST_VBYTES / sizeof(T);
    constexpr int RW = 64 * VEC;
    __shared__ __align__(16) T lds[2][ST_SLOTS][4][RW];
    if (mask && !*mask) return;
    const int wid = blockIdx.x;
    int rc, t;
    if ((nrc & 7) == 0) {
        // dispatch is round-robin over the 8 XCDs: the tiles of one row chunk run on one XCD, so the
        // halo columns a tile reads are the plane its neighbours are loading into that XCD's L2
        const int xcd = wid & 7, local = wid >> 3;
        rc = (local / ntiles) * 8 + xcd;
        t = local % ntiles;
    } else {
        rc = wid / ntiles;
        t = wid % ntiles;
    }
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int bx = (t % ntx) * ST_TX, by = (t / ntx) * ST_TY;
    const int tx = w % ST_TX, ty = w / ST_TX;
    const int ox = bx + tx, oy = by + ty;  // own site of wave w
    // halo site of wave w: edges x = bx-1, x = bx+4 (k = y offset), y = by-1, y = by+4 (k = x offset)
    const int side = w / 4, k = w % 4;
    int hx, hy;
    if (side == 0) {
        hx = bx - 1; hy = by + k;
    } else if (side == 1) {
        hx = bx + ST_TX; hy = by + k;
    } else if (side == 2) {
        hx = bx + k; hy = by - 1;
    } else {
        hx = bx + k; hy = by + ST_TY;
    }
    hx = (hx + Lx) % Lx;
    hy = (hy + Ly) % Ly;
    const int plane = Lx * Ly;
    const int own_xy = ox + Lx * oy, halo_xy = hx + Lx * hy;
    const int i0 = rc * RW + lane * VEC;
    const bool full = (i0 + VEC <= no);
    const bool any = (i0 < no);
    // in-plane neighbours of the own site: their LDS slots (own tile or halo edge)
    const int slot_xm = (tx > 0) ? w - 1 : ST_WAVES + 0 * 4 + ty;
    const int slot_xp = (tx < ST_TX - 1) ? w + 1 : ST_WAVES + 1 * 4 + ty;
    const int slot_ym = (ty > 0) ? w - ST_TX : ST_WAVES + 2 * 4 + tx;
    const int slot_yp = (ty < ST_TY - 1) ? w + ST_TX : ST_WAVES + 3 * 4 + tx;
    const int xy_xm = (ox + Lx - 1) % Lx + Lx * oy, xy_xp = (ox + 1) % Lx + Lx * oy;
    const int xy_ym = ox + Lx * ((oy + Ly - 1) % Ly), xy_yp = ox + Lx * ((oy + 1) % Ly);

    auto load = [&](int site, T (&x)[4][VEC]) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const T *src = X + (int64_t)(4 * site + r) * ldx + i0;
            if (full) {
                vload<T, VEC>(src, x[r]);
            } else {
#pragma unroll
                for (int q = 0; q < VEC; ++q) x[r][q] = (any && i0 + q < no) ? src[q] : (T)0;
            }
        }
    };
    auto to_lds = [&](int buf, int slot, const T (&x)[4][VEC]) {
#pragma unroll
        for (int r = 0; r < 4; ++r) vstore<T, VEC>(&lds[buf][slot][r][lane * VEC], x[r]);
    };
    auto fma_block = [&](const T (&x)[4][VEC], const double *__restrict__ v, T (&acc)[4][VEC]) {
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const T a = (T)v[r + 4 * c];
#pragma unroll
                for (int q = 0; q < VEC; ++q) acc[c][q] = fma(x[r][q], a, acc[c][q]);
            }
    };
    // plane z of the march: xm, xc, xp hold the own site's planes z-1, z, z+1, hc the halo of plane
    // z+1 and buffer z&1 plane z of the tile and halo.  The step loads the own plane z+2 into xn and
    // the halo of plane z+2 into hn (a step ahead of their use), computes plane z, then stages plane
    // z+1 (xp, hc) into the other buffer.
    // Register rings, slots fixed at compile time (plane p of the own site in own[p % 5], halo plane
    // p in halo[p % 3]): the step of plane z computes from planes z-1, z, z+1, stages plane z+1 and
    // the halo of plane z+1 into the other LDS buffer, and loads the own plane z+3 and the halo of
    // plane z+3, two steps before their use.
    T own[5][4][VEC], halo[3][4][VEC];
    auto step = [&](int z, auto K) {
        constexpr int k = decltype(K)::value;  // z % 15
        T (&xm)[4][VEC] = own[(k + 4) % 5];
        T (&xc)[4][VEC] = own[k % 5];
        T (&xp)[4][VEC] = own[(k + 1) % 5];
        T (&hc)[4][VEC] = halo[(k + 1) % 3];
        const bool more = (z + 1 < Lz);
        if (z + 2 < Lz) load(own_xy + plane * ((z + 3) % Lz), own[(k + 3) % 5]);
        if (z + 3 < Lz) load(halo_xy + plane * (z + 3), halo[k % 3]);
        const int buf = z & 1;
        const int zo = plane * z;
        const int J = own_xy + zo;
        const int Jzm = own_xy + plane * ((z + Lz - 1) % Lz), Jzp = own_xy + plane * ((z + 1) % Lz);
        T acc[4][VEC];
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int q = 0; q < VEC; ++q) acc[c][q] = (T)0;
        // the stored (ascending block row) order of the 7 blocks: sort the neighbour sites, each
        // tagged with where its basis columns are (0 self, 1 z-1, 2 z+1, 3..6 the LDS slots of x-1,
        // x+1, y-1, y+1)
        int key[7] = {J, Jzm, Jzp, xy_xm + zo, xy_xp + zo, xy_ym + zo, xy_yp + zo};
        int dir[7] = {0, 1, 2, 3, 4, 5, 6};
        sort7(key, dir);
        int dpack = 0;
#pragma unroll
        for (int b = 0; b < 7; ++b) dpack |= dir[b] << (3 * b);
        // The 7 blocks of block column J are stored at 7J..7J+6 (every column of a recognised
        // lattice holds exactly 7), so their 112 values are contiguous at val + 112 J; J is
        // wave-uniform, so a block's 16 values come through the scalar cache.  A loop, not
        // unrolled: one block's values (32 scalar registers) live at a time.
        const double *vals = val + 112 * (int64_t)J;
#pragma unroll 1
        for (int b = 0; b < 7; ++b) {
            const double *a = vals + 16 * b;
            const int d = (dpack >> (3 * b)) & 7;
            if (d == 0) {
                fma_block(xc, a, acc);
            } else if (d == 1) {
                fma_block(xm, a, acc);
            } else if (d == 2) {
                fma_block(xp, a, acc);
            } else {
                const int slot = (d == 3) ? slot_xm : (d == 4) ? slot_xp : (d == 5) ? slot_ym : slot_yp;
                T x[4][VEC];
#pragma unroll
                for (int r = 0; r < 4; ++r) vload<T, VEC>(&lds[buf][slot][r][lane * VEC], x[r]);
                fma_block(x, a, acc);
            }
        }
        if (any) {
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                T *dst = Y + (int64_t)(4 * J + c) * ldy + i0;
                T out[VEC];
                if (full) {
                    if (beta != 0.0) {
                        T old[VEC];
                        vload<T, VEC>(dst, old);
#pragma unroll
                        for (int q = 0; q < VEC; ++q) out[q] = spmm_out(alpha, acc[c][q], beta, old[q]);
                    } else {
#pragma unroll
                        for (int q = 0; q < VEC; ++q) out[q] = spmm_out(alpha, acc[c][q], 0.0, (T)0);
                    }
                    vstore<T, VEC>(dst, out);
                } else {
#pragma unroll
                    for (int q = 0; q < VEC; ++q)
                        if (i0 + q < no) dst[q] = spmm_out(alpha, acc[c][q], beta, beta != 0.0 ? dst[q] : (T)0);
                }
            }
        }
        if (more) {
            // buffer (z+1)&1 was last read in the step of plane z-1, which every wave finished
            // before the barrier that ended it
            to_lds(buf ^ 1, w, xp);
            to_lds(buf ^ 1, ST_WAVES + w, hc);
        }
        lds_barrier();
    };

    load(own_xy + plane * (Lz - 1), own[4]);
    load(own_xy, own[0]);
    load(own_xy + plane, own[1]);
    load(own_xy + plane * 2, own[2]);
    load(halo_xy, halo[0]);
    load(halo_xy + plane, halo[1]);
    if (Lz > 2) load(halo_xy + plane * 2, halo[2]);
    to_lds(0, w, own[0]);
    to_lds(0, ST_WAVES + w, halo[0]);
    lds_barrier();
    for (int z0 = 0; z0 < Lz; z0 += 15) unroll_steps(step, z0, Lz, std::make_integer_sequence<int, 15>{});
}

template <typename T>
static void launch_spmm_stencil(omx_ctx ctx, int no, double alpha, const T *X, int64_t ldx, omx_bsr A, double beta,
                                T *Y, int64_t ldy) {
    constexpr int RW = 64 * (ST_VBYTES / sizeof(T));
    const int nrc = (no + RW - 1) / RW;
    const int ntx = A->lat[0] / ST_TX, nty = A->lat[1] / ST_TY;
    const unsigned grid = (unsigned)((int64_t)nrc * ntx * nty);
    hipLaunchKernelGGL((k_spmm_stencil<T, 2>), dim3(grid), dim3(64 * ST_WAVES), 0, ctx->stream, no, A->lat[0], A->lat[1],
                       A->lat[2], alpha, X, ldx, A->d_colptr, A->d_rowidx, A->d_val, beta, Y, ldy, nrc, ntx,
                       ntx * nty, ctx->launch_mask);
}

// The stencil kernel applies to a recognised lattice with Lx, Ly multiples of the 4x4 tile, with
// its 8-byte vectors aligned (FP32: no, the leading dimensions and the base addresses even).
template <typename T>
static bool stencil_applies(const omx_ctx ctx, const omx_bsr A, int no, const T *X, int64_t ldx, const T *Y,
                            int64_t ldy) {
    constexpr int V = ST_VBYTES / sizeof(T);
    const bool aligned = (no % V == 0) && (ldx % V == 0) && (ldy % V == 0) &&
                         ((uintptr_t)X % ST_VBYTES == 0) && ((uintptr_t)Y % ST_VBYTES == 0);
    return A->lat[0] > 0 && A->lat[0] % ST_TX == 0 && A->lat[1] % ST_TY == 0 && A->lat[2] >= 3 && !A->d_walk &&
           A->group == 0 && !ctx->spmm_forced && ctx->spmm_stencil && aligned && A->nb == 4 * A->nbc &&
           (int64_t)((no + 64 * V - 1) / (64 * V)) * (A->lat[0] / ST_TX) * (A->lat[1] / ST_TY) < (1LL << 31);
}

typedef omx_spmm_shape_s SpmmShape;

template <typename T, int VEC, int NB>
static void launch_spmm(omx_ctx ctx, const SpmmShape &sh, int no, int nb, double alpha, const T *X, int64_t ldx,
                        omx_bsr A, double beta, T *Y, int64_t ldy) {
    const int nrc = (no + 64 * VEC * sh.wpg - 1) / (64 * VEC * sh.wpg);
    const int njg = (A->nbc + sh.jper - 1) / sh.jper;
    const unsigned grid = (unsigned)((int64_t)nrc * njg);
#define OMX_SPMM_LAUNCH(W)                                                                                       \
    hipLaunchKernelGGL((k_spmm_bsr4<T, VEC, NB, W>), dim3(grid), dim3(64 * W), 0, ctx->stream, no, nb, A->nbc, \
                       alpha, X, ldx, A->d_colptr, A->d_rowidx, A->d_val, beta, Y, ldy, nrc, njg, sh.jper,   \
                       sh.mode, A->d_walk, ctx->launch_mask)
    if (sh.wpg == 1)
        OMX_SPMM_LAUNCH(1);
    else if (sh.wpg == 2)
        OMX_SPMM_LAUNCH(2);
    else
        OMX_SPMM_LAUNCH(4);
#undef OMX_SPMM_LAUNCH
}

template <typename T, int VEC>
static void launch_spmm_shape(omx_ctx ctx, const SpmmShape &sh, int no, int nb, double alpha, const T *X,
                              int64_t ldx, omx_bsr A, double beta, T *Y, int64_t ldy) {
    if (sh.batch >= 8)
        launch_spmm<T, VEC, 8>(ctx, sh, no, nb, alpha, X, ldx, A, beta, Y, ldy);
    else if (sh.batch >= 4)
        launch_spmm<T, VEC, 4>(ctx, sh, no, nb, alpha, X, ldx, A, beta, Y, ldy);
    else
        launch_spmm<T, VEC, 1>(ctx, sh, no, nb, alpha, X, ldx, A, beta, Y, ldy);
}

static int64_t spmm_grid(const SpmmShape &sh, int rows_per_wave, int no, int nbc) {
    const int64_t nrc = (no + (int64_t)rows_per_wave * sh.wpg - 1) / ((int64_t)rows_per_wave * sh.wpg);
    return nrc * ((nbc + sh.jper - 1) / sh.jper);
}

// Autotuner for the SpMM launch shape: the first time an operator shape is seen with tuning on,
// the candidate shapes are timed into a scratch result (one warm-up and one timed launch each)
// and the fastest is remembered.
template <typename T, int VEC>
static int tuned_spmm_shape(omx_ctx ctx, int no, int nb, const T *X, int64_t ldx, omx_bsr A, SpmmShape &out) {
    char key[160];
    snprintf(key, sizeof key, "spmm es=%d vec=%d no=%d nb=%d nnzb=%lld walk=%d", (int)sizeof(T), VEC, no, nb,
             (long long)A->nnzb, A->d_walk ? 1 : 0);
    auto it = ctx->tune_spmm.find(key);
    if (it != ctx->tune_spmm.end()) {
        out = it->second.first;
        return OMX_OK;
    }
    std::vector<SpmmShape> cands;
    for (int mode : {0, 1})
        for (int w : {1, 2, 4})
            for (int b : {4, 8})
                for (int jp : {1, 4}) {
                    SpmmShape sh;
                    sh.batch = b;
                    sh.wpg = w;
                    sh.jper = jp;
                    sh.mode = mode;
                    if (spmm_grid(sh, 64 * VEC, no, A->nbc) < (1LL << 31)) cands.push_back(sh);
                }
    OMX_REQUIRE(!cands.empty(), "omx_spmm: grid too large");
    T *tmp = nullptr;
    OMX_CHECK_HIP(hipMalloc((void **)&tmp, sizeof(T) * (size_t)no * nb));
    hipEvent_t e0, e1;
    OMX_CHECK_HIP(hipEventCreate(&e0));
    OMX_CHECK_HIP(hipEventCreate(&e1));
    SpmmShape best = cands[0];
    float best_ms = -1.0f;
    for (const SpmmShape &sh : cands) {
        float ms = 0.0f;
        for (int rep = 0; rep < 2; ++rep) {
            (void)hipEventRecord(e0, ctx->stream);
            launch_spmm_shape<T, VEC>(ctx, sh, no, nb, 1.0, X, ldx, A, 0.0, tmp, no);
            (void)hipEventRecord(e1, ctx->stream);
            (void)hipEventSynchronize(e1);
            (void)hipEventElapsedTime(&ms, e0, e1);
        }
        if (best_ms < 0.0f || ms < best_ms) {
            best = sh;
            best_ms = ms;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(tmp);
    OMX_CHECK_HIP(hipGetLastError());
    ctx->tune_spmm[key] = {best, (double)best_ms};
    out = best;
    return OMX_OK;
}

// LDS-staged gathers (the operator carries a group plan, omx_bsr_set_walk_groups).  The work
// items are (chunk of PR = 64 states, group of block columns consecutive in walk order).  For an
// item, a workgroup loads the distinct basis-column blocks the whole group gathers, for its 64
// states, into LDS (16-byte loads along the states); then each wave computes block columns of the
// group from LDS with one state per lane.  Each block column's blocks are taken in ascending
// (block, r) order with the same FMAs as k_spmm_bsr4, so the results are bitwise equal; what
// changes is that a block shared by several block columns of the group crosses from L2 once.
// The grid is persistent: 8*K workgroups, workgroup (x, i) on XCD x (dispatch is round-robin over
// the 8 XCDs) takes the items of that XCD's state chunks (x, x+8, ...) with stride K, so the K
// workgroups of an XCD walk neighbouring groups of one chunk together and share its L2.  The next
// item's gathers are loaded into registers while the current item computes from LDS.
constexpr int PATCH_ROWS = 64, PATCH_UMAX = 40, PATCH_NT = 256;

template <typename T>
__global__ __launch_bounds__(PATCH_NT, 2) void k_spmm_patch(int no, int nb, int nbc, double alpha,
                                                            const T *__restrict__ X, int64_t ldx,
                                                            const int *__restrict__ colptr,
                                                            const double *__restrict__ val,
                                                            const int *__restrict__ walk, int group, int ngroups,
                                                            const int *__restrict__ gptr,
                                                            const int *__restrict__ uidx,
                                                            const unsigned char *__restrict__ slot, double beta,
                                                            T *__restrict__ Y, int64_t ldy, int nrc,
                                                            const int *__restrict__ mask) {
    constexpr int PR = PATCH_ROWS;
    constexpr int VEC = 16 / sizeof(T);
    constexpr int VPC = PR / VEC;                               // vectors per block column of a chunk
    constexpr int NV = PATCH_UMAX * 4 * VPC / PATCH_NT;         // vectors per thread at most
    static_assert(NV * PATCH_NT == PATCH_UMAX * 4 * VPC, "LDS tile not divisible");
    typedef typename VecT<T, VEC>::type V;
    __shared__ V xs[PATCH_UMAX * 4 * VPC];
    if (mask && !*mask) return;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const bool xcd_map = (nrc & 7) == 0 && (gridDim.x & 7) == 0;
    const int K = xcd_map ? (int)gridDim.x >> 3 : (int)gridDim.x;
    const int x = xcd_map ? (int)blockIdx.x & 7 : 0, first = xcd_map ? (int)blockIdx.x >> 3 : (int)blockIdx.x;
    const int64_t nitems = xcd_map ? (int64_t)(nrc >> 3) * ngroups : (int64_t)nrc * ngroups;
    auto decode = [&](int64_t k, int &rc, int &g) {
        if (xcd_map) {
            rc = (int)(k / ngroups) * 8 + x;
            g = (int)(k % ngroups);
        } else {
            rc = (int)(k / ngroups);
            g = (int)(k % ngroups);
        }
    };
    const bool xal = (ldx % VEC == 0) && ((reinterpret_cast<uintptr_t>(X) & 15) == 0);
    V regs[NV];
    auto fetch = [&](int rc, int g) {
        const int r0 = rc * PR, u0 = gptr[g], nv = (gptr[g + 1] - u0) * 4 * VPC;
        const bool full = xal && (r0 + PR <= no);
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            const int e = tid + PATCH_NT * q;
            if (e < nv) {
                const int rv = (e % VPC) * VEC, uc = e / VPC;
                const int col = 4 * uidx[u0 + (uc >> 2)] + (uc & 3);
                const T *src = X + (int64_t)col * ldx + r0 + rv;
                if (col < nb && full) {
                    regs[q] = *reinterpret_cast<const V *>(src);
                } else {
                    T t[VEC];
#pragma unroll
                    for (int h = 0; h < VEC; ++h) t[h] = (col < nb && r0 + rv + h < no) ? src[h] : (T)0;
                    regs[q] = *reinterpret_cast<const V *>(t);
                }
            }
        }
    };
    int64_t k = first;
    int rc = 0, g = 0;
    if (k < nitems) {
        decode(k, rc, g);
        fetch(rc, g);
    }
    while (k < nitems) {
        const int nv = (gptr[g + 1] - gptr[g]) * 4 * VPC;
        __syncthreads();  // the previous item's reads of xs are done
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            const int e = tid + PATCH_NT * q;
            if (e < nv) xs[e] = regs[q];
        }
        __syncthreads();
        const int crc = rc, cg = g;
        const int64_t kn = k + K;
        if (kn < nitems) {  // next item's gathers in flight while this one computes
            decode(kn, rc, g);
            fetch(rc, g);
        }
        const T *xt = reinterpret_cast<const T *>(xs);
        const int i = crc * PR + lane;
        const int jb = cg * group, je = min(nbc, jb + group);
        for (int jj = jb + wv; jj < je; jj += PATCH_NT / 64) {
            const int J = walk ? walk[jj] : jj;
            T acc[4] = {(T)0, (T)0, (T)0, (T)0};
            const int lb = colptr[J], le = colptr[J + 1];
            // blocks in batches of 8: their slots and LDS reads are issued together, then the FMAs
            // run in ascending (block, r) order
            for (int l0 = lb; l0 < le; l0 += 8) {
                const int nblk = min(8, le - l0);
                int sl[8];
#pragma unroll
                for (int b = 0; b < 8; ++b) sl[b] = b < nblk ? (int)slot[l0 + b] : 0;
                T xr[8][4];
#pragma unroll
                for (int b = 0; b < 8; ++b)
#pragma unroll
                    for (int r = 0; r < 4; ++r) xr[b][r] = b < nblk ? xt[(sl[b] * 4 + r) * PR + lane] : (T)0;
#pragma unroll
                for (int b = 0; b < 8; ++b) {
                    if (b < nblk) {
                        const double *v = val + 16 * (int64_t)(l0 + b);
#pragma unroll
                        for (int r = 0; r < 4; ++r)
#pragma unroll
                            for (int c = 0; c < 4; ++c) acc[c] = fma(xr[b][r], (T)v[r + 4 * c], acc[c]);
                    }
                }
            }
            if (i >= no) continue;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int jcol = 4 * J + c;
                if (jcol >= nb) break;
                T *dst = Y + (int64_t)jcol * ldy + i;
                *dst = spmm_out(alpha, acc[c], beta, beta != 0.0 ? *dst : (T)0);
            }
        }
        k = kn;
    }
}

template <typename T>
static void launch_spmm_patch(omx_ctx ctx, int no, int nb, double alpha, const T *X, int64_t ldx, omx_bsr A,
                              double beta, T *Y, int64_t ldy) {
    const int nrc = (no + PATCH_ROWS - 1) / PATCH_ROWS;
    const int64_t items = (int64_t)nrc * A->ngroups;
    // two workgroups per CU, all resident; fewer when there are fewer items
    const int64_t want = 8LL * std::max(1, 2 * ctx->num_cus / 8);
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(want, items));
    hipLaunchKernelGGL(k_spmm_patch<T>, dim3(grid), dim3(PATCH_NT), 0, ctx->stream, no, nb, A->nbc, alpha, X, ldx,
                       A->d_colptr, A->d_val, A->d_walk, A->group, A->ngroups, A->d_gptr, A->d_uidx, A->d_slot, beta,
                       Y, ldy, nrc, ctx->launch_mask);
}

template <typename T, int VEC>
static int run_spmm(omx_ctx ctx, int no, int nb, double alpha, const T *X, int64_t ldx, omx_bsr A, double beta, T *Y,
                    int64_t ldy) {
    if (A->group > 0 && !ctx->spmm_forced) {
        OMX_REQUIRE((int64_t)((no + PATCH_ROWS - 1) / PATCH_ROWS) * A->ngroups < (1LL << 31), "omx_spmm: grid too large");
        hipEvent_t ev;
        prof_begin(ctx, OMX_K_SPMM, &ev);
        prof_work(ctx, OMX_K_SPMM, 2.0 * 16.0 * A->nnzb * no,
                  sizeof(T) * 2.0 * (double)no * nb * (beta != 0.0 ? 1.5 : 1.0) + A->nnzb * (16.0 * 8 + 4) +
                      4.0 * (A->nbc + 1));
        launch_spmm_patch<T>(ctx, no, nb, alpha, X, ldx, A, beta, Y, ldy);
        prof_end(ctx, OMX_K_SPMM, ev);
        OMX_CHECK_HIP(hipGetLastError());
        return OMX_OK;
    }
    if (stencil_applies<T>(ctx, A, no, X, ldx, Y, ldy)) {
        hipEvent_t ev;
        prof_begin(ctx, OMX_K_SPMM, &ev);
        prof_work(ctx, OMX_K_SPMM, 2.0 * 16.0 * A->nnzb * no,
                  sizeof(T) * 2.0 * (double)no * nb * (beta != 0.0 ? 1.5 : 1.0) + A->nnzb * (16.0 * 8 + 4) +
                      4.0 * (A->nbc + 1));
        launch_spmm_stencil<T>(ctx, no, alpha, X, ldx, A, beta, Y, ldy);
        prof_end(ctx, OMX_K_SPMM, ev);
        OMX_CHECK_HIP(hipGetLastError());
        return OMX_OK;
    }
    SpmmShape sh;
    if (ctx->spmm_forced) {
        sh = ctx->spmm_shape;  // forced (omx_spmm_set_tile)
    } else if (ctx->autotune) {
        int rc = tuned_spmm_shape<T, VEC>(ctx, no, nb, X, ldx, A, sh);
        if (rc) return rc;
    }
    OMX_REQUIRE(spmm_grid(sh, 64 * VEC, no, A->nbc) < (1LL << 31), "omx_spmm: grid too large");
    hipEvent_t ev;
    prof_begin(ctx, OMX_K_SPMM, &ev);
    prof_work(ctx, OMX_K_SPMM, 2.0 * 16.0 * A->nnzb * no,
              sizeof(T) * 2.0 * (double)no * nb * (beta != 0.0 ? 1.5 : 1.0) + A->nnzb * (16.0 * 8 + 4) + 4.0 * (A->nbc + 1));
    launch_spmm_shape<T, VEC>(ctx, sh, no, nb, alpha, X, ldx, A, beta, Y, ldy);
    prof_end(ctx, OMX_K_SPMM, ev);
    OMX_CHECK_HIP(hipGetLastError());
    return OMX_OK;
}

// Complex operator and complex X, Y (interleaved (re, im), element type R per part): the same
// band-major mapping; each lane holds VEC complex states.
template <typename R, int VEC, int WPG>
__global__ __launch_bounds__(64 * WPG) void k_spmm_bsr4_z(int no, int nb, int nbc, double alpha,
                                                        const R *__restrict__ X, int64_t ldx,
                                                        const int *__restrict__ colptr, const int *__restrict__ rowidx,
                                                        const double *__restrict__ val, double beta,
                                                        R *__restrict__ Y, int64_t ldy, int nbands, int jgroups,
                                                        int64_t sX, int64_t sY, int64_t sV,
                                                        const int *__restrict__ active) {
    constexpr int ROWS = 64 * VEC;
    constexpr int NB = 4;
    // batched use: blockIdx.y selects the system (operator values, X and Y offset by its strides)
    const int bsys = blockIdx.y;
    if (active && !active[bsys]) return;
    X += bsys * sX;
    Y += bsys * sY;
    val += bsys * sV;
    const int w = blockIdx.x;
    int band, jg;
    if ((nbands & 7) == 0) {
        const int xcd = w & 7, local = w >> 3;
        band = (local / jgroups) * 8 + xcd;
        jg = local % jgroups;
    } else {
        band = w / jgroups;
        jg = w % jgroups;
    }
    const int J = jg * WPG + (threadIdx.x >> 6);
    if (J >= nbc) return;  // whole wave; the kernel has no barriers
    const int i0 = band * ROWS + (threadIdx.x & 63) * VEC;
    const bool full = (i0 + VEC <= no);
    const bool any = (i0 < no);
    R ar[4][VEC], ai[4][VEC];
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int q = 0; q < VEC; ++q) ar[c][q] = ai[c][q] = (R)0;
    const int lb = colptr[J], le = colptr[J + 1];
    for (int l0 = lb; l0 < le; l0 += NB) {
        const int nblk = min(NB, le - l0);
        R xr[NB][4][VEC], xi[NB][4][VEC];
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            if (b < nblk) {
                const int I = rowidx[l0 + b];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int kcol = 4 * I + r;
                    const R *src = X + 2 * ((int64_t)kcol * ldx + i0);
                    R tmp[2 * VEC];
                    if (kcol < nb && full) {
                        vload<R, 2 * VEC>(src, tmp);
                    } else {
#pragma unroll
                        for (int q = 0; q < 2 * VEC; ++q) tmp[q] = (kcol < nb && any && i0 + q / 2 < no) ? src[q] : (R)0;
                    }
#pragma unroll
                    for (int q = 0; q < VEC; ++q) {
                        xr[b][r][q] = tmp[2 * q];
                        xi[b][r][q] = tmp[2 * q + 1];
                    }
                }
            }
        }
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            if (b < nblk) {
                const double *v = val + 32 * (int64_t)(l0 + b);
#pragma unroll
                for (int r = 0; r < 4; ++r)
#pragma unroll
                    for (int c = 0; c < 4; ++c) {
                        const R vr = (R)v[2 * (r + 4 * c)], vi = (R)v[2 * (r + 4 * c) + 1];
#pragma unroll
                        for (int q = 0; q < VEC; ++q) {
                            ar[c][q] = fma(xr[b][r][q], vr, ar[c][q]);
                            ar[c][q] = fma(-xi[b][r][q], vi, ar[c][q]);
                            ai[c][q] = fma(xr[b][r][q], vi, ai[c][q]);
                            ai[c][q] = fma(xi[b][r][q], vr, ai[c][q]);
                        }
                    }
            }
        }
    }
    if (!any) return;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int jcol = 4 * J + c;
        if (jcol >= nb) break;
        R *dst = Y + 2 * ((int64_t)jcol * ldy + i0);
#pragma unroll
        for (int q = 0; q < VEC; ++q) {
            if (!full && i0 + q >= no) continue;
            double orr = alpha * (double)ar[c][q], oii = alpha * (double)ai[c][q];
            if (beta != 0.0) {
                orr += beta * (double)dst[2 * q];
                oii += beta * (double)dst[2 * q + 1];
            }
            dst[2 * q] = (R)orr;
            dst[2 * q + 1] = (R)oii;
        }
    }
}

template <typename R, int VEC>
static int run_spmm_z(omx_ctx ctx, int no, int nb, double alpha, const R *X, int64_t ldx, omx_bsr A, double beta,
                      R *Y, int64_t ldy) {
    constexpr int ROWS = 64 * VEC, WPG = 4;
    const int nbands = (no + ROWS - 1) / ROWS;
    const int jgroups = (A->nbc + WPG - 1) / WPG;
    const int64_t total = (int64_t)nbands * jgroups;
    OMX_REQUIRE(total < (1LL << 31), "omx_spmm: grid too large");
    hipEvent_t ev;
    prof_begin(ctx, OMX_K_SPMM, &ev);
    prof_work(ctx, OMX_K_SPMM, 8.0 * 16.0 * A->nnzb * no,
              2.0 * sizeof(R) * 2.0 * (double)no * nb * (beta != 0.0 ? 1.5 : 1.0) + A->nnzb * (32.0 * 8 + 4) +
                  4.0 * (A->nbc + 1));
    hipLaunchKernelGGL((k_spmm_bsr4_z<R, VEC, WPG>), dim3((unsigned)total), dim3(64 * WPG), 0, ctx->stream, no, nb,
                       A->nbc, alpha, X, ldx, A->d_colptr, A->d_rowidx, A->d_val, beta, Y, ldy, nbands, jgroups,
                       (int64_t)0, (int64_t)0, (int64_t)0, (const int *)nullptr);
    prof_end(ctx, OMX_K_SPMM, ev);
    OMX_CHECK_HIP(hipGetLastError());
    return OMX_OK;
}

static bool aligned16(const void *p) { return ((uintptr_t)p & 15) == 0; }

// Batched complex SpMM (complex128, or complex64 with the FP64 operator values rounded to FP32)
// over nbatch systems sharing one block pattern: Y_s := X_s * A_s (beta 0), A_s values at
// val + s*sV (doubles), X_s, Y_s at s*sX, s*sY (complex elements).
int launch_spmm_z_batched(omx_ctx ctx, int dt, int no, int nb, int nbc, const int *colptr, const int *rowidx,
                          const double *val, int64_t sV, const void *X, int64_t ldx, int64_t sX, void *Y, int64_t ldy,
                          int64_t sY, int nbatch, const int *active) {
    constexpr int VEC = 1, ROWS = 64 * VEC, WPG = 4;
    const int nbands = (no + ROWS - 1) / ROWS;
    const int jgroups = (nbc + WPG - 1) / WPG;
    const int64_t total = (int64_t)nbands * jgroups;
    OMX_REQUIRE(total < (1LL << 31) && nbatch < 65536, "batched spmm: grid too large");
    if (dt == OMX_C64)
        hipLaunchKernelGGL((k_spmm_bsr4_z<float, VEC, WPG>), dim3((unsigned)total, nbatch), dim3(64 * WPG), 0,
                           ctx->stream, no, nb, nbc, 1.0, (const float *)X, ldx, colptr, rowidx, val, 0.0, (float *)Y,
                           ldy, nbands, jgroups, 2 * sX, 2 * sY, sV, active);
    else
        hipLaunchKernelGGL((k_spmm_bsr4_z<double, VEC, WPG>), dim3((unsigned)total, nbatch), dim3(64 * WPG), 0,
                           ctx->stream, no, nb, nbc, 1.0, (const double *)X, ldx, colptr, rowidx, val, 0.0,
                           (double *)Y, ldy, nbands, jgroups, 2 * sX, 2 * sY, sV, active);
    OMX_CHECK_HIP(hipGetLastError());
    return OMX_OK;
}

}  // namespace omx

using namespace omx;

extern "C" int omx_spmm_stencil(omx_ctx ctx, int enable) {
    OMX_REQUIRE(ctx, "omx_spmm_stencil: null context");
    ctx->spmm_stencil = (enable != 0);
    return OMX_OK;
}

extern "C" int omx_spmm_set_tile(omx_ctx ctx, int batch, int wpg, int jper, int mode) {
    OMX_REQUIRE(ctx, "omx_spmm_set_tile: null context");
    if (batch == 0) {
        ctx->spmm_forced = false;
        return OMX_OK;
    }
    OMX_REQUIRE((batch == 1 || batch == 4 || batch == 8) && (wpg == 1 || wpg == 2 || wpg == 4) && jper >= 1 &&
                    jper <= 4096 && (mode == 0 || mode == 1),
                "omx_spmm_set_tile: bad shape (batch %d, wpg %d, jper %d, mode %d)", batch, wpg, jper, mode);
    ctx->spmm_shape.batch = batch;
    ctx->spmm_shape.wpg = wpg;
    ctx->spmm_shape.jper = jper;
    ctx->spmm_shape.mode = mode;
    ctx->spmm_forced = true;
    return OMX_OK;
}

static void clear_patch_plan(omx_bsr A) {
    if (A->d_gptr) (void)hipFree(A->d_gptr);
    if (A->d_uidx) (void)hipFree(A->d_uidx);
    if (A->d_slot) (void)hipFree(A->d_slot);
    A->d_gptr = A->d_uidx = nullptr;
    A->d_slot = nullptr;
    A->group = A->ngroups = 0;
}

extern "C" int omx_bsr_set_walk(omx_bsr A, const int *order) {
    OMX_REQUIRE(A, "omx_bsr_set_walk: null operator");
    omx_ctx ctx = A->ctx;
    OMX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    clear_patch_plan(A);
    if (!order) {
        if (A->d_walk) OMX_CHECK_HIP(hipFree(A->d_walk));
        A->d_walk = nullptr;
        return OMX_OK;
    }
    std::vector<char> seen(A->nbc, 0);
    for (int j = 0; j < A->nbc; ++j) {
        OMX_REQUIRE(order[j] >= 0 && order[j] < A->nbc && !seen[order[j]],
                    "omx_bsr_set_walk: order is not a permutation of the %d block columns (entry %d)", A->nbc, j);
        seen[order[j]] = 1;
    }
    if (!A->d_walk) OMX_CHECK_HIP(hipMalloc((void **)&A->d_walk, sizeof(int) * std::max(A->nbc, 1)));
    OMX_CHECK_HIP(hipMemcpyAsync(A->d_walk, order, sizeof(int) * A->nbc, hipMemcpyHostToDevice, ctx->stream));
    OMX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    return OMX_OK;
}

extern "C" int omx_spmm(omx_ctx ctx, int dt, int no, int nb, double alpha, const void *X, int64_t ldx, omx_bsr A,
                        double beta, void *Y, int64_t ldy) {
    OMX_REQUIRE(ctx && A, "omx_spmm: null argument");
    OMX_REQUIRE(nb == A->nb, "omx_spmm: operator is %d x %d but X has %d columns", A->nb, A->nb, nb);
    OMX_REQUIRE(no >= 0 && ldx >= std::max(1, no) && ldy >= std::max(1, no), "omx_spmm: bad leading dimension");
    const bool zdt = (dt == OMX_C128 || dt == OMX_C64);
    OMX_REQUIRE(!A->is_complex || zdt, "omx_spmm: complex operator needs a complex element type");
    if (no == 0) return OMX_OK;
    OMX_REQUIRE(X && Y, "omx_spmm: null operand");
    OMX_REQUIRE(X != Y, "omx_spmm: X and Y must not alias");
    if (zdt) {
        OMX_REQUIRE(A->is_complex, "omx_spmm: complex X needs a complex operator");
        if (dt == OMX_C128) return run_spmm_z<double, 1>(ctx, no, nb, alpha, (const double *)X, ldx, A, beta, (double *)Y, ldy);
        const bool vec = (no % 2 == 0) && (ldx % 2 == 0) && (ldy % 2 == 0) && aligned16(X) && aligned16(Y);
        if (vec) return run_spmm_z<float, 2>(ctx, no, nb, alpha, (const float *)X, ldx, A, beta, (float *)Y, ldy);
        return run_spmm_z<float, 1>(ctx, no, nb, alpha, (const float *)X, ldx, A, beta, (float *)Y, ldy);
    }
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

// =============================================================================================
// Dense x dense -> block-sparse pattern (SDDMM): Cp := alpha * (A^T B) restricted to the stored
// blocks of Cp + beta * Cp, with A, B no x nb column-major.  Block (I,J), entry (r,c) is the dot
// product over the state index of basis columns 4I+r of A and 4J+c of B.  This is the
// reference's dense*dense -> CSC product (MatrixSwitch/src/MatrixSwitch_mm_multiply.F90:334-398,
// op = 'c','n'), used for the density matrix D = C^T (QW C) on the pattern of H
// (libOMM/src/omm_ops.F90:104-124 with a sparse D_min) and for dE/dH.
//
// Mapping: a workgroup owns a chunk of SD_NT*VEC*SD_STR states and one block column J; it holds
// B's four columns 4J..4J+3 of the chunk in registers and, per stored block, reads A's four
// columns, accumulates the 16 products per thread over its SD_STR strides and reduces them over
// each wave (a reduce-scatter, no barrier).  Per-wave partial blocks go to the workspace and a
// second kernel adds them in (chunk, wave) order (deterministic).
// =============================================================================================
namespace omx {

constexpr int SD_NT = 256;
// strides of SD_NT*VEC states per thread: one workgroup reduction per stored block per SD_STR strides
constexpr int SD_STR = 4;

template <typename T, int VEC>
__global__ __launch_bounds__(SD_NT) void k_sddmm_bsr4_partial(int no, int nb, const T *__restrict__ A, int64_t lda,
                                                              const T *__restrict__ B, int64_t ldb,
                                                              const int *__restrict__ colptr,
                                                              const int *__restrict__ rowidx, int rchunks,
                                                              int64_t nnzb, double *__restrict__ partial) {
    constexpr int ROWS = SD_NT * VEC * SD_STR;
    const int rc = blockIdx.x % rchunks, J = blockIdx.x / rchunks;
    double bcol[SD_STR][4][VEC];
#pragma unroll
    for (int st = 0; st < SD_STR; ++st) {
        const int i0 = rc * ROWS + st * SD_NT * VEC + threadIdx.x * VEC;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int j = 4 * J + c;
#pragma unroll
            for (int q = 0; q < VEC; ++q)
                bcol[st][c][q] = (j < nb && i0 + q < no) ? (double)B[(int64_t)j * ldb + i0 + q] : 0.0;
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int l = colptr[J]; l < colptr[J + 1]; ++l) {
        const int I = rowidx[l];
        double s[16];
#pragma unroll
        for (int e = 0; e < 16; ++e) s[e] = 0.0;
#pragma unroll
        for (int st = 0; st < SD_STR; ++st) {
            const int i0 = rc * ROWS + st * SD_NT * VEC + threadIdx.x * VEC;
            const bool full = (i0 + VEC <= no);
            double acol[4][VEC];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int k = 4 * I + r;
                const T *src = A + (int64_t)k * lda + i0;
#pragma unroll
                for (int q = 0; q < VEC; ++q) acol[r][q] = (k < nb && (full || i0 + q < no)) ? (double)src[q] : 0.0;
            }
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int c = 0; c < 4; ++c)
#pragma unroll
                    for (int q = 0; q < VEC; ++q) s[r + 4 * c] = fma(acol[r][q], bcol[st][c][q], s[r + 4 * c]);
        }
        // Reduce-scatter of the 16 sums over the wave by recursive halving: at offset 32 a lane keeps
        // half of its values and adds its partner's copy of that half (8 shuffles), then 4, 2 and 1;
        // lane l ends with output (l >> 2) & 15, and two more steps add the four lanes that share
        // it.  17 shuffles instead of 16 x 6; the order is fixed, and each wave writes its own
        // partial block, so no barrier is needed.
        double h8[8], h4[4], h2[2];
        const bool b5 = lane & 32, b4 = lane & 16, b3 = lane & 8, b2 = lane & 4;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const double keep = b5 ? s[e + 8] : s[e], send = b5 ? s[e] : s[e + 8];
            h8[e] = keep + __shfl_xor(send, 32, 64);
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const double keep = b4 ? h8[e + 4] : h8[e], send = b4 ? h8[e] : h8[e + 4];
            h4[e] = keep + __shfl_xor(send, 16, 64);
        }
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            const double keep = b3 ? h4[e + 2] : h4[e], send = b3 ? h4[e] : h4[e + 2];
            h2[e] = keep + __shfl_xor(send, 8, 64);
        }
        double t = (b2 ? h2[1] : h2[0]) + __shfl_xor(b2 ? h2[0] : h2[1], 4, 64);
        t += __shfl_xor(t, 1, 64);
        t += __shfl_xor(t, 2, 64);
        if ((lane & 3) == 0)
            partial[(((int64_t)rc * (SD_NT / 64) + wave) * nnzb + l) * 16 + ((lane >> 2) & 15)] = t;
    }
}

__global__ void k_sddmm_bsr4_final(int64_t nnzb, int nparts, const double *__restrict__ partial, double alpha,
                                   double beta, double *__restrict__ vals) {
    const int64_t total = nnzb * 16;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += (int64_t)gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int p = 0; p < nparts; ++p) s += partial[(int64_t)p * total + q];
        vals[q] = (beta == 0.0) ? alpha * s : alpha * s + beta * vals[q];
    }
}

template <typename T, int VEC>
static int run_sddmm(omx_ctx ctx, int no, int nb, double alpha, const T *A, int64_t lda, const T *B, int64_t ldb,
                     double beta, omx_bsr C) {
    constexpr int ROWS = SD_NT * VEC * SD_STR;
    const int rchunks = (no + ROWS - 1) / ROWS;
    const int64_t grid = (int64_t)rchunks * C->nbc;
    OMX_REQUIRE(grid < (1LL << 31), "omx_sddmm_bsr: grid too large");
    const int nparts = rchunks * (SD_NT / 64);  // one partial block per wave of every chunk
    const size_t bytes = sizeof(double) * 16 * (size_t)C->nnzb * nparts;
    int rc = ensure_ws(ctx, bytes);
    if (rc) return rc;
    double *partial = (double *)ctx->ws;
    hipEvent_t ev;
    prof_begin(ctx, OMX_K_SPMM, &ev);
    prof_work(ctx, OMX_K_SPMM, 2.0 * 16.0 * C->nnzb * no, sizeof(T) * 2.0 * (double)no * nb + C->nnzb * (16.0 * 8 + 4));
    hipLaunchKernelGGL((k_sddmm_bsr4_partial<T, VEC>), dim3((unsigned)grid), dim3(SD_NT), 0, ctx->stream, no, nb, A,
                       lda, B, ldb, C->d_colptr, C->d_rowidx, rchunks, C->nnzb, partial);
    const int64_t total = C->nnzb * 16;
    const int g2 = (int)std::max<int64_t>(1, std::min<int64_t>((total + 255) / 256, 8192));
    hipLaunchKernelGGL(k_sddmm_bsr4_final, dim3(g2), dim3(256), 0, ctx->stream, C->nnzb, nparts, partial, alpha, beta,
                       C->d_val);
    prof_end(ctx, OMX_K_SPMM, ev);
    OMX_CHECK_HIP(hipGetLastError());
    return OMX_OK;
}

}  // namespace omx

extern "C" int omx_sddmm_bsr(omx_ctx ctx, int dt, int no, int nb, double alpha, const void *A, int64_t lda,
                             const void *B, int64_t ldb, double beta, omx_bsr C) {
    OMX_REQUIRE(ctx && C, "omx_sddmm_bsr: null argument");
    OMX_REQUIRE(nb == C->nb && !C->is_complex, "omx_sddmm_bsr: pattern does not match nb");
    OMX_REQUIRE(no >= 0 && lda >= std::max(1, no) && ldb >= std::max(1, no), "omx_sddmm_bsr: bad leading dimension");
    if (C->nnzb == 0) return OMX_OK;
    if (no == 0) {  // empty contraction: Cp := beta * Cp
        const int g2 = (int)std::max<int64_t>(1, std::min<int64_t>((C->nnzb * 16 + 255) / 256, 8192));
        hipLaunchKernelGGL(k_sddmm_bsr4_final, dim3(g2), dim3(256), 0, ctx->stream, C->nnzb, 0, (const double *)nullptr,
                           alpha, beta, C->d_val);
        OMX_CHECK_HIP(hipGetLastError());
        return OMX_OK;
    }
    OMX_REQUIRE(A && B, "omx_sddmm_bsr: null operand");
    if (dt == OMX_F64) return run_sddmm<double, 2>(ctx, no, nb, alpha, (const double *)A, lda, (const double *)B, ldb, beta, C);
    if (dt == OMX_F32) return run_sddmm<float, 2>(ctx, no, nb, alpha, (const float *)A, lda, (const float *)B, ldb, beta, C);
    set_error("omx_sddmm_bsr: unsupported element type %d", dt);
    return OMX_ERR_UNSUPPORTED;
}

extern "C" int omx_bsr_set_walk_groups(omx_bsr A, const int *order, int group) {
    OMX_REQUIRE(A && group >= 0, "omx_bsr_set_walk_groups: bad arguments");
    int rc = omx_bsr_set_walk(A, order);
    if (rc || group == 0) return rc;
    OMX_REQUIRE(!A->is_complex, "omx_bsr_set_walk_groups: real operators only");
    omx_ctx ctx = A->ctx;
    const int nbc = A->nbc;
    std::vector<int> colptr(nbc + 1), rowidx(std::max<int64_t>(A->nnzb, 1));
    OMX_CHECK_HIP(hipMemcpy(colptr.data(), A->d_colptr, sizeof(int) * (nbc + 1), hipMemcpyDeviceToHost));
    if (A->nnzb) OMX_CHECK_HIP(hipMemcpy(rowidx.data(), A->d_rowidx, sizeof(int) * A->nnzb, hipMemcpyDeviceToHost));
    const int ngroups = (nbc + group - 1) / group;
    std::vector<int> gptr(ngroups + 1, 0), uidx;
    std::vector<unsigned char> slot(std::max<int64_t>(A->nnzb, 1), 0);
    std::vector<int> where(nbc, -1);
    for (int g = 0; g < ngroups; ++g) {
        const int jb = g * group, je = std::min(nbc, jb + group);
        std::vector<int> u;
        for (int jj = jb; jj < je; ++jj) {
            const int J = order ? order[jj] : jj;
            for (int l = colptr[J]; l < colptr[J + 1]; ++l) u.push_back(rowidx[l]);
        }
        std::sort(u.begin(), u.end());
        u.erase(std::unique(u.begin(), u.end()), u.end());
        if ((int)u.size() > PATCH_UMAX) {
            (void)omx_bsr_set_walk(A, order);  // keep the walk, without the plan
            set_error("omx_bsr_set_walk_groups: group %d gathers %d distinct blocks (at most %d)", g, (int)u.size(),
                      PATCH_UMAX);
            return OMX_ERR_ARG;
        }
        for (size_t q = 0; q < u.size(); ++q) where[u[q]] = (int)q;
        for (int jj = jb; jj < je; ++jj) {
            const int J = order ? order[jj] : jj;
            for (int l = colptr[J]; l < colptr[J + 1]; ++l) slot[l] = (unsigned char)where[rowidx[l]];
        }
        uidx.insert(uidx.end(), u.begin(), u.end());
        gptr[g + 1] = (int)uidx.size();
    }
    bool ok = hipMalloc((void **)&A->d_gptr, sizeof(int) * (ngroups + 1)) == hipSuccess &&
              hipMalloc((void **)&A->d_uidx, sizeof(int) * std::max<size_t>(uidx.size(), 1)) == hipSuccess &&
              hipMalloc((void **)&A->d_slot, slot.size()) == hipSuccess;
    if (!ok) {
        clear_patch_plan(A);
        set_error("omx_bsr_set_walk_groups: device allocation failed");
        return OMX_ERR_NOMEM;
    }
    OMX_CHECK_HIP(hipMemcpy(A->d_gptr, gptr.data(), sizeof(int) * (ngroups + 1), hipMemcpyHostToDevice));
    if (!uidx.empty())
        OMX_CHECK_HIP(hipMemcpy(A->d_uidx, uidx.data(), sizeof(int) * uidx.size(), hipMemcpyHostToDevice));
    OMX_CHECK_HIP(hipMemcpy(A->d_slot, slot.data(), slot.size(), hipMemcpyHostToDevice));
    A->group = group;
    A->ngroups = ngroups;
    (void)ctx;
    return OMX_OK;
}
