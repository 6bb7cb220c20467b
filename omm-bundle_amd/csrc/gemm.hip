// Dense products on the CDNA4 matrix cores: C := alpha*op(A)*op(B) + beta*C.
//
// One kernel template covers the OMM products
//   Gram      (No x Nb)(Nb x No)  -> op = ('n','t'): HWd = HC*D^T, SWd = SC*D^T, HWdd = HG*D^T, ...
//   gradient  (No x No)(No x Nb)  -> op = ('n','n'): SW*HC, HW*SC (fused pair: calc_G)
//   density   (Nb x No)(No x Nb)  -> op = ('t','n'): D_min = C^T*(QW*C)
// and the three precisions:
//   FP64: v_mfma_f64_16x16x4f64, FP64 accumulate (the oracle-grade path);
//   FP32: operands rounded to FP32, v_mfma_f32_16x16x4f32, FP32 accumulate;
//   BF16: operands rounded to BF16 (one rounding from the stored value),
//         v_mfma_f32_16x16x32_bf16, FP32 accumulate.
//
// Work decomposition: a workgroup of WGM x WGN wave64s owns a BM x BN tile of C; each wave owns
// a (BM/WGM) x (BN/WGN) sub-tile held as 16x16 MFMA accumulator blocks.  K is consumed in BK
// slices staged through LDS and double-buffered: the next slice is fetched from HBM/L2 into
// registers (16-byte vector loads along the contiguous index) while the current slice feeds the
// MFMAs.  When the tile count cannot fill the chip, K is split over grid.z into partial tiles
// that a second kernel adds in split order (deterministic).  The linear workgroup id is remapped
// so that each XCD (ids are dealt round-robin over the 8 XCDs) walks a contiguous band of tiles,
// grouped GM tile-rows at a time, so operand panels are shared inside one XCD's L2.
#include <algorithm>
#include <type_traits>

#include "omx_internal.h"

namespace omx {

typedef double d4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));

// ---------------------------------------------------------------------------------------------
// Precision traits: K slice, LDS element type, accumulator type and result lane layout.
// ---------------------------------------------------------------------------------------------
template <int PREC>
struct Prec;

template <>
struct Prec<OMX_PREC_F64> {
    static constexpr int BK = 16;
    typedef double lds_t;
    typedef double acc_t;
    typedef d4v accv;
    // v_mfma_f64_16x16x4f64 result: lane l, register pair v -> row (l/16) + 4v, column l%16
    __device__ static inline int out_row(int lq, int v) { return lq + 4 * v; }
};
template <>
struct Prec<OMX_PREC_F32> {
    static constexpr int BK = 32;
    typedef float lds_t;
    typedef float acc_t;
    typedef f4v accv;
    // v_mfma_f32_16x16x4f32 result: lane l, register v -> row 4(l/16) + v, column l%16
    __device__ static inline int out_row(int lq, int v) { return 4 * lq + v; }
};
template <>
struct Prec<OMX_PREC_BF16> {
    static constexpr int BK = 64;
    typedef __bf16 lds_t;
    typedef float acc_t;
    typedef f4v accv;
    // v_mfma_f32_16x16x32_bf16 result: lane l, register v -> row 4(l/16) + v, column l%16
    __device__ static inline int out_row(int lq, int v) { return 4 * lq + v; }
};

template <typename TS, int PREC>
__device__ inline typename Prec<PREC>::lds_t to_lds(TS v) {
    return (typename Prec<PREC>::lds_t)v;
}

// ---------------------------------------------------------------------------------------------
// Tile configurations.
// ---------------------------------------------------------------------------------------------
template <int BM_, int BN_, int WGM_, int WGN_, int MINW_>
struct Cfg {
    static constexpr int BM = BM_, BN = BN_, WGM = WGM_, WGN = WGN_;
    static constexpr int NT = 64 * WGM * WGN;
    static constexpr int WM = BM / WGM, WN = BN / WGN;
    static constexpr int FBM = WM / 16, FBN = WN / 16;
    static constexpr int MINW = MINW_;  // waves per SIMD the register budget must allow
};
// Ids used by the selector and by omx_gemm_set_tile().
typedef Cfg<128, 128, 2, 2, 2> Cfg0;  // 4 waves, 64x64 per wave, two workgroups per CU
typedef Cfg<128, 128, 2, 2, 1> Cfg1;  // 4 waves, 64x64 per wave, one workgroup per CU
typedef Cfg<256, 128, 2, 2, 1> Cfg2;  // 4 waves, 128x64 per wave
typedef Cfg<128, 256, 2, 2, 1> Cfg3;  // 4 waves, 64x128 per wave
typedef Cfg<256, 256, 2, 4, 1> Cfg4;  // 8 waves, 128x64 per wave (BF16 / FP32 products)
typedef Cfg<128, 128, 4, 2, 2> Cfg5;  // 8 waves, 32x64 per wave, two workgroups per CU (FP64)
typedef Cfg<128, 128, 2, 4, 2> Cfg6;  // 8 waves, 64x32 per wave, two workgroups per CU (FP64)
constexpr int kNumCfg = 8;  // 7: Cfg5 tiles with direct-to-LDS staging (FP64 products, FP32 Grams)

// LDS layouts.  FP64/FP32: k-major ([k][i], i contiguous): one element per lane per MFMA step,
// lanes 0..15 read 16 consecutive i.  BF16: i-major ([i][k], k contiguous): each lane feeds 8
// consecutive k values per MFMA step as one 16-byte read.
template <int PREC, int ROWS>
struct LdsT {
    static constexpr int BK = Prec<PREC>::BK;
    static constexpr bool KMAJOR = (PREC != OMX_PREC_BF16);
    static constexpr int PAD = KMAJOR ? 4 : 8;
    static constexpr int ELEMS = KMAJOR ? BK * (ROWS + PAD) : ROWS * (BK + PAD);
    __device__ static inline int off(int r, int k) { return KMAJOR ? k * (ROWS + PAD) + r : r * (BK + PAD) + k; }
};

// ---------------------------------------------------------------------------------------------
// Global -> register staging of one BK slice of op(A) (BM x BK) or op(B)^T (BN x BK).
// "Row" index r (i for A, j for B) and k.  CONTIG_R: r is contiguous in memory (A 'n', B 't');
// otherwise k is (A 't', B 'n').  Loads are VEC-wide vectors along the contiguous index when
// the vector lies inside the matrix and the leading dimension keeps it 16-byte aligned.
// ---------------------------------------------------------------------------------------------
template <typename TS, int PREC, bool CONTIG_R, int ROWS, int NT>
struct Stage {
    static constexpr int BK = Prec<PREC>::BK;
    static constexpr int VEC = 16 / sizeof(TS);
    // BF16 with contiguous rows: a thread owns a micro-tile of TR rows x TK consecutive k, loaded
    // as TK vectors along the rows and written to the i-major LDS tile as TR runs of TK k values
    // (a register transpose).  Every other case: vectors along the contiguous index.
    static constexpr bool MICRO = (PREC == OMX_PREC_BF16 && CONTIG_R);
    static constexpr int TR = 4, TK = 8;
    static constexpr int NE = ROWS * BK / NT;  // elements per thread
    static constexpr int NV = NE / VEC;
    static_assert(NE % VEC == 0, "slice not divisible into vectors");
    static_assert(!MICRO || (NE == TR * TK && TR % VEC == 0), "micro-tile shape");
    typename Prec<PREC>::lds_t v[NE];

    // Vector q of this thread -> (r, k) of its first element (non-MICRO modes).
    __device__ static inline void coord(int t, int q, int &r, int &k) {
        if constexpr (CONTIG_R) {
            int idx = t + NT * q;
            r = (idx % (ROWS / VEC)) * VEC;
            k = idx / (ROWS / VEC);
        } else {
            int idx = t + NT * q;
            k = (idx % (BK / VEC)) * VEC;
            r = idx / (BK / VEC);
        }
    }
    // MICRO mode: micro-tile origin of thread t.
    __device__ static inline void micro_origin(int t, int &r, int &k) {
        r = (t % (ROWS / TR)) * TR;
        k = (t / (ROWS / TR)) * TK;
    }

    __device__ inline void load(const TS *__restrict__ P, int64_t ld, int r0, int k0, int rlim, int klim, bool vec_ok,
                                int t) {
        if constexpr (MICRO) {
            int r, k;
            micro_origin(t, r, k);
            const int gr = r0 + r;
#pragma unroll
            for (int kk = 0; kk < TK; ++kk) {
                const int gk = k0 + k + kk;
                const TS *src = P + (int64_t)gr + (int64_t)gk * ld;
                TS x[TR];
                if (vec_ok && gr + TR <= rlim && gk < klim && TR % VEC == 0) {
                    typedef TS V __attribute__((ext_vector_type(VEC)));
#pragma unroll
                    for (int h = 0; h < TR / VEC; ++h) {
                        V w = *reinterpret_cast<const V *>(src + h * VEC);
#pragma unroll
                        for (int e = 0; e < VEC; ++e) x[h * VEC + e] = w[e];
                    }
                } else {
#pragma unroll
                    for (int e = 0; e < TR; ++e) x[e] = (gr + e < rlim && gk < klim) ? src[e] : (TS)0;
                }
#pragma unroll
                for (int e = 0; e < TR; ++e) v[e * TK + kk] = to_lds<TS, PREC>(x[e]);
            }
        } else {
#pragma unroll
            for (int q = 0; q < NV; ++q) {
                int r, k;
                coord(t, q, r, k);
                const int gr = r0 + r, gk = k0 + k;
                const bool full = CONTIG_R ? (gr + VEC <= rlim && gk < klim) : (gr < rlim && gk + VEC <= klim);
                const TS *src = CONTIG_R ? P + (int64_t)gr + (int64_t)gk * ld : P + (int64_t)gk + (int64_t)gr * ld;
                TS x[VEC];
                if (vec_ok && full) {
                    typedef TS V __attribute__((ext_vector_type(VEC)));
                    V w = *reinterpret_cast<const V *>(src);
#pragma unroll
                    for (int e = 0; e < VEC; ++e) x[e] = w[e];
                } else {
#pragma unroll
                    for (int e = 0; e < VEC; ++e) {
                        const bool in = CONTIG_R ? (gr + e < rlim && gk < klim) : (gr < rlim && gk + e < klim);
                        x[e] = in ? src[e] : (TS)0;
                    }
                }
#pragma unroll
                for (int e = 0; e < VEC; ++e) v[q * VEC + e] = to_lds<TS, PREC>(x[e]);
            }
        }
    }

    __device__ inline void store(typename Prec<PREC>::lds_t *lds, int t) const {
        if constexpr (MICRO) {
            // one 16-byte write of TK = 8 consecutive k per row of the micro-tile
            int r, k;
            micro_origin(t, r, k);
#pragma unroll
            for (int e = 0; e < TR; ++e) {
                bf8v w;
#pragma unroll
                for (int kk = 0; kk < TK; ++kk) w[kk] = v[e * TK + kk];
                *reinterpret_cast<bf8v *>(&lds[LdsT<PREC, ROWS>::off(r + e, k)]) = w;
            }
        } else {
#pragma unroll
            for (int q = 0; q < NV; ++q) {
                int r, k;
                coord(t, q, r, k);
#pragma unroll
                for (int e = 0; e < VEC; ++e) {
                    if (CONTIG_R)
                        lds[LdsT<PREC, ROWS>::off(r + e, k)] = v[q * VEC + e];
                    else
                        lds[LdsT<PREC, ROWS>::off(r, k + e)] = v[q * VEC + e];
                }
            }
        }
    }
};

// ---------------------------------------------------------------------------------------------
// MFMA step over one BK slice held in LDS.
// ---------------------------------------------------------------------------------------------
template <int PREC, class CF>
__device__ inline void mma_slice(const typename Prec<PREC>::lds_t *As, const typename Prec<PREC>::lds_t *Bs,
                                 typename Prec<PREC>::accv (&acc)[CF::FBM][CF::FBN], int wm, int wn, int lane) {
    constexpr int BK = Prec<PREC>::BK;
    typedef LdsT<PREC, CF::BM> LA;
    typedef LdsT<PREC, CF::BN> LB;
    const int li = lane & 15, lk = lane >> 4;
    if constexpr (PREC == OMX_PREC_F64 || PREC == OMX_PREC_F32) {
        typedef typename Prec<PREC>::lds_t T;
#pragma unroll
        for (int ks = 0; ks < BK; ks += 4) {
            T a[CF::FBM], b[CF::FBN];
#pragma unroll
            for (int f = 0; f < CF::FBM; ++f) a[f] = As[LA::off(wm * CF::WM + f * 16 + li, ks + lk)];
#pragma unroll
            for (int f = 0; f < CF::FBN; ++f) b[f] = Bs[LB::off(wn * CF::WN + f * 16 + li, ks + lk)];
#pragma unroll
            for (int fm = 0; fm < CF::FBM; ++fm)
#pragma unroll
                for (int fn = 0; fn < CF::FBN; ++fn) {
                    if constexpr (PREC == OMX_PREC_F64)
                        acc[fm][fn] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[fm], b[fn], acc[fm][fn], 0, 0, 0);
                    else
                        acc[fm][fn] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[fm], b[fn], acc[fm][fn], 0, 0, 0);
                }
        }
    } else {
#pragma unroll
        for (int ks = 0; ks < BK; ks += 32) {
            bf8v a[CF::FBM], b[CF::FBN];
#pragma unroll
            for (int f = 0; f < CF::FBM; ++f)
                a[f] = *reinterpret_cast<const bf8v *>(&As[LA::off(wm * CF::WM + f * 16 + li, ks + 8 * lk)]);
#pragma unroll
            for (int f = 0; f < CF::FBN; ++f)
                b[f] = *reinterpret_cast<const bf8v *>(&Bs[LB::off(wn * CF::WN + f * 16 + li, ks + 8 * lk)]);
#pragma unroll
            for (int fm = 0; fm < CF::FBM; ++fm)
#pragma unroll
                for (int fn = 0; fn < CF::FBN; ++fn)
                    acc[fm][fn] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[fm], b[fn], acc[fm][fn], 0, 0, 0);
        }
    }
}

// Pair mode (NPAIR = 2): C := alpha*(A1*B1 + A2*B2) + beta0*C0 + beta*C over a K range made of
// the two contractions back to back (calc_G).  NPAIR = 1: the plain product.
template <typename TS, int PREC, bool TRA, bool TRB, int NPAIR, class CF>
__global__ __launch_bounds__(CF::NT, CF::MINW) void k_gemm(int m, int n, int k, double alpha, const TS *__restrict__ A1,
                                                 const TS *__restrict__ A2, int64_t lda, const TS *__restrict__ B1,
                                                 const TS *__restrict__ B2, int64_t ldb, double beta0,
                                                 const TS *__restrict__ C0, double beta, TS *__restrict__ C,
                                                 int64_t ldc, int k_per_split, int vec_a, int vec_b, int sym,
                                                 typename Prec<PREC>::acc_t *__restrict__ partial,
                                                 const int *__restrict__ mask) {
    typedef typename Prec<PREC>::lds_t LT;
    typedef typename Prec<PREC>::acc_t AT;
    typedef typename Prec<PREC>::accv AV;
    typedef LdsT<PREC, CF::BM> LA;
    typedef LdsT<PREC, CF::BN> LB;
    constexpr int BK = Prec<PREC>::BK;
    __shared__ LT smem[2 * (LA::ELEMS + LB::ELEMS)];
    if (mask && !*mask) return;  // whole workgroup, before any barrier

    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = wave / CF::WGN, wn = wave % CF::WGN;
    const int tiles_m = (m + CF::BM - 1) / CF::BM, tiles_n = (n + CF::BN - 1) / CF::BN;
    int tm, tn;
    if (sym)
        tile_of_lower(blockIdx.x, tiles_m, tm, tn);
    else
        tile_of(blockIdx.x, tiles_m, tiles_n, tm, tn);
    const int i0 = tm * CF::BM, j0 = tn * CF::BN;

    // K range of this split over the virtual contraction [0, NPAIR*k).
    const int ktot = NPAIR * k;
    const int kb = blockIdx.z * k_per_split;
    const int ke = min(ktot, kb + k_per_split);

    AV acc[CF::FBM][CF::FBN];
#pragma unroll
    for (int a = 0; a < CF::FBM; ++a)
#pragma unroll
        for (int b = 0; b < CF::FBN; ++b) acc[a][b] = AV{0, 0, 0, 0};

    Stage<TS, PREC, !TRA, CF::BM, CF::NT> sa;
    Stage<TS, PREC, TRB, CF::BN, CF::NT> sb;

    auto fetch = [&](int kk) {
        // kk is a virtual K index; pick the operand pair it falls in.
        const TS *Ap = A1, *Bp = B1;
        int kl = kk;
        if (NPAIR == 2 && kk >= k) {
            Ap = A2;
            Bp = B2;
            kl = kk - k;
        }
        int klim = min(k, kl + (ke - kk));  // stop at the split end and at the pair end
        sa.load(Ap, lda, i0, kl, m, klim, vec_a != 0, t);
        sb.load(Bp, ldb, j0, kl, n, klim, vec_b != 0, t);
    };

    LT *As0 = smem, *Bs0 = smem + LA::ELEMS;
    LT *As1 = smem + LA::ELEMS + LB::ELEMS, *Bs1 = As1 + LA::ELEMS;

    // In pair mode the host guarantees k % BK == 0, so no slice straddles the pair boundary.
    const int nslices = (ke - kb + BK - 1) / BK;
    if (nslices > 0) {
        fetch(kb);
        sa.store(As0, t);
        sb.store(Bs0, t);
        __syncthreads();
        for (int s = 0; s < nslices; ++s) {
            LT *Ac = (s & 1) ? As1 : As0, *Bc = (s & 1) ? Bs1 : Bs0;
            LT *An = (s & 1) ? As0 : As1, *Bn = (s & 1) ? Bs0 : Bs1;
            const bool more = (s + 1 < nslices);
            if (more) fetch(kb + (s + 1) * BK);
            mma_slice<PREC, CF>(Ac, Bc, acc, wm, wn, lane);
            if (more) {
                sa.store(An, t);
                sb.store(Bn, t);
            }
            __syncthreads();
        }
    }

    // Epilogue.  Accumulator v of block (fm, fn) holds column lane%16 and row out_row(lane/16, v).
    const int li = lane & 15, lq = lane >> 4;
#pragma unroll
    for (int fm = 0; fm < CF::FBM; ++fm)
#pragma unroll
        for (int fn = 0; fn < CF::FBN; ++fn) {
            const int j = j0 + wn * CF::WN + fn * 16 + li;
            if (j >= n) continue;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = i0 + wm * CF::WM + fm * 16 + Prec<PREC>::out_row(lq, v);
                if (i >= m) continue;
                const AT val = acc[fm][fn][v];
                if (partial) {
                    partial[(int64_t)blockIdx.z * m * n + (int64_t)i + (int64_t)j * m] = val;
                } else {
                    double r = alpha * (double)val;
                    if (C0) r += beta0 * (double)C0[(int64_t)i + (int64_t)j * ldc];
                    if (beta != 0.0) r += beta * (double)C[(int64_t)i + (int64_t)j * ldc];
                    C[(int64_t)i + (int64_t)j * ldc] = (TS)r;
                }
            }
        }
}

// Reduction of split-K partials in split order, then the alpha/beta epilogue.
template <typename TS, typename AT>
__global__ void k_splitk_reduce(int m, int n, int splits, const AT *__restrict__ partial, double alpha, double beta0,
                                const TS *__restrict__ C0, double beta, TS *__restrict__ C, int64_t ldc) {
    int64_t total = (int64_t)m * n;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += (int64_t)gridDim.x * blockDim.x) {
        int64_t i = q % m, j = q / m;
        AT s = partial[q];
        for (int z = 1; z < splits; ++z) s += partial[(int64_t)z * total + q];
        double r = alpha * (double)s;
        if (C0) r += beta0 * (double)C0[i + j * ldc];
        if (beta != 0.0) r += beta * (double)C[i + j * ldc];
        C[i + j * ldc] = (TS)r;
    }
}

static bool aligned16(const void *p) { return ((uintptr_t)p & 15) == 0; }

// Output area (elements) of the lower-triangle tiles of an n x n product with bm x bm tiles: the
// flop count of a symmetric product is 2*k times this (the MFMA work actually issued).
static double sym_area(int n, int bm) {
    double area = 0.0;
    for (int a = 0; a * bm < n; ++a) area += (double)std::min(bm, n - a * bm) * std::min(n, (a + 1) * bm);
    return area;
}

// ---------------------------------------------------------------------------------------------
// FP64 products with direct-to-LDS staging (configuration 7): C := alpha*sum_p A_p*op(B_p) + beta0*C0
// + beta*C over NPAIR products back to back along K (the Gram: NPAIR = 1, op 't'; calc_G's pair:
// NPAIR = 2, op 'n'; single 'n','n' products: NPAIR = 1, op 'n').  A is m x k with the row index
// contiguous, so every k-row of a 128-wide tile is 1 KB, which one wave moves with one
// buffer_load_dwordx4 ... lds (16 bytes per lane, lane l landing at byte 16*l of the destination)
// straight into the k-major LDS tile.  The descriptors are rebased at each slice, so the lanes'
// byte offsets are small and fixed; the buffer form keeps the compiler's LDS waits counted (see
// k_bgemm_tn), and the fragments of k step ks + 4 are read before the MFMAs of step ks.
//   B 't' (n x k, row index contiguous): the same k-major rows.
//   B 'n' (k x n, k contiguous): a j-major tile, one 128-byte row of 16 k values per column j, 8
//   columns per wave load.  Logical 16-byte chunk c of column j is stored at physical chunk
//   c ^ ((j >> 1) & 7).  A fragment read takes one k of 16 columns, 8 bytes at the same place in
//   16 different 16-byte chunks, so it spans at most 8 of the 16-byte bank groups: two-way bank
//   conflicts that no placement of whole chunks avoids (measured: conflict cycles on 40 % of the
//   LDS-active cycles with this swizzle and with c ^ (j & 7); none on the 't' path).
// No staging registers and no LDS write instructions; two LDS stages, the loads of slice s+1 in
// flight while slice s feeds the MFMAs.  Tiles, waves and the MFMA chain per element are those of
// the staged kernel with Cfg5, so the results are bitwise the same.  Requires m, n % 128 == 0,
// k % 16 == 0, even leading dimensions and 16-byte aligned operands.
// ---------------------------------------------------------------------------------------------
typedef __attribute__((address_space(3))) void lds_void;
// s_waitcnt immediate (gfx9 encoding): vmcnt(vm), expcnt and lgkmcnt left at their maxima
constexpr int waitcnt_vmcnt(int vm) { return (vm & 0xF) | ((vm >> 4) << 14) | (7 << 4) | (0xF << 8); }
__device__ inline int dl_swz(int j) { return (j >> 1) & 7; }
// Raw (unstrided) buffer descriptor over `bytes` bytes from p, for the buffer loads into LDS of
// the direct-to-LDS products; word 3 is gfx9's for 32-bit data without swizzling.  Reads at or
// past `bytes` return zero instead of touching memory.
template <typename T>
__device__ inline __amdgpu_buffer_rsrc_t raw_rsrc(const T *p, int bytes = 0x7fffffff) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<T *>(p), (short)0, bytes, 0x00020000);
}

template <class CF, bool TRB, int NPAIR>
__global__ __launch_bounds__(CF::NT, CF::MINW) void k_dgemm_dl(int m, int n, int k, double alpha,
                                                               const double *__restrict__ A1,
                                                               const double *__restrict__ A2, int64_t lda,
                                                               const double *__restrict__ B1,
                                                               const double *__restrict__ B2, int64_t ldb,
                                                               double beta0, const double *__restrict__ C0,
                                                               double beta, double *__restrict__ C, int64_t ldc,
                                                               int sym, int tbase, int nsplit,
                                                               double *__restrict__ partial,
                                                               const int *__restrict__ mask) {
    static_assert(CF::BM == 128 && CF::BN == 128, "one 1 KB row per wave load");
    if (mask && !*mask) return;  // system masked off (grouped solver; whole workgroup, before any barrier)
    constexpr int BK = Prec<OMX_PREC_F64>::BK, NW = CF::WGM * CF::WGN, INS = 2 * BK, LPW = INS / NW;
    constexpr int STAGE = INS * 128;  // doubles: A rows (BK x 128), then B (BK x 128 or 128 x BK)
    static_assert(LPW * NW == INS, "loads per wave");
    __shared__ double smem[2 * STAGE];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = wave / CF::WGN, wn = wave % CF::WGN;
    // tile w of the launch's tile order and, with nsplit > 1, the part z of K it takes
    const int w = tbase + (int)blockIdx.x / nsplit, z = (int)blockIdx.x % nsplit;
    int tm, tn;
    if (sym)
        tile_of_lower(w, m / 128, tm, tn);
    else
        tile_of(w, m / 128, n / 128, tm, tn);
    const int i0 = tm * 128, j0 = tn * 128;
    d4v acc[CF::FBM][CF::FBN];
#pragma unroll
    for (int a = 0; a < CF::FBM; ++a)
#pragma unroll
        for (int b = 0; b < CF::FBN; ++b) acc[a][b] = d4v{0, 0, 0, 0};
    const int kpart = NPAIR * k / nsplit, s0 = z * (kpart / BK);  // this part's first slice
    // Each lane's byte offset from its slice's descriptor base (A + i0 + kl*lda; B 't': B + j0 +
    // kl*ldb; B 'n': B + j0*ldb + kl), the same for every slice.
    const int uwave = __builtin_amdgcn_readfirstlane(wave);
    int voff[LPW];
#pragma unroll
    for (int q = 0; q < LPW; ++q) {
        const int ins = uwave * LPW + q;  // < BK: row k of A; else the B part
        if (ins < BK) {
            voff[q] = 8 * (2 * lane + ins * (int)lda);
        } else if (TRB) {
            voff[q] = 8 * (2 * lane + (ins - BK) * (int)ldb);
        } else {
            const int j = (ins - BK) * 8 + (lane >> 3);  // column of the tile
            const int c = (lane & 7) ^ dl_swz(j);         // logical chunk this lane holds
            voff[q] = 8 * (j * (int)ldb + 2 * c);
        }
    }
    auto issue = [&](int s) {
        double *st = smem + (s & 1) * STAGE;
        int kl = (s0 + s) * BK;
        const double *A = A1, *B = B1;
        if (NPAIR == 2 && kl >= k) {
            A = A2;
            B = B2;
            kl -= k;
        }
        const __amdgpu_buffer_rsrc_t ra = raw_rsrc(A + i0 + (int64_t)kl * lda);
        const __amdgpu_buffer_rsrc_t rb = raw_rsrc(TRB ? B + j0 + (int64_t)kl * ldb : B + (int64_t)j0 * ldb + kl);
#pragma unroll
        for (int q = 0; q < LPW; ++q) {
            const int ins = uwave * LPW + q;
            const int vo = voff[q];  // (a local: see k_bgemm_tn)
            if (ins < BK)
                __builtin_amdgcn_raw_ptr_buffer_load_lds(ra, (lds_void *)(st + ins * 128), 16, vo, 0, 0, 0);
            else
                __builtin_amdgcn_raw_ptr_buffer_load_lds(rb, (lds_void *)(st + ins * 128), 16, vo, 0, 0, 0);
        }
    };
    const int li = lane & 15, lk = lane >> 4;
    const int ns = kpart / BK;
    issue(0);
    for (int s = 0; s < ns; ++s) {
        __builtin_amdgcn_s_waitcnt(waitcnt_vmcnt(0));  // this wave's rows of slice s have landed
        __syncthreads();                               // everyone's have; slice s-1 is no longer read
        if (s + 1 < ns) issue(s + 1);
        const double *As = smem + (s & 1) * STAGE, *Bs = As + BK * 128;
        // the fragments of k step ks + 4 are read before the MFMAs of step ks
        double a[2][CF::FBM], b[2][CF::FBN];
        auto frag = [&](int ks, double (&fa)[CF::FBM], double (&fb)[CF::FBN]) {
#pragma unroll
            for (int f = 0; f < CF::FBM; ++f) fa[f] = As[(ks + lk) * 128 + wm * CF::WM + f * 16 + li];
#pragma unroll
            for (int f = 0; f < CF::FBN; ++f) {
                const int j = wn * CF::WN + f * 16 + li, kk = ks + lk;
                fb[f] = TRB ? Bs[kk * 128 + j] : Bs[j * BK + ((kk >> 1) ^ dl_swz(j)) * 2 + (kk & 1)];
            }
        };
        frag(0, a[0], b[0]);
#pragma unroll
        for (int ks = 0; ks < BK; ks += 4) {
            const int cur = (ks / 4) & 1;
            if (ks + 4 < BK) frag(ks + 4, a[cur ^ 1], b[cur ^ 1]);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int fm = 0; fm < CF::FBM; ++fm)
#pragma unroll
                for (int fn = 0; fn < CF::FBN; ++fn)
                    acc[fm][fn] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[cur][fm], b[cur][fn], acc[fm][fn], 0, 0, 0);
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    const int lq = lane >> 4;
    if (partial) {  // a K part of a split tile: the raw sums, reduced in part order later
        double *pt = partial + (int64_t)blockIdx.x * (128 * 128);
#pragma unroll
        for (int fm = 0; fm < CF::FBM; ++fm)
#pragma unroll
            for (int fn = 0; fn < CF::FBN; ++fn)
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    pt[(wm * CF::WM + fm * 16 + Prec<OMX_PREC_F64>::out_row(lq, v)) +
                       (wn * CF::WN + fn * 16 + li) * 128] = acc[fm][fn][v];
        return;
    }
#pragma unroll
    for (int fm = 0; fm < CF::FBM; ++fm)
#pragma unroll
        for (int fn = 0; fn < CF::FBN; ++fn) {
            const int j = j0 + wn * CF::WN + fn * 16 + li;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = i0 + wm * CF::WM + fm * 16 + Prec<OMX_PREC_F64>::out_row(lq, v);
                double r = alpha * acc[fm][fn][v];
                if (C0) r += beta0 * C0[(int64_t)i + (int64_t)j * ldc];
                if (beta != 0.0) r += beta * C[(int64_t)i + (int64_t)j * ldc];
                C[(int64_t)i + (int64_t)j * ldc] = r;
            }
        }
}

// C tile := alpha * (sum of its nsplit K parts, in part order) + beta0*C0 + beta*C for the split
// tiles tbase.. of a symmetric product's lower-triangle order (blockIdx.x: tile, blockIdx.y: a
// 1024-element slice of it).
__global__ __launch_bounds__(256) void k_dl_split_reduce(int n, int tbase, int nsplit, const double *__restrict__ partial,
                                                         double alpha, double beta0, const double *__restrict__ C0,
                                                         double beta, double *__restrict__ C, int64_t ldc) {
    int tm, tn;
    tile_of_lower(tbase + (int)blockIdx.x, n / 128, tm, tn);
    const double *pt = partial + (int64_t)blockIdx.x * nsplit * (128 * 128);
    const int e0 = blockIdx.y * 1024;
    for (int e = e0 + threadIdx.x; e < e0 + 1024; e += 256) {
        double sum = 0.0;
        for (int zz = 0; zz < nsplit; ++zz) sum += pt[(int64_t)zz * (128 * 128) + e];
        const int64_t off = (int64_t)(tm * 128 + e % 128) + (int64_t)(tn * 128 + e / 128) * ldc;
        double r = alpha * sum;
        if (C0) r += beta0 * C0[off];
        if (beta != 0.0) r += beta * C[off];
        C[off] = r;
    }
}

// Launches configuration 7 when the product qualifies; returns OMX_ERR_UNSUPPORTED otherwise.
// A symmetric product's tiles that do not fill a whole round of resident workgroups (two per CU)
// are split along K over up to 16 workgroups each and reduced in part order, so the round that
// would hold only them costs a fraction of a tile.
template <bool TRB, int NPAIR>
static int run_dgemm_dl(omx_ctx ctx, int m, int n, int k, double alpha, const double *A1, const double *A2,
                        int64_t lda, const double *B1, const double *B2, int64_t ldb, double beta0, const double *C0,
                        double beta, double *C, int64_t ldc) {
    typedef Cfg5 CF;
    const bool sym = ctx->gemm_sym && m == n && NPAIR == 1;
    const int tiles = sym ? (m / 128) * (m / 128 + 1) / 2 : (m / 128) * (n / 128);
    if (m <= 0 || n <= 0 || k <= 0 || m % 128 || n % 128 || k % Prec<OMX_PREC_F64>::BK || lda % 2 || ldb % 2 ||
        !aligned16(A1) || !aligned16(A2) || !aligned16(B1) || !aligned16(B2) || tiles < 2 * ctx->num_cus ||
        lda > (1 << 20) || ldb > (1 << 20))  // 32-bit lane offsets from the slice descriptors
        return OMX_ERR_UNSUPPORTED;
    const int cls = NPAIR == 2 ? OMX_K_GEMM_PAIR : OMX_K_GEMM;
    hipEvent_t ev;
    prof_begin(ctx, cls, &ev);
    prof_work(ctx, cls, 2.0 * (sym ? sym_area(m, 128) : (double)m * n) * (double)NPAIR * k,
              sizeof(double) * ((double)NPAIR * ((double)m * k + (double)k * n) +
                                (double)m * n * (1 + (C0 ? 1 : 0) + (beta != 0.0 ? 1 : 0))));
    const int slots = 2 * ctx->num_cus;
    const int whole = sym ? tiles / slots * slots : tiles, rest = tiles - whole;
    int nsplit = 1;
    if (rest > 0 && !ctx->launch_mask) {  // masked (grouped, captured) launches use no workspace
        nsplit = std::max(1, std::min(16, slots / rest));
        while (nsplit > 1 && (k / Prec<OMX_PREC_F64>::BK) % nsplit) --nsplit;
    }
    if (rest > 0 && nsplit > 1) {
        int rc = ensure_ws(ctx, sizeof(double) * (size_t)rest * nsplit * 128 * 128);
        if (rc) return rc;
        hipLaunchKernelGGL((k_dgemm_dl<CF, TRB, NPAIR>), dim3(whole), dim3(CF::NT), 0, ctx->stream, m, n, k, alpha, A1,
                           A2, lda, B1, B2, ldb, beta0, C0, beta, C, ldc, (int)sym, 0, 1, (double *)nullptr,
                           (const int *)nullptr);
        hipLaunchKernelGGL((k_dgemm_dl<CF, TRB, NPAIR>), dim3(rest * nsplit), dim3(CF::NT), 0, ctx->stream, m, n, k,
                           alpha, A1, A2, lda, B1, B2, ldb, beta0, C0, beta, C, ldc, (int)sym, whole, nsplit,
                           (double *)ctx->ws, (const int *)nullptr);
        hipLaunchKernelGGL(k_dl_split_reduce, dim3(rest, 128 * 128 / 1024), dim3(256), 0, ctx->stream, m, whole, nsplit,
                           (const double *)ctx->ws, alpha, beta0, C0, beta, C, ldc);
    } else {
        hipLaunchKernelGGL((k_dgemm_dl<CF, TRB, NPAIR>), dim3(tiles), dim3(CF::NT), 0, ctx->stream, m, n, k, alpha, A1,
                           A2, lda, B1, B2, ldb, beta0, C0, beta, C, ldc, (int)sym, 0, 1, (double *)nullptr,
                           ctx->launch_mask);
    }
    prof_end(ctx, cls, ev);
    OMX_CHECK_HIP(hipGetLastError());
    return OMX_OK;
}

// ---------------------------------------------------------------------------------------------
// FP32 Gram C := alpha*A*B^T + beta0*C0 + beta*C with direct-to-LDS staging (configuration 7 on
// FP32 storage): A (m x k) and B (n x k) with the row index contiguous, 128x128 tiles, 8 waves of
// 32x64, v_mfma_f32_16x16x4f32, BK = 32.  A k-row of a tile is 512 B, so one
// buffer_load_dwordx4 ... lds per wave moves two consecutive k-rows (lanes 0-31 the first, 32-63
// the second), through descriptors rebased at every slice.  An MFMA step's fragment read takes 16
// consecutive i from each of four consecutive k-rows, which lie 512 B apart, on the same banks:
// 16-byte chunk c of k-row r is therefore stored at chunk c ^ (4 (r mod 4)), which places the four
// 64-byte runs in four different quarters of 256 bytes.  The fragments of k step ks + 4 are read
// before the MFMAs of step ks.  The MFMA chain per element (k in order) and the epilogue are the
// staged kernel's, so the results are bitwise the same.  Requires m, n % 128 == 0, k % 32 == 0,
// leading dimensions that are multiples of 4 and 16-byte aligned operands.
// ---------------------------------------------------------------------------------------------
__device__ inline int sl_swz(int r) { return 4 * (r & 3); }

template <class CF>
__global__ __launch_bounds__(CF::NT, CF::MINW) void k_sgemm_dl(int m, int n, int k, double alpha,
                                                               const float *__restrict__ A, int64_t lda,
                                                               const float *__restrict__ B, int64_t ldb, double beta0,
                                                               const float *__restrict__ C0, double beta,
                                                               float *__restrict__ C, int64_t ldc,
                                                               const int *__restrict__ mask) {
    static_assert(CF::BM == 128 && CF::BN == 128, "two 512-byte rows per wave load");
    if (mask && !*mask) return;  // system masked off (whole workgroup, before any barrier)
    constexpr int BK = Prec<OMX_PREC_F32>::BK, NW = CF::WGM * CF::WGN, INS = BK, LPW = INS / NW;
    constexpr int STAGE = 2 * BK * 128;  // floats: A rows (BK x 128), then B rows
    static_assert(LPW * NW == INS, "loads per wave");
    __shared__ float smem[2 * STAGE];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = wave / CF::WGN, wn = wave % CF::WGN;
    int tm, tn;
    tile_of((int)blockIdx.x, m / 128, n / 128, tm, tn);
    const int i0 = tm * 128, j0 = tn * 128;
    f4v acc[CF::FBM][CF::FBN];
#pragma unroll
    for (int a = 0; a < CF::FBM; ++a)
#pragma unroll
        for (int b = 0; b < CF::FBN; ++b) acc[a][b] = f4v{0, 0, 0, 0};
    // wave load `ins` < BK/2 fills A rows 2 ins, 2 ins + 1; the others B rows; lane l the physical
    // chunk l % 32 of row 2 ins + l / 32 (of its operand)
    const int uwave = __builtin_amdgcn_readfirstlane(wave);
    int voff[LPW];
#pragma unroll
    for (int q = 0; q < LPW; ++q) {
        const int ins = uwave * LPW + q;
        const int r = 2 * (ins % (BK / 2)) + (lane >> 5);
        const int c = (lane & 31) ^ sl_swz(r);  // logical chunk: i = 4c .. 4c+3
        voff[q] = 4 * (4 * c + r * (int)(ins < BK / 2 ? lda : ldb));
    }
    auto issue = [&](int s) {
        float *st = smem + (s & 1) * STAGE;
        const int kl = s * BK;
        const __amdgpu_buffer_rsrc_t ra = raw_rsrc(A + i0 + (int64_t)kl * lda);
        const __amdgpu_buffer_rsrc_t rb = raw_rsrc(B + j0 + (int64_t)kl * ldb);
#pragma unroll
        for (int q = 0; q < LPW; ++q) {
            const int ins = uwave * LPW + q;
            const int vo = voff[q];  // (a local: see k_bgemm_tn)
            if (ins < BK / 2)
                __builtin_amdgcn_raw_ptr_buffer_load_lds(ra, (lds_void *)(st + ins * 256), 16, vo, 0, 0, 0);
            else
                __builtin_amdgcn_raw_ptr_buffer_load_lds(rb, (lds_void *)(st + ins * 256), 16, vo, 0, 0, 0);
        }
    };
    const int li = lane & 15, lk = lane >> 4;
    const int ns = k / BK;
    issue(0);
    for (int s = 0; s < ns; ++s) {
        __builtin_amdgcn_s_waitcnt(waitcnt_vmcnt(0));  // this wave's rows of slice s have landed
        __syncthreads();                               // everyone's have; slice s-1 is no longer read
        if (s + 1 < ns) issue(s + 1);
        const float *As = smem + (s & 1) * STAGE, *Bs = As + BK * 128;
        float a[2][CF::FBM], b[2][CF::FBN];
        auto frag = [&](int ks, float (&fa)[CF::FBM], float (&fb)[CF::FBN]) {
            const int r = ks + lk;
#pragma unroll
            for (int f = 0; f < CF::FBM; ++f) {
                const int i = wm * CF::WM + f * 16 + li;
                fa[f] = As[r * 128 + (((i >> 2) ^ sl_swz(r)) << 2) + (i & 3)];
            }
#pragma unroll
            for (int f = 0; f < CF::FBN; ++f) {
                const int j = wn * CF::WN + f * 16 + li;
                fb[f] = Bs[r * 128 + (((j >> 2) ^ sl_swz(r)) << 2) + (j & 3)];
            }
        };
        frag(0, a[0], b[0]);
#pragma unroll
        for (int ks = 0; ks < BK; ks += 4) {
            const int cur = (ks / 4) & 1;
            if (ks + 4 < BK) frag(ks + 4, a[cur ^ 1], b[cur ^ 1]);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int fm = 0; fm < CF::FBM; ++fm)
#pragma unroll
                for (int fn = 0; fn < CF::FBN; ++fn)
                    acc[fm][fn] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[cur][fm], b[cur][fn], acc[fm][fn], 0, 0, 0);
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    const int lq = lane >> 4;
#pragma unroll
    for (int fm = 0; fm < CF::FBM; ++fm)
#pragma unroll
        for (int fn = 0; fn < CF::FBN; ++fn) {
            const int j = j0 + wn * CF::WN + fn * 16 + li;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = i0 + wm * CF::WM + fm * 16 + Prec<OMX_PREC_F32>::out_row(lq, v);
                const float val = acc[fm][fn][v];
                double r = alpha * (double)val;
                if (C0) r += beta0 * (double)C0[(int64_t)i + (int64_t)j * ldc];
                if (beta != 0.0) r += beta * (double)C[(int64_t)i + (int64_t)j * ldc];
                C[(int64_t)i + (int64_t)j * ldc] = (float)r;
            }
        }
}

// Launches the FP32 direct-to-LDS Gram when the product qualifies (full products whose tiles fill
// the chip; symmetric and small products take the staged kernel); OMX_ERR_UNSUPPORTED otherwise.
static int run_sgemm_dl(omx_ctx ctx, int m, int n, int k, double alpha, const float *A, int64_t lda, const float *B,
                        int64_t ldb, double beta0, const float *C0, double beta, float *C, int64_t ldc) {
    typedef Cfg5 CF;
    const int tiles = (m / 128) * (n / 128);
    if (m <= 0 || n <= 0 || k <= 0 || m % 128 || n % 128 || k % Prec<OMX_PREC_F32>::BK || lda % 4 || ldb % 4 ||
        !aligned16(A) || !aligned16(B) || tiles < 2 * ctx->num_cus || ctx->gemm_sym || lda > (1 << 20) ||
        ldb > (1 << 20))  // 32-bit lane offsets from the slice descriptors
        return OMX_ERR_UNSUPPORTED;
    hipEvent_t ev;
    prof_begin(ctx, OMX_K_GEMM, &ev);
    prof_work(ctx, OMX_K_GEMM, 2.0 * (double)m * n * k,
              sizeof(float) * ((double)m * k + (double)k * n + (double)m * n * (1 + (C0 ? 1 : 0) + (beta != 0.0 ? 1 : 0))));
    hipLaunchKernelGGL((k_sgemm_dl<CF>), dim3(tiles), dim3(CF::NT), 0, ctx->stream, m, n, k, alpha, A, lda, B, ldb,
                       beta0, C0, beta, C, ldc, ctx->launch_mask);
    prof_end(ctx, OMX_K_GEMM, ev);
    OMX_CHECK_HIP(hipGetLastError());
    return OMX_OK;
}

template <typename TS, int PREC, bool TRA, bool TRB, int NPAIR, class CF>
static int run_cfg(omx_ctx ctx, int m, int n, int k, double alpha, const TS *A1, const TS *A2, int64_t lda,
                   const TS *B1, const TS *B2, int64_t ldb, double beta0, const TS *C0, double beta, TS *C,
                   int64_t ldc) {
    typedef typename Prec<PREC>::acc_t AT;
    constexpr int BK = Prec<PREC>::BK;
    constexpr int VEC = 16 / sizeof(TS);
    const int tiles = ((m + CF::BM - 1) / CF::BM) * ((n + CF::BN - 1) / CF::BN);
    const int ktot = NPAIR * k;
    // Split K when the tile count cannot fill the chip (aim for >= 2 workgroups per CU).
    int splits = 1;
    const int target = 2 * ctx->num_cus;
    if (tiles < target && !ctx->launch_mask) {
        int max_splits = std::max(1, ktot / (4 * BK));
        splits = std::min(max_splits, (target + tiles - 1) / tiles);
        splits = std::max(1, std::min(splits, 64));
    }
    int kps = (ktot + splits - 1) / splits;
    kps = (kps + BK - 1) / BK * BK;  // slices never straddle a split
    splits = (ktot + kps - 1) / kps;
    AT *partial = nullptr;
    if (splits > 1) {
        size_t bytes = sizeof(AT) * (size_t)splits * m * n;
        int rc = ensure_ws(ctx, bytes);
        if (rc) return rc;
        partial = (AT *)ctx->ws;
    }
    const int vec_a = (lda % VEC == 0) && aligned16(A1) && aligned16(A2);
    const int vec_b = (ldb % VEC == 0) && aligned16(B1) && aligned16(B2);
    // products known to be symmetric (omx_gemm_sym): lower-triangle tiles only, when one split
    // covers K and the tiles are square
    const int tiles_m = (m + CF::BM - 1) / CF::BM;
    const bool sym = ctx->gemm_sym && m == n && CF::BM == CF::BN && splits == 1;
    const int grid = sym ? tiles_m * (tiles_m + 1) / 2 : tiles;
    hipEvent_t ev;
    prof_begin(ctx, NPAIR == 2 ? OMX_K_GEMM_PAIR : OMX_K_GEMM, &ev);
    prof_work(ctx, NPAIR == 2 ? OMX_K_GEMM_PAIR : OMX_K_GEMM, 2.0 * (sym ? sym_area(m, CF::BM) : (double)m * n) * (double)ktot,
              sizeof(TS) * ((double)NPAIR * ((double)m * k + (double)k * n) +
                            (double)m * n * (1 + (C0 ? 1 : 0) + (beta != 0.0 ? 1 : 0))));
    hipLaunchKernelGGL((k_gemm<TS, PREC, TRA, TRB, NPAIR, CF>), dim3(grid, 1, splits), dim3(CF::NT), 0, ctx->stream,
                       m, n, k, alpha, A1, A2, lda, B1, B2, ldb, beta0, C0, beta, C, ldc, kps, vec_a, vec_b, (int)sym,
                       partial, ctx->launch_mask);
    if (splits > 1) {
        int64_t total = (int64_t)m * n;
        int grid = (int)std::min<int64_t>((total + 255) / 256, 8192);
        hipLaunchKernelGGL((k_splitk_reduce<TS, AT>), dim3(grid), dim3(256), 0, ctx->stream, m, n, splits, partial,
                           alpha, beta0, C0, beta, C, ldc);
    }
    prof_end(ctx, NPAIR == 2 ? OMX_K_GEMM_PAIR : OMX_K_GEMM, ev);
    OMX_CHECK_HIP(hipGetLastError());
    return OMX_OK;
}

// Tile configurations that apply to a product: FP64 products with A not transposed ('n','t' Grams,
// 'n','n' products and the calc_G pair) carry 0-3 and 5-7 (7 on FP64 storage); the FP32 and BF16
// ones on FP32 storage 0 and 4, and the FP32 'n','t' Gram also 7; every other product uses 0.
template <typename TS, int PREC, bool TRA, bool TRB, int NPAIR>
struct CfgSet {
    static constexpr bool f64 = (PREC == OMX_PREC_F64) && !TRA;
    static constexpr bool dl = f64 && std::is_same<TS, double>::value;
    static constexpr bool lowp = (PREC != OMX_PREC_F64) && std::is_same<TS, float>::value && !TRA && (TRB || NPAIR == 2);
    static constexpr bool sdl = lowp && PREC == OMX_PREC_F32 && TRB && NPAIR == 1;
    static std::vector<int> ids() {
        if (dl) return {7, 0, 1, 2, 3, 5, 6};
        if (f64) return {0, 1, 2, 3, 5, 6};
        if (sdl) return {7, 0, 4};
        if (lowp) return {0, 4};
        return {0};
    }
};

template <typename TS, int PREC, bool TRA, bool TRB, int NPAIR>
static int run_cfg_id(int id, omx_ctx ctx, int m, int n, int k, double alpha, const TS *A1, const TS *A2,
                      int64_t lda, const TS *B1, const TS *B2, int64_t ldb, double beta0, const TS *C0, double beta,
                      TS *C, int64_t ldc) {
    typedef CfgSet<TS, PREC, TRA, TRB, NPAIR> CS;
    if constexpr (CS::dl) {
        if (id == 7) {
            int rc = run_dgemm_dl<TRB, NPAIR>(ctx, m, n, k, alpha, A1, A2, lda, B1, B2, ldb, beta0, C0, beta, C, ldc);
            if (rc != OMX_ERR_UNSUPPORTED) return rc;
            id = 5;  // the staged kernel with the same tiles covers every other shape
        }
    }
    if constexpr (CS::f64) {
        switch (id) {
            case 1:
                return run_cfg<TS, PREC, TRA, TRB, NPAIR, Cfg1>(ctx, m, n, k, alpha, A1, A2, lda, B1, B2, ldb, beta0,
                                                                C0, beta, C, ldc);
            case 2:
                return run_cfg<TS, PREC, TRA, TRB, NPAIR, Cfg2>(ctx, m, n, k, alpha, A1, A2, lda, B1, B2, ldb, beta0,
                                                                C0, beta, C, ldc);
            case 3:
                return run_cfg<TS, PREC, TRA, TRB, NPAIR, Cfg3>(ctx, m, n, k, alpha, A1, A2, lda, B1, B2, ldb, beta0,
                                                                C0, beta, C, ldc);
            case 5:
                return run_cfg<TS, PREC, TRA, TRB, NPAIR, Cfg5>(ctx, m, n, k, alpha, A1, A2, lda, B1, B2, ldb, beta0,
                                                                C0, beta, C, ldc);
            case 6:
                return run_cfg<TS, PREC, TRA, TRB, NPAIR, Cfg6>(ctx, m, n, k, alpha, A1, A2, lda, B1, B2, ldb, beta0,
                                                                C0, beta, C, ldc);
            default:
                break;
        }
    }
    if constexpr (CS::sdl) {
        if (id == 7) {
            int rc = run_sgemm_dl(ctx, m, n, k, alpha, A1, lda, B1, ldb, beta0, C0, beta, C, ldc);
            if (rc != OMX_ERR_UNSUPPORTED) return rc;
            id = 0;  // the staged kernel covers every other shape
        }
    }
    if constexpr (CS::lowp) {
        if (id == 4)
            return run_cfg<TS, PREC, TRA, TRB, NPAIR, Cfg4>(ctx, m, n, k, alpha, A1, A2, lda, B1, B2, ldb, beta0, C0,
                                                            beta, C, ldc);
    }
    return run_cfg<TS, PREC, TRA, TRB, NPAIR, Cfg0>(ctx, m, n, k, alpha, A1, A2, lda, B1, B2, ldb, beta0, C0, beta,
                                                    C, ldc);
}

// Autotuner: the first time a product shape is seen with tuning on, every applicable tile
// configuration is timed (one warm-up and one timed launch each, with HIP events on the library
// stream, into a scratch result so the caller's C is untouched) and the fastest is remembered for
// that shape on this context.  Tuning synchronises and allocates, so it must not run under graph
// capture.
template <typename TS, int PREC, bool TRA, bool TRB, int NPAIR>
static int tuned_cfg(omx_ctx ctx, int m, int n, int k, const TS *A1, const TS *A2, int64_t lda, const TS *B1,
                     const TS *B2, int64_t ldb, int &out) {
    typedef CfgSet<TS, PREC, TRA, TRB, NPAIR> CS;
    char key[128];
    snprintf(key, sizeof key, "gemm prec=%d es=%d op=%c%c pair=%d m=%d n=%d k=%d", PREC, (int)sizeof(TS),
             TRA ? 't' : 'n', TRB ? 't' : 'n', NPAIR, m, n, k);
    auto it = ctx->tune_gemm.find(key);
    if (it != ctx->tune_gemm.end()) {
        out = it->second.first;
        return OMX_OK;
    }
    const std::vector<int> ids = CS::ids();
    if (ids.size() == 1) {
        ctx->tune_gemm[key] = {ids[0], 0.0};
        out = ids[0];
        return OMX_OK;
    }
    TS *tmp = nullptr;
    OMX_CHECK_HIP(hipMalloc((void **)&tmp, sizeof(TS) * (size_t)std::max(m, 1) * std::max(n, 1)));
    hipEvent_t e0, e1;
    OMX_CHECK_HIP(hipEventCreate(&e0));
    OMX_CHECK_HIP(hipEventCreate(&e1));
    const bool prof = ctx->prof_on;
    ctx->prof_on = false;
    int best = ids[0], rc = OMX_OK;
    float best_ms = 0.0f;
    for (int id : ids) {
        float ms = 0.0f;
        for (int rep = 0; rep < 2 && rc == OMX_OK; ++rep) {
            (void)hipEventRecord(e0, ctx->stream);
            rc = run_cfg_id<TS, PREC, TRA, TRB, NPAIR>(id, ctx, m, n, k, 1.0, A1, A2, lda, B1, B2, ldb, 0.0, nullptr,
                                                       0.0, tmp, std::max(m, 1));
            (void)hipEventRecord(e1, ctx->stream);
            (void)hipEventSynchronize(e1);
            (void)hipEventElapsedTime(&ms, e0, e1);
        }
        if (rc) break;
        if (id == ids[0] || ms < best_ms) {
            best = id;
            best_ms = ms;
        }
    }
    ctx->prof_on = prof;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(tmp);
    if (rc) return rc;
    ctx->tune_gemm[key] = {best, (double)best_ms};
    out = best;
    return OMX_OK;
}

template <typename TS, int PREC, bool TRA, bool TRB, int NPAIR>
static int run_gemm(omx_ctx ctx, int m, int n, int k, double alpha, const TS *A1, const TS *A2, int64_t lda,
                    const TS *B1, const TS *B2, int64_t ldb, double beta0, const TS *C0, double beta, TS *C,
                    int64_t ldc) {
    // default: 128x128 with 8 waves and two workgroups per CU for the FP64 Gram and gradient products
    // (3-4 % ahead of the 4-wave tile at n=16384, bitwise the same result), with direct-to-LDS
    // staging where it applies (bitwise the same); the FP32 Gram's direct-to-LDS kernel where it
    // applies; 0 otherwise
    typedef CfgSet<TS, PREC, TRA, TRB, NPAIR> CS;
    int id = (CS::dl || CS::sdl) ? 7 : CS::f64 ? 5 : 0;
    if (ctx->gemm_cfg >= 0) {
        id = ctx->gemm_cfg;  // forced (omx_gemm_set_tile); a configuration that does not apply falls back to 0
    } else if (ctx->autotune) {
        int rc = tuned_cfg<TS, PREC, TRA, TRB, NPAIR>(ctx, m, n, k, A1, A2, lda, B1, B2, ldb, id);
        if (rc) return rc;
    }
    return run_cfg_id<TS, PREC, TRA, TRB, NPAIR>(id, ctx, m, n, k, alpha, A1, A2, lda, B1, B2, ldb, beta0, C0, beta,
                                                 C, ldc);
}

// ---------------------------------------------------------------------------------------------
// BF16 products of FP32 matrices through k-contiguous BF16 copies.
//
// The operands are first rounded to BF16 (round to nearest even, the same rounding as the
// staging conversion of k_gemm) into k-contiguous panels: op(A) as At[kk + i*K] and op(B) as
// Bt[kk + j*K].  The GEMM then moves half the bytes of FP32 operands, loads 16-byte vectors of 8
// consecutive k straight into the i-major LDS tiles, and skips the per-tile conversion.  With
// split = true each operand is also written as a second BF16 term, lo = bf16(a - hi), and the
// product runs over a K three times longer, [Ah Ah Al] x [Bh; Bl; Bh]: hi*hi + hi*lo + lo*hi,
// which keeps about 16 significant bits of each operand (the BF16x3 product).
// ---------------------------------------------------------------------------------------------

constexpr int BG_T = 64;  // conversion tile (rows x cols)

// Destinations of one operand's conversion: BgLay<L>::term[q] (1 = hi = bf16(a), 2 = mid =
// bf16(a - hi), 3 = lo = bf16(a - hi - mid), each difference taken exactly in the source
// precision) goes to p[q], q < BgLay<L>::n.  For FP32 sources hi + mid + lo is the value exactly;
// hi + mid keeps about 16 significant bits.  Layouts: 0 = BF16; 1, 2 = BF16x3 A and B; 3, 4 =
// BF16x6 A and B (the k blocks of the panel product, in order).
struct BgDst {
    __bf16 *p[6];
};
template <int L>
struct BgLay;
template <>
struct BgLay<0> {
    static constexpr int n = 1;
    __device__ static constexpr int term(int) { return 1; }
};
template <>
struct BgLay<1> {
    static constexpr int n = 3;
    __device__ static constexpr int term(int q) { return q == 2 ? 2 : 1; }  // hi hi mid
};
template <>
struct BgLay<2> {
    static constexpr int n = 3;
    __device__ static constexpr int term(int q) { return q == 1 ? 2 : 1; }  // hi mid hi
};
template <>
struct BgLay<3> {
    static constexpr int n = 6;
    __device__ static constexpr int term(int q) { return q == 1 ? 3 : (q == 2 || q == 4) ? 2 : 1; }  // 1 3 2 1 2 1
};
template <>
struct BgLay<4> {
    static constexpr int n = 6;
    __device__ static constexpr int term(int q) { return q == 0 ? 3 : (q == 2 || q == 3) ? 2 : 1; }  // 3 1 2 2 1 1
};

// dst terms of src (rows x cols, ld lds, FP32 or FP64), each rounded once from the stored value:
// TRANS: element (r, c) goes to dst[c + r*ldd], else to dst[r + c*ldd].  One 64 x 64 tile per
// workgroup: 16-byte loads along r, and 16-byte stores of 8 BF16 along the destination's
// contiguous index (through an LDS transpose for TRANS).  Edges and unaligned operands take
// element-wise paths.
template <bool TRANS, typename TS, int L>
__global__ __launch_bounds__(256) void k_to_bf16(int rows, int cols, const TS *__restrict__ src, int64_t lds,
                                                 BgDst dst, int64_t ldd) {
    constexpr int ND = BgLay<L>::n;
    constexpr int VEC = 16 / sizeof(TS);  // elements per 16-byte load
    const int t = threadIdx.x;
    const int r0 = blockIdx.x * BG_T, c0 = blockIdx.y * BG_T;
    const bool vld = (lds % VEC) == 0 && ((uintptr_t)src & 15) == 0;
    bool vst = (ldd % 8) == 0;
#pragma unroll
    for (int q = 0; q < ND; ++q) vst = vst && ((uintptr_t)dst.p[q] & 15) == 0;
    auto load = [&](int gr, int gc, int count, TS *x) {  // x[0..count) = src(gr.., gc), zero outside
        const TS *p = src + (int64_t)gr + (int64_t)gc * lds;
        if (vld && gc < cols && gr + count <= rows) {
#pragma unroll
            for (int e = 0; e < count; e += VEC) {
                typedef TS vec_t __attribute__((ext_vector_type(VEC)));
                const vec_t v = *reinterpret_cast<const vec_t *>(p + e);
#pragma unroll
                for (int q = 0; q < VEC; ++q) x[e + q] = v[q];
            }
        } else {
            for (int e = 0; e < count; ++e) x[e] = (gc < cols && gr + e < rows) ? p[e] : (TS)0;
        }
    };
    auto store8 = [&](const TS *x, int64_t off, int avail) {  // the terms of 8 values to every destination
        bf8v t1, t2, t3;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const __bf16 h = (__bf16)x[e];
            const TS r1 = x[e] - (TS)h;
            const __bf16 md = (__bf16)r1;
            t1[e] = h;
            t2[e] = md;
            t3[e] = (__bf16)(r1 - (TS)md);
        }
#pragma unroll
        for (int q = 0; q < ND; ++q) {
            const int tq = BgLay<L>::term(q);
            const bf8v v = tq == 1 ? t1 : (tq == 2 ? t2 : t3);
            if (vst && avail >= 8) {
                *reinterpret_cast<bf8v *>(dst.p[q] + off) = v;
            } else {
                for (int e = 0; e < 8 && e < avail; ++e) dst.p[q][off + e] = v[e];
            }
        }
    };
    if constexpr (TRANS) {
        __shared__ TS tile[BG_T][BG_T + 1];  // [c][r]
        constexpr int VPC = BG_T / VEC;      // load vectors per tile column
        for (int v = t; v < BG_T * VPC; v += 256) {
            const int c = v / VPC, rv = (v % VPC) * VEC;
            TS x[VEC];
            load(r0 + rv, c0 + c, VEC, x);
#pragma unroll
            for (int e = 0; e < VEC; ++e) tile[c][rv + e] = x[e];
        }
        __syncthreads();
        for (int u = t; u < BG_T * (BG_T / 8); u += 256) {
            const int r = u / (BG_T / 8), c8 = (u % (BG_T / 8)) * 8;
            const int gr = r0 + r, gc = c0 + c8;
            if (gr >= rows || gc >= cols) continue;
            TS x[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) x[e] = tile[c8 + e][r];
            store8(x, (int64_t)gc + (int64_t)gr * ldd, cols - gc);
        }
    } else {
        for (int u = t; u < BG_T * (BG_T / 8); u += 256) {
            const int c = u / (BG_T / 8), r8 = (u % (BG_T / 8)) * 8;
            const int gr = r0 + r8, gc = c0 + c;
            if (gr >= rows || gc >= cols) continue;
            TS x[8];
            load(gr, gc, 8, x);
            store8(x, (int64_t)gr + (int64_t)gc * ldd, rows - gr);
        }
    }
}

// LDS tile of the BF16 panel product: i-major rows of BK = 64 BF16 (128 bytes, 8 chunks of 16 B),
// no padding.  Logical chunk c of row r is stored at physical chunk c ^ bg_swz(r).  A fragment
// read takes one chunk of 16 consecutive rows; with this swizzle those 16 reads cover 16
// distinct 16-byte bank groups, so the read is free of bank conflicts.
__device__ inline int bg_swz(int r) { return (r & 7) ^ ((r >> 3) & 1); }


// One BK = 64 slice of the BF16 panel product, software-pipelined: the wave's MFMAs run in groups
// (one A fragment x the FBN B fragments of a 32-wide k step), and the LDS reads for group g + LEAD
// are issued before the MFMAs of group g, so that LEAD groups of reads are in flight behind the
// matrix core (the waits become lgkmcnt(n > 0) instead of a full drain after every read).  The
// scheduling fences keep the compiler from regrouping the reads.
template <class CF, int LEAD>
__device__ inline void mma_slice_bf16_pipe(const __bf16 *As, const __bf16 *Bs, f4v (&acc)[CF::FBM][CF::FBN], int wm,
                                           int wn, int lane) {
    constexpr int BK = Prec<OMX_PREC_BF16>::BK, NKS = BK / 32, NG = NKS * CF::FBM;
    const int li = lane & 15, lk = lane >> 4;
    bf8v fa[NKS][CF::FBM], fb[NKS][CF::FBN];
#pragma unroll
    for (int g = -LEAD; g < NG; ++g) {
        const int lg = g + LEAD;
        if (lg >= 0 && lg < NG) {
            const int ks = lg / CF::FBM, fm = lg % CF::FBM;
            const int c = ks * 4 + lk;  // logical 16-byte chunk: k = 8c .. 8c+7
            if (fm == 0) {
#pragma unroll
                for (int f = 0; f < CF::FBN; ++f) {
                    const int r = wn * CF::WN + f * 16 + li;
                    fb[ks][f] = *reinterpret_cast<const bf8v *>(&Bs[r * BK + (c ^ bg_swz(r)) * 8]);
                }
            }
            const int r = wm * CF::WM + fm * 16 + li;
            fa[ks][fm] = *reinterpret_cast<const bf8v *>(&As[r * BK + (c ^ bg_swz(r)) * 8]);
        }
        __builtin_amdgcn_sched_barrier(0);
        if (g >= 0) {
            const int ks = g / CF::FBM, fm = g % CF::FBM;
#pragma unroll
            for (int fn = 0; fn < CF::FBN; ++fn)
                acc[fm][fn] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[ks][fm], fb[ks][fn], acc[fm][fn], 0, 0, 0);
        }
        __builtin_amdgcn_sched_barrier(0);
    }
}

// C := alpha * sum_kk At[kk + i*lda] * Bt[kk + j*ldb] + beta0*C0 + beta*C over kk < K (K % 64 == 0,
// lda, ldb multiples of 8, 16-byte aligned panels), FP32 accumulation; C, C0 of type TO (the
// storage type, FP32 or FP64).
//
// Operand rows go straight from the panels into LDS: one buffer_load_dwordx4 ... lds per wave moves
// 8 rows of a slice (1 KB, lane l landing at byte 16*l), each lane fetching the logical chunk that
// belongs at its physical position under bg_swz.  No staging registers and no LDS write
// instructions; two LDS stages, the loads of slice s+1 in flight while slice s feeds the MFMAs.
// The buffer form rather than global_load_lds: while a FLAT-encoded load into LDS is in flight the
// compiler's wait insertion cannot count LDS reads in order and drains every one of them
// (lgkmcnt(0)) before each use, which serialises the fragment reads the slice pipeline issues ahead;
// with the buffer form the waits count only the reads a group needs.  Requires BM * K * 2 bytes
// below 2^31 (byte offsets inside the tile's rows are 32-bit).
// Rows past m or n are fetched from the last valid row and their results are not stored.
template <class CF, typename TO>
__global__ __launch_bounds__(CF::NT, CF::MINW) void k_bgemm_tn(int m, int n, int K, double alpha,
                                                     const __bf16 *__restrict__ At, int64_t lda,
                                                     const __bf16 *__restrict__ Bt, int64_t ldb, double beta0,
                                                     const TO *__restrict__ C0, double beta, TO *__restrict__ C,
                                                     int64_t ldc, int sym) {
    constexpr int BK = Prec<OMX_PREC_BF16>::BK;
    constexpr int NW = CF::WGM * CF::WGN;
    constexpr int INS = (CF::BM + CF::BN) / 8;  // wave loads per slice, 8 rows each
    constexpr int LPW = INS / NW;
    constexpr int STAGE = (CF::BM + CF::BN) * BK;  // BF16 per LDS stage
    static_assert(LPW * NW == INS && CF::BM % 8 == 0, "rows per wave");
    __shared__ __bf16 smem[2 * STAGE];

    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = wave / CF::WGN, wn = wave % CF::WGN;
    const int tiles_m = (m + CF::BM - 1) / CF::BM, tiles_n = (n + CF::BN - 1) / CF::BN;
    int tm, tn;
    if (sym)
        tile_of_lower(blockIdx.x, tiles_m, tm, tn);
    else
        tile_of(blockIdx.x, tiles_m, tiles_n, tm, tn);
    const int i0 = tm * CF::BM, j0 = tn * CF::BN;

    f4v acc[CF::FBM][CF::FBN];
#pragma unroll
    for (int a = 0; a < CF::FBM; ++a)
#pragma unroll
        for (int b = 0; b < CF::FBN; ++b) acc[a][b] = f4v{0, 0, 0, 0};

    // Buffer descriptors over the tile's A rows and B rows; each lane's byte offset of its chunk
    // in slice 0 is fixed for the whole K loop (slice s adds 2*s*BK).
    const __amdgpu_buffer_rsrc_t ra = raw_rsrc(At + (int64_t)i0 * lda, (int)(2 * ((int64_t)min(CF::BM, m - i0) * lda)));
    const __amdgpu_buffer_rsrc_t rb = raw_rsrc(Bt + (int64_t)j0 * ldb, (int)(2 * ((int64_t)min(CF::BN, n - j0) * ldb)));
    const int uwave = __builtin_amdgcn_readfirstlane(wave);
    int voff[LPW];
#pragma unroll
    for (int q = 0; q < LPW; ++q) {
        const int ins = uwave * LPW + q;       // A rows first, then B rows
        const int rr = ins * 8 + (lane >> 3);  // tile row this lane fills
        const int p = lane & 7;                // physical chunk it fills
        if (ins * 8 < CF::BM)
            voff[q] = 2 * (int)(min(rr, m - 1 - i0) * lda + (p ^ bg_swz(rr)) * 8);
        else
            voff[q] = 2 * (int)(min(rr - CF::BM, n - 1 - j0) * ldb + (p ^ bg_swz(rr - CF::BM)) * 8);
    }
    auto issue = [&](int s) {
        __bf16 *st = smem + (s & 1) * STAGE;
        const int kb = 2 * s * BK;
#pragma unroll
        for (int q = 0; q < LPW; ++q) {
            const int ins = uwave * LPW + q;
            const int vo = voff[q] + kb;
            // An if over two calls, and the offset in a local: with a conditional expression over the
            // two descriptors, or an array element in the argument, the host compilation pass drops
            // the kernel's launch stub without a diagnostic (build.sh checks for that).
            if (ins * 8 < CF::BM)
                __builtin_amdgcn_raw_ptr_buffer_load_lds(ra, (lds_void *)(st + ins * 8 * BK), 16, vo, 0, 0, 0);
            else
                __builtin_amdgcn_raw_ptr_buffer_load_lds(rb, (lds_void *)(st + ins * 8 * BK), 16, vo, 0, 0, 0);
        }
    };

    const int nslices = K / BK;
    if (nslices > 0) issue(0);
    for (int s = 0; s < nslices; ++s) {
        __builtin_amdgcn_s_waitcnt(waitcnt_vmcnt(0));  // this wave's rows of slice s have landed
        __syncthreads();                               // everyone's have; slice s-1 is no longer read
        if (s + 1 < nslices) issue(s + 1);
        const __bf16 *As = smem + (s & 1) * STAGE, *Bs = As + CF::BM * BK;
        mma_slice_bf16_pipe<CF, 2>(As, Bs, acc, wm, wn, lane);
    }

    const int li = lane & 15, lq = lane >> 4;
#pragma unroll
    for (int fm = 0; fm < CF::FBM; ++fm)
#pragma unroll
        for (int fn = 0; fn < CF::FBN; ++fn) {
            const int j = j0 + wn * CF::WN + fn * 16 + li;
            if (j >= n) continue;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = i0 + wm * CF::WM + fm * 16 + Prec<OMX_PREC_BF16>::out_row(lq, v);
                if (i >= m) continue;
                double r = alpha * (double)acc[fm][fn][v];
                if (C0) r += beta0 * (double)C0[(int64_t)i + (int64_t)j * ldc];
                if (beta != 0.0) r += beta * (double)C[(int64_t)i + (int64_t)j * ldc];
                C[(int64_t)i + (int64_t)j * ldc] = (TO)r;
            }
        }
}

// One operand term of a BF16 panel product: op(X) with X rows x cols (ld ldx); as a k-contiguous
// panel it is `extent` x K (extent = m for A, n for B).
struct BgTerm {
    const void *X;  // FP32 or FP64 (the product's storage type)
    int64_t ldx;
    bool trans;  // true: X is stored K x extent (already k-contiguous)
};

template <typename TS, int L>
static void bg_convert_l(omx_ctx ctx, const BgTerm &tm, int extent, int K, const BgDst &dst, int64_t ldd) {
    const TS *X = (const TS *)tm.X;
    if (tm.trans) {
        // X is K x extent with k contiguous: same layout, dst[kk + e*ldd]
        dim3 grid((K + BG_T - 1) / BG_T, (extent + BG_T - 1) / BG_T);
        hipLaunchKernelGGL((k_to_bf16<false, TS, L>), grid, dim3(256), 0, ctx->stream, K, extent, X, tm.ldx, dst, ldd);
    } else {
        // X is extent x K: transpose so that dst[kk + e*ldd] = X(e, kk)
        dim3 grid((extent + BG_T - 1) / BG_T, (K + BG_T - 1) / BG_T);
        hipLaunchKernelGGL((k_to_bf16<true, TS, L>), grid, dim3(256), 0, ctx->stream, extent, K, X, tm.ldx, dst, ldd);
    }
}

// Converts one operand into the panel blocks of layout `lay` (see BgLay).
template <typename TS>
static int bg_convert(omx_ctx ctx, const BgTerm &tm, int extent, int K, const BgDst &dst, int64_t ldd, int lay) {
    switch (lay) {
        case 0: bg_convert_l<TS, 0>(ctx, tm, extent, K, dst, ldd); break;
        case 1: bg_convert_l<TS, 1>(ctx, tm, extent, K, dst, ldd); break;
        case 2: bg_convert_l<TS, 2>(ctx, tm, extent, K, dst, ldd); break;
        case 3: bg_convert_l<TS, 3>(ctx, tm, extent, K, dst, ldd); break;
        default: bg_convert_l<TS, 4>(ctx, tm, extent, K, dst, ldd); break;
    }
    OMX_CHECK_HIP(hipGetLastError());
    return OMX_OK;
}

// C := alpha * sum_t op(A_t) op(B_t) + beta0*C0 + beta*C for nterm products of inner size k on
// TS storage (FP32 or FP64) through k-contiguous BF16 panels in the workspace.  nsplit = 1: BF16
// operands.  nsplit = 2 (BF16x3): hi*hi + hi*mid + mid*hi, about 16 significant bits per operand.
// nsplit = 3 (BF16x6): the six term products of order >= 2^-16 of the operands' three-term splits,
// smallest first (hi*lo + lo*hi + mid*mid + hi*mid + mid*hi + hi*hi), FP32-grade products.
// Returns OMX_ERR_UNSUPPORTED (nothing launched) when the shape does not suit the panels.
template <typename TS>
static int bgemm_panels(omx_ctx ctx, int m, int n, int k, int nterm, const BgTerm *A, const BgTerm *B, int nsplit,
                        double alpha, double beta0, const TS *C0, double beta, TS *C, int64_t ldc) {
    constexpr int BK = Prec<OMX_PREC_BF16>::BK;
    if (ctx->launch_mask || k % BK != 0 || m < 256 || n < 256 || k < 4 * BK) return OMX_ERR_UNSUPPORTED;
    // k blocks per product term and the conversion layouts of A and B (BgLay)
    static const int nblk[3] = {1, 3, 6}, lay_a[3] = {0, 1, 3}, lay_b[3] = {0, 2, 4};
    OMX_REQUIRE(nsplit >= 1 && nsplit <= 3, "bf16 panels: split %d", nsplit);
    const int reps = nblk[nsplit - 1];
    const int64_t K = (int64_t)k * nterm * reps;
    if ((int64_t)Cfg4::BM * K * (int64_t)sizeof(__bf16) >= (1LL << 31)) return OMX_ERR_UNSUPPORTED;  // 32-bit offsets
    const size_t bytes_a = sizeof(__bf16) * (size_t)K * m, bytes_b = sizeof(__bf16) * (size_t)K * n;
    int rc = ensure_ws(ctx, bytes_a + bytes_b + 256);
    if (rc) return rc;
    __bf16 *At = (__bf16 *)ctx->ws;
    __bf16 *Bt = (__bf16 *)((char *)ctx->ws + ((bytes_a + 255) / 256) * 256);
    const bool sym = ctx->gemm_sym && m == n && nterm == 1;  // lower-triangle tiles of a symmetric product
    hipEvent_t ev;
    prof_begin(ctx, nterm == 2 ? OMX_K_GEMM_PAIR : OMX_K_GEMM, &ev);
    prof_work(ctx, nterm == 2 ? OMX_K_GEMM_PAIR : OMX_K_GEMM,
              2.0 * (sym ? sym_area(m, Cfg4::BM) : (double)m * n) * (double)k * nterm,
              sizeof(TS) * (nterm * ((double)m * k + (double)k * n) + (double)m * n * (1 + (C0 ? 1 : 0) + (beta != 0.0))));
    // Rows of At are m panels of length K (ld K), rows of Bt n of length K; product term tt takes
    // the K range [tt*reps*k, (tt+1)*reps*k), block b of it holding the A and B terms of pair b.
    for (int tt = 0; tt < nterm && !rc; ++tt) {
        const int64_t kb = (int64_t)tt * reps * k;
        BgDst da{}, db{};
        for (int b = 0; b < reps; ++b) {
            da.p[b] = At + kb + (int64_t)b * k;
            db.p[b] = Bt + kb + (int64_t)b * k;
        }
        rc = bg_convert<TS>(ctx, A[tt], m, k, da, K, lay_a[nsplit - 1]);
        if (!rc) rc = bg_convert<TS>(ctx, B[tt], n, k, db, K, lay_b[nsplit - 1]);
    }
    if (!rc) {
        typedef Cfg4 CF;
        const int tm = (m + CF::BM - 1) / CF::BM;
        const int tiles = sym ? tm * (tm + 1) / 2 : tm * ((n + CF::BN - 1) / CF::BN);
        hipLaunchKernelGGL((k_bgemm_tn<CF, TS>), dim3(tiles), dim3(CF::NT), 0, ctx->stream, m, n, (int)K, alpha, At, K,
                           Bt, K, beta0, C0, beta, C, ldc, (int)sym);
        rc = hipGetLastError() == hipSuccess ? OMX_OK : OMX_ERR_HIP;
        if (rc) set_error("bf16 panel GEMM launch failed");
    }
    prof_end(ctx, nterm == 2 ? OMX_K_GEMM_PAIR : OMX_K_GEMM, ev);
    return rc;
}

// BF16 panel split of a product precision: 1 term (BF16), 2 (BF16x3) or 3 (BF16x6).
static int split_of(int prec) { return prec == OMX_PREC_BF16X6 ? 3 : prec == OMX_PREC_BF16X3 ? 2 : 1; }

template <typename TS, int PREC>
static int dispatch_ops(omx_ctx ctx, bool trA, bool trB, int m, int n, int k, double alpha, const void *A,
                        int64_t lda, const void *B, int64_t ldb, double beta, void *C, int64_t ldc) {
    const TS *a = (const TS *)A, *b = (const TS *)B;
    TS *c = (TS *)C;
    if (!trA && !trB)
        return run_gemm<TS, PREC, false, false, 1>(ctx, m, n, k, alpha, a, a, lda, b, b, ldb, 0.0, nullptr, beta, c, ldc);
    if (!trA && trB)
        return run_gemm<TS, PREC, false, true, 1>(ctx, m, n, k, alpha, a, a, lda, b, b, ldb, 0.0, nullptr, beta, c, ldc);
    if (trA && !trB)
        return run_gemm<TS, PREC, true, false, 1>(ctx, m, n, k, alpha, a, a, lda, b, b, ldb, 0.0, nullptr, beta, c, ldc);
    return run_gemm<TS, PREC, true, true, 1>(ctx, m, n, k, alpha, a, a, lda, b, b, ldb, 0.0, nullptr, beta, c, ldc);
}

int launch_gemm(omx_ctx ctx, int prec, int dt, bool trA, bool trB, int m, int n, int k, double alpha, const void *A,
                int64_t lda, const void *B, int64_t ldb, double beta, void *C, int64_t ldc) {
    if (dt == OMX_F64) {
        if (prec == OMX_PREC_BF16 || prec == OMX_PREC_BF16X3 || prec == OMX_PREC_BF16X6) {
            const BgTerm a{A, lda, trA}, b{B, ldb, !trB};
            int rc = bgemm_panels<double>(ctx, m, n, k, 1, &a, &b, split_of(prec), alpha, 0.0, nullptr, beta,
                                          (double *)C, ldc);
            if (rc != OMX_ERR_UNSUPPORTED) return rc;
            if (prec != OMX_PREC_BF16) prec = OMX_PREC_F32;  // the panels do not cover this shape
        }
        if (prec == OMX_PREC_F64)
            return dispatch_ops<double, OMX_PREC_F64>(ctx, trA, trB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
        if (prec == OMX_PREC_F32)
            return dispatch_ops<double, OMX_PREC_F32>(ctx, trA, trB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
        if (prec == OMX_PREC_BF16)
            return dispatch_ops<double, OMX_PREC_BF16>(ctx, trA, trB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    } else if (dt == OMX_F32) {
        if (prec == OMX_PREC_BF16 || prec == OMX_PREC_BF16X3 || prec == OMX_PREC_BF16X6) {
            // k-contiguous BF16 panels (A 't' and B 'n' are already k-contiguous)
            const BgTerm a{A, lda, trA}, b{B, ldb, !trB};
            int rc = bgemm_panels<float>(ctx, m, n, k, 1, &a, &b, split_of(prec), alpha, 0.0, nullptr, beta,
                                         (float *)C, ldc);
            if (rc != OMX_ERR_UNSUPPORTED) return rc;
            if (prec != OMX_PREC_BF16) prec = OMX_PREC_F32;  // the panels do not cover this shape
        }
        if (prec == OMX_PREC_F32)
            return dispatch_ops<float, OMX_PREC_F32>(ctx, trA, trB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
        if (prec == OMX_PREC_BF16)
            return dispatch_ops<float, OMX_PREC_BF16>(ctx, trA, trB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    }
    set_error("omx_gemm: unsupported (dt=%d, prec=%d)", dt, prec);
    return OMX_ERR_UNSUPPORTED;
}

// C := alpha*(A1*B1 + A2*B2) + beta0*C0 + beta*C with A* m x k ('n'), B* k x n ('n').
int launch_gemm_pair(omx_ctx ctx, int prec, int dt, int m, int n, int k, double alpha, const void *A1,
                     const void *A2, int64_t lda, const void *B1, const void *B2, int64_t ldb, double beta0,
                     const void *C0, double beta, void *C, int64_t ldc) {
    if ((dt == OMX_F32 || dt == OMX_F64) &&
        (prec == OMX_PREC_BF16 || prec == OMX_PREC_BF16X3 || prec == OMX_PREC_BF16X6)) {
        const BgTerm a[2] = {{A1, lda, false}, {A2, lda, false}};
        const BgTerm b[2] = {{B1, ldb, true}, {B2, ldb, true}};
        const int split = split_of(prec);
        int rc = dt == OMX_F32 ? bgemm_panels<float>(ctx, m, n, k, 2, a, b, split, alpha, beta0, (const float *)C0,
                                                     beta, (float *)C, ldc)
                               : bgemm_panels<double>(ctx, m, n, k, 2, a, b, split, alpha, beta0, (const double *)C0,
                                                      beta, (double *)C, ldc);
        if (rc != OMX_ERR_UNSUPPORTED) return rc;
        if (prec != OMX_PREC_BF16) prec = OMX_PREC_F32;
    }
    const int bk = (prec == OMX_PREC_F64)   ? Prec<OMX_PREC_F64>::BK
                   : (prec == OMX_PREC_F32) ? Prec<OMX_PREC_F32>::BK
                                            : Prec<OMX_PREC_BF16>::BK;
    // The pair boundary must fall on a K-slice boundary; otherwise the caller runs separate passes.
    if (k % bk != 0) return OMX_ERR_UNSUPPORTED;
    if (dt == OMX_F64) {
        const double *a1 = (const double *)A1, *a2 = (const double *)A2, *b1 = (const double *)B1,
                     *b2 = (const double *)B2, *c0 = (const double *)C0;
        if (prec == OMX_PREC_F64)
            return run_gemm<double, OMX_PREC_F64, false, false, 2>(ctx, m, n, k, alpha, a1, a2, lda, b1, b2, ldb,
                                                                   beta0, c0, beta, (double *)C, ldc);
        if (prec == OMX_PREC_F32)
            return run_gemm<double, OMX_PREC_F32, false, false, 2>(ctx, m, n, k, alpha, a1, a2, lda, b1, b2, ldb,
                                                                   beta0, c0, beta, (double *)C, ldc);
        if (prec == OMX_PREC_BF16)
            return run_gemm<double, OMX_PREC_BF16, false, false, 2>(ctx, m, n, k, alpha, a1, a2, lda, b1, b2, ldb,
                                                                    beta0, c0, beta, (double *)C, ldc);
    } else if (dt == OMX_F32) {
        const float *a1 = (const float *)A1, *a2 = (const float *)A2, *b1 = (const float *)B1,
                    *b2 = (const float *)B2, *c0 = (const float *)C0;
        if (prec == OMX_PREC_F32)
            return run_gemm<float, OMX_PREC_F32, false, false, 2>(ctx, m, n, k, alpha, a1, a2, lda, b1, b2, ldb, beta0,
                                                                  c0, beta, (float *)C, ldc);
        if (prec == OMX_PREC_BF16)
            return run_gemm<float, OMX_PREC_BF16, false, false, 2>(ctx, m, n, k, alpha, a1, a2, lda, b1, b2, ldb,
                                                                   beta0, c0, beta, (float *)C, ldc);
    }
    set_error("calc_G: unsupported (dt=%d, prec=%d)", dt, prec);
    return OMX_ERR_UNSUPPORTED;
}

// C(i, j) := C(j, i) for i < j (n x n, ld ldc): copies the lower triangle onto the upper one.  One
// 64 x 64 tile pair per workgroup (the upper tile and the lower one it mirrors, the diagonal tiles
// on themselves), through an LDS transpose so that reads and writes run along columns.
template <typename TS>
__global__ __launch_bounds__(256) void k_sym_mirror(int n, TS *__restrict__ C, int64_t ldc,
                                                    const int *__restrict__ mask) {
    constexpr int T = 64;
    __shared__ TS tile[T][T + 1];
    if (mask && !*mask) return;  // system masked off (whole workgroup, before the barrier)
    int a, b;  // lower tile (a, b), a >= b; it fills the upper tile (b, a)
    tile_of_lower(blockIdx.x, (n + T - 1) / T, a, b);
    const int r0 = a * T, c0 = b * T;  // lower tile origin (rows r0.., columns c0..)
    for (int q = threadIdx.x; q < T * T; q += 256) {
        const int r = q % T, c = q / T;  // r fastest: along a column
        if (r0 + r < n && c0 + c < n) tile[c][r] = C[(int64_t)(r0 + r) + (int64_t)(c0 + c) * ldc];
    }
    __syncthreads();
    for (int q = threadIdx.x; q < T * T; q += 256) {
        const int i = q % T, j = q / T;  // upper element (c0 + i, r0 + j), i fastest: along a column
        const int gi = c0 + i, gj = r0 + j;
        if (gi < gj && gj < n) C[(int64_t)gi + (int64_t)gj * ldc] = tile[i][j];
    }
}

// C := alpha * A * B^T for a product known to be symmetric (n x n; A, B n x k, 'n' and 't'), as
// AW = (C*H)*C^T with H symmetric in omm()'s calc_AW (omm_ops.F90:44-64).  The kernels compute the
// lower-triangle tiles only where they can (ctx->gemm_sym), otherwise the full product; the lower
// triangle is then mirrored, so C is exactly symmetric.
int launch_gemm_sym(omx_ctx ctx, int prec, int dt, int n, int k, double alpha, const void *A, int64_t lda,
                    const void *B, int64_t ldb, void *C, int64_t ldc) {
    ctx->gemm_sym = true;
    int rc = launch_gemm(ctx, prec, dt, false, true, n, n, k, alpha, A, lda, B, ldb, 0.0, C, ldc);
    ctx->gemm_sym = false;
    if (rc) return rc;
    const int nt = (n + 63) / 64;
    const unsigned grid = (unsigned)(nt * (nt + 1) / 2);
    if (dt == OMX_F64)
        hipLaunchKernelGGL(k_sym_mirror<double>, dim3(grid), dim3(256), 0, ctx->stream, n, (double *)C, ldc,
                           ctx->launch_mask);
    else
        hipLaunchKernelGGL(k_sym_mirror<float>, dim3(grid), dim3(256), 0, ctx->stream, n, (float *)C, ldc,
                           ctx->launch_mask);
    OMX_CHECK_HIP(hipGetLastError());
    return OMX_OK;
}

}  // namespace omx

using namespace omx;

extern "C" int omx_gemm(omx_ctx ctx, int prec, int dt, char opA, char opB, int m, int n, int k, double alpha,
                        const void *A, int64_t lda, const void *B, int64_t ldb, double beta, void *C, int64_t ldc) {
    OMX_REQUIRE(ctx, "omx_gemm: null context");
    OMX_REQUIRE(valid_op(opA) && valid_op(opB), "omx_gemm: invalid op ('%c','%c')", opA, opB);
    OMX_REQUIRE(m >= 0 && n >= 0 && k >= 0, "omx_gemm: negative dimension");
    bool trA = is_trans(opA), trB = is_trans(opB);
    OMX_REQUIRE(lda >= std::max(1, trA ? k : m) && ldb >= std::max(1, trB ? n : k) && ldc >= std::max(1, m),
                "omx_gemm: leading dimension too small");
    if (m == 0 || n == 0) return OMX_OK;
    if (k == 0) return omx_scale(ctx, dt, m, n, beta, C, ldc);
    OMX_REQUIRE(A && B && C, "omx_gemm: null operand");
    if (dt == OMX_C128 || dt == OMX_C64)  // complex: FP64 / FP32 MFMA by element type
        return launch_zgemm(ctx, dt, opA, opB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    return launch_gemm(ctx, prec, dt, trA, trB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

extern "C" int omx_gemm_sym(omx_ctx ctx, int prec, int dt, int n, int k, double alpha, const void *A, int64_t lda,
                            const void *B, int64_t ldb, void *C, int64_t ldc) {
    OMX_REQUIRE(ctx, "omx_gemm_sym: null context");
    OMX_REQUIRE(n >= 0 && k >= 0, "omx_gemm_sym: negative dimension");
    OMX_REQUIRE(lda >= std::max(1, n) && ldb >= std::max(1, n) && ldc >= std::max(1, n),
                "omx_gemm_sym: leading dimension too small");
    OMX_REQUIRE(dt == OMX_F64 || dt == OMX_F32 || dt == OMX_C128 || dt == OMX_C64,
                "omx_gemm_sym: unsupported element type %d", dt);
    if (n == 0) return OMX_OK;
    if (k == 0) return omx_scale(ctx, dt, n, n, 0.0, C, ldc);
    OMX_REQUIRE(A && B && C, "omx_gemm_sym: null operand");
    if (dt == OMX_C128 || dt == OMX_C64)  // Hermitian: C = alpha A B^H
        return launch_zgemm_herm(ctx, dt, n, k, alpha, A, lda, B, ldb, C, ldc);
    return launch_gemm_sym(ctx, prec, dt, n, k, alpha, A, lda, B, ldb, C, ldc);
}

extern "C" int omx_gemm_set_tile(omx_ctx ctx, int cfg) {
    OMX_REQUIRE(ctx && cfg >= -1 && cfg < kNumCfg, "omx_gemm_set_tile: configuration %d not in [-1, %d)", cfg,
                kNumCfg);
    ctx->gemm_cfg = cfg;
    return OMX_OK;
}

extern "C" int omx_omm_calc_G(omx_ctx ctx, int prec, int dt, int n, int m, const void *HW, const void *SW,
                              int64_t ldw, const void *HC, const void *SC, void *G, int64_t ldc) {
    OMX_REQUIRE(ctx && HW && SW && HC && SC && G, "omx_omm_calc_G: null argument");
    OMX_REQUIRE(n > 0 && m > 0 && ldw >= n && ldc >= n, "omx_omm_calc_G: bad shape");
    OMX_REQUIRE(G != HC && G != SC, "omx_omm_calc_G: G must not alias HC or SC");
    const bool cplx = (dt == OMX_C128 || dt == OMX_C64);
    // G = 4*HC - 2*(SW*HC + HW*SC): pair product with the 4*HC term folded into the epilogue.
    int rc = cplx ? launch_zgemm_pair(ctx, dt, n, m, n, -2.0, SW, HW, ldw, HC, SC, ldc, 4.0, HC, 0.0, G, ldc)
                  : launch_gemm_pair(ctx, prec, dt, n, m, n, -2.0, SW, HW, ldw, HC, SC, ldc, 4.0, HC, 0.0, G, ldc);
    if (rc != OMX_ERR_UNSUPPORTED) return rc;
    // Unaligned n: three passes (G = 4HC; G -= 2 SW HC; G -= 2 HW SC), as calc_G in omm_ops.F90:35-37.
    rc = omx_add(ctx, dt, 'n', n, m, 4.0, HC, ldc, 0.0, G, ldc);
    if (rc) return rc;
    if (cplx) {
        rc = launch_zgemm(ctx, dt, 'n', 'n', n, m, n, -2.0, SW, ldw, HC, ldc, 1.0, G, ldc);
        if (rc) return rc;
        return launch_zgemm(ctx, dt, 'n', 'n', n, m, n, -2.0, HW, ldw, SC, ldc, 1.0, G, ldc);
    }
    rc = launch_gemm(ctx, prec, dt, false, false, n, m, n, -2.0, SW, ldw, HC, ldc, 1.0, G, ldc);
    if (rc) return rc;
    return launch_gemm(ctx, prec, dt, false, false, n, m, n, -2.0, HW, ldw, SC, ldc, 1.0, G, ldc);
}
