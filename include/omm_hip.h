/*
 * omm_hip.h -- C-ABI of the MI355X (gfx950) OMM hot-path library, libommhip.so.
 *
 * The library is the device half of a MatrixSwitch storage format that keeps matrices resident
 * in HBM ("g" storage labels, operation label "hip") and of libOMM's CG loop driven over it.
 * It is called from the Fortran host (MatrixSwitch and omm() drop-ins, through ISO_C_BINDING
 * interfaces) and, for tests and benchmarks, from Python through ctypes.  The signatures use
 * only plain C types: device pointers are void*, sizes are int / int64_t, scalars are double.
 *
 * Conventions
 *   - Every function returns 0 (OMX_OK) or a negative OMX_ERR_* code; omx_last_error() gives
 *     the message.  Invalid shapes or leading dimensions are rejected before any launch.
 *   - Dense matrices are column-major with a leading dimension, as MatrixSwitch's dval(:,:)
 *     (MatrixSwitch/src/MatrixSwitch_ops.F90:84).  Element types are OMX_F64, OMX_F32,
 *     OMX_BF16, OMX_C128, OMX_C64.
 *   - Work is enqueued on the context's stream and is asynchronous, except functions that
 *     return a host scalar (traces, line-search coefficients, dots), which synchronise.
 *   - A context is used from one host thread at a time; the reference is non-reentrant too
 *     (libOMM/src/omm.F90:62-106 keeps module SAVE state).
 */
#ifndef OMM_HIP_H
#define OMM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------------------------ */
#define OMX_OK 0
#define OMX_ERR_ARG (-1)         /* invalid argument / incompatible shapes */
#define OMX_ERR_HIP (-2)         /* HIP runtime error */
#define OMX_ERR_NOMEM (-3)       /* device allocation failed */
#define OMX_ERR_UNSUPPORTED (-4) /* combination not implemented */
#define OMX_ERR_STATE (-5)       /* call sequence error */

/* ---- element types and product precisions ------------------------------------------------ */
#define OMX_F64 0
#define OMX_F32 1
#define OMX_BF16 2
#define OMX_C128 3
#define OMX_C64 4

#define OMX_PREC_F64 0  /* FP64 MFMA, FP64 accumulate (the oracle-grade path) */
#define OMX_PREC_F32 1  /* operands rounded to FP32, FP32 MFMA, FP32 accumulate */
#define OMX_PREC_BF16 2 /* operands rounded to BF16, BF16 MFMA, FP32 accumulate */
/* Operands split into two BF16 terms (hi = bf16(a), lo = bf16(a - hi)); hi*hi + hi*lo + lo*hi on
 * BF16 MFMA with FP32 accumulation, about 16 significant bits per operand (real FP32 or FP64
 * storage; shapes the BF16 panels do not cover run as OMX_PREC_F32) */
#define OMX_PREC_BF16X3 3
/* Operands split into three BF16 terms (hi, mid = bf16(a - hi), lo = bf16(a - hi - mid); exact for
 * FP32 values); the six term products of order >= 2^-16 on BF16 MFMA with FP32 accumulation:
 * FP32-grade products (real FP32 or FP64 storage; shapes the BF16 panels do not cover run as
 * OMX_PREC_F32) */
#define OMX_PREC_BF16X6 4

/* ---- counter-based generator streams (lattice inputs) ------------------------------------- */
#define OMX_STREAM_ONSITE 1
#define OMX_STREAM_HOP 2
#define OMX_STREAM_OVL 3
#define OMX_STREAM_C0 4
#define OMX_STREAM_THETA 5

/* ---- kernel classes for event timing (omx_prof_*) ----------------------------------------- */
#define OMX_K_SPMM 0
#define OMX_K_GEMM 1
#define OMX_K_COEFF 2
#define OMX_K_ELEM 3
#define OMX_K_DOT 4
#define OMX_K_GEMM_PAIR 5 /* calc_G's fused pair products (OMX_K_GEMM: every other product) */
#define OMX_K_COUNT 6

typedef struct omx_ctx_s *omx_ctx;
typedef struct omx_bsr_s *omx_bsr;

/* ==== context, memory ====================================================================== */

/* Number of visible HIP devices (0 without a GPU; does not fail). */
int omx_device_count(int *count);
/* Create a context on `device` with its own (default-priority) stream. */
int omx_init(int device, omx_ctx *ctx);
int omx_finalize(omx_ctx ctx);
/* Wait for all work enqueued on the context's stream. */
int omx_sync(omx_ctx ctx);
/* Message of the last error on this host thread ("" if none). */
const char *omx_last_error(void);
/* hipStream_t of the context, as void*, for callers that enqueue their own work. */
void *omx_stream(omx_ctx ctx);

/* Device allocation, zero-filled, as m_allocate (MatrixSwitch/src/MatrixSwitch.F90:394-415). */
int omx_malloc(omx_ctx ctx, size_t bytes, void **dptr);
int omx_free(omx_ctx ctx, void *dptr);
/* Zero `bytes` of device memory with a fill kernel (safe under stream capture). */
int omx_zero(omx_ctx ctx, void *dptr, size_t bytes);
int omx_h2d(omx_ctx ctx, void *dst, const void *src, size_t bytes);
int omx_d2h(omx_ctx ctx, void *dst, const void *src, size_t bytes);
int omx_d2d(omx_ctx ctx, void *dst, const void *src, size_t bytes);
/* Strided 2-D host<->device copies of an m x n column-major matrix of element size esize. */
int omx_h2d_2d(omx_ctx ctx, void *dst, int64_t ldd, const void *src, int64_t lds, int m, int n, int esize);
int omx_d2h_2d(omx_ctx ctx, void *dst, int64_t ldd, const void *src, int64_t lds, int m, int n, int esize);
/* Element-type conversion of an m x n matrix (any real type pair). */
int omx_convert(omx_ctx ctx, int m, int n, const void *src, int dt_src, int64_t lds, void *dst, int dt_dst,
                int64_t ldd);

/* ==== BSR-4 sparse operators (H, S) ======================================================== */

/*
 * Upload a basis-size nb operator given as 0-based BSR-4 block columns (block-CSC: bcolptr has
 * ceil(nb/4)+1 entries, browidx block rows ascending, 16 column-major values per block).
 * Replaces the CSC registration m_register_pcsc / psp_register_spm
 * (MatrixSwitch/src/MatrixSwitch_m_register.F90:394-443, pspBLAS/src/pspVariable/pspMat.F90:87-140):
 * like pspBLAS, the data are copied, here into HBM.
 */
int omx_bsr_create(omx_ctx ctx, int nb, const int *bcolptr, const int *browidx, const double *bval, omx_bsr *out);
/*
 * Build the BSR-4 operator from a CSC matrix (index_base 0 or 1; colptr 0- or 1-based with
 * the same base).  Accepts both MatrixSwitch's 'sdcsc' (1-based rows iaux3, 0-based offsets
 * iaux4: pass rowbase=1, ptrbase=0; MatrixSwitch/src/MatrixSwitch_m_copy.F90:574-599) and
 * pspBLAS's pdcsc (both 1-based; pspMat.F90:31-60).  Blocks that hold no entry are not stored.
 */
int omx_bsr_from_csc(omx_ctx ctx, int nb, const int *colptr, const int *rowidx, const double *val, int ptrbase,
                     int rowbase, omx_bsr *out);
/* The host half of omx_bsr_from_csc / omx_bsr_from_csc_z (no device, no context): the BSR-4 block
 * columns of a CSC matrix.  colptr must start at ptrbase and never decrease (checked before any
 * entry is read).  Call with null outputs to get *nnzb, then with bcolptr[ceil(nb/4)+1],
 * browidx[*nnzb] and bval[*nnzb * (is_complex ? 32 : 16)]. */
int omx_csc_to_bsr(int nb, const int *colptr, const int *rowidx, const double *val, int ptrbase, int rowbase,
                   int is_complex, int64_t *nnzb, int *bcolptr, int *browidx, double *bval);
/* Complex (k-point) operators: 32 doubles per block, (re, im) interleaved, column-major inside
 * the block; CSC values (re, im) interleaved per entry.  Same conventions as the real calls
 * (szcsc / pzcsc storage, MatrixSwitch_m_copy.F90, pspMat.F90). */
int omx_bsr_create_z(omx_ctx ctx, int nb, const int *bcolptr, const int *browidx, const double *bval,
                     omx_bsr *out);
int omx_bsr_from_csc_z(omx_ctx ctx, int nb, const int *colptr, const int *rowidx, const double *val, int ptrbase,
                       int rowbase, omx_bsr *out);
int omx_bsr_destroy(omx_bsr a);
int omx_bsr_info(omx_bsr a, int *nb, int64_t *nnzb, int *is_complex);
/* Replace the block values (same pattern), e.g. after a parameter change. */
int omx_bsr_update_values(omx_bsr a, const double *bval);
/* Copy the block values (16 per stored block, column-major) / the block pattern to the host. */
int omx_bsr_get_values(omx_bsr a, double *bval);
int omx_bsr_get_pattern(omx_bsr a, int *bcolptr, int *browidx);
/* Element (i, j), 1-based, of the operator (0 when its block is not stored): m_get_element on a
 * sparse matrix (MatrixSwitch/src/MatrixSwitch.F90:2917-3086, 'sdcsc' branch). */
int omx_bsr_get_element(omx_bsr a, int i, int j, double *out);

/* ==== MatrixSwitch operations on device matrices =========================================== */

/*
 * Y := alpha*X*A + beta*Y, X and Y no x nb dense (element type dt: OMX_F64, OMX_F32, OMX_C128,
 * OMX_C64), A a BSR-4 operator.  The dense x sparse product mm_multiply(C,'n',H,'n',HC)
 * (libOMM/src/omm_ops.F90:61 -> MatrixSwitch/src/MatrixSwitch_mm_multiply.F90:263-327 'sdcsc',
 * pspBLAS/src/pspLevel3/pspMspm_nn.F90:58-212 'pdcsc').
 */
int omx_spmm(omx_ctx ctx, int dt, int no, int nb, double alpha, const void *X, int64_t ldx, omx_bsr A, double beta,
             void *Y, int64_t ldy);

/*
 * Cp := alpha*(A^T B) on the stored blocks of Cp + beta*Cp, A and B no x nb dense (dt OMX_F64 or
 * OMX_F32), Cp a BSR-4 pattern (its values are overwritten): the dense x dense -> sparse product
 * mm_multiply(A,'c',B,'n',Cp) (MatrixSwitch/src/MatrixSwitch_mm_multiply.F90:334-398).  With
 * A = C and B = (2I - SW) C it is the density matrix on the pattern of H, which is also dE/dH.
 */
int omx_sddmm_bsr(omx_ctx ctx, int dt, int no, int nb, double alpha, const void *A, int64_t lda, const void *B,
                  int64_t ldb, double beta, omx_bsr Cp);
/*
 * Adjoint contraction dE/dtheta_p = sum_e J_p(e) * Dp(e) over the stored values of Dp (dE/dH on
 * H's pattern, from omx_sddmm_bsr or omm() with a 'gdcsc' D_min) for a Hamiltonian linear in
 * theta.  seg (nparam+1), ent, coef: device arrays in the layout of omx_lattice_theta_jacobian.
 * grad: host, nparam <= 256.  No counterpart in the reference (libOMM has no derivatives); it is
 * the backward pass of omm() with respect to the Hamiltonian's parameters.
 */
int omx_bsr_param_grad(omx_ctx ctx, omx_bsr Dp, int nparam, const int64_t *seg, const int64_t *ent,
                       const double *coef, double *grad);

/*
 * C := alpha*op(A)*op(B) + beta*C  (mm_multiply on dense operands,
 * MatrixSwitch/src/MatrixSwitch.F90:1046-1390; 'ref' semantics of
 * MatrixSwitch_mm_multiply.F90:19-79).  op = 'n', 't' or 'c' ('c' = 't' for real data).
 * A, B and C share the element type dt (OMX_F64 / OMX_F32); prec selects the MFMA path.
 * Complex dt (OMX_C128 on FP64 MFMA, OMX_C64 on FP32 MFMA; prec ignored): 'c' is the conjugate
 * transpose (mm_zmultiply, MatrixSwitch.F90:1395-1726).  alpha, beta are real.
 * beta == 0 overwrites C (BLAS convention).
 */
int omx_gemm(omx_ctx ctx, int prec, int dt, char opA, char opB, int m, int n, int k, double alpha, const void *A,
             int64_t lda, const void *B, int64_t ldb, double beta, void *C, int64_t ldc);
/* C := alpha*A*B^T for a product the caller knows to be symmetric, n x n with A and B n x k ('n',
 * 't'): omm()'s calc_AW Gram AW = (C*H)*C^T with H symmetric (libOMM/src/omm_ops.F90:44-64, the
 * mm_multiply(AC,'n',C,'c',AW) of :62 with beta = 0).  Only the lower-triangle tiles are computed
 * where the kernel allows it, and the lower triangle is mirrored, so C is exactly symmetric.
 * Real element types (OMX_F64, OMX_F32) with prec as in omx_gemm; complex types (OMX_C128,
 * OMX_C64; prec ignored) compute C := alpha*A*B^H, Hermitian, with the upper triangle the
 * conjugate of the lower.  The reference computes the full product with dgemm/zgemm; this entry
 * point is the symmetric-result form of it. */
int omx_gemm_sym(omx_ctx ctx, int prec, int dt, int n, int k, double alpha, const void *A, int64_t lda, const void *B,
                 int64_t ldb, void *C, int64_t ldc);
/* SpMM launch shape of later real products on this context: blocks whose loads are issued
 * together (batch 1, 4 or 8), waves stacked in rows per workgroup (wpg 1, 2 or 4), block columns
 * walked per workgroup (jper >= 1) and XCD map (mode 0: each XCD owns a contiguous eighth of the
 * rows; 1: each XCD walks one row chunk at a time).  batch = 0 returns to the default shape (or
 * the autotuner's).  Tuning hook; results do not depend on it. */
int omx_spmm_set_tile(omx_ctx ctx, int batch, int wpg, int jper, int mode);
/* With enable != 0, real operators whose pattern is the periodic 7-point stencil of an
 * Lx x Ly x Lz lattice of 4x4 blocks (recognised when the operator is created; Lx, Ly multiples of
 * 4) run through a 2.5-D kernel that keeps z-planes of X in registers and LDS instead of gathering
 * each basis column 7 times: 1.2x the algorithmic HBM bytes instead of 2.1x, but slower than the
 * general kernel at n = 16384 (0.89 ms against 0.83 ms), so off by default.  Tuning hook; the two
 * kernels give the same bits. */
int omx_spmm_stencil(omx_ctx ctx, int enable);
/* Lattice edges (Lx, Ly, Lz) recognised in the operator's pattern, all 0 when it is not such a
 * stencil. */
int omx_bsr_lattice(omx_bsr a, int *dims);
/* Attach to an operator the order in which SpMMs visit its block columns (a permutation of
 * 0..nbc-1, host array copied to the device; NULL clears it).  A schedule hint only: every
 * output column is computed the same way in any order (no counterpart in the reference). */
int omx_bsr_set_walk(omx_bsr A, const int *order);
/* The same with LDS-staged gathers: later real SpMMs take the block columns `group` at a time (in
 * walk order) and load the union of the basis-column blocks a group gathers into LDS once, for
 * 64-state chunks.  Fails if a group gathers more than 40 distinct blocks (FP64; 80 for FP32 is
 * not used).  group = 0 is omx_bsr_set_walk.  Results are bitwise those of the default kernel. */
int omx_bsr_set_walk_groups(omx_bsr A, const int *order, int group);
/* Force the GEMM tile configuration of later products on this context (-1 restores the default)
 * for the Gram and gradient products: 0 = 128x128 tiles, 4 waves, two workgroups per CU;
 * 1 = 128x128, one per CU; 2 = 256x128; 3 = 128x256; 5 = 128x128 with 8 waves of 32x64 (the FP64
 * default); 6 = 128x128 with 8 waves of 64x32 (FP64 products); 4 = 256x256 with 8 waves (FP32 and
 * BF16 products on FP32 storage); 7 = the tiles of 5 with the operands loaded straight into LDS
 * (FP64 Gram and gradient products with m, n % 128 == 0 and k % 16 == 0, the default there; other
 * shapes take 5; and FP32 'n','t' Grams on FP32 storage with m, n % 128 == 0 and k % 32 == 0 whose
 * tiles fill the chip, the default there; other shapes take 0).  A configuration that does not
 * apply to a product falls back to 0.  Used by the tile autotuner; results do not depend on it. */
int omx_gemm_set_tile(omx_ctx ctx, int cfg);
/* Autotuner (project-written; no counterpart in the reference): with enable != 0, the first
 * product or SpMM of each shape on this context times every applicable tile configuration / SpMM
 * launch shape into a scratch result and keeps the fastest for that shape.  Tuning allocates and
 * synchronises: do not enable it while a stream is being captured.  Forced shapes
 * (omx_gemm_set_tile, omx_spmm_set_tile) take precedence. */
int omx_autotune(omx_ctx ctx, int enable);
/* The choices made so far, one line per shape, NUL-terminated into buf (len bytes); returns the
 * full text length. */
int omx_autotune_report(omx_ctx ctx, char *buf, int64_t len);
/* C := alpha*op(A) + beta*C  (m_add, MatrixSwitch.F90:1731-1902; MatrixSwitch_m_add.F90:95-175). */
int omx_add(omx_ctx ctx, int dt, char opA, int m, int n, double alpha, const void *A, int64_t lda, double beta,
            void *C, int64_t ldc);
/* C := beta*C  (m_scale, MatrixSwitch.F90:2496-2621). */
int omx_scale(omx_ctx ctx, int dt, int m, int n, double beta, void *C, int64_t ldc);
/* C_ij := alpha (i /= j), beta (i == j) on 'a'll, 'u'pper or 'l'ower part
 * (m_set, MatrixSwitch.F90:2747-2827; MatrixSwitch_m_set.F90:19-75). */
int omx_set(omx_ctx ctx, int dt, char seC, int m, int n, double alpha, double beta, void *C, int64_t ldc);
/* alpha := tr(A), FP64 accumulation (m_trace, MatrixSwitch.F90:2060-2156; complex: out = (re, im),
 * :2161-2248). Synchronises. */
int omx_trace(omx_ctx ctx, int dt, int n, const void *A, int64_t lda, double *out);
/* alpha := sum_ij A_ij*B_ij = tr(A^T B), FP64 accumulation (mm_trace, MatrixSwitch.F90:2253-2377;
 * complex: sum conj(A_ij)*B_ij, :2471, out = (re, im)). Synchronises. */
int omx_mm_trace(omx_ctx ctx, int dt, int m, int n, const void *A, int64_t lda, const void *B, int64_t ldb,
                 double *out);
/* C_ij := alpha + beta*C_ij for one element, 1-based (m_set_element, MatrixSwitch.F90:2917-3086). */
int omx_set_element(omx_ctx ctx, int dt, void *C, int64_t ldc, int i, int j, double alpha, double beta);
/* alpha := C_ij, 1-based (m_get_element, MatrixSwitch.F90:3309-3417); complex types write
 * (re, im) to out[0..1]. Synchronises. */
int omx_get_element(omx_ctx ctx, int dt, const void *C, int64_t ldc, int i, int j, double *out);
/* C_ij := alpha + beta*C_ij with complex alpha, 1-based (m_set_element on complex matrices). */
int omx_set_element_z(omx_ctx ctx, int dt, void *C, int64_t ldc, int i, int j, double alpha_re, double alpha_im,
                      double beta);

/* ==== fused libOMM kernels ================================================================== */

/*
 * The five quartic line-search coefficients from one fused FP64 reduction over HW, SW, HWd,
 * SWd, HWdd, SWdd (n x n, element type dt, leading dimension ld): replaces calc_coeff's three
 * m_trace, ten mm_trace and the SWd^T copy (libOMM/src/omm_ops.F90:156-219).  The reduction
 * is two-stage and deterministic.  Synchronises and returns coeff[0..4] (coeff[0] is the energy).
 */
int omx_omm_coeff(omx_ctx ctx, int dt, int n, int64_t ld, const void *HW, const void *SW, const void *HWd,
                  const void *SWd, const void *HWdd, const void *SWdd, double *coeff);

/*
 * G := 4*HC - 2*SW*HC - 2*HW*SC in one MFMA product with a K = 2n contraction
 * (calc_G, libOMM/src/omm_ops.F90:17-39).  HW, SW: n x n; HC, SC, G: n x m.
 */
int omx_omm_calc_G(omx_ctx ctx, int prec, int dt, int n, int m, const void *HW, const void *SW, int64_t ldw,
                   const void *HC, const void *SC, void *G, int64_t ldc);

/* out[0] = <G, G - G_p>, out[1] = <G_p, G_p> in FP64 (omm.F90:536-546, Polak-Ribiere). Synchronises. */
int omx_omm_pr_dots(omx_ctx ctx, int dt, int n, int m, const void *G, const void *G_p, int64_t ld, double *out);

/* omx_omm_pr_dots without any allocation.  No counterpart in the reference: for callers that
 * cannot allocate device memory (code under stream capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The partial sums live in a file-scope device array of 32 KB (2048 workgroups'
 * partials and the 2 final sums) that each GPU holds from omx_init until the process ends; inputs
 * of any size use at most 2048 workgroups, the same launch as omx_omm_pr_dots, so the two return
 * the same bits.  Every call on one GPU uses that GPU's single array: calls that overlap on one GPU
 * (other host threads, other contexts on the device) race and must be serialised by the caller.
 * Synchronises. */
int omx_omm_pr_dots_noalloc(omx_ctx ctx, int dt, int n, int m, const void *G, const void *G_p, int64_t ld,
                            double *out);

/* D := G + lambda*D and G_p := G in one pass (omm.F90:440-445). */
int omx_omm_cg_direction(omx_ctx ctx, int dt, int n, int m, double lambda, const void *G, void *D, void *G_p,
                         int64_t ld);

/* W := W + x*(Wd^T + Wd) + x^2*Wdd in one pass (omm.F90:496-502 and 511-517). */
int omx_omm_update_w(omx_ctx ctx, int dt, int n, int64_t ld, double x, const void *Wd, const void *Wdd, void *W);

/* ==== batched k-point solver ============================================================== */

/*
 * Many complex128 systems of one shape whose operators share one BSR-4 block pattern (the
 * k-points of one cell, H(k) and S(k) from omx_lattice_bsr_k) solved together by omm()'s
 * flavour-0 CG loop (libOMM/src/omm.F90:350-552), all decisions (quartic line search,
 * convergence, rescaling, Polak-Ribiere) taken on the device: one CG step of all systems is a
 * fixed launch sequence, captured once in a hipGraph and replayed.  No counterpart in the
 * reference, which loops over the k-points (np, ip) one omm() call at a time.
 * hval, sval: nsys * nnzb * 32 doubles (system s at s*32*nnzb, H already shifted by -eta*S).
 */
typedef struct omx_zbatch_s *omx_zbatch;
int omx_zbatch_create(omx_ctx ctx, int nsys, int no, int nb, const int *bcolptr, const int *browidx,
                      const double *hval, const double *sval, omx_zbatch *out);
/* The same with the state in dt = OMX_C128 or OMX_C64 (complex64: FP32 MFMA products, operator
 * values rounded to FP32 in the SpMM; traces, coefficients and the line search stay FP64).
 * omx_zbatch_create is omx_zbatch_create_dt with OMX_C128. */
int omx_zbatch_create_dt(omx_ctx ctx, int dt, int nsys, int no, int nb, const int *bcolptr, const int *browidx,
                         const double *hval, const double *sval, omx_zbatch *out);
int omx_zbatch_destroy(omx_zbatch b);
/* Coefficient matrices of all systems (host, in the batch's element type, (re, im), no x nb
 * column-major each). */
int omx_zbatch_set_c(omx_zbatch b, const void *C);
int omx_zbatch_get_c(omx_zbatch b, void *C);
/* Run up to max_steps CG steps per system.  use_graph captures check_every steps as one graph
 * (re-captured only when check_every changes) and launches it until every system has converged
 * or reached max_steps, reading the convergence flags once per launch; cg_tol and max_steps live
 * in device memory, so every solve replays the graph with its own values.  Without use_graph the
 * same steps are launched eagerly.  Returns per system e_min + eta*TrQS, the step count icg and
 * the converged flag (e_diff <= cg_tol, cg_tol < 0: 1e-9); *steps_run counts the step sequences
 * issued, including those that were no-ops for every system. */
int omx_zbatch_solve(omx_zbatch b, double eta, double cg_tol, int max_steps, int check_every, int use_graph,
                     double *e_min, int *icg, int *converged, int *steps_run);

/* ==== uneven-size group of real systems ===================================================== */
/* Replaces the reference's sequence of independent omm() calls, one per system (np/ip state,
 * libOMM/src/omm.F90:34-35, 96-106), for real flavour-0 systems of different sizes: system s has
 * no[s] states and the operators H[s] (already shifted by -eta_s S as omm() expects) and S[s]
 * (created with omx_bsr_create on ctx; the group keeps references, not copies).  Every CG
 * decision is taken on the device; one step of all systems is captured as one graph: the systems
 * whose products fill the GPU on their own in sequence, each other system on a parallel branch
 * beside them.  With omx_autotune on, each system's shapes are tuned in the eager
 * initial pass of the first solve, before the capture. */
typedef struct omx_group_s *omx_group;
int omx_group_create(omx_ctx ctx, int nsys, const int *no, const omx_bsr *H, const omx_bsr *S, omx_group *out);
int omx_group_destroy(omx_group g);
/* C of system s (no[s] x nb[s] column-major, ld = no[s]), host memory. */
int omx_group_set_c(omx_group g, int s, const double *C);
int omx_group_get_c(omx_group g, int s, double *C);
/* Solve all systems from their current C (init_C semantics): e_min[s] (with + eta[s]*TrQS),
 * icg[s], converged[s]; graphs of check_every steps and the device-side cg_tol and max_steps as in
 * omx_zbatch_solve. */
int omx_group_solve(omx_group g, const double *eta, double cg_tol, int max_steps, int check_every, int use_graph,
                    double *e_min, int *icg, int *converged, int *steps_run);

/* ==== dense factorisations (Cholesky and preconditioned flavours) ========================= */
/* All FP64 column-major; U is upper triangular and only its upper triangle is read. */

/* A = U^T U in place, upper (m_factorize, libOMM/src/omm_ops.F90:246-315, dpotrf 'u').  *info > 0
 * names the first non-positive pivot (1-based).  As with dpotrf 'u', only the upper triangle is read
 * and written; the strictly lower triangle is left as it was (the factorisation runs on a transposed
 * copy of the upper triangle in a temporary device buffer of order n rounded up to 128).
 * Synchronises. */
int omx_potrf(omx_ctx ctx, int dt, int n, void *A, int64_t lda, int *info);
/* B := alpha*op(U)^-1*B (side 'l') or alpha*B*op(U)^-1 (side 'r'), op 'n' or 't' (dtrsm; used by
 * m_back_transform, omm_ops.F90:829-897). */
int omx_trsm(omx_ctx ctx, int dt, char side, char trans, int m, int n, double alpha, const void *U, int64_t ldu,
             void *B, int64_t ldb);
/* B := alpha*op(U)*B (side 'l') or alpha*B*op(U) (side 'r') (dtrmm; m_transform, omm_ops.F90:662-730). */
int omx_trmm(omx_ctx ctx, int dt, char side, char trans, int m, int n, double alpha, const void *U, int64_t ldu,
             void *B, int64_t ldb);
/* H := U^-T H U^-1, symmetric result in both triangles (m_reduce: dsygst itype 1 'u' and the
 * mirror copy, omm_ops.F90:417-524). */
int omx_sygst(omx_ctx ctx, int dt, int n, void *H, int64_t ldh, const void *U, int64_t ldu);
/* A := A^-1 for symmetric positive definite A (m_inverse of the preconditioner S + T/scale_T,
 * omm_ops.F90:992-1106; the reference factors with dsytrf, here Cholesky).  Synchronises. */
int omx_inverse_spd(omx_ctx ctx, int dt, int n, void *A, int64_t lda, int *info);

/* ==== event timing ========================================================================= */

/* Start (enable=1) or stop accumulating HIP-event durations of launches of each kernel class. */
int omx_prof_enable(omx_ctx ctx, int enable);
/* Total milliseconds and launch count accumulated for a kernel class (synchronises). */
int omx_prof_read(omx_ctx ctx, int kclass, double *ms, int64_t *launches);
int omx_prof_reset(omx_ctx ctx);
/* Algorithmic flops and minimum HBM bytes of the launches timed for a kernel class:
 * products 2*m*n*k flops (operands read once, result written once); SpMM 2*nnz*no flops,
 * X read + Y written + block values and indices; reductions and element-wise passes: each
 * operand element read or written once. */
int omx_prof_work(omx_ctx ctx, int kclass, double *flops, double *bytes);

/* ==== synthetic inputs ====================================================================== */

/*
 * Periodic Lx x Ly x Lz simple-cubic lattice, 4 orbitals per site, nearest-neighbour blocks.
 * which = 0: H (onsite diag(-5,1,1,1) + sym. U(-0.1,0.1) noise, hopping U(-1,1) blocks);
 * which = 1: S (onsite I, bonds 0.03*U(-1,1)).  First call with bcolptr == NULL returns the
 * block count in *nnzb; second call fills bcolptr (Lx*Ly*Lz+1), browidx (nnzb), bval (16*nnzb).
 */
int omx_lattice_bsr(int Lx, int Ly, int Lz, uint64_t seed, int which, int64_t *nnzb, int *bcolptr, int *browidx,
                    double *bval);
/* Bloch operators H(k) (which = 0) / S(k) (which = 1) of the periodic cell: the blocks of
 * omx_lattice_bsr, with the bonds that cross the cell boundary along d multiplied by
 * exp(2 pi i phi[d]) (forward) and exp(-2 pi i phi[d]) (backward), i.e. sum_R H_R exp(i k.R) with
 * k.R = 2 pi phi[d] per cell translation (a uniform N-point grid along d: phi = n/N).  Hermitian.
 * Values: 32 doubles per block, (re, im) interleaved.  Two-call protocol as omx_lattice_bsr. */
int omx_lattice_bsr_k(int Lx, int Ly, int Lz, uint64_t seed, int which, const double *phi, int64_t *nnzb,
                      int *bcolptr, int *browidx, double *bval);
/* H(theta): one hopping block theta[0..15] (column-major) on every bond, onsite diag(theta[16..19]). */
int omx_lattice_bsr_theta(int Lx, int Ly, int Lz, const double *theta, int64_t *nnzb, int *bcolptr, int *browidx,
                          double *bval);
/* Number of generator parameters theta (16 hopping-block entries + 4 onsite levels). */
#define OMX_THETA_COUNT 20
/*
 * Jacobian of H(theta) on its block pattern (H is linear in theta): for parameter p, entries
 * seg[p]..seg[p+1]-1 of (ent, coef) give the stored value index e = 16*l + r + 4*c and dH(e)/dtheta_p.
 * First call with seg == NULL returns the entry count in *nent; second call fills seg
 * (OMX_THETA_COUNT+1), ent and coef (*nent each).  No counterpart in the reference (its inputs
 * are files); it is the generator side of the adjoint.
 */
int omx_lattice_theta_jacobian(int Lx, int Ly, int Lz, int64_t *nent, int64_t *seg, int64_t *ent, double *coef);
/* Initial parameters: hopping U(-1,1) from the seed, onsite levels (-6, -6, 6, 6) (half filling
 * then has a gap of about 3.8). */
int omx_lattice_theta0(uint64_t seed, double *theta);
/* Tiled walk order of the lattice sites (x-y tiles of tx x ty, each swept through z), a
 * permutation of 0..Lx*Ly*Lz-1 for omx_bsr_set_walk on the lattice operators. */
int omx_lattice_walk(int Lx, int Ly, int Lz, int tx, int ty, int *order);
/* Patch walk order: tx x ty x tz patches of sites listed one after another (groups of tx*ty*tz
 * consecutive entries are compact patches), for omx_bsr_set_walk_groups. */
int omx_lattice_walk3(int Lx, int Ly, int Lz, int tx, int ty, int tz, int *order);
/* C0(i,k) = U(-0.5,0.5)/sqrt(nb), no x nb column-major. */
int omx_random_c0(int no, int nb, uint64_t seed, double *C, int64_t ldc);
/* The underlying counter-based uniform variate in [0,1). */
double omx_u01(uint64_t seed, uint64_t stream, uint64_t idx);

#ifdef __cplusplus
}
#endif
#endif /* OMM_HIP_H */
