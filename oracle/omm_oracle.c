/*
 * CPU oracle for the OMM conjugate-gradient hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * This file restates, in plain C, the algorithm of libOMM's omm() (basic flavour 0, with or
 * without S) and the MatrixSwitch 'sdden'/'sdcsc' + 'ref' operations it drives.  Only tests,
 * the smoke check and the CPU-baseline leg of the benchmark may load it, and only as the
 * checker.  The product path (the HIP library and the Fortran host) never calls it.
 *
 * Every per-element floating-point sequence follows the reference's 'ref' loops so that,
 * compiled without FMA contraction, results are bit-identical to the reference build:
 *   - dense products: C := beta*C first, then C(i,j) += (alpha*A(i,k))*B(k,j) for k ascending
 *     (MatrixSwitch/src/MatrixSwitch_mm_multiply.F90:43-77);
 *   - dense x CSC product: C := beta*C, then per column j, per stored entry l ascending,
 *     C(i,j) += (alpha*A(i,row(l)))*val(l)  (MatrixSwitch_mm_multiply.F90:287-297);
 *   - add: C := beta*C then C += alpha*op(A)  (MatrixSwitch/src/MatrixSwitch_m_add.F90:117-127);
 *   - traces: sequential sums, mm_trace row-outer (MatrixSwitch/src/MatrixSwitch.F90:2137-2141,
 *     2346-2352).
 * The loop nests are blocked and threaded for speed, but k (or l) is always consumed in
 * ascending order per output element, so the rounding sequence is unchanged.
 *
 * Parity pins (tests/test_oracle_*.py): the reference serial build in oracle/_ref (same
 * inputs, bit-level comparison), example_gamma / example_sdden known answers, closed forms
 * (generalised eigenproblem, quartic brute force, finite-difference gradient, polynomial
 * identity of the line-search coefficients) and the LCG recurrence.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define IDX(i, j, ld) ((size_t)(i) + (size_t)(j) * (size_t)(ld))

/* ------------------------------------------------------------------------------------------ */
/* MatrixSwitch 'ref' operations                                                               */
/* ------------------------------------------------------------------------------------------ */

static int is_trans(char op) { return op == 't' || op == 'T' || op == 'c' || op == 'C'; }

/* C := beta*C  (MatrixSwitch_mm_multiply.F90:43, MatrixSwitch_m_add.F90:117) */
static void scale_inplace(int m, int n, double beta, double *C, int ldc) {
#pragma omp parallel for schedule(static)
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < m; ++i) C[IDX(i, j, ldc)] = beta * C[IDX(i, j, ldc)];
}

/*
 * C := alpha*op(A)*op(B) + beta*C  with the 'ref' rounding sequence
 * (MatrixSwitch_mm_multiply.F90:19-79).  m, n = dims of C, k = contraction length.
 * Blocked: packed panels of alpha*op(A) (alpha applied first, as in the reference expression
 * alpha*A(i,k)*B(k,j)) and op(B); per element the k order is ascending.
 */
#define OR_MC 64
#define OR_NC 64
#define OR_KC 256
void or_gemm(char opA, char opB, int m, int n, int k, double alpha, const double *A, int lda,
             const double *B, int ldb, double beta, double *C, int ldc) {
    int trA = is_trans(opA), trB = is_trans(opB);
    scale_inplace(m, n, beta, C, ldc);
    if (m <= 0 || n <= 0 || k <= 0) return;
    int nib = (m + OR_MC - 1) / OR_MC, njb = (n + OR_NC - 1) / OR_NC;
#pragma omp parallel
    {
        double *ap = (double *)malloc(sizeof(double) * OR_MC * OR_KC);
        double *bp = (double *)malloc(sizeof(double) * OR_KC * OR_NC);
#pragma omp for schedule(dynamic) collapse(2)
        for (int jb = 0; jb < njb; ++jb)
            for (int ib = 0; ib < nib; ++ib) {
                int i0 = ib * OR_MC, j0 = jb * OR_NC;
                int mb = m - i0 < OR_MC ? m - i0 : OR_MC;
                int nb = n - j0 < OR_NC ? n - j0 : OR_NC;
                for (int k0 = 0; k0 < k; k0 += OR_KC) {
                    int kb = k - k0 < OR_KC ? k - k0 : OR_KC;
                    /* pack ap[kk*OR_MC + ii] = alpha*op(A)(i0+ii, k0+kk) */
                    for (int kk = 0; kk < kb; ++kk)
                        for (int ii = 0; ii < mb; ++ii) {
                            double a = trA ? A[IDX(k0 + kk, i0 + ii, lda)] : A[IDX(i0 + ii, k0 + kk, lda)];
                            ap[kk * OR_MC + ii] = alpha * a;
                        }
                    /* pack bp[jj*OR_KC + kk] = op(B)(k0+kk, j0+jj) */
                    for (int jj = 0; jj < nb; ++jj)
                        for (int kk = 0; kk < kb; ++kk)
                            bp[jj * OR_KC + kk] = trB ? B[IDX(j0 + jj, k0 + kk, ldb)] : B[IDX(k0 + kk, j0 + jj, ldb)];
                    for (int jj = 0; jj < nb; ++jj) {
                        double *c = &C[IDX(i0, j0 + jj, ldc)];
                        const double *b = &bp[jj * OR_KC];
                        if (mb == OR_MC) {
                            double acc[OR_MC];
                            for (int ii = 0; ii < OR_MC; ++ii) acc[ii] = c[ii];
                            for (int kk = 0; kk < kb; ++kk) {
                                const double bk = b[kk];
                                const double *a = &ap[kk * OR_MC];
                                for (int ii = 0; ii < OR_MC; ++ii) acc[ii] = acc[ii] + a[ii] * bk;
                            }
                            for (int ii = 0; ii < OR_MC; ++ii) c[ii] = acc[ii];
                        } else {
                            for (int kk = 0; kk < kb; ++kk) {
                                const double bk = b[kk];
                                const double *a = &ap[kk * OR_MC];
                                for (int ii = 0; ii < mb; ++ii) c[ii] = c[ii] + a[ii] * bk;
                            }
                        }
                    }
                }
            }
        free(ap);
        free(bp);
    }
}

/*
 * C := alpha*op(A)*B + beta*C with B a CSC matrix (0-based colptr/rowidx), op(B)='n'
 * (MatrixSwitch_mm_multiply.F90:263-297, both trA branches).  C is m x nb, B is kdim x nb.
 */
void or_spmm_csc(char opA, int m, int nb, double alpha, const double *A, int lda, const int *colptr,
                 const int *rowidx, const double *val, double beta, double *C, int ldc) {
    int trA = is_trans(opA);
    scale_inplace(m, nb, beta, C, ldc);
#pragma omp parallel for schedule(static)
    for (int j = 0; j < nb; ++j) {
        double *c = &C[IDX(0, j, ldc)];
        for (int l = colptr[j]; l < colptr[j + 1]; ++l) {
            int k = rowidx[l];
            double v = val[l];
            if (!trA) {
                const double *a = &A[IDX(0, k, lda)];
                for (int i = 0; i < m; ++i) c[i] = c[i] + alpha * a[i] * v;
            } else {
                for (int i = 0; i < m; ++i) c[i] = c[i] + alpha * A[IDX(k, i, lda)] * v;
            }
        }
    }
}

/*
 * C (m x n) := alpha*op(A)*op(B) + beta*C restricted to the stored pattern of C (CSC, 0-based):
 * the dense x dense -> sparse product (MatrixSwitch_mm_multiply.F90:334-398).  cval is
 * updated in place.
 */
void or_sddmm_csc(char opA, char opB, int k, double alpha, const double *A, int lda, const double *B,
                  int ldb, int n, const int *colptr, const int *rowidx, double beta, double *cval) {
    int trA = is_trans(opA), trB = is_trans(opB);
    int nnz = colptr[n];
    for (int l = 0; l < nnz; ++l) cval[l] = beta * cval[l];
#pragma omp parallel for schedule(static)
    for (int j = 0; j < n; ++j)
        for (int l = colptr[j]; l < colptr[j + 1]; ++l) {
            int i = rowidx[l];
            double c = cval[l];
            for (int kk = 0; kk < k; ++kk) {
                double a = trA ? A[IDX(kk, i, lda)] : A[IDX(i, kk, lda)];
                double b = trB ? B[IDX(j, kk, ldb)] : B[IDX(kk, j, ldb)];
                c = c + alpha * a * b;
            }
            cval[l] = c;
        }
}

/* C := alpha*op(A) + beta*C  (MatrixSwitch_m_add.F90:95-129) */
void or_add(char opA, int m, int n, double alpha, const double *A, int lda, double beta, double *C, int ldc) {
    int trA = is_trans(opA);
    scale_inplace(m, n, beta, C, ldc);
#pragma omp parallel for schedule(static)
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < m; ++i) {
            double a = trA ? A[IDX(j, i, lda)] : A[IDX(i, j, lda)];
            C[IDX(i, j, ldc)] = C[IDX(i, j, ldc)] + alpha * a;
        }
}

/* C := beta*C  (MatrixSwitch.F90:2496-2621, ref branch) */
void or_scale(int m, int n, double beta, double *C, int ldc) { scale_inplace(m, n, beta, C, ldc); }

/* C_ij := alpha (i /= j), beta (i == j) over 'a'll / 'u'pper / 'l'ower (MatrixSwitch_m_set.F90:19-75) */
void or_set(char seC, int m, int n, double alpha, double beta, double *C, int ldc) {
    int lu = (seC == 'l' || seC == 'L') ? 2 : ((seC == 'u' || seC == 'U') ? 1 : 0);
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < n; ++j) {
            if (lu == 1 && j < i) continue;
            if (lu == 2 && j > i) continue;
            C[IDX(i, j, ldc)] = (i != j) ? alpha : beta;
        }
}

/* tr(A), sequential over the diagonal (MatrixSwitch.F90:2137-2141) */
double or_trace(int n, const double *A, int lda) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s = s + A[IDX(i, i, lda)];
    return s;
}

/* sum_ij A_ij*B_ij, i outer and j inner (MatrixSwitch.F90:2346-2352) */
double or_mm_trace(int m, int n, const double *A, int lda, const double *B, int ldb) {
    double s = 0.0;
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < n; ++j) s = s + A[IDX(i, j, lda)] * B[IDX(i, j, ldb)];
    return s;
}

/* ------------------------------------------------------------------------------------------ */
/* libOMM pieces                                                                               */
/* ------------------------------------------------------------------------------------------ */

/* x <- (1103515245*x + 12345) mod 2^31, r = x/2^31  (libOMM/src/omm_rand.F90:37-55) */
double or_bsd_lcg(int32_t *x) {
    int64_t t = (int64_t)1103515245 * (int64_t)(*x) + 12345;
    int64_t m = (int64_t)1 << 31;
    int64_t r = t % m; /* Fortran mod: sign of the dividend */
    *x = (int32_t)r;
    return (double)(*x) / (double)m;
}

/* Minimum of the quartic sum c(i) x^i along the line (libOMM/src/omm_quartic.F90:60-153). */
void or_solve_quartic(const double *c, double *x_min, int *fail) {
    const double Pi = 3.141592653589793238462643383279502884197;
    double t[3], z[3];
    *fail = 0;
    double a = 3.0 * c[3] / (4.0 * c[4]);
    double b = 2.0 * c[2] / (4.0 * c[4]);
    if (fabs(b) >= 1.0e11 || fabs(c[4]) <= 1.0e-11) {
        *x_min = -0.5 * c[1] / c[2];
        return;
    }
    double d = c[1] / (4.0 * c[4]);
    double Q = (a * a - 3.0 * b) / 9.0;
    double R = (2.0 * a * a * a - 9.0 * a * b + 27.0 * d) / 54.0;
    double Q3 = Q * Q * Q;
    if (R * R < Q3) {
        double theta = acos(R / sqrt(Q3));
        double sq = sqrt(Q);
        t[0] = -2.0 * sq * cos(theta / 3.0) - a / 3.0;
        t[1] = -2.0 * sq * cos((theta + 2.0 * Pi) / 3.0) - a / 3.0;
        t[2] = -2.0 * sq * cos((theta - 2.0 * Pi) / 3.0) - a / 3.0;
        for (int i = 0; i < 3; ++i) {
            double ti = t[i], t2 = ti * ti;
            z[i] = c[4] * (t2 * t2) + c[3] * (t2 * ti) + c[2] * t2 + c[1] * ti + c[0];
        }
        if (c[4] > 0.0) {
            int o0, o1, o2;
            if (z[0] >= z[1] && z[0] >= z[2]) {
                o0 = 0; o1 = 1; o2 = 2;
            } else if (z[1] > z[2]) {
                o0 = 1; o1 = 2; o2 = 0;
            } else {
                o0 = 2; o1 = 0; o2 = 1;
            }
            if (0.0 <= t[o0] && t[o1] <= t[o0])
                *x_min = t[o1];
            else
                *x_min = t[o2];
        } else {
            if (z[0] <= z[1] && z[0] <= z[2])
                *x_min = t[0];
            else if (z[1] < z[2])
                *x_min = t[1];
            else
                *x_min = t[2];
        }
    } else {
        double S = -copysign(1.0, R) * pow(fabs(R) + sqrt(R * R - Q3), 1.0 / 3.0);
        double U = (S == 0.0) ? 0.0 : Q / S;
        *x_min = (S + U) - (a / 3.0);
        if (c[4] < 0.0) *fail = 1;
    }
}

/* Five line-search coefficients from 13 traces (libOMM/src/omm_ops.F90:156-219). */
void or_calc_coeff(int n, const double *HW, const double *SW, const double *HWd, const double *SWd,
                   const double *HWdd, const double *SWdd, double *work1, double *coeff) {
    double TrH = or_trace(n, HW, n);
    double TrHS = or_mm_trace(n, n, HW, n, SW, n);
    double TrHd = or_trace(n, HWd, n);
    double TrHdS = or_mm_trace(n, n, SW, n, HWd, n);
    double TrHSd = or_mm_trace(n, n, HW, n, SWd, n);
    double TrHdd = or_trace(n, HWdd, n);
    double TrHddS = or_mm_trace(n, n, HWdd, n, SW, n);
    double TrHSdd = or_mm_trace(n, n, HW, n, SWdd, n);
    or_add('c', n, n, 1.0, SWd, n, 0.0, work1, n);
    double TrHdSd = or_mm_trace(n, n, work1, n, HWd, n);
    double TrHdSdH = or_mm_trace(n, n, SWd, n, HWd, n);
    double TrHddSd = or_mm_trace(n, n, HWdd, n, SWd, n);
    double TrHdSdd = or_mm_trace(n, n, SWdd, n, HWd, n);
    double TrHddSdd = or_mm_trace(n, n, HWdd, n, SWdd, n);
    coeff[0] = 2.0 * TrH - TrHS;
    coeff[1] = 2.0 * (2.0 * TrHd - TrHdS - TrHSd);
    coeff[2] = 2.0 * (TrHdd - TrHdSd - TrHdSdH) - TrHddS - TrHSdd;
    coeff[3] = -2.0 * (TrHddSd + TrHdSdd);
    coeff[4] = -TrHddSdd;
}

/* ------------------------------------------------------------------------------------------ */
/* omm() flavour 0                                                                             */
/* ------------------------------------------------------------------------------------------ */

/* An operator matrix in the basis: dense (column-major m x m) or CSC (0-based). */
typedef struct {
    int kind; /* 0 = absent, 1 = dense, 2 = CSC */
    const double *dense;
    const int *colptr;
    const int *rowidx;
    const double *val;
} or_op;

/* Per-(spin,k)-point state kept between calls (libOMM/src/omm.F90:63,84,96-106). */
typedef struct {
    int allocated, first_call;
    int m, n;
    double x_min;
    double *HW, *HWd, *HWdd, *SW, *QW; /* n x n */
    double *SC, *SG, *work2, *CD;      /* n x m */
} or_ip_state;

typedef struct {
    int np;
    or_ip_state *st;
} or_omm_ctx;

or_omm_ctx *or_omm_create(int np) {
    or_omm_ctx *c = (or_omm_ctx *)calloc(1, sizeof(or_omm_ctx));
    c->np = np;
    c->st = (or_ip_state *)calloc((size_t)np, sizeof(or_ip_state));
    for (int i = 0; i < np; ++i) c->st[i].first_call = 1;
    return c;
}

static void free_state(or_ip_state *s) {
    free(s->HW); free(s->HWd); free(s->HWdd); free(s->SW); free(s->QW);
    free(s->SC); free(s->SG); free(s->work2); free(s->CD);
    memset(s, 0, sizeof(*s));
    s->first_call = 1;
}

void or_omm_destroy(or_omm_ctx *c) {
    if (!c) return;
    for (int i = 0; i < c->np; ++i) free_state(&c->st[i]);
    free(c->st);
    free(c);
}

static double *zalloc(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }

/* AC := C*A (operator product), the first half of calc_AW (omm_ops.F90:61). */
static void op_mul(const or_op *A, int n, int m, const double *C, double *AC) {
    if (A->kind == 1)
        or_gemm('n', 'n', n, m, m, 1.0, C, n, A->dense, m, 0.0, AC, n);
    else
        or_spmm_csc('n', n, m, 1.0, C, n, A->colptr, A->rowidx, A->val, 0.0, AC, n);
}

/* calc_AW: AC = C*A; AW = AC*C^T  (omm_ops.F90:44-64) */
static void calc_AW(const or_op *A, int n, int m, const double *C, double *AW, double *AC) {
    op_mul(A, n, m, C, AC);
    or_gemm('n', 'c', n, n, m, 1.0, AC, n, C, n, 0.0, AW, n);
}

/* calc_G: G = 4HC - 2 SW HC - 2 HW SC  (omm_ops.F90:17-39) */
static void calc_G(int n, int m, const double *HW, const double *SW, double *G, const double *HC,
                   const double *SC) {
    or_add('n', n, m, 4.0, HC, n, 0.0, G, n);
    or_gemm('n', 'n', n, m, n, -2.0, SW, n, HC, n, 1.0, G, n);
    or_gemm('n', 'n', n, m, n, -2.0, HW, n, SC, n, 1.0, G, n);
}

/*
 * omm(), flavour 0 (libOMM/src/omm.F90:5-672).  m = basis size, n = occupied states.
 * H, S: operators (S->kind == 0 means S absent, the no_S branch).  C_min: n x m in/out.
 * D_min: m x m out (may be NULL to skip the dense density matrix).  Extensions for tests:
 *   max_iter > 0 caps the total CG steps (icg), 0 keeps the reference caps;
 *   e_log (may be NULL, length log_cap) receives e_min after each CG step;
 *   info[0] = icg, info[1] = converged, info[2] = number of ls_fail rescalings,
 *   out[0] = TrQS (Tr[(2I-SW)SW]).
 * Returns 0, or a negative error for a flavour other than 0.
 */
int or_omm(or_omm_ctx *ctx, int m, int n, const or_op *H, const or_op *S, int new_S, double *e_min_out,
           double *D_min, int calc_ED, double eta, double *C_min, int init_C, int flavour, int ip,
           double cg_tol, int dealloc, int max_iter, double *e_log, int log_cap, int *info, double *out) {
    if (flavour != 0) return -1;
    if (ip < 1 || ip > ctx->np) return -2;
    or_ip_state *st = &ctx->st[ip - 1];
    int no_S = (S == NULL || S->kind == 0);
    double cg_tol_internal = cg_tol < 0.0 ? 1.0e-9 : cg_tol;
    size_t nn = (size_t)n * n, nm = (size_t)n * m;

    if (calc_ED) {
        if (!st->allocated || st->first_call) return -3; /* omm.F90:165 */
        double x = st->x_min;
        or_add('c', n, n, x, st->HWd, n, 1.0, st->HW, n);
        or_add('n', n, n, x, st->HWd, n, 1.0, st->HW, n);
        or_add('n', n, n, x * x, st->HWdd, n, 1.0, st->HW, n);
        if (eta != 0.0) or_add('n', n, n, eta, st->SW, n, 1.0, st->HW, n);
        /* calc_A2(HW, C_min, CD, D_min, work2): work2 = HW*CD; D_min = C^T*work2 (omm_ops.F90:129-150) */
        or_gemm('n', 'n', n, m, n, 1.0, st->HW, n, st->CD, n, 0.0, st->work2, n);
        if (D_min) or_gemm('c', 'n', m, m, n, 1.0, C_min, n, st->work2, n, 0.0, D_min, m);
        if (!dealloc && eta != 0.0) or_add('n', n, n, -eta, st->SW, n, 1.0, st->HW, n);
        if (dealloc)
            for (int i = 0; i < ctx->np; ++i) free_state(&ctx->st[i]);
        return 0;
    }

    if (!st->allocated) {
        st->m = m; st->n = n;
        st->HW = zalloc(nn); st->HWd = zalloc(nn); st->HWdd = zalloc(nn); st->SW = zalloc(nn);
        st->QW = zalloc(nn);
        st->SC = zalloc(nm); st->SG = zalloc(nm); st->work2 = zalloc(nm); st->CD = zalloc(nm);
        st->allocated = 1;
    }
    double *SWd = zalloc(nn), *SWdd = zalloc(nn), *work1 = zalloc(nn);
    double *G = zalloc(nm), *G_p = zalloc(nm), *D = zalloc(nm), *HC = zalloc(nm), *HG = zalloc(nm);
    double *HW = st->HW, *HWd = st->HWd, *HWdd = st->HWdd, *SW = st->SW, *SC = st->SC, *SG = st->SG;
    double coeff[5];
    int nfail = 0;

    /* random initial guess (omm.F90:329-343) */
    if (st->first_call && !init_C) {
        int32_t seed = 123456; /* NORAND seed (omm_rand.F90:27-28) */
        for (int i = 0; i < m; ++i)
            for (int j = 0; j < n; ++j) {
                double r1 = or_bsd_lcg(&seed), r2 = or_bsd_lcg(&seed);
                double el = copysign(0.5 * r1, r2 - 0.5);
                C_min[IDX(j, i, n)] = el;
            }
        or_scale(n, m, 1.0e-2 / sqrt((double)m), C_min, n);
    }

    /* initial energy and gradient (omm.F90:350-413) */
    calc_AW(H, n, m, C_min, HW, HC);
    if (no_S) {
        if (st->first_call || init_C) or_gemm('n', 'c', n, n, m, 1.0, C_min, n, C_min, n, 0.0, SW, n);
    } else {
        if (st->first_call || init_C || new_S)
            calc_AW(S, n, m, C_min, SW, SC);
        else
            or_add('n', n, m, st->x_min, SG, n, 1.0, SC, n);
    }
    calc_G(n, m, HW, SW, G, HC, no_S ? C_min : SC);
    or_gemm('n', 'c', n, n, m, 1.0, HC, n, G, n, 0.0, HWd, n);
    or_gemm('n', 'c', n, n, m, 1.0, no_S ? C_min : SC, n, G, n, 0.0, SWd, n);
    calc_AW(H, n, m, G, HWdd, HG);
    if (no_S)
        or_gemm('n', 'c', n, n, m, 1.0, G, n, G, n, 0.0, SWdd, n);
    else
        calc_AW(S, n, m, G, SWdd, SG);
    or_calc_coeff(n, HW, SW, HWd, SWd, HWdd, SWdd, work1, coeff);
    double e_min = coeff[0], e_min_old;

    /* CG loop (omm.F90:429-552) */
    int icg = 0, conv = 0, nlog = 0;
    long n_step_max_inner = (m < 10000) ? (long)m * n - 1 : 10000000L;
    int outer;
    int capped = 0;
    for (outer = 1; outer <= 100; ++outer) {
        double lambda = 0.0;
        for (long inner = 1; inner <= n_step_max_inner; ++inner) {
            or_add('n', n, m, 1.0, G, n, lambda, D, n);
            or_add('n', n, m, 1.0, G, n, 0.0, G_p, n);
            e_min_old = e_min;
            if (icg > 0) {
                or_gemm('n', 'c', n, n, m, 1.0, HC, n, D, n, 0.0, HWd, n);
                or_gemm('n', 'c', n, n, m, 1.0, no_S ? C_min : SC, n, D, n, 0.0, SWd, n);
                calc_AW(H, n, m, D, HWdd, HG);
                if (no_S)
                    or_gemm('n', 'c', n, n, m, 1.0, D, n, D, n, 0.0, SWdd, n);
                else
                    calc_AW(S, n, m, D, SWdd, SG);
                or_calc_coeff(n, HW, SW, HWd, SWd, HWdd, SWdd, work1, coeff);
            }
            int ls_fail, ls_conv;
            or_solve_quartic(coeff, &st->x_min, &ls_fail);
            double x = st->x_min;
            if (ls_fail) {
                e_min = 3.0 * e_min;
                or_scale(n, m, 0.5, C_min, n);
                ls_conv = 0;
                ++nfail;
            } else {
                double x2 = x * x;
                e_min = coeff[4] * (x2 * x2) + coeff[3] * (x2 * x) + coeff[2] * x2 + coeff[1] * x + coeff[0];
                or_add('n', n, m, x, D, n, 1.0, C_min, n);
                ls_conv = 1;
            }
            if (ls_fail) {
                or_scale(n, n, 0.25, SW, n);
            } else {
                or_add('c', n, n, x, SWd, n, 1.0, SW, n);
                or_add('n', n, n, x, SWd, n, 1.0, SW, n);
                or_add('n', n, n, x * x, SWdd, n, 1.0, SW, n);
            }
            double e_diff = 2.0 * fabs((e_min - e_min_old) / (e_min + e_min_old));
            if (e_log && nlog < log_cap) e_log[nlog++] = e_min;
            ++icg;
            if (e_diff <= cg_tol_internal) {
                conv = 1;
                break;
            }
            if (max_iter > 0 && icg >= max_iter) {
                capped = 1;
                break;
            }
            if (ls_fail) {
                or_scale(n, n, 0.25, HW, n);
            } else {
                or_add('c', n, n, x, HWd, n, 1.0, HW, n);
                or_add('n', n, n, x, HWd, n, 1.0, HW, n);
                or_add('n', n, n, x * x, HWdd, n, 1.0, HW, n);
            }
            if (ls_fail) {
                or_scale(n, m, 0.5, HC, n);
                if (!no_S) or_scale(n, m, 0.5, SC, n);
                or_add('n', n, m, 1.0, G_p, n, 0.0, G, n);
                or_add('n', n, m, 1.5, HC, n, 1.0, G, n);
            } else {
                or_add('n', n, m, x, HG, n, 1.0, HC, n);
                if (no_S) {
                    calc_G(n, m, HW, SW, G, HC, C_min);
                } else {
                    or_add('n', n, m, x, SG, n, 1.0, SC, n);
                    calc_G(n, m, HW, SW, G, HC, SC);
                }
            }
            if (ls_conv) {
                or_add('n', n, m, 1.0, G, n, 0.0, st->work2, n);
                or_add('n', n, m, -1.0, G_p, n, 1.0, st->work2, n);
                double lambda_n = or_mm_trace(n, m, G, n, st->work2, n);
                double lambda_d = or_mm_trace(n, m, G_p, n, G_p, n);
                lambda = lambda_n / lambda_d;
            } else {
                break;
            }
        }
        if (conv || capped) break;
    }

    /* density matrix (omm.F90:575-593) */
    or_set('a', n, n, 0.0, 2.0, st->QW, n);
    or_add('n', n, n, -1.0, SW, n, 1.0, st->QW, n);
    /* calc_A(QW, C_min, D_min, CD): CD = QW*C; D_min = C^T*CD (omm_ops.F90:104-124) */
    or_gemm('n', 'n', n, m, n, 1.0, st->QW, n, C_min, n, 0.0, st->CD, n);
    if (D_min) or_gemm('c', 'n', m, m, n, 1.0, C_min, n, st->CD, n, 0.0, D_min, m);
    double TrQS = or_mm_trace(n, n, st->QW, n, SW, n);
    e_min = e_min + TrQS * eta;

    free(SWd); free(SWdd); free(work1); free(G); free(G_p); free(D); free(HC); free(HG);
    if (dealloc) {
        for (int i = 0; i < ctx->np; ++i) free_state(&ctx->st[i]);
    } else {
        st->first_call = 0;
    }
    *e_min_out = e_min;
    if (info) {
        info[0] = icg;
        info[1] = conv;
        info[2] = nfail;
        info[3] = nlog;
    }
    if (out) out[0] = TrQS;
    return 0;
}

/* Convenience entry for ctypes: separate fields instead of the or_op struct. */
int or_omm_flat(or_omm_ctx *ctx, int m, int n, int h_kind, const double *h_dense, const int *h_colptr,
                const int *h_rowidx, const double *h_val, int s_kind, const double *s_dense, const int *s_colptr,
                const int *s_rowidx, const double *s_val, int new_S, double *e_min, double *D_min, int calc_ED,
                double eta, double *C_min, int init_C, int flavour, int ip, double cg_tol, int dealloc,
                int max_iter, double *e_log, int log_cap, int *info, double *out) {
    or_op H = {h_kind, h_dense, h_colptr, h_rowidx, h_val};
    or_op S = {s_kind, s_dense, s_colptr, s_rowidx, s_val};
    return or_omm(ctx, m, n, &H, &S, new_S, e_min, D_min, calc_ED, eta, C_min, init_C, flavour, ip, cg_tol,
                  dealloc, max_iter, e_log, log_cap, info, out);
}

int or_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
