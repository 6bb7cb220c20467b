/*
 * CPU oracle for the OMM conjugate-gradient hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * This file restates, in plain C, the algorithm of libOMM's omm() (basic flavour 0, with or
 * without S) and the MatrixSwitch 'ref' operations it drives, on real ('sdden'/'sdcsc') and on
 * complex ('szden'/'szcsc') storage: one CG driver runs on either operation table.  Only tests,
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
 * Parity pins (tests/test_oracle.py): the reference serial build in oracle/_ref (same inputs,
 * bit-level comparison; real dense and CSC, complex dense on example_kpoints and a Bloch
 * lattice), example_gamma / example_kpoints / example_sdden known answers, closed forms
 * (generalised eigenproblem, real and Hermitian; quartic brute force, finite-difference gradient,
 * polynomial identity of the line-search coefficients, real and complex), the complex table on
 * real data reproducing the real one bit for bit, the complex CSC product equal to the dense one,
 * and the LCG recurrence.
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
/* MatrixSwitch 'ref' operations on complex storage ('szden', 'szcsc')                         */
/* ------------------------------------------------------------------------------------------ */
/*
 * Complex elements are stored as interleaved (re, im) pairs (Fortran complex(dp) layout);
 * leading dimensions count complex elements.  Real scalars reach these routines as complex
 * ones with a zero imaginary part (MatrixSwitch.F90 converts alpha/beta with cmplx(x, 0) under
 * CONV), and every product is the textbook one, (a+ib)(c+id) = (ac - bd) + i(ad + bc), each
 * product rounded before the sum, so the rounding sequence is that of the reference build
 * compiled without FMA contraction.
 */
typedef struct {
    double re, im;
} or_z;

static inline or_z zmul(or_z a, or_z b) {
    or_z r;
    r.re = a.re * b.re - a.im * b.im;
    r.im = a.re * b.im + a.im * b.re;
    return r;
}
static inline or_z zadd(or_z a, or_z b) {
    or_z r = {a.re + b.re, a.im + b.im};
    return r;
}
static inline or_z zget(const double *A, size_t i) {
    or_z r = {A[2 * i], A[2 * i + 1]};
    return r;
}
static inline void zput(double *A, size_t i, or_z v) {
    A[2 * i] = v.re;
    A[2 * i + 1] = v.im;
}
static inline or_z zconj(or_z a) {
    or_z r = {a.re, -a.im};
    return r;
}
/* op(A)(i, k) for op 'n', 't' or 'c' */
static inline or_z zop(char op, const double *A, int lda, int i, int k) {
    if (op == 'n' || op == 'N') return zget(A, IDX(i, k, lda));
    or_z a = zget(A, IDX(k, i, lda));
    return (op == 'c' || op == 'C') ? zconj(a) : a;
}

/* C := beta*C  (C%zval = beta*C%zval with beta = cmplx(beta, 0)) */
void or_zscale(int m, int n, double beta, double *C, int ldc) {
    const or_z b = {beta, 0.0};
#pragma omp parallel for schedule(static)
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < m; ++i) zput(C, IDX(i, j, ldc), zmul(b, zget(C, IDX(i, j, ldc))));
}

/*
 * C := alpha*op(A)*op(B) + beta*C, op in {'n','t','c'} (mm_multiply_szdenref,
 * MatrixSwitch_mm_multiply.F90:85-185): C := beta*C, then per element
 * C(i,j) = C(i,j) + (alpha*op(A)(i,k))*op(B)(k,j) for k ascending.  alpha*op(A) is formed once
 * into a packed copy (the same values the reference forms per term).
 */
void or_zgemm(char opA, char opB, int m, int n, int k, double alpha, const double *A, int lda, const double *B,
              int ldb, double beta, double *C, int ldc) {
    or_zscale(m, n, beta, C, ldc);
    if (m <= 0 || n <= 0 || k <= 0) return;
    const or_z al = {alpha, 0.0};
    double *ap = (double *)malloc(sizeof(double) * 2 * (size_t)m * k);
    double *bp = (double *)malloc(sizeof(double) * 2 * (size_t)k * n);
#pragma omp parallel for schedule(static)
    for (int kk = 0; kk < k; ++kk)
        for (int i = 0; i < m; ++i) zput(ap, IDX(i, kk, m), zmul(al, zop(opA, A, lda, i, kk)));
#pragma omp parallel for schedule(static)
    for (int j = 0; j < n; ++j)
        for (int kk = 0; kk < k; ++kk) zput(bp, IDX(kk, j, k), zop(opB, B, ldb, kk, j));
#pragma omp parallel for schedule(static)
    for (int j = 0; j < n; ++j)
        for (int kk = 0; kk < k; ++kk) {
            const or_z b = zget(bp, IDX(kk, j, k));
            for (int i = 0; i < m; ++i)
                zput(C, IDX(i, j, ldc), zadd(zget(C, IDX(i, j, ldc)), zmul(zget(ap, IDX(i, kk, m)), b)));
        }
    free(ap);
    free(bp);
}

/*
 * C := alpha*op(A)*B + beta*C with B a complex CSC matrix (0-based), op(A) in {'n','t','c'}
 * (mm_multiply_szdenszcscszdenref, MatrixSwitch_mm_multiply.F90:734-778): per column j, per
 * stored entry l ascending, C(i,j) = C(i,j) + (alpha*op(A)(i,row(l)))*val(l).
 */
void or_zspmm_csc(char opA, int m, int nb, double alpha, const double *A, int lda, const int *colptr,
                  const int *rowidx, const double *val, double beta, double *C, int ldc) {
    or_zscale(m, nb, beta, C, ldc);
    const or_z al = {alpha, 0.0};
#pragma omp parallel for schedule(static)
    for (int j = 0; j < nb; ++j)
        for (int l = colptr[j]; l < colptr[j + 1]; ++l) {
            const int k = rowidx[l];
            const or_z v = zget(val, (size_t)l);
            for (int i = 0; i < m; ++i)
                zput(C, IDX(i, j, ldc), zadd(zget(C, IDX(i, j, ldc)), zmul(zmul(al, zop(opA, A, lda, i, k)), v)));
        }
}

/* C := beta*C, then C := C + alpha*op(A)  (m_add_szdenref, MatrixSwitch_m_add.F90:135-175) */
void or_zadd(char opA, int m, int n, double alpha, const double *A, int lda, double beta, double *C, int ldc) {
    or_zscale(m, n, beta, C, ldc);
    const or_z al = {alpha, 0.0};
#pragma omp parallel for schedule(static)
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < m; ++i)
            zput(C, IDX(i, j, ldc), zadd(zget(C, IDX(i, j, ldc)), zmul(al, zop(opA, A, lda, i, j))));
}

/* off-diagonal := alpha, diagonal := beta (m_set_szdenref, MatrixSwitch_m_set.F90:81-137) */
void or_zset(char seC, int m, int n, double alpha, double beta, double *C, int ldc) {
    int lu = (seC == 'u' || seC == 'U') ? 1 : ((seC == 'l' || seC == 'L') ? 2 : 0);
    const or_z a = {alpha, 0.0}, b = {beta, 0.0};
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < n; ++j) {
            if (lu == 1 && j < i) continue;
            if (lu == 2 && j > i) continue;
            zput(C, IDX(i, j, ldc), (i != j) ? a : b);
        }
}

/* tr(A) = sum of the diagonal, sequential (m_ztrace, MatrixSwitch.F90:2161-2248); out[0..1] = re, im.
 * Returns the real part, which is what the real-valued callers receive (m_dtrace under CONV). */
double or_ztrace(int n, const double *A, int lda, double *out) {
    or_z s = {0.0, 0.0};
    for (int i = 0; i < n; ++i) s = zadd(s, zget(A, IDX(i, i, lda)));
    if (out) {
        out[0] = s.re;
        out[1] = s.im;
    }
    return s.re;
}

/* sum_ij conjg(A_ij)*B_ij, i outer and j inner (mm_ztrace, MatrixSwitch.F90:2382-2491);
 * out[0..1] = re, im; returns the real part (mm_dtrace under CONV). */
double or_zmm_trace(int m, int n, const double *A, int lda, const double *B, int ldb, double *out) {
    or_z s = {0.0, 0.0};
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < n; ++j) s = zadd(s, zmul(zconj(zget(A, IDX(i, j, lda))), zget(B, IDX(i, j, ldb))));
    if (out) {
        out[0] = s.re;
        out[1] = s.im;
    }
    return s.re;
}

/* ------------------------------------------------------------------------------------------ */
/* The operation table the omm() driver runs on: real ('sdden'/'sdcsc') or complex storage.    */
/* ------------------------------------------------------------------------------------------ */

typedef struct {
    int cw; /* doubles per element: 1 real, 2 complex */
    void (*gemm)(char, char, int, int, int, double, const double *, int, const double *, int, double, double *, int);
    void (*spmm)(char, int, int, double, const double *, int, const int *, const int *, const double *, double,
                 double *, int);
    void (*add)(char, int, int, double, const double *, int, double, double *, int);
    void (*scale)(int, int, double, double *, int);
    void (*set)(char, int, int, double, double, double *, int);
    double (*trace)(int, const double *, int);
    double (*mm_trace)(int, int, const double *, int, const double *, int);
} or_ops;

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
static void calc_coeff_x(const or_ops *X, int n, const double *HW, const double *SW, const double *HWd,
                         const double *SWd, const double *HWdd, const double *SWdd, double *work1, double *coeff) {
    double TrH = X->trace(n, HW, n);
    double TrHS = X->mm_trace(n, n, HW, n, SW, n);
    double TrHd = X->trace(n, HWd, n);
    double TrHdS = X->mm_trace(n, n, SW, n, HWd, n);
    double TrHSd = X->mm_trace(n, n, HW, n, SWd, n);
    double TrHdd = X->trace(n, HWdd, n);
    double TrHddS = X->mm_trace(n, n, HWdd, n, SW, n);
    double TrHSdd = X->mm_trace(n, n, HW, n, SWdd, n);
    X->add('c', n, n, 1.0, SWd, n, 0.0, work1, n);
    double TrHdSd = X->mm_trace(n, n, work1, n, HWd, n);
    double TrHdSdH = X->mm_trace(n, n, SWd, n, HWd, n);
    double TrHddSd = X->mm_trace(n, n, HWdd, n, SWd, n);
    double TrHdSdd = X->mm_trace(n, n, SWdd, n, HWd, n);
    double TrHddSdd = X->mm_trace(n, n, HWdd, n, SWdd, n);
    coeff[0] = 2.0 * TrH - TrHS;
    coeff[1] = 2.0 * (2.0 * TrHd - TrHdS - TrHSd);
    coeff[2] = 2.0 * (TrHdd - TrHdSd - TrHdSdH) - TrHddS - TrHSdd;
    coeff[3] = -2.0 * (TrHddSd + TrHdSdd);
    coeff[4] = -TrHddSdd;
}

static double ztrace_re(int n, const double *A, int lda) { return or_ztrace(n, A, lda, NULL); }
static double zmm_trace_re(int m, int n, const double *A, int lda, const double *B, int ldb) {
    return or_zmm_trace(m, n, A, lda, B, ldb, NULL);
}

static const or_ops OPS_REAL = {1, or_gemm, or_spmm_csc, or_add, or_scale, or_set, or_trace, or_mm_trace};
static const or_ops OPS_COMPLEX = {2, or_zgemm, or_zspmm_csc, or_zadd, or_zscale, or_zset, ztrace_re, zmm_trace_re};

void or_calc_coeff(int n, const double *HW, const double *SW, const double *HWd, const double *SWd,
                   const double *HWdd, const double *SWdd, double *work1, double *coeff) {
    calc_coeff_x(&OPS_REAL, n, HW, SW, HWd, SWd, HWdd, SWdd, work1, coeff);
}

/* The same coefficients from complex WF-basis matrices: real parts of the complex traces (calc_coeff
 * declares them real(dp), so m_dtrace/mm_dtrace take real(...) of m_ztrace/mm_ztrace under CONV). */
void or_zcalc_coeff(int n, const double *HW, const double *SW, const double *HWd, const double *SWd,
                    const double *HWdd, const double *SWdd, double *work1, double *coeff) {
    calc_coeff_x(&OPS_COMPLEX, n, HW, SW, HWd, SWd, HWdd, SWdd, work1, coeff);
}

/* ------------------------------------------------------------------------------------------ */
/* omm() flavour 0                                                                             */
/* ------------------------------------------------------------------------------------------ */

/* An operator matrix in the basis: dense (column-major m x m) or CSC (0-based), real or complex
 * (interleaved) as the driver's operation table says. */
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
    int m, n, cw;
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
static void op_mul(const or_ops *X, const or_op *A, int n, int m, const double *C, double *AC) {
    if (A->kind == 1)
        X->gemm('n', 'n', n, m, m, 1.0, C, n, A->dense, m, 0.0, AC, n);
    else
        X->spmm('n', n, m, 1.0, C, n, A->colptr, A->rowidx, A->val, 0.0, AC, n);
}

/* calc_AW: AC = C*A; AW = AC*C^H  (omm_ops.F90:44-64) */
static void calc_AW(const or_ops *X, const or_op *A, int n, int m, const double *C, double *AW, double *AC) {
    op_mul(X, A, n, m, C, AC);
    X->gemm('n', 'c', n, n, m, 1.0, AC, n, C, n, 0.0, AW, n);
}

/* calc_G: G = 4HC - 2 SW HC - 2 HW SC  (omm_ops.F90:17-39) */
static void calc_G(const or_ops *X, int n, int m, const double *HW, const double *SW, double *G, const double *HC,
                   const double *SC) {
    X->add('n', n, m, 4.0, HC, n, 0.0, G, n);
    X->gemm('n', 'n', n, m, n, -2.0, SW, n, HC, n, 1.0, G, n);
    X->gemm('n', 'n', n, m, n, -2.0, HW, n, SC, n, 1.0, G, n);
}

/*
 * omm(), flavour 0 (libOMM/src/omm.F90:5-672), on the operation table X (real or complex storage).
 * m = basis size, n = occupied states.
 * H, S: operators (S->kind == 0 means S absent, the no_S branch).  C_min: n x m in/out.
 * D_min: m x m out (may be NULL to skip the dense density matrix).  Extensions for tests:
 *   max_iter > 0 caps the total CG steps (icg), 0 keeps the reference caps;
 *   e_log (may be NULL, length log_cap) receives e_min after each CG step;
 *   info[0] = icg, info[1] = converged, info[2] = number of ls_fail rescalings,
 *   out[0] = TrQS (Tr[(2I-SW)SW]).
 * Returns 0, or a negative error for a flavour other than 0.
 */
static int omm_run(const or_ops *X, or_omm_ctx *ctx, int m, int n, const or_op *H, const or_op *S, int new_S,
                   double *e_min_out, double *D_min, int calc_ED, double eta, double *C_min, int init_C, int flavour,
                   int ip, double cg_tol, int dealloc, int max_iter, double *e_log, int log_cap, int *info,
                   double *out) {
    if (flavour != 0) return -1;
    if (ip < 1 || ip > ctx->np) return -2;
    or_ip_state *st = &ctx->st[ip - 1];
    if (st->allocated && st->cw != X->cw) return -4; /* one ip keeps one storage type */
    int no_S = (S == NULL || S->kind == 0);
    double cg_tol_internal = cg_tol < 0.0 ? 1.0e-9 : cg_tol;
    const size_t nn = (size_t)X->cw * n * n, nm = (size_t)X->cw * n * m;

    if (calc_ED) {
        if (!st->allocated || st->first_call) return -3; /* omm.F90:165 */
        double x = st->x_min;
        X->add('c', n, n, x, st->HWd, n, 1.0, st->HW, n);
        X->add('n', n, n, x, st->HWd, n, 1.0, st->HW, n);
        X->add('n', n, n, x * x, st->HWdd, n, 1.0, st->HW, n);
        if (eta != 0.0) X->add('n', n, n, eta, st->SW, n, 1.0, st->HW, n);
        /* calc_A2(HW, C_min, CD, D_min, work2): work2 = HW*CD; D_min = C^T*work2 (omm_ops.F90:129-150) */
        X->gemm('n', 'n', n, m, n, 1.0, st->HW, n, st->CD, n, 0.0, st->work2, n);
        if (D_min) X->gemm('c', 'n', m, m, n, 1.0, C_min, n, st->work2, n, 0.0, D_min, m);
        if (!dealloc && eta != 0.0) X->add('n', n, n, -eta, st->SW, n, 1.0, st->HW, n);
        if (dealloc)
            for (int i = 0; i < ctx->np; ++i) free_state(&ctx->st[i]);
        return 0;
    }

    if (!st->allocated) {
        st->m = m; st->n = n; st->cw = X->cw;
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
                /* m_set_element(C_min, j, i, el, 0): el + 0*C_min(j, i) with C_min zero-allocated */
                C_min[X->cw * IDX(j, i, n)] = el;
                if (X->cw == 2) C_min[2 * IDX(j, i, n) + 1] = 0.0;
            }
        X->scale(n, m, 1.0e-2 / sqrt((double)m), C_min, n);
    }

    /* initial energy and gradient (omm.F90:350-413) */
    calc_AW(X, H, n, m, C_min, HW, HC);
    if (no_S) {
        if (st->first_call || init_C) X->gemm('n', 'c', n, n, m, 1.0, C_min, n, C_min, n, 0.0, SW, n);
    } else {
        if (st->first_call || init_C || new_S)
            calc_AW(X, S, n, m, C_min, SW, SC);
        else
            X->add('n', n, m, st->x_min, SG, n, 1.0, SC, n);
    }
    calc_G(X, n, m, HW, SW, G, HC, no_S ? C_min : SC);
    X->gemm('n', 'c', n, n, m, 1.0, HC, n, G, n, 0.0, HWd, n);
    X->gemm('n', 'c', n, n, m, 1.0, no_S ? C_min : SC, n, G, n, 0.0, SWd, n);
    calc_AW(X, H, n, m, G, HWdd, HG);
    if (no_S)
        X->gemm('n', 'c', n, n, m, 1.0, G, n, G, n, 0.0, SWdd, n);
    else
        calc_AW(X, S, n, m, G, SWdd, SG);
    calc_coeff_x(X, n, HW, SW, HWd, SWd, HWdd, SWdd, work1, coeff);
    double e_min = coeff[0], e_min_old;

    /* CG loop (omm.F90:429-552) */
    int icg = 0, conv = 0, nlog = 0;
    long n_step_max_inner = (m < 10000) ? (long)m * n - 1 : 10000000L;
    int outer;
    int capped = 0;
    for (outer = 1; outer <= 100; ++outer) {
        double lambda = 0.0;
        for (long inner = 1; inner <= n_step_max_inner; ++inner) {
            X->add('n', n, m, 1.0, G, n, lambda, D, n);
            X->add('n', n, m, 1.0, G, n, 0.0, G_p, n);
            e_min_old = e_min;
            if (icg > 0) {
                X->gemm('n', 'c', n, n, m, 1.0, HC, n, D, n, 0.0, HWd, n);
                X->gemm('n', 'c', n, n, m, 1.0, no_S ? C_min : SC, n, D, n, 0.0, SWd, n);
                calc_AW(X, H, n, m, D, HWdd, HG);
                if (no_S)
                    X->gemm('n', 'c', n, n, m, 1.0, D, n, D, n, 0.0, SWdd, n);
                else
                    calc_AW(X, S, n, m, D, SWdd, SG);
                calc_coeff_x(X, n, HW, SW, HWd, SWd, HWdd, SWdd, work1, coeff);
            }
            int ls_fail, ls_conv;
            or_solve_quartic(coeff, &st->x_min, &ls_fail);
            double x = st->x_min;
            if (ls_fail) {
                e_min = 3.0 * e_min;
                X->scale(n, m, 0.5, C_min, n);
                ls_conv = 0;
                ++nfail;
            } else {
                double x2 = x * x;
                e_min = coeff[4] * (x2 * x2) + coeff[3] * (x2 * x) + coeff[2] * x2 + coeff[1] * x + coeff[0];
                X->add('n', n, m, x, D, n, 1.0, C_min, n);
                ls_conv = 1;
            }
            if (ls_fail) {
                X->scale(n, n, 0.25, SW, n);
            } else {
                X->add('c', n, n, x, SWd, n, 1.0, SW, n);
                X->add('n', n, n, x, SWd, n, 1.0, SW, n);
                X->add('n', n, n, x * x, SWdd, n, 1.0, SW, n);
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
                X->scale(n, n, 0.25, HW, n);
            } else {
                X->add('c', n, n, x, HWd, n, 1.0, HW, n);
                X->add('n', n, n, x, HWd, n, 1.0, HW, n);
                X->add('n', n, n, x * x, HWdd, n, 1.0, HW, n);
            }
            if (ls_fail) {
                X->scale(n, m, 0.5, HC, n);
                if (!no_S) X->scale(n, m, 0.5, SC, n);
                X->add('n', n, m, 1.0, G_p, n, 0.0, G, n);
                X->add('n', n, m, 1.5, HC, n, 1.0, G, n);
            } else {
                X->add('n', n, m, x, HG, n, 1.0, HC, n);
                if (no_S) {
                    calc_G(X, n, m, HW, SW, G, HC, C_min);
                } else {
                    X->add('n', n, m, x, SG, n, 1.0, SC, n);
                    calc_G(X, n, m, HW, SW, G, HC, SC);
                }
            }
            if (ls_conv) {
                X->add('n', n, m, 1.0, G, n, 0.0, st->work2, n);
                X->add('n', n, m, -1.0, G_p, n, 1.0, st->work2, n);
                double lambda_n = X->mm_trace(n, m, G, n, st->work2, n);
                double lambda_d = X->mm_trace(n, m, G_p, n, G_p, n);
                lambda = lambda_n / lambda_d;
            } else {
                break;
            }
        }
        if (conv || capped) break;
    }

    /* density matrix (omm.F90:575-593) */
    X->set('a', n, n, 0.0, 2.0, st->QW, n);
    X->add('n', n, n, -1.0, SW, n, 1.0, st->QW, n);
    /* calc_A(QW, C_min, D_min, CD): CD = QW*C; D_min = C^T*CD (omm_ops.F90:104-124) */
    X->gemm('n', 'n', n, m, n, 1.0, st->QW, n, C_min, n, 0.0, st->CD, n);
    if (D_min) X->gemm('c', 'n', m, m, n, 1.0, C_min, n, st->CD, n, 0.0, D_min, m);
    double TrQS = X->mm_trace(n, n, st->QW, n, SW, n);
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

int or_omm(or_omm_ctx *ctx, int m, int n, const or_op *H, const or_op *S, int new_S, double *e_min_out,
           double *D_min, int calc_ED, double eta, double *C_min, int init_C, int flavour, int ip,
           double cg_tol, int dealloc, int max_iter, double *e_log, int log_cap, int *info, double *out) {
    return omm_run(&OPS_REAL, ctx, m, n, H, S, new_S, e_min_out, D_min, calc_ED, eta, C_min, init_C, flavour, ip,
                   cg_tol, dealloc, max_iter, e_log, log_cap, info, out);
}

/* Convenience entry for ctypes: separate fields instead of the or_op struct.  is_complex selects
 * the complex ('szden'/'szcsc') operation table; every matrix argument is then complex128. */
int or_omm_flat(or_omm_ctx *ctx, int is_complex, int m, int n, int h_kind, const double *h_dense, const int *h_colptr,
                const int *h_rowidx, const double *h_val, int s_kind, const double *s_dense, const int *s_colptr,
                const int *s_rowidx, const double *s_val, int new_S, double *e_min, double *D_min, int calc_ED,
                double eta, double *C_min, int init_C, int flavour, int ip, double cg_tol, int dealloc,
                int max_iter, double *e_log, int log_cap, int *info, double *out) {
    or_op H = {h_kind, h_dense, h_colptr, h_rowidx, h_val};
    or_op S = {s_kind, s_dense, s_colptr, s_rowidx, s_val};
    return omm_run(is_complex ? &OPS_COMPLEX : &OPS_REAL, ctx, m, n, &H, &S, new_S, e_min, D_min, calc_ED, eta,
                   C_min, init_C, flavour, ip, cg_tol, dealloc, max_iter, e_log, log_cap, info, out);
}

int or_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
