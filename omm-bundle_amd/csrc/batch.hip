// Batched k-point solver: many complex128 systems of one shape (the k-points of one cell, which
// share H's block pattern) advanced by omm()'s flavour-0 conjugate-gradient loop in lockstep, with
// every decision taken on the device, so that one CG step of all systems is a fixed sequence of
// launches.  That sequence is captured once into a hipGraph and replayed; the host only reads the
// per-system convergence flags every few steps.
//
// Per step (all systems; libOMM/src/omm.F90:436-552 for one system):
//   D = G + lambda*D, G_p = G                         batched element-wise (lambda per system)
//   HG = D*H_s, SG = D*S_s                            batched complex BSR SpMM
//   HWd = HC*D^H, SWd = SC*D^H, HWdd = HG*D^H,       batched complex MFMA GEMMs
//   SWdd = SG*D^H
//   coefficients                                      batched fused reduction, on the device
//   quartic, energy, e_diff, convergence, ls_fail     one thread per system (k_batch_quartic)
//   C, SW updates; then (not converged) HW, HC, SC    per-system element-wise with x_min
//   G = 4HC - 2SW*HC - 2HW*SC                         batched pair GEMM (systems still running)
//   Polak-Ribiere lambda                              batched dots, divided on the device
// The first step of a CG run recomputes the line-search products from D = G, which is what the
// reference's initial pass computes (omm.F90:350-413).  A converged system is frozen: its
// `active` flag masks every later launch, and its final state is the reference's (the HW, HC and
// G updates after the convergence exit are skipped, omm.F90:506-509).
//
// Uneven-size group (omx_group_*): real systems of different sizes, each with its own H and S,
// advanced by the same device-side step.  A system's step is its own sequence of library launches
// (SpMM with its shape, GEMMs with its tuned tiles, fused reductions) masked by its flags.  The
// systems whose products fill the chip on their own run one after another on the library stream;
// the others go to separate streams forked from it before that chain and joined after it, so the
// captured graph holds the chain plus one parallel branch per small system, and a single graph
// launch advances all of them.
#include <algorithm>
#include <array>
#include <cmath>
#include <vector>

#include "omx_internal.h"

struct omx_zbatch_s {
    omx_ctx ctx = nullptr;
    int nsys = 0, no = 0, nb = 0, nbc = 0;
    int64_t nnzb = 0;
    int *colptr = nullptr, *rowidx = nullptr;
    double *hval = nullptr, *sval = nullptr;  // nsys * nnzb * 32
    int dt = OMX_C128;  // element type of the state: OMX_C128 or OMX_C64
    // state, system s at s * (no*nb) or s * (no*no) complex elements, ld = no
    void *C = nullptr, *HC = nullptr, *SC = nullptr, *G = nullptr, *Gp = nullptr, *D = nullptr, *HG = nullptr,
         *SG = nullptr;
    void *HW = nullptr, *SW = nullptr, *HWd = nullptr, *SWd = nullptr, *HWdd = nullptr, *SWdd = nullptr;
    // per-system scalars
    double *coeff = nullptr;  // 18 per system
    double *dots = nullptr;   // 2 per system
    double *sc = nullptr;     // [x, lambda, e_min, e_diff] per system, then the solve's [cg_tol, max_steps]
    int *active = nullptr;    // per-system flags, F_COUNT arrays of nsys (see flag())
    double *partial = nullptr, *partial2 = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int graph_steps = 0;  // CG steps captured in exec
};

namespace omx {

constexpr int BT_NT = 256;

// The quartic line search of omm_solve_quartic (libOMM/src/omm_quartic.F90:60-153), the energy at
// the minimum and the convergence test, for one system.
__device__ inline void quartic_min(const double *c, double &x_min, bool &fail) {
    const double Pi = 3.141592653589793238462643383279502884197;
    fail = false;
    const double a = 3.0 * c[3] / (4.0 * c[4]);
    const double b = 2.0 * c[2] / (4.0 * c[4]);
    if (fabs(b) >= 1.0e11 || fabs(c[4]) <= 1.0e-11) {
        x_min = -0.5 * c[1] / c[2];
        return;
    }
    const double d = c[1] / (4.0 * c[4]);
    const double Q = (a * a - 3.0 * b) / 9.0;
    const double R = (2.0 * a * a * a - 9.0 * a * b + 27.0 * d) / 54.0;
    if (R * R < Q * Q * Q) {
        const double theta = acos(R / sqrt(Q * Q * Q));
        double t[3], z[3];
        t[0] = -2.0 * sqrt(Q) * cos(theta / 3.0) - a / 3.0;
        t[1] = -2.0 * sqrt(Q) * cos((theta + 2.0 * Pi) / 3.0) - a / 3.0;
        t[2] = -2.0 * sqrt(Q) * cos((theta - 2.0 * Pi) / 3.0) - a / 3.0;
        for (int q = 0; q < 3; ++q) {
            const double x = t[q];
            z[q] = c[4] * x * x * x * x + c[3] * x * x * x + c[2] * x * x + c[1] * x + c[0];
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
            x_min = (0.0 <= t[o0] && t[o1] <= t[o0]) ? t[o1] : t[o2];
        } else {
            if (z[0] <= z[1] && z[0] <= z[2])
                x_min = t[0];
            else if (z[1] < z[2])
                x_min = t[1];
            else
                x_min = t[2];
        }
    } else {
        const double sr = (R > 0.0 || (R == 0.0 && !signbit(R))) ? 1.0 : -1.0;
        const double S = -sr * cbrt(fabs(R) + sqrt(R * R - Q * Q * Q));
        const double U = (S == 0.0) ? 0.0 : Q / S;
        x_min = (S + U) - (a / 3.0);
        if (c[4] < 0.0) fail = true;
    }
}

// Per system and step, written by k_batch_quartic:
//   failf  the line search failed (coefficients rescaled, omm.F90:470-477)
//   cont   stepped, not converged and below the step cap: HW, HC, SC and G are updated
//          (omm.F90:506-534)
//   grad   cont and not failf: the gradient is recomputed and lambda follows (omm.F90:536-546)
// sc per system: [0] x_min, [1] lambda, [2] e_min, [3] e_diff.  params = [cg_tol, max_steps] of the
// current solve, in device memory so a captured step reads the values of the solve that replays it.
// A system that reaches max_steps stops like a converged one (before the updates) but keeps
// conv = 0; the steps that follow in the same graph are no-ops for it.
__global__ void k_batch_quartic(int nsys, const double *__restrict__ coeff, double *__restrict__ sc,
                                const int *__restrict__ active, int *__restrict__ failf, int *__restrict__ cont,
                                int *__restrict__ grad, int *__restrict__ icg, int *__restrict__ conv,
                                int *__restrict__ nfail, const double *__restrict__ params) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsys) return;
    if (!active[s]) {
        failf[s] = 0;
        cont[s] = 0;
        grad[s] = 0;
        return;
    }
    const double *c = coeff + 18 * s;
    double *v = sc + 4 * s;
    const double e_old = (icg[s] == 0) ? c[0] : v[2];
    double x;
    bool fail;
    quartic_min(c, x, fail);
    double e;
    if (fail) {
        e = 3.0 * e_old;
        nfail[s] += 1;
    } else {
        e = c[4] * x * x * x * x + c[3] * x * x * x + c[2] * x * x + c[1] * x + c[0];
    }
    const double ed = 2.0 * fabs((e - e_old) / (e + e_old));
    v[0] = x;
    v[2] = e;
    v[3] = ed;
    icg[s] += 1;
    const bool done = (ed <= params[0]);
    const bool capped = icg[s] >= (int)params[1];
    if (done) conv[s] = 1;
    failf[s] = fail ? 1 : 0;
    cont[s] = (done || capped) ? 0 : 1;
    grad[s] = (!done && !capped && !fail) ? 1 : 0;
}

// The element-wise kernels below are written over El<T> (double for real systems, double2 for
// complex128); system s = blockIdx.y (blockIdx.z for update_w) of a batch with contiguous
// per-system blocks of len elements.

// D := G + lambda*D; G_p := G (active systems)
template <typename T>
__global__ __launch_bounds__(BT_NT) void k_batch_direction(int64_t len, const T *__restrict__ G, T *__restrict__ D,
                                                           T *__restrict__ Gp, const double *__restrict__ sc,
                                                           const int *__restrict__ active) {
    typedef El<T> E;
    const int s = blockIdx.y;
    if (!active[s]) return;
    const double lam = sc[4 * s + 1];
    const int64_t o = (int64_t)s * len;
    for (int64_t q = (int64_t)blockIdx.x * BT_NT + threadIdx.x; q < len; q += (int64_t)gridDim.x * BT_NT) {
        const T g = G[o + q];
        T d;
        if (lam == 0.0) {
            d = g;
        } else {
            const T dd = D[o + q];
            d = E::make(lam * E::re(dd) + E::re(g), E::cplx ? lam * E::im(dd) + E::im(g) : 0.0);
        }
        D[o + q] = d;
        Gp[o + q] = g;
    }
}

// For systems with mask[s]: Y := Y + x*X, or Y := fail_scale*Y after a failed line search.
template <typename T>
__global__ __launch_bounds__(BT_NT) void k_batch_axpy(int64_t len, const T *__restrict__ X, T *__restrict__ Y,
                                                      const double *__restrict__ sc, const int *__restrict__ mask,
                                                      const int *__restrict__ failf, double fail_scale) {
    typedef El<T> E;
    const int s = blockIdx.y;
    if (!mask[s]) return;
    const bool failed = failf[s] != 0;
    const double x = sc[4 * s];
    const int64_t o = (int64_t)s * len;
    for (int64_t q = (int64_t)blockIdx.x * BT_NT + threadIdx.x; q < len; q += (int64_t)gridDim.x * BT_NT) {
        const T y = Y[o + q];
        if (failed) {
            Y[o + q] = E::make(E::re(y) * fail_scale, E::im(y) * fail_scale);
        } else {
            const T xx = X[o + q];
            Y[o + q] = E::make(E::re(y) + x * E::re(xx), E::cplx ? E::im(y) + x * E::im(xx) : 0.0);
        }
    }
}

// For systems with mask[s]: W := W + x*Wd^H + x*Wd + x^2*Wdd (the reference's order), or
// W := 0.25*W after a failed line search.
template <typename T>
__global__ __launch_bounds__(BT_NT) void k_batch_update_w(int n, const T *__restrict__ Wd, const T *__restrict__ Wdd,
                                                          T *__restrict__ W, const double *__restrict__ sc,
                                                          const int *__restrict__ mask, const int *__restrict__ failf) {
    typedef El<T> E;
    constexpr int TS = 32;
    __shared__ double tr[TS][TS + 1];
    __shared__ double ti[E::cplx ? TS : 1][TS + 1];
    const int s = blockIdx.z;
    if (!mask[s]) return;  // whole workgroup, before the barrier
    const int64_t o = (int64_t)s * n * n;
    const bool failed = failf[s] != 0;
    const double x = sc[4 * s], x2 = x * x;
    const int tx = threadIdx.x % TS, ty = threadIdx.x / TS;
    const int i0 = blockIdx.x * TS, j0 = blockIdx.y * TS;
    for (int b = ty; b < TS; b += BT_NT / TS) {
        const int r = j0 + tx, c = i0 + b;
        const T w = (r < n && c < n) ? Wd[o + r + (int64_t)c * n] : E::make(0.0, 0.0);
        tr[tx][b] = E::re(w);
        if constexpr (E::cplx) ti[tx][b] = -E::im(w);  // conjugate transpose
    }
    __syncthreads();
    for (int c = ty; c < TS; c += BT_NT / TS) {
        const int i = i0 + tx, j = j0 + c;
        if (i >= n || j >= n) continue;
        const int64_t q = o + i + (int64_t)j * n;
        const T w = W[q];
        if (failed) {
            W[q] = E::make(E::re(w) * 0.25, E::im(w) * 0.25);
        } else {
            const T d = Wd[q], dd = Wdd[q];
            double wr = E::re(w);
            wr = wr + x * tr[c][tx];
            wr = wr + x * E::re(d);
            wr = wr + x2 * E::re(dd);
            double wi = 0.0;
            if constexpr (E::cplx) {
                wi = E::im(w);
                wi = wi + x * ti[c][tx];
                wi = wi + x * E::im(d);
                wi = wi + x2 * E::im(dd);
            }
            W[q] = E::make(wr, wi);
        }
    }
}

// After a rescaling (omm.F90:519-523): G_p := G; G := 1.5*HC + G (HC already halved).
template <typename T>
__global__ __launch_bounds__(BT_NT) void k_batch_fail_grad(int64_t len, const T *__restrict__ HC, T *__restrict__ G,
                                                           T *__restrict__ Gp, const int *__restrict__ cont,
                                                           const int *__restrict__ failf) {
    typedef El<T> E;
    const int s = blockIdx.y;
    if (!cont[s] || !failf[s]) return;
    const int64_t o = (int64_t)s * len;
    for (int64_t q = (int64_t)blockIdx.x * BT_NT + threadIdx.x; q < len; q += (int64_t)gridDim.x * BT_NT) {
        const T g = G[o + q], h = HC[o + q];
        Gp[o + q] = g;
        G[o + q] = E::make(1.5 * E::re(h) + E::re(g), E::cplx ? 1.5 * E::im(h) + E::im(g) : 0.0);
    }
}

// lambda := <G, G - G_p> / <G_p, G_p> where the gradient was recomputed, 0 after a rescaling
// (the restart of the inner loop); the next step's active set is the running set.
__global__ void k_batch_lambda(int nsys, const double *__restrict__ dots, double *__restrict__ sc,
                               const int *__restrict__ grad, const int *__restrict__ failf,
                               const int *__restrict__ cont, int *__restrict__ active) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsys) return;
    if (grad[s])
        sc[4 * s + 1] = dots[2 * s] / dots[2 * s + 1];
    else if (failf[s])
        sc[4 * s + 1] = 0.0;
    active[s] = cont[s];
}

// Start of a solve: every system active, counters cleared, and the solve's parameters
// [cg_tol, max_steps] stored after the per-system scalars (params = sc + 4*nsys).
__global__ void k_batch_init_flags(int nsys, int *__restrict__ active, int *__restrict__ icg, int *__restrict__ conv,
                                   int *__restrict__ nfail, double *__restrict__ sc, double tol, int max_steps) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s == 0) {
        sc[4 * nsys] = tol;
        sc[4 * nsys + 1] = (double)max_steps;
    }
    if (s >= nsys) return;
    active[s] = 1;
    icg[s] = 0;
    conv[s] = 0;
    nfail[s] = 0;
    sc[4 * s + 0] = 0.0;
    sc[4 * s + 1] = 0.0;
    sc[4 * s + 2] = 0.0;
    sc[4 * s + 3] = 0.0;
}

// G := 4*HC (systems with mask[s]); the first pass of calc_G when the pair launch does not apply.
template <typename T>
__global__ __launch_bounds__(BT_NT) void k_batch_scale_copy(int64_t len, const T *__restrict__ X, T *__restrict__ Y,
                                                            double a, const int *__restrict__ mask) {
    typedef El<T> E;
    const int s = blockIdx.y;
    if (!mask[s]) return;
    const int64_t o = (int64_t)s * len;
    for (int64_t q = (int64_t)blockIdx.x * BT_NT + threadIdx.x; q < len; q += (int64_t)gridDim.x * BT_NT) {
        const T x = X[o + q];
        Y[o + q] = E::make(a * E::re(x), a * E::im(x));
    }
}

static inline dim3 bt_grid(int64_t len, int nsys) {
    return dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>((len + BT_NT - 1) / BT_NT, 4096)), nsys);
}

}  // namespace omx

using namespace omx;

// per-system flag arrays inside one allocation
static int *flag(omx_zbatch b, int which) { return b->active + which * b->nsys; }
enum { F_ACTIVE = 0, F_FAIL, F_CONT, F_GRAD, F_ICG, F_CONV, F_NFAIL, F_COUNT };

// G = 4HC - 2SW*HC - 2HW*SC for the systems with mask[s] (calc_G, omm_ops.F90:17-39): one pair
// launch, or three passes when no is not a multiple of the K slice.  T: double2 or Cx<float>.
template <typename T>
static int batch_calc_G(omx_zbatch b, const int *mask) {
    omx_ctx ctx = b->ctx;
    const int ns = b->nsys, no = b->no, nb = b->nb;
    const int64_t lenC = (int64_t)no * nb, lenW = (int64_t)no * no;
    int rc = launch_zgemm_pair_batched(ctx, b->dt, no, nb, no, -2.0, b->SW, b->HW, no, lenW, b->HC, b->SC, no, lenC,
                                       4.0, b->HC, 0.0, b->G, no, lenC, ns, mask);
    if (rc != OMX_ERR_UNSUPPORTED) return rc;
    hipLaunchKernelGGL(k_batch_scale_copy<T>, bt_grid(lenC, ns), dim3(BT_NT), 0, ctx->stream, lenC, (const T *)b->HC,
                       (T *)b->G, 4.0, mask);
    if ((rc = launch_zgemm_batched(ctx, b->dt, 'n', 'n', no, nb, no, -2.0, b->SW, no, lenW, b->HC, no, lenC, 1.0, b->G,
                                   no, lenC, ns, mask)))
        return rc;
    return launch_zgemm_batched(ctx, b->dt, 'n', 'n', no, nb, no, -2.0, b->HW, no, lenW, b->SC, no, lenC, 1.0, b->G,
                                no, lenC, ns, mask);
}

// One CG step of every system (the sequence that is captured).
template <typename T>
static int batch_step(omx_zbatch b) {
    omx_ctx ctx = b->ctx;
    const int ns = b->nsys, no = b->no, nb = b->nb, dt = b->dt;
    const int64_t lenC = (int64_t)no * nb, lenW = (int64_t)no * no, sV = 32 * b->nnzb;
    int *act = flag(b, F_ACTIVE), *failf = flag(b, F_FAIL), *cont = flag(b, F_CONT), *grad = flag(b, F_GRAD);
    int rc;
    hipLaunchKernelGGL(k_batch_direction<T>, bt_grid(lenC, ns), dim3(BT_NT), 0, ctx->stream, lenC, (const T *)b->G,
                       (T *)b->D, (T *)b->Gp, b->sc, act);
    if ((rc = launch_spmm_z_batched(ctx, dt, no, nb, b->nbc, b->colptr, b->rowidx, b->hval, sV, b->D, no, lenC, b->HG,
                                    no, lenC, ns, act)))
        return rc;
    if ((rc = launch_spmm_z_batched(ctx, dt, no, nb, b->nbc, b->colptr, b->rowidx, b->sval, sV, b->D, no, lenC, b->SG,
                                    no, lenC, ns, act)))
        return rc;
    const void *src[4] = {b->HC, b->SC, b->HG, b->SG};
    void *dst[4] = {b->HWd, b->SWd, b->HWdd, b->SWdd};
    // HG D^H = D H D^H and SG D^H are Hermitian: lower-triangle tiles and a mirror
    for (int q = 0; q < 4; ++q)
        if ((rc = q < 2 ? launch_zgemm_batched(ctx, dt, 'n', 'c', no, no, nb, 1.0, src[q], no, lenC, b->D, no, lenC,
                                               0.0, dst[q], no, lenW, ns, act)
                        : launch_zgemm_herm_batched(ctx, dt, no, nb, 1.0, src[q], no, lenC, b->D, no, lenC, dst[q],
                                                    no, lenW, ns, act)))
            return rc;
    if ((rc = launch_coeff_batched(ctx, dt, no, no, lenW, b->HW, b->SW, b->HWd, b->SWd, b->HWdd, b->SWdd, ns,
                                   b->partial, b->coeff, nullptr)))
        return rc;
    hipLaunchKernelGGL(k_batch_quartic, dim3((ns + 63) / 64), dim3(64), 0, ctx->stream, ns, b->coeff, b->sc, act,
                       failf, cont, grad, flag(b, F_ICG), flag(b, F_CONV), flag(b, F_NFAIL), b->sc + 4 * ns);
    // C and SW for every system that stepped (omm.F90:470-502)
    hipLaunchKernelGGL(k_batch_axpy<T>, bt_grid(lenC, ns), dim3(BT_NT), 0, ctx->stream, lenC, (const T *)b->D,
                       (T *)b->C, b->sc, act, failf, 0.5);
    const dim3 wgrid((no + 31) / 32, (no + 31) / 32, ns);
    hipLaunchKernelGGL(k_batch_update_w<T>, wgrid, dim3(BT_NT), 0, ctx->stream, no, (const T *)b->SWd,
                       (const T *)b->SWdd, (T *)b->SW, b->sc, act, failf);
    // HW, HC, SC for the systems still running (omm.F90:511-527)
    hipLaunchKernelGGL(k_batch_update_w<T>, wgrid, dim3(BT_NT), 0, ctx->stream, no, (const T *)b->HWd,
                       (const T *)b->HWdd, (T *)b->HW, b->sc, cont, failf);
    hipLaunchKernelGGL(k_batch_axpy<T>, bt_grid(lenC, ns), dim3(BT_NT), 0, ctx->stream, lenC, (const T *)b->HG,
                       (T *)b->HC, b->sc, cont, failf, 0.5);
    hipLaunchKernelGGL(k_batch_axpy<T>, bt_grid(lenC, ns), dim3(BT_NT), 0, ctx->stream, lenC, (const T *)b->SG,
                       (T *)b->SC, b->sc, cont, failf, 0.5);
    hipLaunchKernelGGL(k_batch_fail_grad<T>, bt_grid(lenC, ns), dim3(BT_NT), 0, ctx->stream, lenC, (const T *)b->HC,
                       (T *)b->G, (T *)b->Gp, cont, failf);
    // G = 4HC - 2SW*HC - 2HW*SC where the line search succeeded, then Polak-Ribiere
    if ((rc = batch_calc_G<T>(b, grad))) return rc;
    if ((rc = launch_pr_batched(ctx, dt, no, nb, no, lenC, b->G, b->Gp, ns, b->partial2, b->dots, nullptr)))
        return rc;
    hipLaunchKernelGGL(k_batch_lambda, dim3((ns + 63) / 64), dim3(64), 0, ctx->stream, ns, b->dots, b->sc, grad, failf,
                       cont, act);
    OMX_CHECK_HIP(hipGetLastError());
    return OMX_OK;
}

static int batch_step_any(omx_zbatch b) {
    return b->dt == OMX_C64 ? batch_step<Cx<float>>(b) : batch_step<double2>(b);
}
static int batch_calc_G_any(omx_zbatch b, const int *mask) {
    return b->dt == OMX_C64 ? batch_calc_G<Cx<float>>(b, mask) : batch_calc_G<double2>(b, mask);
}
static size_t batch_esize(omx_zbatch b) { return b->dt == OMX_C64 ? 8 : 16; }

static void batch_free(omx_zbatch b) {
    if (!b) return;
    if (b->ctx) (void)hipStreamSynchronize(b->ctx->stream);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    void *ptrs[] = {b->colptr, b->rowidx, b->hval, b->sval, b->C, b->HC, b->SC, b->G, b->Gp, b->D, b->HG, b->SG,
                    b->HW, b->SW, b->HWd, b->SWd, b->HWdd, b->SWdd, b->coeff, b->dots, b->sc, b->active,
                    b->partial, b->partial2};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    delete b;
}

extern "C" {

int omx_zbatch_create_dt(omx_ctx ctx, int dt, int nsys, int no, int nb, const int *bcolptr, const int *browidx,
                         const double *hval, const double *sval, omx_zbatch *out) {
    OMX_REQUIRE(ctx && out && nsys > 0 && no > 0 && nb > 0 && bcolptr && browidx && hval && sval,
                "omx_zbatch_create: bad arguments");
    OMX_REQUIRE(dt == OMX_C128 || dt == OMX_C64, "omx_zbatch_create: element type %d is not complex", dt);
    OMX_REQUIRE(nsys < 65536, "omx_zbatch_create: too many systems");
    const int nbc = (nb + 3) / 4;
    OMX_REQUIRE(bcolptr[0] == 0, "omx_zbatch_create: bcolptr must be 0-based");
    for (int j = 0; j < nbc; ++j) {
        OMX_REQUIRE(bcolptr[j + 1] >= bcolptr[j], "omx_zbatch_create: bcolptr not monotone");
        for (int l = bcolptr[j]; l < bcolptr[j + 1]; ++l)
            OMX_REQUIRE(browidx[l] >= 0 && browidx[l] < nbc && (l == bcolptr[j] || browidx[l] > browidx[l - 1]),
                        "omx_zbatch_create: bad block rows in column %d", j);
    }
    omx_zbatch b = new omx_zbatch_s();
    b->ctx = ctx;
    b->dt = dt;
    b->nsys = nsys;
    b->no = no;
    b->nb = nb;
    b->nbc = nbc;
    b->nnzb = bcolptr[nbc];
    const size_t lenC = (size_t)no * nb * nsys, lenW = (size_t)no * no * nsys;
    bool ok = hipMalloc(&b->colptr, sizeof(int) * (nbc + 1)) == hipSuccess &&
              hipMalloc(&b->rowidx, sizeof(int) * std::max<int64_t>(b->nnzb, 1)) == hipSuccess &&
              hipMalloc(&b->hval, sizeof(double) * 32 * std::max<int64_t>(b->nnzb, 1) * nsys) == hipSuccess &&
              hipMalloc(&b->sval, sizeof(double) * 32 * std::max<int64_t>(b->nnzb, 1) * nsys) == hipSuccess;
    const size_t es = batch_esize(b);
    void **nm[] = {&b->C, &b->HC, &b->SC, &b->G, &b->Gp, &b->D, &b->HG, &b->SG};
    for (auto p : nm) ok = ok && hipMalloc(p, es * lenC) == hipSuccess;
    void **nn[] = {&b->HW, &b->SW, &b->HWd, &b->SWd, &b->HWdd, &b->SWdd};
    for (auto p : nn) ok = ok && hipMalloc(p, es * lenW) == hipSuccess;
    ok = ok && hipMalloc(&b->coeff, sizeof(double) * 18 * nsys) == hipSuccess &&
         hipMalloc(&b->dots, sizeof(double) * 2 * nsys) == hipSuccess &&
         hipMalloc(&b->sc, sizeof(double) * (4 * nsys + 2)) == hipSuccess &&
         hipMalloc(&b->active, sizeof(int) * F_COUNT * nsys) == hipSuccess &&
         hipMalloc(&b->partial, sizeof(double) * coeff_partial_doubles(no, nsys)) == hipSuccess &&
         hipMalloc(&b->partial2, sizeof(double) * pr_partial_doubles(no, nb, nsys)) == hipSuccess;
    if (!ok) {
        batch_free(b);
        set_error("omx_zbatch_create: device allocation failed");
        return OMX_ERR_NOMEM;
    }
    hipStream_t st = ctx->stream;
    (void)hipMemcpyAsync(b->colptr, bcolptr, sizeof(int) * (nbc + 1), hipMemcpyHostToDevice, st);
    if (b->nnzb) {
        (void)hipMemcpyAsync(b->rowidx, browidx, sizeof(int) * b->nnzb, hipMemcpyHostToDevice, st);
        (void)hipMemcpyAsync(b->hval, hval, sizeof(double) * 32 * b->nnzb * nsys, hipMemcpyHostToDevice, st);
        (void)hipMemcpyAsync(b->sval, sval, sizeof(double) * 32 * b->nnzb * nsys, hipMemcpyHostToDevice, st);
    }
    OMX_CHECK_HIP(hipStreamSynchronize(st));
    *out = b;
    return OMX_OK;
}

int omx_zbatch_create(omx_ctx ctx, int nsys, int no, int nb, const int *bcolptr, const int *browidx,
                      const double *hval, const double *sval, omx_zbatch *out) {
    return omx_zbatch_create_dt(ctx, OMX_C128, nsys, no, nb, bcolptr, browidx, hval, sval, out);
}

int omx_zbatch_destroy(omx_zbatch b) {
    batch_free(b);
    return OMX_OK;
}

int omx_zbatch_set_c(omx_zbatch b, const void *C) {
    OMX_REQUIRE(b && C, "omx_zbatch_set_c: bad arguments");
    OMX_CHECK_HIP(hipMemcpyAsync(b->C, C, batch_esize(b) * (size_t)b->no * b->nb * b->nsys, hipMemcpyHostToDevice,
                                 b->ctx->stream));
    OMX_CHECK_HIP(hipStreamSynchronize(b->ctx->stream));
    return OMX_OK;
}

int omx_zbatch_get_c(omx_zbatch b, void *C) {
    OMX_REQUIRE(b && C, "omx_zbatch_get_c: bad arguments");
    OMX_CHECK_HIP(hipMemcpyAsync(C, b->C, batch_esize(b) * (size_t)b->no * b->nb * b->nsys, hipMemcpyDeviceToHost,
                                 b->ctx->stream));
    OMX_CHECK_HIP(hipStreamSynchronize(b->ctx->stream));
    return OMX_OK;
}

int omx_zbatch_solve(omx_zbatch b, double eta, double cg_tol, int max_steps, int check_every, int use_graph,
                     double *e_min, int *icg, int *converged, int *steps_run) {
    OMX_REQUIRE(b && e_min && icg && converged && max_steps > 0 && check_every > 0, "omx_zbatch_solve: bad arguments");
    omx_ctx ctx = b->ctx;
    hipStream_t st = ctx->stream;
    const int ns = b->nsys, no = b->no, nb = b->nb;
    const int64_t lenC = (int64_t)no * nb, lenW = (int64_t)no * no, sV = 32 * b->nnzb;
    const double tol = cg_tol < 0.0 ? 1.0e-9 : cg_tol;
    int rc;
    // initial state (omm.F90:350-370): HC = C H, SC = C S, HW = HC C^H, SW = SC C^H, G
    hipLaunchKernelGGL(k_batch_init_flags, dim3((ns + 63) / 64), dim3(64), 0, st, ns, flag(b, F_ACTIVE),
                       flag(b, F_ICG), flag(b, F_CONV), flag(b, F_NFAIL), b->sc, tol, max_steps);
    int *act = flag(b, F_ACTIVE);
    if ((rc = launch_spmm_z_batched(ctx, b->dt, no, nb, b->nbc, b->colptr, b->rowidx, b->hval, sV, b->C, no, lenC, b->HC, no,
                                    lenC, ns, act)))
        return rc;
    if ((rc = launch_spmm_z_batched(ctx, b->dt, no, nb, b->nbc, b->colptr, b->rowidx, b->sval, sV, b->C, no, lenC, b->SC, no,
                                    lenC, ns, act)))
        return rc;
    if ((rc = launch_zgemm_herm_batched(ctx, b->dt, no, nb, 1.0, b->HC, no, lenC, b->C, no, lenC, b->HW, no, lenW, ns,
                                        act)))
        return rc;
    if ((rc = launch_zgemm_herm_batched(ctx, b->dt, no, nb, 1.0, b->SC, no, lenC, b->C, no, lenC, b->SW, no, lenW, ns,
                                        act)))
        return rc;
    if ((rc = batch_calc_G_any(b, act))) return rc;
    // the CG steps: check_every steps captured as one graph (kept while check_every is unchanged),
    // one graph launch per flag read; the step cap and cg_tol are read from device memory, so
    // steps past a system's convergence or cap do nothing for it
    if (use_graph && (!b->exec || b->graph_steps != check_every)) {
        if (b->exec) (void)hipGraphExecDestroy(b->exec);
        if (b->graph) (void)hipGraphDestroy(b->graph);
        b->exec = nullptr;
        b->graph = nullptr;
        b->graph_steps = 0;
        OMX_CHECK_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        rc = OMX_OK;
        for (int q = 0; q < check_every && !rc; ++q) rc = batch_step_any(b);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(st, &g);
        if (rc || e != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            if (rc) return rc;
            OMX_CHECK_HIP(e);
        }
        b->graph = g;
        OMX_CHECK_HIP(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
        b->graph_steps = check_every;
    }
    std::vector<int> h_act(ns);
    int done_steps = 0;
    while (done_steps < max_steps) {
        if (use_graph) {
            OMX_CHECK_HIP(hipGraphLaunch(b->exec, st));
            done_steps += check_every;
        } else {
            const int burst = std::min(check_every, max_steps - done_steps);
            for (int q = 0; q < burst; ++q)
                if ((rc = batch_step_any(b))) return rc;
            done_steps += burst;
        }
        OMX_CHECK_HIP(hipMemcpyAsync(h_act.data(), act, sizeof(int) * ns, hipMemcpyDeviceToHost, st));
        OMX_CHECK_HIP(hipStreamSynchronize(st));
        bool any = false;
        for (int s = 0; s < ns; ++s) any = any || h_act[s];
        if (!any) break;
    }
    if (steps_run) *steps_run = done_steps;  // step sequences issued (masked ones included)
    // results: e_min + eta*TrQS with TrQS = Re tr[(2I - SW) SW] = 2 Re tr(SW) - <SW, SW>
    std::vector<double> sc(4 * ns);
    std::vector<int> h_icg(ns), h_conv(ns);
    OMX_CHECK_HIP(hipMemcpyAsync(sc.data(), b->sc, sizeof(double) * 4 * ns, hipMemcpyDeviceToHost, st));
    OMX_CHECK_HIP(hipMemcpyAsync(h_icg.data(), flag(b, F_ICG), sizeof(int) * ns, hipMemcpyDeviceToHost, st));
    OMX_CHECK_HIP(hipMemcpyAsync(h_conv.data(), flag(b, F_CONV), sizeof(int) * ns, hipMemcpyDeviceToHost, st));
    OMX_CHECK_HIP(hipStreamSynchronize(st));
    for (int s = 0; s < ns; ++s) {
        double tr[2], dd[2];
        const void *W = (const char *)b->SW + batch_esize(b) * (size_t)s * lenW;
        if ((rc = omx_trace(ctx, b->dt, no, W, no, tr))) return rc;
        if ((rc = omx_mm_trace(ctx, b->dt, no, no, W, no, W, no, dd))) return rc;
        const double TrQS = 2.0 * tr[0] - dd[0];
        e_min[s] = sc[4 * s + 2] + eta * TrQS;
        icg[s] = h_icg[s];
        converged[s] = h_conv[s];
    }
    return OMX_OK;
}

}  // extern "C"

// ---- uneven-size group of real systems -------------------------------------------------------

struct omx_group_s {
    omx_ctx ctx = nullptr;
    int nsys = 0;
    std::vector<int> no, nb;
    std::vector<omx_bsr> H, S;
    // per system: C, HC, SC, G, Gp, D, HG, SG (no x nb), HW, SW, HWd, SWd, HWdd, SWdd (no x no), ld = no
    std::vector<std::array<double *, 14>> m;
    std::vector<double *> partial, partial2;
    double *coeff = nullptr, *dots = nullptr, *sc = nullptr;
    int *flags = nullptr;  // F_COUNT arrays of nsys
    std::vector<hipStream_t> streams;
    std::vector<hipEvent_t> joins;
    std::vector<char> chain;  // system runs in the serial chain on the library stream (fills the chip)
    hipEvent_t fork = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int graph_steps = 0;  // CG steps captured in exec
};

enum { M_C = 0, M_HC, M_SC, M_G, M_GP, M_D, M_HG, M_SG, M_HW, M_SW, M_HWD, M_SWD, M_HWDD, M_SWDD };

static int *gflag(omx_group g, int which, int s) { return g->flags + which * g->nsys + s; }

// G = 4HC - 2SW*HC - 2HW*SC for system s when *mask (calc_G, omm_ops.F90:17-39).
static int group_calc_G(omx_group g, int s, const int *mask) {
    omx_ctx ctx = g->ctx;
    const int no = g->no[s], nb = g->nb[s];
    auto &M = g->m[s];
    ctx->launch_mask = mask;
    int rc = launch_gemm_pair(ctx, OMX_PREC_F64, OMX_F64, no, nb, no, -2.0, M[M_SW], M[M_HW], no, M[M_HC], M[M_SC],
                              no, 4.0, M[M_HC], 0.0, M[M_G], no);
    if (rc == OMX_ERR_UNSUPPORTED) {
        const int64_t lenC = (int64_t)no * nb;
        hipLaunchKernelGGL(k_batch_scale_copy<double>, bt_grid(lenC, 1), dim3(BT_NT), 0, ctx->stream, lenC, M[M_HC],
                           M[M_G], 4.0, mask);
        rc = launch_gemm(ctx, OMX_PREC_F64, OMX_F64, false, false, no, nb, no, -2.0, M[M_SW], no, M[M_HC], no, 1.0,
                         M[M_G], no);
        if (!rc)
            rc = launch_gemm(ctx, OMX_PREC_F64, OMX_F64, false, false, no, nb, no, -2.0, M[M_HW], no, M[M_SC], no,
                             1.0, M[M_G], no);
    }
    ctx->launch_mask = nullptr;
    return rc;
}

// Line-search products X*Y^T (no x no) of system s, masked; sym: the product is symmetric (X = Y*A
// with A = H or S, calc_AW, omm_ops.F90:44-64), computed on its lower-triangle tiles and mirrored,
// as omm()'s symmetric Grams.
static int group_gram(omx_group g, int s, const double *X, const double *Y, double *W, bool sym = false) {
    const int no = g->no[s], nb = g->nb[s];
    if (sym) return launch_gemm_sym(g->ctx, OMX_PREC_F64, OMX_F64, no, nb, 1.0, X, no, Y, no, W, no);
    return launch_gemm(g->ctx, OMX_PREC_F64, OMX_F64, false, true, no, no, nb, 1.0, X, no, Y, no, 0.0, W, no);
}

// One CG step of system s on the current library stream (the batch step, per system).
static int group_sys_step(omx_group g, int s) {
    omx_ctx ctx = g->ctx;
    const int no = g->no[s], nb = g->nb[s];
    const int64_t lenC = (int64_t)no * nb;
    auto &M = g->m[s];
    int *act = gflag(g, F_ACTIVE, s), *failf = gflag(g, F_FAIL, s), *cont = gflag(g, F_CONT, s),
        *grad = gflag(g, F_GRAD, s);
    double *scs = g->sc + 4 * s;
    int rc;
    hipLaunchKernelGGL(k_batch_direction<double>, bt_grid(lenC, 1), dim3(BT_NT), 0, ctx->stream, lenC, M[M_G],
                       M[M_D], M[M_GP], scs, act);
    ctx->launch_mask = act;
    rc = omx_spmm(ctx, OMX_F64, no, nb, 1.0, M[M_D], no, g->H[s], 0.0, M[M_HG], no);
    if (!rc) rc = omx_spmm(ctx, OMX_F64, no, nb, 1.0, M[M_D], no, g->S[s], 0.0, M[M_SG], no);
    if (!rc) rc = group_gram(g, s, M[M_HC], M[M_D], M[M_HWD]);
    if (!rc) rc = group_gram(g, s, M[M_SC], M[M_D], M[M_SWD]);
    if (!rc) rc = group_gram(g, s, M[M_HG], M[M_D], M[M_HWDD], true);
    if (!rc) rc = group_gram(g, s, M[M_SG], M[M_D], M[M_SWDD], true);
    if (!rc)
        rc = launch_coeff_batched(ctx, OMX_F64, no, no, 0, M[M_HW], M[M_SW], M[M_HWD], M[M_SWD], M[M_HWDD],
                                  M[M_SWDD], 1, g->partial[s], g->coeff + 18 * s, act);
    ctx->launch_mask = nullptr;
    if (rc) return rc;
    hipLaunchKernelGGL(k_batch_quartic, dim3(1), dim3(64), 0, ctx->stream, 1, g->coeff + 18 * s, scs, act, failf,
                       cont, grad, gflag(g, F_ICG, s), gflag(g, F_CONV, s), gflag(g, F_NFAIL, s),
                       g->sc + 4 * g->nsys);
    // C and SW when the system stepped; HW, HC, SC when it is still running (omm.F90:470-527)
    hipLaunchKernelGGL(k_batch_axpy<double>, bt_grid(lenC, 1), dim3(BT_NT), 0, ctx->stream, lenC, M[M_D], M[M_C],
                       scs, act, failf, 0.5);
    const dim3 wgrid((no + 31) / 32, (no + 31) / 32, 1);
    hipLaunchKernelGGL(k_batch_update_w<double>, wgrid, dim3(BT_NT), 0, ctx->stream, no, M[M_SWD], M[M_SWDD], M[M_SW],
                       scs, act, failf);
    hipLaunchKernelGGL(k_batch_update_w<double>, wgrid, dim3(BT_NT), 0, ctx->stream, no, M[M_HWD], M[M_HWDD], M[M_HW],
                       scs, cont, failf);
    hipLaunchKernelGGL(k_batch_axpy<double>, bt_grid(lenC, 1), dim3(BT_NT), 0, ctx->stream, lenC, M[M_HG], M[M_HC],
                       scs, cont, failf, 0.5);
    hipLaunchKernelGGL(k_batch_axpy<double>, bt_grid(lenC, 1), dim3(BT_NT), 0, ctx->stream, lenC, M[M_SG], M[M_SC],
                       scs, cont, failf, 0.5);
    hipLaunchKernelGGL(k_batch_fail_grad<double>, bt_grid(lenC, 1), dim3(BT_NT), 0, ctx->stream, lenC, M[M_HC],
                       M[M_G], M[M_GP], cont, failf);
    if ((rc = group_calc_G(g, s, grad))) return rc;
    if ((rc = launch_pr_batched(ctx, OMX_F64, no, nb, no, 0, M[M_G], M[M_GP], 1, g->partial2[s], g->dots + 2 * s,
                                grad)))
        return rc;
    hipLaunchKernelGGL(k_batch_lambda, dim3(1), dim3(64), 0, ctx->stream, 1, g->dots + 2 * s, scs, grad, failf, cont,
                       act);
    OMX_CHECK_HIP(hipGetLastError());
    return OMX_OK;
}

// Runs f(s) for every system: the systems whose products fill the chip on their own one after
// another on the library stream (concurrent, they would only contend for the same CUs and L2), the
// others each on its own stream, forked from the library stream before the chain and joined back
// after it, so they overlap the chain.
template <typename F>
static int group_fanout(omx_group g, F &&f) {
    omx_ctx ctx = g->ctx;
    hipStream_t main = ctx->stream;
    const bool prof = ctx->prof_on;
    ctx->prof_on = false;
    int rc = OMX_OK;
    if (hipEventRecord(g->fork, main) != hipSuccess) rc = OMX_ERR_HIP;
    for (int s = 0; s < g->nsys && !rc; ++s) {
        if (g->chain[s]) continue;
        if (hipStreamWaitEvent(g->streams[s], g->fork, 0) != hipSuccess) {
            rc = OMX_ERR_HIP;
            break;
        }
        ctx->stream = g->streams[s];
        rc = f(s);
        ctx->stream = main;
        if (!rc && hipEventRecord(g->joins[s], g->streams[s]) != hipSuccess) rc = OMX_ERR_HIP;
    }
    for (int s = 0; s < g->nsys && !rc; ++s)
        if (g->chain[s]) rc = f(s);
    ctx->stream = main;
    for (int s = 0; s < g->nsys && !rc; ++s)
        if (!g->chain[s] && hipStreamWaitEvent(main, g->joins[s], 0) != hipSuccess) rc = OMX_ERR_HIP;
    ctx->launch_mask = nullptr;
    ctx->prof_on = prof;
    if (rc == OMX_ERR_HIP) set_error("omx_group: stream fork/join failed");
    return rc;
}

static void group_free(omx_group g) {
    if (!g) return;
    if (g->ctx) (void)hipStreamSynchronize(g->ctx->stream);
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    for (auto &M : g->m)
        for (double *p : M)
            if (p) (void)hipFree(p);
    for (double *p : g->partial)
        if (p) (void)hipFree(p);
    for (double *p : g->partial2)
        if (p) (void)hipFree(p);
    void *ptrs[] = {g->coeff, g->dots, g->sc, g->flags};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (hipStream_t st : g->streams)
        if (st) (void)hipStreamDestroy(st);
    for (hipEvent_t e : g->joins)
        if (e) (void)hipEventDestroy(e);
    if (g->fork) (void)hipEventDestroy(g->fork);
    delete g;
}

extern "C" {

int omx_group_create(omx_ctx ctx, int nsys, const int *no, const omx_bsr *H, const omx_bsr *S, omx_group *out) {
    OMX_REQUIRE(ctx && out && nsys > 0 && no && H && S, "omx_group_create: bad arguments");
    for (int s = 0; s < nsys; ++s) {
        OMX_REQUIRE(H[s] && S[s] && H[s]->ctx == ctx && S[s]->ctx == ctx, "omx_group_create: system %d: operators "
                    "missing or on another context", s);
        OMX_REQUIRE(!H[s]->is_complex && !S[s]->is_complex, "omx_group_create: system %d: complex operator", s);
        OMX_REQUIRE(H[s]->nb == S[s]->nb && no[s] > 0 && no[s] <= H[s]->nb,
                    "omx_group_create: system %d: bad sizes (no %d, H %d, S %d)", s, no[s], H[s]->nb, S[s]->nb);
    }
    omx_group g = new omx_group_s();
    g->ctx = ctx;
    g->nsys = nsys;
    g->no.assign(no, no + nsys);
    g->H.assign(H, H + nsys);
    g->S.assign(S, S + nsys);
    g->m.resize(nsys);
    g->partial.assign(nsys, nullptr);
    g->partial2.assign(nsys, nullptr);
    g->streams.assign(nsys, nullptr);
    g->joins.assign(nsys, nullptr);
    g->chain.assign(nsys, 0);
    for (int s = 0; s < nsys; ++s) {
        const int64_t t = (no[s] + 127) / 128;
        g->chain[s] = (t * t >= 2 * ctx->num_cus);  // its Gram alone fills every workgroup slot
    }
    bool ok = true;
    for (int s = 0; s < nsys && ok; ++s) {
        const int n = H[s]->nb, o = no[s];
        g->nb.push_back(n);
        g->m[s].fill(nullptr);
        for (int q = 0; q < 14 && ok; ++q) {
            const size_t elems = (size_t)o * (q < M_HW ? n : o);
            ok = hipMalloc(&g->m[s][q], sizeof(double) * elems) == hipSuccess;
            if (ok) ok = hipMemsetAsync(g->m[s][q], 0, sizeof(double) * elems, ctx->stream) == hipSuccess;
        }
        ok = ok && hipMalloc(&g->partial[s], sizeof(double) * coeff_partial_doubles(o, 1)) == hipSuccess &&
             hipMalloc(&g->partial2[s], sizeof(double) * pr_partial_doubles(o, n, 1)) == hipSuccess &&
             hipStreamCreateWithFlags(&g->streams[s], hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&g->joins[s], hipEventDisableTiming) == hipSuccess;
    }
    ok = ok && hipMalloc(&g->coeff, sizeof(double) * 18 * nsys) == hipSuccess &&
         hipMalloc(&g->dots, sizeof(double) * 2 * nsys) == hipSuccess &&
         hipMalloc(&g->sc, sizeof(double) * (4 * nsys + 2)) == hipSuccess &&
         hipMalloc(&g->flags, sizeof(int) * F_COUNT * nsys) == hipSuccess &&
         hipEventCreateWithFlags(&g->fork, hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        group_free(g);
        set_error("omx_group_create: device allocation failed");
        return OMX_ERR_NOMEM;
    }
    OMX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    *out = g;
    return OMX_OK;
}

int omx_group_destroy(omx_group g) {
    group_free(g);
    return OMX_OK;
}

int omx_group_set_c(omx_group g, int s, const double *C) {
    OMX_REQUIRE(g && C && s >= 0 && s < g->nsys, "omx_group_set_c: bad arguments");
    OMX_CHECK_HIP(hipMemcpyAsync(g->m[s][M_C], C, sizeof(double) * (size_t)g->no[s] * g->nb[s],
                                 hipMemcpyHostToDevice, g->ctx->stream));
    OMX_CHECK_HIP(hipStreamSynchronize(g->ctx->stream));
    return OMX_OK;
}

int omx_group_get_c(omx_group g, int s, double *C) {
    OMX_REQUIRE(g && C && s >= 0 && s < g->nsys, "omx_group_get_c: bad arguments");
    OMX_CHECK_HIP(hipMemcpyAsync(C, g->m[s][M_C], sizeof(double) * (size_t)g->no[s] * g->nb[s],
                                 hipMemcpyDeviceToHost, g->ctx->stream));
    OMX_CHECK_HIP(hipStreamSynchronize(g->ctx->stream));
    return OMX_OK;
}

int omx_group_solve(omx_group g, const double *eta, double cg_tol, int max_steps, int check_every, int use_graph,
                    double *e_min, int *icg, int *converged, int *steps_run) {
    OMX_REQUIRE(g && eta && e_min && icg && converged && max_steps > 0 && check_every > 0,
                "omx_group_solve: bad arguments");
    omx_ctx ctx = g->ctx;
    hipStream_t st = ctx->stream;
    const int ns = g->nsys;
    const double tol = cg_tol < 0.0 ? 1.0e-9 : cg_tol;
    int rc;
    hipLaunchKernelGGL(k_batch_init_flags, dim3((ns + 63) / 64), dim3(64), 0, st, ns, g->flags + F_ACTIVE * ns,
                       g->flags + F_ICG * ns, g->flags + F_CONV * ns, g->flags + F_NFAIL * ns, g->sc, tol, max_steps);
    // initial state of every system (omm.F90:350-370), eager: HC = C H, SC = C S, HW = HC C^T,
    // SW = SC C^T, G; with tuning on, each system's shapes are tuned here, before any capture
    rc = group_fanout(g, [&](int s) {
        auto &M = g->m[s];
        const int no = g->no[s], nb = g->nb[s];
        ctx->launch_mask = gflag(g, F_ACTIVE, s);
        int r = omx_spmm(ctx, OMX_F64, no, nb, 1.0, M[M_C], no, g->H[s], 0.0, M[M_HC], no);
        if (!r) r = omx_spmm(ctx, OMX_F64, no, nb, 1.0, M[M_C], no, g->S[s], 0.0, M[M_SC], no);
        if (!r) r = group_gram(g, s, M[M_HC], M[M_C], M[M_HW], true);
        if (!r) r = group_gram(g, s, M[M_SC], M[M_C], M[M_SW], true);
        ctx->launch_mask = nullptr;
        if (!r) r = group_calc_G(g, s, gflag(g, F_ACTIVE, s));
        return r;
    });
    if (rc) return rc;
    // check_every steps of every system captured as one graph (one branch per system and step,
    // joined after each step), one launch per flag read; see omx_zbatch_solve
    if (use_graph && (!g->exec || g->graph_steps != check_every)) {
        if (g->exec) (void)hipGraphExecDestroy(g->exec);
        if (g->graph) (void)hipGraphDestroy(g->graph);
        g->exec = nullptr;
        g->graph = nullptr;
        g->graph_steps = 0;
        OMX_CHECK_HIP(hipStreamSynchronize(st));  // tuning (if on) has finished before the capture
        OMX_CHECK_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        rc = OMX_OK;
        for (int q = 0; q < check_every && !rc; ++q) rc = group_fanout(g, [&](int s) { return group_sys_step(g, s); });
        hipGraph_t gr = nullptr;
        hipError_t e = hipStreamEndCapture(st, &gr);
        if (rc || e != hipSuccess) {
            if (gr) (void)hipGraphDestroy(gr);
            if (rc) return rc;
            OMX_CHECK_HIP(e);
        }
        g->graph = gr;
        OMX_CHECK_HIP(hipGraphInstantiate(&g->exec, gr, nullptr, nullptr, 0));
        g->graph_steps = check_every;
    }
    std::vector<int> h_act(ns);
    int done_steps = 0;
    while (done_steps < max_steps) {
        if (use_graph) {
            OMX_CHECK_HIP(hipGraphLaunch(g->exec, st));
            done_steps += check_every;
        } else {
            const int burst = std::min(check_every, max_steps - done_steps);
            for (int q = 0; q < burst; ++q)
                if ((rc = group_fanout(g, [&](int s) { return group_sys_step(g, s); }))) return rc;
            done_steps += burst;
        }
        OMX_CHECK_HIP(hipMemcpyAsync(h_act.data(), g->flags + F_ACTIVE * ns, sizeof(int) * ns, hipMemcpyDeviceToHost,
                                     st));
        OMX_CHECK_HIP(hipStreamSynchronize(st));
        bool any = false;
        for (int s = 0; s < ns; ++s) any = any || h_act[s];
        if (!any) break;
    }
    if (steps_run) *steps_run = done_steps;  // step sequences issued (masked ones included)
    // e_min + eta*TrQS, TrQS = tr[(2I - SW) SW] = 2 tr(SW) - <SW, SW> (omm.F90:576-593)
    std::vector<double> sc(4 * ns);
    std::vector<int> h_icg(ns), h_conv(ns);
    OMX_CHECK_HIP(hipMemcpyAsync(sc.data(), g->sc, sizeof(double) * 4 * ns, hipMemcpyDeviceToHost, st));
    OMX_CHECK_HIP(hipMemcpyAsync(h_icg.data(), g->flags + F_ICG * ns, sizeof(int) * ns, hipMemcpyDeviceToHost, st));
    OMX_CHECK_HIP(hipMemcpyAsync(h_conv.data(), g->flags + F_CONV * ns, sizeof(int) * ns, hipMemcpyDeviceToHost, st));
    OMX_CHECK_HIP(hipStreamSynchronize(st));
    for (int s = 0; s < ns; ++s) {
        double tr[2], dd[2];
        const double *W = g->m[s][M_SW];
        if ((rc = omx_trace(ctx, OMX_F64, g->no[s], W, g->no[s], tr))) return rc;
        if ((rc = omx_mm_trace(ctx, OMX_F64, g->no[s], g->no[s], W, g->no[s], W, g->no[s], dd))) return rc;
        e_min[s] = sc[4 * s + 2] + eta[s] * (2.0 * tr[0] - dd[0]);
        icg[s] = h_icg[s];
        converged[s] = h_conv[s];
    }
    return OMX_OK;
}

}  // extern "C"
