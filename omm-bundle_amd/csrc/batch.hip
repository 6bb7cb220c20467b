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
#include <algorithm>
#include <cmath>
#include <vector>

#include "omx_internal.h"

struct omx_zbatch_s {
    omx_ctx ctx = nullptr;
    int nsys = 0, no = 0, nb = 0, nbc = 0;
    int64_t nnzb = 0;
    int *colptr = nullptr, *rowidx = nullptr;
    double *hval = nullptr, *sval = nullptr;  // nsys * nnzb * 32
    // state, system s at s * (no*nb) or s * (no*no) complex elements, ld = no
    double2 *C = nullptr, *HC = nullptr, *SC = nullptr, *G = nullptr, *Gp = nullptr, *D = nullptr, *HG = nullptr,
            *SG = nullptr;
    double2 *HW = nullptr, *SW = nullptr, *HWd = nullptr, *SWd = nullptr, *HWdd = nullptr, *SWdd = nullptr;
    // per-system scalars
    double *coeff = nullptr;  // 18 per system
    double *dots = nullptr;   // 2 per system
    double *sc = nullptr;     // [x, lambda, e_min, e_diff] per system
    int *active = nullptr;    // per-system flags, F_COUNT arrays of nsys (see flag())
    double *partial = nullptr, *partial2 = nullptr;
    double tol = 1e-9;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
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
//   cont   stepped and not converged: HW, HC, SC and G are updated (omm.F90:506-534)
//   grad   cont and not failf: the gradient is recomputed and lambda follows (omm.F90:536-546)
// sc per system: [0] x_min, [1] lambda, [2] e_min, [3] e_diff
__global__ void k_batch_quartic(int nsys, const double *__restrict__ coeff, double *__restrict__ sc,
                                const int *__restrict__ active, int *__restrict__ failf, int *__restrict__ cont,
                                int *__restrict__ grad, int *__restrict__ icg, int *__restrict__ conv,
                                int *__restrict__ nfail, double tol) {
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
    const bool done = (ed <= tol);
    if (done) conv[s] = 1;
    failf[s] = fail ? 1 : 0;
    cont[s] = done ? 0 : 1;
    grad[s] = (!done && !fail) ? 1 : 0;
}

// D := G + lambda*D; G_p := G (active systems)
__global__ __launch_bounds__(BT_NT) void k_batch_direction(int64_t len, const double2 *__restrict__ G,
                                                           double2 *__restrict__ D, double2 *__restrict__ Gp,
                                                           const double *__restrict__ sc,
                                                           const int *__restrict__ active) {
    const int s = blockIdx.y;
    if (!active[s]) return;
    const double lam = sc[4 * s + 1];
    const int64_t o = (int64_t)s * len;
    for (int64_t q = (int64_t)blockIdx.x * BT_NT + threadIdx.x; q < len; q += (int64_t)gridDim.x * BT_NT) {
        const double2 g = G[o + q];
        double2 d;
        if (lam == 0.0) {
            d = g;
        } else {
            const double2 dd = D[o + q];
            d.x = lam * dd.x + g.x;
            d.y = lam * dd.y + g.y;
        }
        D[o + q] = d;
        Gp[o + q] = g;
    }
}

// For systems with mask[s]: Y := Y + x*X, or Y := fail_scale*Y after a failed line search.
__global__ __launch_bounds__(BT_NT) void k_batch_axpy(int64_t len, const double2 *__restrict__ X,
                                                      double2 *__restrict__ Y, const double *__restrict__ sc,
                                                      const int *__restrict__ mask, const int *__restrict__ failf,
                                                      double fail_scale) {
    const int s = blockIdx.y;
    if (!mask[s]) return;
    const bool failed = failf[s] != 0;
    const double x = sc[4 * s];
    const int64_t o = (int64_t)s * len;
    for (int64_t q = (int64_t)blockIdx.x * BT_NT + threadIdx.x; q < len; q += (int64_t)gridDim.x * BT_NT) {
        double2 y = Y[o + q];
        if (failed) {
            y.x *= fail_scale;
            y.y *= fail_scale;
        } else {
            const double2 xx = X[o + q];
            y.x = y.x + x * xx.x;
            y.y = y.y + x * xx.y;
        }
        Y[o + q] = y;
    }
}

// For systems with mask[s]: W := W + x*Wd^H + x*Wd + x^2*Wdd (the reference's order), or
// W := 0.25*W after a failed line search.
__global__ __launch_bounds__(BT_NT) void k_batch_update_w(int n, const double2 *__restrict__ Wd,
                                                          const double2 *__restrict__ Wdd, double2 *__restrict__ W,
                                                          const double *__restrict__ sc, const int *__restrict__ mask,
                                                          const int *__restrict__ failf) {
    constexpr int T = 32;
    __shared__ double tr[T][T + 1], ti[T][T + 1];
    const int s = blockIdx.z;
    if (!mask[s]) return;  // whole workgroup, before the barrier
    const int64_t o = (int64_t)s * n * n;
    const bool failed = failf[s] != 0;
    const double x = sc[4 * s], x2 = x * x;
    const int tx = threadIdx.x % T, ty = threadIdx.x / T;
    const int i0 = blockIdx.x * T, j0 = blockIdx.y * T;
    for (int b = ty; b < T; b += BT_NT / T) {
        const int r = j0 + tx, c = i0 + b;
        double2 w = {0.0, 0.0};
        if (r < n && c < n) w = Wd[o + r + (int64_t)c * n];
        tr[tx][b] = w.x;
        ti[tx][b] = -w.y;  // conjugate transpose
    }
    __syncthreads();
    for (int c = ty; c < T; c += BT_NT / T) {
        const int i = i0 + tx, j = j0 + c;
        if (i >= n || j >= n) continue;
        const int64_t q = o + i + (int64_t)j * n;
        double2 w = W[q];
        if (failed) {
            w.x *= 0.25;
            w.y *= 0.25;
        } else {
            const double2 d = Wd[q], dd = Wdd[q];
            w.x = w.x + x * tr[c][tx];
            w.x = w.x + x * d.x;
            w.x = w.x + x2 * dd.x;
            w.y = w.y + x * ti[c][tx];
            w.y = w.y + x * d.y;
            w.y = w.y + x2 * dd.y;
        }
        W[q] = w;
    }
}

// After a rescaling (omm.F90:519-523): G_p := G; G := 1.5*HC + G (HC already halved).
__global__ __launch_bounds__(BT_NT) void k_batch_fail_grad(int64_t len, const double2 *__restrict__ HC,
                                                           double2 *__restrict__ G, double2 *__restrict__ Gp,
                                                           const int *__restrict__ cont, const int *__restrict__ failf) {
    const int s = blockIdx.y;
    if (!cont[s] || !failf[s]) return;
    const int64_t o = (int64_t)s * len;
    for (int64_t q = (int64_t)blockIdx.x * BT_NT + threadIdx.x; q < len; q += (int64_t)gridDim.x * BT_NT) {
        const double2 g = G[o + q], h = HC[o + q];
        Gp[o + q] = g;
        double2 r;
        r.x = 1.5 * h.x + g.x;
        r.y = 1.5 * h.y + g.y;
        G[o + q] = r;
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

__global__ void k_batch_init_flags(int nsys, int *__restrict__ active, int *__restrict__ icg, int *__restrict__ conv,
                                   int *__restrict__ nfail, double *__restrict__ sc) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
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
__global__ __launch_bounds__(BT_NT) void k_batch_scale_copy(int64_t len, const double2 *__restrict__ X,
                                                            double2 *__restrict__ Y, double a,
                                                            const int *__restrict__ mask) {
    const int s = blockIdx.y;
    if (!mask[s]) return;
    const int64_t o = (int64_t)s * len;
    for (int64_t q = (int64_t)blockIdx.x * BT_NT + threadIdx.x; q < len; q += (int64_t)gridDim.x * BT_NT) {
        const double2 x = X[o + q];
        double2 y;
        y.x = a * x.x;
        y.y = a * x.y;
        Y[o + q] = y;
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
// launch, or three passes when no is not a multiple of the K slice.
static int batch_calc_G(omx_zbatch b, const int *mask) {
    omx_ctx ctx = b->ctx;
    const int ns = b->nsys, no = b->no, nb = b->nb;
    const int64_t lenC = (int64_t)no * nb, lenW = (int64_t)no * no;
    int rc = launch_zgemm_pair_batched(ctx, no, nb, no, -2.0, b->SW, b->HW, no, lenW, b->HC, b->SC, no, lenC, 4.0,
                                       b->HC, 0.0, b->G, no, lenC, ns, mask);
    if (rc != OMX_ERR_UNSUPPORTED) return rc;
    hipLaunchKernelGGL(k_batch_scale_copy, bt_grid(lenC, ns), dim3(BT_NT), 0, ctx->stream, lenC, b->HC, b->G, 4.0,
                       mask);
    if ((rc = launch_zgemm_batched(ctx, 'n', 'n', no, nb, no, -2.0, b->SW, no, lenW, b->HC, no, lenC, 1.0, b->G, no,
                                   lenC, ns, mask)))
        return rc;
    return launch_zgemm_batched(ctx, 'n', 'n', no, nb, no, -2.0, b->HW, no, lenW, b->SC, no, lenC, 1.0, b->G, no, lenC,
                                ns, mask);
}

// One CG step of every system (the sequence that is captured).
static int batch_step(omx_zbatch b) {
    omx_ctx ctx = b->ctx;
    const int ns = b->nsys, no = b->no, nb = b->nb;
    const int64_t lenC = (int64_t)no * nb, lenW = (int64_t)no * no, sV = 32 * b->nnzb;
    int *act = flag(b, F_ACTIVE), *failf = flag(b, F_FAIL), *cont = flag(b, F_CONT), *grad = flag(b, F_GRAD);
    int rc;
    hipLaunchKernelGGL(k_batch_direction, bt_grid(lenC, ns), dim3(BT_NT), 0, ctx->stream, lenC, b->G, b->D, b->Gp,
                       b->sc, act);
    if ((rc = launch_spmm_z_batched(ctx, no, nb, b->nbc, b->colptr, b->rowidx, b->hval, sV, b->D, no, lenC, b->HG, no,
                                    lenC, ns, act)))
        return rc;
    if ((rc = launch_spmm_z_batched(ctx, no, nb, b->nbc, b->colptr, b->rowidx, b->sval, sV, b->D, no, lenC, b->SG, no,
                                    lenC, ns, act)))
        return rc;
    const double2 *src[4] = {b->HC, b->SC, b->HG, b->SG};
    double2 *dst[4] = {b->HWd, b->SWd, b->HWdd, b->SWdd};
    for (int q = 0; q < 4; ++q)
        if ((rc = launch_zgemm_batched(ctx, 'n', 'c', no, no, nb, 1.0, src[q], no, lenC, b->D, no, lenC, 0.0, dst[q],
                                       no, lenW, ns, act)))
            return rc;
    if ((rc = launch_coeff_z_batched(ctx, no, no, lenW, b->HW, b->SW, b->HWd, b->SWd, b->HWdd, b->SWdd, ns,
                                     b->partial, b->coeff)))
        return rc;
    hipLaunchKernelGGL(k_batch_quartic, dim3((ns + 63) / 64), dim3(64), 0, ctx->stream, ns, b->coeff, b->sc, act,
                       failf, cont, grad, flag(b, F_ICG), flag(b, F_CONV), flag(b, F_NFAIL), b->tol);
    // C and SW for every system that stepped (omm.F90:470-502)
    hipLaunchKernelGGL(k_batch_axpy, bt_grid(lenC, ns), dim3(BT_NT), 0, ctx->stream, lenC, b->D, b->C, b->sc, act,
                       failf, 0.5);
    const dim3 wgrid((no + 31) / 32, (no + 31) / 32, ns);
    hipLaunchKernelGGL(k_batch_update_w, wgrid, dim3(BT_NT), 0, ctx->stream, no, b->SWd, b->SWdd, b->SW, b->sc, act,
                       failf);
    // HW, HC, SC for the systems still running (omm.F90:511-527)
    hipLaunchKernelGGL(k_batch_update_w, wgrid, dim3(BT_NT), 0, ctx->stream, no, b->HWd, b->HWdd, b->HW, b->sc, cont,
                       failf);
    hipLaunchKernelGGL(k_batch_axpy, bt_grid(lenC, ns), dim3(BT_NT), 0, ctx->stream, lenC, b->HG, b->HC, b->sc, cont,
                       failf, 0.5);
    hipLaunchKernelGGL(k_batch_axpy, bt_grid(lenC, ns), dim3(BT_NT), 0, ctx->stream, lenC, b->SG, b->SC, b->sc, cont,
                       failf, 0.5);
    hipLaunchKernelGGL(k_batch_fail_grad, bt_grid(lenC, ns), dim3(BT_NT), 0, ctx->stream, lenC, b->HC, b->G, b->Gp,
                       cont, failf);
    // G = 4HC - 2SW*HC - 2HW*SC where the line search succeeded, then Polak-Ribiere
    if ((rc = batch_calc_G(b, grad))) return rc;
    if ((rc = launch_pr_z_batched(ctx, no, nb, no, lenC, b->G, b->Gp, ns, b->partial2, b->dots))) return rc;
    hipLaunchKernelGGL(k_batch_lambda, dim3((ns + 63) / 64), dim3(64), 0, ctx->stream, ns, b->dots, b->sc, grad, failf,
                       cont, act);
    OMX_CHECK_HIP(hipGetLastError());
    return OMX_OK;
}

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

int omx_zbatch_create(omx_ctx ctx, int nsys, int no, int nb, const int *bcolptr, const int *browidx,
                      const double *hval, const double *sval, omx_zbatch *out) {
    OMX_REQUIRE(ctx && out && nsys > 0 && no > 0 && nb > 0 && bcolptr && browidx && hval && sval,
                "omx_zbatch_create: bad arguments");
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
    double2 **nm[] = {&b->C, &b->HC, &b->SC, &b->G, &b->Gp, &b->D, &b->HG, &b->SG};
    for (auto p : nm) ok = ok && hipMalloc(p, sizeof(double2) * lenC) == hipSuccess;
    double2 **nn[] = {&b->HW, &b->SW, &b->HWd, &b->SWd, &b->HWdd, &b->SWdd};
    for (auto p : nn) ok = ok && hipMalloc(p, sizeof(double2) * lenW) == hipSuccess;
    ok = ok && hipMalloc(&b->coeff, sizeof(double) * 18 * nsys) == hipSuccess &&
         hipMalloc(&b->dots, sizeof(double) * 2 * nsys) == hipSuccess &&
         hipMalloc(&b->sc, sizeof(double) * 4 * nsys) == hipSuccess &&
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

int omx_zbatch_destroy(omx_zbatch b) {
    batch_free(b);
    return OMX_OK;
}

int omx_zbatch_set_c(omx_zbatch b, const double *C) {
    OMX_REQUIRE(b && C, "omx_zbatch_set_c: bad arguments");
    OMX_CHECK_HIP(hipMemcpyAsync(b->C, C, sizeof(double2) * (size_t)b->no * b->nb * b->nsys, hipMemcpyHostToDevice,
                                 b->ctx->stream));
    OMX_CHECK_HIP(hipStreamSynchronize(b->ctx->stream));
    return OMX_OK;
}

int omx_zbatch_get_c(omx_zbatch b, double *C) {
    OMX_REQUIRE(b && C, "omx_zbatch_get_c: bad arguments");
    OMX_CHECK_HIP(hipMemcpyAsync(C, b->C, sizeof(double2) * (size_t)b->no * b->nb * b->nsys, hipMemcpyDeviceToHost,
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
    b->tol = cg_tol < 0.0 ? 1.0e-9 : cg_tol;
    int rc;
    // initial state (omm.F90:350-370): HC = C H, SC = C S, HW = HC C^H, SW = SC C^H, G
    hipLaunchKernelGGL(k_batch_init_flags, dim3((ns + 63) / 64), dim3(64), 0, st, ns, flag(b, F_ACTIVE),
                       flag(b, F_ICG), flag(b, F_CONV), flag(b, F_NFAIL), b->sc);
    int *act = flag(b, F_ACTIVE);
    if ((rc = launch_spmm_z_batched(ctx, no, nb, b->nbc, b->colptr, b->rowidx, b->hval, sV, b->C, no, lenC, b->HC, no,
                                    lenC, ns, act)))
        return rc;
    if ((rc = launch_spmm_z_batched(ctx, no, nb, b->nbc, b->colptr, b->rowidx, b->sval, sV, b->C, no, lenC, b->SC, no,
                                    lenC, ns, act)))
        return rc;
    if ((rc = launch_zgemm_batched(ctx, 'n', 'c', no, no, nb, 1.0, b->HC, no, lenC, b->C, no, lenC, 0.0, b->HW, no,
                                   lenW, ns, act)))
        return rc;
    if ((rc = launch_zgemm_batched(ctx, 'n', 'c', no, no, nb, 1.0, b->SC, no, lenC, b->C, no, lenC, 0.0, b->SW, no,
                                   lenW, ns, act)))
        return rc;
    if ((rc = batch_calc_G(b, act))) return rc;
    // the CG steps: captured once, replayed; flags read every check_every steps
    if (use_graph && !b->exec) {
        OMX_CHECK_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        rc = batch_step(b);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(st, &g);
        if (rc) return rc;
        OMX_CHECK_HIP(e);
        b->graph = g;
        OMX_CHECK_HIP(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
    }
    std::vector<int> h_act(ns);
    int done_steps = 0;
    while (done_steps < max_steps) {
        const int burst = std::min(check_every, max_steps - done_steps);
        for (int q = 0; q < burst; ++q) {
            if (use_graph)
                OMX_CHECK_HIP(hipGraphLaunch(b->exec, st));
            else if ((rc = batch_step(b)))
                return rc;
        }
        done_steps += burst;
        OMX_CHECK_HIP(hipMemcpyAsync(h_act.data(), act, sizeof(int) * ns, hipMemcpyDeviceToHost, st));
        OMX_CHECK_HIP(hipStreamSynchronize(st));
        bool any = false;
        for (int s = 0; s < ns; ++s) any = any || h_act[s];
        if (!any) break;
    }
    if (steps_run) *steps_run = done_steps;
    // results: e_min + eta*TrQS with TrQS = Re tr[(2I - SW) SW] = 2 Re tr(SW) - <SW, SW>
    std::vector<double> sc(4 * ns);
    std::vector<int> h_icg(ns), h_conv(ns);
    OMX_CHECK_HIP(hipMemcpyAsync(sc.data(), b->sc, sizeof(double) * 4 * ns, hipMemcpyDeviceToHost, st));
    OMX_CHECK_HIP(hipMemcpyAsync(h_icg.data(), flag(b, F_ICG), sizeof(int) * ns, hipMemcpyDeviceToHost, st));
    OMX_CHECK_HIP(hipMemcpyAsync(h_conv.data(), flag(b, F_CONV), sizeof(int) * ns, hipMemcpyDeviceToHost, st));
    OMX_CHECK_HIP(hipStreamSynchronize(st));
    for (int s = 0; s < ns; ++s) {
        double tr[2], dd[2];
        const double2 *W = b->SW + (size_t)s * lenW;
        if ((rc = omx_trace(ctx, OMX_C128, no, W, no, tr))) return rc;
        if ((rc = omx_mm_trace(ctx, OMX_C128, no, no, W, no, W, no, dd))) return rc;
        const double TrQS = 2.0 * tr[0] - dd[0];
        e_min[s] = sc[4 * s + 2] + eta * TrQS;
        icg[s] = h_icg[s];
        converged[s] = h_conv[s];
    }
    return OMX_OK;
}

}  // extern "C"
