#!/usr/bin/env python3
"""OMM conjugate-gradient benchmark on one MI355X.

Workload: the periodic 16x16x16 simple-cubic lattice, 4 orbitals per site (basis n = 16384,
m = 8192 occupied states), block-sparse H and S from the seeded generator, C0 ~ U(-0.5,0.5)/sqrt(n),
H shifted by the mid-gap eta (H - eta*S), solved by omm() (flavour 0) through the Fortran
MatrixSwitch drop-in with device storage ('gdden' / 'gdcsc').

A step is one CG iteration of omm() (2 BSR SpMMs, 4 Gram products, the fused gradient product,
the fused line-search reduction, the quartic solve and the vector updates).  Inputs are resident
in HBM before timing.  `value` = FP64 CG iterations/s over exactly --steps complete iterations,
timed between the stream-idle points that end the coefficient readbacks of CG steps 1 and
steps+1 of one omm() call.
The same measurement is repeated with BF16 and FP32 products ('hbf', 'h32').  Then (unless
--no-ttc) the time-to-converge (the reference's e_diff <= 1e-10) of the FP64 and the mixed
('hmx': BF16 phase + FP64 refinement) solves from the same C0, with the energy and density-matrix
errors of the mixed run against FP64 and the exact eigenvalue sum (ttc_* keys); and the same pair
on the parameterised lattice of the same size with a wide gap, where both solves reach the exact
sum (gapped_ttc_* keys).

Roofline: the dominant kernel is the FP64 MFMA GEMM (Gram and gradient products); `achieved` =
algorithmic flops of all its launches in the timed call / their summed HIP-event durations on the
library stream.  The CPU baseline times the reference's 'lap' path (host BLAS dgemm for the dense
products, the CSC product in 'ref' order) on a row sample of one CG iteration's products, with
the oracle's 'ref' loops for every product beside it (cpu_baseline.ref_loops).
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "omm-bundle_amd"), os.path.join(ROOT, "oracle")]

import numpy as np  # noqa: E402

# Mid-gap shifts eta = (lambda_m + lambda_{m+1})/2 of the generalised problem H x = lambda S x
# for the default seed, from `python bench.py --compute-eta L` (scipy generalised eigensolver).
ETA_MID_GAP = {8: 0.04129763547490679, 16: 0.05171565082133889}

PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0


def compute_eta(L):
    import scipy.linalg as sla

    import ommhip as oh

    nb = 4 * L ** 3
    no = nb // 2
    H = oh.bsr_to_dense(nb, *oh.lattice_bsr(L, 0))
    S = oh.bsr_to_dense(nb, *oh.lattice_bsr(L, 1))
    w = sla.eigh(H, S, eigvals_only=True, subset_by_index=[no - 1, no], driver="gvx")
    return 0.5 * (w[0] + w[1]), w[1] - w[0]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def prof_totals(oh, ctxh):
    """Per kernel class: (summed event ms, launches, algorithmic flops, algorithmic bytes)."""
    out = {}
    for k, name in ((oh.K_GEMM, "gemm"), (oh.K_GEMM_PAIR, "gemm_pair"), (oh.K_SPMM, "spmm"), (oh.K_COEFF, "coeff"),
                    (oh.K_ELEM, "elem"), (oh.K_DOT, "dot")):
        ms, n = ctypes.c_double(0), ctypes.c_int64(0)
        fl, by = ctypes.c_double(0), ctypes.c_double(0)
        oh.check(oh.lib().omx_prof_read(ctxh, k, ctypes.byref(ms), ctypes.byref(n)), "prof_read")
        oh.check(oh.lib().omx_prof_work(ctxh, k, ctypes.byref(fl), ctypes.byref(by)), "prof_work")
        out[name] = (ms.value, n.value, fl.value, by.value)
    return out


def run_steps(F, oh, ctxh, mats, nb, no, eta, C0, operation, warmup, steps):
    """Warm-up call, then a timed call of exactly `steps` CG steps.  Returns (iters/s, prof)."""
    H, S, Cm = mats
    F.set_option("max_iter", max(1, warmup))
    Cm.upload(C0)
    F.omm(nb, no, H, S, Cm, init_C=True, eta=eta, cg_tol=1e-300, operation=operation, dealloc=True)
    # The first CG step reuses the line-search data of the initial energy evaluation and the last
    # one stops before the gradient update, so the timed window runs from the end of step 1 to
    # the end of step steps+1: exactly `steps` complete iterations (gradient, direction, 2 SpMM,
    # 4 Gram products, fused coefficients, line search, updates).
    F.set_option("max_iter", steps + 1)
    Cm.upload(C0)
    F.sync()
    oh.check(oh.lib().omx_prof_reset(ctxh), "prof_reset")
    oh.check(oh.lib().omx_prof_enable(ctxh, 1), "prof_enable")
    r = F.omm(nb, no, H, S, Cm, init_C=True, eta=eta, cg_tol=1e-300, operation=operation, dealloc=True)
    oh.check(oh.lib().omx_prof_enable(ctxh, 0), "prof_disable")
    F.set_option("max_iter", 0)
    assert r["icg"] == steps + 1, (r["icg"], steps)
    dt = r["t_log"][steps] - r["t_log"][0]
    return steps / dt, dt, prof_totals(oh, ctxh), r


def cpu_baseline(nb, no, seconds_hint=20.0):
    """Oracle timing on a row sample of one CG iteration's products (scaled by the sample factor)."""
    import oracle as O

    import ommhip as oh

    L = round((nb // 4) ** (1 / 3))
    frac = 16  # rows of the sample = no/16
    rs = max(1, no // frac)
    rng = np.random.default_rng(0)
    A = np.asfortranarray(rng.standard_normal((rs, nb)))
    B = np.asfortranarray(rng.standard_normal((no, nb)))
    Cg = np.zeros((rs, no), order="F")
    t = time.perf_counter()
    O.gemm("n", "c", A, B, Cg)  # Gram sample: (no/16 x nb)(nb x no)
    t_gram = (time.perf_counter() - t) * frac
    W = np.asfortranarray(rng.standard_normal((rs, no)))
    Gs = np.zeros((rs, nb), order="F")
    t = time.perf_counter()
    O.gemm("n", "n", W, B, Gs, -2.0, 1.0)  # gradient sample: (no/16 x no)(no x nb)
    t_grad = (time.perf_counter() - t) * frac
    cp, ri, bv = oh.lattice_bsr(L, 0)
    colptr, rowidx, val = bsr_to_csc_fast(nb, cp, ri, bv)
    Y = np.zeros((rs, nb), order="F")
    t = time.perf_counter()
    O.spmm_csc(A, colptr, rowidx, val, Y)  # SpMM sample
    t_spmm = (time.perf_counter() - t) * frac
    # per iteration: 4 Gram + 2 gradient GEMMs + 2 SpMMs (element-wise passes and traces not counted)
    t_iter = 4 * t_gram + 2 * t_grad + 2 * t_spmm
    ref_loops = {"value": 1.0 / t_iter, "unit": "CG iters/s", "cores": O.num_threads(), "kind": "port",
                 "sample": (f"oracle (C restatement of omm()/MatrixSwitch 'ref', OpenMP) on rows 1..{rs} of one "
                            f"Gram (x4), one gradient GEMM (x2) and one CSC SpMM (x2) of a CG iteration at n={nb}, "
                            f"scaled by {frac}; element-wise passes and traces excluded (favours the CPU)"),
                 "seconds_per_iter": t_iter}
    # the reference's 'lap' operation label: the dense products are dgemm calls of the host BLAS
    # (MatrixSwitch.F90:1303); the CSC product has no BLAS form and stays the 'ref' loop
    # (MatrixSwitch_mm_multiply.F90:263-327), timed above
    from scipy.linalg import blas
    from threadpoolctl import threadpool_info

    bl = [x for x in threadpool_info() if x.get("user_api") == "blas"]
    blas_name = (f"{bl[0].get('internal_api', 'blas')} {bl[0].get('version', '')} (scipy's)" if bl else "scipy BLAS")
    blas_threads = int(bl[0].get("num_threads", 1)) if bl else 1
    lfrac = 4
    lr = max(1, no // lfrac)
    A2 = np.asfortranarray(rng.standard_normal((lr, nb)))
    W2 = np.asfortranarray(rng.standard_normal((lr, no)))
    G2 = np.zeros((lr, nb), order="F")
    blas.dgemm(1.0, A2[:64], B, trans_b=True)  # thread pool start-up outside the timing
    t = time.perf_counter()
    blas.dgemm(1.0, A2, B, trans_b=True)  # Gram sample: (no/4 x nb)(nb x no)
    l_gram = (time.perf_counter() - t) * lfrac
    t = time.perf_counter()
    blas.dgemm(-2.0, W2, B, beta=1.0, c=G2, overwrite_c=True)  # gradient sample: (no/4 x no)(no x nb)
    l_grad = (time.perf_counter() - t) * lfrac
    l_iter = 4 * l_gram + 2 * l_grad + 2 * t_spmm
    return {"value": 1.0 / l_iter, "unit": "CG iters/s", "cores": blas_threads, "kind": "port",
            "sample": (f"the reference's 'lap' path on the host: dgemm of {blas_name} on rows 1..{lr} of one Gram (x4) "
                       f"and one gradient product (x2), scaled by {lfrac}, plus the CSC SpMM in 'ref' order (x2, the "
                       f"oracle, {O.num_threads()} threads) of a CG iteration at n={nb}; element-wise passes and "
                       f"traces excluded (favours the CPU)"),
            "blas": blas_name, "seconds_per_iter": l_iter, "gram_TFLOPs": 2.0 * no * no * nb / l_gram / 1e12,
            "ref_loops": ref_loops}


def bsr_to_csc_fast(nb, cp, ri, bv):
    """0-based CSC of a BSR-4 operator (vectorised)."""
    nbc = len(cp) - 1
    counts = np.diff(cp)
    J = np.repeat(np.arange(nbc), counts)
    # entries (l, r, c): row 4*I + r, column 4*J + c, value bv[l, r + 4c]
    l_idx = np.arange(len(ri))
    rows = (4 * ri[:, None, None] + np.arange(4)[None, :, None]).repeat(4, axis=2)  # [l, r, c]
    cols = (4 * J[:, None, None] + np.arange(4)[None, None, :]).repeat(4, axis=1)
    vals = bv.reshape(-1, 4, 4, order="F")  # [l, r, c] = bv[l, r + 4c]
    rows, cols, vals = rows.ravel(), cols.ravel(), vals.ravel()
    keep = (rows < nb) & (cols < nb)
    rows, cols, vals = rows[keep], cols[keep], vals[keep]
    order = np.lexsort((rows, cols))
    rows, cols, vals = rows[order], cols[order], vals[order]
    colptr = np.zeros(nb + 1, np.int64)
    np.add.at(colptr, cols + 1, 1)
    colptr = np.cumsum(colptr)
    del l_idx
    return colptr.astype(np.int32), rows.astype(np.int32), vals


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--L", type=int, default=16, help="lattice edge: n = 4 L^3 basis functions, m = n/2")
    ap.add_argument("--eta", type=float, default=None)
    ap.add_argument("--ttc", action="store_true", help="(default) time-to-converge of the FP64 and mixed solves")
    ap.add_argument("--no-ttc", action="store_true", help="skip the time-to-converge solves")
    ap.add_argument("--c0-seed", type=int, default=None,
                    help="seed of the initial coefficients (default: the generator's default seed)")
    ap.add_argument("--no-gapped", action="store_true",
                    help="skip the time-to-converge pair on the gapped parameterised lattice of the same size")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-lowp", action="store_true", help="skip the BF16 / FP32 step-rate runs")
    ap.add_argument("--compute-eta", type=int, default=None, help="print the mid-gap shift for lattice edge L")
    ap.add_argument("--ttc-fp32-gradient", action="store_true",
                    help="with --ttc, also time the mixed solve with an FP32 gradient (the default is BF16x3)")
    ap.add_argument("--ttc-state-bf16x3", action="store_true",
                    help="with --ttc, also time the mixed solve with the periodic state recomputation in BF16x3")
    ap.add_argument("--ttc-log", default=None, help="write the per-step energies of the --ttc solves (JSON) here")
    ap.add_argument("--kpoints", action="store_true",
                    help="k-point mode: 64 complex k-points (4x4x4 grid) of the 10x10x10 cell, batched vs sequential")
    ap.add_argument("--kpoint-steps", type=int, default=4, help="CG steps timed per k-point in --kpoints mode")
    ap.add_argument("--kpoint-L", type=int, default=10, help="cell edge of the k-point systems (n = 4 L^3)")
    ap.add_argument("--la", action="store_true",
                    help="time the SDDMM and the dense factorisations of the Cholesky/preconditioned flavours")
    ap.add_argument("--group", action="store_true",
                    help="uneven-size mode: 24 real systems, edges L drawn from U{6..20}, one graph launch per step "
                         "with autotuned tiles, against one omm() call per system")
    ap.add_argument("--group-steps", type=int, default=3, help="CG steps timed in --group mode")
    ap.add_argument("--group-count", type=int, default=24, help="systems in --group mode")
    ap.add_argument("--dims", default=None, help="non-cubic lattice Lx,Ly,Lz (overrides --L), e.g. 32,32,16")
    ap.add_argument("--kernel", choices=["gram", "spmm", "pair"], default=None,
                    help="counter mode: --launches launches of one kernel at the workload's shape (the FP64 Gram "
                         "HW = HC*C^T, the BSR SpMM Y = X*H, the calc_G pair), then 3 plain copies of a no x n matrix "
                         "that calibrate FETCH_SIZE / WRITE_SIZE against known bytes")
    ap.add_argument("--launches", type=int, default=3, help="launches in --kernel mode")
    ap.add_argument("--kernel-prec", choices=["f64", "f32", "bf16", "bf16x3", "bf16x6"], default="f64",
                    help="product precision in --kernel mode (all but f64 on FP32 storage)")
    ap.add_argument("--spmm-stencil", action="store_true",
                    help="run lattice operators through the 2.5-D stencil SpMM instead of the general BSR kernel")
    ap.add_argument("--blaslt", action="store_true",
                    help="time the vendor BF16 GEMM (torch.mm, hipBLASLt) at the BF16 Gram and pair shapes, "
                         "as a reference point for k_bgemm_tn")
    ap.add_argument("--mixed-steps", type=int, default=0,
                    help="also time this many low-precision steps of the mixed (hmx) solve, FP32 storage")
    args = ap.parse_args()

    if args.compute_eta:
        eta, gap = compute_eta(args.compute_eta)
        print(json.dumps({"L": args.compute_eta, "eta": eta, "gap": gap}))
        return 0
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              f"(multi-GPU sharding of the OMM loop is a design note, not a measured path)", file=sys.stderr)
        return 2

    import ommhip as oh
    import pyomm as F

    if args.kpoints:
        print(json.dumps(kpoint_bench(oh, F, args.kpoint_steps, L=args.kpoint_L)), flush=True)
        return 0
    if args.group:
        print(json.dumps(group_bench(oh, F, args.group_steps, args.group_count)), flush=True)
        return 0
    if args.la:
        print(json.dumps(la_bench(oh, args.L)), flush=True)
        return 0
    if args.blaslt:
        print(json.dumps(blaslt_bench(args.L, args.launches)), flush=True)
        return 0
    if args.kernel:
        print(json.dumps(kernel_mode(oh, args.kernel, args.L, args.launches, args.kernel_prec,
                                     stencil=args.spmm_stencil)), flush=True)
        return 0

    if args.dims:
        dims = tuple(int(x) for x in args.dims.split(","))
        L = dims if len(set(dims)) > 1 else dims[0]
    else:
        dims, L = (args.L,) * 3, args.L
    nb = 4 * dims[0] * dims[1] * dims[2]
    no = nb // 2
    eta = args.eta if args.eta is not None else ETA_MID_GAP.get(L)
    if eta is None and nb <= 16384:
        eta, _ = compute_eta(L)
    if eta is None:
        eta = GROUP_ETA  # no dense eigensolve at this size; the shift near half filling of this generator
    log(f"bench: lattice {dims} n={nb} m={no} eta={eta:.6f}")
    t = time.time()
    Hb = oh.lattice_bsr(L, 0)
    Sb = oh.lattice_bsr(L, 1)
    # H - eta*S on the union pattern (same pattern: onsite + 6 neighbour blocks)
    assert np.array_equal(Hb[0], Sb[0]) and np.array_equal(Hb[1], Sb[1])
    Hs = (Hb[0], Hb[1], Hb[2] - eta * Sb[2])
    C0 = oh.random_c0(no, nb) if args.c0_seed is None else oh.random_c0(no, nb, args.c0_seed)
    log(f"bench: inputs generated in {time.time() - t:.1f}s")

    F.setup(0)
    ctxh = F.ctx_handle()
    H = F.Mat.from_bsr(nb, *Hs)
    S = F.Mat.from_bsr(nb, *Sb)
    Cm = F.Mat(no, nb)
    mats = (H, S, Cm)
    nnzb = len(Hb[1])

    res = {}
    ips, dt, prof, r64 = run_steps(F, oh, ctxh, mats, nb, no, eta, C0, "hip", args.warmup, args.steps)
    log(f"bench: FP64 {ips:.3f} CG it/s ({dt:.3f}s for {args.steps} steps) e={r64['e_min']:.10f}")
    res["fp64"] = (ips, dt, prof)
    if not args.no_lowp:
        for lab in ("hbf", "h32"):
            i2, d2, p2, r2 = run_steps(F, oh, ctxh, mats, nb, no, eta, C0, lab, args.warmup, args.steps)
            log(f"bench: {lab} {i2:.3f} CG it/s ({d2:.3f}s)")
            res[lab] = (i2, d2, p2)

    if args.mixed_steps > 0:
        # the mixed solve's low-precision phase (FP32 state, BF16 Grams, BF16x3 gradient): K steps,
        # then one FP64 refinement step; the rate comes from the per-step timestamps
        K = args.mixed_steps
        F.set_option("lowp_max_iter", K + 1)
        F.set_option("max_iter", 1)
        Cm.upload(C0)
        F.sync()
        rm = F.omm(nb, no, H, S, Cm, init_C=True, eta=eta, cg_tol=1e-300, operation="hmx", dealloc=True)
        F.set_option("lowp_max_iter", 0)
        F.set_option("max_iter", 0)
        assert rm["icg_lowp"] == K + 1, (rm["icg_lowp"], K)
        tl = rm["t_log"]
        res["mixed_lowp_ips"] = K / (tl[K] - tl[0])
        log(f"bench: mixed low-precision phase {res['mixed_lowp_ips']:.3f} CG it/s ({K} steps)")

    # roofline of the FP64 GEMM (dominant kernel) from the timed FP64 call: algorithmic flops
    # (2 m n k per product, counted by the library per launch) over the summed event durations
    steps = args.steps
    gemm_ms, gemm_n, gemm_fl, _ = prof["gemm"]
    achieved_tf = gemm_fl / (gemm_ms * 1e-3) / 1e12
    spmm_ms, spmm_n, _, spmm_by = prof["spmm"]
    spmm_gbs = spmm_by / (spmm_ms * 1e-3) / 1e9
    coeff_ms, coeff_n, _, coeff_by = prof["coeff"]
    coeff_gbs = coeff_by / (coeff_ms * 1e-3) / 1e9

    line = {
        "metric": (f"OMM time-to-converge (|\u0394E|/E<1e-10) & CG iters/s, n={nb}, mixed vs FP64; value = FP64 CG "
                   f"iters/s, time-to-converge in ttc_*"),
        "value": round(ips, 4),
        "unit": "CG iters/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * dt / steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded lattice generator)",
        "config": {"workload": f"omm flavour 0, {'x'.join(map(str, dims))} lattice, n={nb}, m={no}, eta={eta:.6f}"
                               + (" (mid-gap)" if L in ETA_MID_GAP else ""),
                   "n": nb, "m": no, "nnz_H": 16 * nnzb, "storage": "gdden/gdcsc", "operation": "hip",
                   "c0_seed": args.c0_seed if args.c0_seed is not None else oh.DEFAULT_SEED},
        "roofline": {"bound": "mfma",
                     "kernel": "FP64 Gram and state products (k_dgemm_dl direct-to-LDS; the symmetric calc_AW Grams on their "
                               "lower-triangle tiles; v_mfma_f64_16x16x4f64, 128x128 tiles)",
                     "achieved": round(achieved_tf, 3), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved_tf / PEAK_FP64_TFLOPS, 4),
                     "traffic": profiled_traffic(GEMM_PMC_FILE) if L == 16 else None,
                     "traffic_source": GEMM_PMC_FILE + " (FETCH_SIZE + WRITE_SIZE of the 8192 x 8192 x 16384 "
                                       "FP64 Gram launch, scaled by a copy kernel)",
                     "launches": gemm_n, "flops_per_launch": gemm_fl / max(gemm_n, 1),
                     "avg_launch_ms": round(gemm_ms / max(gemm_n, 1), 4)},
        "spmm": {"achieved_GBs": round(spmm_gbs, 1), "frac_hbm": round(spmm_gbs / PEAK_HBM_GBS, 4),
                 "avg_launch_ms": round(spmm_ms / max(spmm_n, 1), 4), "bytes_per_launch": spmm_by / max(spmm_n, 1)},
        "gemm_pair": {"achieved_TFLOPs": round(prof["gemm_pair"][2] / max(prof["gemm_pair"][0], 1e-9) / 1e9, 3),
                      "avg_launch_ms": round(prof["gemm_pair"][0] / max(prof["gemm_pair"][1], 1), 4),
                      "launches": prof["gemm_pair"][1]},
        "coeff_reduction": {"achieved_GBs": round(coeff_gbs, 1), "avg_launch_ms": round(coeff_ms / max(coeff_n, 1), 4)},
        "kernel_ms_in_timed_call": {k: round(v[0], 3) for k, v in prof.items()},
    }
    for lab, key in (("hbf", "bf16"), ("h32", "fp32")):
        if lab in res:
            line[f"{key}_products_iters_per_s"] = round(res[lab][0], 4)
            g = res[lab][2]["gemm"]
            line[f"{key}_products_gram_TFLOPs"] = round(g[2] / max(g[0] * 1e-3, 1e-12) / 1e12, 2)
            g = res[lab][2]["gemm_pair"]  # the gradient pair: BF16x3 under 'hbf', FP32 under 'h32'
            line[f"{key}_products_gradient_TFLOPs"] = round(g[2] / max(g[0] * 1e-3, 1e-12) / 1e12, 2)

    if "mixed_lowp_ips" in res:
        line["mixed_low_precision_iters_per_s"] = round(res["mixed_lowp_ips"], 4)
    if not args.no_ttc:
        tt = time_to_converge(F, oh, ctxh, mats, nb, no, eta, C0, log_path=args.ttc_log,
                              variants=("mixed",)
                              + (("mixed_fp32_gradient",) if args.ttc_fp32_gradient else ())
                              + (("mixed_state_bf16x3",) if args.ttc_state_bf16x3 else ()))
        line["time_to_converge"] = tt
        line.update(ttc_summary(tt, ""))
        gexact = exact_eigs(nb, f"theta0_L{L}") if isinstance(L, int) else None
        if gexact is not None and not args.no_gapped:
            # the same solves on the parameterised lattice at theta0 (one shared hopping block,
            # onsite levels -6, -6, 6, 6): the gap at half filling is 3.9 instead of 3e-4, so both
            # reach the exact eigenvalue sum
            geta = gexact["eta_mid"]
            cpg, rig, hg = oh.lattice_bsr_theta(L, oh.lattice_theta0())
            assert np.array_equal(cpg, Sb[0]) and np.array_equal(rig, Sb[1])
            Hg = F.Mat.from_bsr(nb, cpg, rig, hg - geta * Sb[2])
            log(f"bench: gapped lattice theta0 eta={geta:.6f} gap={gexact['gap_at_m']:.3f}")
            tg = time_to_converge(F, oh, ctxh, (Hg, S, Cm), nb, no, geta, C0, exact_key=f"theta0_L{L}")
            tg["workload"] = (f"omm flavour 0, parameterised {L}x{L}x{L} lattice (theta0: shared hopping block, "
                              f"onsite -6,-6,6,6), n={nb}, m={no}, eta={geta:.6f} (mid-gap)")
            line["time_to_converge_gapped"] = tg
            line.update(ttc_summary(tg, "gapped_"))
            Hg.free()

    if not args.no_cpu:
        t = time.time()
        line["cpu_baseline"] = cpu_baseline(nb, no)
        log(f"bench: cpu baseline {line['cpu_baseline']['value']:.5f} it/s (sampled in {time.time() - t:.1f}s)")
    for M in mats:
        M.free()
    print(json.dumps(line), flush=True)
    return 0


def ttc_summary(tt, prefix):
    """Flat time-to-converge figures of one FP64/mixed pair for the bench line."""
    mx = tt["mixed"]
    out = {f"{prefix}ttc_fp64_s": round(tt["fp64"]["seconds"], 3), f"{prefix}ttc_mixed_s": round(mx["seconds"], 3),
           f"{prefix}ttc_speedup": round(mx["speedup_vs_fp64"], 3),
           f"{prefix}ttc_icg_fp64": tt["fp64"]["icg"], f"{prefix}ttc_icg_mixed": mx["icg"],
           f"{prefix}mixed_energy_rel_err_vs_fp64": mx["energy_rel_err_vs_fp64"],
           f"{prefix}mixed_density_rel_fro_err_vs_fp64": mx["density_rel_fro_err_vs_fp64"],
           f"{prefix}ttc_converged": bool(tt["fp64"]["converged"] and mx["converged"])}
    if "exact" in tt:
        out[f"{prefix}fp64_energy_rel_err_vs_exact"] = tt["exact"]["fp64_energy_rel_err"]
        out[f"{prefix}mixed_energy_rel_err_vs_exact"] = mx["energy_rel_err_vs_exact"]
    return out


EIGS_FILE = "tests/golden/lattice_eigs.json"
GEMM_PMC_FILE = "profiles/k_gemm_fp64_gram_n16384_pmc.md"


def profiled_traffic(rel):
    """HBM bytes per launch recorded in a committed counter summary ('HBM bytes per launch: N')."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), rel)
    if not os.path.exists(path):
        return None
    for ln in open(path):
        if ln.strip().startswith("HBM bytes per launch:"):
            return float(ln.split(":", 1)[1].split()[0])
    return None


def exact_eigs(nb, key=None):
    """Tabulated exact eigenvalue data of the seeded lattice with nb basis functions (or the entry
    `key`, e.g. 'theta0_L16' for the parameterised lattice), or None."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), EIGS_FILE)
    if not os.path.exists(path):
        return None
    tab = json.load(open(path))
    if key is not None:
        return tab.get(key)
    for k, v in tab.items():
        if v.get("n") == nb and not k.startswith("theta"):
            return v
    return None


def time_to_converge(F, oh, ctxh, mats, nb, no, eta, C0, tol=1e-10, progress=50, log_path=None, variants=("mixed",),
                     exact_key=None):
    """FP64 omm() and the mixed-precision variants ("mixed": the default BF16x3 gradient,
    "mixed_fp32_gradient": FP32 gradient, "mixed_state_bf16x3": BF16x3 periodic state recomputation)
    from the same C0 to the reference's e_diff <= tol;
    energies and density matrices compared with the FP64 result and, when tabulated, with the exact
    eigenvalue sum."""
    H, S, Cm = mats
    out = {"cg_tol": tol}
    Dref = F.Mat(nb, nb)
    F.set_option("progress", progress)
    Cm.upload(C0)
    F.sync()
    t = time.perf_counter()
    r = F.omm(nb, no, H, S, Cm, D_min=Dref, init_C=True, eta=eta, cg_tol=tol, operation="hip", dealloc=True)
    F.sync()
    t64 = time.perf_counter() - t
    out["fp64"] = {"seconds": t64, "icg": r["icg"], "e_min": r["e_min"], "converged": r["converged"]}
    log(f"bench: ttc FP64 {t64:.2f}s icg={r['icg']} e={r['e_min']:.12f}")
    p1, ld = Dref.dptr()
    fro_ref = ctypes.c_double(0)
    oh.check(oh.lib().omx_mm_trace(ctxh, oh.OMX_F64, nb, nb, p1, ld, p1, ld, ctypes.byref(fro_ref)), "fro")
    exact = exact_eigs(nb, exact_key)
    if exact is not None:
        out["exact"] = {"sum_lowest_m": exact["sum_lowest_m"], "gap_at_m": exact["gap_at_m"], "source": EIGS_FILE,
                        "fp64_energy_rel_err": abs(r["e_min"] - exact["sum_lowest_m"]) / abs(exact["sum_lowest_m"])}
    logs = {"fp64": {"e": r["e_log"].tolist(), "t": (r["t_log"] - r["t0"]).tolist()}}
    for name in variants:
        F.set_option("lowp_grad_bf16x3", 0 if name == "mixed_fp32_gradient" else 1)
        F.set_option("lowp_state_bf16x3", 1 if name == "mixed_state_bf16x3" else 0)
        Dmx = F.Mat(nb, nb)
        Cm.upload(C0)
        F.sync()
        t = time.perf_counter()
        rm = F.omm(nb, no, H, S, Cm, D_min=Dmx, init_C=True, eta=eta, cg_tol=tol, operation="hmx", dealloc=True)
        F.sync()
        tmx = time.perf_counter() - t
        F.set_option("lowp_grad_bf16x3", 1)
        F.set_option("lowp_state_bf16x3", 0)
        # ||D_mx - D_64||_F on the device
        p2, _ = Dmx.dptr()
        oh.check(oh.lib().omx_add(ctxh, oh.OMX_F64, b"n", nb, nb, -1.0, p1, ld, 1.0, p2, ld), "add")
        fro = ctypes.c_double(0)
        oh.check(oh.lib().omx_mm_trace(ctxh, oh.OMX_F64, nb, nb, p2, ld, p2, ld, ctypes.byref(fro)), "fro")
        v = {"seconds": tmx, "icg": rm["icg"], "icg_low_precision": rm["icg_lowp"],
             "icg_fp64": rm["icg"] - rm["icg_lowp"], "e_min": rm["e_min"], "converged": rm["converged"],
             "refine_gradient_start": rm["refine_g0"], "refine_gradient_end": rm["refine_gend"],
             "gradient": "fp32" if name == "mixed_fp32_gradient" else "bf16x3",
             "state_recompute": "bf16x3" if name == "mixed_state_bf16x3" else "fp32",
             "energy_rel_err_vs_fp64": abs(rm["e_min"] - r["e_min"]) / abs(r["e_min"]),
             "density_fro_err_vs_fp64": float(np.sqrt(max(fro.value, 0.0))),
             "density_rel_fro_err_vs_fp64": float(np.sqrt(max(fro.value, 0.0) / max(fro_ref.value, 1e-300))),
             "speedup_vs_fp64": t64 / tmx}
        if exact is not None:
            v["energy_rel_err_vs_exact"] = abs(rm["e_min"] - exact["sum_lowest_m"]) / abs(exact["sum_lowest_m"])
        out[name] = v
        log(f"bench: ttc {name} {tmx:.2f}s icg={rm['icg']} (low precision {rm['icg_lowp']}) e={rm['e_min']:.12f} "
            f"speedup {t64 / tmx:.2f}x")
        logs[name] = {"e": rm["e_log"].tolist(), "t": (rm["t_log"] - rm["t0"]).tolist(),
                      "icg_low_precision": rm["icg_lowp"]}
        Dmx.free()
    F.set_option("progress", 0)
    Dref.free()
    if log_path:
        with open(log_path, "w") as f:
            json.dump(logs, f)
    return out


GROUP_ETA = 0.0517


def blaslt_bench(L, launches):
    """torch.mm in BF16 with an FP32 result (hipBLASLt on this stack) at the shapes k_bgemm_tn runs
    for the workload: the Gram (m x m, K = n) and the BF16x3 gradient pair (m x n, K = 6 m).  A
    reference point for the library's own BF16 kernel, not a path of the library."""
    import torch
    n = 4 * L ** 3
    m = n // 2
    dev = "cuda:0"
    g = torch.Generator(device=dev).manual_seed(0)
    out = {"mode": "blaslt", "L": L, "launches": launches}
    for name, (M, N, K) in (("gram", (m, m, n)), ("pair", (m, n, 6 * m))):
        a = (torch.randn(M, K, device=dev, generator=g) * 1e-2).to(torch.bfloat16)
        b = (torch.randn(N, K, device=dev, generator=g) * 1e-2).to(torch.bfloat16)
        torch.mm(a, b.t(), out_dtype=torch.float32)
        torch.cuda.synchronize()
        ts = []
        for _ in range(launches):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            torch.mm(a, b.t(), out_dtype=torch.float32)
            e1.record()
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1))
        out[name] = {"shape": [M, N, K], "ms": [round(t, 4) for t in ts],
                     "avg_PFs": round(2.0 * M * N * K / (sum(ts) / len(ts) * 1e-3) / 1e15, 4)}
        del a, b
        torch.cuda.empty_cache()
    return out


def kernel_mode(oh, kernel, L, launches, prec="f64", copies=3, stencil=False):
    """Fixed launch sequences of one library kernel at the L^3 workload's shape (X, Y no x n),
    followed by `copies` device copies of a no x n FP64 matrix (no*n*8 bytes read and written per
    copy): under rocprofv3 --pmc the copies' counter totals give the scale between a counter and
    bytes, the kernel's dispatches then convert to HBM bytes per launch.  prec other than f64 runs
    the product on FP32 storage (f32, bf16, bf16x3 or bf16x6 products)."""
    nb = 4 * L ** 3
    no = nb // 2
    dt = oh.OMX_F64 if prec == "f64" else oh.OMX_F32
    pr = {"f64": oh.PREC_F64, "f32": oh.PREC_F32, "bf16": oh.PREC_BF16, "bf16x3": oh.PREC_BF16X3,
          "bf16x6": oh.PREC_BF16X6}[prec]
    es = 8 if prec == "f64" else 4
    ctx = oh.Ctx(0)
    ctx.spmm_stencil(stencil)
    mats = []
    rng = np.random.default_rng(0)
    if kernel == "spmm":
        cp, ri, bv = oh.lattice_bsr(L, 0)
        H = oh.Bsr(ctx, nb, cp, ri, bv)
        X, Y = oh.DMat(ctx, no, nb, dt), oh.DMat(ctx, no, nb, dt)
        mats += [X, Y]
        run = lambda: oh.spmm(ctx, X, H, Y)  # noqa: E731
        work = {"bytes_per_launch": 2.0 * no * nb * es + len(ri) * (128 + 4) + 4 * (nb // 4 + 1)}
    elif kernel == "gram":
        A, B, G = oh.DMat(ctx, no, nb, dt), oh.DMat(ctx, no, nb, dt), oh.DMat(ctx, no, no, dt)
        mats += [A, B, G]
        run = lambda: oh.gemm(ctx, "n", "t", A, B, G, prec=pr)  # noqa: E731
        work = {"flops_per_launch": 2.0 * no * no * nb, "bytes_per_launch": 2.0 * no * nb * es + no * no * es}
    else:
        W1, W2 = oh.DMat(ctx, no, no, dt), oh.DMat(ctx, no, no, dt)
        HC, SC, G = (oh.DMat(ctx, no, nb, dt) for _ in range(3))
        mats += [W1, W2, HC, SC, G]
        run = lambda: oh.omm_calc_G(ctx, W1, W2, HC, SC, G, prec=pr)  # noqa: E731
        work = {"flops_per_launch": 4.0 * no * no * nb}
    for M in mats:  # finite operands (the values do not change the launch sequence)
        M.upload(np.asfortranarray(rng.standard_normal((M.rows, M.cols)) * 1e-2))
    ctx.sync()
    t = time.perf_counter()
    for _ in range(launches):
        run()
    ctx.sync()
    dt = (time.perf_counter() - t) / launches
    Sx, Dx = oh.DMat(ctx, no, nb), oh.DMat(ctx, no, nb)
    for _ in range(copies):
        oh.add(ctx, "n", Sx, Dx, 1.0, 0.0)
    ctx.sync()
    out = {"mode": "kernel", "kernel": kernel, "prec": prec, "L": L, "n": nb, "m": no, "launches": launches,
           "spmm_kernel": "stencil" if stencil else "general",
           "wall_ms_per_launch": dt * 1e3, "copy_bytes_read": no * nb * 8, "copy_bytes_written": no * nb * 8}
    out.update(work)
    for M in mats + [Sx, Dx]:
        M.free()
    return out


def group_bench(oh, F, steps, count, seed=None, stream=6):
    """Uneven-size batch: `count` real systems with cubic edges L drawn from U{6..20} by the
    counter-based generator (stream 6), H shifted by a common eta near half filling.  One
    omx_group solve (one graph launch per CG step of all systems, tiles autotuned per shape in the
    eager first pass) against one omm() call per system, both over the same CG steps."""
    seed = oh.DEFAULT_SEED if seed is None else seed
    Ls = [6 + int(15 * oh.u01(seed, stream, s)) for s in range(count)]
    eta = GROUP_ETA
    mem = sum(44.0 * (4 * L ** 3) ** 2 for L in Ls)
    log(f"bench group: {count} systems, L = {Ls}, state {mem / 1e9:.1f} GB")
    res = {"systems": count, "L": Ls, "n": [4 * L ** 3 for L in Ls], "eta": eta, "state_GB": mem / 1e9,
           "steps": steps, "dtype": "f64"}
    ctx = oh.Ctx(0)
    ctx.autotune(True)
    mats = []
    for L in Ls:
        cp, ri, h = oh.lattice_bsr(L, 0)
        _, _, sv = oh.lattice_bsr(L, 1)
        nb = 4 * L ** 3
        mats.append((oh.Bsr(ctx, nb, cp, ri, h - eta * sv), oh.Bsr(ctx, nb, cp, ri, sv)))
    g = oh.Group(ctx, [2 * L ** 3 for L in Ls], [m[0] for m in mats], [m[1] for m in mats])

    def run_group(nsteps):
        for k, L in enumerate(Ls):
            g.set_c(k, oh.random_c0(2 * L ** 3, 4 * L ** 3))
        ctx.sync()
        t0 = time.perf_counter()
        g.solve([eta] * count, cg_tol=1e-14, max_steps=nsteps, check_every=nsteps, use_graph=True)
        return time.perf_counter() - t0

    t_first = run_group(1)  # tuning, capture and one step
    t1 = run_group(1)
    tn = run_group(1 + steps)
    res["group_first_call_s"] = t_first
    res["group_s_per_step_all_systems"] = (tn - t1) / steps
    res["tuned_shapes"] = ctx.autotune_report().count("\n")
    log(f"bench group: {res['group_s_per_step_all_systems']:.3f} s per CG step of all systems "
        f"(first call with tuning {t_first:.1f} s)")
    g.free()
    for h, sv in mats:
        h.free()
        sv.free()
    ctx.close()
    # sequential: one omm() call per system through the Fortran drop-in, tuned tiles as well
    F.setup(0)
    oh.check(oh.lib().omx_autotune(F.ctx_handle(), 1), "autotune")

    def run_seq(nsteps):
        F.set_option("max_iter", nsteps)
        tot = 0.0
        for L in Ls:
            nb, no = 4 * L ** 3, 2 * L ** 3
            cp, ri, h = oh.lattice_bsr(L, 0)
            _, _, sv = oh.lattice_bsr(L, 1)
            Hm, Sm = F.Mat.from_bsr(nb, cp, ri, h - eta * sv), F.Mat.from_bsr(nb, cp, ri, sv)
            Cm = F.Mat(no, nb)
            Cm.upload(oh.random_c0(no, nb))
            F.sync()
            t0 = time.perf_counter()
            F.omm(nb, no, Hm, Sm, Cm, init_C=True, eta=eta, cg_tol=1e-14, dealloc=True)
            F.sync()
            tot += time.perf_counter() - t0
            for M in (Hm, Sm, Cm):
                M.free()
        F.set_option("max_iter", 0)
        return tot

    run_seq(1)  # tuning
    s1 = run_seq(1)
    sn = run_seq(1 + steps)
    res["sequential_s_per_step_all_systems"] = (sn - s1) / steps
    res["group_speedup"] = res["sequential_s_per_step_all_systems"] / res["group_s_per_step_all_systems"]
    flops = sum(12.0 * (2 * L ** 3) ** 2 * (4 * L ** 3) for L in Ls)
    res["group_TFLOPs"] = flops / res["group_s_per_step_all_systems"] / 1e12
    log(f"bench group: sequential {res['sequential_s_per_step_all_systems']:.3f} s per step; "
        f"group speedup {res['group_speedup']:.2f}x, {res['group_TFLOPs']:.1f} TF/s")
    return {"group_batch": res}


def la_bench(oh, L=16, reps=5):
    """The kernels outside the flavour-0 CG loop, at the n = 4 L^3 lattice (L = 16: n = 16384,
    m = 8192), FP64, inputs resident in HBM, wall clock around synchronised launches:
    * the SDDMM Cp := C^T B on H's pattern (the density matrix / dE/dH of a 'gdcsc' D_min), against
      the HBM roofline on its algorithmic bytes (C and B read once, the blocks written once);
    * the dense factorisations of the Cholesky and preconditioned flavours on the workload's own S
      and H stored dense (potrf of S, sygst of H, the SPD inverse of S, the back-transform trsm of
      an m x n state), in TFLOP/s at LAPACK's flop counts (potrf n^3/3, sygst n^3, potrf + potri
      n^3, trsm m n^2)."""
    import ctypes as C
    nb = 4 * L ** 3
    no = nb // 2
    rng = np.random.default_rng(5)
    ctx = oh.Ctx(0)
    res = {"n": nb, "m": no, "dtype": "f64"}

    def timed(fn, n_rep, prep=None):
        best = []
        for _ in range(n_rep):
            if prep is not None:
                prep()
            ctx.sync()
            t0 = time.perf_counter()
            fn()
            ctx.sync()
            best.append(time.perf_counter() - t0)
        return float(np.median(best))

    # SDDMM on H's pattern
    cp, ri, hv = oh.lattice_bsr(L, 0)
    Cp = oh.Bsr(ctx, nb, cp, ri, hv)
    dA = oh.DMat.from_host(ctx, rng.standard_normal((no, nb)))
    dB = oh.DMat.from_host(ctx, rng.standard_normal((no, nb)))
    nnzb = len(ri)
    oh.sddmm(ctx, dA, dB, Cp)  # warm-up
    t = timed(lambda: oh.sddmm(ctx, dA, dB, Cp), reps)
    alg = 2.0 * no * nb * 8 + nnzb * 128.0 + 4.0 * (nnzb + nb // 4 + 1)
    res["sddmm"] = {"ms": t * 1e3, "bytes_per_launch": alg, "achieved_GBs": alg / t / 1e9,
                    "frac_hbm": alg / t / 8.0e12, "nnzb": nnzb}
    log(f"bench la: sddmm {t * 1e3:.3f} ms, {alg / t / 1e9:.0f} GB/s")
    dA.free()
    dB.free()

    S = oh.bsr_to_dense(nb, *oh.lattice_bsr(L, 1))
    H = oh.bsr_to_dense(nb, *oh.lattice_bsr(L, 0))
    dS = oh.DMat.from_host(ctx, S)
    dU = oh.DMat(ctx, nb, nb, oh.OMX_F64)
    info = C.c_int(0)
    n3 = float(nb) ** 3

    def potrf(M):
        oh.check(oh.lib().omx_potrf(ctx.h, oh.OMX_F64, nb, M.p, M.ld, C.byref(info)), "potrf")
        assert info.value == 0, info.value

    t = timed(lambda: potrf(dU), 3, prep=lambda: dU.upload(S))
    res["potrf"] = {"ms": t * 1e3, "TFLOPs": n3 / 3 / t / 1e12}
    log(f"bench la: potrf {t * 1e3:.1f} ms")
    dH = oh.DMat(ctx, nb, nb, oh.OMX_F64)
    t = timed(lambda: oh.check(oh.lib().omx_sygst(ctx.h, oh.OMX_F64, nb, dH.p, dH.ld, dU.p, dU.ld), "sygst"), 3,
              prep=lambda: dH.upload(H))
    res["sygst"] = {"ms": t * 1e3, "TFLOPs": n3 / t / 1e12}
    log(f"bench la: sygst {t * 1e3:.1f} ms")
    dH.free()
    dX = oh.DMat.from_host(ctx, rng.standard_normal((no, nb)))
    t = timed(lambda: oh.check(oh.lib().omx_trsm(ctx.h, oh.OMX_F64, b"r", b"t", no, nb, 1.0, dU.p, dU.ld, dX.p,
                                                 dX.ld), "trsm"), 3)
    res["trsm_back_transform"] = {"ms": t * 1e3, "TFLOPs": float(no) * nb * nb / t / 1e12}
    log(f"bench la: trsm {t * 1e3:.1f} ms")
    dX.free()

    def inv():
        oh.check(oh.lib().omx_inverse_spd(ctx.h, oh.OMX_F64, nb, dS.p, dS.ld, C.byref(info)), "inverse_spd")
        assert info.value == 0, info.value

    t = timed(inv, 3, prep=lambda: dS.upload(S))
    res["inverse_spd"] = {"ms": t * 1e3, "TFLOPs": n3 / t / 1e12}
    log(f"bench la: inverse {t * 1e3:.1f} ms")
    for M in (dS, dU):
        M.free()
    return {"dense_la_and_sddmm": res}


def kpoint_bench(oh, F, steps, L=10, nk1=4, eta=0.05):
    """CG steps of the 64 k-points of the L x L x L cell (L = 10: n = 4000, m = 2000, complex128): one
    batched solve (all systems per graph-captured step) against one omm() call per k-point, both
    timed over the same number of CG steps per system, from the same resident inputs."""
    nb, no = 4 * L ** 3, 2 * L ** 3
    g = [i / nk1 for i in range(nk1)]
    phis = [(a, b, c) for a in g for b in g for c in g]
    t = time.time()
    hv, sv = [], []
    for phi in phis:
        cp, ri, h = oh.lattice_bsr_k(L, 0, phi)
        _, _, sk = oh.lattice_bsr_k(L, 1, phi)
        hv.append(h - eta * sk)
        sv.append(sk)
    C0 = oh.random_c0(no, nb).astype(np.complex128)
    log(f"bench kpoints: {len(phis)} systems n={nb} m={no}, inputs in {time.time() - t:.1f}s")
    F.setup(0)
    ctx = oh.Ctx(0)
    B = oh.ZBatch(ctx, no, nb, cp, ri, hv, sv)
    res = {"systems": len(phis), "L": L, "n": nb, "m": no, "dtype": "c128", "steps_per_system": steps}

    def run_batch(nsteps):
        B.set_c([C0] * len(phis))
        ctx.sync()
        t0 = time.perf_counter()
        B.solve(eta=eta, cg_tol=1e-14, max_steps=nsteps, check_every=nsteps, use_graph=True)
        return time.perf_counter() - t0

    run_batch(1)  # graph capture and warm-up
    t1 = run_batch(1)
    tn = run_batch(1 + steps)
    res["batched_s_per_step_all_systems"] = (tn - t1) / steps
    log(f"bench kpoints: batched {res['batched_s_per_step_all_systems'] * 1e3:.1f} ms per CG step of all systems")
    B.free()
    # sequential: one omm() call per k-point (np = 64), the same number of steps each
    Hm = [F.Mat.from_bsr_z(nb, cp, ri, h) for h in hv]
    Sm = [F.Mat.from_bsr_z(nb, cp, ri, sk) for sk in sv]
    Cm = F.Mat(no, nb, "gzden")

    def run_seq(nsteps):
        F.set_option("max_iter", nsteps)
        tot = 0.0
        for k in range(len(phis)):
            Cm.upload(C0)
            F.sync()
            t0 = time.perf_counter()
            F.omm(nb, no, Hm[k], Sm[k], Cm, init_C=True, eta=eta, cg_tol=1e-14, storage="gzden", np_=len(phis),
                  ip=k + 1, dealloc=(k == len(phis) - 1))
            F.sync()
            tot += time.perf_counter() - t0
        F.set_option("max_iter", 0)
        return tot

    run_seq(1)
    s1 = run_seq(1)
    sn = run_seq(1 + steps)
    res["sequential_s_per_step_all_systems"] = (sn - s1) / steps
    res["batched_speedup"] = res["sequential_s_per_step_all_systems"] / res["batched_s_per_step_all_systems"]
    log(f"bench kpoints: sequential {res['sequential_s_per_step_all_systems'] * 1e3:.1f} ms per CG step of all "
        f"systems; batched speedup {res['batched_speedup']:.2f}x")
    for M in Hm + Sm + [Cm]:
        M.free()
    return {"kpoint_batch": res}


if __name__ == "__main__":
    sys.exit(main())
