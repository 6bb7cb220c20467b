"""Fixed launch sequences of single kernels for hardware-counter runs.

Runs `--launches` launches of one library kernel at the benchmark shape (the 16x16x16 lattice:
X, Y 8192 x 16384), then `--copies` plain device copies of a known size (1 GiB read + 1 GiB written
per copy at L=16, FP64).  Counter totals per dispatch of the copy give the scale between a counter
and bytes; the kernel's dispatches then convert to HBM bytes per launch.

    rocprofv3 --pmc FETCH_SIZE -d OUT -o pmc -- python3 omm-bundle_amd/kernel_probe.py spmm
    rocprofv3 --pmc WRITE_SIZE -d OUT -o pmc -- python3 omm-bundle_amd/kernel_probe.py spmm
    rocprofv3 --kernel-trace --stats -d OUT -o run -- python3 omm-bundle_amd/kernel_probe.py gemm
"""
from __future__ import annotations

import argparse
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import ommhip as oh  # noqa: E402


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("kernel", choices=["spmm", "gemm", "calc_g"])
    ap.add_argument("--L", type=int, default=16)
    ap.add_argument("--dtype", choices=["f64", "f32"], default="f64")
    ap.add_argument("--prec", choices=["f64", "f32", "bf16", "bf16x3", "bf16x6"], default=None)
    ap.add_argument("--pad", type=int, default=0, help="extra rows in the leading dimension of X and Y")
    ap.add_argument("--shape", default=None, help="SpMM shape batch,wpg,jper,mode (default: library default)")
    ap.add_argument("--launches", type=int, default=3)
    ap.add_argument("--copies", type=int, default=3)
    a = ap.parse_args(argv)
    dt = oh.OMX_F64 if a.dtype == "f64" else oh.OMX_F32
    prec = {"f64": oh.PREC_F64, "f32": oh.PREC_F32, "bf16": oh.PREC_BF16, "bf16x3": oh.PREC_BF16X3,
            "bf16x6": oh.PREC_BF16X6}[a.prec or a.dtype]
    nb = 4 * a.L ** 3
    no = nb // 2
    ld = no + a.pad
    ctx = oh.Ctx(0)
    mats = []
    if a.kernel == "spmm":
        cp, ri, bv = oh.lattice_bsr(a.L, 0)
        H = oh.Bsr(ctx, nb, cp, ri, bv)
        if a.shape:
            ctx.spmm_shape(*[int(x) for x in a.shape.split(",")])
        X, Y = oh.DMat(ctx, no, nb, dt, ld=ld), oh.DMat(ctx, no, nb, dt, ld=ld)
        mats += [X, Y]
        run = lambda: oh.spmm(ctx, X, H, Y)  # noqa: E731
    elif a.kernel == "gemm":
        # Gram product HW = HC * C^T: (no x nb)(nb x no)
        A, B, G = oh.DMat(ctx, no, nb, dt, ld=ld), oh.DMat(ctx, no, nb, dt, ld=ld), oh.DMat(ctx, no, no, dt, ld=ld)
        mats += [A, B, G]
        run = lambda: oh.gemm(ctx, "n", "t", A, B, G, prec=prec)  # noqa: E731
    else:
        # gradient pair G = 4 HC - 2 SW HC - 2 HW SC
        W1, W2 = oh.DMat(ctx, no, no, dt, ld=ld), oh.DMat(ctx, no, no, dt, ld=ld)
        HC, SC, G = (oh.DMat(ctx, no, nb, dt, ld=ld) for _ in range(3))
        mats += [W1, W2, HC, SC, G]
        run = lambda: oh.omm_calc_G(ctx, W1, W2, HC, SC, G, prec=prec)  # noqa: E731
    for _ in range(a.launches):
        run()
    ctx.sync()
    # reference copies: Y := X over no x nb elements of the storage type
    S, D = oh.DMat(ctx, no, nb, dt), oh.DMat(ctx, no, nb, dt)
    for _ in range(a.copies):
        oh.add(ctx, "n", S, D, 1.0, 0.0)
    ctx.sync()
    copy_bytes = no * nb * (8 if dt == oh.OMX_F64 else 4)
    print(f"kernel_probe {a.kernel} L={a.L} dtype={a.dtype} pad={a.pad} launches={a.launches}; "
          f"copy reads {copy_bytes} B and writes {copy_bytes} B per launch", flush=True)
    return 0  # buffers and context are released at process exit


if __name__ == "__main__":
    sys.exit(main())
