"""Oracle fixtures for omm() at the benchmark sizes, written once in the build container.

Both runs use the seeded lattice generator, C0 ~ U(-0.5,0.5)/sqrt(n) from the same generator,
H shifted by the mid-gap eta of lattice_eigs.json (H - eta*S on the shared block pattern) and the
CPU oracle (oracle/omm_oracle.c, the reference's 'ref' loops restated in C and bit-identical to the
reference serial build on the small cases in tests/test_oracle.py):

  cfg0  8x8x8 lattice, n=2048, m=1024: the whole solve, cg_tol 1e-9 with a 100-step cap
        (omm.F90:429-552 with the cap), D_min and TrQS.  About two minutes on 8 cores.
  cfg1  16x16x16 lattice, n=16384, m=8192: the first three CG steps (max_iter 3) and the
        TrQS-corrected energy.  About 21 products of 2.2 TFLOP in 'ref' order: 30-60 minutes
        on 8 cores and about 16 GB of host memory.
  cfg3  one k-point of the 10x10x10 cell's 4x4x4 grid, k = (1/4, 1/2, 3/4): complex Bloch
        operators H(k) - eta S(k), S(k) (n=4000, m=2000, eta 0.05 as in the k-point workload),
        complex start C0 = the real generator's C0, the first three CG steps through the oracle's
        complex ('szden'/'szcsc') table.

Output: omm_cfg0_L8.npz, omm_cfg1_L16_3steps.npz, omm_cfg3_L10k_3steps.npz (numpy arrays, no
pickles).  D_min at n=2048 is
kept as probes (D @ P for a seeded +-1 matrix P, its diagonal and its first 8 columns) so the file
stays small.
Usage: python make_cfg_fixtures.py cfg0 [cfg1] [cfg3]
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "omm-bundle_amd"), os.path.join(ROOT, "oracle")]

import ommhip as oh  # noqa: E402  (CPU-side generator functions only)
import oracle as O  # noqa: E402


def csc_of_bsr(nb, cp, ri, bv):
    """0-based CSC (rows ascending within a column) of a BSR-4 operator."""
    nbc = len(cp) - 1
    J = np.repeat(np.arange(nbc), np.diff(cp))
    rows = (4 * ri[:, None, None] + np.arange(4)[None, :, None]).repeat(4, axis=2)
    cols = (4 * J[:, None, None] + np.arange(4)[None, None, :]).repeat(4, axis=1)
    vals = bv.reshape(-1, 4, 4, order="F")
    rows, cols, vals = rows.ravel(), cols.ravel(), vals.ravel()
    order = np.lexsort((rows, cols))
    rows, cols, vals = rows[order], cols[order], vals[order]
    colptr = np.zeros(nb + 1, np.int64)
    np.add.at(colptr, cols + 1, 1)
    return np.cumsum(colptr).astype(np.int32), rows.astype(np.int32), np.ascontiguousarray(vals)


def inputs(L):
    eigs = json.load(open(os.path.join(HERE, "lattice_eigs.json")))[f"L{L}"]
    eta = eigs["eta_mid"]
    nb = 4 * L ** 3
    no = nb // 2
    Hb, Sb = oh.lattice_bsr(L, 0), oh.lattice_bsr(L, 1)
    assert np.array_equal(Hb[0], Sb[0]) and np.array_equal(Hb[1], Sb[1])
    Hs = csc_of_bsr(nb, Hb[0], Hb[1], Hb[2] - eta * Sb[2])
    S = csc_of_bsr(nb, *Sb)
    C0 = oh.random_c0(no, nb)
    return nb, no, eta, Hs, S, C0, Hb[2] - eta * Sb[2]


KPOINT_PHI = (0.25, 0.5, 0.75)
KPOINT_ETA = 0.05


def inputs_k(L):
    nb = 4 * L ** 3
    no = nb // 2
    Hb, Sb = oh.lattice_bsr_k(L, 0, KPOINT_PHI), oh.lattice_bsr_k(L, 1, KPOINT_PHI)
    assert np.array_equal(Hb[0], Sb[0]) and np.array_equal(Hb[1], Sb[1])
    hv = Hb[2] - KPOINT_ETA * Sb[2]
    Hs = csc_of_bsr(nb, Hb[0], Hb[1], hv)
    S = csc_of_bsr(nb, *Sb)
    C0 = oh.random_c0(no, nb).astype(np.complex128)
    return nb, no, KPOINT_ETA, Hs, S, C0, hv


def probes(nb):
    return np.where(np.random.default_rng(7).random((nb, 8)) < 0.5, -1.0, 1.0)


def make(which):
    L, cap, want_D = {"cfg0": (8, 100, True), "cfg1": (16, 3, False), "cfg3": (10, 3, False)}[which]
    nb, no, eta, Hs, S, C0, hvals = inputs_k(L) if which == "cfg3" else inputs(L)
    cg_tol = 1e-14 if which == "cfg3" else 1e-9
    Cc = np.asfortranarray(C0.copy())
    t = time.time()
    print(f"{which}: L={L} n={nb} m={no} eta={eta!r} cap={cap} threads={O.num_threads()}", flush=True)
    r = O.Omm()(nb, no, Hs, S, Cc, init_C=True, eta=eta, cg_tol=cg_tol, max_iter=cap, want_D=want_D)
    dt = time.time() - t
    print(f"{which}: icg={r['icg']} converged={r['converged']} e_min={r['e_min']!r} ({dt:.0f}s)", flush=True)
    out = dict(L=L, n=nb, m=no, eta=eta, cg_tol=cg_tol, max_iter=cap, e_min=r["e_min"], e_log=r["e_log"],
               icg=r["icg"], converged=int(r["converged"]), nfail=r["nfail"], TrQS=r["TrQS"],
               c0_sum=C0.sum(), h_sum=hvals.sum(), oracle_seconds=dt,
               oracle_threads=O.num_threads())
    if want_D:
        D = r["D"]
        P = probes(nb)
        out.update(D_probe=D @ P, D_diag=np.diag(D).copy(), D_cols=D[:, :8].copy(), C_probe=Cc @ P)
    if which == "cfg3":
        out.update(phi=np.array(KPOINT_PHI), C_probe=Cc @ probes(nb))
    name = {"cfg0": "omm_cfg0_L8.npz", "cfg1": "omm_cfg1_L16_3steps.npz", "cfg3": "omm_cfg3_L10k_3steps.npz"}[which]
    np.savez(os.path.join(HERE, name), **{k: np.asarray(v) for k, v in out.items()})
    print(f"{which}: wrote {name}", flush=True)


if __name__ == "__main__":
    for w in sys.argv[1:] or ["cfg0"]:
        make(w)
