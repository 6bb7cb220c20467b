"""Exact generalized eigenvalues of the seeded lattice (H, S) pairs, written to lattice_eigs.json.

H and S come from lattice_gen.py (the numpy restatement of the generator, seed 0x0EEDC0DE); the
m = n/2 lowest eigenvalues of H x = e S x are computed with scipy.linalg.eigh (LAPACK, FP64).
For each edge L the file holds the sum of the m lowest (the converged OMM energy when every
occupied shifted level is negative), the eigenvalues around the m-th and the mid-gap shift.
Usage: python make_lattice_eigs.py 4 8 16   (L=16 takes a few minutes and about 8 GB)."""
import json
import os
import sys

import numpy as np
import scipy.linalg as sl

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import lattice_gen as G  # noqa: E402

SEED = 0x0EEDC0DE


def main(edges):
    path = os.path.join(HERE, "lattice_eigs.json")
    out = json.load(open(path)) if os.path.exists(path) else {}
    for L in edges:
        n = 4 * L ** 3
        m = n // 2
        H = G.dense_lattice(L, L, L, SEED, 0)
        S = G.dense_lattice(L, L, L, SEED, 1)
        w = sl.eigh(H, S, eigvals_only=True, subset_by_index=[0, m + 1], overwrite_a=True, overwrite_b=True)
        out[f"L{L}"] = {
            "L": L, "n": n, "m": m, "seed": SEED,
            "sum_lowest_m": float(np.sum(w[:m])),
            "eig_m_minus_2_to_m_plus_1": [float(x) for x in w[m - 2:m + 2]],
            "gap_at_m": float(w[m] - w[m - 1]),
            "eta_mid": float(0.5 * (w[m - 1] + w[m])),
        }
        print(L, out[f"L{L}"], flush=True)
        json.dump(out, open(path, "w"), indent=1)


if __name__ == "__main__":
    main([int(a) for a in sys.argv[1:]] or [4, 8])
