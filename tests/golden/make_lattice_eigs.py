"""Exact generalized eigenvalues of the seeded lattice (H, S) pairs, written to lattice_eigs.json.

H and S come from lattice_gen.py (the numpy restatement of the generator, seed 0x0EEDC0DE); the
m = n/2 lowest eigenvalues of H x = e S x are computed with scipy.linalg.eigh (LAPACK, FP64).
For each edge L the file holds the sum of the m lowest (the converged OMM energy when every
occupied shifted level is negative), the eigenvalues around the m-th and the mid-gap shift.
Usage: python make_lattice_eigs.py 4 8 16   (L=16 takes a few minutes and about 8 GB).
       python make_lattice_eigs.py theta 4 16  (the parameterised lattice at theta0, keys theta0_L<L>)."""
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


def theta_lattice(L, theta, seed):
    """The parameterised ('gapped') lattice: one hopping block T = theta[0:16] (column-major) on every
    forward bond (H[s, s+e] = T, H[s+e, s] = T^T), onsite diag(theta[16:20]); S as dense_lattice."""
    ns = L ** 3
    n = 4 * ns
    T = np.asarray(theta[:16], float).reshape(4, 4, order="F")
    H = np.zeros((n, n))
    for s in range(ns):
        H[4 * s:4 * s + 4, 4 * s:4 * s + 4] += np.diag(theta[16:20])
        x, y, z = s % L, (s // L) % L, s // (L * L)
        for t in (((x + 1) % L) + L * (y + L * z), x + L * (((y + 1) % L) + L * z), x + L * (y + L * ((z + 1) % L))):
            H[4 * s:4 * s + 4, 4 * t:4 * t + 4] += T
            H[4 * t:4 * t + 4, 4 * s:4 * s + 4] += T.T
    return H, G.dense_lattice(L, L, L, seed, 1)


def main_theta(edges, theta):
    """Same data for the parameterised lattice at theta (onsite levels -6, -6, 6, 6: a wide gap at
    half filling), keyed 'theta0_L<L>'."""
    path = os.path.join(HERE, "lattice_eigs.json")
    out = json.load(open(path)) if os.path.exists(path) else {}
    for L in edges:
        n = 4 * L ** 3
        m = n // 2
        H, S = theta_lattice(L, theta, SEED)
        w = sl.eigh(H, S, eigvals_only=True, subset_by_index=[0, m + 1], overwrite_a=True, overwrite_b=True)
        out[f"theta0_L{L}"] = {
            "L": L, "n": n, "m": m, "seed": SEED, "theta": [float(x) for x in theta],
            "sum_lowest_m": float(np.sum(w[:m])),
            "eig_m_minus_2_to_m_plus_1": [float(x) for x in w[m - 2:m + 2]],
            "gap_at_m": float(w[m] - w[m - 1]),
            "eta_mid": float(0.5 * (w[m - 1] + w[m])),
        }
        print("theta0", L, out[f"theta0_L{L}"], flush=True)
        json.dump(out, open(path, "w"), indent=1)


if __name__ == "__main__":
    # python make_lattice_eigs.py 4 8 16 | python make_lattice_eigs.py theta 4 16 (theta0 of the
    # default seed, from the library's generator omx_lattice_theta0)
    if sys.argv[1:2] == ["theta"]:
        sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "omm-bundle_amd"))
        import ommhip as oh  # noqa: E402

        main_theta([int(a) for a in sys.argv[2:]], oh.lattice_theta0(SEED))
    else:
        main([int(a) for a in sys.argv[1:]] or [4, 8])
