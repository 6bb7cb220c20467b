"""numpy restatement of the lattice generator (independent of the C++ one in
omm-bundle_amd/csrc/lattice.cpp): counter-based uniforms u = splitmix64(key(seed,stream) + idx) >> 11
scaled by 2^-53; H/S blocks of a periodic simple-cubic lattice with 4 orbitals per site."""
import numpy as np

M64 = (1 << 64) - 1


def mix64(z):
    z = (z + 0x9E3779B97F4A7C15) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def u01(seed, stream, idx):
    key = mix64(seed ^ ((stream * 0xD1B54A32D192ED03) & M64))
    return (mix64((key + idx) & M64) >> 11) * 2.0 ** -53


def dense_lattice(Lx, Ly, Lz, seed, which):
    ns = Lx * Ly * Lz
    nb = 4 * ns
    A = np.zeros((nb, nb))
    for s in range(ns):
        if which == 0:
            R = np.array([[-0.1 + 0.2 * u01(seed, 1, 16 * s + 4 * r + q) for q in range(4)] for r in range(4)])
            A[4 * s:4 * s + 4, 4 * s:4 * s + 4] += np.diag([-5.0, 1, 1, 1]) + 0.5 * (R + R.T)
        else:
            A[4 * s:4 * s + 4, 4 * s:4 * s + 4] += np.eye(4)
        x, y, z = s % Lx, (s // Lx) % Ly, s // (Lx * Ly)
        for d in range(3):
            nx, ny, nz = x, y, z
            if d == 0:
                nx = (x + 1) % Lx
            elif d == 1:
                ny = (y + 1) % Ly
            else:
                nz = (z + 1) % Lz
            t = nx + Lx * (ny + Ly * nz)
            T = np.zeros((4, 4))
            for r in range(4):
                for q in range(4):
                    idx = (s * 3 + d) * 16 + 4 * r + q
                    if which == 0:
                        T[r, q] = -1.0 + 2.0 * u01(seed, 2, idx)
                    else:
                        T[r, q] = 0.03 * (-1.0 + 2.0 * u01(seed, 3, idx))
            A[4 * s:4 * s + 4, 4 * t:4 * t + 4] += T
            A[4 * t:4 * t + 4, 4 * s:4 * s + 4] += T.T
    return A
