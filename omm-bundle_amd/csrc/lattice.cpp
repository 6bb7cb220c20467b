// Synthetic tight-binding inputs for the OMM benchmarks: periodic simple-cubic lattices with
// 4 orbitals per site, block-sparse H and S (one 4x4 block per site and per nearest-neighbour
// bond), a random initial coefficient matrix C0, and Bloch-phase k-point variants.
//
// Site s = x + Lx*(y + Ly*z); orbital index 4*s + a.  Blocks are stored as BSR-4 by block
// column (block-CSC): bcolptr[J]..bcolptr[J+1] index the stored blocks of block column J, with
// block-row indices ascending and 16 values per block in column-major order
// (val[r + 4*c] = A(4I+r, 4J+c)).
//
// Random numbers come from a counter-based generator (splitmix64 finaliser applied to
// seed-derived keys), so every entry is a pure function of (seed, stream, index) and any
// subset can be regenerated independently on the host or in tests.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>
#include <vector>

#include "omm_hip.h"

namespace {

inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

inline double u01(uint64_t seed, uint64_t stream, uint64_t idx) {
    uint64_t key = mix64(seed ^ (stream * 0xD1B54A32D192ED03ULL));
    return (double)(mix64(key + idx) >> 11) * 0x1.0p-53;
}

inline double uab(uint64_t seed, uint64_t stream, uint64_t idx, double a, double b) {
    return a + (b - a) * u01(seed, stream, idx);
}

struct Contrib {
    int64_t I, J;   // block row, block column
    double v[16];   // column-major 4x4
};

// Neighbour of site s one step along direction d (periodic).
inline int64_t neighbour(int64_t s, int d, int Lx, int Ly, int Lz) {
    int64_t x = s % Lx, y = (s / Lx) % Ly, z = s / ((int64_t)Lx * Ly);
    if (d == 0) x = (x + 1) % Lx;
    if (d == 1) y = (y + 1) % Ly;
    if (d == 2) z = (z + 1) % Lz;
    return x + (int64_t)Lx * (y + (int64_t)Ly * z);
}

// Builds the block contributions of H (which=0), S (which=1) or of the shared-parameter
// Hamiltonian H(theta) (which=2: theta[0..15] = hopping block, column-major; theta[16..19] =
// onsite levels), then merges duplicates (small lattices where +d and -d bonds coincide).
std::vector<Contrib> build(int Lx, int Ly, int Lz, uint64_t seed, int which, const double *theta) {
    const int64_t ns = (int64_t)Lx * Ly * Lz;
    std::vector<Contrib> out;
    out.reserve((size_t)ns * 7);
    for (int64_t s = 0; s < ns; ++s) {
        Contrib c{};
        c.I = c.J = s;
        if (which == 0) {
            double A[16];
            for (int k = 0; k < 16; ++k) A[k] = uab(seed, OMX_STREAM_ONSITE, (uint64_t)s * 16 + k, -0.1, 0.1);
            for (int r = 0; r < 4; ++r)
                for (int q = 0; q < 4; ++q) {
                    // A[4*r+q] is A(r,q); symmetric part (A+A^T)/2
                    double sym = 0.5 * (A[4 * r + q] + A[4 * q + r]);
                    double diag = (r == q) ? (r == 0 ? -5.0 : 1.0) : 0.0;
                    c.v[r + 4 * q] = diag + sym;
                }
        } else if (which == 1) {
            for (int r = 0; r < 4; ++r) c.v[r + 4 * r] = 1.0;
        } else {
            for (int r = 0; r < 4; ++r) c.v[r + 4 * r] = theta[16 + r];
        }
        out.push_back(c);
        for (int d = 0; d < 3; ++d) {
            int64_t t = neighbour(s, d, Lx, Ly, Lz);
            double T[16];  // T[r + 4*q] = block (s -> t) entry (r, q)
            for (int r = 0; r < 4; ++r)
                for (int q = 0; q < 4; ++q) {
                    uint64_t idx = ((uint64_t)s * 3 + d) * 16 + 4 * r + q;
                    double v;
                    if (which == 0)
                        v = uab(seed, OMX_STREAM_HOP, idx, -1.0, 1.0);
                    else if (which == 1)
                        v = 0.03 * uab(seed, OMX_STREAM_OVL, idx, -1.0, 1.0);
                    else
                        v = theta[r + 4 * q];
                    T[r + 4 * q] = v;
                }
            Contrib f{}, b{};
            f.I = s; f.J = t;
            b.I = t; b.J = s;
            for (int r = 0; r < 4; ++r)
                for (int q = 0; q < 4; ++q) {
                    f.v[r + 4 * q] = T[r + 4 * q];
                    b.v[q + 4 * r] = T[r + 4 * q];
                }
            out.push_back(f);
            out.push_back(b);
        }
    }
    std::stable_sort(out.begin(), out.end(), [](const Contrib &a, const Contrib &b) {
        return a.J != b.J ? a.J < b.J : a.I < b.I;
    });
    std::vector<Contrib> merged;
    merged.reserve(out.size());
    for (const Contrib &c : out) {
        if (!merged.empty() && merged.back().I == c.I && merged.back().J == c.J) {
            for (int k = 0; k < 16; ++k) merged.back().v[k] += c.v[k];
        } else {
            merged.push_back(c);
        }
    }
    return merged;
}

// Bloch operator H(k) (which = 0) or S(k) (which = 1) of the periodic Lx x Ly x Lz cell: the
// same blocks as build(), with every bond that crosses the cell boundary in direction d
// multiplied by the phase exp(2 pi i phi_d) (forward block) and its conjugate (backward block),
// i.e. H(k) = sum_R H_R exp(i k.R) with k.R = 2 pi phi_d for a translation by one cell along d.
// Values are complex, (re, im) interleaved: 32 doubles per block.
struct ContribZ {
    int64_t I, J;
    double v[32];
};

std::vector<ContribZ> build_k(int Lx, int Ly, int Lz, uint64_t seed, int which, const double *phi) {
    const int64_t ns = (int64_t)Lx * Ly * Lz;
    const int L[3] = {Lx, Ly, Lz};
    std::vector<ContribZ> out;
    out.reserve((size_t)ns * 7);
    for (int64_t s = 0; s < ns; ++s) {
        ContribZ c{};
        c.I = c.J = s;
        if (which == 0) {
            double A[16];
            for (int k = 0; k < 16; ++k) A[k] = uab(seed, OMX_STREAM_ONSITE, (uint64_t)s * 16 + k, -0.1, 0.1);
            for (int r = 0; r < 4; ++r)
                for (int q = 0; q < 4; ++q) {
                    double sym = 0.5 * (A[4 * r + q] + A[4 * q + r]);
                    double diag = (r == q) ? (r == 0 ? -5.0 : 1.0) : 0.0;
                    c.v[2 * (r + 4 * q)] = diag + sym;
                }
        } else {
            for (int r = 0; r < 4; ++r) c.v[2 * (r + 4 * r)] = 1.0;
        }
        out.push_back(c);
        const int64_t coord[3] = {s % Lx, (s / Lx) % Ly, s / ((int64_t)Lx * Ly)};
        for (int d = 0; d < 3; ++d) {
            const int64_t t = neighbour(s, d, Lx, Ly, Lz);
            const bool wraps = (coord[d] == L[d] - 1);
            const double ang = wraps ? 2.0 * M_PI * phi[d] : 0.0;
            const double cr = std::cos(ang), ci = std::sin(ang);
            ContribZ f{}, b{};
            f.I = s;
            f.J = t;
            b.I = t;
            b.J = s;
            for (int r = 0; r < 4; ++r)
                for (int q = 0; q < 4; ++q) {
                    const uint64_t idx = ((uint64_t)s * 3 + d) * 16 + 4 * r + q;
                    const double v = (which == 0) ? uab(seed, OMX_STREAM_HOP, idx, -1.0, 1.0)
                                                  : 0.03 * uab(seed, OMX_STREAM_OVL, idx, -1.0, 1.0);
                    // forward block (s -> t): T e^{i ang}; backward block (t -> s): T^T e^{-i ang}
                    f.v[2 * (r + 4 * q)] = v * cr;
                    f.v[2 * (r + 4 * q) + 1] = v * ci;
                    b.v[2 * (q + 4 * r)] = v * cr;
                    b.v[2 * (q + 4 * r) + 1] = -v * ci;
                }
            out.push_back(f);
            out.push_back(b);
        }
    }
    std::stable_sort(out.begin(), out.end(), [](const ContribZ &a, const ContribZ &b) {
        return a.J != b.J ? a.J < b.J : a.I < b.I;
    });
    std::vector<ContribZ> merged;
    merged.reserve(out.size());
    for (const ContribZ &c : out) {
        if (!merged.empty() && merged.back().I == c.I && merged.back().J == c.J) {
            for (int k = 0; k < 32; ++k) merged.back().v[k] += c.v[k];
        } else {
            merged.push_back(c);
        }
    }
    return merged;
}

int fill(const std::vector<Contrib> &blocks, int64_t nbc, int64_t *nnzb, int *bcolptr, int *browidx, double *bval) {
    *nnzb = (int64_t)blocks.size();
    if (!bcolptr) return 0;
    std::fill(bcolptr, bcolptr + nbc + 1, 0);
    for (const Contrib &c : blocks) bcolptr[c.J + 1]++;
    for (int64_t j = 0; j < nbc; ++j) bcolptr[j + 1] += bcolptr[j];
    for (size_t l = 0; l < blocks.size(); ++l) {
        browidx[l] = (int)blocks[l].I;
        std::memcpy(bval + 16 * l, blocks[l].v, sizeof(double) * 16);
    }
    return 0;
}

}  // namespace

extern "C" {

int omx_lattice_bsr(int Lx, int Ly, int Lz, uint64_t seed, int which, int64_t *nnzb, int *bcolptr,
                    int *browidx, double *bval) {
    if (Lx < 1 || Ly < 1 || Lz < 1 || (which != 0 && which != 1) || !nnzb) return OMX_ERR_ARG;
    auto blocks = build(Lx, Ly, Lz, seed, which, nullptr);
    return fill(blocks, (int64_t)Lx * Ly * Lz, nnzb, bcolptr, browidx, bval);
}

// Jacobian of H(theta) (linear, no constant term): column p is H(e_p) on the same block pattern,
// listed as its nonzero stored entries (index 16*l + r + 4*c into the block values).
int omx_lattice_theta_jacobian(int Lx, int Ly, int Lz, int64_t *nent, int64_t *seg, int64_t *ent, double *coef) {
    if (Lx < 1 || Ly < 1 || Lz < 1 || !nent) return OMX_ERR_ARG;
    int64_t k = 0;
    if (seg) seg[0] = 0;
    for (int p = 0; p < OMX_THETA_COUNT; ++p) {
        double th[OMX_THETA_COUNT] = {0};
        th[p] = 1.0;
        auto blocks = build(Lx, Ly, Lz, 0, 2, th);
        for (size_t l = 0; l < blocks.size(); ++l)
            for (int q = 0; q < 16; ++q)
                if (blocks[l].v[q] != 0.0) {
                    if (ent) {
                        ent[k] = (int64_t)l * 16 + q;
                        coef[k] = blocks[l].v[q];
                    }
                    ++k;
                }
        if (seg) seg[p + 1] = k;
    }
    *nent = k;
    return 0;
}

int omx_lattice_bsr_k(int Lx, int Ly, int Lz, uint64_t seed, int which, const double *phi, int64_t *nnzb,
                      int *bcolptr, int *browidx, double *bval) {
    if (Lx < 1 || Ly < 1 || Lz < 1 || (which != 0 && which != 1) || !phi || !nnzb) return OMX_ERR_ARG;
    auto blocks = build_k(Lx, Ly, Lz, seed, which, phi);
    *nnzb = (int64_t)blocks.size();
    if (!bcolptr) return 0;
    const int64_t nbc = (int64_t)Lx * Ly * Lz;
    std::fill(bcolptr, bcolptr + nbc + 1, 0);
    for (const ContribZ &c : blocks) bcolptr[c.J + 1]++;
    for (int64_t j = 0; j < nbc; ++j) bcolptr[j + 1] += bcolptr[j];
    for (size_t l = 0; l < blocks.size(); ++l) {
        browidx[l] = (int)blocks[l].I;
        std::memcpy(bval + 32 * l, blocks[l].v, sizeof(double) * 32);
    }
    return 0;
}

int omx_lattice_bsr_theta(int Lx, int Ly, int Lz, const double *theta, int64_t *nnzb, int *bcolptr,
                          int *browidx, double *bval) {
    if (Lx < 1 || Ly < 1 || Lz < 1 || !theta || !nnzb) return OMX_ERR_ARG;
    auto blocks = build(Lx, Ly, Lz, 0, 2, theta);
    return fill(blocks, (int64_t)Lx * Ly * Lz, nnzb, bcolptr, browidx, bval);
}

int omx_lattice_theta0(uint64_t seed, double *theta) {
    if (!theta) return OMX_ERR_ARG;
    for (int k = 0; k < 16; ++k) theta[k] = uab(seed, OMX_STREAM_THETA, (uint64_t)k, -1.0, 1.0);
    // two low and two high onsite levels: the half-filled system has a gap of about 3.8 (with
    // one low level, e.g. (-5, 1, 1, 1), half filling lands inside a band)
    theta[16] = theta[17] = -6.0;
    theta[18] = theta[19] = 6.0;
    return 0;
}

// Patch walk over the sites: tx x ty x tz patches listed one after another (inside a patch z, y,
// x with x fastest), patches swept through z inside each x-y column of patches.  Groups of
// tx*ty*tz consecutive entries are then compact 3D patches, whose neighbour blocks overlap most.
int omx_lattice_walk3(int Lx, int Ly, int Lz, int tx, int ty, int tz, int *order) {
    if (Lx < 1 || Ly < 1 || Lz < 1 || tx < 1 || ty < 1 || tz < 1 || !order) return OMX_ERR_ARG;
    int64_t k = 0;
    for (int y0 = 0; y0 < Ly; y0 += ty)
        for (int x0 = 0; x0 < Lx; x0 += tx)
            for (int z0 = 0; z0 < Lz; z0 += tz)
                for (int z = z0; z < std::min(Lz, z0 + tz); ++z)
                    for (int y = y0; y < std::min(Ly, y0 + ty); ++y)
                        for (int x = x0; x < std::min(Lx, x0 + tx); ++x) order[k++] = x + Lx * (y + Ly * z);
    return 0;
}

int omx_random_c0(int no, int nb, uint64_t seed, double *C, int64_t ldc) {
    if (no < 0 || nb < 1 || !C || ldc < no) return OMX_ERR_ARG;
    const double scale = 1.0 / std::sqrt((double)nb);
    for (int64_t k = 0; k < nb; ++k)
        for (int64_t i = 0; i < no; ++i)
            C[i + k * ldc] = (u01(seed, OMX_STREAM_C0, (uint64_t)(i + (int64_t)no * k)) - 0.5) * scale;
    return 0;
}

double omx_u01(uint64_t seed, uint64_t stream, uint64_t idx) { return u01(seed, stream, idx); }

// Tiled walk over the sites: x-y tiles of tx x ty sites taken in row-major tile order, each
// swept through all z (and within one z-plane row-major).  A site's +-x and +-y neighbours inside
// the tile come within tx*ty positions of it and its +-z neighbours exactly tx*ty positions away,
// instead of Lx*Ly in the natural order.
int omx_lattice_walk(int Lx, int Ly, int Lz, int tx, int ty, int *order) {
    if (Lx < 1 || Ly < 1 || Lz < 1 || tx < 1 || ty < 1 || !order) return OMX_ERR_ARG;
    int64_t k = 0;
    for (int y0 = 0; y0 < Ly; y0 += ty)
        for (int x0 = 0; x0 < Lx; x0 += tx)
            for (int z = 0; z < Lz; ++z)
                for (int y = y0; y < std::min(Ly, y0 + ty); ++y)
                    for (int x = x0; x < std::min(Lx, x0 + tx); ++x) order[k++] = x + Lx * (y + Ly * z);
    return 0;
}

}  // extern "C"
