// Consumers of the vertex rows on the device: the linear tetrahedron delta
// function (src/delta.f90:61-225) on the tetrahedra of src/delta.f90:494-646,
// or the triangle method (648-882); the transition probabilities X+/- of
// calculate_eph_interaction_ibzk key 'X' (src/interactions.f90:2278-2354);
// the e-ph RTA scattering rates of calculate_el_rta_rates
// (src/interactions.f90:2957-3063); and the electron RTA transport
// coefficients (src/bte.f90:912-1011, src/bz_sums.f90:484-611).
//
// Layout in HBM:
//   simplices   tcount[nk], tmap[nk][24] = 4 (it - 1) + (iv - 1) (it the
//               tetrahedron or triangle, iv the wave vector's vertex label in it),
//               tev[ntet][nb] the vertex energies sorted (32 bytes; a triangle's
//               highest energy repeated);
//   delta table per (wave vector, band): a 128-byte header (energy range, piece
//               count, search starts of 64 uniform bins) and up to 96 pieces of 64
//               bytes (bounds, cubic coefficients of the vertex-weight sum);
//   processes   four per thread of k_ibzk_X, consecutive in a g2 row: each reads
//               its |g|^2 (8 B), writes X+ and X- (16 B) and optionally the final
//               electron and phonon state (16 B).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "ctx.h"
#include "kernels.h"

using namespace elph;

constexpr int kMaxPieces = 96;
constexpr int kBins = 64;  // uniform bins over a (point, band)'s energy range

// Table header of one (point, band), one 128-byte line: the energy range of
// the point's tetrahedra, the number of pieces m and, per uniform bin of the
// range, the piece holding the bin's left edge (where the search starts).
struct alignas(128) DeltaHead {
    double lo, hi, binscale;  // binscale = kBins / (hi - lo)
    int m, pad;
    unsigned char bins[kBins];
};
// One piece [lo, hi] of a (point, band): the weight sum there is
// c0 + c1 x + c2 x^2 + c3 x^3 with x = (e - mid) * ih.
struct alignas(64) DeltaPiece {
    double lo, hi, mid, ih, c0, c1, c2, c3;
};

struct elph_tetra {
    elph_ctx *ctx = nullptr;
    int64_t nk = 0, nb = 0, ntet = 0, mesh[3] = {0, 0, 0};
    int nvert = 4;  // 4: tetrahedra, 3: triangles
    // host copies in the reference's layouts (tetra (ntet,4), tetracount (nk),
    // tetramap (2,nk,24), tetra_evals (ntet,nb,4))
    std::vector<int64_t> tetra, tetracount, tetramap;
    std::vector<double> tetra_evals;
    DBuf<int> tcount, tmap;
    DBuf<d4> tev;
    // tabulated delta function (see k_delta_table)
    DBuf<DeltaHead> dhead;
    DBuf<DeltaPiece> dpiece;
};

// Static inputs of the electron RTA transport on the device.
struct elph_el_transport {
    elph_ctx *ctx = nullptr;
    int64_t nk = 0, nb = 0, nk_ibz = 0, deg = 0, mesh[3] = {0, 0, 0};
    double volume = 0;
    std::vector<double> crot;  // (3,3,nrot) Cartesian rotations, host
    bool symmetrize_k = false;
    DBuf<double> ens;          // [nk][nb]
    DBuf<double> vels;         // [nk][nb][3]
    DBuf<double> sym;          // [nk][9] symmetrizer (column-major 3x3)
    DBuf<int> src;             // [nk] IBZ point whose rate feeds the list point (-1: none)
    DBuf<double> part, sums;   // block partials [nb][nblk][36], totals [nb][36]
};

namespace {

// Constants of src/params.f90:27-43.
constexpr double kQe = 1.602176634e-19;
constexpr double kHbar = 1.05457172647e-22;
constexpr double kBoltz = 1.380649e-23 / kQe;  // eV/K

// Weight of vertex label iv (1..4) of a tetrahedron with sorted energies
// e1 <= e2 <= e3 <= e4 at energy e, for e1 <= e <= e4: the case expressions of
// src/delta.f90:118-210, with their degenerate-energy replacements and the
// equal-energies value 1/4.  Evaluated without contraction so the rounding
// follows the expressions as written.
__device__ __noinline__ double tetra_weight(int iv, double e, double e1, double e2, double e3, double e4) {
#pragma clang fp contract(off)
    const double e1e = e1 - e, e2e = e2 - e, e3e = e3 - e, e4e = e4 - e;
    const double e21 = e2 - e1, e31 = e3 - e1, e41 = e4 - e1, e32 = e3 - e2, e42 = e4 - e2, e43 = e4 - e3;
    const bool c1 = e1 <= e && e <= e2, c2 = e2 <= e && e <= e3, c3 = e3 <= e && e <= e4;
    double w = 0.0;
    if (iv == 1) {
        if (c1) {
            w = e1 == e2 ? 0.0 : (e2e / e21 + e3e / e31 + e4e / e41) * (e1e * e1e) / e41 / e31 / e21;
        } else if (c2) {
            if (e2 == e3)
                w = -0.5 * (e4e * e1e / e41 / e42 + e1e / e41 + e4e / (e41 * e41) * (e4e * e1e / e42 / e31 + e4e / e42 + e1e / e31));
            else
                w = -0.5 * (e3e / (e31 * e31) * (e3e * e2e / e42 / e32 + e4e * e1e / e41 / e42 + e3e * e1e / e32 / e41) +
                            e4e / (e41 * e41) * (e4e * e1e / e42 / e31 + e4e * e2e / e42 / e32 + e3e * e1e / e31 / e32));
        } else if (c3) {
            w = e3 == e4 ? (e4e * e4e) / (e41 * e41) / e42 : (e4e * e4e * e4e) / (e41 * e41) / e42 / e43;
        }
    } else if (iv == 2) {
        if (c1) {
            w = e1 == e2 ? 0.0 : -(e1e * e1e * e1e) / (e21 * e21) / e31 / e41;
        } else if (c2) {
            if (e2 == e3)
                w = -0.5 * (0.0 + e4e / e42 / e41 + 0.0 + e4e / (e42 * e42) * (0.0 + e4e * e1e / e41 / e31 + 1.0));
            else
                w = -0.5 * (e3e / (e32 * e32) * (e3e * e2e / e42 / e31 + e4e * e2e / e42 / e41 + e3e * e1e / e31 / e41) +
                            e4e / (e42 * e42) * (e3e * e2e / e32 / e31 + e4e * e1e / e41 / e31 + e4e * e2e / e32 / e41));
        } else if (c3) {
            w = e3 == e4 ? 0.0 : (e4e * e4e * e4e) / e41 / (e42 * e42) / e43;
        }
    } else if (iv == 3) {
        if (c1) {
            w = e1 == e2 ? 0.0 : -(e1e * e1e * e1e) / e21 / (e31 * e31) / e41;
        } else if (c2) {
            if (e2 == e3)
                w = 0.5 * (0.0 + e4e / e42 / e41 + e1e / e31 / e41 + e1e / (e31 * e31) * (0.0 + e4e * e1e / e41 / e42 + e1e / e41));
            else
                w = 0.5 * (e2e / (e32 * e32) * (e3e * e2e / e42 / e31 + e4e * e2e / e42 / e41 + e3e * e1e / e31 / e41) +
                           e1e / (e31 * e31) * (e3e * e2e / e42 / e32 + e4e * e1e / e41 / e42 + e3e * e1e / e32 / e41));
        } else if (c3) {
            w = e3 == e4 ? 0.0 : (e4e * e4e * e4e) / e41 / e42 / (e43 * e43);
        }
    } else {
        if (c1) {
            w = e1 == e2 ? 0.0 : -(e1e * e1e * e1e) / e21 / e31 / (e41 * e41);
        } else if (c2) {
            if (e2 == e3)
                w = 0.5 * (0.0 + e1e / (e41 * e41) * (e4e * e1e / e42 / e31 + e4e / e42 + e1e / e31));
            else
                w = 0.5 * (e2e / (e42 * e42) * (e3e * e2e / e32 / e31 + e4e * e1e / e41 / e31 + e4e * e2e / e32 / e41) +
                           e1e / (e41 * e41) * (e4e * e1e / e42 / e31 + e4e * e2e / e42 / e32 + e3e * e1e / e31 / e32));
        } else if (c3) {
            w = e3 == e4 ? 0.0 : -(e3e / e43 + e2e / e42 + e1e / e41) * (e4e * e4e) / e41 / e42 / e43;
        }
    }
    if (e1 == e2 && e1 == e3 && e1 == e4 && e == e1) w = 0.25;
    return w;
}

// Weight of vertex label iv (1..3) of a triangle with sorted energies
// e1 <= e2 <= e3 at energy e, e1 < e <= e3: the case expressions of
// src/delta.f90:842-871 (triangle method of Kurganskii et al.).
__device__ __noinline__ double triang_weight(int iv, double e, double e1, double e2, double e3) {
#pragma clang fp contract(off)
    const bool c2 = e1 < e && e <= e2, c3 = e2 < e && e <= e3;
    const double E12 = (e - e2) / (e1 - e2), E21 = (e - e1) / (e2 - e1), E13 = (e - e3) / (e1 - e3);
    const double E31 = (e - e1) / (e3 - e1), E23 = (e - e3) / (e2 - e3), E32 = (e - e2) / (e3 - e2);
    double w = 0.0;
    if (iv == 1) {
        if (c2) w = E21 * (E12 + E13) / (e3 - e1);
        else if (c3) w = E23 * E13 / (e3 - e1);
    } else if (iv == 2) {
        if (c2) w = E21 * E21 / (e3 - e1);
        else if (c3) w = E23 * E23 / (e3 - e1);
    } else {
        if (c2) w = E21 * E31 / (e3 - e1);
        else if (c3) w = E23 * (E31 + E32) / (e3 - e1);
    }
    return w;
}

// Intervals per (list point, band) of the tabulated delta function: the
// distinct vertex energies of a point's (at most 24) tetrahedra split the
// energy axis into at most 95 pieces, on each of which every vertex weight is
// one cubic polynomial of e.
struct TetraView {
    const int *tcount, *tmap;
    const d4 *tev;
    int nb;
    bool tri;         // triangles (delta_fn_triang) instead of tetrahedra
    double numtetra;  // number of simplices of the mesh
    const DeltaHead *head;    // [pair]
    const DeltaPiece *piece;  // [pair][kMaxPieces]
};

// Sum over the tetrahedra of list point r of the weight of r's vertex at
// energy e, band n: the loop of src/delta.f90:92-216 before its clamp and
// normalization.
__device__ double weight_sum(const TetraView &t, double e, int r, int n) {
    double d = 0.0;
    const int num = t.tcount[r];
    for (int j = 0; j < num; ++j) {
        const int code = t.tmap[r * 24 + j];
        const d4 v = t.tev[(size_t)(code >> 2) * t.nb + n];
        if (t.tri) {
            if (e <= v.x || e > v.z) continue;
            d += triang_weight((code & 3) + 1, e, v.x, v.y, v.z);
        } else {
            if (e < v.x || e > v.w) continue;
            d += tetra_weight((code & 3) + 1, e, v.x, v.y, v.z, v.w);
        }
    }
    return d;
}

// Weight sum at e in [lo, hi] of a pair with m > 0 pieces from the table:
// start at e's bin, step to the piece holding e (correct from any start),
// evaluate its cubic.  False when e is a piece bound (a vertex energy): the
// sum can jump there (a tetrahedron with two equal vertex energies takes its
// degenerate-case value at exactly that energy), so the caller sums the
// tetrahedra directly.
__device__ __forceinline__ bool table_sum(const DeltaHead *h, const DeltaPiece *pc, int m, double lo, double scale,
                                          double e, double &d) {
    int b = (int)((e - lo) * scale);
    b = max(0, min(kBins - 1, b));
    int j = h->bins[b];
    DeltaPiece p = pc[j];
    while (j > 0 && e < p.lo) p = pc[--j];
    while (j < m - 1 && e >= p.hi) p = pc[++j];
    if (e == p.lo || e == p.hi) return false;
    const double x = (e - p.mid) * p.ih;
    d = ((p.c3 * x + p.c2) * x + p.c1) * x + p.c0;
    return true;
}

// delta_fn_tetra(e, r + 1, n + 1) (src/delta.f90:61-225) from the table: the
// weight sum, zero below 1e-12, divided by 6 prod(mesh); zero outside the
// energy range of r's tetrahedra.  Points whose tetrahedra all have one
// energy (no piece) and energies at a piece bound are summed directly.
__device__ __forceinline__ double delta_tetra(const TetraView &t, double e, int r, int n) {
    const size_t pr = (size_t)r * t.nb + n;
    const DeltaHead *h = t.head + pr;
    const double lo = h->lo, hi = h->hi;
    if (e < lo || e > hi) return 0.0;
    const int m = h->m;
    double d;
    if (!(m > 0 && table_sum(h, t.piece + pr * kMaxPieces, m, lo, h->binscale, e, d))) d = weight_sum(t, e, r, n);
    return d < 1.0e-12 ? 0.0 : d / t.numtetra;
}

struct ChebFit {
    double minv[16];  // inverse Vandermonde of the nodes, row-major
    double node[4];   // cos((2k+1) pi / 8)
};

// Builds the table, one thread per (point, band): the distinct vertex
// energies in ascending order, and on each piece [a, b] the cubic through the
// exact weight sums at the four Chebyshev nodes of the piece.
__global__ __launch_bounds__(256) void k_delta_table(TetraView t, int npairs, ChebFit f, DeltaHead *head,
                                                     DeltaPiece *piece) {
    const int pr = blockIdx.x * 256 + threadIdx.x;
    if (pr >= npairs) return;
    const int r = pr / t.nb, n = pr - r * t.nb;
    double b[kMaxPieces];
    int nbp = 0;
    const int num = t.tcount[r];
    for (int j = 0; j < num; ++j) {
        const d4 v = t.tev[(size_t)(t.tmap[r * 24 + j] >> 2) * t.nb + n];
        const double ev[4] = {v.x, v.y, v.z, v.w};
        for (int i = 0; i < 4; ++i) {
            // insertion into the ascending list of distinct values
            const double x = ev[i];
            int k = nbp;
            bool dup = false;
            while (k > 0 && b[k - 1] >= x) {
                if (b[k - 1] == x) dup = true;
                --k;
            }
            if (dup) continue;
            for (int q = nbp; q > k; --q) b[q] = b[q - 1];
            b[k] = x;
            ++nbp;
        }
    }
    const int m = nbp > 1 ? nbp - 1 : 0;
    DeltaHead h;
    h.lo = nbp > 0 ? b[0] : 0.0;
    h.hi = nbp > 0 ? b[nbp - 1] : -1.0;  // no tetrahedra: an empty range
    h.binscale = m > 0 ? (double)kBins / (h.hi - h.lo) : 0.0;
    h.m = m;
    h.pad = 0;
    for (int k = 0; k < kBins; ++k) {
        int j = 0;
        if (m > 0) {
            const double edge = h.lo + (h.hi - h.lo) * ((double)k / (double)kBins);
            while (j < m - 1 && b[j + 1] <= edge) ++j;
        }
        h.bins[k] = (unsigned char)j;
    }
    head[pr] = h;
    for (int j = 0; j < m; ++j) {
        const double a = b[j], z = b[j + 1];
        const double mid = 0.5 * (a + z), half = 0.5 * (z - a);
        double sv[4];
        for (int k = 0; k < 4; ++k) sv[k] = weight_sum(t, mid + half * f.node[k], r, n);
        double c[4];
        for (int i = 0; i < 4; ++i)
            c[i] = f.minv[4 * i] * sv[0] + f.minv[4 * i + 1] * sv[1] + f.minv[4 * i + 2] * sv[2] + f.minv[4 * i + 3] * sv[3];
        piece[(size_t)pr * kMaxPieces + j] = DeltaPiece{a, z, mid, 1.0 / half, c[0], c[1], c[2], c[3]};
    }
}

// Inverse of V[k][i] = x_k^i for the four Chebyshev nodes (Gauss-Jordan with
// partial pivoting; V is well conditioned).
ChebFit cheb_fit() {
    ChebFit f;
    double a[4][8];
    for (int k = 0; k < 4; ++k) {
        f.node[k] = std::cos((2 * k + 1) * (4.0 * std::atan(1.0)) / 8.0);
        double p = 1.0;
        for (int i = 0; i < 4; ++i) {
            a[k][i] = p;
            a[k][4 + i] = i == k ? 1.0 : 0.0;
            p *= f.node[k];
        }
    }
    for (int c = 0; c < 4; ++c) {
        int piv = c;
        for (int r = c + 1; r < 4; ++r)
            if (std::fabs(a[r][c]) > std::fabs(a[piv][c])) piv = r;
        for (int j = 0; j < 8; ++j) std::swap(a[c][j], a[piv][j]);
        const double d = a[c][c];
        for (int j = 0; j < 8; ++j) a[c][j] /= d;
        for (int r = 0; r < 4; ++r)
            if (r != c) {
                const double g = a[r][c];
                for (int j = 0; j < 8; ++j) a[r][j] -= g * a[c][j];
            }
    }
    for (int i = 0; i < 4; ++i)
        for (int k = 0; k < 4; ++k) f.minv[4 * i + k] = a[i][4 + k];
    return f;
}

__global__ void k_delta_samples(TetraView t, int n, const double *e, const int64_t *ik, const int64_t *ib,
                                double *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = delta_tetra(t, e[i], (int)ik[i] - 1, (int)ib[i] - 1);
}

struct XArgs {
    TetraView t;
    int nb, nbr, mesh_ref, Nel0, Nel1, Nel2, Nph0, Nph1, Nph2;
    int64_t nprocs;
    const int2 *vmap;
    const int *kp_int;
    const int64_t *offs;
    const double *om, *omf;
    int nblk;       // blocks of kXPer x 256 processes per row
    int states_x;   // 1: blockIdx.x = state, blockIdx.y = block of the row; 0: the reverse
    int pairs16;    // 1: g2 and the outputs start 16-byte aligned (pairs of processes in 16-byte accesses)
    const int4 *st_k;      // per in-window state: k (electron mesh ints)
    const double *st_e;    // its energy
    const int *st_row;     // its row in g2 / X
    double T, chempot, cnst, ikT;  // ikT = 1 / (kB T)
    const double *bose;            // Bose factor of every coarse, then fine, mode
    size_t ncoarse;
    const double *g2;
    double *xp, *xm;
    int64_t *se, *sp;
    double *partial;  // [state][block] sums of X+ + X-
};

// Bose factor 1 / expm1(e / kB / T) (src/misc.f90:992-998) of n phonon
// energies (0 for e == 0, as src/interactions.f90:2287-2291 sets it).
__global__ void k_bose(int n, const double *en, double T, double *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double e = en[i];
    out[i] = e != 0.0 ? 1.0 / expm1(e / kBoltz / T) : 0.0;
}

constexpr int kXPer = 2;  // consecutive processes per thread of k_ibzk_X

// Processes of in-window states (the state's index among them and the block
// of its row from the grid, st_row its row in g2 / X).  With the states along
// x, consecutive workgroups are the same processes of consecutive states, so
// the tables of those k' rows (the delta-table heads, phonon energies and Bose
// factors, the process map) are read from L2 by all states instead of being
// streamed in full once per state.  Per process the loop body of
// src/interactions.f90:2278-2332, X scaled by 2 pi / hbar (2351-2354), and the
// block's sum of X+ + X- for the rates.  Block: 256 threads x kXPer
// consecutive processes of one row.  The few energies that fall inside a (k', n)
// tetrahedra range are gathered in LDS and their deltas evaluated by the
// whole block, so no wave runs the table search for one or two lanes.
// |g|^2 is read and X+-, and the state maps written, with nontemporal hints
// (each is touched once; -12% against cached accesses, with kXPer 2 rather
// than 4: -3%).  The kernel is bound by the latency of its dependent table
// reads (process map -> k' row -> phonon mode; delta-table head -> bin ->
// piece), so each step's loads of all of a thread's processes, and of its
// queued deltas, are issued together (addresses of processes past the row
// end clamped into it, their results unused).
__global__ __launch_bounds__(256) void k_ibzk_X(XArgs a) {
    constexpr int QMAX = 2 * kXPer * 256;
    __shared__ double qe[QMAX];  // queued energies, then their delta values
    __shared__ int qp[QMAX];     // (point, band) pair of each queued energy
    __shared__ int qn;
    const int ia = a.states_x ? blockIdx.x : blockIdx.y;
    const int64_t blk = a.states_x ? blockIdx.y : blockIdx.x;
    const int4 k = a.st_k[ia];
    const size_t row = (size_t)a.st_row[ia] * a.nprocs;
    const double en_el = a.st_e[ia];
    if (threadIdx.x == 0) qn = 0;
    __syncthreads();
    // 1. per process: phonon mode, occupations, state maps; the energies
    // e +- omega that fall inside their (k', n) tetrahedra range are queued
    // the thread's kXPer consecutive processes; with an even row length every
    // row starts 16-byte aligned and a pair is one 16-byte access per array
    const int64_t cb = (blk * 256 + threadIdx.x) * kXPer;
    const bool pair = kXPer == 2 && cb + 1 < a.nprocs;
    const bool vrow = pair && a.pairs16 && (a.nprocs & 1) == 0;
    int64_t cc[kXPer];
    bool live[kXPer];
    int2 vm[kXPer];
#pragma unroll
    for (int u = 0; u < kXPer; ++u) {
        const int64_t c = cb + u;
        live[u] = c < a.nprocs;
        cc[u] = live[u] ? c : a.nprocs - 1;
    }
    if (pair) {
        const v4i v = *reinterpret_cast<const v4i *>(a.vmap + cb);
        vm[0] = make_int2(v.x, v.y);
        vm[kXPer - 1] = make_int2(v.z, v.w);
    } else {
#pragma unroll
        for (int u = 0; u < kXPer; ++u) vm[u] = a.vmap[cc[u]];
    }
    int kr[kXPer][3];
#pragma unroll
    for (int u = 0; u < kXPer; ++u) {
        const int r = vm[u].x;
        kr[u][0] = a.kp_int[3 * r];
        kr[u][1] = a.kp_int[3 * r + 1];
        kr[u][2] = a.kp_int[3 * r + 2];
    }
    double en_ph[kXPer], bose[kXPer], gw[kXPer], lo[kXPer], hi[kXPer];
    int64_t sph[kXPer];
#pragma unroll
    for (int u = 0; u < kXPer; ++u) {
        const int r = vm[u].x, n = vm[u].y >> 8, s = vm[u].y & 255;
        // q = k' - k on the electron mesh (both components in [0, N))
        int q0 = kr[u][0] - k.x, q1 = kr[u][1] - k.y, q2 = kr[u][2] - k.z;
        q0 += q0 < 0 ? a.Nel0 : 0;
        q1 += q1 < 0 ? a.Nel1 : 0;
        q2 += q2 < 0 ? a.Nel2 : 0;
        // the mode's energy and Bose factor: [coarse modes][fine modes]
        const bool coarse =
            a.mesh_ref == 1 || (q0 % a.mesh_ref == 0 && q1 % a.mesh_ref == 0 && q2 % a.mesh_ref == 0);
        const int iqc = a.mesh_ref == 1 ? (q2 * a.Nph1 + q1) * a.Nph0 + q0
                                        : ((q2 / a.mesh_ref) * a.Nph1 + q1 / a.mesh_ref) * a.Nph0 + q0 / a.mesh_ref;
        const int iqf = (q2 * a.Nel1 + q1) * a.Nel0 + q0;
        const size_t mi = coarse ? (size_t)iqc * a.nbr + s : (size_t)iqf * a.nbr + s;
        en_ph[u] = (coarse ? a.om : a.omf)[mi];
        bose[u] = a.bose[coarse ? mi : a.ncoarse + mi];
        sph[u] = coarse ? (int64_t)iqc * a.nbr + s + 1 : -((int64_t)iqf * a.nbr + s + 1);
        if (!vrow) gw[u] = __builtin_nontemporal_load(a.g2 + row + cc[u]);
        const DeltaHead *h = a.t.head + (r * a.t.nb + n);
        lo[u] = h->lo;
        hi[u] = h->hi;
    }
    if (vrow) {
        const dv2 g = __builtin_nontemporal_load(reinterpret_cast<const dv2 *>(a.g2 + row + cb));
        gw[0] = g.x;
        gw[kXPer - 1] = g.y;
    }
    if (a.se) {
        int64_t sel[kXPer];
#pragma unroll
        for (int u = 0; u < kXPer; ++u) sel[u] = (int64_t)vm[u].x * a.nb + (vm[u].y >> 8) + 1;
        if (vrow) {
            typedef long long l2 __attribute__((ext_vector_type(2)));
            __builtin_nontemporal_store(l2{sel[0], sel[kXPer - 1]}, reinterpret_cast<l2 *>(a.se + row + cb));
            __builtin_nontemporal_store(l2{sph[0], sph[kXPer - 1]}, reinterpret_cast<l2 *>(a.sp + row + cb));
        } else {
#pragma unroll
            for (int u = 0; u < kXPer; ++u)
                if (live[u]) {
                    __builtin_nontemporal_store(sel[u], a.se + row + cc[u]);
                    __builtin_nontemporal_store(sph[u], a.sp + row + cc[u]);
                }
        }
    }
    double occp[kXPer], occm[kXPer];
    int slp[kXPer], slm[kXPer];
#pragma unroll
    for (int u = 0; u < kXPer; ++u) {
        slp[u] = slm[u] = -1;
        occp[u] = occm[u] = 0.0;
        if (!live[u]) continue;
        const int r = vm[u].x, n = vm[u].y >> 8;
        if (en_ph[u] >= 0.5e-3) {
            occp[u] = bose[u] + 1.0 / (exp((en_el + en_ph[u] - a.chempot) * a.ikT) + 1.0);
            occm[u] = 1.0 + bose[u] - 1.0 / (exp((en_el - en_ph[u] - a.chempot) * a.ikT) + 1.0);
            const int pr = r * a.t.nb + n;
            const double ep = en_el + en_ph[u], em = en_el - en_ph[u];
            if (!(ep < lo[u] || ep > hi[u])) {
                slp[u] = atomicAdd(&qn, 1);
                qe[slp[u]] = ep;
                qp[slp[u]] = pr;
            }
            if (!(em < lo[u] || em > hi[u])) {
                slm[u] = atomicAdd(&qn, 1);
                qe[slm[u]] = em;
                qp[slm[u]] = pr;
            }
        } else {
            gw[u] = 0.0;
        }
    }
    __syncthreads();
    // 2. the block's queued deltas (each e known to lie in its pair's range),
    // all lanes busy, two per thread and round with each step's loads issued
    // for both: the heads, the starting piece of each energy's bin, the piece
    // (then the rare steps to the piece holding e; delta_tetra's lookup in
    // that order)
    const int nq = qn;
    for (int base = 0; base < nq; base += 512) {
        constexpr int NV = 2;
        int pr[NV], m[NV], j[NV];
        double e[NV], hlo[NV], bs[NV];
        bool act[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const int i = base + threadIdx.x + 256 * v;
            act[v] = i < nq;
            const int ii = act[v] ? i : 0;
            e[v] = qe[ii];
            pr[v] = qp[ii];
            const DeltaHead *h = a.t.head + pr[v];
            m[v] = h->m;
            hlo[v] = h->lo;
            bs[v] = h->binscale;
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            int b = (int)((e[v] - hlo[v]) * bs[v]);
            b = max(0, min(kBins - 1, b));
            // bins of a pair without pieces are not used (and kept in range)
            j[v] = min((int)a.t.head[pr[v]].bins[b], kMaxPieces - 1);
        }
        double plo[NV], phi[NV], pmid[NV], pih[NV], c0[NV], c1[NV], c2[NV], c3[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const DeltaPiece q = a.t.piece[(size_t)pr[v] * kMaxPieces + j[v]];
            plo[v] = q.lo;
            phi[v] = q.hi;
            pmid[v] = q.mid;
            pih[v] = q.ih;
            c0[v] = q.c0;
            c1[v] = q.c1;
            c2[v] = q.c2;
            c3[v] = q.c3;
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            if (!act[v]) continue;
            double d;
            if (m[v] > 0) {
                const DeltaPiece *pc = a.t.piece + (size_t)pr[v] * kMaxPieces;
                int jj = j[v];
                auto take = [&](int t) {
                    const DeltaPiece q = pc[t];
                    plo[v] = q.lo;
                    phi[v] = q.hi;
                    pmid[v] = q.mid;
                    pih[v] = q.ih;
                    c0[v] = q.c0;
                    c1[v] = q.c1;
                    c2[v] = q.c2;
                    c3[v] = q.c3;
                };
                while (jj > 0 && e[v] < plo[v]) take(--jj);
                while (jj < m[v] - 1 && e[v] >= phi[v]) take(++jj);
                const double x = (e[v] - pmid[v]) * pih[v];
                d = e[v] == plo[v] || e[v] == phi[v] ? weight_sum(a.t, e[v], pr[v] / a.t.nb, pr[v] % a.t.nb)
                                                    : ((c3[v] * x + c2[v]) * x + c1[v]) * x + c0[v];
            } else {
                d = weight_sum(a.t, e[v], pr[v] / a.t.nb, pr[v] % a.t.nb);
            }
            qe[base + threadIdx.x + 256 * v] = d >= 1.0e-12 ? d / a.t.numtetra : 0.0;
        }
    }
    __syncthreads();
    // 3. X+- = 2 pi / hbar g2 occ delta
    double contrib = 0.0;
    double xpv[kXPer], xmv[kXPer];
#pragma unroll
    for (int u = 0; u < kXPer; ++u) {
        const double dp = slp[u] >= 0 ? qe[slp[u]] : 0.0, dm = slm[u] >= 0 ? qe[slm[u]] : 0.0;
        xpv[u] = a.cnst * (gw[u] * occp[u] * dp);
        xmv[u] = a.cnst * (gw[u] * occm[u] * dm);
        if (live[u]) contrib += xpv[u] + xmv[u];
    }
    if (a.xp) {
        if (vrow) {
            __builtin_nontemporal_store(dv2{xpv[0], xpv[kXPer - 1]}, reinterpret_cast<dv2 *>(a.xp + row + cb));
            __builtin_nontemporal_store(dv2{xmv[0], xmv[kXPer - 1]}, reinterpret_cast<dv2 *>(a.xm + row + cb));
        } else {
#pragma unroll
            for (int u = 0; u < kXPer; ++u)
                if (live[u]) {
                    __builtin_nontemporal_store(xpv[u], a.xp + row + cc[u]);
                    __builtin_nontemporal_store(xmv[u], a.xm + row + cc[u]);
                }
        }
    }
    // block sum (fixed order: wave shuffles, then the 4 waves in order)
    for (int off = 32; off > 0; off >>= 1) contrib += __shfl_down(contrib, off, 64);
    __shared__ double ws[4];
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = contrib;
    __syncthreads();
    if (threadIdx.x == 0) a.partial[(size_t)ia * a.nblk + blk] = ((ws[0] + ws[1]) + ws[2]) + ws[3];
}

// rates[t] = sum over the blocks of state t's partial sums, 0 for a state
// outside the window (st_idx[t] = -1); one block per state.
__global__ __launch_bounds__(256) void k_rates(const double *partial, int nblk, const int *st_idx, double *rates) {
    const int t = blockIdx.x, ia = st_idx[t];
    double v = 0.0;
    if (ia >= 0)
        for (int i = threadIdx.x; i < nblk; i += 256) v += partial[(size_t)ia * nblk + i];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    __shared__ double ws[4];
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) rates[t] = ((ws[0] + ws[1]) + ws[2]) + ws[3];
}

// Electron RTA transport sums (src/bte.f90:912-1011 field terms, then
// src/bz_sums.f90:559-583): for list point ik and band b with IBZ rate r,
//   F_E = qe v / r,  F_T = (1/T) v (e - mu) / r   (0 when r == 0 or no IBZ image),
// optionally F <- S_k F (the per-k symmetrizer, src/bte.f90:241-244, 262-265),
// and with d = f (1 - f) the 36 sums per band
//   hcE(j,l) += (e - mu) d v_j F_E,l   ccE(j,l) += d v_j F_E,l
//   hcT(j,l) += (e - mu) d v_j F_T,l   ccT(j,l) += d v_j F_T,l
// (grid: blocks x bands; each block writes its 36 partial sums).
__global__ __launch_bounds__(256) void k_el_transport(int nk, int nb, const double *ens, const double *vels,
                                                      const double *sym, const int *src, const double *rates, double T,
                                                      double chempot, double qe, double *part) {
    const int b = blockIdx.y;
    double acc[36];
#pragma unroll
    for (int i = 0; i < 36; ++i) acc[i] = 0.0;
    for (int ik = blockIdx.x * 256 + threadIdx.x; ik < nk; ik += gridDim.x * 256) {
        const int i = src[ik];
        const double r = i >= 0 ? rates[(size_t)i * nb + b] : 0.0;
        if (r == 0.0) continue;
        const double e = ens[(size_t)ik * nb + b];
        const double *v = vels + ((size_t)ik * nb + b) * 3;
        double fE[3], fT[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            fE[d] = qe * v[d] * 1.0 / r;
            fT[d] = 1.0 / T * v[d] * (e - chempot) / r;
        }
        if (sym) {
            const double *S = sym + (size_t)ik * 9;
            double gE[3], gT[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                gE[d] = S[d] * fE[0] + S[d + 3] * fE[1] + S[d + 6] * fE[2];
                gT[d] = S[d] * fT[0] + S[d + 3] * fT[1] + S[d + 6] * fT[2];
            }
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                fE[d] = gE[d];
                fT[d] = gT[d];
            }
        }
        double f = 1.0 / (exp((e - chempot) / kBoltz / T) + 1.0);
        f = f * (1.0 - f);
        const double w = e - chempot;
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int l = 0; l < 3; ++l) {
                acc[j + 3 * l] += w * f * v[j] * fE[l];
                acc[9 + j + 3 * l] += f * v[j] * fE[l];
                acc[18 + j + 3 * l] += w * f * v[j] * fT[l];
                acc[27 + j + 3 * l] += f * v[j] * fT[l];
            }
    }
    __shared__ double ws[4][36];
#pragma unroll
    for (int i = 0; i < 36; ++i) {
        double x = acc[i];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
        if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6][i] = x;
    }
    __syncthreads();
    if (threadIdx.x < 36)
        part[((size_t)b * gridDim.x + blockIdx.x) * 36 + threadIdx.x] =
            ((ws[0][threadIdx.x] + ws[1][threadIdx.x]) + ws[2][threadIdx.x]) + ws[3][threadIdx.x];
}

// sums[b][i] = sum over the blocks of part[b][blk][i] (one thread per (b, i)).
__global__ void k_el_transport_sum(int nb, int nblk, const double *part, double *sums) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nb * 36) return;
    const int b = t / 36, i = t % 36;
    double x = 0.0;
    for (int k = 0; k < nblk; ++k) x += part[((size_t)b * nblk + k) * 36 + i];
    sums[t] = x;
}

TetraView view_of(const elph_tetra *t) {
    TetraView v;
    v.tcount = t->tcount.p;
    v.tmap = t->tmap.p;
    v.tev = t->tev.p;
    v.head = t->dhead.p;
    v.piece = t->dpiece.p;
    v.nb = (int)t->nb;
    v.tri = t->nvert == 3;
    v.numtetra = (double)(t->mesh[0] * t->mesh[1] * t->mesh[2] * (t->nvert == 3 ? 2 : 6));
    return v;
}

}  // namespace

extern "C" {

namespace {
// form_tetrahedra_3d + fill_tetrahedra_3d (src/delta.f90:494-646), or with
// triangles form_triangles + fill_triangles (648-788): the simplices of each
// list point's subcell, the (simplex, vertex label) map of each point, the
// sorted vertex energies, then the delta table.
int create_simplices(elph_ctx *ctx, int64_t nk, const int64_t *mesh, const int64_t *indexlist, int64_t nb,
                     const double *ens, bool triangles, elph_tetra **out) {
    if (!ctx || !mesh || !ens || !out || nk < 1 || nb < 1) return ELPH_EINVAL;
    *out = nullptr;
    auto *t = new elph_tetra();
    try {
        HCHECK(hipSetDevice(ctx->device));
        for (int d = 0; d < 3; ++d) {
            if (mesh[d] < 1) throw Error{ELPH_EINVAL, "mesh components must be positive"};
            t->mesh[d] = mesh[d];
        }
        const int64_t nmesh = mesh[0] * mesh[1] * mesh[2];
        const bool blocks = indexlist != nullptr;
        if (!blocks && nk != nmesh) throw Error{ELPH_EINVAL, "without an index list nk must be the mesh size"};
        if (blocks) {
            for (int64_t i = 0; i < nk; ++i) {
                if (indexlist[i] < 1 || indexlist[i] > nmesh) throw Error{ELPH_EINVAL, "index list entry off the mesh"};
                if (i > 0 && indexlist[i] <= indexlist[i - 1])
                    throw Error{ELPH_EINVAL, "index list must be strictly ascending"};
            }
        }
        // vertex offsets of a subcell (src/delta.f90:565-573; the first four are
        // the triangles' subcell of 708-713), the six tetrahedra (522-529) and
        // the two triangles (675-679) as labels into them
        static const int corner[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0},
                                         {0, 0, 1}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
        static const int tlabel[6][4] = {{1, 2, 3, 6}, {1, 3, 5, 6}, {3, 5, 6, 7}, {3, 6, 7, 8}, {3, 4, 6, 8}, {2, 3, 4, 6}};
        static const int flabel[2][4] = {{1, 2, 3, 0}, {1, 3, 4, 0}};
        const int nper = triangles ? 2 : 6, nv = triangles ? 3 : 4, maxper = triangles ? 6 : 24;
        if (nper * nk >= (int64_t)1 << 28) throw Error{ELPH_EINVAL, "too many simplices"};
        t->ctx = ctx;
        t->nk = nk;
        t->nb = nb;
        t->nvert = nv;
        t->ntet = nper * nk;
        const int64_t ntet = t->ntet;
        t->tetra.assign(ntet * nv, 0);
        t->tetracount.assign(nk, 0);
        t->tetramap.assign(2 * nk * maxper, 0);
        std::vector<int> hcount(nk, 0), hmap(nk * 24, 0);
        for (int64_t p = 0; p < nk; ++p) {
            int64_t v[3];
            demux_vector0(blocks ? indexlist[p] : p + 1, v, mesh);
            for (int tk = 0; tk < nper; ++tk) {
                const int64_t it = nper * p + tk;  // 0-based simplex
                for (int tl = 0; tl < nv; ++tl) {
                    const int *cn = corner[(triangles ? flabel[tk][tl] : tlabel[tk][tl]) - 1];
                    int64_t w[3];
                    for (int d = 0; d < 3; ++d) w[d] = (v[d] + cn[d]) % mesh[d];
                    const int64_t muxed = (w[2] * mesh[1] + w[1]) * mesh[0] + w[0] + 1;
                    int64_t idx = muxed;
                    if (blocks) {
                        const int64_t *f = std::lower_bound(indexlist, indexlist + nk, muxed);
                        idx = (f != indexlist + nk && *f == muxed) ? (f - indexlist) + 1 : -1;
                    }
                    t->tetra[it + ntet * tl] = idx;
                    if (idx > 0) {
                        const int64_t c = t->tetracount[idx - 1]++;
                        if (c >= maxper) throw Error{ELPH_EINVAL, "too many simplices at a wave vector"};
                        t->tetramap[0 + 2 * ((idx - 1) + nk * c)] = it + 1;
                        t->tetramap[1 + 2 * ((idx - 1) + nk * c)] = tl + 1;
                        hmap[(idx - 1) * 24 + c] = (int)(4 * it + tl);
                        hcount[idx - 1] = (int)(c + 1);
                    }
                }
            }
        }
        // vertex energies (0 off the list), sorted per simplex and band
        // (src/delta.f90:607-646, 739-788); a triangle's highest energy is
        // repeated in the fourth slot
        t->tetra_evals.assign(ntet * nb * nv, 0.0);
        std::vector<d4> hev(ntet * nb);
        for (int64_t it = 0; it < ntet; ++it)
            for (int64_t b = 0; b < nb; ++b) {
                double e[4];
                for (int iv = 0; iv < nv; ++iv) {
                    const int64_t idx = t->tetra[it + ntet * iv];
                    e[iv] = idx > 0 ? ens[(idx - 1) + nk * b] : 0.0;
                }
                std::sort(e, e + nv);
                if (nv == 3) e[3] = e[2];
                for (int iv = 0; iv < nv; ++iv) t->tetra_evals[it + ntet * (b + nb * iv)] = e[iv];
                hev[it * nb + b] = d4{e[0], e[1], e[2], e[3]};
            }
        hipStream_t s = ctx->stream;
        t->tcount.upload(hcount, s);
        t->tmap.upload(hmap, s);
        t->tev.upload(hev, s);
        const int64_t npairs = nk * nb;
        t->dhead.alloc(npairs);
        t->dpiece.alloc(npairs * kMaxPieces);
        hipLaunchKernelGGL(k_delta_table, dim3((unsigned)((npairs + 255) / 256)), dim3(256), 0, s, view_of(t),
                           (int)npairs, cheb_fit(), t->dhead.p, t->dpiece.p);
        HCHECK(hipGetLastError());
        HCHECK(hipStreamSynchronize(s));
    } catch (const Error &e) {
        delete t;
        return ret_error(ctx, e);
    }
    *out = t;
    return ELPH_OK;
}
}  // namespace

int elph_tetra_create(elph_ctx *ctx, int64_t nk, const int64_t *mesh, const int64_t *indexlist, int64_t nb,
                      const double *ens, elph_tetra **out) {
    return create_simplices(ctx, nk, mesh, indexlist, nb, ens, false, out);
}

int elph_triang_create(elph_ctx *ctx, int64_t nk, const int64_t *mesh, const int64_t *indexlist, int64_t nb,
                       const double *ens, elph_tetra **out) {
    return create_simplices(ctx, nk, mesh, indexlist, nb, ens, true, out);
}

int elph_tetra_tables(const elph_tetra *t, int64_t *tetra, int64_t *tetracount, int64_t *tetramap,
                      double *tetra_evals) {
    if (!t) return ELPH_EINVAL;
    if (tetra) std::memcpy(tetra, t->tetra.data(), t->tetra.size() * sizeof(int64_t));
    if (tetracount) std::memcpy(tetracount, t->tetracount.data(), t->tetracount.size() * sizeof(int64_t));
    if (tetramap) std::memcpy(tetramap, t->tetramap.data(), t->tetramap.size() * sizeof(int64_t));
    if (tetra_evals) std::memcpy(tetra_evals, t->tetra_evals.data(), t->tetra_evals.size() * sizeof(double));
    return ELPH_OK;
}

int elph_tetra_delta(elph_tetra *t, int64_t n, const double *e, const int64_t *ik, const int64_t *ib, double *out) {
    if (!t || n < 0 || (n > 0 && (!e || !ik || !ib || !out))) return ELPH_EINVAL;
    if (n == 0) return ELPH_OK;
    elph_ctx *c = t->ctx;
    try {
        HCHECK(hipSetDevice(c->device));
        for (int64_t i = 0; i < n; ++i)
            if (ik[i] < 1 || ik[i] > t->nk || ib[i] < 1 || ib[i] > t->nb)
                throw Error{ELPH_EINVAL, "wave vector or band index out of range"};
        if (n > INT32_MAX / 2) throw Error{ELPH_EINVAL, "too many samples"};
        hipStream_t s = c->stream;
        DBuf<double> de, dout;
        DBuf<int64_t> dk, db;
        de.upload(e, n, s);
        dk.upload(ik, n, s);
        db.upload(ib, n, s);
        dout.alloc(n);
        hipLaunchKernelGGL(k_delta_samples, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, view_of(t), (int)n,
                           de.p, dk.p, db.p, dout.p);
        HCHECK(hipGetLastError());
        HCHECK(hipMemcpyAsync(out, dout.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
        HCHECK(hipStreamSynchronize(s));
    } catch (const Error &err) {
        return ret_error(c, err);
    }
    return ELPH_OK;
}

int elph_tetra_destroy(elph_tetra *t) {
    delete t;
    return ELPH_OK;
}

int elph_ibzk_plan_X(elph_ibzk_plan *plan, const elph_tetra *tet, double T, double chempot, int64_t first,
                     int64_t last, const double *g2, double *xplus, double *xminus, int64_t *istate_el,
                     int64_t *istate_ph, double *rates) {
    if (!plan || !tet || !g2) return ELPH_EINVAL;
    elph_ctx *c = tet->ctx;
    try {
        HCHECK(hipSetDevice(c->device));
        IbzkRows v = ibzk_rows(plan);
        c = v.ctx;
        if (!(T > 0.0)) throw Error{ELPH_EINVAL, "temperature must be positive"};
        if ((xplus == nullptr) != (xminus == nullptr)) throw Error{ELPH_EINVAL, "pass both X+ and X- or neither"};
        if ((istate_el == nullptr) != (istate_ph == nullptr))
            throw Error{ELPH_EINVAL, "pass both process maps or neither"};
        if (!xplus && !istate_el && !rates) throw Error{ELPH_EINVAL, "no output requested"};
        if (tet->ctx->device != c->device) throw Error{ELPH_EINVAL, "tetrahedra and plan on different devices"};
        if (tet->nk != v.nwv || tet->nb != v.nb)
            throw Error{ELPH_EINVAL, "tetrahedra not formed on the plan's electron list and bands"};
        for (int d = 0; d < 3; ++d)
            if (tet->mesh[d] != v.Nel[d]) throw Error{ELPH_EINVAL, "tetrahedra mesh differs from the electron mesh"};
        if (first < 1 || last > v.nirr * v.nb || last < first)
            throw Error{ELPH_EINVAL, "istate range outside 1..nwv_irred*numwannbands"};
        const int64_t nst = last - first + 1;
        const int64_t nblk = (v.nprocs + 256 * kXPer - 1) / (256 * kXPer);
        if (nst > 65535) throw Error{ELPH_EINVAL, "at most 65535 states per call"};
        // in-window states only (the reference skips the others, 2188-2189)
        std::vector<int4> sk;
        std::vector<double> se;
        std::vector<int> srow, sidx(nst, -1);
        for (int64_t i = 0; i < nst; ++i) {
            int64_t m, ik;
            demux_state(first + i, v.nb, &m, &ik);
            const double e = v.ens_irred[(ik - 1) + v.nirr * (m - 1)];
            if (std::fabs(e - v.enref) > v.fsthick) continue;
            const int *kk = v.kint_irr + 3 * (ik - 1);
            sidx[i] = (int)sk.size();
            sk.push_back(make_int4(kk[0], kk[1], kk[2], 1));
            se.push_back(e);
            srow.push_back((int)i);
        }
        const int64_t nact = (int64_t)sk.size();
        hipStream_t s = c->stream;
        ConsumerScratch &scr = *v.scr;
        if (nact > 0) {
            scr.st_k.upload(sk, s);
            scr.st_e.upload(se, s);
            scr.st_row.upload(srow, s);
        }
        scr.st_idx.upload(sidx, s);
        scr.partial.alloc(std::max<int64_t>(1, nact * nblk));
        // Bose factors of the phonon modes the processes use
        const int64_t ncoarse = (int64_t)v.Nph[0] * v.Nph[1] * v.Nph[2] * v.nbr;
        const int64_t nfine = v.mesh_ref > 1 ? (int64_t)v.Nel[0] * v.Nel[1] * v.Nel[2] * v.nbr : 0;
        scr.bose.alloc(ncoarse + nfine);
        hipLaunchKernelGGL(k_bose, dim3((unsigned)((ncoarse + 255) / 256)), dim3(256), 0, s, (int)ncoarse, v.om, T,
                           scr.bose.p);
        if (nfine > 0)
            hipLaunchKernelGGL(k_bose, dim3((unsigned)((nfine + 255) / 256)), dim3(256), 0, s, (int)nfine, v.omf, T,
                               scr.bose.p + ncoarse);
        HCHECK(hipGetLastError());
        XArgs a;
        a.t = view_of(tet);
        a.bose = scr.bose.p;
        a.ncoarse = (size_t)ncoarse;
        a.ikT = 1.0 / (kBoltz * T);
        a.nb = v.nb;
        a.nbr = v.nbr;
        a.mesh_ref = v.mesh_ref;
        a.Nel0 = v.Nel[0];
        a.Nel1 = v.Nel[1];
        a.Nel2 = v.Nel[2];
        a.Nph0 = v.Nph[0];
        a.Nph1 = v.Nph[1];
        a.Nph2 = v.Nph[2];
        a.nprocs = v.nprocs;
        a.vmap = scr.vmap.p;
        a.kp_int = v.kp_int;
        a.offs = v.offs;
        a.om = v.om;
        a.omf = v.omf;
        a.st_k = scr.st_k.p;
        a.st_e = scr.st_e.p;
        a.st_row = scr.st_row.p;
        a.T = T;
        a.chempot = chempot;
        a.cnst = 2.0 * (4.0 * std::atan(1.0)) / (kHbar / kQe);
        a.g2 = g2;
        a.xp = xplus;
        a.xm = xminus;
        a.se = istate_el;
        a.sp = istate_ph;
        a.partial = scr.partial.p;
        HCHECK(hipEventRecord(c->ev_a, s));
        a.nblk = (int)nblk;
        a.states_x = nblk <= 65535 ? 1 : 0;  // grid y holds at most 65535 blocks
        {
            auto a16 = [](const void *q) { return q == nullptr || ((uintptr_t)q & 15) == 0; };
            a.pairs16 = a16(a.g2) && a16(a.xp) && a16(a.xm) && a16(a.se) && a16(a.sp) ? 1 : 0;
        }
        if (nblk > 0 && nact > 0) {
            const dim3 grid = a.states_x ? dim3((unsigned)nact, (unsigned)nblk) : dim3((unsigned)nblk, (unsigned)nact);
            hipLaunchKernelGGL(k_ibzk_X, grid, dim3(256), 0, s, a);
            HCHECK(hipGetLastError());
        }
        HCHECK(hipEventRecord(c->ev_b, s));
        if (rates) {
            hipLaunchKernelGGL(k_rates, dim3((unsigned)nst), dim3(256), 0, s, scr.partial.p, (int)nblk, scr.st_idx.p,
                               rates);
            HCHECK(hipGetLastError());
        }
        HCHECK(hipStreamSynchronize(s));
        float ms = 0;
        HCHECK(hipEventElapsedTime(&ms, c->ev_a, c->ev_b));
        c->t_dominant_ms = ms;
        c->n_dominant = 1;
    } catch (const Error &err) {
        return ret_error(c, err);
    }
    return ELPH_OK;
}

int elph_el_transport_create(elph_ctx *ctx, int64_t nk, int64_t nb, const double *ens, const double *vels,
                             const int64_t *indexlist, int64_t nk_ibz, int64_t maxeq, const int64_t *nequiv,
                             const int64_t *ibz2fbz_map, const double *symmetrizers, int64_t nrot,
                             const double *crotations, double volume, const int64_t *mesh, int64_t spindeg,
                             elph_el_transport **out) {
    if (!ctx || !out || !ens || !vels || !indexlist || !nequiv || !ibz2fbz_map || !mesh || nk < 1 || nb < 1 ||
        nk_ibz < 1 || maxeq < 1 || nrot < 0 || (nrot > 0 && !crotations))
        return ELPH_EINVAL;
    *out = nullptr;
    auto *t = new elph_el_transport();
    try {
        HCHECK(hipSetDevice(ctx->device));
        if (!(volume > 0.0)) throw Error{ELPH_EINVAL, "volume must be positive"};
        if (nk > INT32_MAX / 36 || nb > 65535) throw Error{ELPH_EINVAL, "list too large"};
        t->ctx = ctx;
        t->nk = nk;
        t->nb = nb;
        t->nk_ibz = nk_ibz;
        t->deg = spindeg;
        t->volume = volume;
        for (int d = 0; d < 3; ++d) {
            if (mesh[d] < 1) throw Error{ELPH_EINVAL, "mesh components must be positive"};
            t->mesh[d] = mesh[d];
        }
        t->crot.assign(crotations, crotations + 9 * nrot);
        // IBZ point feeding each list point: the images ibz2fbz_map(ieq, i, 2)
        // located in the list (src/bte.f90:984-998), later entries win
        std::vector<int> src(nk, -1);
        for (int64_t i = 0; i < nk_ibz; ++i) {
            if (nequiv[i] < 0 || nequiv[i] > maxeq) throw Error{ELPH_EINVAL, "nequiv out of range"};
            for (int64_t l = 0; l < nequiv[i]; ++l) {
                const int64_t muxed = ibz2fbz_map[l + maxeq * (i + nk_ibz * 1)];
                const int64_t *f = std::lower_bound(indexlist, indexlist + nk, muxed);
                if (f == indexlist + nk || *f != muxed) throw Error{ELPH_EINVAL, "IBZ image not in the index list"};
                src[f - indexlist] = (int)i;
            }
        }
        std::vector<double> he(nk * nb), hv(nk * nb * 3);
        for (int64_t k = 0; k < nk; ++k)
            for (int64_t b = 0; b < nb; ++b) {
                he[k * nb + b] = ens[k + nk * b];
                for (int d = 0; d < 3; ++d) hv[(k * nb + b) * 3 + d] = vels[k + nk * (b + nb * d)];
            }
        hipStream_t s = ctx->stream;
        t->ens.upload(he, s);
        t->vels.upload(hv, s);
        t->src.upload(src, s);
        if (symmetrizers) {
            t->symmetrize_k = true;
            t->sym.upload(symmetrizers, nk * 9, s);
        }
        HCHECK(hipStreamSynchronize(s));
    } catch (const Error &e) {
        delete t;
        return ret_error(ctx, e);
    }
    *out = t;
    return ELPH_OK;
}

int elph_el_transport_rta(elph_el_transport *t, const double *rates_ibz, double T, double chempot, double *kappa0,
                          double *sigmaS, double *sigma, double *alphabyT) {
    if (!t || !rates_ibz) return ELPH_EINVAL;
    elph_ctx *c = t->ctx;
    try {
        HCHECK(hipSetDevice(c->device));
        if (!(T > 0.0)) throw Error{ELPH_EINVAL, "temperature must be positive"};
        const int nb = (int)t->nb;
        const int nblk = (int)std::min<int64_t>(64, (t->nk + 255) / 256);
        hipStream_t s = c->stream;
        t->part.alloc((size_t)nb * nblk * 36);
        t->sums.alloc((size_t)nb * 36);
        hipLaunchKernelGGL(k_el_transport, dim3(nblk, nb), dim3(256), 0, s, (int)t->nk, nb, t->ens.p, t->vels.p,
                           t->symmetrize_k ? t->sym.p : nullptr, t->src.p, rates_ibz, T, chempot, kQe, t->part.p);
        HCHECK(hipGetLastError());
        hipLaunchKernelGGL(k_el_transport_sum, dim3((nb * 36 + 255) / 256), dim3(256), 0, s, nb, nblk, t->part.p,
                           t->sums.p);
        HCHECK(hipGetLastError());
        std::vector<double> h((size_t)nb * 36);
        HCHECK(hipMemcpyAsync(h.data(), t->sums.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, s));
        HCHECK(hipStreamSynchronize(s));
        // unit factors of src/bz_sums.f90:540-556 (the source's 1.0e21 is a
        // single-precision literal) and the tensor symmetrization of 594-608
        const double fac = (double)1.0e21f / kBoltz / T / t->volume / (double)(t->mesh[0] * t->mesh[1] * t->mesh[2]);
        const double A_cc_E = t->deg * fac, A_hc_E = -A_cc_E;
        const double A_cc_T = -t->deg * kQe * fac, A_hc_T = t->deg * kQe * fac;
        const double A[4] = {A_hc_E, A_cc_E, A_hc_T, A_cc_T};
        double *dst[4] = {alphabyT, sigma, kappa0, sigmaS};
        const int64_t nrot = (int64_t)t->crot.size() / 9;
        for (int q = 0; q < 4; ++q) {
            if (!dst[q]) continue;
            for (int b = 0; b < nb; ++b) {
                double m[9];
                for (int i = 0; i < 9; ++i) m[i] = A[q] * h[(size_t)b * 36 + 9 * q + i];
                if (nrot > 0) {
                    double aux[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
                    for (int64_t r = 0; r < nrot; ++r) {
                        const double *C = &t->crot[9 * r];
                        double mc[9];
                        for (int i = 0; i < 3; ++i)
                            for (int j = 0; j < 3; ++j)
                                mc[i + 3 * j] = m[i] * C[j] + m[i + 3] * C[j + 3] + m[i + 6] * C[j + 6];
                        for (int i = 0; i < 3; ++i)
                            for (int j = 0; j < 3; ++j)
                                aux[i + 3 * j] += C[i] * mc[3 * j] + C[i + 3] * mc[1 + 3 * j] + C[i + 6] * mc[2 + 3 * j];
                    }
                    for (int i = 0; i < 9; ++i) m[i] = aux[i] / (double)nrot;
                }
                // alpha/T: the E-field heat coefficient divided by T (src/bte.f90:275)
                for (int i = 0; i < 9; ++i) dst[q][b + nb * i] = q == 0 ? m[i] / T : m[i];
            }
        }
    } catch (const Error &err) {
        return ret_error(c, err);
    }
    return ELPH_OK;
}

int elph_el_transport_destroy(elph_el_transport *t) {
    delete t;
    return ELPH_OK;
}

}  // extern "C"
