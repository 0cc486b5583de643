// elph_vertex: electron-phonon vertex path for MI355X (gfx950).
//
// Reference path (elphbolt): Wannier%el_wann/ph_wann/gkRp/gReq/g2
// (src/wannier.f90:430-1188) driven by calculate_gkRp and
// calculate_eph_interaction_ibzk/ibzq (src/interactions.f90:1603-2400).
//
// Device pipeline of one IBZ-k batch (FP64):
//   1. k_phase_list    phi(k,Re) = e^{i 2pi k.Re} / elwsdeg(Re)
//   2. k_gkRp          gkRp[k] = sum_Re phi(k,Re) gwann(:,:,Re,:,:)    (FP64 MFMA GEMM)
//   3. k_rotate_h      h[k,m](np,sp,Rp) = sum_mp U_k(m,mp) gkRp[k](np,mp,sp,Rp) / phwsdeg(Rp)
//   4. k_pair_fp64     Y[k'](np,sp) = sum_Rp e^{i 2pi q.Rp} h[k,m](np,sp,Rp), q = k'-k   (MFMA)
//                      fused epilogue: g = sum_sp u_q(s,sp) sum_np conj(U_k'(n,np)) Y(np,sp),
//                      |g|^2 = 1/2 |g|^2 / omega * Ryd2eV^3 Ryd2amu, written in the
//                      reference's (k', n, s) order.
// The order of the linear contractions differs from the reference's per-vertex
// g2 (overlap, phonon dot, then Rp sum); the results agree to rounding.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

#include "common.h"
#include "ctx.h"
#include "kernels.h"

using namespace elph;

// Scratch of the allocation-free entry point (elph_ibzk_plan_run_noalloc):
// one array per device, held from the library's first use on that device
// until the process ends; every call on a device shares it.  Host code never
// takes the name as an address: it uses the pointer hipGetSymbolAddress
// returns with the device current (elph_ctx::noalloc_scratch).
constexpr size_t kNoAllocScratchBytes = (size_t)128 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

// Device address of `scratch` on `device` (current), looked up once per device.
unsigned char *noalloc_scratch_address(int device) {
    static std::mutex mu;
    static std::map<int, unsigned char *> addr;
    std::lock_guard<std::mutex> lock(mu);
    auto it = addr.find(device);
    if (it != addr.end()) return it->second;
    void *ptr = nullptr;
    HCHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) throw Error{ELPH_EDEVICE, "hipGetSymbolAddress returned no address for the scratch array"};
    addr[device] = static_cast<unsigned char *>(ptr);
    return addr[device];
}

// Bump allocator over a fixed device region: hands out 256-byte aligned
// views, throws ELPH_ENOMEM past the end.
struct Carve {
    unsigned char *base;
    size_t cap, used = 0;
    template <typename T>
    void take(DBuf<T> &b, size_t count) {
        const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
        const size_t off = (used + 255) & ~(size_t)255;
        if (off + bytes > cap)
            throw Error{ELPH_ENOMEM, "the request needs more than the " + std::to_string(cap >> 20) +
                                         " MiB scratch of the allocation-free entry point"};
        b.borrow(reinterpret_cast<T *>(base + off), std::max<size_t>(count, 1));
        used = off + bytes;
    }
};

// ---------------------------------------------------------------------------
// Phases
// ---------------------------------------------------------------------------

// out[v][c] = expi(2 pi vecs(v,:) . cells(c,:)) / deg(c), the factor of
// src/wannier.f90:466-467 / 1096-1099 / 1165.  vecs is (n,3) column-major.
// Element (v, c) goes to out[v * ld_v + c * ld_c]; vecs(v, d) is vecs[v + vstride * d].
__global__ void k_phase_list(int n, int ncell, const double *__restrict__ vecs, int64_t vstride,
                             const int *__restrict__ cells, const int *__restrict__ deg, double2 *__restrict__ out,
                             int64_t ld_v, int64_t ld_c) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int v = blockIdx.y;
    if (c >= ncell || v >= n) return;
    double dot = vecs[v] * (double)cells[3 * c] + vecs[v + vstride] * (double)cells[3 * c + 1] +
                 vecs[v + 2 * vstride] * (double)cells[3 * c + 2];
    double s, co;
    sincos(kTwoPi * dot, &s, &co);
    const double d = (double)deg[c];
    out[(size_t)v * ld_v + (size_t)c * ld_c] = make_double2(co / d, s / d);
}

// Cell-pair phase table of the Re contraction: out[v * ld_v + K2 * ld_c] is
// cos (K2 = 2p) or sin (K2 = 2p + 1) of 2 pi vecs(v,:) . R_p, the factor of
// src/wannier.f90:466-467 without the degeneracy (carried by the folded
// gwann, gw_pair).  vecs is (n,3) column-major with stride vstride.
__global__ void k_phase_pairs(int n, int npair, const double *__restrict__ vecs, int64_t vstride,
                              const int *__restrict__ pcell, double *__restrict__ out, int64_t ld_v, int64_t ld_c) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int v = blockIdx.y;
    if (p >= npair || v >= n) return;
    const double dot = vecs[v] * (double)pcell[3 * p] + vecs[v + vstride] * (double)pcell[3 * p + 1] +
                       vecs[v + 2 * vstride] * (double)pcell[3 * p + 2];
    double sn, co;
    sincos(kTwoPi * dot, &sn, &co);
    out[(size_t)v * ld_v + (size_t)(2 * p) * ld_c] = co;
    out[(size_t)v * ld_v + (size_t)(2 * p + 1) * ld_c] = sn;
}

// Row K2 of the folded gwann of one (sp,Rp) slice (gs = slice base, [Re][i]):
//   G(2p) = g(R_p)/d(R_p) + g(-R_p)/d(-R_p),  G(2p+1) = i (g(R_p)/d(R_p) - g(-R_p)/d(-R_p)),
// so that sum_Re e^{i theta(Re)} g(Re) / d(Re) = sum_p cos(theta_p) G(2p) + sin(theta_p) G(2p+1)
// (zero past the pairs; a cell without its opposite has g(-R_p) = 0).  invdeg
// holds 1/d per cell.
__device__ __forceinline__ double2 gw_pair(const double2 *__restrict__ gs, int nb2, int i, int k2, int npair,
                                           const int2 *__restrict__ pairs, const double *__restrict__ invdeg) {
    const int p = k2 >> 1;
    if (p >= npair) return make_double2(0.0, 0.0);
    const int2 pr = pairs[p];
    const double2 ga = gs[(size_t)pr.x * nb2 + i];
    const double da = invdeg[pr.x];
    double2 a = make_double2(ga.x * da, ga.y * da), b = make_double2(0.0, 0.0);
    if (pr.y >= 0) {
        const double2 gb = gs[(size_t)pr.y * nb2 + i];
        const double db = invdeg[pr.y];
        b = make_double2(gb.x * db, gb.y * db);
    }
    return (k2 & 1) ? make_double2(b.y - a.y, a.x - b.x) : make_double2(a.x + b.x, a.y + b.y);
}

// Zero fill (used instead of hipMemsetAsync so the launches stay graph-safe).
__global__ void k_zero2(double2 *__restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = make_double2(0.0, 0.0);
}

// ---------------------------------------------------------------------------
// Re contraction (gkRp) as an FP64 MFMA GEMM per (sp,Rp) slice in cell-pair
// form (real phase operand, gwann folded on the fly by gw_pair):
//   out[k][slice][i] = sum_K2 G[i][K2][slice] * phi[K2][k]
// One workgroup per slice; wave w takes row blocks w, w+4, ...
// ---------------------------------------------------------------------------
template <int NCB>
__global__ __launch_bounds__(256) void k_gkRp(const double2 *__restrict__ gw, int nb2, int nwsk, int nslices, int npair,
                                              const int2 *__restrict__ pairs, const double *__restrict__ invdeg,
                                              const double *__restrict__ phi, int ldphi, int nk,
                                              double2 *__restrict__ out, int64_t out_kstride) {
    // pair table with the cells' 1/deg in LDS: (index, 1/deg) of R_p and -R_p
    extern __shared__ double4 ptab[];
    const int slice = blockIdx.x;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int li = lane & 15, lk = lane >> 4;
    for (int q = threadIdx.x; q < npair; q += 256) {
        const int2 pr = pairs[q];
        ptab[q] = make_double4((double)pr.x, invdeg[pr.x], (double)pr.y, pr.y >= 0 ? invdeg[pr.y] : 0.0);
    }
    __syncthreads();
    const double2 *A = gw + (size_t)slice * nb2 * nwsk;
    const int nrb = (nb2 + 15) >> 4, K2 = 2 * npair;
    for (int rb = wave; rb < nrb; rb += 4) {
        d4 cr[NCB], ci[NCB];
#pragma unroll
        for (int c = 0; c < NCB; ++c) { cr[c] = d4{0, 0, 0, 0}; ci[c] = d4{0, 0, 0, 0}; }
        const int row = rb * 16 + li;
#pragma unroll 4
        for (int k0 = 0; k0 < K2; k0 += 4) {
            const int kk = k0 + lk;
            // G(kk) of gw_pair from the LDS pair table
            double2 a = make_double2(0.0, 0.0);
            if (row < nb2 && kk < K2) {
                const double4 t = ptab[kk >> 1];
                const double2 ga = A[(size_t)(int)t.x * nb2 + row];
                double2 u = make_double2(ga.x * t.y, ga.y * t.y), v = make_double2(0.0, 0.0);
                if (t.z >= 0.0) {
                    const double2 gb = A[(size_t)(int)t.z * nb2 + row];
                    v = make_double2(gb.x * t.w, gb.y * t.w);
                }
                a = (kk & 1) ? make_double2(v.y - u.y, u.x - v.x) : make_double2(u.x + v.x, u.y + v.y);
            }
#pragma unroll
            for (int c = 0; c < NCB; ++c) {
                const int col = c * 16 + li;
                const double b = kk < K2 ? phi[(size_t)kk * ldphi + col] : 0.0;
                cr[c] = mfma_f64(a.x, b, cr[c]);
                ci[c] = mfma_f64(a.y, b, ci[c]);
            }
        }
#pragma unroll
        for (int c = 0; c < NCB; ++c) {
            const int col = c * 16 + li;
            if (col >= nk) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = rb * 16 + lk + 4 * r;
                if (i < nb2) out[(size_t)col * out_kstride + (size_t)slice * nb2 + i] = make_double2(cr[c][r], ci[c][r]);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Band rotation of gkRp for IBZ state (k, m), folded over pairs of opposite
// cells.  With
//   h(R) = sum_mp U_k(m,mp) gkRp[k](np,mp,sp,R) / phwsdeg(R),  c = np + nb sp
// (the initial-band half of the overlap of src/wannier.f90:960-965) and the
// pair (R_p, -R_p) of the Rp list,
//   F[st][2p][c] = h(R_p) + h(-R_p),   F[st][2p+1][c] = i (h(R_p) - h(-R_p)),
// so that sum_R e^{i theta(R)} h(R) = sum_p cos(theta_p) F[2p] + sin(theta_p) F[2p+1],
// theta_p = theta(R_p): the phase operand of the contraction is real.  A cell
// whose opposite is not in the list (R = 0) has h(-R_p) = 0.
// ---------------------------------------------------------------------------
__global__ void k_rotate_h(int nst, int nb, int nbr, int npair, int Kp, int CP, const double2 *__restrict__ gk,
                           int64_t gk_kstride, const int *__restrict__ st_kb, const int *__restrict__ st_m,
                           const double2 *__restrict__ Uk /*[kb][m][mp]*/, const int *__restrict__ degq,
                           const int2 *__restrict__ pairs, double2 *__restrict__ h) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t per = (int64_t)(Kp / 2) * CP;
    if (idx >= per * nst) return;
    const int st = (int)(idx / per);
    const int rem = (int)(idx % per);
    const int p = rem / CP, c = rem % CP;
    double2 f0 = make_double2(0.0, 0.0), f1 = make_double2(0.0, 0.0);
    if (p < npair && c < nb * nbr) {
        const int np = c % nb, sp = c / nb;
        const int kb = st_kb[st], m = st_m[st];
        const double2 *u = Uk + ((size_t)kb * nb + m) * nb;
        auto rot = [&](int rp) {
            const double2 *g = gk + (size_t)kb * gk_kstride + ((size_t)(sp + nbr * rp)) * nb * nb + np;
            double2 acc = make_double2(0.0, 0.0);
            for (int mp = 0; mp < nb; ++mp) cfma(acc, u[mp], g[(size_t)mp * nb]);
            const double d = (double)degq[rp];
            return make_double2(acc.x / d, acc.y / d);
        };
        const int2 pr = pairs[p];
        const double2 hi = rot(pr.x);
        const double2 hj = pr.y >= 0 ? rot(pr.y) : make_double2(0.0, 0.0);
        f0 = make_double2(hi.x + hj.x, hi.y + hj.y);
        f1 = make_double2(hj.y - hi.y, hi.x - hj.x);
    }
    double2 *o = h + ((size_t)st * Kp + 2 * p) * CP + c;
    o[0] = f0;
    o[CP] = f1;
}

// ---------------------------------------------------------------------------
// Fused pair kernel (FP64): Rp contraction on MFMA + band/mode rotations +
// |g|^2 epilogue.  Block = 4 waves; each wave owns 32 k' rows (2 row blocks)
// and all CP = 16*CB columns (np, sp) of one IBZ state (k, m).
// ---------------------------------------------------------------------------
struct PairArgs {
    const double2 *h;          // [st][Kp][CP]
    int Kp, CP;
    const int *st_kint;        // [st][3] IBZ k on the electron mesh (0-based ints)
    const int64_t *st_slot;    // output row of the state
    int nwv;                   // k' rows
    const int *kp_int;         // [nwv][3]
    int Nel0, Nel1, Nel2;      // electron mesh
    int mesh_ref;
    int Nph0, Nph1, Nph2;      // phonon mesh
    int L, Lf0, Lf1, Lf2;      // phase table size and L / Nel
    unsigned Lmul;             // ceil(2^32 / L) (0 for L = 1): index mod L by multiply-high
    const double2 *T;          // [L] e^{i 2pi j / L}
    const int *Rq;             // [Kp][4] cell (padded rows are 0)
    const double2 *Vkp;        // [nwv][n][np] = evecs(ikp, n, np)
    const int *nwin;           // [nwv] in-window band count
    const int *wband;          // [nwv][nb] in-window bands (0-based)
    const int64_t *offs;       // [nwv] offset of k' in a row
    const double2 *u;          // [nq][s][sp] phonon eigvec / sqrt(mass)
    const double *om;          // [nq][s] eV
    const double2 *uf;         // fine-mesh phonons (mesh_ref > 1) [nq_el][s][sp]
    const double *omf;
    // polar crystals: long-range prefactor per mode (null otherwise), [nq][s] and
    // [nq_el][s], and the IBZ eigenvector U_k(m,:) of each state
    const double2 *lp, *lpf;
    const double2 *Ukb;        // [k of batch][m][mp]
    const int *st_kb, *st_m;
    int nb, nbr;
    double *out;
    int64_t nprocs;
};

// ibzk epilogue of one k' row (row `row`, Y row `yi` of the staged 16 x CP
// tile): for every in-window final band n (ranks lk, lk+4, ...) and branch s,
//   g = sum_sp u_q(s,sp) sum_np conj(U_k'(n,np)) Y(np + nb sp),
//   |g|^2 = 1/2 |g|^2 / omega * Ryd2eV^3 Ryd2amu   (0 when omega == 0),
// written at the reference's position (k' block, rank of n, s).  Polar
// crystals add lp(q,s) * sum_i conj(U_k'(n,i)) U_k(m,i) to g first
// (src/wannier.f90:998-1006).
// NBM > 0: nb <= NBM and nbr <= NBRM; the loops over bands and branches are
// unrolled to those bounds (guarded by the runtime counts), the row of U_k'
// is held in registers and every load of a rank is issued before its
// arithmetic.  NBM == 0: loops over the runtime counts.
// The phonon row of q = k' - k (mesh_ref > 1: the fine-mesh table for q off
// the phonon mesh).
struct PhRow {
    const double2 *uq = nullptr, *lq = nullptr;
    const double *omq = nullptr;
};
__device__ __forceinline__ PhRow phonon_row(const PairArgs &a, int q0, int q1, int q2) {
    PhRow r;
    if (q0 % a.mesh_ref == 0 && q1 % a.mesh_ref == 0 && q2 % a.mesh_ref == 0) {
        const int iq = ((q2 / a.mesh_ref) * a.Nph1 + q1 / a.mesh_ref) * a.Nph0 + q0 / a.mesh_ref;
        r.uq = a.u + (size_t)iq * a.nbr * a.nbr;
        r.omq = a.om + (size_t)iq * a.nbr;
        if (a.lp) r.lq = a.lp + (size_t)iq * a.nbr;
    } else {
        const int iq = (q2 * a.Nel1 + q1) * a.Nel0 + q0;
        r.uq = a.uf + (size_t)iq * a.nbr * a.nbr;
        r.omq = a.omf + (size_t)iq * a.nbr;
        if (a.lpf) r.lq = a.lpf + (size_t)iq * a.nbr;
    }
    return r;
}

template <int NBM, int NBRM>
__device__ __forceinline__ void ibzk_epilogue_row(const PairArgs &a, const double2 *Y, int ys, int row, int yi, int lk,
                                                  int st, int k0i, int k1i, int k2i, double *outrow) {
    if (row >= a.nwv) return;
    const int q0 = pos_mod(a.kp_int[3 * row] - k0i, a.Nel0);
    const int q1 = pos_mod(a.kp_int[3 * row + 1] - k1i, a.Nel1);
    const int q2 = pos_mod(a.kp_int[3 * row + 2] - k2i, a.Nel2);
    const PhRow ph = phonon_row(a, q0, q1, q2);
    const double2 *uq = ph.uq, *lq = ph.lq;
    const double *omq = ph.omq;
    const double2 *Ukm = lq ? a.Ukb + ((size_t)a.st_kb[st] * a.nb + a.st_m[st]) * a.nb : nullptr;
    const int nn = a.nwin[row];
    const double2 *Yr = Y + yi * ys;
    if constexpr (NBM > 0) {
        const int nb = a.nb, nbr = a.nbr;
        for (int rank = lk; rank < nn; rank += 4) {
            const int n = a.wband[row * nb + rank];
            const double2 *V = a.Vkp + ((size_t)row * nb + n) * nb;
            double2 v[NBM];
#pragma unroll
            for (int np = 0; np < NBM; ++np) v[np] = np < nb ? V[np] : make_double2(0.0, 0.0);
            // z(sp) = sum_np conj(U_k'(n,np)) Y(np + nb sp)
            double2 z[NBRM];
#pragma unroll
            for (int sp = 0; sp < NBRM; ++sp) {
                double2 t = make_double2(0.0, 0.0);
                if (sp < nbr)
#pragma unroll
                    for (int np = 0; np < NBM; ++np)
                        if (np < nb) {
                            const double2 y = Yr[np + nb * sp];
                            t.x = fma(v[np].x, y.x, t.x);
                            t.x = fma(v[np].y, y.y, t.x);
                            t.y = fma(v[np].x, y.y, t.y);
                            t.y = fma(-v[np].y, y.x, t.y);
                        }
                z[sp] = t;
            }
            double2 unm = make_double2(0.0, 0.0);
            if (lq)
#pragma unroll
                for (int i = 0; i < NBM; ++i)
                    if (i < nb) {
                        const double2 t = cmulc(v[i], Ukm[i]);
                        unm.x += t.x;
                        unm.y += t.y;
                    }
            // g(s) = sum_sp u_q(s,sp) z(sp) (+ polar); 8-byte stores at the
            // branch stride go through L2, which merges them into whole lines
            // (nontemporal stores did not change the time but wrote 2.4x the
            // bytes)
            double *o = outrow + a.offs[row] + (int64_t)rank * nbr;
#pragma unroll
            for (int s = 0; s < NBRM; ++s)
                if (s < nbr) {
                    double2 gb = make_double2(0.0, 0.0);
#pragma unroll
                    for (int sp = 0; sp < NBRM; ++sp)
                        if (sp < nbr) cfma(gb, uq[s * nbr + sp], z[sp]);
                    if (lq) cfma(gb, lq[s], unm);
                    const double w = omq[s];
                    o[s] = w != 0.0 ? 0.5 * (gb.x * gb.x + gb.y * gb.y) / w * kG2Unit : 0.0;
                }
        }
        return;
    }
    for (int rank = lk; rank < nn; rank += 4) {
        const int n = a.wband[row * a.nb + rank];
        const double2 *V = a.Vkp + ((size_t)row * a.nb + n) * a.nb;
        double *o = outrow + a.offs[row] + (int64_t)rank * a.nbr;
        double2 unm = make_double2(0.0, 0.0);
        if (lq)
            for (int i = 0; i < a.nb; ++i) {
                const double2 z = cmulc(V[i], Ukm[i]);
                unm.x += z.x;
                unm.y += z.y;
            }
        for (int s0 = 0; s0 < a.nbr; s0 += 6) {
            double2 gb[6];
#pragma unroll
            for (int j = 0; j < 6; ++j) gb[j] = make_double2(0.0, 0.0);
            for (int sp = 0; sp < a.nbr; ++sp) {
                double2 z = make_double2(0.0, 0.0);
                for (int np = 0; np < a.nb; ++np) {
                    const double2 v = V[np];
                    const double2 y = Yr[np + a.nb * sp];
                    z.x = fma(v.x, y.x, z.x);  // conj(v) * y
                    z.x = fma(v.y, y.y, z.x);
                    z.y = fma(v.x, y.y, z.y);
                    z.y = fma(-v.y, y.x, z.y);
                }
#pragma unroll
                for (int j = 0; j < 6; ++j)
                    if (s0 + j < a.nbr) cfma(gb[j], uq[(s0 + j) * a.nbr + sp], z);
            }
            if (lq)
#pragma unroll
                for (int j = 0; j < 6; ++j)
                    if (s0 + j < a.nbr) cfma(gb[j], lq[s0 + j], unm);
#pragma unroll
            for (int j = 0; j < 6; ++j) {
                const int sb = s0 + j;
                if (sb < a.nbr) {
                    const double w = omq[sb];
                    double g2 = 0.0;
                    if (w != 0.0) g2 = 0.5 * (gb[j].x * gb[j].x + gb[j].y * gb[j].y) / w * kG2Unit;
                    o[sb] = g2;
                }
            }
        }
    }
}

// Real phase operand of the cell-pair rows (k_rotate_h): row K2 = 2p + f of
// pair p is cos (f = 0) or sin (f = 1) of 2 pi (coef . R_p) / L, read from the
// table T[j] = e^{i 2 pi j / L}; Rsh[p] holds R_p.
__device__ __forceinline__ double pair_phase(const double2 *Tsh, const int4 *Rsh, const int (&coef)[3], int L, int k) {
    const int4 R = Rsh[k >> 1];
    const double2 t = Tsh[pos_mod(coef[0] * R.x + coef[1] * R.y + coef[2] * R.z, L)];
    return (k & 1) ? t.y : t.x;
}
// The same for L <= 255 and electron meshes up to 255: Rsh[p].w holds R_p
// reduced modulo the mesh, one byte per component, and cpk the byte-packed
// coefficients (phase_index).
__device__ __forceinline__ double pair_phase_packed(const double2 *Tsh, const int4 *Rsh, unsigned cpk, unsigned lmul,
                                                    int negL, int k) {
    const double2 t = Tsh[phase_index(cpk, (unsigned)Rsh[k >> 1].w, lmul, negL)];
    return (k & 1) ? t.y : t.x;
}

// Phase-contraction core of the mesh drivers: the phase of row r and cell R_K
// is the exact root of unity T[(coef_r . R_K) mod L] (see gemm_core_f64).
// PACKED (L and the mesh up to 255): the index by pair_phase_packed from the
// byte-packed coefficients cpk and Rsh[p].w = R_p mod the mesh.
template <int CB, int RB, int KC, bool PACKED>
__device__ __forceinline__ void phase_gemm_core(const double2 *__restrict__ hst, int ldb, int Kp, const int (&coef)[RB][3],
                                                int L, unsigned lmul, const double2 *Tsh, const int4 *Rsh,
                                                double2 *buf0, double2 *buf1, d4 (&cr)[RB][CB], d4 (&ci)[RB][CB]) {
    unsigned cpk[RB];
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
        cpk[rb] = (unsigned)coef[rb][0] | ((unsigned)coef[rb][1] << 8) | ((unsigned)coef[rb][2] << 16);
    // row 2p: cos(2 pi k.R_p), row 2p + 1: sin(2 pi k.R_p) (Rsh[p] = R_p)
    auto aop = [&](int rb, int k) {
        if constexpr (PACKED) {
            return pair_phase_packed(Tsh, Rsh, cpk[rb], lmul, -L, k);
        } else {
            const int4 R = Rsh[k >> 1];
            const int x = coef[rb][0] * R.x + coef[rb][1] * R.y + coef[rb][2] * R.z;
            const double2 t = Tsh[pos_mod(x, L)];
            return (k & 1) ? t.y : t.x;
        }
    };
    gemm_core_f64<CB, RB, KC>(hst, ldb, Kp, aop, buf0, buf1, cr, ci);
}

// One tile (state st, k' rows 128 rblk .. 128 rblk + 127) of the FP64 fused pair
// kernel; PACKED: phase indices by pair_phase_packed (L and the mesh up to
// 255).  The caller's workgroup is 256 threads with the dynamic LDS of
// fused_launch; the LDS is free again when the call returns.
template <int CB, int NBM, int NBRM, bool PACKED>
__device__ __forceinline__ void pair_fp64_tile(const PairArgs &a, int st, int rblk) {
    constexpr int KC = 32;
    constexpr int RB = 2;
    const int CP = 16 * CB;
    extern __shared__ double2 smem[];
    double2 *buf0 = smem;
    double2 *buf1 = smem + KC * CP;
    double2 *Tsh = smem + 2 * KC * CP;
    int4 *Rsh = reinterpret_cast<int4 *>(Tsh + a.L);

    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int li = lane & 15, lk = lane >> 4;
    const int row_base = rblk * 128 + wave * 32;
    const double2 *hst = a.h + (size_t)st * a.Kp * CP;

    for (int i = tid; i < a.L; i += 256) Tsh[i] = a.T[i];
    for (int i = tid; i < a.Kp / 2; i += 256)
        Rsh[i] = make_int4(a.Rq[4 * i], a.Rq[4 * i + 1], a.Rq[4 * i + 2], a.Rq[4 * i + 3]);

    const int k0i = a.st_kint[3 * st], k1i = a.st_kint[3 * st + 1], k2i = a.st_kint[3 * st + 2];
    int coef[RB][3];
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
        const int r = row_base + rb * 16 + li;
        coef[rb][0] = coef[rb][1] = coef[rb][2] = 0;
        if (r < a.nwv) {
            // q = k' - k folded onto the electron mesh (vector_allreps_sub, src/vector.f90:86-102)
            const int q0 = pos_mod(a.kp_int[3 * r] - k0i, a.Nel0);
            const int q1 = pos_mod(a.kp_int[3 * r + 1] - k1i, a.Nel1);
            const int q2 = pos_mod(a.kp_int[3 * r + 2] - k2i, a.Nel2);
            coef[rb][0] = (q0 * a.Lf0) % a.L;
            coef[rb][1] = (q1 * a.Lf1) % a.L;
            coef[rb][2] = (q2 * a.Lf2) % a.L;
        }
    }
    d4 cr[RB][CB], ci[RB][CB];
    if constexpr (PACKED) {
        unsigned cpk[RB];
#pragma unroll
        for (int rb = 0; rb < RB; ++rb)
            cpk[rb] = (unsigned)coef[rb][0] | ((unsigned)coef[rb][1] << 8) | ((unsigned)coef[rb][2] << 16);
        const unsigned lmul = a.Lmul;
        const int negL = -a.L;
        auto aop = [&](int rb, int k) { return pair_phase_packed(Tsh, Rsh, cpk[rb], lmul, negL, k); };
        gemm_core_f64<CB, RB, KC>(hst, CP, a.Kp, aop, buf0, buf1, cr, ci);
    } else {
        const int L = a.L;
        auto aop = [&](int rb, int k) { return pair_phase(Tsh, Rsh, coef[rb], L, k); };
        gemm_core_f64<CB, RB, KC>(hst, CP, a.Kp, aop, buf0, buf1, cr, ci);
    }

    // ---- epilogue: per row block, stage Y (16 x CP) of this wave in LDS; the
    // odd row stride YS keeps the 16 rows on distinct LDS banks ----
    const int YS = CP + 1;
    double2 *Y = smem + wave * 16 * YS;
    double *outrow = a.out + a.st_slot[st] * a.nprocs;
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
#pragma unroll
        for (int c = 0; c < CB; ++c)
#pragma unroll
            for (int r = 0; r < 4; ++r) Y[(lk + 4 * r) * YS + c * 16 + li] = make_double2(cr[rb][c][r], ci[rb][c][r]);
        __syncthreads();
        ibzk_epilogue_row<NBM, NBRM>(a, Y, YS, row_base + rb * 16 + li, li, lk, st, k0i, k1i, k2i, outrow);
        __syncthreads();
    }
}
template <int CB, int NBM, int NBRM, bool PACKED>
__global__ __launch_bounds__(256) void k_pair_fp64(PairArgs a) {
    pair_fp64_tile<CB, NBM, NBRM, PACKED>(a, blockIdx.y, blockIdx.x);
}

// ---------------------------------------------------------------------------
// Mixed-precision phase contraction.  Both operands are split into kNS int8
// signed digits (base 254, split_digits); the kNS (kNS+1)/2 digit products with
// i + j < kNS run on the int8 matrix cores with exact int32 accumulation over
// the whole K, one accumulator per level i + j, recombined in FP64 with the
// level weights 1 / (127^2 254^l) and the column scale of B.  The band/mode
// rotations and |g|^2 stay FP64.
// ---------------------------------------------------------------------------
// Column scale of B = h[st] (max over K of max(|Re|, |Im|); 1 for zero columns).
// Block (state, 64-column tile): 4 row groups of 64 lanes read rows k = g,
// g + 4, ... (one coalesced row segment per wave), reduced in LDS.
__global__ __launch_bounds__(256) void k_colscale(int nst, int Kp, int CP, const double2 *__restrict__ h,
                                                  double *__restrict__ scale) {
    __shared__ double red[4][64];
    const int st = blockIdx.x, lane = threadIdx.x & 63, g = threadIdx.x >> 6;
    const int c = blockIdx.y * 64 + lane;
    double m = 0.0;
    if (c < CP) {
        const double2 *p = h + (size_t)st * Kp * CP + c;
        for (int k = g; k < Kp; k += 4) {
            const double2 v = p[(size_t)k * CP];
            m = fmax(m, fmax(fabs(v.x), fabs(v.y)));
        }
    }
    red[g][lane] = m;
    __syncthreads();
    if (g == 0 && c < CP) {
        m = fmax(fmax(red[0][lane], red[1][lane]), fmax(red[2][lane], red[3][lane]));
        scale[(size_t)st * CP + c] = m > 0.0 ? m : 1.0;
    }
}

// Bsl[st][kc][cb][p][s][lane] (16 bytes): digits of B / scale in the int8 MFMA
// B-operand order, p = Re, Im, -Im.
__global__ void k_slice_b(int nst, int nkc, int CB, int Kp, int CP, const double2 *__restrict__ h,
                          const double *__restrict__ scale, v4i *__restrict__ bsl) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t total = (int64_t)nst * nkc * CB * 64;
    if (t >= total) return;
    const int lane = (int)(t % 64);
    int64_t r = t / 64;
    const int cb = (int)(r % CB);
    r /= CB;
    const int kc = (int)(r % nkc);
    const int st = (int)(r / nkc);
    const int col = cb * 16 + (lane & 15), g = lane >> 4;
    const double inv = 1.0 / scale[(size_t)st * CP + col];
    unsigned ure[kNS][4] = {}, uim[kNS][4] = {}, unim[kNS][4] = {};
#pragma unroll
    for (int b = 0; b < 16; ++b) {
        const int k = kc * 64 + g * 16 + b;
        const double2 x = h[((size_t)st * Kp + k) * CP + col];
        signed char dr[kNS], di[kNS];
        split_digits<kNS>(x.x * inv, dr);
        split_digits<kNS>(x.y * inv, di);
#pragma unroll
        for (int sd = 0; sd < kNS; ++sd) {
            put_byte(ure[sd], b, dr[sd]);
            put_byte(uim[sd], b, di[sd]);
            put_byte(unim[sd], b, (signed char)(-di[sd]));
        }
    }
    v4i *o = bsl + ((((size_t)st * nkc + kc) * CB + cb) * 3) * kNS * 64 + lane;
#pragma unroll
    for (int sd = 0; sd < kNS; ++sd) {
        o[(0 * kNS + sd) * 64] = as_v4i(ure[sd]);
        o[(1 * kNS + sd) * 64] = as_v4i(uim[sd]);
        o[(2 * kNS + sd) * 64] = as_v4i(unim[sd]);
    }
}

// Bsl[st][kc][cb][p][s][lane] (16 bytes): digits of B / scale in the int8 MFMA
// B-operand order, p = Re, Im (real phase operand: the cell-pair layout of
// k_rotate_h).
__global__ void k_slice_b_ri(int nst, int nkc, int CB, int Kp, int CP, const double2 *__restrict__ h,
                             const double *__restrict__ scale, v4i *__restrict__ bsl) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t total = (int64_t)nst * nkc * CB * 64;
    if (t >= total) return;
    const int lane = (int)(t % 64);
    int64_t r = t / 64;
    const int cb = (int)(r % CB);
    r /= CB;
    const int kc = (int)(r % nkc);
    const int st = (int)(r / nkc);
    const int col = cb * 16 + (lane & 15), g = lane >> 4;
    const double inv = 1.0 / scale[(size_t)st * CP + col];
    unsigned ure[kNS][4] = {}, uim[kNS][4] = {};
#pragma unroll
    for (int b = 0; b < 16; ++b) {
        const int k = kc * 64 + g * 16 + b;
        const double2 x = h[((size_t)st * Kp + k) * CP + col];
        signed char dr[kNS], di[kNS];
        split_digits<kNS>(x.x * inv, dr);
        split_digits<kNS>(x.y * inv, di);
#pragma unroll
        for (int sd = 0; sd < kNS; ++sd) {
            put_byte(ure[sd], b, dr[sd]);
            put_byte(uim[sd], b, di[sd]);
        }
    }
    v4i *o = bsl + ((((size_t)st * nkc + kc) * CB + cb) * 2) * kNS * 64 + lane;
#pragma unroll
    for (int sd = 0; sd < kNS; ++sd) {
        o[(0 * kNS + sd) * 64] = as_v4i(ure[sd]);
        o[(1 * kNS + sd) * 64] = as_v4i(uim[sd]);
    }
}

// Mixed-precision B operand straight from g(k,Rp), for the fused and general
// ibzk paths: the folded h of k_rotate_h (same expressions, so the same
// values), its column scale as k_colscale takes it and its int8 digits in
// k_slice_b_ri's layout ([st][kc][cb][Re, Im][digit][lane]), without h going
// through memory.  Workgroup (k, 16-column block, z) holds all in-window
// states of IBZ k `kb` (st_kb is nondecreasing: they are consecutive).  A run
// is the 16 K2 rows (8 cell pairs) of one lane's digit dwords, run = 4 kc + lg;
// thread (run, w, column) takes dword w of it -- cell pairs 2w and 2w + 1 --
// for up to kRotMC states at a time, so each g(k,Rp) value is loaded once for
// all of them (NBX > 0: the band loop unrolled to NBX, loads issued together).
// Pass MAX writes the column maxima of the workgroup's runs to
// hmax[st][z][col]; pass DIGITS reduces them over z into the scale,
// recomputes h and writes the digits.
constexpr int kRotRuns = 4;  // runs per workgroup (x 4 dwords x 16 columns = 256 threads)
constexpr int kRotMC = 4;    // states per register set

__device__ __forceinline__ int lower_bound_int(const int *a, int n, int v) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

template <bool DIGITS, int NBX>
__global__ __launch_bounds__(256) void k_rotate_digits(int nb, int nbr, int npair, int nkc, int CB, int CP, int nst,
                                                       const double2 *__restrict__ gk, int64_t gk_kstride,
                                                       const int *__restrict__ st_kb, const int *__restrict__ st_m,
                                                       const double2 *__restrict__ Uk, const int *__restrict__ degq,
                                                       const int2 *__restrict__ pairs, double *__restrict__ hmax,
                                                       double *__restrict__ scale, v4i *__restrict__ bsl) {
    constexpr int NBL = NBX > 0 ? NBX : 1;  // register copies of g (runtime loop when NBX == 0)
    extern __shared__ double2 ush[];        // u rows of the k's states, [m][mp]
    __shared__ double red[4 * kRotRuns][16];
    static_assert(kRotMC == 4 && kRotRuns >= kRotMC, "the scale reduction maps 4 waves x 4 states onto red");
    const int kb = blockIdx.x, cb = blockIdx.y, z = blockIdx.z, nz = gridDim.z, t = threadIdx.x;
    const int li = t & 15, w = (t >> 4) & 3, run = z * kRotRuns + (t >> 6);
    const int s0 = lower_bound_int(st_kb, nst, kb), s1 = lower_bound_int(st_kb, nst, kb + 1);
    const int nm = s1 - s0;
    for (int i = t; i < nm * nb; i += blockDim.x) {
        const int m = i / nb, mp = i % nb;
        ush[i] = Uk[((size_t)kb * nb + st_m[s0 + m]) * nb + mp];
    }
    __syncthreads();
    const int col = cb * 16 + li;
    const bool colok = col < nb * nbr;
    const int np = colok ? col % nb : 0, sp = colok ? col / nb : 0;
    const bool live = colok && run < nkc * 4;
    const int kc = run >> 2, lg = run & 3;
    const double2 *gbase = gk + (size_t)kb * gk_kstride + (size_t)sp * nb * nb + np;
    const int64_t rstride = (int64_t)nbr * nb * nb;  // next Rp
    for (int m0 = 0; m0 < nm; m0 += kRotMC) {
        const int mc = min(kRotMC, nm - m0);
        double mx[kRotMC], inv[kRotMC];
        unsigned ure[kRotMC][kNS], uim[kRotMC][kNS];
#pragma unroll
        for (int m = 0; m < kRotMC; ++m) {
            mx[m] = 0.0;
            inv[m] = 1.0;
#pragma unroll
            for (int sd = 0; sd < kNS; ++sd) ure[m][sd] = uim[m][sd] = 0u;
        }
        if constexpr (DIGITS) {
            // column scale of each state = max over the nz partial maxima:
            // wave `part` reduces partials part, part + 4, ... of state pm,
            // then every thread combines the four waves' results
            const int pm = (t >> 4) & 3, part = t >> 6;
            double v = 0.0;
            if (pm < mc)
                for (int zz = part; zz < nz; zz += 4) v = fmax(v, hmax[((size_t)(s0 + m0 + pm) * nz + zz) * CP + col]);
            red[part * kRotMC + pm][li] = v;
            __syncthreads();
#pragma unroll
            for (int m = 0; m < kRotMC; ++m)
                if (m < mc) {
                    double sc = fmax(fmax(red[m][li], red[kRotMC + m][li]),
                                     fmax(red[2 * kRotMC + m][li], red[3 * kRotMC + m][li]));
                    sc = sc > 0.0 ? sc : 1.0;
                    inv[m] = 1.0 / sc;
                    if (z == 0 && t < 16) scale[(size_t)(s0 + m0 + m) * CP + col] = sc;
                }
            __syncthreads();  // red is written again for the next states
        }
        if (live) {
#pragma unroll 1
            for (int j = 0; j < 2; ++j) {
                const int pidx = kc * 32 + lg * 8 + 2 * w + j;
                if (pidx < npair) {
                    const int2 pr = pairs[pidx];
                    const double2 *gi = gbase + pr.x * rstride;
                    const double2 *gj = gbase + (pr.y >= 0 ? pr.y : 0) * rstride;
                    double2 ga[NBL], gb[NBL];
                    if constexpr (NBX > 0) {
#pragma unroll
                        for (int mp = 0; mp < NBX; ++mp) {
                            ga[mp] = mp < nb ? gi[(size_t)mp * nb] : make_double2(0.0, 0.0);
                            gb[mp] = mp < nb && pr.y >= 0 ? gj[(size_t)mp * nb] : make_double2(0.0, 0.0);
                        }
                    }
                    const double di = (double)degq[pr.x], dj = pr.y >= 0 ? (double)degq[pr.y] : 1.0;
#pragma unroll
                    for (int m = 0; m < kRotMC; ++m)
                        if (m < mc) {
                            const double2 *u = ush + (m0 + m) * nb;
                            double2 hi = make_double2(0.0, 0.0), hj = make_double2(0.0, 0.0);
                            if constexpr (NBX > 0) {
#pragma unroll
                                for (int mp = 0; mp < NBX; ++mp)
                                    if (mp < nb) {
                                        cfma(hi, u[mp], ga[mp]);
                                        cfma(hj, u[mp], gb[mp]);
                                    }
                            } else {
                                for (int mp = 0; mp < nb; ++mp) {
                                    cfma(hi, u[mp], gi[(size_t)mp * nb]);
                                    cfma(hj, u[mp], pr.y >= 0 ? gj[(size_t)mp * nb] : make_double2(0.0, 0.0));
                                }
                            }
                            const double2 a = make_double2(hi.x / di, hi.y / di);
                            const double2 b = pr.y >= 0 ? make_double2(hj.x / dj, hj.y / dj) : make_double2(0.0, 0.0);
                            const double2 f0 = make_double2(a.x + b.x, a.y + b.y);
                            const double2 f1 = make_double2(b.y - a.y, a.x - b.x);
                            if constexpr (!DIGITS) {
                                mx[m] = fmax(mx[m], fmax(fmax(fabs(f0.x), fabs(f0.y)), fmax(fabs(f1.x), fabs(f1.y))));
                            } else {
                                signed char dr[kNS], dim[kNS];
                                split_digits<kNS>(f0.x * inv[m], dr);
                                split_digits<kNS>(f0.y * inv[m], dim);
#pragma unroll
                                for (int sd = 0; sd < kNS; ++sd) {
                                    ure[m][sd] |= (unsigned)(unsigned char)dr[sd] << (16 * j);
                                    uim[m][sd] |= (unsigned)(unsigned char)dim[sd] << (16 * j);
                                }
                                split_digits<kNS>(f1.x * inv[m], dr);
                                split_digits<kNS>(f1.y * inv[m], dim);
#pragma unroll
                                for (int sd = 0; sd < kNS; ++sd) {
                                    ure[m][sd] |= (unsigned)(unsigned char)dr[sd] << (16 * j + 8);
                                    uim[m][sd] |= (unsigned)(unsigned char)dim[sd] << (16 * j + 8);
                                }
                            }
                        }
                }
            }
        }
        if constexpr (!DIGITS) {
            // column maxima over the workgroup's runs and dwords
#pragma unroll
            for (int m = 0; m < kRotMC; ++m) {
                if (m >= mc) break;
                red[t >> 4][li] = mx[m];
                __syncthreads();
                if (t < 16) {
                    double v = 0.0;
#pragma unroll
                    for (int r = 0; r < 4 * kRotRuns; ++r) v = fmax(v, red[r][li]);
                    hmax[((size_t)(s0 + m0 + m) * nz + z) * CP + col] = v;
                }
                __syncthreads();
            }
        } else if (run < nkc * 4) {
            const int lane = lg * 16 + li;
#pragma unroll
            for (int m = 0; m < kRotMC; ++m)
                if (m < mc) {
                    unsigned *o = reinterpret_cast<unsigned *>(
                                      bsl + ((((size_t)(s0 + m0 + m) * nkc + kc) * CB + cb) * 2) * kNS * 64 + lane) + w;
#pragma unroll
                    for (int sd = 0; sd < kNS; ++sd) {
                        o[(0 * kNS + sd) * 64 * 4] = ure[m][sd];
                        o[(1 * kNS + sd) * 64 * 4] = uim[m][sd];
                    }
                }
        }
    }
}

// Mixed-precision contraction over the cell-pair rows (K2 = 2p: cos, 2p+1:
// sin; k_rotate_h) for one block of 4 waves:
//   acc[c][l][0|1] (the wave's 16 rows x column block c) = sum over K2 and
//   digit pairs i + j = l of A_i(row, K2) B_j(K2, col) for Re|Im of B,
// exact in int32.  A digits are gathered from the digit table of the roots of
// unity: Tsh[j] holds the kNS digits of cos and sin of 2 pi j / L as byte
// pairs (cos_l, sin_l) for levels l = 0..3 (bytes 2l, 2l + 1), so a level's
// dword of two pairs is one byte permute of their table entries; Rsh[p] the
// cell of pair p reduced modulo the electron mesh
// (bytes 0-2), so the phase index of row q is (c . Rsh[p]) mod L with
// c_d = q_d L / N_d (cpk, byte-packed; phase_index).  Lane (li, lg) holds
// K2 = 64 kk + 16 lg + 4 w + b of chunk kk: pair 32 kk + 8 lg + 2 w + b / 2.
// B: chunk kc of the block's CB column blocks at bst + kc * bkc
// ([c][Re, Im][digit][lane] v4i), copied global -> LDS asynchronously one
// chunk ahead into two buffers of BCH v4i (buf); the next chunk's A digits
// are gathered between the matrix instructions of the current one, into the
// second of two register sets (the K loop is unrolled by two).  The caller
// fills Tsh and Rsh before the call (the first barrier here publishes them)
// and owns buf again after it.  Two buffers keep the LDS of a block small
// enough for three blocks per CU (CB <= 3), which hides the copy latency
// better than a deeper prefetch with two blocks.  The caller issues chunk 0
// (i8_copy_chunk(bst, bkc, 0, buf)) before it fills the tables, so that copy
// overlaps the table loads.
// (NBM: the epilogue's unrolled band bound, whose registers fit three blocks
// per CU up to 8 bands)
__host__ __device__ constexpr int pair_blocks_per_cu(int CB, int NBM = 0) { return CB <= 3 && NBM <= 8 ? 3 : 2; }
// global -> LDS copy of B chunk kn into buffer kn & 1 (the whole block)
template <int CB>
__device__ __forceinline__ void i8_copy_chunk(const v4i *__restrict__ bst, size_t bkc, int kn, v4i *buf) {
    constexpr int BCH = CB * 2 * kNS * 64;
    constexpr int NR = BCH / 256;  // copy instructions per wave and chunk
    static_assert(BCH % 256 == 0, "a chunk is a whole number of 1 KiB rows per wave");
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const v4i *src = bst + (size_t)kn * bkc;
    v4i *dst = buf + (kn & 1) * BCH;
#pragma unroll
    for (int j = 0; j < NR; ++j) {
        const int r = 4 * j + wave;
        lds_dma16(src + r * 64 + lane, dst + r * 64);
    }
}
// Dword w of the four level operands from the table entries of pairs 2w
// (t0) and 2w + 1 (t1): level l = bytes (cos_l, sin_l) of each.
__device__ __forceinline__ void level_dwords(uint2 t0, uint2 t1, v4i (&A)[kNS], int w) {
    A[0][w] = (int)__builtin_amdgcn_perm(t1.x, t0.x, 0x05040100u);
    A[1][w] = (int)__builtin_amdgcn_perm(t1.x, t0.x, 0x07060302u);
    A[2][w] = (int)__builtin_amdgcn_perm(t1.y, t0.y, 0x05040100u);
    A[3][w] = (int)__builtin_amdgcn_perm(t1.y, t0.y, 0x07060302u);
}
template <int CB>
__device__ __forceinline__ void i8_pair_core(const v4i *__restrict__ bst, size_t bkc, int nkc, const uint2 *Tsh,
                                             const unsigned *Rsh, unsigned cpk, unsigned lmul, int negL, v4i *buf,
                                             v4i (&acc)[CB][kNS][2]) {
    constexpr int NS = kNS;
    static_assert(NS == 4, "digit packing assumes 4 digits per part");
    constexpr int BCH = CB * 2 * NS * 64;
    const int lane = threadIdx.x & 63, lg = lane >> 4;
    auto bufp = [&](int kc) { return buf + (kc & 1) * BCH; };
    auto copy_b = [&](int kn) { i8_copy_chunk<CB>(bst, bkc, kn, buf); };
    // chunk kc + 1 complete in LDS for the whole block.  No LDS stores are
    // pending here, so the barrier needs no fence; the memory clobber keeps the
    // compiler's LDS reads on their side of it.
    auto next_chunk_ready = [&]() { asm volatile("s_waitcnt vmcnt(0)\n\ts_barrier" ::: "memory"); };
#pragma unroll
    for (int c = 0; c < CB; ++c)
#pragma unroll
        for (int l = 0; l < NS; ++l) { acc[c][l][0] = v4i{0, 0, 0, 0}; acc[c][l][1] = v4i{0, 0, 0, 0}; }
    lds_dma_wait();
    __syncthreads();
    // dword w of the A operands of chunk kk (4 levels)
    auto gather = [&](int kk, int w, v4i (&A)[NS]) {
        const uint2 rp = *reinterpret_cast<const uint2 *>(Rsh + kk * 32 + lg * 8 + 2 * w);
        const uint2 t0 = Tsh[phase_index(cpk, rp.x, lmul, negL)];
        const uint2 t1 = Tsh[phase_index(cpk, rp.y, lmul, negL)];
        level_dwords(t0, t1, A, w);
    };
    // all products of chunk B with A, with chunk kn's A operands gathered in
    // three stages between them, so the two dependent LDS reads of the
    // gather (cells, then table entries) are in flight while the matrix
    // cores work: the cells of the 8 pairs before digit 0, the table entries
    // after its products, the byte transposes after those of digits 1 and 2
    auto chunk = [&](const v4i *B, const v4i (&A)[NS], int kn, v4i (&nA)[NS]) {
        const uint4 ra = *reinterpret_cast<const uint4 *>(Rsh + kn * 32 + lg * 8);
        const uint4 rb = *reinterpret_cast<const uint4 *>(Rsh + kn * 32 + lg * 8 + 4);
        uint2 t[8];
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            v4i Bre[CB], Bim[CB];
#pragma unroll
            for (int c = 0; c < CB; ++c) {
                Bre[c] = B[((c * 2 + 0) * NS + j) * 64 + lane];
                Bim[c] = B[((c * 2 + 1) * NS + j) * 64 + lane];
            }
#pragma unroll
            for (int i = 0; i < NS - j; ++i)
#pragma unroll
                for (int c = 0; c < CB; ++c) {
                    acc[c][i + j][0] = mfma_i8(A[i], Bre[c], acc[c][i + j][0]);
                    acc[c][i + j][1] = mfma_i8(A[i], Bim[c], acc[c][i + j][1]);
                }
            if (j == 0) {
                const unsigned rr[8] = {ra.x, ra.y, ra.z, ra.w, rb.x, rb.y, rb.z, rb.w};
#pragma unroll
                for (int e = 0; e < 8; ++e) t[e] = Tsh[phase_index(cpk, rr[e], lmul, negL)];
            } else if (j < 3) {
#pragma unroll
                for (int w = 2 * (j - 1); w < 2 * j; ++w) level_dwords(t[2 * w], t[2 * w + 1], nA, w);
            }
        }
    };
    v4i A0[NS], A1[NS];
#pragma unroll
    for (int w = 0; w < 4; ++w) gather(0, w, A0);
    // chunk kc: the copy of chunk kc + 1 goes out first (its buffer was last
    // read in chunk kc - 1, before the preceding barrier); past the last chunk
    // the gather re-reads it (no branch in the chunk body)
    for (int kc = 0;; kc += 2) {
        if (kc + 1 < nkc) copy_b(kc + 1);
        chunk(bufp(kc), A0, min(kc + 1, nkc - 1), A1);
        if (kc + 1 >= nkc) break;
        next_chunk_ready();
        if (kc + 2 < nkc) copy_b(kc + 2);
        chunk(bufp(kc + 1), A1, min(kc + 2, nkc - 1), A0);
        if (kc + 2 >= nkc) break;
        next_chunk_ready();
    }
    lds_dma_wait();
    __syncthreads();
}

// Epilogue of the fused mixed-precision kernel for up to 8 bands: lane
// (li, lk) takes in-window ranks lk and lk + 4 of its epilogue row together,
// so the row's Y values (LDS) and u_q(s, sp) are read once for both.  The
// rows of a tile are handed to the lanes in the order of their in-window
// count (pair_i8_tile), so the waves whose rows have at most four leave the
// second rank out.  Everything that does not depend on Y -- the phonon row of
// q, the in-window count and bands, the output offset -- is set up before the
// contraction (EpiRow), so those loads hide behind the matrix work.  Same
// expressions and order as ibzk_epilogue_row up to the last step, where
// 0.5 / omega * unit is formed once for both ranks.
struct EpiRow {
    int ph = 0;                  // phonon_code of q
    int nn = 0, na = 0, nb = 0;  // in-window count; bands of ranks lk and lk + 4
    int64_t off = 0;
};
// The phonon row of q as one integer (its pointers are formed again in the
// epilogue, so the row's state held across the contraction stays small):
// iq on the phonon mesh, or -(iq + 1) for a q of the electron mesh off it
// (the choice of phonon_row).
__device__ __forceinline__ int phonon_code(const PairArgs &a, int q0, int q1, int q2) {
    if (q0 % a.mesh_ref == 0 && q1 % a.mesh_ref == 0 && q2 % a.mesh_ref == 0)
        return ((q2 / a.mesh_ref) * a.Nph1 + q1 / a.mesh_ref) * a.Nph0 + q0 / a.mesh_ref;
    return -((q2 * a.Nel1 + q1) * a.Nel0 + q0) - 1;
}
__device__ __forceinline__ PhRow phonon_ptrs(const PairArgs &a, int code) {
    PhRow r;
    if (code >= 0) {
        r.uq = a.u + (size_t)code * a.nbr * a.nbr;
        r.omq = a.om + (size_t)code * a.nbr;
        if (a.lp) r.lq = a.lp + (size_t)code * a.nbr;
    } else {
        const int iq = -code - 1;
        r.uq = a.uf + (size_t)iq * a.nbr * a.nbr;
        r.omq = a.omf + (size_t)iq * a.nbr;
        if (a.lpf) r.lq = a.lpf + (size_t)iq * a.nbr;
    }
    return r;
}
__device__ __forceinline__ void epi_row_setup(const PairArgs &a, int row, int q0, int q1, int q2, int lk, EpiRow &e) {
    if (row >= a.nwv) return;
    e.ph = phonon_code(a, q0, q1, q2);
    e.nn = a.nwin[row];
    // wband holds nb entries per row (0 past the in-window count)
    e.na = lk < a.nb ? a.wband[row * a.nb + lk] : 0;
    e.nb = lk + 4 < a.nb ? a.wband[row * a.nb + lk + 4] : 0;
    e.off = a.offs[row];
}
template <int NBM>
__device__ __forceinline__ void epi_load_v(const PairArgs &a, int row, const EpiRow &e, int lk, double2 (&va)[NBM],
                                           double2 (&vb)[NBM]) {
    const int nb = a.nb;
    const bool ha = lk < e.nn, hb = lk + 4 < e.nn;
    const double2 *Va = a.Vkp + ((size_t)row * nb + e.na) * nb, *Vb = a.Vkp + ((size_t)row * nb + e.nb) * nb;
#pragma unroll
    for (int np = 0; np < NBM; ++np) {
        va[np] = np < nb && ha ? Va[np] : make_double2(0.0, 0.0);
        vb[np] = np < nb && hb ? Vb[np] : make_double2(0.0, 0.0);
    }
}
// |g|^2 of in-window ranks lk (ga) and, with TWO, lk + 4 (gb) of the lane's
// row over the branches; the caller knows lk < e.nn.  Without TWO (no lane of
// the wave has a second rank) the second rank's work is left out.
template <int NBM, int NBRM, bool TWO>
__device__ __forceinline__ void epi_pair_rows(const PairArgs &a, int st, int row, const double2 *Yr, const EpiRow &e,
                                              int lk, double (&ga)[NBRM], double (&gb)[NBRM]) {
    const int nb = a.nb, nbr = a.nbr;
    const PhRow ph = phonon_ptrs(a, e.ph);
    const double2 *Ukm = ph.lq ? a.Ukb + ((size_t)a.st_kb[st] * a.nb + a.st_m[st]) * a.nb : nullptr;
    double2 va[NBM], vb[NBM];
    epi_load_v<NBM>(a, row, e, lk, va, vb);
    double2 za[NBRM], zb[NBRM];
#pragma unroll
    for (int sp = 0; sp < NBRM; ++sp) {
        double2 ta = make_double2(0.0, 0.0), tb = make_double2(0.0, 0.0);
        if (sp < nbr)
#pragma unroll
            for (int np = 0; np < NBM; ++np)
                if (np < nb) {
                    const double2 y = Yr[np + nb * sp];
                    ta.x = fma(va[np].x, y.x, ta.x);
                    ta.x = fma(va[np].y, y.y, ta.x);
                    ta.y = fma(va[np].x, y.y, ta.y);
                    ta.y = fma(-va[np].y, y.x, ta.y);
                    if constexpr (TWO) {
                        tb.x = fma(vb[np].x, y.x, tb.x);
                        tb.x = fma(vb[np].y, y.y, tb.x);
                        tb.y = fma(vb[np].x, y.y, tb.y);
                        tb.y = fma(-vb[np].y, y.x, tb.y);
                    }
                }
        za[sp] = ta;
        zb[sp] = tb;
    }
    double2 ua = make_double2(0.0, 0.0), ub = make_double2(0.0, 0.0);
    if (ph.lq)
#pragma unroll
        for (int i = 0; i < NBM; ++i)
            if (i < nb) {
                const double2 t = cmulc(va[i], Ukm[i]);
                ua.x += t.x;
                ua.y += t.y;
                if constexpr (TWO) {
                    const double2 t2 = cmulc(vb[i], Ukm[i]);
                    ub.x += t2.x;
                    ub.y += t2.y;
                }
            }
#pragma unroll
    for (int s = 0; s < NBRM; ++s) {
        ga[s] = gb[s] = 0.0;
        if (s < nbr) {
            double2 gA = make_double2(0.0, 0.0), gB = make_double2(0.0, 0.0);
#pragma unroll
            for (int sp = 0; sp < NBRM; ++sp)
                if (sp < nbr) {
                    const double2 u = ph.uq[s * nbr + sp];
                    cfma(gA, u, za[sp]);
                    if constexpr (TWO) cfma(gB, u, zb[sp]);
                }
            if (ph.lq) {
                cfma(gA, ph.lq[s], ua);
                if constexpr (TWO) cfma(gB, ph.lq[s], ub);
            }
            // one division for both ranks (the mixed path's tolerance covers
            // the reordered rounding of 0.5 |g|^2 / omega * unit)
            const double w = ph.omq[s];
            if (w != 0.0) {
                const double f = 0.5 * kG2Unit / w;
                ga[s] = (gA.x * gA.x + gA.y * gA.y) * f;
                if constexpr (TWO) gb[s] = (gB.x * gB.x + gB.y * gB.y) * f;
            }
        }
    }
}

// Byte offset of the fused int8 tile's column scales in its LDS: past the two
// B chunk buffers and the digit and cell tables, and past the epilogue's Y
// tile (64 rows of 16 CB + 1 complex), which reuses the buffers and tables.
__host__ __device__ constexpr size_t pair_i8_scale_off(int CB, int L, int nkc) {
    const size_t tables = (size_t)2 * CB * 2 * kNS * 64 * 16 + (size_t)((L + 1) & ~1) * 8 + (size_t)nkc * 32 * 4;
    const size_t ytile = (size_t)64 * (16 * CB + 1) * 16;
    return ((tables > ytile ? tables : ytile) + 15) & ~(size_t)15;
}
// Fused mixed-precision pair tile (state st, k' rows 64 rblk .. 64 rblk +
// 63): the cell-pair contraction of i8_pair_core, FP64 recombination of the
// digit levels with the column scale of B, and the FP64 rotation / |g|^2
// epilogue.  One wave owns 16 k' rows and all CP columns of the state.  The
// caller's workgroup is 256 threads with the dynamic LDS of fused_launch; the
// LDS is free again when the call returns.
template <int CB, int NBM, int NBRM>
__device__ __forceinline__ void pair_i8_tile(const PairArgs &a, const uint2 *__restrict__ Tdig,
                                             const unsigned *__restrict__ Rpk, const v4i *__restrict__ bsl,
                                             const double *__restrict__ bscale, int nkc, int st, int rblk) {
    constexpr int NS = kNS;
    constexpr int CP = 16 * CB;
    constexpr int BCH = CB * 2 * NS * 64;  // v4i per K2 chunk of one state
    constexpr bool STAGED = NBM > 0 && NBM <= 8;
    extern __shared__ v4i smem8[];
    uint2 *Tsh = reinterpret_cast<uint2 *>(smem8 + 2 * BCH);
    unsigned *Rsh = reinterpret_cast<unsigned *>(Tsh + ((a.L + 1) & ~1));
    // the state's column scales, past the tables and the epilogue's Y tile
    double *Ssh = reinterpret_cast<double *>(reinterpret_cast<char *>(smem8) + pair_i8_scale_off(CB, a.L, nkc));
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int li = lane & 15, lg = lane >> 4;
    const int row_base = rblk * 64 + wave * 16;
    // The prologue's global loads complete in issue order, so the row data
    // whose uses come first (this lane's k' row and, staged, the tile's
    // in-window counts and output range) are issued before the chunk-0 copy
    // and the table loads, and the tables are loaded into registers and
    // written to LDS only after the rows' dependent loads are issued.
    const int k0i = a.st_kint[3 * st], k1i = a.st_kint[3 * st + 1], k2i = a.st_kint[3 * st + 2];
    const v4i *bst = bsl + (size_t)st * nkc * BCH;
    const int r = row_base + li;
    int kr0 = 0, kr1 = 0, kr2 = 0;
    if (r < a.nwv) {
        kr0 = a.kp_int[3 * r];
        kr1 = a.kp_int[3 * r + 1];
        kr2 = a.kp_int[3 * r + 2];
    }
    const int R0 = rblk * 64;  // first k' row of the tile (R0 < nwv)
    int nraw = 0;              // in-window count of tile row `lane`
    int64_t B0 = 0, B1 = 0;    // the tile's range of the output row
    if constexpr (STAGED) {
        nraw = R0 + lane < a.nwv ? a.nwin[R0 + lane] : 0;
        B0 = a.offs[R0];
        B1 = R0 + 64 < a.nwv ? a.offs[R0 + 64] : a.nprocs;
    }
    i8_copy_chunk<CB>(bst, BCH, 0, smem8);
    // L <= 255 and CP <= 256 (one entry per thread); the cell table in two
    // registers per thread, any rest written by a loop below
    const int nrp = nkc * 32;
    const uint2 tv = tid < a.L ? Tdig[tid] : make_uint2(0u, 0u);
    const unsigned rv0 = tid < nrp ? Rpk[tid] : 0u, rv1 = tid + 256 < nrp ? Rpk[tid + 256] : 0u;
    const double sv = tid < CP ? bscale[(size_t)st * CP + tid] : 0.0;
    unsigned cpk = 0;
    if (r < a.nwv) {
        const int q0 = pos_mod(kr0 - k0i, a.Nel0);
        const int q1 = pos_mod(kr1 - k1i, a.Nel1);
        const int q2 = pos_mod(kr2 - k2i, a.Nel2);
        cpk = (unsigned)((q0 * a.Lf0) % a.L) | ((unsigned)((q1 * a.Lf1) % a.L) << 8) |
              ((unsigned)((q2 * a.Lf2) % a.L) << 16);
    }
    EpiRow er;
    int re = 0;  // the lane's epilogue row
    if constexpr (STAGED) {
        // epilogue rows in the order in-window count descending, then row:
        // the rows with a second rank (count > 4) fill the first waves' slots,
        // the others' waves leave the second rank out; the waves' shares of
        // the order rotate with the tile, spreading the second ranks over the
        // SIMDs.  Slot p goes to lane li of wave (p / 16 - rblk - st) mod 4.
        const int nnl = min(nraw, NBM);  // (at most nb <= NBM)
        const int p = 16 * ((wave + rblk + st) & 3) + li;
        int before = 0, rank = 0;
        unsigned long long mv = 0;
#pragma unroll
        for (int v = NBM; v >= 0; --v) {
            const unsigned long long m = __ballot(nnl == v);
            const int c = __popcll(m);
            if (p >= before && p < before + c) {
                mv = m;
                rank = p - before;
            }
            before += c;
        }
        // the rank-th set bit of mv: the largest lo with popc(mv below lo) <= rank
        int lo = 0;
#pragma unroll
        for (int step = 32; step > 0; step >>= 1)
            if (__popcll(mv & ((1ull << (lo + step)) - 1)) <= rank) lo += step;
        re = R0 + lo;
        int p0 = 0, p1 = 0, p2 = 0;
        if (re < a.nwv) {
            p0 = pos_mod(a.kp_int[3 * re] - k0i, a.Nel0);
            p1 = pos_mod(a.kp_int[3 * re + 1] - k1i, a.Nel1);
            p2 = pos_mod(a.kp_int[3 * re + 2] - k2i, a.Nel2);
        }
        epi_row_setup(a, re, p0, p1, p2, lg, er);
    }
    if (tid < a.L) Tsh[tid] = tv;
    if (tid < nrp) Rsh[tid] = rv0;
    if (tid + 256 < nrp) Rsh[tid + 256] = rv1;
    for (int i = tid + 512; i < nrp; i += 256) Rsh[i] = Rpk[i];
    if (tid < CP) Ssh[tid] = sv;
    v4i acc[CB][NS][2];
    i8_pair_core<CB>(bst, BCH, nkc, Tsh, Rsh, cpk, a.Lmul, -a.L, smem8, acc);
    // ---- recombine levels in FP64 and run the FP64 epilogue ----
    constexpr int YS = CP + 1;
    double2 *Y = reinterpret_cast<double2 *>(smem8) + wave * 16 * YS;
    double *outrow = a.out + a.st_slot[st] * a.nprocs;
    double wl[NS];
    wl[0] = 1.0 / (127.0 * 127.0);
#pragma unroll
    for (int l = 1; l < NS; ++l) wl[l] = wl[l - 1] / 254.0;
#pragma unroll
    for (int c = 0; c < CB; ++c) {
        const double s = Ssh[c * 16 + li];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            double re = 0.0, im = 0.0;
#pragma unroll
            for (int l = NS - 1; l >= 0; --l) {
                re += wl[l] * (double)acc[c][l][0][r];
                im += wl[l] * (double)acc[c][l][1][r];
            }
            Y[(4 * lg + r) * YS + c * 16 + li] = make_double2(re * s, im * s);
        }
    }
    __syncthreads();
    if constexpr (STAGED) {
        // tile row i's Y at Yt + i YS (the waves' row blocks are consecutive)
        const double2 *Yt = reinterpret_cast<const double2 *>(smem8);
        double ga[NBRM], gb[NBRM];
        const bool has = re < a.nwv && lg < er.nn;
        if (__any(lg + 4 < er.nn)) {
            if (has) epi_pair_rows<NBM, NBRM, true>(a, st, re, Yt + (re - R0) * YS, er, lg, ga, gb);
        } else {
            if (has) epi_pair_rows<NBM, NBRM, false>(a, st, re, Yt + (re - R0) * YS, er, lg, ga, gb);
        }
        // the |g|^2 of the tile's 64 rows are one contiguous range of the
        // output: staged over the tile's Y rows once every wave has read
        // them, then written out by consecutive lanes, so HBM sees whole
        // lines rather than 8-byte pieces at the branch stride
        __syncthreads();
        double *stage = reinterpret_cast<double *>(smem8);
        if (has) {
            const int nbr = a.nbr;
            double *o = stage + (er.off - B0) + (int64_t)lg * nbr;
#pragma unroll
            for (int s = 0; s < NBRM; ++s)
                if (s < nbr) o[s] = ga[s];
            if (lg + 4 < er.nn)
#pragma unroll
                for (int s = 0; s < NBRM; ++s)
                    if (s < nbr) o[4 * nbr + s] = gb[s];
        }
        __syncthreads();
        const int cnt = (int)(B1 - B0);
        for (int i = tid; i < cnt; i += 256) __builtin_nontemporal_store(stage[i], outrow + B0 + i);
    } else {
        ibzk_epilogue_row<NBM, NBRM>(a, Y, YS, row_base + li, li, lg, st, k0i, k1i, k2i, outrow);
    }
}
template <int CB, int NBM, int NBRM>
__global__ __launch_bounds__(256, pair_blocks_per_cu(CB, NBM)) void k_pair_i8(PairArgs a, const uint2 *__restrict__ Tdig,
                                                    const unsigned *__restrict__ Rpk, const v4i *__restrict__ bsl,
                                                    const double *__restrict__ bscale, int nkc) {
    pair_i8_tile<CB, NBM, NBRM>(a, Tdig, Rpk, bsl, bscale, nkc, blockIdx.y, blockIdx.x);
}

// One group's share of a multi-material pair launch: its arguments (plan,
// group tables, operands, output), the mixed-precision tables, its row blocks
// and the first tile of its range in the launch's tile numbering (tiles of a
// group: state-major, then row block).
struct PairJob {
    PairArgs a;
    const uint2 *Tdig;
    const unsigned *Rpk;
    const v4i *bsl;
    const double *bscale;
    int nkc, nrb;
    int64_t tile0;
};
// the job whose tile range holds tile t (jobs ascending in tile0, tile0[0] = 0)
__device__ __forceinline__ int job_of(const PairJob *__restrict__ jobs, int njobs, int64_t t) {
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].tile0 <= t) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}
// Multi-material forms of the fused pair kernels: persistent workgroups walk
// the tiles t = blockIdx.x, blockIdx.x + gridDim.x, ... of all jobs of one
// launch (the groups of every plan whose pair stage has this kernel shape);
// the barrier frees the LDS for the next tile.
template <int CB, int NBM, int NBRM>
__global__ __launch_bounds__(256, pair_blocks_per_cu(CB, NBM)) void k_pair_i8_jobs(const PairJob *__restrict__ jobs,
                                                                                    int njobs, int64_t ntiles) {
    for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const PairJob &J = jobs[job_of(jobs, njobs, t)];
        const int64_t local = t - J.tile0;
        pair_i8_tile<CB, NBM, NBRM>(J.a, J.Tdig, J.Rpk, J.bsl, J.bscale, J.nkc, (int)(local / J.nrb),
                                    (int)(local % J.nrb));
        __syncthreads();
    }
}
template <int CB, int NBM, int NBRM, bool PACKED>
__global__ __launch_bounds__(256) void k_pair_fp64_jobs(const PairJob *__restrict__ jobs, int njobs, int64_t ntiles) {
    for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const PairJob &J = jobs[job_of(jobs, njobs, t)];
        const int64_t local = t - J.tile0;
        pair_fp64_tile<CB, NBM, NBRM, PACKED>(J.a, (int)(local / J.nrb), (int)(local % J.nrb));
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// General ibzk path for any numwannbands * numbranches: the Rp contraction
// runs in column tiles of 16*CB (grid.z) and writes Y[st][k'][CPT] to HBM for
// a chunk of states; k_ibzk_epi then applies the rotations and |g|^2.  Used
// when nb*nbr exceeds what the fused kernels hold in registers.
// ---------------------------------------------------------------------------

// phase coefficients of row r for state st: (q . e_d) L / N_d, q = k' - k
__device__ __forceinline__ void ibzk_coef(const PairArgs &a, int st, int r, int &c0, int &c1, int &c2) {
    c0 = c1 = c2 = 0;
    if (r < a.nwv) {
        c0 = (pos_mod(a.kp_int[3 * r] - a.st_kint[3 * st], a.Nel0) * a.Lf0) % a.L;
        c1 = (pos_mod(a.kp_int[3 * r + 1] - a.st_kint[3 * st + 1], a.Nel1) * a.Lf1) % a.L;
        c2 = (pos_mod(a.kp_int[3 * r + 2] - a.st_kint[3 * st + 2], a.Nel2) * a.Lf2) % a.L;
    }
}

// FP64: block = 4 waves x 16*RB rows, grid (rows, states of the chunk, column tiles)
template <int CB, int RB>
__global__ __launch_bounds__(256) void k_ibzk_y_fp64(PairArgs a, int st0, double2 *__restrict__ Y) {
    constexpr int KC = 32, CPt = 16 * CB;
    extern __shared__ double2 smem[];
    double2 *buf0 = smem, *buf1 = smem + KC * CPt;
    double2 *Tsh = smem + 2 * KC * CPt;
    int4 *Rsh = reinterpret_cast<int4 *>(Tsh + a.L);
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int li = lane & 15, lk = lane >> 4;
    const int stl = blockIdx.y, st = st0 + stl, col0 = blockIdx.z * CPt;
    const int row_base = blockIdx.x * 64 * RB + wave * 16 * RB;
    for (int i = tid; i < a.L; i += 256) Tsh[i] = a.T[i];
    for (int i = tid; i < a.Kp / 2; i += 256) Rsh[i] = make_int4(a.Rq[4 * i], a.Rq[4 * i + 1], a.Rq[4 * i + 2], 0);
    int coef[RB][3];
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) ibzk_coef(a, st, row_base + rb * 16 + li, coef[rb][0], coef[rb][1], coef[rb][2]);
    __syncthreads();
    d4 cr[RB][CB], ci[RB][CB];
    const int L = a.L;
    auto aop = [&](int rb, int k) { return pair_phase(Tsh, Rsh, coef[rb], L, k); };
    gemm_core_f64<CB, RB, KC>(a.h + (size_t)st * a.Kp * a.CP + col0, a.CP, a.Kp, aop, buf0, buf1, cr, ci);
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
        for (int c = 0; c < CB; ++c)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = row_base + rb * 16 + lk + 4 * r;
                if (row < a.nwv)
                    Y[((size_t)stl * a.nwv + row) * a.CP + col0 + c * 16 + li] = make_double2(cr[rb][c][r], ci[rb][c][r]);
            }
}

// Mixed: block = 4 waves x 16 rows, the cell-pair contraction of i8_pair_core
// over the block's CB column blocks of B digits in bsl
// ([st][kc][CBT col blocks][Re, Im][digit][lane]).
template <int CB>
__global__ __launch_bounds__(256, pair_blocks_per_cu(CB)) void k_ibzk_y_i8(PairArgs a, int st0, const uint2 *__restrict__ Tdig,
                                                      const unsigned *__restrict__ Rpk, const v4i *__restrict__ bsl,
                                                      const double *__restrict__ bscale, int nkc, int CBT,
                                                      double2 *__restrict__ Y) {
    constexpr int NS = kNS;
    constexpr int BCH = CB * 2 * NS * 64;
    extern __shared__ v4i smem8[];
    uint2 *Tsh = reinterpret_cast<uint2 *>(smem8 + 2 * BCH);
    unsigned *Rsh = reinterpret_cast<unsigned *>(Tsh + ((a.L + 1) & ~1));
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int li = lane & 15, lg = lane >> 4;
    const int stl = blockIdx.y, st = st0 + stl, cb0 = blockIdx.z * CB;
    const int row_base = blockIdx.x * 64 + wave * 16;
    const size_t bkc = (size_t)CBT * 2 * NS * 64;
    const v4i *bst = bsl + (size_t)st * nkc * bkc + (size_t)cb0 * 2 * NS * 64;
    i8_copy_chunk<CB>(bst, bkc, 0, smem8);
    for (int i = tid; i < a.L; i += 256) Tsh[i] = Tdig[i];
    for (int i = tid; i < nkc * 32; i += 256) Rsh[i] = Rpk[i];
    int c0, c1, c2;
    ibzk_coef(a, st, row_base + li, c0, c1, c2);
    const unsigned cpk = (unsigned)c0 | ((unsigned)c1 << 8) | ((unsigned)c2 << 16);
    v4i acc[CB][NS][2];
    i8_pair_core<CB>(bst, bkc, nkc, Tsh, Rsh, cpk, a.Lmul, -a.L, smem8, acc);
    double wl[NS];
    wl[0] = 1.0 / (127.0 * 127.0);
#pragma unroll
    for (int l = 1; l < NS; ++l) wl[l] = wl[l - 1] / 254.0;
    const double *sc = bscale + (size_t)st * a.CP + cb0 * 16;
#pragma unroll
    for (int c = 0; c < CB; ++c) {
        const double s = sc[c * 16 + li];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = row_base + 4 * lg + r;
            double re = 0.0, im = 0.0;
#pragma unroll
            for (int l = NS - 1; l >= 0; --l) {
                re += wl[l] * (double)acc[c][l][0][r];
                im += wl[l] * (double)acc[c][l][1][r];
            }
            if (row < a.nwv) Y[((size_t)stl * a.nwv + row) * a.CP + (cb0 + c) * 16 + li] = make_double2(re * s, im * s);
        }
    }
}

// General-path epilogue for nb <= 16 and nbr <= 12: the fused kernels' row
// epilogue (ibzk_epilogue_row) on Y rows read from HBM, Y[st - st0][k'][CP].
// Block = 4 waves x 16 rows of one state, 4 lanes per row (final-band ranks
// lk, lk + 4, ...); NBM / NBRM as in ibzk_epilogue_row.
template <int NBM, int NBRM>
__global__ __launch_bounds__(256) void k_ibzk_epi(PairArgs a, int st0, const double2 *__restrict__ Y) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int stl = blockIdx.y, st = st0 + stl;
    const int row = blockIdx.x * 64 + wave * 16 + li;
    double *outrow = a.out + a.st_slot[st] * a.nprocs;
    ibzk_epilogue_row<NBM, NBRM>(a, Y + (size_t)stl * a.nwv * a.CP, a.CP, row, row, lk, st, a.st_kint[3 * st],
                                 a.st_kint[3 * st + 1], a.st_kint[3 * st + 2], outrow);
}

// General-path epilogue, one wave per k' row (four rows of one state per
// workgroup, no workgroup barriers): the lanes take the (rank, sp) entries of
//   Z(rank, sp) = sum_np conj(U_k'(n_rank, np)) Y(np + nb sp)
// put Z (and the polar overlaps) in the wave's LDS area, then take the
// (rank, s) entries of
//   g(rank, s)  = sum_sp u_q(s, sp) Z(rank, sp) [+ lp(q,s) sum_i conj(U_k'(n,i)) U_k(m,i)]
// and write |g|^2 with consecutive lanes on consecutive outputs.  Products
// in the order of ibzk_epilogue_row.  STAGE: the row's Y and the U_k' rows of
// its in-window bands are first copied into the wave's LDS area with
// consecutive lanes on consecutive elements, so the Z products read LDS
// instead of issuing two scattered global loads per multiply-add; without
// STAGE (shapes whose staging would not leave two workgroups per CU) they are
// read straight from memory.  LDS: epi_wave_lds.
template <bool STAGE>
__global__ __launch_bounds__(256) void k_ibzk_epi_wave(PairArgs a, int st0, const double2 *__restrict__ Y) {
    extern __shared__ double2 zsm[];
    const int nb = a.nb, nbr = a.nbr;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int stl = blockIdx.y, st = st0 + stl;
    const int row = blockIdx.x * 4 + wave;
    if (row >= a.nwv) return;  // no workgroup barrier follows
    const int ncol = nb * nbr;
    double2 *Zs = zsm + wave * (STAGE ? 2 * ncol + nb * nb + nb : ncol + nb), *un = Zs + ncol;
    double2 *Ys = un + nb, *Vs = Ys + ncol;  // STAGE only: the row's Y, its U_k' rows [rank][np]
    const int q0 = pos_mod(a.kp_int[3 * row] - a.st_kint[3 * st], a.Nel0);
    const int q1 = pos_mod(a.kp_int[3 * row + 1] - a.st_kint[3 * st + 1], a.Nel1);
    const int q2 = pos_mod(a.kp_int[3 * row + 2] - a.st_kint[3 * st + 2], a.Nel2);
    const PhRow ph = phonon_row(a, q0, q1, q2);
    const int nn = a.nwin[row];
    const double2 *Yr = Y + ((size_t)stl * a.nwv + row) * a.CP;
    const double2 *Vr = a.Vkp + (size_t)row * nb * nb;
    if constexpr (STAGE) {
        for (int i = lane; i < ncol; i += 64) Ys[i] = Yr[i];
        for (int i = lane; i < nn * nb; i += 64) {
            const int rank = i / nb, np = i - rank * nb;
            Vs[i] = Vr[(size_t)a.wband[row * nb + rank] * nb + np];
        }
        // the wave's LDS writes complete before its reads (a wave's LDS
        // requests are served in order; the wait keeps it explicit)
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
        __builtin_amdgcn_wave_barrier();
    }
    auto vrow = [&](int rank) {
        if constexpr (STAGE) return (const double2 *)(Vs + rank * nb);
        else return Vr + (size_t)a.wband[row * nb + rank] * nb;
    };
    const double2 *Ysrc = STAGE ? (const double2 *)Ys : Yr;
    for (int i = lane; i < nn * nbr; i += 64) {
        const int rank = i % nn, sp = i / nn;
        const double2 *V = vrow(rank);
        double2 z = make_double2(0.0, 0.0);
        for (int np = 0; np < nb; ++np) {
            const double2 v = V[np];
            const double2 y = Ysrc[np + nb * sp];
            z.x = fma(v.x, y.x, z.x);  // conj(v) * y
            z.x = fma(v.y, y.y, z.x);
            z.y = fma(v.x, y.y, z.y);
            z.y = fma(-v.y, y.x, z.y);
        }
        Zs[rank + nb * sp] = z;
    }
    if (ph.lq) {
        const double2 *Ukm = a.Ukb + ((size_t)a.st_kb[st] * nb + a.st_m[st]) * nb;
        for (int rank = lane; rank < nn; rank += 64) {
            const double2 *V = vrow(rank);
            double2 z = make_double2(0.0, 0.0);
            for (int i = 0; i < nb; ++i) {
                const double2 t = cmulc(V[i], Ukm[i]);
                z.x += t.x;
                z.y += t.y;
            }
            un[rank] = z;
        }
    }
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    __builtin_amdgcn_wave_barrier();
    double *o = a.out + a.st_slot[st] * a.nprocs + a.offs[row];
    for (int i = lane; i < nn * nbr; i += 64) {
        const int rank = i / nbr, s = i % nbr;
        double2 g = make_double2(0.0, 0.0);
        for (int sp = 0; sp < nbr; ++sp) cfma(g, ph.uq[s * nbr + sp], Zs[rank + nb * sp]);
        if (ph.lq) cfma(g, ph.lq[s], un[rank]);
        const double w = ph.omq[s];
        o[i] = w != 0.0 ? 0.5 * (g.x * g.x + g.y * g.y) / w * kG2Unit : 0.0;
    }
}
// Dynamic LDS of k_ibzk_epi_wave (4 waves) and whether it stages Y and U_k'
// (when that leaves room for two workgroups per CU).
bool epi_wave_stage(int64_t nb, int64_t nbr) { return (size_t)4 * (2 * nb * nbr + nb * nb + nb) * 16 <= 80 * 1024; }
size_t epi_wave_lds(int64_t nb, int64_t nbr) {
    return (size_t)4 * (epi_wave_stage(nb, nbr) ? 2 * nb * nbr + nb * nb + nb : nb * nbr + nb) * sizeof(double2);
}

// ---------------------------------------------------------------------------
// Mixed-precision Re contraction (gkRp) in cell-pair form: the folded gwann
// (gw_pair) split once per context into kNS int8 digits per (row i = (np,mp),
// slice = (sp,Rp)) with a per-row scale, the real cos/sin phases split per
// batch; 10 digit products x (Re, Im) of G on the int8 matrix cores with exact
// int32 accumulation, recombined in FP64 into the FP64 g(k,Rp) layout.
// ---------------------------------------------------------------------------

// gscale[slice][i] = max over K2 of max(|Re G|, |Im G|) (1 for zero rows).
__global__ void k_gwann_scale(int nb2, int nwsk, int nslices, int nrb, int npair, const int2 *__restrict__ pairs,
                              const double *__restrict__ invdeg, const double2 *__restrict__ gw,
                              double *__restrict__ gscale) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nslices * nrb * 16) return;
    const int i = (int)(t % (nrb * 16));
    const int64_t slice = t / (nrb * 16);
    double m = 0.0;
    if (i < nb2) {
        const double2 *gs = gw + (size_t)slice * nb2 * nwsk;
        for (int k2 = 0; k2 < 2 * npair; ++k2) {
            const double2 v = gw_pair(gs, nb2, i, k2, npair, pairs, invdeg);
            m = fmax(m, fmax(fabs(v.x), fabs(v.y)));
        }
    }
    gscale[t] = m > 0.0 ? m : 1.0;
}

// Gdig[slice][rb][kc][p][s][lane]: digits of G / scale in the int8 MFMA
// A-operand order (lane i + 16 g holds rows rb*16 + i, K2 = 64 kc + 16 g .. +15),
// p = Re, Im.
__global__ void k_gwann_digits(int nb2, int nwsk, int nslices, int nrb, int nkc, int npair,
                               const int2 *__restrict__ pairs, const double *__restrict__ invdeg,
                               const double2 *__restrict__ gw, const double *__restrict__ gscale,
                               v4i *__restrict__ gdig) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t total = (int64_t)nslices * nrb * nkc * 64;
    if (t >= total) return;
    const int lane = (int)(t % 64);
    int64_t r = t / 64;
    const int kc = (int)(r % nkc);
    r /= nkc;
    const int rb = (int)(r % nrb);
    const int64_t slice = r / nrb;
    const int i = rb * 16 + (lane & 15), g = lane >> 4;
    const double inv = 1.0 / gscale[slice * nrb * 16 + i];
    const double2 *gs = gw + (size_t)slice * nb2 * nwsk;
    unsigned ure[kNS][4] = {}, uim[kNS][4] = {};
#pragma unroll
    for (int b = 0; b < 16; ++b) {
        const int k2 = kc * 64 + g * 16 + b;
        double2 x = make_double2(0.0, 0.0);
        if (i < nb2) x = gw_pair(gs, nb2, i, k2, npair, pairs, invdeg);
        signed char dr[kNS], di[kNS];
        split_digits<kNS>(x.x * inv, dr);
        split_digits<kNS>(x.y * inv, di);
#pragma unroll
        for (int sd = 0; sd < kNS; ++sd) { put_byte(ure[sd], b, dr[sd]); put_byte(uim[sd], b, di[sd]); }
    }
    v4i *o = gdig + ((((size_t)slice * nrb + rb) * nkc + kc) * 2) * kNS * 64 + lane;
#pragma unroll
    for (int sd = 0; sd < kNS; ++sd) {
        o[(0 * kNS + sd) * 64] = as_v4i(ure[sd]);
        o[(1 * kNS + sd) * 64] = as_v4i(uim[sd]);
    }
}

// Bphi[kc][cb][s][lane]: digits of the real cell-pair phases phi[K2][col]
// (|phi| <= 1, no scale) in the int8 MFMA B-operand order.
__global__ void k_phase_digits(int nkc, int ncb, const double *__restrict__ phi, int ldphi, v4i *__restrict__ bphi) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nkc * ncb * 64) return;
    const int lane = t % 64, cb = (t / 64) % ncb, kc = t / (64 * ncb);
    const int col = cb * 16 + (lane & 15), g = lane >> 4;
    unsigned u[kNS][4] = {};
#pragma unroll
    for (int b = 0; b < 16; ++b) {
        const int k = kc * 64 + g * 16 + b;
        signed char d[kNS];
        split_digits<kNS>(phi[(size_t)k * ldphi + col], d);
#pragma unroll
        for (int sd = 0; sd < kNS; ++sd) put_byte(u[sd], b, d[sd]);
    }
    v4i *o = bphi + ((size_t)kc * ncb + cb) * kNS * 64 + lane;
#pragma unroll
    for (int sd = 0; sd < kNS; ++sd) o[sd * 64] = as_v4i(u[sd]);
}

// Persistent workgroups: work item = (sp,Rp) slice x round of 4 / CS row
// blocks; wave w takes row block (4 / CS) round + w / CS and column group
// w % CS (NCBW column blocks of 16 k each, CS NCBW blocks per launch), so the
// CS waves of one row block share its A loads through the caches and one pass
// over the digits of gwann (the launch's HBM traffic) serves 16 CS NCBW k.
// Workgroup b takes items b, b + grid, ... and walks them as one stream of K
// chunks (item j, chunk kc), so the copies of the next chunk run across item
// boundaries and an item's recombination and stores overlap the next item's
// first loads.  out[k][slice][i] FP64 complex.  The phase digits (the same
// for every slice) go global -> LDS one chunk ahead into two buffers shared by
// the 4 waves; the wave's A digits (its rows of the folded gwann, streamed
// from HBM) are loaded one chunk ahead into the second of two register sets.
template <int NCBW, int CS>
__global__ __launch_bounds__(256, NCBW <= 4 ? 2 : 1) void k_gkRp_i8(const v4i *__restrict__ gdig, const double *__restrict__ gscale,
                                                    int nb2, int nrb, int nkc, int nslices,
                                                    const v4i *__restrict__ bphi, int nk, double2 *__restrict__ out,
                                                    int64_t out_kstride) {
    constexpr int NS = kNS;
    constexpr int NCB = NCBW * CS;      // column blocks of the launch
    constexpr int RBI = 4 / CS;         // row blocks of an item
    static_assert(RBI * CS == 4, "the 4 waves split into row blocks x column groups");
    constexpr int BCH = NCB * NS * 64;  // v4i per chunk of phase digits
    static_assert(BCH % 256 == 0, "a chunk is a whole number of 1 KiB rows per wave");
    extern __shared__ v4i bsm[];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int li = lane & 15, lg = lane >> 4;
    const int cg = wave % CS;           // the wave's column group
    const int nround = (nrb + RBI - 1) / RBI;
    const int64_t nitems = (int64_t)nslices * nround;
    if ((int64_t)blockIdx.x >= nitems) return;  // uniform over the workgroup
    const int64_t nmine = (nitems - blockIdx.x + gridDim.x - 1) / gridDim.x;
    const int64_t nsteps = nmine * nkc;
    // step st: item j = st / nkc, chunk kc = st % nkc
    auto item = [&](int64_t st, int &slice, int &rb) {
        const int64_t it = blockIdx.x + (st / nkc) * (int64_t)gridDim.x;
        slice = (int)(it / nround);
        rb = (int)(it % nround) * RBI + wave / CS;
    };
    auto copy_b = [&](int64_t st) {
        const v4i *src = bphi + (size_t)(st % nkc) * BCH;
        v4i *dst = bsm + (st & 1) * BCH;
#pragma unroll
        for (int q = 0; q < BCH / 256; ++q) {
            const int r = 4 * q + wave;
            lds_dma16(src + r * 64 + lane, dst + r * 64);
        }
    };
    auto load_a = [&](int64_t st, v4i (&Ar)[NS], v4i (&Ai)[NS]) {
        int slice, rb;
        item(st, slice, rb);
        if (rb >= nrb) return;
        const int kc = (int)(st % nkc);
        const v4i *ga = gdig + (((size_t)slice * nrb + rb) * nkc) * 2 * NS * 64 + lane;
#pragma unroll
        for (int sd = 0; sd < NS; ++sd) {
            Ar[sd] = ga[((size_t)kc * 2 * NS + 0 * NS + sd) * 64];
            Ai[sd] = ga[((size_t)kc * 2 * NS + 1 * NS + sd) * 64];
        }
    };
    v4i acc[NCBW][NS][2];
    auto zero_acc = [&]() {
#pragma unroll
        for (int c = 0; c < NCBW; ++c)
#pragma unroll
            for (int l = 0; l < NS; ++l) { acc[c][l][0] = v4i{0, 0, 0, 0}; acc[c][l][1] = v4i{0, 0, 0, 0}; }
    };
    auto chunk = [&](int64_t st, const v4i (&Ar)[NS], const v4i (&Ai)[NS]) {
        const v4i *bb = bsm + (st & 1) * BCH + cg * NCBW * NS * 64 + lane;
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            v4i B[NCBW];
#pragma unroll
            for (int c = 0; c < NCBW; ++c) B[c] = bb[(c * NS + j) * 64];
#pragma unroll
            for (int i = 0; i < NS - j; ++i)
#pragma unroll
                for (int c = 0; c < NCBW; ++c) {
                    acc[c][i + j][0] = mfma_i8(Ar[i], B[c], acc[c][i + j][0]);
                    acc[c][i + j][1] = mfma_i8(Ai[i], B[c], acc[c][i + j][1]);
                }
        }
    };
    // row scales of the lane's 4 rows when st is an item's last chunk, loaded
    // before the next step's copies are issued (a use waits for every
    // vector-memory access issued before it)
    auto load_sc = [&](int64_t st, double (&sc)[4]) {
        int slice, rb;
        item(st, slice, rb);
        if (st % nkc != nkc - 1 || rb >= nrb) return;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int i = rb * 16 + 4 * lg + r;
            sc[r] = i < nb2 ? gscale[(size_t)slice * nrb * 16 + i] : 0.0;
        }
    };
    // the step's chunk, and after an item's last chunk its recombination,
    // stores and a fresh accumulator
    auto step = [&](int64_t st, const v4i (&Ar)[NS], const v4i (&Ai)[NS], const double (&scv)[4]) {
        int slice, rb;
        item(st, slice, rb);
        const bool live = rb < nrb;
        if (live) chunk(st, Ar, Ai);
        if (st % nkc != nkc - 1) return;
        if (live) {
            double wl[NS];
            wl[0] = 1.0 / (127.0 * 127.0);
#pragma unroll
            for (int l = 1; l < NS; ++l) wl[l] = wl[l - 1] / 254.0;
#pragma unroll
            for (int c = 0; c < NCBW; ++c) {
                const int col = (cg * NCBW + c) * 16 + li;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int i = rb * 16 + 4 * lg + r;
                    if (col < nk && i < nb2) {
                        double re = 0.0, im = 0.0;
#pragma unroll
                        for (int l = NS - 1; l >= 0; --l) {
                            re += wl[l] * (double)acc[c][l][0][r];
                            im += wl[l] * (double)acc[c][l][1][r];
                        }
                        const double sc = scv[r];
                        out[(size_t)col * out_kstride + (size_t)slice * nb2 + i] = make_double2(re * sc, im * sc);
                    }
                }
            }
        }
        zero_acc();
    };
    // step st + 1 in LDS for the whole workgroup (no LDS stores pending: the
    // barrier needs no fence; the clobber keeps LDS reads on their side)
    auto next_ready = [&]() { asm volatile("s_waitcnt vmcnt(0)\n\ts_barrier" ::: "memory"); };
    zero_acc();
    v4i Ar0[NS], Ai0[NS], Ar1[NS], Ai1[NS];
    copy_b(0);
    load_a(0, Ar0, Ai0);
    lds_dma_wait();
    __syncthreads();
    for (int64_t st = 0;; st += 2) {
        double sc0[4] = {0.0, 0.0, 0.0, 0.0};
        load_sc(st, sc0);
        if (st + 1 < nsteps) {
            copy_b(st + 1);
            load_a(st + 1, Ar1, Ai1);
        }
        step(st, Ar0, Ai0, sc0);
        if (st + 1 >= nsteps) break;
        next_ready();
        double sc1[4] = {0.0, 0.0, 0.0, 0.0};
        load_sc(st + 1, sc1);
        if (st + 2 < nsteps) {
            copy_b(st + 2);
            load_a(st + 2, Ar0, Ai0);
        }
        step(st + 1, Ar1, Ai1, sc1);
        if (st + 2 >= nsteps) break;
        next_ready();
    }
    lds_dma_wait();
    __syncthreads();
}

// ---------------------------------------------------------------------------
// Rp contraction of the IBZ-q side (gReq) as an FP64 MFMA GEMM per branch s:
//   out[q][(i, s, Re)] = sum_Rp gwann[(i,Re)][s][Rp] * phiT[Rp][q]
// rows X = i + nb2 * Re; one wave per 16-row block and all 16 NCB q columns,
// so one pass over gwann serves 16 NCB q.  Per block of 16 K rows each wave
// copies the gwann rows of its row block into its own LDS buffer and a
// quarter of the block's phase rows into the workgroup's shared one (LDS DMA,
// one block ahead, two buffers each), so the 4 waves read the phase rows from
// LDS rather than each from L2; the pair table and the per-pair weights sit
// in LDS for the whole launch, so no global load waits behind the copies in
// flight (the vector-memory counter completes in order).  Waves past
// the last row block copy phase rows and meet the barriers but do no
// products.  K2 is a multiple of 16, rows past 2 npair have zero weight and
// phases; phiR rows are 16 NCB doubles (ldphi = 16 NCB).
// ---------------------------------------------------------------------------
constexpr int kGReqKI = 16;  // K rows (8 cell pairs) per block of the K loop
// phase-row stride in LDS (doubles): consecutive rows start 128 bytes apart
// modulo 256, so the four K rows a wave reads at once fall on both halves of
// the banks
__host__ __device__ constexpr int greq_ldb(int NCB) { return 16 * NCB + 16; }
__host__ __device__ constexpr size_t greq_lds(int NCB, int K2) {
    return (size_t)4 * 2 * (kGReqKI / 2) * 2 * 16 * 16 + (size_t)2 * kGReqKI * greq_ldb(NCB) * 8 + (size_t)K2 / 2 * (16 + 8);
}
template <int NCB>
__global__ __launch_bounds__(256) void k_gReq(const double2 *__restrict__ gw, int nb2, int nwsk, int nbr, int npair,
                                              const int2 *__restrict__ pairs, const double2 *__restrict__ pwt,
                                              const double *__restrict__ phiR, int ldphi, int K2, int nq,
                                              double2 *__restrict__ out, int64_t out_qstride) {
    constexpr int KI = kGReqKI;
    constexpr int LDB = greq_ldb(NCB);
    static_assert(16 * NCB == 2 * 64, "a phase row is one 16-byte copy per lane");
    extern __shared__ double2 gsm[];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int li = lane & 15, lk = lane >> 4;
    const int s = blockIdx.y;
    const int64_t nrows = (int64_t)nb2 * nwsk;
    const int64_t rb = (int64_t)blockIdx.x * 4 + wave;
    const bool live = rb * 16 < nrows;
    const int64_t sstride = nrows * nbr;  // stride of Rp in gwann
    const double2 *A = gw + (size_t)s * nrows;
    double2 *buf = gsm + wave * 2 * (KI / 2) * 2 * 16;  // [2][pair j][cell R_p | -R_p][row]
    double *bsh = reinterpret_cast<double *>(gsm + 4 * 2 * (KI / 2) * 2 * 16);  // [2][K row][LDB]
    double2 *wsh = reinterpret_cast<double2 *>(bsh + 2 * KI * LDB);             // [K2 / 2] pair weights
    int2 *psh = reinterpret_cast<int2 *>(wsh + K2 / 2);                          // [K2 / 2] pair cells
    for (int i = threadIdx.x; i < K2 / 2; i += 256) {
        wsh[i] = i < npair ? pwt[i] : make_double2(0.0, 0.0);
        psh[i] = i < npair ? pairs[i] : make_int2(0, -1);  // pairs past npair copy cell 0
    }
    d4 cr[NCB], ci[NCB];
#pragma unroll
    for (int c = 0; c < NCB; ++c) { cr[c] = d4{0, 0, 0, 0}; ci[c] = d4{0, 0, 0, 0}; }
    // global -> LDS copy of the gwann rows of K rows k0 .. k0 + KI - 1 (cells
    // R_p and -R_p of pairs k0/2 .. k0/2 + 7, the wave's 16 rows) into buffer
    // b; a missing -R_p copies R_p (its weight is 0 below), rows past the end
    // copy the tile's first row, pairs past npair copy cell 0 (zero phases)
    auto copy = [&](int k0, int b) {
        double2 *dst = buf + b * (KI / 2) * 2 * 16;
#pragma unroll
        for (int t = 0; t < (KI / 2) * 2 * 16 / 64; ++t) {
            const int e = t * 64 + lane;
            const int r = e & 15, c = (e >> 4) & 1, jp = e >> 5;
            const int2 pr = psh[(k0 >> 1) + jp];
            const int cell = (c && pr.y >= 0) ? pr.y : pr.x;
            const int64_t rr = rb * 16 + r < nrows ? rb * 16 + r : rb * 16;
            lds_dma16(A + (size_t)cell * sstride + rr, dst + t * 64);
        }
    };
    // the wave's rows 4 j + wave of the phase block k0 .. k0 + KI - 1
    auto copy_b = [&](int k0, int b) {
        double *dst = bsh + b * KI * LDB;
#pragma unroll
        for (int j = 0; j < KI / 4; ++j) {
            const int r = 4 * j + wave;
            lds_dma16(phiR + (size_t)(k0 + r) * ldphi + 2 * lane, dst + r * LDB);
        }
    };
    // the 4 K steps of buffer b: G(2p) = g(R_p)/d + g(-R_p)/d',
    // G(2p+1) = i (g(R_p)/d - g(-R_p)/d')
    auto steps = [&](int k0, int b) {
        const double2 *src = buf + b * (KI / 2) * 2 * 16;
        const double *bs = bsh + b * KI * LDB;
#pragma unroll
        for (int u = 0; u < KI / 4; ++u) {
            const int kk = k0 + 4 * u + lk;
            const int p = kk >> 1, jp = (4 * u + lk) >> 1;
            const double2 w = wsh[p];
            const double dx = w.x, dy = w.y;
            const double2 gx = src[(jp * 2 + 0) * 16 + li], gy = src[(jp * 2 + 1) * 16 + li];
            const double2 x = make_double2(gx.x * dx, gx.y * dx), y = make_double2(gy.x * dy, gy.y * dy);
            const double2 a = (kk & 1) ? make_double2(y.y - x.y, x.x - y.x) : make_double2(x.x + y.x, x.y + y.y);
#pragma unroll
            for (int c = 0; c < NCB; ++c) {
                const double bv = bs[(4 * u + lk) * LDB + c * 16 + li];
                cr[c] = mfma_f64(a.x, bv, cr[c]);
                ci[c] = mfma_f64(a.y, bv, ci[c]);
            }
        }
    };
    __syncthreads();  // the weights
    if (live) copy(0, 0);
    copy_b(0, 0);
    for (int k0 = 0; k0 < K2; k0 += KI) {
        const int b = (k0 / KI) & 1;
        // block k0's copies are complete in every wave, and every wave is
        // done with buffer b ^ 1 (block k0 - KI), which the next copies fill;
        // no LDS stores are pending, so the barrier needs no fence, and the
        // clobber keeps the LDS reads on their side of it
        asm volatile("s_waitcnt vmcnt(0)\n\ts_barrier" ::: "memory");
        if (k0 + KI < K2) {
            if (live) copy(k0 + KI, b ^ 1);
            copy_b(k0 + KI, b ^ 1);
        }
        if (live) steps(k0, b);
    }
#pragma unroll
    for (int c = 0; c < NCB; ++c) {
        const int col = c * 16 + li;
        if (col >= nq) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int64_t X = rb * 16 + lk + 4 * r;
            if (X < nrows) {
                const int64_t i = X % nb2, re = X / nb2;
                out[(size_t)col * out_qstride + i + (size_t)nb2 * (s + (size_t)nbr * re)] = make_double2(cr[c][r], ci[c][r]);
            }
        }
    }
}

// Phonon-branch rotation of gReq for phonon state (q, s):
//   hq[st][Re][c] = sum_sp u(s,sp) gReq[q](np,mp,sp,Re) / elwsdeg(Re),  c = np + nb mp
// hq[st][K][CP] for the ibzq contraction in cell-pair form over Re: with
// h(Re) = sum_sp u(st, sp) g(Re, q)[:, sp] / elwsdeg(Re) (column c = np + nb mp),
// row 2p = h(R_p) + h(-R_p), row 2p + 1 = i (h(R_p) - h(-R_p)); pairs[p] =
// (index of R_p, index of -R_p or -1); rows past 2 npair are zero.
__global__ void k_rotate_hq(int nst, int nb, int nbr, int npair, int Kp, int CP, const double2 *__restrict__ gq,
                            int64_t gq_stride, const int *__restrict__ st_qb, const double2 *__restrict__ st_u,
                            const int2 *__restrict__ pairs, const double *__restrict__ invdeg,
                            double2 *__restrict__ h) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t per = (int64_t)Kp * CP;
    if (idx >= per * nst) return;
    const int st = (int)(idx / per);
    const int rem = (int)(idx % per);
    const int row = rem / CP, c = rem % CP, pr = row >> 1;
    const int nb2 = nb * nb;
    double2 out = make_double2(0.0, 0.0);
    if (pr < npair && c < nb2) {
        const double2 *u = st_u + (size_t)st * nbr;
        const double2 *gb = gq + (size_t)st_qb[st] * gq_stride + c;
        auto hcell = [&](int re) {
            double2 acc = make_double2(0.0, 0.0);
            const double2 *g = gb + (size_t)nb2 * nbr * re;
            for (int sp = 0; sp < nbr; ++sp) cfma(acc, u[sp], g[(size_t)sp * nb2]);
            const double d = invdeg[re];
            return make_double2(acc.x * d, acc.y * d);
        };
        const int2 q = pairs[pr];
        const double2 hi = hcell(q.x);
        const double2 hj = q.y >= 0 ? hcell(q.y) : make_double2(0.0, 0.0);
        out = (row & 1) ? make_double2(-(hi.y - hj.y), hi.x - hj.x) : make_double2(hi.x + hj.x, hi.y + hj.y);
    }
    h[idx] = out;
}

// Per state and k row: k' = k + q on the electron mesh (vector_allreps_add,
// src/vector.f90:68-84), its position in the window list (binsearch of
// src/interactions.f90:1817, here a mesh lookup table) and the number of
// vertices the row contributes (in-window m at k times in-window n at k').
__global__ void k_ibzq_rows(int nst, int nwv, const int *__restrict__ st_qe, const int *__restrict__ kint, int N0, int N1,
                            int N2, const int *__restrict__ lut, const int *__restrict__ nwin, int *__restrict__ rowmap,
                            int64_t *__restrict__ rowcnt) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int st = blockIdx.y;
    if (r >= nwv || st >= nst) return;
    const int a0 = (kint[3 * r] + st_qe[3 * st]) % N0;
    const int a1 = (kint[3 * r + 1] + st_qe[3 * st + 1]) % N1;
    const int a2 = (kint[3 * r + 2] + st_qe[3 * st + 2]) % N2;
    const int w = lut[(a2 * N1 + a1) * N0 + a0];
    rowmap[(size_t)st * nwv + r] = w;
    rowcnt[(size_t)st * nwv + r] = (w < 0) ? 0 : (int64_t)nwin[r] * nwin[w];
}

// Exclusive scan of rowcnt per state (one block per state); total -> counts.
__global__ __launch_bounds__(1024) void k_scan_rows(int nwv, int64_t *__restrict__ rowcnt, int64_t *__restrict__ total) {
    __shared__ int64_t part[1024];
    const int st = blockIdx.x, t = threadIdx.x;
    int64_t *v = rowcnt + (size_t)st * nwv;
    const int per = (nwv + 1023) / 1024;
    const int b = t * per, e = min(nwv, b + per);
    int64_t sum = 0;
    for (int i = b; i < e; ++i) sum += v[i];
    part[t] = sum;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        int64_t add = (t >= o) ? part[t - o] : 0;
        __syncthreads();
        part[t] += add;
        __syncthreads();
    }
    int64_t run = (t == 0) ? 0 : part[t - 1];
    for (int i = b; i < e; ++i) {
        const int64_t c = v[i];
        v[i] = run;
        run += c;
    }
    if (t == 1023) total[st] = part[1023];
}

struct PairQArgs {
    const double2 *h;          // [st][Kp][CP]
    int Kp, CP;
    int nwv;
    const int *kint;           // [nwv][3]
    int L, Lf0, Lf1, Lf2;
    const double2 *T;
    const int *Rk;             // [npair][3]: first cell R_p of each Re pair
    int npair;
    int Nel0, Nel1, Nel2;      // electron mesh
    unsigned Lmul;             // ceil(2^32 / L): index mod L by multiply-high
    const double2 *Vk;         // [nwv][n][np]
    const int *nwin, *wband;
    const int *rowmap;         // [st][nwv]
    const int64_t *rowoff;     // [st][nwv]
    const double *st_om;       // [st] phonon energy (eV)
    const double2 *st_lp;      // [st] polar long-range prefactor, or null
    const int64_t *st_slot;
    int nb;
    double *out;
    int64_t stride;
};

constexpr int kPairQKC = 32;  // K rows per LDS chunk of the fused ibzq kernel

// ibzq epilogue of one k row (k' row kp), in-window band ranks mr = lane0,
// lane0 + nlanes, ...: W(np) = sum_mp U_k(m,mp) Y(np + nb mp) once per m, then
// for every in-window n, g = sum_np conj(U_k'(n,np)) W(np) (+ polar term,
// src/wannier.f90:998-1006) and |g|^2 = 1/2 |g|^2 / omega * Ryd2eV^3 Ryd2amu
// (0 when omega == 0): the per-vertex sum sum_mp U_k(m,mp) sum_np
// conj(U_k'(n,np)) Y re-associated, nb^2 + nn nb products per m instead of
// nn (nb^2 + nb).  out: the row's vertices (mr * nn + nr).  nb <= NBX.
template <int NBX>
__device__ __forceinline__ void ibzq_row(const PairQArgs &a, const double2 *Yr, int row, int kp, int nm, int nn,
                                         int lane0, int nlanes, double w, int st, double *out) {
    const int nb = a.nb;
    for (int mr = lane0; mr < nm; mr += nlanes) {
        const int m = a.wband[row * nb + mr];
        const double2 *Uk = a.Vk + ((size_t)row * nb + m) * nb;
        double2 W[NBX];
#pragma unroll
        for (int np = 0; np < NBX; ++np) W[np] = make_double2(0.0, 0.0);
        for (int mp = 0; mp < nb; ++mp) {
            const double2 u = Uk[mp];
            const double2 *Ym = Yr + nb * mp;
#pragma unroll
            for (int np = 0; np < NBX; ++np)
                if (np < nb) cfma(W[np], u, Ym[np]);
        }
        for (int nr = 0; nr < nn; ++nr) {
            const int n = a.wband[kp * nb + nr];
            const double2 *Ukp = a.Vk + ((size_t)kp * nb + n) * nb;
            double2 g = make_double2(0.0, 0.0);
#pragma unroll
            for (int np = 0; np < NBX; ++np)
                if (np < nb) {
                    const double2 z = cmulc(Ukp[np], W[np]);
                    g.x += z.x;
                    g.y += z.y;
                }
            if (a.st_lp) {  // src/wannier.f90:998-1006
                double2 unm = make_double2(0.0, 0.0);
                for (int i = 0; i < nb; ++i) {
                    const double2 z = cmulc(Ukp[i], Uk[i]);
                    unm.x += z.x;
                    unm.y += z.y;
                }
                cfma(g, a.st_lp[st], unm);
            }
            out[(int64_t)mr * nn + nr] = w != 0.0 ? 0.5 * (g.x * g.x + g.y * g.y) / w * kG2Unit : 0.0;
        }
    }
}

// General ibzq path (nb^2 beyond the fused kernel's columns): Y of a chunk of
// states into HBM ([state of the chunk][k row][CP] complex), one 16*CB column
// tile per blockIdx.z, then k_ibzq_epi.
template <int CB, bool PACKED>
__global__ __launch_bounds__(256) void k_ibzq_y(PairQArgs a, int st0, double2 *__restrict__ Y) {
    constexpr int KC = 32;
    constexpr int RB = 2;
    constexpr int CPt = 16 * CB;
    extern __shared__ double2 smem[];
    double2 *buf0 = smem;
    double2 *buf1 = smem + KC * CPt;
    double2 *Tsh = smem + 2 * KC * CPt;
    int4 *Rsh = reinterpret_cast<int4 *>(Tsh + a.L);
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int li = lane & 15, lk = lane >> 4;
    const int sl = blockIdx.y, st = st0 + sl, tile = blockIdx.z;
    const int row_base = blockIdx.x * 128 + wave * 32;
    for (int i = tid; i < a.L; i += 256) Tsh[i] = a.T[i];
    for (int i = tid; i < a.Kp / 2; i += 256) {
        int4 R = make_int4(0, 0, 0, 0);
        if (i < a.npair) {
            R.x = a.Rk[3 * i];
            R.y = a.Rk[3 * i + 1];
            R.z = a.Rk[3 * i + 2];
            R.w = pos_mod(R.x, a.Nel0) | (pos_mod(R.y, a.Nel1) << 8) | (pos_mod(R.z, a.Nel2) << 16);
        }
        Rsh[i] = R;
    }
    int coef[RB][3];
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
        const int r = row_base + rb * 16 + li;
        coef[rb][0] = coef[rb][1] = coef[rb][2] = 0;
        if (r < a.nwv) {
            coef[rb][0] = (a.kint[3 * r] * a.Lf0) % a.L;
            coef[rb][1] = (a.kint[3 * r + 1] * a.Lf1) % a.L;
            coef[rb][2] = (a.kint[3 * r + 2] * a.Lf2) % a.L;
        }
    }
    d4 cr[RB][CB], ci[RB][CB];
    phase_gemm_core<CB, RB, KC, PACKED>(a.h + (size_t)st * a.Kp * a.CP + tile * CPt, a.CP, a.Kp, coef, a.L, a.Lmul,
                                        Tsh, Rsh, buf0, buf1, cr, ci);
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
        for (int c = 0; c < CB; ++c)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = row_base + rb * 16 + lk + 4 * r;
                if (row < a.nwv)
                    Y[((size_t)sl * a.nwv + row) * a.CP + tile * CPt + c * 16 + li] =
                        make_double2(cr[rb][c][r], ci[rb][c][r]);
            }
}

// Epilogue of the general ibzq path: 4 lanes per k row, Y read from HBM.
__global__ __launch_bounds__(256) void k_ibzq_epi(PairQArgs a, int st0, const double2 *__restrict__ Y) {
    const int row = blockIdx.x * 64 + (threadIdx.x >> 2), lk = threadIdx.x & 3;
    const int sl = blockIdx.y, st = st0 + sl;
    if (row >= a.nwv) return;
    const int kp = a.rowmap[(size_t)st * a.nwv + row];
    if (kp < 0) return;
    const int nm = a.nwin[row], nn = a.nwin[kp];
    const int64_t base = a.rowoff[(size_t)st * a.nwv + row];
    const double2 *Yr = Y + ((size_t)sl * a.nwv + row) * a.CP;
    double *outrow = a.out + a.st_slot[st] * a.stride;
    ibzq_row<32>(a, Yr, row, kp, nm, nn, lk, 4, a.st_om[st], st, outrow + base);
}

template <int CB, bool PACKED>
__global__ __launch_bounds__(256) void k_pair_q_fp64(PairQArgs a) {
    constexpr int KC = kPairQKC;
    constexpr int RB = 2;
    const int CP = 16 * CB;
    extern __shared__ double2 smem[];
    double2 *buf0 = smem;
    double2 *buf1 = smem + KC * CP;
    double2 *Tsh = smem + 2 * KC * CP;
    int4 *Rsh = reinterpret_cast<int4 *>(Tsh + a.L);
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int li = lane & 15, lk = lane >> 4;
    const int st = blockIdx.y;
    const int row_base = blockIdx.x * 128 + wave * 32;
    for (int i = tid; i < a.L; i += 256) Tsh[i] = a.T[i];
    for (int i = tid; i < a.Kp / 2; i += 256) {
        int4 R = make_int4(0, 0, 0, 0);
        if (i < a.npair) {
            R.x = a.Rk[3 * i];
            R.y = a.Rk[3 * i + 1];
            R.z = a.Rk[3 * i + 2];
            R.w = pos_mod(R.x, a.Nel0) | (pos_mod(R.y, a.Nel1) << 8) | (pos_mod(R.z, a.Nel2) << 16);
        }
        Rsh[i] = R;
    }
    int coef[RB][3];
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
        const int r = row_base + rb * 16 + li;
        coef[rb][0] = coef[rb][1] = coef[rb][2] = 0;
        if (r < a.nwv) {
            coef[rb][0] = (a.kint[3 * r] * a.Lf0) % a.L;
            coef[rb][1] = (a.kint[3 * r + 1] * a.Lf1) % a.L;
            coef[rb][2] = (a.kint[3 * r + 2] * a.Lf2) % a.L;
        }
    }
    d4 cr[RB][CB], ci[RB][CB];
    phase_gemm_core<CB, RB, KC, PACKED>(a.h + (size_t)st * a.Kp * CP, CP, a.Kp, coef, a.L, a.Lmul, Tsh, Rsh, buf0, buf1,
                                        cr, ci);

    const int YS = CP + 1;
    double2 *Y = smem + wave * 16 * YS;
    double *outrow = a.out + a.st_slot[st] * a.stride;
    const double w = a.st_om[st];
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
#pragma unroll
        for (int c = 0; c < CB; ++c)
#pragma unroll
            for (int r = 0; r < 4; ++r) Y[(lk + 4 * r) * YS + c * 16 + li] = make_double2(cr[rb][c][r], ci[rb][c][r]);
        __syncthreads();
        const int row = row_base + rb * 16 + li;
        if (row < a.nwv) {
            const int kp = a.rowmap[(size_t)st * a.nwv + row];
            if (kp >= 0) {
                const int nm = a.nwin[row], nn = a.nwin[kp];
                const int64_t base = a.rowoff[(size_t)st * a.nwv + row];
                ibzq_row<9>(a, Y + li * YS, row, kp, nm, nn, lk, 4, w, st, outrow + base);
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Hamiltonian / dynamical matrix assembly and the batched eigensolver
// ---------------------------------------------------------------------------

// H[k][ij] = sum_R phi[k][R] Hw[R][ij]; dH[k][p][ij] = sum_R i rcart_p(R) phi Hw
// (src/wannier.f90:462-478).
__global__ void k_build_H(int nk, int nb2, int nws, const double2 *__restrict__ phi, const double2 *__restrict__ HwT,
                          const double *__restrict__ rcart, double2 *__restrict__ H, double2 *__restrict__ dH) {
    const int k = blockIdx.y;
    const int ij = blockIdx.x * blockDim.x + threadIdx.x;
    if (ij >= nb2 || k >= nk) return;
    double2 acc = make_double2(0.0, 0.0);
    double2 d0 = acc, d1 = acc, d2 = acc;
    const double2 *ph = phi + (size_t)k * nws;
    for (int r = 0; r < nws; ++r) {
        const double2 t = cmul(ph[r], HwT[(size_t)r * nb2 + ij]);
        acc.x += t.x;
        acc.y += t.y;
        if (dH) {
            const double x = rcart[3 * r], y = rcart[3 * r + 1], z = rcart[3 * r + 2];
            d0.x -= x * t.y; d0.y += x * t.x;
            d1.x -= y * t.y; d1.y += y * t.x;
            d2.x -= z * t.y; d2.y += z * t.x;
        }
    }
    H[(size_t)k * nb2 + ij] = acc;
    if (dH) {
        dH[((size_t)k * 3 + 0) * nb2 + ij] = d0;
        dH[((size_t)k * 3 + 1) * nb2 + ij] = d1;
        dH[((size_t)k * 3 + 2) * nb2 + ij] = d2;
    }
}

// D(q) of src/wannier.f90:631-665: for each cell, the WS-wrapped images of the
// atom pair, weighted 1 / (phwsdeg * num_wrap), times Dphwann.
__global__ void k_build_D(int nq, int nbr, int nat, const double *__restrict__ qv, const int *__restrict__ toff,
                          const int *__restrict__ tiuc, const int *__restrict__ tT, const int *__restrict__ tdiv,
                          const double2 *__restrict__ DphT, const double *__restrict__ rcart, double2 *__restrict__ D,
                          double2 *__restrict__ dD) {
    const int q = blockIdx.y;
    const int nbr2 = nbr * nbr;
    const int ij = blockIdx.x * blockDim.x + threadIdx.x;
    if (ij >= nbr2 || q >= nq) return;
    const int i = ij % nbr, j = ij / nbr;
    const int ia = i / 3, ja = j / 3;
    const double qx = qv[q], qy = qv[q + nq], qz = qv[q + 2 * nq];
    double2 acc = make_double2(0.0, 0.0), d0 = acc, d1 = acc, d2 = acc;
    const int pair = ia + nat * ja;
    int t = toff[pair];
    const int tend = toff[pair + 1];
    while (t < tend) {
        const int iuc = tiuc[t];
        double2 cm = make_double2(0.0, 0.0);
        for (; t < tend && tiuc[t] == iuc; ++t) {
            const double dot = qx * (double)tT[3 * t] + qy * (double)tT[3 * t + 1] + qz * (double)tT[3 * t + 2];
            double s, c;
            sincos(kTwoPi * dot, &s, &c);
            const double dv = (double)tdiv[t];
            cm.x += c / dv;
            cm.y += s / dv;
        }
        const double2 v = cmul(cm, DphT[(size_t)iuc * nbr2 + ij]);
        acc.x += v.x;
        acc.y += v.y;
        if (dD) {
            const double x = rcart[3 * iuc], y = rcart[3 * iuc + 1], z = rcart[3 * iuc + 2];
            d0.x -= x * v.y; d0.y += x * v.x;
            d1.x -= y * v.y; d1.y += y * v.x;
            d2.x -= z * v.y; d2.y += z * v.x;
        }
    }
    D[(size_t)q * nbr2 + ij] = acc;
    if (dD) {
        dD[((size_t)q * 3 + 0) * nbr2 + ij] = d0;
        dD[((size_t)q * 3 + 1) * nbr2 + ij] = d1;
        dD[((size_t)q * 3 + 2) * nbr2 + ij] = d2;
    }
}

// Polar (Froehlich) long-range terms at a list of q, src/wannier.f90:784-901
// (dyn_nonanalytic) and 1013-1071 (long_range_prefac).  With g = G + q in
// 1/bohr, G over the (2N+1)^3 box of reciprocal vectors (N = coarse q mesh),
// kept where g.eps.g > 0 and g.eps.g / (4 alpha) < gmax, f(g) =
// exp(-g.eps.g / (4 alpha)) / g.eps.g and zg_a = g.Z_a:
//   D_L(3a+i, 3b+j) = D0 + fac sum_g f e^{i g.(tau_a - tau_b)} zg_a(i) zg_b(j)
//   dD_L(., ., c)   = dscale fac sum_g f e^{i g.(tau_a - tau_b)} [zg_b(j) Z_a(c,i)
//                     + zg_a(i) Z_b(c,j) + zg_a(i) zg_b(j) (i (tau_a - tau_b)_c
//                     - (1/(4 alpha) + 1/g.eps.g) ((eps + eps^T) g)_c)]
//   Lq(3a+i)        = i fac sum_g f e^{-i g.tau_a} zg_a(i)
// D and dD are accumulated into (the short-range D(q) is already there); D0 is
// the q-independent self term (the G-only sum of dyn_nonanalytic).
constexpr int kPolMaxAt = 10, kPolChunk = 64;

struct PolarArgs {
    int nq;
    const double *qv;     // (nq,3) column-major, crystal coordinates
    int nat, nbr, nG;
    const double *G;      // [nG][3] (1/bohr)
    const double *Z;      // Born charges (3,3,nat), Fortran order: Z[c + 3 i + 9 a]
    const double *tau;    // [nat][3] (bohr)
    double eps[9], rec[9];  // column-major; rec in 1/bohr, 2 pi included
    double alpha, fac, gmax, dscale;
    const double2 *D0;    // [nbr^2]
    double2 *D;           // [q][nbr^2] or null
    double2 *dD;          // [q][3][nbr^2] or null
    double2 *Lq;          // [q][nbr] or null
};

__global__ __launch_bounds__(256) void k_polar(PolarArgs a) {
    const int q = blockIdx.x;
    if (q >= a.nq) return;
    __shared__ double sf[kPolChunk], sdc[kPolChunk], sdg[kPolChunk][3];
    __shared__ double2 sph[kPolChunk][kPolMaxAt];
    __shared__ double sz[kPolChunk][kPolMaxAt][3];
    __shared__ double sZ[9 * kPolMaxAt], stau[3 * kPolMaxAt];
    const int t = threadIdx.x, nbr = a.nbr, nbr2 = nbr * nbr, nat = a.nat;
    for (int i = t; i < 9 * nat; i += 256) sZ[i] = a.Z[i];
    for (int i = t; i < 3 * nat; i += 256) stau[i] = a.tau[i];
    const double q0 = a.qv[q], q1 = a.qv[q + a.nq], q2 = a.qv[q + 2 * a.nq];
    double qc[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) qc[r] = a.rec[r] * q0 + a.rec[r + 3] * q1 + a.rec[r + 6] * q2;
    const bool wantd = a.dD != nullptr;
    double2 accD[4], accd[4][3], accL = make_double2(0.0, 0.0);
#pragma unroll
    for (int e4 = 0; e4 < 4; ++e4) {
        accD[e4] = make_double2(0.0, 0.0);
#pragma unroll
        for (int c = 0; c < 3; ++c) accd[e4][c] = make_double2(0.0, 0.0);
    }
    for (int g0 = 0; g0 < a.nG; g0 += kPolChunk) {
        __syncthreads();
        if (t < kPolChunk) {
            const int gi = g0 + t;
            double f = 0.0, g[3] = {0.0, 0.0, 0.0};
            if (gi < a.nG) {
#pragma unroll
                for (int r = 0; r < 3; ++r) g[r] = a.G[3 * gi + r] + qc[r];
                double geg = 0.0;
#pragma unroll
                for (int r = 0; r < 3; ++r) geg += g[r] * (a.eps[r] * g[0] + a.eps[r + 3] * g[1] + a.eps[r + 6] * g[2]);
                if (geg > 0.0 && geg / a.alpha / 4.0 < a.gmax) {
                    f = exp(-geg / a.alpha / 4.0) / geg;
                    sdc[t] = 0.25 / a.alpha + 1.0 / geg;
#pragma unroll
                    for (int r = 0; r < 3; ++r)
                        sdg[t][r] = (a.eps[r] + a.eps[3 * r]) * g[0] + (a.eps[r + 3] + a.eps[3 * r + 1]) * g[1] +
                                    (a.eps[r + 6] + a.eps[3 * r + 2]) * g[2];
                }
            }
            sf[t] = f;
            for (int at = 0; at < nat; ++at) {
                double s = 0.0, c = 0.0;
                if (f != 0.0) sincos(g[0] * stau[3 * at] + g[1] * stau[3 * at + 1] + g[2] * stau[3 * at + 2], &s, &c);
                sph[t][at] = make_double2(c, s);
#pragma unroll
                for (int i = 0; i < 3; ++i)
                    sz[t][at][i] = g[0] * sZ[9 * at + 3 * i] + g[1] * sZ[9 * at + 3 * i + 1] + g[2] * sZ[9 * at + 3 * i + 2];
            }
        }
        __syncthreads();
        const int cnt = min(kPolChunk, a.nG - g0);
        if (a.D) {
#pragma unroll
            for (int e4 = 0; e4 < 4; ++e4) {
                const int e = t + 256 * e4;
                if (e < nbr2) {
                    const int ii = e % nbr, jj = e / nbr;
                    const int ia = ii / 3, ip = ii % 3, ja = jj / 3, jp = jj % 3;
                    double rr[3];
#pragma unroll
                    for (int c = 0; c < 3; ++c) rr[c] = stau[3 * ia + c] - stau[3 * ja + c];
                    for (int k = 0; k < cnt; ++k) {
                        const double f = sf[k];
                        if (f == 0.0) continue;
                        const double2 ph = cmulc(sph[k][ja], sph[k][ia]);  // e^{i g.(tau_a - tau_b)}
                        const double2 fq = make_double2(f * ph.x, f * ph.y);
                        const double zi = sz[k][ia][ip], zj = sz[k][ja][jp];
                        accD[e4].x = fma(fq.x, zi * zj, accD[e4].x);
                        accD[e4].y = fma(fq.y, zi * zj, accD[e4].y);
                        if (wantd) {
#pragma unroll
                            for (int c = 0; c < 3; ++c) {
                                // zj Z_a(c,i) + zi Z_b(c,j) + zi zj (i rr_c - dcoef dgeg_c)
                                const double re = zj * sZ[9 * ia + 3 * ip + c] + zi * sZ[9 * ja + 3 * jp + c] -
                                                  zi * zj * sdc[k] * sdg[k][c];
                                const double im = zi * zj * rr[c];
                                cfma(accd[e4][c], fq, make_double2(re, im));
                            }
                        }
                    }
                }
            }
        }
        if (a.Lq && t < nbr) {
            const int at = t / 3, ip = t % 3;
            for (int k = 0; k < cnt; ++k) {
                const double f = sf[k];
                if (f == 0.0) continue;
                const double2 ph = sph[k][at];
                const double w = f * sz[k][at][ip];
                accL.x = fma(w, ph.x, accL.x);   // f zg e^{-i g.tau}
                accL.y = fma(-w, ph.y, accL.y);
            }
        }
    }
    if (a.D) {
#pragma unroll
        for (int e4 = 0; e4 < 4; ++e4) {
            const int e = t + 256 * e4;
            if (e < nbr2) {
                double2 &d = a.D[(size_t)q * nbr2 + e];
                d.x += a.fac * accD[e4].x + a.D0[e].x;
                d.y += a.fac * accD[e4].y + a.D0[e].y;
                if (wantd)
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        double2 &dd = a.dD[((size_t)q * 3 + c) * nbr2 + e];
                        dd.x += a.dscale * a.fac * accd[e4][c].x;
                        dd.y += a.dscale * a.fac * accd[e4][c].y;
                    }
            }
        }
    }
    if (a.Lq && t < nbr) a.Lq[(size_t)q * nbr + t] = make_double2(-a.fac * accL.y, a.fac * accL.x);  // i fac acc
}

// lp[i] = sum_p Lq[r][p] u[i][p] with r = ridx[i] (or i / div): the long-range
// vertex prefactor of one phonon mode, u = e / sqrt(M) (src/wannier.f90:1002-1003).
__global__ void k_polar_lp(int n, int nbr, int div, const int *__restrict__ ridx, const double2 *__restrict__ Lq,
                           const double2 *__restrict__ u, double2 *__restrict__ lp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int r = ridx ? ridx[i] : i / div;
    double2 acc = make_double2(0.0, 0.0);
    for (int p = 0; p < nbr; ++p) cfma(acc, Lq[(size_t)r * nbr + p], u[(size_t)i * nbr + p]);
    lp[i] = acc;
}

// Batched Hermitian eigensolver, one wavefront per matrix: parallel cyclic
// Jacobi (round-robin pairing, disjoint rotations applied together) in FP64.
// Input A (column-major n x n); the matrix used is the Hermitian one whose
// upper triangle is the force-Hermitised upper triangle of A (zheev('V','U')
// of src/wannier.f90:481-497, 708-738), optionally scaled elementwise by
// scale (real symmetric, e.g. the mass normalisation of 716-729).
// Output: W ascending, V column j = eigenvector j.
template <int NMAX>
__global__ __launch_bounds__(64) void k_jacobi(int nmat, int n, const double2 *__restrict__ Ain,
                                               const double *__restrict__ scale, double *__restrict__ W,
                                               double2 *__restrict__ Vout) {
    __shared__ double2 A[NMAX * NMAX];
    __shared__ double2 V[NMAX * NMAX];
    __shared__ double rc[NMAX / 2 + 1], rs[NMAX / 2 + 1];
    __shared__ double2 rph[NMAX / 2 + 1];
    __shared__ int rp[NMAX / 2 + 1], rq[NMAX / 2 + 1];
    __shared__ double sred[2];
    const int mat = blockIdx.x;
    const int lane = threadIdx.x;
    if (mat >= nmat) return;
    const double2 *a = Ain + (size_t)mat * n * n;
    for (int idx = lane; idx < NMAX * NMAX; idx += 64) {
        const int i = idx % NMAX, j = idx / NMAX;
        double2 v = make_double2(0.0, 0.0);
        if (i < n && j < n) {
            if (i < j) {
                const double2 u = a[i + n * j], l = a[j + n * i];
                v = make_double2((u.x + l.x) * 0.5, (u.y - l.y) * 0.5);
            } else if (i == j) {
                v = make_double2(a[i + n * i].x, 0.0);
            } else {
                const double2 u = a[j + n * i], l = a[i + n * j];
                v = make_double2((u.x + l.x) * 0.5, -(u.y - l.y) * 0.5);
            }
            if (scale) {
                const double sc = scale[i + n * j];
                v.x *= sc;
                v.y *= sc;
            }
        }
        A[idx] = v;
        V[idx] = make_double2(i == j ? 1.0 : 0.0, 0.0);
    }
    __syncthreads();
    const int m = n + (n & 1);
    const int npair = m / 2;
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int idx = lane; idx < n * n; idx += 64) {
            const int i = idx % n, j = idx / n;
            const double2 v = A[i + NMAX * j];
            const double mag = v.x * v.x + v.y * v.y;
            tot += mag;
            if (i != j) off += mag;
        }
        for (int o = 32; o > 0; o >>= 1) {
            off += __shfl_xor(off, o);
            tot += __shfl_xor(tot, o);
        }
        if (lane == 0) { sred[0] = off; sred[1] = tot; }
        __syncthreads();
        off = sred[0];
        tot = sred[1];
        __syncthreads();
        if (off <= 1e-32 * tot || off == 0.0) break;
        for (int step = 0; step < m - 1; ++step) {
            if (lane < npair) {
                int p, q;
                if (lane == 0) { p = step; q = m - 1; }
                else { p = (step + lane) % (m - 1); q = (step - lane + (m - 1)) % (m - 1); }
                if (p > q) { int t = p; p = q; q = t; }
                double cs = 1.0, sn = 0.0;
                double2 ph = make_double2(1.0, 0.0);
                if (q < n) {
                    const double2 c = A[p + NMAX * q];
                    const double ac = sqrt(c.x * c.x + c.y * c.y);
                    if (ac > 0.0) {
                        const double app = A[p + NMAX * p].x, aqq = A[q + NMAX * q].x;
                        const double th = (aqq - app) / (2.0 * ac);
                        const double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
                        cs = 1.0 / sqrt(t * t + 1.0);
                        sn = t * cs;
                        ph = make_double2(c.x / ac, c.y / ac);
                    }
                } else {
                    q = -1;
                }
                rp[lane] = p; rq[lane] = q; rc[lane] = cs; rs[lane] = sn; rph[lane] = ph;
            }
            __syncthreads();
            // columns: A <- A U, V <- V U
            for (int w = lane; w < npair * n; w += 64) {
                const int k = w / n, row = w % n;
                const int p = rp[k], q = rq[k];
                if (q < 0) continue;
                const double cs = rc[k], sn = rs[k];
                const double2 cph = make_double2(rph[k].x, -rph[k].y);  // conj(ph)
                const double2 uqp = make_double2(-sn * cph.x, -sn * cph.y), uqq = make_double2(cs * cph.x, cs * cph.y);
                double2 xp = A[row + NMAX * p], xq = A[row + NMAX * q];
                double2 t1 = cmul(xq, uqp), t2 = cmul(xq, uqq);
                A[row + NMAX * p] = make_double2(xp.x * cs + t1.x, xp.y * cs + t1.y);
                A[row + NMAX * q] = make_double2(xp.x * sn + t2.x, xp.y * sn + t2.y);
                xp = V[row + NMAX * p];
                xq = V[row + NMAX * q];
                t1 = cmul(xq, uqp);
                t2 = cmul(xq, uqq);
                V[row + NMAX * p] = make_double2(xp.x * cs + t1.x, xp.y * cs + t1.y);
                V[row + NMAX * q] = make_double2(xp.x * sn + t2.x, xp.y * sn + t2.y);
            }
            __syncthreads();
            // rows: A <- U^H A
            for (int w = lane; w < npair * n; w += 64) {
                const int k = w / n, col = w % n;
                const int p = rp[k], q = rq[k];
                if (q < 0) continue;
                const double cs = rc[k], sn = rs[k];
                const double2 ph = rph[k];
                // conj(u_qp) = -sn ph, conj(u_qq) = cs ph
                const double2 xp = A[p + NMAX * col], xq = A[q + NMAX * col];
                const double2 t1 = cmul(make_double2(-sn * ph.x, -sn * ph.y), xq);
                const double2 t2 = cmul(make_double2(cs * ph.x, cs * ph.y), xq);
                A[p + NMAX * col] = make_double2(cs * xp.x + t1.x, cs * xp.y + t1.y);
                A[q + NMAX * col] = make_double2(sn * xp.x + t2.x, sn * xp.y + t2.y);
            }
            __syncthreads();
            if (lane < npair && rq[lane] >= 0) {
                const int p = rp[lane], q = rq[lane];
                A[p + NMAX * q] = make_double2(0.0, 0.0);
                A[q + NMAX * p] = make_double2(0.0, 0.0);
                A[p + NMAX * p].y = 0.0;
                A[q + NMAX * q].y = 0.0;
            }
            __syncthreads();
        }
    }
    // sort ascending, ties by index
    if (lane < n) {
        const double wi = A[lane + NMAX * lane].x;
        int rank = 0;
        for (int j = 0; j < n; ++j) {
            const double wj = A[j + NMAX * j].x;
            if (wj < wi || (wj == wi && j < lane)) rank++;
        }
        W[(size_t)mat * n + rank] = wi;
        for (int r = 0; r < n; ++r) Vout[(size_t)mat * n * n + r + (size_t)n * rank] = V[r + NMAX * lane];
    }
}

// v(k,b,p) = Re(V_b^H dH_p V_b) (Hellmann-Feynman, src/wannier.f90:504-510).
__global__ void k_quadform(int nk, int n, const double2 *__restrict__ V, const double2 *__restrict__ dH,
                           double *__restrict__ out /*[k][b][3]*/) {
    const int k = blockIdx.x;
    const int t = threadIdx.x;
    if (t >= 3 * n) return;
    const int b = t / 3, p = t % 3;
    const double2 *v = V + (size_t)k * n * n + (size_t)b * n;
    const double2 *M = dH + ((size_t)k * 3 + p) * n * n;
    double acc = 0.0;
    for (int r = 0; r < n; ++r) {
        double2 mv = make_double2(0.0, 0.0);
        for (int c = 0; c < n; ++c) cfma(mv, M[r + n * c], v[c]);
        const double2 z = cmulc(v[r], mv);
        acc += z.x;
    }
    out[((size_t)k * n + b) * 3 + p] = acc;
}

// ---------------------------------------------------------------------------
// Generic vertex list (Wannier%g2, src/wannier.f90:903-1011): one workgroup
// per vertex, literal order of contractions.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_g2_list(int n, int nb, int nbr, int nws, const double2 *__restrict__ phase,
                                                 const double2 *__restrict__ ek, const double2 *__restrict__ ekp,
                                                 const double2 *__restrict__ uq, const double *__restrict__ en,
                                                 const double2 *__restrict__ gm, const int64_t *__restrict__ gidx,
                                                 const double2 *__restrict__ lp, double *__restrict__ out) {
    const int v = blockIdx.x;
    if (v >= n) return;
    __shared__ double2 ov[32 * 32];
    __shared__ double red[2][256];
    const int t = threadIdx.x;
    for (int i = t; i < nb * nb; i += 256) {
        const int mp = i % nb, np = i / nb;
        ov[i] = cmulc(ekp[(size_t)v * nb + np], ek[(size_t)v * nb + mp]);  // overlap(mp,np)
    }
    __syncthreads();
    const double2 *g = gm + (size_t)gidx[v] * nb * nb * nbr * nws;
    double2 acc = make_double2(0.0, 0.0);
    for (int iws = t; iws < nws; iws += 256) {
        double2 B = make_double2(0.0, 0.0);
        for (int sp = 0; sp < nbr; ++sp) {
            double2 caux = make_double2(0.0, 0.0);
            for (int np = 0; np < nb; ++np)
                for (int mp = 0; mp < nb; ++mp)
                    cfma(caux, ov[mp + nb * np], g[np + nb * (mp + nb * (sp + (size_t)nbr * iws))]);
            cfma(B, uq[(size_t)v * nbr + sp], caux);
        }
        cfma(acc, phase[(size_t)v * nws + iws], B);
    }
    red[0][t] = acc.x;
    red[1][t] = acc.y;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (t < o) { red[0][t] += red[0][t + o]; red[1][t] += red[1][t + o]; }
        __syncthreads();
    }
    if (t == 0) {
        double2 g = make_double2(red[0][0], red[1][0]);
        if (lp) {  // polar: + glprefac * dot_product(U_k', U_k) (src/wannier.f90:998-1006)
            double2 unm = make_double2(0.0, 0.0);
            for (int i = 0; i < nb; ++i) {
                unm.x += ov[i + nb * i].x;
                unm.y += ov[i + nb * i].y;
            }
            cfma(g, lp[v], unm);
        }
        const double w = en[v];
        out[v] = (w == 0.0) ? 0.0 : 0.5 * (g.x * g.x + g.y * g.y) / w * kG2Unit;
    }
}

// Fixed-layout probe of the fp64 16x16x4 MFMA operand mapping.
__global__ void k_mfma_probe(const double *A /*16x4 row-major*/, const double *B /*4x16*/, double *D /*16x16*/) {
    const int lane = threadIdx.x;
    const double a = A[(lane & 15) * 4 + (lane >> 4)];
    const double b = B[(lane >> 4) * 16 + (lane & 15)];
    d4 c = {0, 0, 0, 0};
    c = mfma_f64(a, b, c);
    for (int r = 0; r < 4; ++r) D[((lane >> 4) + 4 * r) * 16 + (lane & 15)] = c[r];
}

// ---------------------------------------------------------------------------
// Host helpers
// ---------------------------------------------------------------------------
}  // namespace

void elph::require_lds(const elph_ctx *c, size_t bytes, const char *what) {
    if (c->lds_max > 0 && (int64_t)bytes > c->lds_max)
        throw Error{ELPH_EINVAL, std::string(what) + " needs " + std::to_string(bytes) +
                                     " bytes of LDS per workgroup, above the device's " +
                                     std::to_string(c->lds_max) +
                                     " (the phase table has lcm(mesh) entries: use a mesh with a smaller lcm)"};
}

void elph::timing_flush(elph_ctx *c) {
    if (c->tev_used == 0) return;
    HCHECK(hipEventSynchronize(c->tev_stop[c->tev_used - 1]));
    for (int i = 0; i < c->tev_used; ++i) {
        float ms = 0;
        HCHECK(hipEventElapsedTime(&ms, c->tev_start[i], c->tev_stop[i]));
        c->t_dominant_ms += ms;
        c->n_dominant += 1;
    }
    c->tev_used = 0;
}

void elph::timed_begin(elph_ctx *c, hipStream_t st) {
    if (c->tev_used == (int)c->tev_start.size()) timing_flush(c);
    HCHECK(hipEventRecord(c->tev_start[c->tev_used], st));
}

void elph::timed_end(elph_ctx *c, hipStream_t st) { HCHECK(hipEventRecord(c->tev_stop[c->tev_used++], st)); }

int elph::ret_error(elph_ctx *ctx, const Error &e) {
    if (ctx) ctx->err = e.msg;
    return e.code;
}

int64_t elph::lcm64(int64_t a, int64_t b) { return a / std::gcd(a, b) * b; }

std::vector<int2> elph::cell_pairs(const int64_t *cells, int64_t ld, int64_t n) {
    std::map<std::array<int64_t, 3>, int64_t> at;
    for (int64_t i = n - 1; i >= 0; --i) at[{cells[i], cells[i + ld], cells[i + 2 * ld]}] = i;
    std::vector<char> used(n, 0);
    std::vector<int2> out;
    for (int64_t i = 0; i < n; ++i) {
        if (used[i]) continue;
        used[i] = 1;
        int j = -1;
        const auto it = at.find({-cells[i], -cells[i + ld], -cells[i + 2 * ld]});
        if (it != at.end() && !used[it->second]) {
            j = (int)it->second;
            used[j] = 1;
        }
        out.push_back(make_int2((int)i, j));
    }
    return out;
}

namespace {

void upload_cells(elph_ctx *c, const std::vector<int64_t> &cells, int64_t n, DBuf<int> &dst, DBuf<double> *cart) {
    std::vector<int> v(3 * n);
    std::vector<double> cc(3 * n);
    for (int64_t i = 0; i < n; ++i) {
        for (int d = 0; d < 3; ++d) v[3 * i + d] = (int)cells[i + n * d];
        double R[3] = {(double)cells[i], (double)cells[i + n], (double)cells[i + 2 * n]};
        for (int r = 0; r < 3; ++r)
            cc[3 * i + r] = c->lattvecs[r] * R[0] + c->lattvecs[r + 3] * R[1] + c->lattvecs[r + 6] * R[2];
    }
    dst.upload(v, c->stream);
    if (cart) cart->upload(cc, c->stream);
}

void build_phonon_terms(elph_ctx *c) {
    const int64_t nat = c->natoms, nws = c->nwsq;
    std::vector<int> off(nat * nat + 1, 0), tiuc, tT, tdiv;
    std::vector<int> wr;
    for (int64_t ja = 0; ja < nat; ++ja)
        for (int64_t ia = 0; ia < nat; ++ia) {
            off[ia + nat * ja] = (int)tiuc.size();
            for (int64_t iuc = 0; iuc < nws; ++iuc) {
                double dist[3];
                for (int d = 0; d < 3; ++d) {
                    dist[d] = (double)c->rcells_q[iuc + nws * d] - c->basis[3 * ia + d] + c->basis[3 * ja + d];
                    dist[d] = dist[d] / (double)c->coarse_qmesh[d];
                }
                const int nw = ws_wrapping_vectors(c->lattvecs, dist, 1.0e-6, wr);
                for (int w = 0; w < nw; ++w) {
                    tiuc.push_back((int)iuc);
                    for (int d = 0; d < 3; ++d)
                        tT.push_back((int)(c->rcells_q[iuc + nws * d] + (int64_t)wr[3 * w + d] * c->coarse_qmesh[d]));
                    tdiv.push_back((int)(c->phwsdeg[iuc] * nw));
                }
            }
        }
    off[nat * nat] = (int)tiuc.size();
    // pairs are laid out (ia + nat*ja); reorder offsets to that index
    std::vector<int> off2(nat * nat + 1);
    for (int64_t p = 0; p <= nat * nat; ++p) off2[p] = off[p];
    c->term_off.upload(off2, c->stream);
    c->term_iuc.upload(tiuc, c->stream);
    c->term_T.upload(tT, c->stream);
    c->term_div.upload(tdiv, c->stream);
}

// Constants of the polar terms (src/wannier.f90:795-800, 1027-1033) and the
// G-only self term of dyn_nonanalytic (src/wannier.f90:814-846), times fac.
void build_polar_terms(elph_ctx *c) {
    const int64_t nat = c->natoms, nbr = c->nbr;
    const double *L = c->lattvecs, *rec = c->reclattvecs, *eps = c->epsilon;
    const double a1 = std::sqrt(L[0] * L[0] + L[1] * L[1] + L[2] * L[2]);
    const double tpiba = kTwoPi / a1 * kBohr2nm;
    c->pol_alpha = tpiba * tpiba;
    c->pol_fac = 4.0 * kTwoPi / (c->volume / (kBohr2nm * kBohr2nm * kBohr2nm));
    const double gmax = 14.0, alph = c->pol_alpha;
    std::vector<double> G, tau(3 * nat);
    for (int64_t a = 0; a < nat; ++a)
        for (int r = 0; r < 3; ++r)
            tau[3 * a + r] = (L[r] * c->basis[3 * a] + L[r + 3] * c->basis[3 * a + 1] + L[r + 6] * c->basis[3 * a + 2]) /
                             kBohr2nm;
    const double *Z = c->born.data();
    std::vector<double2> D0(nbr * nbr, make_double2(0.0, 0.0));
    const int64_t *N = c->coarse_qmesh;
    for (int64_t m1 = -N[0]; m1 <= N[0]; ++m1)
        for (int64_t m2 = -N[1]; m2 <= N[1]; ++m2)
            for (int64_t m3 = -N[2]; m3 <= N[2]; ++m3) {
                double g[3];
                for (int r = 0; r < 3; ++r) g[r] = (m1 * rec[r] + m2 * rec[3 + r] + m3 * rec[6 + r]) * kBohr2nm;
                G.insert(G.end(), g, g + 3);
                double geg = 0.0;
                for (int r = 0; r < 3; ++r) geg += g[r] * (eps[r] * g[0] + eps[r + 3] * g[1] + eps[r + 6] * g[2]);
                if (!(geg > 0.0 && geg / alph / 4.0 < gmax)) continue;
                const double f = std::exp(-geg / alph / 4.0) / geg;
                for (int64_t ia = 0; ia < nat; ++ia) {
                    double zi[3];
                    for (int i = 0; i < 3; ++i)
                        zi[i] = g[0] * Z[9 * ia + 3 * i] + g[1] * Z[9 * ia + 3 * i + 1] + g[2] * Z[9 * ia + 3 * i + 2];
                    double2 fn[3] = {make_double2(0, 0), make_double2(0, 0), make_double2(0, 0)};
                    for (int64_t ja = 0; ja < nat; ++ja) {
                        double arg = 0.0;
                        for (int r = 0; r < 3; ++r) arg += g[r] * (tau[3 * ia + r] - tau[3 * ja + r]);
                        const double cs = std::cos(arg), sn = std::sin(arg);
                        for (int j = 0; j < 3; ++j) {
                            const double zj = g[0] * Z[9 * ja + 3 * j] + g[1] * Z[9 * ja + 3 * j + 1] +
                                              g[2] * Z[9 * ja + 3 * j + 2];
                            fn[j].x += zj * cs;
                            fn[j].y += zj * sn;
                        }
                    }
                    for (int i = 0; i < 3; ++i)
                        for (int j = 0; j < 3; ++j) {
                            double2 &d = D0[(3 * ia + i) + nbr * (3 * ia + j)];
                            d.x -= f * zi[i] * fn[j].x;
                            d.y -= f * zi[i] * fn[j].y;
                        }
                }
            }
    for (auto &d : D0) { d.x *= c->pol_fac; d.y *= c->pol_fac; }
    c->pol_nG = (int64_t)G.size() / 3;
    c->pol_G.upload(G, c->stream);
    c->pol_Z.upload(c->born, c->stream);
    c->pol_tau.upload(tau, c->stream);
    c->pol_D0.upload(D0, c->stream);
}

// Polar terms at nq device q vectors (crystal coordinates, (nq,3) column-major):
// D_L added into D (and dD_L into dD), and/or Lq written.
}  // namespace

void elph::polar_launch(elph_ctx *c, hipStream_t st, int64_t nq, const double *d_qvecs, double2 *D, double2 *dD,
                        double2 *Lq) {
    if (c->natoms > kPolMaxAt) throw Error{ELPH_EINVAL, "polar terms support up to 10 atoms"};
    PolarArgs a{};
    a.nq = (int)nq;
    a.qv = d_qvecs;
    a.nat = (int)c->natoms;
    a.nbr = (int)c->nbr;
    a.nG = (int)c->pol_nG;
    a.G = c->pol_G.p;
    a.Z = c->pol_Z.p;
    a.tau = c->pol_tau.p;
    for (int i = 0; i < 9; ++i) {
        a.eps[i] = c->epsilon[i];
        a.rec[i] = c->reclattvecs[i] * kBohr2nm;
    }
    a.alpha = c->pol_alpha;
    a.fac = c->pol_fac;
    a.gmax = 14.0;
    a.dscale = kBohr2nm;
    a.D0 = c->pol_D0.p;
    a.D = D;
    a.dD = dD;
    a.Lq = Lq;
    if (nq > 0) hipLaunchKernelGGL(k_polar, dim3((unsigned)nq), dim3(256), 0, st, a);
    HCHECK(hipGetLastError());
}

namespace {

// lp[q][s] = sum_p Lq(q)[p] u[q][s][p] for nq host q vectors and device u [q][s][p].
void polar_prefactors(elph_ctx *c, int64_t nq, const std::vector<double> &qv, const DBuf<double2> &u,
                      DBuf<double2> &lp) {
    const int64_t nbr = c->nbr;
    DBuf<double> dq;
    DBuf<double2> Lq;
    dq.upload(qv, c->stream);
    Lq.alloc(nq * nbr);
    polar_launch(c, c->stream, nq, dq.p, nullptr, nullptr, Lq.p);
    lp.alloc(nq * nbr);
    const int64_t n = nq * nbr;
    hipLaunchKernelGGL(k_polar_lp, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, (int)n, (int)nbr,
                       (int)nbr, (const int *)nullptr, Lq.p, u.p, lp.p);
    HCHECK(hipGetLastError());
    HCHECK(hipStreamSynchronize(c->stream));
}

// Synthetic Wannier vertex of the benchmark workloads, generated in HBM:
// g(np,mp,Re,sp,Rp) = A e^{-|Re|/2} e^{-|Rp|/2} e^{i 2 pi u}, |R| in Angstrom,
// A = 0.1 eV in Ry, u uniform in [0,1) from splitmix64 of (seed, element index).
__global__ void k_synth_vertex(int64_t n, int nb2, int nwsk, int nbr, const double *__restrict__ rk_nm,
                               const int *__restrict__ cells_g, double L0, double L1, double L2, double L3, double L4,
                               double L5, double L6, double L7, double L8, double amp, uint64_t seed,
                               double2 *__restrict__ g) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int64_t r = t / nb2;
    const int re = (int)(r % nwsk);
    r /= nwsk;
    const int rp = (int)(r / nbr);
    const double ke = sqrt(rk_nm[3 * re] * rk_nm[3 * re] + rk_nm[3 * re + 1] * rk_nm[3 * re + 1] +
                           rk_nm[3 * re + 2] * rk_nm[3 * re + 2]) * 10.0;
    const double x = (double)cells_g[3 * rp], y = (double)cells_g[3 * rp + 1], z = (double)cells_g[3 * rp + 2];
    const double cx = L0 * x + L3 * y + L6 * z, cy = L1 * x + L4 * y + L7 * z, cz = L2 * x + L5 * y + L8 * z;
    const double kg = sqrt(cx * cx + cy * cy + cz * cz) * 10.0;
    uint64_t h = seed + (uint64_t)t * 0x9E3779B97F4A7C15ull;
    h = (h ^ (h >> 30)) * 0xBF58476D1CE4E5B9ull;
    h = (h ^ (h >> 27)) * 0x94D049BB133111EBull;
    h ^= h >> 31;
    const double u = (double)(h >> 11) * (1.0 / 9007199254740992.0);
    double sn, cs;
    sincos(kTwoPi * u, &sn, &cs);
    const double a = amp * exp(-0.5 * ke - 0.5 * kg);
    g[t] = make_double2(a * cs, a * sn);
}

void ctx_init_device(elph_ctx *c, const double *Hwann, const double *Dphwann, const double *gwann) {
    HCHECK(hipSetDevice(c->device));
    int lds = 0;
    HCHECK(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device));
    c->lds_max = lds;
    int ncu = 0;
    HCHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device));
    c->ncu = ncu;
    c->noalloc_scratch = noalloc_scratch_address(c->device);
    HCHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HCHECK(hipEventCreate(&c->ev_a));
    HCHECK(hipEventCreate(&c->ev_b));
    c->tev_start.assign(64, nullptr);
    c->tev_stop.assign(64, nullptr);
    for (int i = 0; i < 64; ++i) {
        HCHECK(hipEventCreate(&c->tev_start[i]));
        HCHECK(hipEventCreate(&c->tev_stop[i]));
    }
    HCHECK(hipEventCreate(&c->ev_c));
    HCHECK(hipEventCreate(&c->ev_d));
    reciprocal_lattice(c->lattvecs, c->reclattvecs, &c->volume);
    const int64_t nb = c->nb, nbr = c->nbr;
    // Hwann (nwsk,nb,nb) -> [R][i + nb j]
    std::vector<double2> HwT(c->nwsk * nb * nb);
    for (int64_t j = 0; j < nb; ++j)
        for (int64_t i = 0; i < nb; ++i)
            for (int64_t r = 0; r < c->nwsk; ++r) {
                const int64_t src = r + c->nwsk * (i + nb * j);
                HwT[r * nb * nb + i + nb * j] = make_double2(Hwann[2 * src], Hwann[2 * src + 1]);
            }
    c->HwT.upload(HwT, c->stream);
    std::vector<double2> DT(c->nwsq * nbr * nbr);
    for (int64_t j = 0; j < nbr; ++j)
        for (int64_t i = 0; i < nbr; ++i)
            for (int64_t r = 0; r < c->nwsq; ++r) {
                const int64_t src = r + c->nwsq * (i + nbr * j);
                DT[r * nbr * nbr + i + nbr * j] = make_double2(Dphwann[2 * src], Dphwann[2 * src + 1]);
            }
    c->DphT.upload(DT, c->stream);
    const size_t ng = (size_t)nb * nb * c->nwsk * nbr * c->nwsg;
    c->gwann.alloc(ng);
    if (gwann)
        HCHECK(hipMemcpyAsync(c->gwann.p, gwann, ng * sizeof(double2), hipMemcpyHostToDevice, c->stream));
    else
        hipLaunchKernelGGL(k_zero2, dim3((unsigned)((ng + 255) / 256)), dim3(256), 0, c->stream, c->gwann.p,
                           (int64_t)ng);
    upload_cells(c, c->rcells_k, c->nwsk, c->rk, &c->rcart_k);
    upload_cells(c, c->rcells_q, c->nwsq, c->rq, &c->rcart_q);
    upload_cells(c, c->rcells_g, c->nwsg, c->rg, nullptr);
    auto deg32 = [](const std::vector<int64_t> &d) { return std::vector<int>(d.begin(), d.end()); };
    c->degk.upload(deg32(c->elwsdeg), c->stream);
    c->degq.upload(deg32(c->phwsdeg), c->stream);
    c->degg.upload(deg32(c->gwsdeg), c->stream);
    {
        // Re cells in pairs (R, -R) for the cell-pair form of the Re contraction
        const std::vector<int2> pr = cell_pairs(c->rcells_k.data(), c->nwsk, c->nwsk);
        c->npairk = (int64_t)pr.size();
        c->gk2p = ((2 * c->npairk + 63) / 64) * 64;
        std::vector<int> pc(3 * c->npairk);
        for (int64_t q = 0; q < c->npairk; ++q)
            for (int d = 0; d < 3; ++d) pc[3 * q + d] = (int)c->rcells_k[pr[q].x + c->nwsk * d];
        c->gpair.upload(pr, c->stream);
        c->gpcell.upload(pc, c->stream);
        std::vector<double> inv(c->nwsk);
        for (int64_t q = 0; q < c->nwsk; ++q) inv[q] = 1.0 / (double)c->elwsdeg[q];
        c->ginvdeg.upload(inv, c->stream);
    }
    {
        // Rp cells of the g set in pairs (R, -R) for the cell-pair form of gReq
        const std::vector<int2> pr = cell_pairs(c->rcells_g.data(), c->nwsg, c->nwsg);
        c->npairg = (int64_t)pr.size();
        std::vector<int> pc(3 * c->npairg);
        for (int64_t q = 0; q < c->npairg; ++q)
            for (int d = 0; d < 3; ++d) pc[3 * q + d] = (int)c->rcells_g[pr[q].x + c->nwsg * d];
        c->rgpair.upload(pr, c->stream);
        c->rgpcell.upload(pc, c->stream);
        // weights (1 / gwsdeg of R_p, of -R_p or 0 when it is missing) per pair
        std::vector<double2> wt(c->npairg);
        for (int64_t q = 0; q < c->npairg; ++q)
            wt[q] = make_double2(1.0 / (double)c->gwsdeg[pr[q].x],
                                 pr[q].y >= 0 ? 1.0 / (double)c->gwsdeg[pr[q].y] : 0.0);
        c->rgpwt.upload(wt, c->stream);
    }
    build_phonon_terms(c);
    if (c->polar) build_polar_terms(c);
    HCHECK(hipStreamSynchronize(c->stream));
}

void phase_list(elph_ctx *c, int64_t n, const double *d_vecs, const DBuf<int> &cells, const DBuf<int> &deg,
                int64_t ncell, DBuf<double2> &out) {
    out.alloc((size_t)n * ncell);
    dim3 grid((unsigned)((ncell + 127) / 128), (unsigned)n);
    hipLaunchKernelGGL(k_phase_list, grid, dim3(128), 0, c->stream, (int)n, (int)ncell, d_vecs, (int64_t)n, cells.p,
                       deg.p, out.p, (int64_t)ncell, (int64_t)1);
}

template <int NMAX>
void launch_jacobi_n(hipStream_t st, int64_t nmat, int n, const double2 *A, const double *scale, double *W,
                     double2 *V) {
    const int64_t chunk = 65535;
    for (int64_t m0 = 0; m0 < nmat; m0 += chunk) {
        const int64_t cnt = std::min(chunk, nmat - m0);
        hipLaunchKernelGGL(k_jacobi<NMAX>, dim3((unsigned)cnt), dim3(64), 0, st, (int)cnt, n,
                           A + (size_t)m0 * n * n, scale, W + (size_t)m0 * n, V + (size_t)m0 * n * n);
    }
}

}  // namespace

void elph::launch_jacobi(hipStream_t st, int64_t nmat, int n, const double2 *A, const double *scale, double *W,
                         double2 *V) {
    // NMAX sizes the LDS copies of A and V (one wave per matrix), so the
    // smallest fitting size keeps the most waves resident per CU
    if (n <= 8) launch_jacobi_n<8>(st, nmat, n, A, scale, W, V);
    else if (n <= 12) launch_jacobi_n<12>(st, nmat, n, A, scale, W, V);
    else if (n <= 16) launch_jacobi_n<16>(st, nmat, n, A, scale, W, V);
    else if (n <= 24) launch_jacobi_n<24>(st, nmat, n, A, scale, W, V);
    else if (n <= 32) launch_jacobi_n<32>(st, nmat, n, A, scale, W, V);
    else throw Error{ELPH_EINVAL, "eigensolver supports matrices up to 32x32"};
}

void elph::launch_build_H(elph_ctx *c, hipStream_t st, int64_t n, const double *d_kvecs, int64_t vstride, double2 *phi,
                          double2 *H) {
    const int64_t nb2 = c->nb * c->nb;
    dim3 gp((unsigned)((c->nwsk + 127) / 128), (unsigned)n);
    hipLaunchKernelGGL(k_phase_list, gp, dim3(128), 0, st, (int)n, (int)c->nwsk, d_kvecs, vstride, c->rk.p, c->degk.p,
                       phi, c->nwsk, (int64_t)1);
    dim3 grid((unsigned)((nb2 + 63) / 64), (unsigned)n);
    hipLaunchKernelGGL(k_build_H, grid, dim3(64), 0, st, (int)n, (int)nb2, (int)c->nwsk, phi, c->HwT.p, c->rcart_k.p,
                       H, (double2 *)nullptr);
    HCHECK(hipGetLastError());
}

void elph::launch_dynmat(elph_ctx *c, hipStream_t st, int64_t nq, const double *d_qvecs, double2 *D) {
    const int64_t nbr2 = c->nbr * c->nbr;
    dim3 grid((unsigned)((nbr2 + 63) / 64), (unsigned)nq);
    hipLaunchKernelGGL(k_build_D, grid, dim3(64), 0, st, (int)nq, (int)c->nbr, (int)c->natoms, d_qvecs,
                       c->term_off.p, c->term_iuc.p, c->term_T.p, c->term_div.p, c->DphT.p, c->rcart_q.p, D,
                       (double2 *)nullptr);
    if (c->polar) polar_launch(c, st, nq, d_qvecs, D, nullptr, nullptr);
    HCHECK(hipGetLastError());
}

std::vector<double> elph::mass_scale(const elph_ctx *c) {
    const int64_t nbr = c->nbr;
    std::vector<double> scale(nbr * nbr);
    for (int64_t j = 0; j < nbr; ++j)
        for (int64_t i = 0; i < nbr; ++i)
            scale[i + nbr * j] = 1.0 / std::sqrt(c->masses[i / 3] * c->masses[j / 3]) * kRyd2amu;
    return scale;
}

void elph::launch_polar_lp(hipStream_t st, int n, int nbr, int div, const int *ridx, const double2 *Lq,
                           const double2 *u, double2 *lp) {
    if (n > 0)
        hipLaunchKernelGGL(k_polar_lp, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, nbr, div, ridx, Lq, u,
                           lp);
    HCHECK(hipGetLastError());
}

void elph::launch_rotate_h(hipStream_t st, int nst, int nb, int nbr, int npair, int Kp, int CP, const double2 *gk,
                           int64_t gk_kstride, const int *st_kb, const int *st_m, const double2 *Uk, const int *degq,
                           const int2 *pairs, double2 *h) {
    const int64_t tot = (int64_t)nst * (Kp / 2) * CP;
    hipLaunchKernelGGL(k_rotate_h, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, nst, nb, nbr, npair, Kp, CP,
                       gk, gk_kstride, st_kb, st_m, Uk, degq, pairs, h);
    HCHECK(hipGetLastError());
}

void elph::launch_colscale(hipStream_t st, int nst, int Kp, int CP, const double2 *h, double *scale) {
    if (nst <= 0) return;
    hipLaunchKernelGGL(k_colscale, dim3((unsigned)nst, (unsigned)((CP + 63) / 64)), dim3(256), 0, st, nst, Kp, CP,
                       h, scale);
    HCHECK(hipGetLastError());
}

void elph::launch_slice_b(hipStream_t st, int nst, int nkc, int CB, int Kp, int CP, const double2 *h,
                          const double *scale, v4i *bsl) {
    const int64_t nt = (int64_t)nst * nkc * CB * 64;
    hipLaunchKernelGGL(k_slice_b, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, st, nst, nkc, CB, Kp, CP, h, scale,
                       bsl);
    HCHECK(hipGetLastError());
}

void elph::launch_slice_b_ri(hipStream_t st, int nst, int nkc, int CB, int Kp, int CP, const double2 *h,
                             const double *scale, v4i *bsl) {
    const int64_t nt = (int64_t)nst * nkc * CB * 64;
    hipLaunchKernelGGL(k_slice_b_ri, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, st, nst, nkc, CB, Kp, CP, h,
                       scale, bsl);
    HCHECK(hipGetLastError());
}

namespace {

// Device el_wann into device outputs W [k][b] (Ry), V [k][col b][row], vel [k][b][3] (raw).
void el_wann_device(elph_ctx *c, int64_t nk, const double *d_kvecs, DBuf<double> &W, DBuf<double2> &V,
                    DBuf<double> *vel) {
    const int64_t nb = c->nb, nb2 = nb * nb;
    DBuf<double2> phi, H, dH;
    phase_list(c, nk, d_kvecs, c->rk, c->degk, c->nwsk, phi);
    H.alloc(nk * nb2);
    if (vel) dH.alloc(nk * 3 * nb2);
    dim3 grid((unsigned)((nb2 + 63) / 64), (unsigned)nk);
    hipLaunchKernelGGL(k_build_H, grid, dim3(64), 0, c->stream, (int)nk, (int)nb2, (int)c->nwsk, phi.p, c->HwT.p,
                       c->rcart_k.p, H.p, vel ? dH.p : (double2 *)nullptr);
    W.alloc(nk * nb);
    V.alloc(nk * nb2);
    launch_jacobi(c->stream, nk, (int)nb, H.p, nullptr, W.p, V.p);
    if (vel) {
        vel->alloc(nk * nb * 3);
        hipLaunchKernelGGL(k_quadform, dim3((unsigned)nk), dim3((unsigned)(((3 * nb) + 63) / 64 * 64)), 0, c->stream,
                           (int)nk, (int)nb, V.p, dH.p, vel->p);
    }
    HCHECK(hipGetLastError());
}

// Device ph_wann (short-range D(q), polar D_L, mass normalisation, eigensolve).
void ph_wann_device(elph_ctx *c, int64_t nq, const double *d_qvecs, DBuf<double> &W, DBuf<double2> &V,
                    DBuf<double> *vel) {
    const int64_t nbr = c->nbr, nbr2 = nbr * nbr;
    DBuf<double2> D, dD;
    D.alloc(nq * nbr2);
    if (vel) dD.alloc(nq * 3 * nbr2);
    dim3 grid((unsigned)((nbr2 + 63) / 64), (unsigned)nq);
    hipLaunchKernelGGL(k_build_D, grid, dim3(64), 0, c->stream, (int)nq, (int)nbr, (int)c->natoms, d_qvecs,
                       c->term_off.p, c->term_iuc.p, c->term_T.p, c->term_div.p, c->DphT.p, c->rcart_q.p, D.p,
                       vel ? dD.p : (double2 *)nullptr);
    if (c->polar) polar_launch(c, c->stream, nq, d_qvecs, D.p, vel ? dD.p : nullptr, nullptr);
    std::vector<double> scale(nbr2);
    for (int64_t j = 0; j < nbr; ++j)
        for (int64_t i = 0; i < nbr; ++i)
            scale[i + nbr * j] = 1.0 / std::sqrt(c->masses[i / 3] * c->masses[j / 3]) * kRyd2amu;
    DBuf<double> dscale;
    dscale.upload(scale, c->stream);
    W.alloc(nq * nbr);
    V.alloc(nq * nbr2);
    launch_jacobi(c->stream, nq, (int)nbr, D.p, dscale.p, W.p, V.p);
    if (vel) {
        // mass-normalise dD on the host path below (small); the quadratic form
        // uses the mass-normalised derivative, as the reference scales ddynmat
        // by the same factor (src/wannier.f90:724-727).
        std::vector<double2> hdD(nq * 3 * nbr2);
        HCHECK(hipMemcpyAsync(hdD.data(), dD.p, hdD.size() * sizeof(double2), hipMemcpyDeviceToHost, c->stream));
        HCHECK(hipStreamSynchronize(c->stream));
        for (int64_t q = 0; q < nq; ++q)
            for (int p = 0; p < 3; ++p)
                for (int64_t ij = 0; ij < nbr2; ++ij) {
                    double2 &v = hdD[(q * 3 + p) * nbr2 + ij];
                    v.x *= scale[ij];
                    v.y *= scale[ij];
                }
        dD.upload(hdD, c->stream);
        vel->alloc(nq * nbr * 3);
        hipLaunchKernelGGL(k_quadform, dim3((unsigned)nq), dim3((unsigned)(((3 * nbr) + 63) / 64 * 64)), 0,
                           c->stream, (int)nq, (int)nbr, V.p, dD.p, vel->p);
    }
    HCHECK(hipGetLastError());
}

// Post-processing of ph_wann eigenvalues (src/wannier.f90:740-780), host side.
int ph_postprocess(int64_t nq, int64_t nbr, const double *qvecs_colmajor, std::vector<double> &w,
                   std::vector<double> *vel) {
    int status = ELPH_OK;
    for (int64_t q = 0; q < nq; ++q) {
        for (int64_t s = 0; s < nbr; ++s) {
            const double o2 = w[q * nbr + s];
            const double a = std::sqrt(std::fabs(o2));
            w[q * nbr + s] = std::signbit(o2) ? -a : a;
        }
        if (vel)
            for (int64_t s = 0; s < nbr; ++s)
                for (int p = 0; p < 3; ++p) (*vel)[(q * nbr + s) * 3 + p] /= (2.0 * w[q * nbr + s]);
        for (int64_t s = 0; s < nbr; ++s) w[q * nbr + s] *= kRyd2eV;
        if (vel)
            for (int64_t i = 0; i < nbr * 3; ++i) (*vel)[q * nbr * 3 + i] *= kRyd2radTHz;
        if (qvecs_colmajor[q] == 0 && qvecs_colmajor[q + nq] == 0 && qvecs_colmajor[q + 2 * nq] == 0) {
            for (int64_t s = 0; s < 3 && s < nbr; ++s) w[q * nbr + s] = 0.0;
            if (vel)
                for (int64_t i = 0; i < nbr * 3; ++i) (*vel)[q * nbr * 3 + i] = 0.0;
        }
        for (int64_t s = 0; s < nbr; ++s) {
            const double e = w[q * nbr + s];
            if (e < -0.005) status = ELPH_ENEGPHONON;
            else if (e < 0 && e > -0.005) w[q * nbr + s] = 0.0;
        }
    }
    return status;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int elph_ctx_create(elph_ctx **out, int device, const elph_dims *d, const int64_t *rcells_k, const int64_t *rcells_q,
                    const int64_t *rcells_g, const int64_t *elwsdeg, const int64_t *phwsdeg, const int64_t *gwsdeg,
                    const double *Hwann, const double *Dphwann, const double *gwann, const double *basis,
                    const double *masses, const double *born) {
    if (!out || !d) return ELPH_EINVAL;
    *out = nullptr;
    auto c = std::make_unique<elph_ctx>();
    try {
        if (d->numwannbands <= 0 || d->numbranches <= 0 || d->numbranches != 3 * d->natoms || d->nwsk <= 0 ||
            d->nwsq <= 0 || d->nwsg <= 0)
            throw Error{ELPH_EINVAL, "inconsistent dimensions (numbranches must be 3*natoms, all sizes > 0)"};
        if (d->numwannbands > 32 || d->numbranches > 32)
            throw Error{ELPH_EINVAL, "numwannbands and numbranches must be <= 32"};
        if (d->polar && !born) throw Error{ELPH_EINVAL, "polar crystal needs Born charges"};
        c->device = device;
        c->nb = d->numwannbands;
        c->nbr = d->numbranches;
        c->nwsk = d->nwsk;
        c->nwsq = d->nwsq;
        c->nwsg = d->nwsg;
        c->natoms = d->natoms;
        for (int i = 0; i < 3; ++i) c->coarse_qmesh[i] = d->coarse_qmesh[i];
        std::memcpy(c->lattvecs, d->lattvecs, sizeof(c->lattvecs));
        std::memcpy(c->epsilon, d->epsilon, sizeof(c->epsilon));
        c->polar = d->polar;
        c->basis.assign(basis, basis + 3 * c->natoms);
        c->masses.assign(masses, masses + c->natoms);
        if (d->polar) c->born.assign(born, born + 9 * c->natoms);
        c->rcells_k.assign(rcells_k, rcells_k + 3 * c->nwsk);
        c->rcells_q.assign(rcells_q, rcells_q + 3 * c->nwsq);
        c->rcells_g.assign(rcells_g, rcells_g + 3 * c->nwsg);
        c->elwsdeg.assign(elwsdeg, elwsdeg + c->nwsk);
        c->phwsdeg.assign(phwsdeg, phwsdeg + c->nwsq);
        c->gwsdeg.assign(gwsdeg, gwsdeg + c->nwsg);
        ctx_init_device(c.get(), Hwann, Dphwann, gwann);
    } catch (const Error &e) {
        std::fprintf(stderr, "elph_ctx_create: %s\n", e.msg.c_str());
        return e.code;
    }
    *out = c.release();
    return ELPH_OK;
}

int elph_ctx_create_epw(elph_ctx **out, int device, const char *dir, int64_t natoms, const double *lattvecs,
                        const double *basis, const double *masses, const int64_t *coarse_qmesh, int32_t polar,
                        const double *epsilon, const double *born) {
    EpwData e;
    std::string err;
    int st = read_epw(dir, e, err);
    if (st != ELPH_OK) {
        std::fprintf(stderr, "elph_ctx_create_epw: %s\n", err.c_str());
        return st;
    }
    elph_dims d{};
    d.numwannbands = e.nb;
    d.numbranches = e.nbr;
    d.nwsk = e.nwsk;
    d.nwsq = e.nwsq;
    d.nwsg = e.nwsg;
    d.natoms = natoms;
    for (int i = 0; i < 3; ++i) d.coarse_qmesh[i] = coarse_qmesh[i];
    std::memcpy(d.lattvecs, lattvecs, sizeof(d.lattvecs));
    d.polar = polar;
    if (epsilon) std::memcpy(d.epsilon, epsilon, sizeof(d.epsilon));
    return elph_ctx_create(out, device, &d, e.rcells_k.data(), e.rcells_q.data(), e.rcells_g.data(), e.elwsdeg.data(),
                           e.phwsdeg.data(), e.gwsdeg.data(), e.Hwann.data(), e.Dphwann.data(), e.gwann.data(), basis,
                           masses, born);
}

int elph_ctx_destroy(elph_ctx *ctx) {
    if (!ctx) return ELPH_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->ev_a) (void)hipEventDestroy(ctx->ev_a);
    if (ctx->ev_b) (void)hipEventDestroy(ctx->ev_b);
    for (hipEvent_t e : ctx->tev_start)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : ctx->tev_stop)
        if (e) (void)hipEventDestroy(e);
    if (ctx->ev_c) (void)hipEventDestroy(ctx->ev_c);
    if (ctx->ev_d) (void)hipEventDestroy(ctx->ev_d);
    hipStream_t s = ctx->stream;
    delete ctx;
    if (s) (void)hipStreamDestroy(s);
    return ELPH_OK;
}

const char *elph_last_error(const elph_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int elph_ctx_dims(const elph_ctx *ctx, int64_t out[6]) {
    if (!ctx) return ELPH_EINVAL;
    out[0] = ctx->nb; out[1] = ctx->nbr; out[2] = ctx->nwsk; out[3] = ctx->nwsq; out[4] = ctx->nwsg; out[5] = ctx->natoms;
    return ELPH_OK;
}

int elph_last_timing(const elph_ctx *ctx, double out[3]) {
    if (!ctx) return ELPH_EINVAL;
    out[0] = ctx->t_dominant_ms;
    out[1] = (double)ctx->n_dominant;
    out[2] = ctx->t_all_ms;
    return ELPH_OK;
}

int elph_el_wann(elph_ctx *ctx, int64_t nk, const double *kvecs, double *energies, double *velocities, double *evecs,
                 const double *scissor) {
    if (!ctx || nk < 0) return ELPH_EINVAL;
    if (velocities && !evecs) {
        ctx->err = "In el_wann, velocity is present but not eigenvecs.";
        return ELPH_EINVAL;
    }
    if (nk == 0) return ELPH_OK;
    try {
        HCHECK(hipSetDevice(ctx->device));
        const int64_t nb = ctx->nb;
        DBuf<double> dk, W, vel;
        DBuf<double2> V;
        dk.upload(kvecs, 3 * nk, ctx->stream);
        el_wann_device(ctx, nk, dk.p, W, V, velocities ? &vel : nullptr);
        std::vector<double> w(nk * nb), hv;
        std::vector<double2> v(nk * nb * nb);
        HCHECK(hipMemcpyAsync(w.data(), W.p, w.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        HCHECK(hipMemcpyAsync(v.data(), V.p, v.size() * sizeof(double2), hipMemcpyDeviceToHost, ctx->stream));
        if (velocities) {
            hv.resize(nk * nb * 3);
            HCHECK(hipMemcpyAsync(hv.data(), vel.p, hv.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        }
        HCHECK(hipStreamSynchronize(ctx->stream));
        for (int64_t k = 0; k < nk; ++k)
            for (int64_t b = 0; b < nb; ++b) {
                double e = w[k * nb + b] * kRyd2eV;
                if (scissor) e += scissor[b];
                energies[k + nk * b] = e;
                if (evecs)
                    for (int64_t j = 0; j < nb; ++j) {
                        const double2 z = v[k * nb * nb + b * nb + j];  // component j of eigenvector b
                        const int64_t o = k + nk * (b + nb * j);
                        evecs[2 * o] = z.x;
                        evecs[2 * o + 1] = z.y;
                    }
                if (velocities)
                    for (int p = 0; p < 3; ++p) velocities[k + nk * (b + nb * p)] = hv[(k * nb + b) * 3 + p] * kRyd2radTHz;
            }
    } catch (const Error &e) {
        return ret_error(ctx, e);
    }
    return ELPH_OK;
}

int elph_ph_wann(elph_ctx *ctx, int64_t nq, const double *qvecs, double *energies, double *evecs, double *velocities) {
    if (!ctx || nq < 0) return ELPH_EINVAL;
    if (velocities && !evecs) {
        ctx->err = "In ph_wann, velocity is present but not eigenvecs.";
        return ELPH_EINVAL;
    }
    if (nq == 0) return ELPH_OK;
    int status = ELPH_OK;
    try {
        HCHECK(hipSetDevice(ctx->device));
        const int64_t nbr = ctx->nbr;
        DBuf<double> dq, W, vel;
        DBuf<double2> V;
        dq.upload(qvecs, 3 * nq, ctx->stream);
        ph_wann_device(ctx, nq, dq.p, W, V, velocities ? &vel : nullptr);
        std::vector<double> w(nq * nbr), hv;
        std::vector<double2> v(nq * nbr * nbr);
        HCHECK(hipMemcpyAsync(w.data(), W.p, w.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        HCHECK(hipMemcpyAsync(v.data(), V.p, v.size() * sizeof(double2), hipMemcpyDeviceToHost, ctx->stream));
        if (velocities) {
            hv.resize(nq * nbr * 3);
            HCHECK(hipMemcpyAsync(hv.data(), vel.p, hv.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        }
        HCHECK(hipStreamSynchronize(ctx->stream));
        status = ph_postprocess(nq, nbr, qvecs, w, velocities ? &hv : nullptr);
        for (int64_t q = 0; q < nq; ++q)
            for (int64_t s = 0; s < nbr; ++s) {
                energies[q + nq * s] = w[q * nbr + s];
                if (evecs)
                    for (int64_t j = 0; j < nbr; ++j) {
                        const double2 z = v[q * nbr * nbr + s * nbr + j];
                        const int64_t o = q + nq * (s + nbr * j);
                        evecs[2 * o] = z.x;
                        evecs[2 * o + 1] = z.y;
                    }
                if (velocities)
                    for (int p = 0; p < 3; ++p) velocities[q + nq * (s + nbr * p)] = hv[(q * nbr + s) * 3 + p];
            }
        if (status == ELPH_ENEGPHONON) ctx->err = "Large negative phonon energy found! Stopping!";
    } catch (const Error &e) {
        return ret_error(ctx, e);
    }
    return status;
}

}  // extern "C"

// int8 digits of gwann (once per context).
void elph::ensure_gwann_digits(elph_ctx *c) {
    if (c->gdig.p) return;
    const int64_t nb2 = c->nb * c->nb, nslices = c->nbr * c->nwsg;
    c->gnrb = (nb2 + 15) / 16;
    c->gnkc = c->gk2p / 64;
    c->gscale.alloc(nslices * c->gnrb * 16);
    const int64_t ns = nslices * c->gnrb * 16;
    hipLaunchKernelGGL(k_gwann_scale, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, c->stream, (int)nb2,
                       (int)c->nwsk, (int)nslices, (int)c->gnrb, (int)c->npairk, c->gpair.p, c->ginvdeg.p, c->gwann.p,
                       c->gscale.p);
    const int64_t nt = nslices * c->gnrb * c->gnkc * 64;
    c->gdig.alloc(nt * 2 * kNS);
    hipLaunchKernelGGL(k_gwann_digits, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, c->stream, (int)nb2,
                       (int)c->nwsk, (int)nslices, (int)c->gnrb, (int)c->gnkc, (int)c->npairk, c->gpair.p, c->ginvdeg.p,
                       c->gwann.p, c->gscale.p, c->gdig.p);
    HCHECK(hipGetLastError());
    // callers enqueue on other streams (batch, server) right after this
    HCHECK(hipStreamSynchronize(c->stream));
}

// k per launch of the Re contraction: FP64 k_gkRp<4> (4 column blocks, one
// wave per 16 rows) and the int8 k_gkRp_i8<4, 2> (two column groups of 4
// blocks per pass over the digits of gwann; k_gkRp_i8<4, 1> for up to 64 k)
constexpr int64_t kGkKT = 64, kGkKTMixed = 128;
int64_t elph::gk_phi_count(const elph_ctx *c) { return c->gk2p * kGkKTMixed / 2; }
int64_t elph::gk_bphi_count(const elph_ctx *c) { return (c->gk2p / 64) * (kGkKTMixed / 16) * kNS * 64; }

// Batched gkRp into device memory: out[k] (nb*nb, nbr*nwsg) for kvecs (nk,3)
// column-major on the device; FP64 or int8-digit (mixed) contraction, both in
// cell-pair form.  phiT (gk_phi_count) and, for mixed, bphi (gk_bphi_count)
// are scratch buffers of the caller.
void elph::launch_gkRp(elph_ctx *c, hipStream_t st, int64_t nk, const double *d_kvecs, double2 *out,
                        int64_t out_kstride, bool mixed, double2 *phiT, v4i *bphi) {
    const int64_t nb2 = c->nb * c->nb, nslices = c->nbr * c->nwsg;
    // the int8 form takes 128 k per pass when there are more than 64 (fewer:
    // the 64-k form, whose pass over the gwann digits costs less)
    const int64_t KT = mixed && nk > kGkKT ? kGkKTMixed : kGkKT;
    double *phi = reinterpret_cast<double *>(phiT);
    if (!mixed && c->npairk * (int64_t)sizeof(double4) > 160 * 1024)
        throw Error{ELPH_EINVAL, "FP64 Re contraction holds up to 5120 cell pairs (10240 electron WS cells)"};
    for (int64_t k0 = 0; k0 < nk; k0 += KT) {
        const int64_t n = std::min(KT, nk - k0);
        // real phases in [K2][k] layout (KT columns; unused rows and columns zero)
        const int64_t nz = c->gk2p * KT / 2;
        hipLaunchKernelGGL(k_zero2, dim3((unsigned)((nz + 255) / 256)), dim3(256), 0, st, phiT, nz);
        dim3 grid((unsigned)((c->npairk + 127) / 128), (unsigned)n);
        hipLaunchKernelGGL(k_phase_pairs, grid, dim3(128), 0, st, (int)n, (int)c->npairk, d_kvecs + k0, nk,
                           c->gpcell.p, phi, (int64_t)1, KT);
        if (mixed) {
            const int nt = (int)(c->gnkc * (KT / 16) * 64);
            hipLaunchKernelGGL(k_phase_digits, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, st, (int)c->gnkc,
                               (int)(KT / 16), phi, (int)KT, bphi);
            constexpr int NCBW = 4;
            auto go = [&](auto kern, int CS) {
                const int64_t nitems = nslices * ((c->gnrb + 4 / CS - 1) / (4 / CS));
                const int64_t grid = std::min<int64_t>(nitems, 2 * std::max<int64_t>(c->ncu, 1));
                hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(256), 2 * NCBW * CS * kNS * 64 * sizeof(v4i), st,
                                   c->gdig.p, c->gscale.p, (int)nb2, (int)c->gnrb, (int)c->gnkc, (int)nslices, bphi,
                                   (int)n, out + (size_t)k0 * out_kstride, out_kstride);
            };
            if (KT == kGkKTMixed) go(k_gkRp_i8<NCBW, (int)(kGkKTMixed / 64)>, (int)(kGkKTMixed / 64));
            else go(k_gkRp_i8<NCBW, 1>, 1);
        } else {
            hipLaunchKernelGGL(k_gkRp<4>, dim3((unsigned)nslices), dim3(256), c->npairk * sizeof(double4), st,
                               c->gwann.p, (int)nb2,
                               (int)c->nwsk, (int)nslices, (int)c->npairk, c->gpair.p, c->ginvdeg.p, phi, (int)KT,
                               (int)n, out + (size_t)k0 * out_kstride, out_kstride);
        }
        HCHECK(hipGetLastError());
    }
}

static void gkRp_device(elph_ctx *c, int64_t nk, const double *d_kvecs, double2 *out, int64_t out_kstride,
                        bool mixed = false) {
    DBuf<double2> phiT;
    DBuf<v4i> bphi;
    phiT.alloc(gk_phi_count(c));
    if (mixed) {
        ensure_gwann_digits(c);
        bphi.alloc(gk_bphi_count(c));
    }
    launch_gkRp(c, c->stream, nk, d_kvecs, out, out_kstride, mixed, phiT.p, bphi.p);
    HCHECK(hipStreamSynchronize(c->stream));
}

extern "C" {

int elph_gkRp(elph_ctx *ctx, int64_t nk, const double *kvecs, double *out) {
    if (!ctx || nk < 0) return ELPH_EINVAL;
    try {
        HCHECK(hipSetDevice(ctx->device));
        const int64_t nb2 = ctx->nb * ctx->nb, nslices = ctx->nbr * ctx->nwsg;
        const int64_t blk = nb2 * ctx->nbr * ctx->nwsq;  // per-k output (nb,nb,nbr,nwsq)
        DBuf<double> dk;
        DBuf<double2> dout;
        dk.upload(kvecs, 3 * nk, ctx->stream);
        dout.alloc(nk * nb2 * nslices);
        gkRp_device(ctx, nk, dk.p, dout.p, nb2 * nslices);
        std::vector<double2> h(nk * nb2 * nslices);
        HCHECK(hipMemcpyAsync(h.data(), dout.p, h.size() * sizeof(double2), hipMemcpyDeviceToHost, ctx->stream));
        HCHECK(hipStreamSynchronize(ctx->stream));
        for (int64_t k = 0; k < nk; ++k) {
            std::memcpy(out + 2 * k * blk, &h[k * nb2 * nslices], nb2 * nslices * sizeof(double2));
            if (blk > nb2 * nslices)
                std::memset(out + 2 * (k * blk + nb2 * nslices), 0, (blk - nb2 * nslices) * sizeof(double2));
        }
    } catch (const Error &e) {
        return ret_error(ctx, e);
    }
    return ELPH_OK;
}

int elph_g2(elph_ctx *ctx, int64_t n, const double *kvecs, const double *qvecs, const double *el_evec_k,
            const double *el_evec_kp, const double *ph_evec_q, const double *ph_en, const double *gmixed,
            int64_t ngmixed, const int64_t *gidx, char wannspace, double *out) {
    if (!ctx) return ELPH_EINVAL;
    if (wannspace != 'e' && wannspace != 'p') {
        ctx->err = "Invalid value of wannspace in call to g2_epw. Exiting.";
        return ELPH_EINVAL;
    }
    if (n == 0) return ELPH_OK;
    try {
        HCHECK(hipSetDevice(ctx->device));
        const int64_t nb = ctx->nb, nbr = ctx->nbr;
        const bool ph = wannspace == 'p';
        const int64_t nws = ph ? ctx->nwsg : ctx->nwsk;
        for (int64_t i = 0; i < n; ++i)
            if (gidx[i] < 0 || gidx[i] >= ngmixed) throw Error{ELPH_EINVAL, "gidx out of range"};
        if (ph && ctx->nwsg > ctx->nwsq) throw Error{ELPH_EINVAL, "g2 'ph' sums nwsg cells of rcells_q: nwsg > nwsq"};
        DBuf<double> dv, den;
        DBuf<double2> phase, ek, ekp, uq, gm;
        DBuf<int64_t> dg;
        dv.upload(ph ? qvecs : kvecs, 3 * n, ctx->stream);
        if (ph) phase_list(ctx, n, dv.p, ctx->rq, ctx->degq, nws, phase);
        else phase_list(ctx, n, dv.p, ctx->rk, ctx->degk, nws, phase);
        // inputs are (n,nb) column-major; the kernel reads [v][i]
        auto rows = [&](const double *src, int64_t w) {
            std::vector<double2> r(n * w);
            for (int64_t v = 0; v < n; ++v)
                for (int64_t i = 0; i < w; ++i) r[v * w + i] = make_double2(src[2 * (v + n * i)], src[2 * (v + n * i) + 1]);
            return r;
        };
        ek.upload(rows(el_evec_k, nb), ctx->stream);
        ekp.upload(rows(el_evec_kp, nb), ctx->stream);
        std::vector<double2> u = rows(ph_evec_q, nbr);
        for (int64_t v = 0; v < n; ++v)
            for (int64_t p = 0; p < nbr; ++p) {
                const double s = 1.0 / std::sqrt(ctx->masses[p / 3]);
                u[v * nbr + p].x *= s;
                u[v * nbr + p].y *= s;
            }
        uq.upload(u, ctx->stream);
        den.upload(ph_en, n, ctx->stream);
        gm.upload(reinterpret_cast<const double2 *>(gmixed), (size_t)ngmixed * nb * nb * nbr * nws, ctx->stream);
        dg.upload(gidx, n, ctx->stream);
        DBuf<double> dout;
        DBuf<double2> Lq, lp;
        if (ctx->polar) {
            DBuf<double> dqv;
            if (ph) dqv.alloc(1);
            else dqv.upload(qvecs, 3 * n, ctx->stream);
            Lq.alloc(n * nbr);
            lp.alloc(n);
            polar_launch(ctx, ctx->stream, n, ph ? dv.p : dqv.p, nullptr, nullptr, Lq.p);
            hipLaunchKernelGGL(k_polar_lp, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, (int)n,
                               (int)nbr, 1, (const int *)nullptr, Lq.p, uq.p, lp.p);
            HCHECK(hipStreamSynchronize(ctx->stream));
        }
        dout.alloc(n);
        hipLaunchKernelGGL(k_g2_list, dim3((unsigned)n), dim3(256), 0, ctx->stream, (int)n, (int)nb, (int)nbr, (int)nws,
                           phase.p, ek.p, ekp.p, uq.p, den.p, gm.p, dg.p, ctx->polar ? lp.p : (const double2 *)nullptr,
                           dout.p);
        HCHECK(hipGetLastError());
        HCHECK(hipMemcpyAsync(out, dout.p, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        HCHECK(hipStreamSynchronize(ctx->stream));
    } catch (const Error &e) {
        return ret_error(ctx, e);
    }
    return ELPH_OK;
}

int elph_device_alloc(elph_ctx *ctx, int64_t bytes, void **dptr) {
    if (!ctx || !dptr || bytes < 0) return ELPH_EINVAL;
    *dptr = nullptr;
    if (hipSetDevice(ctx->device) != hipSuccess) return ELPH_EDEVICE;
    if (hipMalloc(dptr, (size_t)std::max<int64_t>(bytes, 1)) != hipSuccess) {
        ctx->err = "hipMalloc of " + std::to_string(bytes) + " bytes failed";
        *dptr = nullptr;
        return ELPH_ENOMEM;
    }
    return ELPH_OK;
}

int elph_device_free(elph_ctx *ctx, void *dptr) {
    if (!ctx) return ELPH_EINVAL;
    if (!dptr) return ELPH_OK;
    if (hipSetDevice(ctx->device) != hipSuccess) return ELPH_EDEVICE;
    return hipFree(dptr) == hipSuccess ? ELPH_OK : ELPH_EDEVICE;
}

int elph_device_to_host(elph_ctx *ctx, void *host, const void *dptr, int64_t bytes) {
    if (!ctx || bytes < 0) return ELPH_EINVAL;
    if (hipSetDevice(ctx->device) != hipSuccess) return ELPH_EDEVICE;
    if (hipMemcpyAsync(host, dptr, (size_t)bytes, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess) return ELPH_EDEVICE;
    return hipStreamSynchronize(ctx->stream) == hipSuccess ? ELPH_OK : ELPH_EDEVICE;
}

int elph_ctx_synchronize(elph_ctx *ctx) {
    if (!ctx) return ELPH_EINVAL;
    if (hipSetDevice(ctx->device) != hipSuccess) return ELPH_EDEVICE;
    return hipDeviceSynchronize() == hipSuccess ? ELPH_OK : ELPH_EDEVICE;
}

int elph_host_to_device(elph_ctx *ctx, void *dptr, const void *host, int64_t bytes) {
    if (!ctx || bytes < 0 || (bytes > 0 && (!dptr || !host))) return ELPH_EINVAL;
    try {
        HCHECK(hipSetDevice(ctx->device));
        HCHECK(hipMemcpyAsync(dptr, host, (size_t)bytes, hipMemcpyHostToDevice, ctx->stream));
        HCHECK(hipStreamSynchronize(ctx->stream));
    } catch (const Error &e) {
        return ret_error(ctx, e);
    }
    return ELPH_OK;
}

int elph_device_copy(elph_ctx *ctx, void *dst, const void *src, int64_t bytes) {
    if (!ctx || bytes < 0) return ELPH_EINVAL;
    try {
        HCHECK(hipSetDevice(ctx->device));
        HCHECK(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToDevice, ctx->stream));
        HCHECK(hipStreamSynchronize(ctx->stream));
    } catch (const Error &e) {
        return ret_error(ctx, e);
    }
    return ELPH_OK;
}

int elph_device_meminfo(int device, int64_t *free_bytes, int64_t *total_bytes) {
    size_t f = 0, t = 0;
    if (hipSetDevice(device) != hipSuccess || hipMemGetInfo(&f, &t) != hipSuccess) return ELPH_EDEVICE;
    if (free_bytes) *free_bytes = (int64_t)f;
    if (total_bytes) *total_bytes = (int64_t)t;
    return ELPH_OK;
}

int elph_synth_vertex(elph_ctx *ctx, uint64_t seed) {
    if (!ctx) return ELPH_EINVAL;
    try {
        HCHECK(hipSetDevice(ctx->device));
        const int64_t n = ctx->nb * ctx->nb * ctx->nwsk * ctx->nbr * ctx->nwsg;
        const double *L = ctx->lattvecs;
        hipLaunchKernelGGL(k_synth_vertex, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, n,
                           (int)(ctx->nb * ctx->nb), (int)ctx->nwsk, (int)ctx->nbr, ctx->rcart_k.p, ctx->rg.p, L[0],
                           L[1], L[2], L[3], L[4], L[5], L[6], L[7], L[8], 0.1 / kRyd2eV, seed, ctx->gwann.p);
        HCHECK(hipGetLastError());
        ctx->gdig.release();  // digits of the previous vertex, if any
        HCHECK(hipStreamSynchronize(ctx->stream));
    } catch (const Error &e) {
        return ret_error(ctx, e);
    }
    return ELPH_OK;
}

int elph_selftest_mfma(int device, double *max_err) {
    try {
        HCHECK(hipSetDevice(device));
        double A[64], B[64], D[256], R[256];
        for (int i = 0; i < 64; ++i) { A[i] = 0.25 * (i % 7) - 0.5 + 0.01 * i; B[i] = 0.125 * (i % 5) - 0.3 + 0.02 * i; }
        for (int i = 0; i < 16; ++i)
            for (int j = 0; j < 16; ++j) {
                double s = 0;
                for (int k = 0; k < 4; ++k) s += A[i * 4 + k] * B[k * 16 + j];
                R[i * 16 + j] = s;
            }
        DBuf<double> dA, dB, dD;
        dA.upload(A, 64, nullptr);
        dB.upload(B, 64, nullptr);
        dD.alloc(256);
        hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, nullptr, dA.p, dB.p, dD.p);
        HCHECK(hipGetLastError());
        HCHECK(hipMemcpy(D, dD.p, sizeof(D), hipMemcpyDeviceToHost));
        double e = 0;
        for (int i = 0; i < 256; ++i) e = std::max(e, std::fabs(D[i] - R[i]));
        *max_err = e;
    } catch (const Error &e) {
        std::fprintf(stderr, "elph_selftest_mfma: %s\n", e.msg.c_str());
        return e.code;
    }
    return ELPH_OK;
}

}  // extern "C"

// ===========================================================================
// ibzk plan (calculate_eph_interaction_ibzk, key 'g')
// ===========================================================================
namespace {

// Device tables of one group of IBZ k of a run (up to the k of one launch of the
// Re contraction: 64 in FP64, 128 in mixed runs) and of their in-window
// states (built once; kernels then run from them without host work).
struct IbzkGroup {
    int64_t nkb = 0, nst = 0;
    std::vector<int64_t> slot;  // output row of each state (relative to first)
    DBuf<double> kv;            // (nkb,3) column-major
    DBuf<double2> Ukb;          // [kb][m][mp]
    DBuf<int> st_kb, st_m, st_kint;
    DBuf<int64_t> st_slot;
};

// Scratch of one stream running groups (sized by ibzk_reserve).
struct IbzkScratch {
    DBuf<double2> gk, h, phiT, Y;
    DBuf<v4i> bphi, bsl;
    DBuf<double> bscale, hmax;
};

}  // namespace

struct elph_ibzk_plan {
    elph_ctx *ctx = nullptr;
    int32_t precision = ELPH_FP64;
    int64_t mesh_ref = 1;
    int64_t nb = 0, nbr = 0, nwv = 0, nirr = 0, nprocs = 0;
    int64_t Nel[3], Nph[3];
    int64_t L = 1, Lf[3];
    // Rp cells in pairs (R_p, -R_p): Kp = 2 npair rounded up to 64 rows (cos, sin)
    int64_t npair = 0, Kp = 0, CP = 0;
    DBuf<int2> pairs;
    bool packed = false;  // L and the electron mesh up to 255 (byte-packed phase indices)
    std::vector<int> kint_irr;     // [nirr][3]
    std::vector<double> ens_irred; // host (nirr,nb)
    double enref = 0, fsthick = 0;
    std::vector<double2> Uk_host;  // [nirr][m][mp]
    DBuf<int> kp_int, nwin, wband, Rq;
    DBuf<int64_t> offs;
    DBuf<double2> Vkp, u, uf, T, lp, lpf;
    DBuf<double> om, omf;
    // per-run scratch of plan_run
    IbzkScratch scr;
    DBuf<double> g2stage;
    // mixed precision: digit table of the phases (all q of the electron mesh)
    int64_t nkc = 0;
    DBuf<uint2> Tdig;
    DBuf<unsigned> Rpk;
    // pair stage: fused kernels, or the general path (nb*nbr beyond the fused
    // kernels, or chosen by the batch autotuner) with column tile 16*gcb and
    // (FP64) 64*grb rows per block
    bool general = false;
    int gcb = 3, grb = 2;
    int64_t y_bytes_cap = 0;  // general path: bytes of Y per chunk of states (0: 512 MiB)
    int64_t CPf = 0;  // column count of the fused layout (16-rounded nb*nbr)
    ConsumerScratch cons;
};

namespace {

// workgroups along the K runs of k_rotate_digits (4 runs of 8 pairs per chunk)
int64_t rot_digits_nz(const elph_ibzk_plan *p) { return (p->nkc * 4 + kRotRuns - 1) / kRotRuns; }

template <int CB, int RB>
void launch_ibzk_y_fp64(hipStream_t s, const PairArgs &a, int st0, int ns, double2 *Y) {
    const size_t lds = (size_t)(2 * 32 * 16 * CB + a.L) * sizeof(double2) + (size_t)(a.Kp / 2) * sizeof(int4);
    dim3 grid((unsigned)((a.nwv + 64 * RB - 1) / (64 * RB)), (unsigned)ns, (unsigned)(a.CP / (16 * CB)));
    hipLaunchKernelGGL((k_ibzk_y_fp64<CB, RB>), grid, dim3(256), lds, s, a, st0, Y);
}

template <int CB>
void launch_ibzk_y_i8(hipStream_t s, const PairArgs &a, const elph_ibzk_plan *p, const IbzkScratch &sc, int st0,
                      int ns, double2 *Y) {
    const size_t lds = (size_t)2 * CB * 2 * kNS * 64 * sizeof(v4i) +
                       (size_t)((a.L + 1) & ~1) * sizeof(uint2) + (size_t)p->nkc * 32 * sizeof(unsigned);
    dim3 grid((unsigned)((a.nwv + 63) / 64), (unsigned)ns, (unsigned)(a.CP / (16 * CB)));
    hipLaunchKernelGGL(k_ibzk_y_i8<CB>, grid, dim3(256), lds, s, a, st0, p->Tdig.p, p->Rpk.p, sc.bsl.p,
                       sc.bscale.p, (int)p->nkc, (int)(a.CP / 16), Y);
}

int64_t general_chunk(const elph_ibzk_plan *p, int64_t nst) {
    const int64_t per_state = p->nwv * p->CP * (int64_t)sizeof(double2);
    const int64_t cap = p->y_bytes_cap > 0 ? p->y_bytes_cap : ((int64_t)1 << 29);
    return std::max<int64_t>(1, std::min<int64_t>(nst, cap / per_state));
}

// General path: Y of a chunk of states into HBM, then the epilogue.
void general_launch(elph_ibzk_plan *p, const PairArgs &a, int64_t nst, IbzkScratch &sc, hipStream_t s) {
    const int64_t chunk = general_chunk(p, nst);
    for (int64_t s0 = 0; s0 < nst; s0 += chunk) {
        const int ns = (int)std::min(chunk, nst - s0);
        if (p->precision == ELPH_MIXED) {
            switch (p->gcb) {
                case 1: launch_ibzk_y_i8<1>(s, a, p, sc, (int)s0, ns, sc.Y.p); break;
                case 2: launch_ibzk_y_i8<2>(s, a, p, sc, (int)s0, ns, sc.Y.p); break;
                default: launch_ibzk_y_i8<3>(s, a, p, sc, (int)s0, ns, sc.Y.p); break;
            }
        } else {
            const int key = p->gcb * 10 + p->grb;
            switch (key) {
                case 11: launch_ibzk_y_fp64<1, 1>(s, a, (int)s0, ns, sc.Y.p); break;
                case 12: launch_ibzk_y_fp64<1, 2>(s, a, (int)s0, ns, sc.Y.p); break;
                case 21: launch_ibzk_y_fp64<2, 1>(s, a, (int)s0, ns, sc.Y.p); break;
                case 22: launch_ibzk_y_fp64<2, 2>(s, a, (int)s0, ns, sc.Y.p); break;
                case 31: launch_ibzk_y_fp64<3, 1>(s, a, (int)s0, ns, sc.Y.p); break;
                case 32: launch_ibzk_y_fp64<3, 2>(s, a, (int)s0, ns, sc.Y.p); break;
                case 41: launch_ibzk_y_fp64<4, 1>(s, a, (int)s0, ns, sc.Y.p); break;
                default: launch_ibzk_y_fp64<4, 2>(s, a, (int)s0, ns, sc.Y.p); break;
            }
        }
        dim3 eg((unsigned)((p->nwv + 63) / 64), (unsigned)ns);
        if (p->nb <= 8 && p->nbr <= 6) {
            hipLaunchKernelGGL((k_ibzk_epi<8, 6>), eg, dim3(256), 0, s, a, (int)s0, sc.Y.p);
        } else {
            const size_t elds = epi_wave_lds(p->nb, p->nbr);
            const dim3 wg((unsigned)((p->nwv + 3) / 4), (unsigned)ns);
            if (epi_wave_stage(p->nb, p->nbr))
                hipLaunchKernelGGL(k_ibzk_epi_wave<true>, wg, dim3(256), elds, s, a, (int)s0, sc.Y.p);
            else
                hipLaunchKernelGGL(k_ibzk_epi_wave<false>, wg, dim3(256), elds, s, a, (int)s0, sc.Y.p);
        }
        HCHECK(hipGetLastError());
    }
}

void fused_launch(elph_ibzk_plan *p, const PairArgs &a, int64_t nst_batch, IbzkScratch &sc, hipStream_t s, int CB);

// Dynamic LDS of the pair-stage kernels for the plan's current path (the
// sizes fused_launch, launch_ibzk_y_* and general_launch's epilogue request).
size_t ibzk_pair_lds(const elph_ibzk_plan *p) {
    const size_t epi_fused = (size_t)64 * (p->CP + 1) * sizeof(double2);
    if (!p->general) {
        if (p->precision == ELPH_MIXED) {
            const int CB = (int)(p->CP / 16);
            return pair_i8_scale_off(CB, (int)p->L, (int)p->nkc) + (size_t)p->CP * sizeof(double);
        }
        return std::max((size_t)(2 * 32 * p->CP + p->L) * sizeof(double2) + (size_t)(p->Kp / 2) * sizeof(int4),
                        epi_fused);
    }
    const size_t y = p->precision == ELPH_MIXED
                         ? (size_t)2 * p->gcb * 2 * kNS * 64 * sizeof(v4i) + (size_t)((p->L + 1) & ~1) * sizeof(uint2) +
                               (size_t)p->nkc * 32 * sizeof(unsigned)
                         : (size_t)(2 * 32 * 16 * p->gcb + p->L) * sizeof(double2) + (size_t)(p->Kp / 2) * sizeof(int4);
    size_t e = 0;
    if (!(p->nb <= 8 && p->nbr <= 6)) e = epi_wave_lds(p->nb, p->nbr);
    return std::max(y, e);
}

// Arguments of the pair stage of group gr (operands in sc, output rows at out).
PairArgs pair_args(const elph_ibzk_plan *p, const IbzkGroup &gr, const IbzkScratch &sc, double *out) {
    const elph_ctx *c = p->ctx;
    PairArgs a{};
    a.h = sc.h.p;
    a.Kp = (int)p->Kp;
    a.CP = (int)p->CP;
    a.st_kint = gr.st_kint.p;
    a.st_slot = gr.st_slot.p;
    a.nwv = (int)p->nwv;
    a.kp_int = p->kp_int.p;
    a.Nel0 = (int)p->Nel[0]; a.Nel1 = (int)p->Nel[1]; a.Nel2 = (int)p->Nel[2];
    a.mesh_ref = (int)p->mesh_ref;
    a.Nph0 = (int)p->Nph[0]; a.Nph1 = (int)p->Nph[1]; a.Nph2 = (int)p->Nph[2];
    a.L = (int)p->L;
    a.Lf0 = (int)p->Lf[0]; a.Lf1 = (int)p->Lf[1]; a.Lf2 = (int)p->Lf[2];
    a.Lmul = phase_lmul(p->L);
    a.T = p->T.p;
    a.Rq = p->Rq.p;
    a.Vkp = p->Vkp.p;
    a.nwin = p->nwin.p;
    a.wband = p->wband.p;
    a.offs = p->offs.p;
    a.u = p->u.p;
    a.om = p->om.p;
    a.uf = p->uf.p;
    a.omf = p->omf.p;
    a.lp = c->polar ? p->lp.p : nullptr;
    a.lpf = (c->polar && p->mesh_ref > 1) ? p->lpf.p : nullptr;
    a.Ukb = gr.Ukb.p;
    a.st_kb = gr.st_kb.p;
    a.st_m = gr.st_m.p;
    a.nb = (int)p->nb;
    a.nbr = (int)p->nbr;
    a.out = out;
    a.nprocs = p->nprocs;
    return a;
}

// Pair stage of one group on stream s; with `timed`, a pair of the context's
// timing events brackets it on s (timing_flush adds the times to the
// dominant-kernel total; nothing here waits).
void pair_launch(elph_ibzk_plan *p, const IbzkGroup &gr, IbzkScratch &sc, hipStream_t s, double *out, bool timed) {
    elph_ctx *c = p->ctx;
    const int64_t nst_batch = gr.nst;
    const PairArgs a = pair_args(p, gr, sc, out);
    const int CB = (int)(p->CP / 16);
    if (timed) timed_begin(c, s);
    if (p->general) {
        general_launch(p, a, nst_batch, sc, s);
    } else {
        fused_launch(p, a, nst_batch, sc, s, CB);
    }
    HCHECK(hipGetLastError());
    if (timed) timed_end(c, s);
}

// Epilogue variant of the fused kernels: 1 for nb <= 8 and nbr <= 6, 2 for
// nb <= 16 and nbr <= 12 (loops unrolled to those bounds), 0 otherwise.
int epilogue_class(int64_t nb, int64_t nbr) {
    if (nb <= 8 && nbr <= 6) return 1;
    if (nb <= 16 && nbr <= 12) return 2;
    return 0;
}

// Kernel shape of the fused pair stage of a plan: precision, column blocks,
// epilogue class (FP64: the unrolled nb <= 16 form only where it keeps two
// waves per SIMD, CB <= 2) and byte-packed phase indices (FP64).
struct FusedShape {
    int32_t precision = 0;
    int CB = 0, e = 0;
    bool packed = false;
    bool operator<(const FusedShape &o) const {
        return std::tie(precision, CB, e, packed) < std::tie(o.precision, o.CB, o.e, o.packed);
    }
};
FusedShape fused_shape(const elph_ibzk_plan *p) {
    FusedShape f;
    f.precision = p->precision;
    f.CB = (int)(p->CP / 16);
    f.e = epilogue_class(p->nb, p->nbr);
    if (p->precision != ELPH_MIXED && f.e == 2 && f.CB > 2) f.e = 0;
    f.packed = p->precision != ELPH_MIXED && p->packed;
    return f;
}
// Dynamic LDS of the fused pair kernels of plan p.
size_t fused_lds(const elph_ibzk_plan *p) {
    if (p->precision == ELPH_MIXED)
        return pair_i8_scale_off((int)(p->CP / 16), (int)p->L, (int)p->nkc) + (size_t)p->CP * sizeof(double);
    return std::max((size_t)(2 * 32 * p->CP + p->L) * sizeof(double2) + (size_t)(p->Kp / 2) * sizeof(int4),
                    (size_t)64 * (p->CP + 1) * sizeof(double2));
}
// Rows of k' per tile of the fused kernels.
int64_t fused_rows(int32_t precision) { return precision == ELPH_MIXED ? 64 : 128; }

// Calls go(kernel) with the fused pair kernel of shape f: the per-group form
// (JOBS false) or the multi-material form (JOBS true); mixed precision, then FP64.
template <bool JOBS, typename Go>
void fused_dispatch_i8(const FusedShape &f, Go &&go) {
    {
        auto k8 = [&](auto cb, auto nbm, auto nbrm) {
            constexpr int C = decltype(cb)::value, M = decltype(nbm)::value, R = decltype(nbrm)::value;
            if constexpr (JOBS) go(k_pair_i8_jobs<C, M, R>);
            else go(k_pair_i8<C, M, R>);
        };
        using std::integral_constant;
        // (nb <= 8 and nbr <= 6 implies CB <= 3)
        switch (f.CB * 4 + f.e) {
            case 4 + 1: k8(integral_constant<int, 1>{}, integral_constant<int, 8>{}, integral_constant<int, 6>{}); break;
            case 8 + 1: k8(integral_constant<int, 2>{}, integral_constant<int, 8>{}, integral_constant<int, 6>{}); break;
            case 12 + 1: k8(integral_constant<int, 3>{}, integral_constant<int, 8>{}, integral_constant<int, 6>{}); break;
            case 4 + 2: k8(integral_constant<int, 1>{}, integral_constant<int, 16>{}, integral_constant<int, 12>{}); break;
            case 8 + 2: k8(integral_constant<int, 2>{}, integral_constant<int, 16>{}, integral_constant<int, 12>{}); break;
            case 12 + 2: k8(integral_constant<int, 3>{}, integral_constant<int, 16>{}, integral_constant<int, 12>{}); break;
            case 16 + 2: k8(integral_constant<int, 4>{}, integral_constant<int, 16>{}, integral_constant<int, 12>{}); break;
            case 4: k8(integral_constant<int, 1>{}, integral_constant<int, 0>{}, integral_constant<int, 0>{}); break;
            case 8: k8(integral_constant<int, 2>{}, integral_constant<int, 0>{}, integral_constant<int, 0>{}); break;
            case 12: k8(integral_constant<int, 3>{}, integral_constant<int, 0>{}, integral_constant<int, 0>{}); break;
            case 16: k8(integral_constant<int, 4>{}, integral_constant<int, 0>{}, integral_constant<int, 0>{}); break;
            default: throw Error{ELPH_EINVAL, "mixed precision fused kernel holds numwannbands*numbranches <= 64"};
        }
    }
}
template <bool JOBS, typename Go>
void fused_dispatch_fp64(const FusedShape &f, Go &&go) {
    auto launch = [&](auto packed) {
        constexpr bool P = decltype(packed)::value;
        auto k64 = [&](auto cb, auto nbm, auto nbrm) {
            constexpr int C = decltype(cb)::value, M = decltype(nbm)::value, R = decltype(nbrm)::value;
            if constexpr (JOBS) go(k_pair_fp64_jobs<C, M, R, P>);
            else go(k_pair_fp64<C, M, R, P>);
        };
        using std::integral_constant;
        switch (f.CB * 4 + f.e) {
            case 4 + 1: k64(integral_constant<int, 1>{}, integral_constant<int, 8>{}, integral_constant<int, 6>{}); break;
            case 8 + 1: k64(integral_constant<int, 2>{}, integral_constant<int, 8>{}, integral_constant<int, 6>{}); break;
            case 12 + 1: k64(integral_constant<int, 3>{}, integral_constant<int, 8>{}, integral_constant<int, 6>{}); break;
            case 4 + 2: k64(integral_constant<int, 1>{}, integral_constant<int, 16>{}, integral_constant<int, 12>{}); break;
            case 8 + 2: k64(integral_constant<int, 2>{}, integral_constant<int, 16>{}, integral_constant<int, 12>{}); break;
            case 4: k64(integral_constant<int, 1>{}, integral_constant<int, 0>{}, integral_constant<int, 0>{}); break;
            case 8: k64(integral_constant<int, 2>{}, integral_constant<int, 0>{}, integral_constant<int, 0>{}); break;
            case 12: k64(integral_constant<int, 3>{}, integral_constant<int, 0>{}, integral_constant<int, 0>{}); break;
            case 16: k64(integral_constant<int, 4>{}, integral_constant<int, 0>{}, integral_constant<int, 0>{}); break;
            case 20: k64(integral_constant<int, 5>{}, integral_constant<int, 0>{}, integral_constant<int, 0>{}); break;
            case 24: k64(integral_constant<int, 6>{}, integral_constant<int, 0>{}, integral_constant<int, 0>{}); break;
            default: throw Error{ELPH_EINVAL, "FP64 fused kernel holds numwannbands*numbranches <= 96"};
        }
    };
    if (f.packed) launch(std::true_type{});
    else launch(std::false_type{});
}

void fused_launch(elph_ibzk_plan *p, const PairArgs &a, int64_t nst_batch, IbzkScratch &sc, hipStream_t s, int CB) {
    (void)CB;
    const FusedShape f = fused_shape(p);
    const dim3 grid((unsigned)((p->nwv + fused_rows(p->precision) - 1) / fused_rows(p->precision)),
                    (unsigned)nst_batch);
    const size_t lds = fused_lds(p);
    if (p->precision == ELPH_MIXED) {
        const uint2 *at = p->Tdig.p;
        const unsigned *rpk = p->Rpk.p;
        const v4i *bs = sc.bsl.p;
        const double *bsc = sc.bscale.p;
        const int nkc = (int)p->nkc;
        fused_dispatch_i8<false>(f, [&](auto kern) {
            hipLaunchKernelGGL(kern, grid, dim3(256), lds, s, a, at, rpk, bs, bsc, nkc);
        });
    } else {
        fused_dispatch_fp64<false>(f, [&](auto kern) { hipLaunchKernelGGL(kern, grid, dim3(256), lds, s, a); });
    }
}

// One launch of the multi-material fused kernel of shape f over the jobs at
// d_jobs (njobs, ntiles tiles in all), with the largest LDS of their plans.
void fused_jobs_launch(const FusedShape &f, const PairJob *d_jobs, int njobs, int64_t ntiles, size_t lds, int ncu,
                       hipStream_t s) {
    const int per_cu = f.precision == ELPH_MIXED ? pair_blocks_per_cu(f.CB, f.e == 1 ? 8 : (f.e == 2 ? 16 : 0)) : 2;
    const unsigned nblk = (unsigned)std::min<int64_t>(ntiles, (int64_t)std::max(1, ncu) * per_cu);
    auto go = [&](auto kern) { hipLaunchKernelGGL(kern, dim3(nblk), dim3(256), lds, s, d_jobs, njobs, ntiles); };
    if (f.precision == ELPH_MIXED) fused_dispatch_i8<true>(f, go);
    else fused_dispatch_fp64<true>(f, go);
}

// Host side of a group of IBZ k and their in-window states (the contents of
// an IbzkGroup's device tables).
struct GroupHost {
    int64_t nkb = 0, nst = 0;
    std::vector<int64_t> slot;
    std::vector<double> kv;
    std::vector<double2> Ukb;
    std::vector<int> skb, sm, ski;
    std::vector<int64_t> sslot;
};

// Active states of first..last (counts -1 for out-of-window states) grouped by
// up to KB IBZ k; st_slot is the output row: relative to `first` when
// on_device (or, compact, the index among active states), else the row
// within the group.
std::vector<GroupHost> ibzk_group_hosts(const elph_ibzk_plan *p, int64_t first, int64_t last, bool on_device,
                                        int64_t *counts, bool compact, int64_t KB) {
    const int64_t nb = p->nb, nb2 = nb * nb;
    struct St { int64_t ik, m, slot; };
    std::vector<St> act;
    for (int64_t istate = first; istate <= last; ++istate) {
        int64_t m, ik;
        demux_state(istate, nb, &m, &ik);
        const double e = p->ens_irred[(ik - 1) + p->nirr * (m - 1)];
        const int64_t t = istate - first;
        if (std::fabs(e - p->enref) > p->fsthick) { counts[t] = -1; continue; }
        counts[t] = p->nprocs;
        act.push_back({ik - 1, m - 1, compact ? (int64_t)act.size() : t});
    }
    std::vector<GroupHost> out;
    size_t i0 = 0;
    while (i0 < act.size()) {
        std::vector<int64_t> ks;
        size_t i1 = i0;
        while (i1 < act.size()) {
            if (ks.empty() || ks.back() != act[i1].ik) {
                if ((int64_t)ks.size() == KB) break;
                ks.push_back(act[i1].ik);
            }
            ++i1;
        }
        GroupHost g;
        const int64_t nkb = (int64_t)ks.size(), nst = (int64_t)(i1 - i0);
        g.nkb = nkb;
        g.nst = nst;
        g.kv.resize(3 * nkb);
        g.Ukb.resize(nkb * nb2);
        for (int64_t b = 0; b < nkb; ++b) {
            for (int d = 0; d < 3; ++d) g.kv[b + nkb * d] = (double)p->kint_irr[3 * ks[b] + d] / (double)p->Nel[d];
            std::memcpy(&g.Ukb[b * nb2], &p->Uk_host[ks[b] * nb2], nb2 * sizeof(double2));
        }
        g.skb.resize(nst);
        g.sm.resize(nst);
        g.ski.resize(3 * nst);
        g.sslot.resize(nst);
        int64_t bidx = -1, prev = -1;
        for (int64_t j = 0; j < nst; ++j) {
            const St &a = act[i0 + j];
            if (a.ik != prev) { ++bidx; prev = a.ik; }
            g.skb[j] = (int)bidx;
            g.sm[j] = (int)a.m;
            for (int d = 0; d < 3; ++d) g.ski[3 * j + d] = p->kint_irr[3 * a.ik + d];
            g.sslot[j] = on_device ? a.slot : j;
            g.slot.push_back(a.slot);
        }
        out.push_back(std::move(g));
        i0 = i1;
    }
    return out;
}

// Device tables of a group (buffers of `g` that are borrowed and large enough
// are written in place); returns after the copies.
void ibzk_group_upload(const GroupHost &h, IbzkGroup &g, hipStream_t s) {
    g.nkb = h.nkb;
    g.nst = h.nst;
    g.slot = h.slot;
    g.kv.upload(h.kv, s);
    g.Ukb.upload(h.Ukb, s);
    g.st_kb.upload(h.skb, s);
    g.st_m.upload(h.sm, s);
    g.st_kint.upload(h.ski, s);
    g.st_slot.upload(h.sslot, s);
    HCHECK(hipStreamSynchronize(s));  // the host vectors may go out of scope
}

void ibzk_groups(elph_ibzk_plan *p, int64_t first, int64_t last, bool on_device, int64_t *counts, hipStream_t s,
                 std::vector<std::unique_ptr<IbzkGroup>> &groups, bool compact = false) {
    // IBZ k per group: the k of one launch of the Re contraction
    const int64_t KB = p->precision == ELPH_MIXED ? kGkKTMixed : kGkKT;
    for (const GroupHost &h : ibzk_group_hosts(p, first, last, on_device, counts, compact, KB)) {
        auto g = std::make_unique<IbzkGroup>();
        ibzk_group_upload(h, *g, s);
        groups.push_back(std::move(g));
    }
}

// Scratch for groups of up to nkb IBZ k and nst states; get(buffer, count)
// provides each buffer (allocates, or carves from a fixed region).
template <typename Get>
void ibzk_scratch(const elph_ibzk_plan *p, IbzkScratch &sc, int64_t nkb, int64_t nst, Get &&get) {
    elph_ctx *c = p->ctx;
    const int64_t nb2 = p->nb * p->nb, nslices = p->nbr * c->nwsg;
    get(sc.gk, nkb * nb2 * nslices);
    if (p->precision != ELPH_MIXED) get(sc.h, nst * p->Kp * p->CP);  // mixed: digits straight from gk
    get(sc.phiT, gk_phi_count(c));
    if (p->precision == ELPH_MIXED) {
        get(sc.bphi, gk_bphi_count(c));
        get(sc.bscale, nst * p->CP);
        get(sc.hmax, nst * rot_digits_nz(p) * p->CP);
        get(sc.bsl, nst * p->nkc * (p->CP / 16) * 2 * kNS * 64);
    }
    if (p->general) get(sc.Y, general_chunk(p, nst) * p->nwv * p->CP);
}

void ibzk_reserve(const elph_ibzk_plan *p, IbzkScratch &sc, int64_t nkb, int64_t nst) {
    ibzk_scratch(p, sc, nkb, nst, [](auto &buf, int64_t n) { buf.alloc((size_t)n); });
}

// Bytes of scratch and group tables for groups of up to nkb IBZ k and nst
// states (each buffer rounded up to 256 bytes, as Carve places them).
size_t ibzk_noalloc_bytes(const elph_ibzk_plan *p, int64_t nkb, int64_t nst, bool with_y) {
    size_t total = 0;
    auto add = [&total](auto &buf, int64_t n) {
        using T = std::remove_pointer_t<decltype(buf.p)>;
        total += ((std::max<int64_t>(n, 1) * sizeof(T)) + 255) & ~(size_t)255;
    };
    IbzkScratch probe;
    ibzk_scratch(p, probe, nkb, nst, [&](auto &buf, int64_t n) {
        if (with_y || (void *)&buf != (void *)&probe.Y) add(buf, n);
    });
    IbzkGroup g;
    add(g.kv, 3 * nkb);
    add(g.Ukb, nkb * p->nb * p->nb);
    add(g.st_kb, nst);
    add(g.st_m, nst);
    add(g.st_kint, 3 * nst);
    add(g.st_slot, nst);
    return total;
}

// One group: g(k,Rp) of its k, h = U_k g / phwsdeg, (mixed) digits, pair stage.
void ibzk_prep(elph_ibzk_plan *p, const IbzkGroup &g, IbzkScratch &sc, hipStream_t s) {
    elph_ctx *c = p->ctx;
    const int64_t nb = p->nb, nbr = p->nbr, nb2 = nb * nb, nslices = nbr * c->nwsg;
    launch_gkRp(c, s, g.nkb, g.kv.p, sc.gk.p, nb2 * nslices, p->precision == ELPH_MIXED, sc.phiT.p, sc.bphi.p);
    if (p->precision == ELPH_MIXED) {
        if (g.nst > 0) {
            const dim3 grid((unsigned)g.nkb, (unsigned)(p->CP / 16), (unsigned)rot_digits_nz(p));
            const size_t lds = (size_t)nb * nb * sizeof(double2);
            auto go = [&](auto kmax, auto kdig) {
                hipLaunchKernelGGL(kmax, grid, dim3(256), lds, s, (int)nb, (int)nbr, (int)p->npair, (int)p->nkc,
                                   (int)(p->CP / 16), (int)p->CP, (int)g.nst, sc.gk.p, nb2 * nslices, g.st_kb.p,
                                   g.st_m.p, g.Ukb.p, c->degq.p, p->pairs.p, sc.hmax.p, sc.bscale.p, sc.bsl.p);
                HCHECK(hipGetLastError());
                hipLaunchKernelGGL(kdig, grid, dim3(256), lds, s, (int)nb, (int)nbr, (int)p->npair, (int)p->nkc,
                                   (int)(p->CP / 16), (int)p->CP, (int)g.nst, sc.gk.p, nb2 * nslices, g.st_kb.p,
                                   g.st_m.p, g.Ukb.p, c->degq.p, p->pairs.p, sc.hmax.p, sc.bscale.p, sc.bsl.p);
            };
            if (nb <= 8) go(k_rotate_digits<false, 8>, k_rotate_digits<true, 8>);
            else if (nb <= 16) go(k_rotate_digits<false, 16>, k_rotate_digits<true, 16>);
            else go(k_rotate_digits<false, 0>, k_rotate_digits<true, 0>);
        }
        HCHECK(hipGetLastError());
    } else {
        launch_rotate_h(s, (int)g.nst, (int)nb, (int)nbr, (int)p->npair, (int)p->Kp, (int)p->CP, sc.gk.p,
                        nb2 * nslices, g.st_kb.p, g.st_m.p, g.Ukb.p, c->degq.p, p->pairs.p, sc.h.p);
    }
}

void ibzk_enqueue(elph_ibzk_plan *p, const IbzkGroup &g, IbzkScratch &sc, hipStream_t s, double *out, bool timed) {
    ibzk_prep(p, g, sc, s);
    pair_launch(p, g, sc, s, out, timed);
}

}  // namespace

extern "C" {

int elph_ibzk_plan_create(elph_ctx *ctx, const elph_electrons *el, const elph_phonons *ph, int64_t mesh_ref,
                          int32_t precision, elph_ibzk_plan **out) {
    if (!ctx || !el || !ph || !out) return ELPH_EINVAL;
    *out = nullptr;
    auto p = std::make_unique<elph_ibzk_plan>();
    try {
        HCHECK(hipSetDevice(ctx->device));
        if (precision != ELPH_FP64 && precision != ELPH_MIXED) throw Error{ELPH_EINVAL, "unknown precision"};
        if (mesh_ref < 1) throw Error{ELPH_EINVAL, "mesh_ref must be >= 1"};
        for (int d = 0; d < 3; ++d)
            if (el->wvmesh[d] != ph->wvmesh[d] * mesh_ref)
                throw Error{ELPH_EINVAL, "electron mesh must equal phonon mesh * mesh_ref"};
        if (ctx->nwsg > ctx->nwsq)
            throw Error{ELPH_EINVAL, "g2 'ph' sums nwsg cells with rcells_q/phwsdeg: needs nwsg <= nwsq"};
        if (el->nwv <= 0) throw Error{ELPH_EWINDOW, "No states found within Fermi window."};
        p->ctx = ctx;
        p->precision = precision;
        p->mesh_ref = mesh_ref;
        const int64_t nb = ctx->nb, nbr = ctx->nbr, nwv = el->nwv;
        p->nb = nb;
        p->nbr = nbr;
        p->nwv = nwv;
        p->nirr = el->nwv_irred;
        p->enref = el->enref;
        p->fsthick = el->fsthick;
        for (int d = 0; d < 3; ++d) { p->Nel[d] = el->wvmesh[d]; p->Nph[d] = ph->wvmesh[d]; }
        p->L = lcm64(lcm64(p->Nel[0], p->Nel[1]), p->Nel[2]);
        for (int d = 0; d < 3; ++d) p->Lf[d] = p->L / p->Nel[d];
        p->CP = ((nb * nbr + 15) / 16) * 16;
        p->CPf = p->CP;
        const std::vector<int2> pairs = cell_pairs(ctx->rcells_q.data(), ctx->nwsq, ctx->nwsg);
        p->npair = (int64_t)pairs.size();
        p->Kp = ((2 * p->npair + 63) / 64) * 64;
        // fused kernels hold all nb*nbr columns of a state in registers up to
        // 96 (FP64) / 64 (mixed); beyond, the general column-tiled path
        p->general = p->CP > (precision == ELPH_MIXED ? 64 : 96);
        if (p->general) {
            p->gcb = std::min<int>(3, (int)(p->CP / 16));
            p->CP = ((nb * nbr + 16 * p->gcb - 1) / (16 * p->gcb)) * (16 * p->gcb);
        }
        hipStream_t s = ctx->stream;
        // k' rows
        std::vector<int> kpi(3 * nwv), nw(nwv), wb(nwv * nb, 0);
        std::vector<int64_t> offs(nwv);
        std::vector<double2> V(nwv * nb * nb);
        int64_t acc = 0, total_win = 0;
        for (int64_t r = 0; r < nwv; ++r) {
            int64_t v[3];
            demux_vector0(el->indexlist[r], v, el->wvmesh);
            for (int d = 0; d < 3; ++d) kpi[3 * r + d] = (int)v[d];
            int cnt = 0;
            for (int64_t n = 0; n < nb; ++n)
                if (std::fabs(el->ens[r + nwv * n] - el->enref) <= el->fsthick) wb[r * nb + cnt++] = (int)n;
            nw[r] = cnt;
            offs[r] = acc;
            acc += (int64_t)cnt * nbr;
            total_win += cnt;
            for (int64_t n = 0; n < nb; ++n)
                for (int64_t j = 0; j < nb; ++j) {
                    const int64_t o = r + nwv * (n + nb * j);
                    V[(r * nb + n) * nb + j] = make_double2(el->evecs[2 * o], el->evecs[2 * o + 1]);
                }
        }
        if (total_win != el->nstates_inwindow)
            throw Error{ELPH_EINVAL, "nstates_inwindow does not match the in-window bands of ens"};
        p->nprocs = el->nstates_inwindow * nbr;
        p->kp_int.upload(kpi, s);
        p->nwin.upload(nw, s);
        p->wband.upload(wb, s);
        p->offs.upload(offs, s);
        p->Vkp.upload(V, s);
        // phonons on the phonon mesh: u = e / sqrt(M), omega
        const int64_t nq = ph->nwv;
        std::vector<double2> u(nq * nbr * nbr);
        std::vector<double> om(nq * nbr);
        for (int64_t q = 0; q < nq; ++q)
            for (int64_t sb = 0; sb < nbr; ++sb) {
                om[q * nbr + sb] = ph->ens[q + nq * sb];
                for (int64_t sp = 0; sp < nbr; ++sp) {
                    const int64_t o = q + nq * (sb + nbr * sp);
                    const double sc = 1.0 / std::sqrt(ctx->masses[sp / 3]);
                    u[(q * nbr + sb) * nbr + sp] = make_double2(ph->evecs[2 * o] * sc, ph->evecs[2 * o + 1] * sc);
                }
            }
        p->u.upload(u, s);
        p->om.upload(om, s);
        if (ctx->polar) {
            // long-range vertex prefactor of every phonon-mesh mode, q = ints / mesh
            std::vector<double> qv(3 * nq);
            for (int64_t i = 0; i < nq; ++i) {
                int64_t v[3];
                demux_vector0(i + 1, v, ph->wvmesh);
                for (int d = 0; d < 3; ++d) qv[i + nq * d] = (double)v[d] / (double)ph->wvmesh[d];
            }
            polar_prefactors(ctx, nq, qv, p->u, p->lp);
        }
        if (mesh_ref > 1) {
            // phonons at every electron-mesh q, interpolated by ph_wann on the device
            // (the reference calls ph_wann per k', src/interactions.f90:2241-2247)
            const int64_t nqf = p->Nel[0] * p->Nel[1] * p->Nel[2];
            std::vector<double> qv(3 * nqf);
            for (int64_t i = 0; i < nqf; ++i) {
                int64_t v[3];
                demux_vector0(i + 1, v, el->wvmesh);
                for (int d = 0; d < 3; ++d) qv[i + nqf * d] = (double)v[d] / (double)el->wvmesh[d];
            }
            std::vector<double> en(nqf * nbr), ev(2 * nqf * nbr * nbr);
            int st = elph_ph_wann(ctx, nqf, qv.data(), en.data(), ev.data(), nullptr);
            if (st != ELPH_OK) throw Error{st, ctx->err};
            std::vector<double2> uf(nqf * nbr * nbr);
            std::vector<double> omf(nqf * nbr);
            for (int64_t q = 0; q < nqf; ++q)
                for (int64_t sb = 0; sb < nbr; ++sb) {
                    omf[q * nbr + sb] = en[q + nqf * sb];
                    for (int64_t sp = 0; sp < nbr; ++sp) {
                        const int64_t o = q + nqf * (sb + nbr * sp);
                        const double sc = 1.0 / std::sqrt(ctx->masses[sp / 3]);
                        uf[(q * nbr + sb) * nbr + sp] = make_double2(ev[2 * o] * sc, ev[2 * o + 1] * sc);
                    }
                }
            p->uf.upload(uf, s);
            p->omf.upload(omf, s);
            if (ctx->polar) polar_prefactors(ctx, nqf, qv, p->uf, p->lpf);
        } else {
            p->uf.alloc(1);
            p->omf.alloc(1);
        }
        // phase table e^{i 2 pi j / L} and the padded Rp cell list (rcells_q(1:nwsg))
        std::vector<double2> T(p->L);
        for (int64_t j = 0; j < p->L; ++j) {
            const double x = kTwoPi * (double)j / (double)p->L;
            T[j] = make_double2(std::cos(x), std::sin(x));
        }
        p->T.upload(T, s);
        // first cell of each pair (rcells_q(1:nwsg) in pairs, padded with 0),
        // and (meshes and L up to 255) that cell reduced modulo the electron
        // mesh, one byte per component
        p->packed = p->L <= 255 && p->Nel[0] <= 255 && p->Nel[1] <= 255 && p->Nel[2] <= 255;
        std::vector<int> R(4 * (p->Kp / 2), 0);
        for (int64_t i = 0; i < p->npair; ++i) {
            unsigned v = 0;
            for (int d = 0; d < 3; ++d) {
                const int64_t c = ctx->rcells_q[pairs[i].x + ctx->nwsq * d];
                R[4 * i + d] = (int)c;
                if (p->packed) v |= (unsigned)fmodulo(c, p->Nel[d]) << (8 * d);
            }
            R[4 * i + 3] = (int)v;
        }
        p->Rq.upload(R, s);
        p->pairs.upload(pairs, s);
        if (precision == ELPH_MIXED) {
            p->nkc = p->Kp / 64;
            std::vector<uint2> tdig(p->L);
            for (int64_t j = 0; j < p->L; ++j) {
                signed char dr[kNS], di[kNS];
                split_digits<kNS>(T[j].x, dr);
                split_digits<kNS>(T[j].y, di);
                // digit pairs (cos, sin) of levels 0, 1 | 2, 3: the 16-bit
                // halves the A operands take (i8_pair_core)
                unsigned w[2] = {0u, 0u};
                for (int sd = 0; sd < kNS; ++sd)
                    w[sd >> 1] |= ((unsigned)(unsigned char)dr[sd] | (unsigned)(unsigned char)di[sd] << 8)
                                  << (16 * (sd & 1));
                tdig[j] = make_uint2(w[0], w[1]);
            }
            p->Tdig.upload(tdig, s);
            // cells reduced modulo the electron mesh, packed one byte per component
            std::vector<unsigned> rpk(p->Kp / 2, 0u);
            for (int64_t i = 0; i < p->npair; ++i) {
                unsigned v = 0;
                for (int d = 0; d < 3; ++d)
                    v |= (unsigned)fmodulo(ctx->rcells_q[pairs[i].x + ctx->nwsq * d], p->Nel[d]) << (8 * d);
                rpk[i] = v;
            }
            for (int d = 0; d < 3; ++d)
                if (p->Nel[d] > 255 || p->L > 255) throw Error{ELPH_EINVAL, "mixed precision needs meshes up to 255"};
            p->Rpk.upload(rpk, s);
            HCHECK(hipGetLastError());
        }
        require_lds(ctx, ibzk_pair_lds(p.get()), "the ibzk pair kernel");
        if (precision == ELPH_MIXED)
            require_lds(ctx, (size_t)nb * nb * sizeof(double2) + 4 * kRotRuns * 16 * sizeof(double),
                        "the mixed-precision B-operand kernel");
        // the int8 digits of gwann, once per context, so that runs allocate no
        // context data
        if (precision == ELPH_MIXED) ensure_gwann_digits(ctx);
        // IBZ data
        p->kint_irr.resize(3 * p->nirr);
        p->ens_irred.assign(el->ens_irred, el->ens_irred + p->nirr * nb);
        p->Uk_host.resize(p->nirr * nb * nb);
        for (int64_t i = 0; i < p->nirr; ++i) {
            int64_t v[3];
            demux_vector0(el->indexlist_irred[i], v, el->wvmesh);
            for (int d = 0; d < 3; ++d) p->kint_irr[3 * i + d] = (int)v[d];
            for (int64_t m = 0; m < nb; ++m)
                for (int64_t j = 0; j < nb; ++j) {
                    const int64_t o = i + p->nirr * (m + nb * j);
                    p->Uk_host[(i * nb + m) * nb + j] = make_double2(el->evecs_irred[2 * o], el->evecs_irred[2 * o + 1]);
                }
        }
        HCHECK(hipStreamSynchronize(s));
    } catch (const Error &e) {
        return ret_error(ctx, e);
    }
    *out = p.release();
    return ELPH_OK;
}

int64_t elph_ibzk_plan_nprocs(const elph_ibzk_plan *p) { return p ? p->nprocs : -1; }

int elph_ibzk_plan_destroy(elph_ibzk_plan *p) {
    delete p;
    return ELPH_OK;
}

int elph_ibzk_plan_run(elph_ibzk_plan *p, int64_t first, int64_t last, int64_t *counts, double *g2, int32_t on_device) {
    if (!p) return ELPH_EINVAL;
    elph_ctx *c = p->ctx;
    try {
        HCHECK(hipSetDevice(c->device));
        if (first < 1 || last > p->nirr * p->nb || last < first - 1)
            throw Error{ELPH_EINVAL, "istate range outside 1..nwv_irred*numwannbands"};
        hipStream_t s = c->stream;
        c->t_dominant_ms = 0;
        c->n_dominant = 0;
        c->tev_used = 0;
        std::vector<std::unique_ptr<IbzkGroup>> groups;
        ibzk_groups(p, first, last, on_device != 0, counts, s, groups);
        int64_t nkb = 1, nst = 1;
        for (auto &g : groups) {
            nkb = std::max(nkb, g->nkb);
            nst = std::max(nst, g->nst);
        }
        if (p->precision == ELPH_MIXED) ensure_gwann_digits(c);
        ibzk_reserve(p, p->scr, nkb, nst);
        if (!on_device) p->g2stage.alloc(nst * p->nprocs);
        HCHECK(hipEventRecord(c->ev_c, s));
        for (auto &g : groups) {
            ibzk_enqueue(p, *g, p->scr, s, on_device ? g2 : p->g2stage.p, true);
            if (!on_device) {
                for (int64_t j = 0; j < g->nst; ++j)
                    HCHECK(hipMemcpyAsync(g2 + g->slot[j] * p->nprocs, p->g2stage.p + j * p->nprocs,
                                          p->nprocs * sizeof(double), hipMemcpyDeviceToHost, s));
                HCHECK(hipStreamSynchronize(s));
            }
        }
        HCHECK(hipEventRecord(c->ev_d, s));
        HCHECK(hipEventSynchronize(c->ev_d));
        float ms = 0;
        HCHECK(hipEventElapsedTime(&ms, c->ev_c, c->ev_d));
        c->t_all_ms = ms;
        timing_flush(c);
    } catch (const Error &e) {
        c->tev_used = 0;
        return ret_error(c, e);
    }
    return ELPH_OK;
}

int elph_ibzk_plan_run_noalloc(elph_ibzk_plan *p, int64_t first, int64_t last, int64_t *counts, double *g2_device) {
    if (!p || !g2_device || !counts) return ELPH_EINVAL;
    elph_ctx *c = p->ctx;
    const int64_t y_cap = p->y_bytes_cap;
    try {
        HCHECK(hipSetDevice(c->device));
        if (first < 1 || last > p->nirr * p->nb || last < first - 1)
            throw Error{ELPH_EINVAL, "istate range outside 1..nwv_irred*numwannbands"};
        if (!c->noalloc_scratch) throw Error{ELPH_EDEVICE, "no scratch array address for this device"};
        if (p->precision == ELPH_MIXED && !c->gdig.p) throw Error{ELPH_EINVAL, "gwann digits missing"};
        const size_t cap = kNoAllocScratchBytes;
        // the largest group of IBZ k whose tables and scratch fit the array
        // (a general-path plan also bounds its chunk of Y by what is left)
        const int64_t nb = p->nb;
        auto need = [&](int64_t nkb) {
            if (p->general) {
                const size_t fixed = ibzk_noalloc_bytes(p, nkb, nkb * nb, false);
                const int64_t per = p->nwv * p->CP * (int64_t)sizeof(double2);
                if (fixed + per + 256 > cap) return cap + 1;
                p->y_bytes_cap = ((int64_t)(cap - fixed) - 256) / per * per;
            }
            return ibzk_noalloc_bytes(p, nkb, nkb * nb, true);
        };
        int64_t KB = 64;
        while (KB > 1 && need(KB) > cap) KB = (KB + 1) / 2;
        if (need(KB) > cap)
            throw Error{ELPH_ENOMEM, "one IBZ k of this material needs more than the " + std::to_string(cap >> 20) +
                                         " MiB scratch of the allocation-free entry point"};
        hipStream_t s = c->stream;
        c->t_dominant_ms = 0;
        c->n_dominant = 0;
        c->tev_used = 0;
        const std::vector<GroupHost> hosts = ibzk_group_hosts(p, first, last, true, counts, false, KB);
        int64_t nkb = 1, nst = 1;
        for (const GroupHost &h : hosts) {
            nkb = std::max(nkb, h.nkb);
            nst = std::max(nst, h.nst);
        }
        Carve cv{c->noalloc_scratch, cap};
        IbzkScratch sc;
        ibzk_scratch(p, sc, nkb, nst, [&cv](auto &buf, int64_t n) { cv.take(buf, (size_t)n); });
        IbzkGroup g;
        cv.take(g.kv, 3 * nkb);
        cv.take(g.Ukb, nkb * nb * nb);
        cv.take(g.st_kb, nst);
        cv.take(g.st_m, nst);
        cv.take(g.st_kint, 3 * nst);
        cv.take(g.st_slot, nst);
        HCHECK(hipEventRecord(c->ev_c, s));
        for (const GroupHost &h : hosts) {
            ibzk_group_upload(h, g, s);  // stream-ordered after the previous group's kernels
            ibzk_enqueue(p, g, sc, s, g2_device, true);
        }
        HCHECK(hipEventRecord(c->ev_d, s));
        HCHECK(hipEventSynchronize(c->ev_d));
        float ms = 0;
        HCHECK(hipEventElapsedTime(&ms, c->ev_c, c->ev_d));
        c->t_all_ms = ms;
        timing_flush(c);
    } catch (const Error &e) {
        p->y_bytes_cap = y_cap;
        c->tev_used = 0;
        return ret_error(c, e);
    }
    p->y_bytes_cap = y_cap;
    return ELPH_OK;
}

}  // extern "C"

namespace elph {
IbzkRows ibzk_rows(elph_ibzk_plan *p) {
    IbzkRows v;
    v.ctx = p->ctx;
    v.nb = (int)p->nb;
    v.nbr = (int)p->nbr;
    v.nwv = (int)p->nwv;
    v.mesh_ref = (int)p->mesh_ref;
    for (int d = 0; d < 3; ++d) {
        v.Nel[d] = (int)p->Nel[d];
        v.Nph[d] = (int)p->Nph[d];
    }
    v.nirr = p->nirr;
    v.nprocs = p->nprocs;
    v.kp_int = p->kp_int.p;
    v.wband = p->wband.p;
    v.offs = p->offs.p;
    v.om = p->om.p;
    v.omf = p->omf.p;
    v.kint_irr = p->kint_irr.data();
    v.ens_irred = p->ens_irred.data();
    v.enref = p->enref;
    v.fsthick = p->fsthick;
    v.scr = &p->cons;
    if (!p->cons.vmap.p && p->nprocs > 0) {
        if (p->nbr > 255) throw Error{ELPH_EINVAL, "more than 255 phonon branches"};
        std::vector<int64_t> offs(p->nwv);
        std::vector<int> nw(p->nwv), wb(p->nwv * p->nb);
        HCHECK(hipStreamSynchronize(p->ctx->stream));
        HCHECK(hipMemcpy(offs.data(), p->offs.p, p->nwv * sizeof(int64_t), hipMemcpyDeviceToHost));
        HCHECK(hipMemcpy(nw.data(), p->nwin.p, p->nwv * sizeof(int), hipMemcpyDeviceToHost));
        HCHECK(hipMemcpy(wb.data(), p->wband.p, wb.size() * sizeof(int), hipMemcpyDeviceToHost));
        std::vector<int2> vm(p->nprocs);
        for (int64_t r = 0; r < p->nwv; ++r)
            for (int64_t j = 0; j < nw[r]; ++j)
                for (int64_t sb = 0; sb < p->nbr; ++sb)
                    vm[offs[r] + j * p->nbr + sb] = make_int2((int)r, (wb[r * p->nb + j] << 8) | (int)sb);
        p->cons.vmap.upload(vm, p->ctx->stream);
        HCHECK(hipStreamSynchronize(p->ctx->stream));
    }
    return v;
}
}  // namespace elph

// ===========================================================================
// Batched multi-material driver: the ibzk work of many plans (materials of
// different band, mode and WS counts) captured once into a hipGraph.  The
// preparations of all groups (g(k,Rp) and the B operand) run on several
// streams; the pair stage of every group on the fused path then runs in one
// multi-material launch per kernel shape (k_pair_i8_jobs / k_pair_fp64_jobs
// over a work list of all those groups), and groups on the column-tiled path
// run their pair stage on their stream.  The pair-stage tile shape of each
// size class is picked by timing the candidates on the class's first group.
// ===========================================================================
struct elph_batch {
    struct Item {
        elph_ibzk_plan *p = nullptr;
        int64_t first = 0, last = 0;
        double *out = nullptr;
        std::vector<std::unique_ptr<IbzkGroup>> groups;
        int stream = 0, cls = -1;
        double work = 0;
        // fused path: each group's pair operands (B digits and column scales,
        // or the folded h) in buffers of its own, read by the shared launch
        struct Ops {
            DBuf<v4i> bsl;
            DBuf<double> bscale;
            DBuf<double2> h;
        };
        std::vector<std::unique_ptr<Ops>> ops;
    };
    // one multi-material pair launch: the jobs of one fused kernel shape
    struct JobList {
        FusedShape shape;
        std::vector<PairJob> host;
        DBuf<PairJob> dev;
        int64_t ntiles = 0;
        size_t lds = 0;
        int ncu = 0;
    };
    struct Class {
        int32_t precision = 0;
        int cb = 0, kb = 0;  // size buckets of nb*nbr and Kp
        int64_t members = 0;
        elph_batch_class_t rep{};
    };
    int device = 0;
    std::vector<hipStream_t> streams;
    std::vector<hipEvent_t> joins, prepped;  // per stream: all work / fused preparations done
    hipStream_t lstream = nullptr;           // the multi-material pair launches
    hipEvent_t fork = nullptr, ev0 = nullptr, ev1 = nullptr;
    std::vector<Item> items;
    std::vector<Class> classes;
    std::vector<std::unique_ptr<IbzkScratch>> scr;  // per stream
    std::vector<std::unique_ptr<JobList>> lists;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::string err;
};

namespace {

int col_bucket(int64_t cp) {
    const int64_t edges[] = {32, 64, 96, 192, 384};
    for (int i = 0; i < 5; ++i)
        if (cp <= edges[i]) return i;
    return 5;
}
int kp_bucket(int64_t kp) {
    const int64_t edges[] = {128, 320, 640};
    for (int i = 0; i < 3; ++i)
        if (kp <= edges[i]) return i;
    return 3;
}

struct TileCand {
    bool general;
    int gcb, grb;
};

void apply_cand(elph_ibzk_plan *p, const TileCand &t) {
    p->general = t.general;
    p->gcb = t.gcb;
    p->grb = t.grb;
    p->CP = t.general ? ((p->nb * p->nbr + 16 * t.gcb - 1) / (16 * t.gcb)) * (16 * t.gcb) : p->CPf;
}

// Times the candidates on the first group of `it` (stream st) and applies the best.
elph_batch_class_t autotune(elph_batch *b, elph_batch::Item &it, hipStream_t st, bool tune) {
    elph_ibzk_plan *p = it.p;
    const bool mixed = p->precision == ELPH_MIXED;
    std::vector<TileCand> cands;
    if (p->CPf <= (mixed ? 64 : 96)) cands.push_back({false, 0, 0});
    if (mixed) {
        for (int cb = 1; cb <= 3; ++cb) cands.push_back({true, cb, 1});
    } else {
        for (int cb = 1; cb <= 4; ++cb)
            for (int rb = 1; rb <= 2; ++rb) cands.push_back({true, cb, rb});
    }
    if (!tune) cands.resize(1);
    const IbzkGroup &g = *it.groups.front();
    IbzkScratch sc;
    double best = 1e300;
    TileCand bc = cands.front();
    for (const TileCand &t : cands) {
        apply_cand(p, t);
        if (p->ctx->lds_max > 0 && (int64_t)ibzk_pair_lds(p) > p->ctx->lds_max) continue;
        ibzk_reserve(p, sc, g.nkb, g.nst);
        ibzk_enqueue(p, g, sc, st, it.out, false);  // warm-up (also builds h for this layout)
        float tmin = 1e30f;
        for (int r = 0; r < 3; ++r) {
            HCHECK(hipEventRecord(b->ev0, st));
            pair_launch(p, g, sc, st, it.out, false);
            HCHECK(hipEventRecord(b->ev1, st));
            HCHECK(hipEventSynchronize(b->ev1));
            float ms = 0;
            HCHECK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
            tmin = std::min(tmin, ms);
        }
        if (tmin < best) {
            best = tmin;
            bc = t;
        }
    }
    apply_cand(p, bc);
    elph_batch_class_t r{};
    r.precision = p->precision;
    r.path = bc.general ? 1 : 0;
    r.cb = bc.general ? bc.gcb : (int)(p->CPf / 16);
    r.rb = bc.general ? bc.grb : (mixed ? 1 : 2);
    r.candidates = (int32_t)cands.size();
    r.cols = p->nb * p->nbr;
    r.kp = p->Kp;
    r.ms = best;
    return r;
}

}  // namespace

extern "C" {

int elph_batch_create(elph_batch **out, int device, int32_t nstreams) {
    if (!out || nstreams < 1 || nstreams > 16) return ELPH_EINVAL;
    *out = nullptr;
    auto b = std::make_unique<elph_batch>();
    try {
        HCHECK(hipSetDevice(device));
        b->device = device;
        for (int i = 0; i < nstreams; ++i) {
            hipStream_t st;
            HCHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
            b->streams.push_back(st);
            hipEvent_t e;
            HCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            b->joins.push_back(e);
            HCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            b->prepped.push_back(e);
            b->scr.push_back(std::make_unique<IbzkScratch>());
        }
        HCHECK(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
        HCHECK(hipStreamCreateWithFlags(&b->lstream, hipStreamNonBlocking));
        HCHECK(hipEventCreate(&b->ev0));
        HCHECK(hipEventCreate(&b->ev1));
    } catch (const Error &e) {
        return e.code;
    }
    *out = b.release();
    return ELPH_OK;
}

int elph_batch_add(elph_batch *b, elph_ibzk_plan *plan, int64_t first, int64_t last, double *g2_device,
                   int64_t *counts) {
    if (!b || !plan || !g2_device || !counts) return ELPH_EINVAL;
    if (b->exec) {
        b->err = "plans must be added before elph_batch_prepare";
        return ELPH_EINVAL;
    }
    try {
        if (plan->ctx->device != b->device) throw Error{ELPH_EINVAL, "plan and batch are on different devices"};
        if (first < 1 || last > plan->nirr * plan->nb || last < first)
            throw Error{ELPH_EINVAL, "istate range outside 1..nwv_irred*numwannbands"};
        HCHECK(hipSetDevice(b->device));
        elph_batch::Item it;
        it.p = plan;
        it.first = first;
        it.last = last;
        it.out = g2_device;
        ibzk_groups(plan, first, last, true, counts, b->streams[0], it.groups, true);
        int64_t nst = 0;
        for (auto &g : it.groups) nst += g->nst;
        it.work = (double)nst * plan->nwv * plan->Kp * plan->nb * plan->nbr;
        if (plan->precision == ELPH_MIXED) ensure_gwann_digits(plan->ctx);
        b->items.push_back(std::move(it));
    } catch (const Error &e) {
        b->err = e.msg;
        return e.code;
    }
    return ELPH_OK;
}

int elph_batch_prepare(elph_batch *b, int32_t autotune_tiles) {
    if (!b) return ELPH_EINVAL;
    try {
        HCHECK(hipSetDevice(b->device));
        // size classes: (precision, bucket of nb*nbr, bucket of Kp)
        std::map<std::tuple<int, int, int>, int> key2cls;
        for (auto &it : b->items) {
            if (it.groups.empty()) continue;
            const auto key = std::make_tuple((int)it.p->precision, col_bucket(it.p->CPf), kp_bucket(it.p->Kp));
            auto f = key2cls.find(key);
            if (f == key2cls.end()) {
                elph_batch::Class c;
                c.precision = it.p->precision;
                c.cb = std::get<1>(key);
                c.kb = std::get<2>(key);
                c.rep = autotune(b, it, b->streams[0], autotune_tiles != 0);
                key2cls[key] = (int)b->classes.size();
                b->classes.push_back(c);
                f = key2cls.find(key);
            } else {
                const elph_batch_class_t &r = b->classes[f->second].rep;
                apply_cand(it.p, {r.path == 1, r.cb, r.rb});
            }
            it.cls = f->second;
            b->classes[f->second].members += 1;
            b->classes[f->second].rep.members = b->classes[f->second].members;
        }
        // longest work first onto the least-loaded stream
        std::vector<size_t> order(b->items.size());
        std::iota(order.begin(), order.end(), 0);
        std::sort(order.begin(), order.end(), [&](size_t x, size_t y) { return b->items[x].work > b->items[y].work; });
        std::vector<double> load(b->streams.size(), 0.0);
        for (size_t i : order) {
            const int s = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            b->items[i].stream = s;
            load[s] += b->items[i].work;
        }
        for (auto &it : b->items)
            for (auto &g : it.groups) ibzk_reserve(it.p, *b->scr[it.stream], g->nkb, g->nst);
        // fused path: per-group operand buffers and the work lists of the
        // multi-material launches (one per kernel shape)
        std::map<FusedShape, size_t> shape2list;
        b->lists.clear();
        for (auto &it : b->items) {
            it.ops.clear();
            if (it.p->general) continue;
            elph_ibzk_plan *p = it.p;
            const FusedShape f = fused_shape(p);
            auto fl = shape2list.find(f);
            if (fl == shape2list.end()) {
                auto jl = std::make_unique<elph_batch::JobList>();
                jl->shape = f;
                jl->ncu = (int)p->ctx->ncu;
                fl = shape2list.emplace(f, b->lists.size()).first;
                b->lists.push_back(std::move(jl));
            }
            elph_batch::JobList &jl = *b->lists[fl->second];
            jl.lds = std::max(jl.lds, fused_lds(p));
            const int nrb = (int)((p->nwv + fused_rows(p->precision) - 1) / fused_rows(p->precision));
            for (auto &g : it.groups) {
                auto o = std::make_unique<elph_batch::Item::Ops>();
                IbzkScratch view;
                if (p->precision == ELPH_MIXED) {
                    o->bsl.alloc((size_t)g->nst * p->nkc * (p->CP / 16) * 2 * kNS * 64);
                    o->bscale.alloc((size_t)g->nst * p->CP);
                    view.bsl.borrow(o->bsl.p, o->bsl.n);
                    view.bscale.borrow(o->bscale.p, o->bscale.n);
                } else {
                    o->h.alloc((size_t)g->nst * p->Kp * p->CP);
                    view.h.borrow(o->h.p, o->h.n);
                }
                PairJob j{};
                j.a = pair_args(p, *g, view, it.out);
                j.Tdig = p->Tdig.p;
                j.Rpk = p->Rpk.p;
                j.bsl = view.bsl.p;
                j.bscale = view.bscale.p;
                j.nkc = (int)p->nkc;
                j.nrb = nrb;
                j.tile0 = jl.ntiles;
                jl.ntiles += g->nst * nrb;
                if (g->nst > 0) jl.host.push_back(j);
                it.ops.push_back(std::move(o));
            }
        }
        for (auto &jl : b->lists) jl->dev.upload(jl->host, b->streams[0]);
        HCHECK(hipDeviceSynchronize());
        // capture: fork from stream 0; every stream first prepares its fused
        // items' groups (event `prepped`), then runs its column-tiled items;
        // the multi-material pair launches run on lstream once every stream's
        // fused preparations are done, alongside the column-tiled work; all
        // join on stream 0
        hipStream_t s0 = b->streams[0];
        HCHECK(hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal));
        HCHECK(hipEventRecord(b->fork, s0));
        for (size_t i = 1; i < b->streams.size(); ++i) HCHECK(hipStreamWaitEvent(b->streams[i], b->fork, 0));
        HCHECK(hipStreamWaitEvent(b->lstream, b->fork, 0));
        std::vector<const elph_batch::Item *> order2;
        for (const auto &it : b->items)
            if (!it.p->general) order2.push_back(&it);
        for (const auto &it : b->items)
            if (it.p->general) order2.push_back(&it);
        bool general_phase = false;
        for (const elph_batch::Item *itp : order2) {
            const elph_batch::Item &it = *itp;
            if (it.p->general && !general_phase) {
                general_phase = true;
                for (size_t i = 0; i < b->streams.size(); ++i) HCHECK(hipEventRecord(b->prepped[i], b->streams[i]));
            }
            hipStream_t st = b->streams[it.stream];
            IbzkScratch &ss = *b->scr[it.stream];
            for (size_t gi = 0; gi < it.groups.size(); ++gi) {
                const IbzkGroup &g = *it.groups[gi];
                if (it.p->general) {
                    ibzk_enqueue(it.p, g, ss, st, it.out, false);
                    continue;
                }
                // the stream's preparation buffers, the group's operand buffers
                IbzkScratch view;
                view.gk.borrow(ss.gk.p, ss.gk.n);
                view.phiT.borrow(ss.phiT.p, ss.phiT.n);
                view.bphi.borrow(ss.bphi.p, ss.bphi.n);
                view.hmax.borrow(ss.hmax.p, ss.hmax.n);
                elph_batch::Item::Ops &o = *it.ops[gi];
                view.bsl.borrow(o.bsl.p, o.bsl.n);
                view.bscale.borrow(o.bscale.p, o.bscale.n);
                view.h.borrow(o.h.p, o.h.n);
                ibzk_prep(it.p, g, view, st);
            }
        }
        if (!general_phase)
            for (size_t i = 0; i < b->streams.size(); ++i) HCHECK(hipEventRecord(b->prepped[i], b->streams[i]));
        for (size_t i = 0; i < b->streams.size(); ++i) HCHECK(hipStreamWaitEvent(b->lstream, b->prepped[i], 0));
        for (auto &jl : b->lists)
            if (!jl->host.empty())
                fused_jobs_launch(jl->shape, jl->dev.p, (int)jl->host.size(), jl->ntiles, jl->lds, jl->ncu,
                                  b->lstream);
        HCHECK(hipGetLastError());
        for (size_t i = 1; i < b->streams.size(); ++i) {
            HCHECK(hipEventRecord(b->joins[i], b->streams[i]));
            HCHECK(hipStreamWaitEvent(s0, b->joins[i], 0));
        }
        HCHECK(hipEventRecord(b->joins[0], b->lstream));
        HCHECK(hipStreamWaitEvent(s0, b->joins[0], 0));
        HCHECK(hipStreamEndCapture(s0, &b->graph));
        HCHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    } catch (const Error &e) {
        b->err = e.msg;
        return e.code;
    }
    return ELPH_OK;
}

int elph_batch_run(elph_batch *b, double *ms) {
    if (!b || !b->exec) return ELPH_EINVAL;
    try {
        HCHECK(hipSetDevice(b->device));
        hipStream_t s0 = b->streams[0];
        HCHECK(hipEventRecord(b->ev0, s0));
        HCHECK(hipGraphLaunch(b->exec, s0));
        HCHECK(hipEventRecord(b->ev1, s0));
        HCHECK(hipEventSynchronize(b->ev1));
        float t = 0;
        HCHECK(hipEventElapsedTime(&t, b->ev0, b->ev1));
        if (ms) *ms = t;
    } catch (const Error &e) {
        b->err = e.msg;
        return e.code;
    }
    return ELPH_OK;
}

int64_t elph_batch_classes(const elph_batch *b, elph_batch_class_t *out, int64_t max) {
    if (!b) return -ELPH_EINVAL;
    for (int64_t i = 0; i < (int64_t)b->classes.size() && i < max; ++i) out[i] = b->classes[i].rep;
    return (int64_t)b->classes.size();
}

const char *elph_batch_error(const elph_batch *b) { return b ? b->err.c_str() : "null batch"; }

int elph_batch_destroy(elph_batch *b) {
    if (!b) return ELPH_EINVAL;
    (void)hipSetDevice(b->device);
    for (auto st : b->streams) (void)hipStreamSynchronize(st);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->lstream) (void)hipStreamSynchronize(b->lstream);
    for (auto e : b->joins) (void)hipEventDestroy(e);
    for (auto e : b->prepped) (void)hipEventDestroy(e);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    for (auto st : b->streams) (void)hipStreamDestroy(st);
    if (b->lstream) (void)hipStreamDestroy(b->lstream);
    delete b;
    return ELPH_OK;
}

int elph_ibzk_g2(elph_ctx *ctx, const elph_electrons *el, const elph_phonons *ph, int64_t mesh_ref, int32_t precision,
                 int64_t first, int64_t last, int64_t *counts, double *g2) {
    elph_ibzk_plan *p = nullptr;
    int st = elph_ibzk_plan_create(ctx, el, ph, mesh_ref, precision, &p);
    if (st != ELPH_OK) return st;
    st = elph_ibzk_plan_run(p, first, last, counts, g2, 0);
    elph_ibzk_plan_destroy(p);
    return st;
}

}  // extern "C"

// ===========================================================================
// gReq and the ibzq plan (calculate_eph_interaction_ibzq, key 'g')
// ===========================================================================
namespace {

// Batched gReq into device memory: out[q] (nb,nb,nbr,nwsk) for q vectors on the device.
// g(Re, q) for nq device q vectors in cell-pair form over Rp: the real
// cos/sin table of the pairs (KT q per launch, zero-padded rows and columns)
// against gwann folded on the fly (1/gwsdeg included).
constexpr int kGReqNCB = 8;  // 128 q per pass over gwann
void gReq_device(elph_ctx *c, int64_t nq, const double *d_qvecs, double2 *out, int64_t out_qstride) {
    const int64_t nb2 = c->nb * c->nb;
    const int64_t KT = 16 * kGReqNCB;
    const int64_t K2 = ((2 * c->npairg + 15) / 16) * 16;  // whole blocks of the K loop
    DBuf<double2> phiR;  // K2 x KT doubles, as double2 pairs for the zero fill
    phiR.alloc((K2 * KT + 1) / 2);
    double *phi = reinterpret_cast<double *>(phiR.p);
    const int64_t nrb = (nb2 * c->nwsk + 15) / 16;
    require_lds(c, greq_lds(kGReqNCB, (int)K2), "k_gReq (vertex and phase blocks, pair weights)");
    for (int64_t q0 = 0; q0 < nq; q0 += KT) {
        const int64_t n = std::min(KT, nq - q0);
        hipLaunchKernelGGL(k_zero2, dim3((unsigned)(((K2 * KT + 1) / 2 + 255) / 256)), dim3(256), 0, c->stream, phiR.p,
                           (K2 * KT + 1) / 2);
        dim3 grid((unsigned)((c->npairg + 127) / 128), (unsigned)n);
        hipLaunchKernelGGL(k_phase_pairs, grid, dim3(128), 0, c->stream, (int)n, (int)c->npairg, d_qvecs + q0, nq,
                           c->rgpcell.p, phi, (int64_t)1, KT);
        hipLaunchKernelGGL((k_gReq<kGReqNCB>), dim3((unsigned)((nrb + 3) / 4), (unsigned)c->nbr), dim3(256),
                           greq_lds(kGReqNCB, (int)K2), c->stream,
                           c->gwann.p, (int)nb2, (int)c->nwsk, (int)c->nbr, (int)c->npairg, c->rgpair.p,
                           c->rgpwt.p, phi, (int)KT, (int)K2, (int)n, out + (size_t)q0 * out_qstride, out_qstride);
        HCHECK(hipGetLastError());
    }
}

}  // namespace

struct elph_ibzq_plan {
    elph_ctx *ctx = nullptr;
    int32_t precision = ELPH_FP64;
    int64_t nb = 0, nbr = 0, nwv = 0, stride = 0;
    int64_t Nel[3], Nph[3], mesh_ref = 1;
    int64_t L = 1, Lf[3];
    int64_t Kp = 0, CP = 0;
    int64_t nirr_q = 0, nq = 0;
    std::vector<int64_t> q_irr;   // IBZ q muxed (phonon mesh)
    std::vector<double> ph_ens;   // (nq, nbr)
    std::vector<double> ph_evecs; // (nq, nbr, nbr) c128
    DBuf<int> kint, nwin, wband, lut;
    DBuf<double2> Vk, T;
    DBuf<double2> gq, h, st_u, Lq, st_lp;
    DBuf<double> qv, st_om;
    DBuf<int> st_qb, st_qe, rowmap;
    DBuf<int64_t> rowoff, st_tot, st_slot;
    DBuf<double> g2stage;
    bool general = false;  // nb^2 beyond the fused kernel: Y tiles in HBM, then the epilogue
    DBuf<double2> Y;
};

extern "C" {

int elph_gReq(elph_ctx *ctx, int64_t nq, const double *qvecs, double *out) {
    if (!ctx || nq < 0) return ELPH_EINVAL;
    try {
        HCHECK(hipSetDevice(ctx->device));
        const int64_t blk = ctx->nb * ctx->nb * ctx->nbr * ctx->nwsk;
        DBuf<double> dq;
        DBuf<double2> dout;
        dq.upload(qvecs, 3 * nq, ctx->stream);
        dout.alloc(nq * blk);
        gReq_device(ctx, nq, dq.p, dout.p, blk);
        HCHECK(hipMemcpyAsync(out, dout.p, nq * blk * sizeof(double2), hipMemcpyDeviceToHost, ctx->stream));
        HCHECK(hipStreamSynchronize(ctx->stream));
    } catch (const Error &e) {
        return ret_error(ctx, e);
    }
    return ELPH_OK;
}

int elph_ibzq_plan_create(elph_ctx *ctx, const elph_electrons *el, const elph_phonons *ph, int32_t precision,
                          elph_ibzq_plan **out) {
    if (!ctx || !el || !ph || !out) return ELPH_EINVAL;
    *out = nullptr;
    auto p = std::make_unique<elph_ibzq_plan>();
    try {
        HCHECK(hipSetDevice(ctx->device));
        if (precision != ELPH_FP64 && precision != ELPH_MIXED) throw Error{ELPH_EINVAL, "unknown precision"};
        const int64_t nb = ctx->nb, nbr = ctx->nbr, nwv = el->nwv;
        if (nwv <= 0) throw Error{ELPH_EWINDOW, "No states found within Fermi window."};
        for (int d = 0; d < 3; ++d) {
            if (ph->wvmesh[d] <= 0 || el->wvmesh[d] % ph->wvmesh[d] != 0)
                throw Error{ELPH_EINVAL, "electron mesh must be a multiple of the phonon mesh"};
            p->Nel[d] = el->wvmesh[d];
            p->Nph[d] = ph->wvmesh[d];
        }
        p->mesh_ref = p->Nel[0] / p->Nph[0];
        for (int d = 1; d < 3; ++d)
            if (p->Nel[d] != p->Nph[d] * p->mesh_ref)
                throw Error{ELPH_EINVAL, "electron mesh must equal phonon mesh * mesh_ref on every axis"};
        p->ctx = ctx;
        p->precision = precision;
        p->nb = nb;
        p->nbr = nbr;
        p->nwv = nwv;
        p->stride = el->nstates_inwindow * nb;
        // the fused kernel holds up to 96 columns; beyond, column tiles of 32
        p->general = nb * nb > 96;
        p->CP = p->general ? ((nb * nb + 31) / 32) * 32 : ((nb * nb + 15) / 16) * 16;
        // cos/sin rows of the Re cell pairs, padded to the chunk of 32 rows
        p->Kp = ((2 * ctx->npairk + 31) / 32) * 32;
        p->L = lcm64(lcm64(p->Nel[0], p->Nel[1]), p->Nel[2]);
        for (int d = 0; d < 3; ++d) p->Lf[d] = p->L / p->Nel[d];
        hipStream_t s = ctx->stream;
        const int64_t nmesh = p->Nel[0] * p->Nel[1] * p->Nel[2];
        std::vector<int> ki(3 * nwv), nw(nwv), wb(nwv * nb, 0), lut(nmesh, -1);
        std::vector<double2> V(nwv * nb * nb);
        int64_t total_win = 0;
        for (int64_t r = 0; r < nwv; ++r) {
            int64_t v[3];
            demux_vector0(el->indexlist[r], v, el->wvmesh);
            for (int d = 0; d < 3; ++d) ki[3 * r + d] = (int)v[d];
            lut[el->indexlist[r] - 1] = (int)r;
            int cnt = 0;
            for (int64_t n = 0; n < nb; ++n)
                if (std::fabs(el->ens[r + nwv * n] - el->enref) <= el->fsthick) wb[r * nb + cnt++] = (int)n;
            nw[r] = cnt;
            total_win += cnt;
            for (int64_t n = 0; n < nb; ++n)
                for (int64_t j = 0; j < nb; ++j) {
                    const int64_t o = r + nwv * (n + nb * j);
                    V[(r * nb + n) * nb + j] = make_double2(el->evecs[2 * o], el->evecs[2 * o + 1]);
                }
        }
        // a row holds sum_k nw(k) nw(k+q) <= nb * sum_k nw(k) entries: the row
        // stride nb * nstates_inwindow bounds it only when the count is right
        if (total_win != el->nstates_inwindow)
            throw Error{ELPH_EINVAL, "nstates_inwindow does not match the in-window bands of ens"};
        {
            const size_t fused = std::max((size_t)(2 * kPairQKC * p->CP + p->L) * sizeof(double2) +
                                              (size_t)(p->Kp / 2) * sizeof(int4),
                                          (size_t)64 * (p->CP + 1) * sizeof(double2));
            const size_t general = (size_t)(2 * 32 * 16 * 2 + p->L) * sizeof(double2) + (size_t)p->Kp * sizeof(int4);
            require_lds(ctx, p->general ? general : fused, "the ibzq pair kernel");
        }
        p->kint.upload(ki, s);
        p->nwin.upload(nw, s);
        p->wband.upload(wb, s);
        p->lut.upload(lut, s);
        p->Vk.upload(V, s);
        std::vector<double2> T(p->L);
        for (int64_t j = 0; j < p->L; ++j) {
            const double x = kTwoPi * (double)j / (double)p->L;
            T[j] = make_double2(std::cos(x), std::sin(x));
        }
        p->T.upload(T, s);
        p->nirr_q = ph->nwv_irred;
        p->nq = ph->nwv;
        p->q_irr.assign(ph->indexlist_irred, ph->indexlist_irred + ph->nwv_irred);
        p->ph_ens.assign(ph->ens, ph->ens + ph->nwv * nbr);
        p->ph_evecs.assign(ph->evecs, ph->evecs + 2 * ph->nwv * nbr * nbr);
        HCHECK(hipStreamSynchronize(s));
    } catch (const Error &e) {
        return ret_error(ctx, e);
    }
    *out = p.release();
    return ELPH_OK;
}

int64_t elph_ibzq_plan_stride(const elph_ibzq_plan *p) { return p ? p->stride : -1; }

int elph_ibzq_plan_destroy(elph_ibzq_plan *p) {
    delete p;
    return ELPH_OK;
}

int elph_ibzq_plan_run(elph_ibzq_plan *p, int64_t first, int64_t last, int64_t *counts, double *g2, int32_t on_device) {
    if (!p) return ELPH_EINVAL;
    elph_ctx *c = p->ctx;
    try {
        HCHECK(hipSetDevice(c->device));
        const int64_t nb = p->nb, nbr = p->nbr, nb2 = nb * nb;
        if (first < 1 || last > p->nirr_q * nbr || last < first - 1)
            throw Error{ELPH_EINVAL, "istate range outside 1..nwv_irred*numbranches"};
        hipStream_t s = c->stream;
        c->t_dominant_ms = 0;
        c->n_dominant = 0;
        HCHECK(hipEventRecord(c->ev_c, s));
        const int64_t QB = 16 * kGReqNCB;  // IBZ q per gReq pass (one pass over gwann)
        const int64_t gblk = nb2 * nbr * c->nwsk;
        int64_t istate = first;
        while (istate <= last) {
            // states of up to QB distinct IBZ q
            std::vector<int64_t> qs, sts;
            int64_t j = istate;
            for (; j <= last; ++j) {
                int64_t sb, iq;
                demux_state(j, nbr, &sb, &iq);
                if (qs.empty() || qs.back() != iq - 1) {
                    if ((int64_t)qs.size() == QB) break;
                    qs.push_back(iq - 1);
                }
                sts.push_back(j);
            }
            const int64_t nqb = (int64_t)qs.size(), nst = (int64_t)sts.size();
            std::vector<double> qv(3 * nqb);
            for (int64_t b = 0; b < nqb; ++b) {
                int64_t v[3];
                demux_vector0(p->q_irr[qs[b]], v, p->Nph);
                for (int d = 0; d < 3; ++d) qv[b + nqb * d] = (double)v[d] / (double)p->Nph[d];
            }
            p->qv.upload(qv, s);
            p->gq.alloc(nqb * gblk);
            gReq_device(c, nqb, p->qv.p, p->gq.p, gblk);
            std::vector<int> sqb(nst), sqe(3 * nst);
            std::vector<double2> su(nst * nbr);
            std::vector<double> som(nst);
            std::vector<int64_t> sslot(nst);
            int64_t bidx = -1, prev = -1;
            for (int64_t t = 0; t < nst; ++t) {
                int64_t sb, iq;
                demux_state(sts[t], nbr, &sb, &iq);
                if (iq - 1 != prev) { ++bidx; prev = iq - 1; }
                sqb[t] = (int)bidx;
                const int64_t iqf = p->q_irr[iq - 1];  // muxed, phonon mesh
                int64_t v[3];
                demux_vector0(iqf, v, p->Nph);
                for (int d = 0; d < 3; ++d) sqe[3 * t + d] = (int)(v[d] * p->mesh_ref);
                som[t] = p->ph_ens[(iqf - 1) + p->nq * (sb - 1)];
                for (int64_t sp = 0; sp < nbr; ++sp) {
                    const int64_t o = (iqf - 1) + p->nq * ((sb - 1) + nbr * sp);
                    const double sc = 1.0 / std::sqrt(c->masses[sp / 3]);
                    su[t * nbr + sp] = make_double2(p->ph_evecs[2 * o] * sc, p->ph_evecs[2 * o + 1] * sc);
                }
                sslot[t] = on_device ? (sts[t] - first) : t;
            }
            p->st_qb.upload(sqb, s);
            p->st_qe.upload(sqe, s);
            p->st_u.upload(su, s);
            p->st_om.upload(som, s);
            p->st_slot.upload(sslot, s);
            if (c->polar) {
                p->Lq.alloc(nqb * nbr);
                p->st_lp.alloc(nst);
                polar_launch(c, s, nqb, p->qv.p, nullptr, nullptr, p->Lq.p);
                hipLaunchKernelGGL(k_polar_lp, dim3((unsigned)((nst + 255) / 256)), dim3(256), 0, s, (int)nst,
                                   (int)nbr, 1, p->st_qb.p, p->Lq.p, p->st_u.p, p->st_lp.p);
            }
            p->h.alloc(nst * p->Kp * p->CP);
            const int64_t tot = nst * p->Kp * p->CP;
            hipLaunchKernelGGL(k_rotate_hq, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, (int)nst, (int)nb,
                               (int)nbr, (int)c->npairk, (int)p->Kp, (int)p->CP, p->gq.p, gblk, p->st_qb.p, p->st_u.p,
                               c->gpair.p, c->ginvdeg.p, p->h.p);
            p->rowmap.alloc(nst * p->nwv);
            p->rowoff.alloc(nst * p->nwv);
            p->st_tot.alloc(nst);
            hipLaunchKernelGGL(k_ibzq_rows, dim3((unsigned)((p->nwv + 255) / 256), (unsigned)nst), dim3(256), 0, s,
                               (int)nst, (int)p->nwv, p->st_qe.p, p->kint.p, (int)p->Nel[0], (int)p->Nel[1],
                               (int)p->Nel[2], p->lut.p, p->nwin.p, p->rowmap.p, p->rowoff.p);
            hipLaunchKernelGGL(k_scan_rows, dim3((unsigned)nst), dim3(1024), 0, s, (int)p->nwv, p->rowoff.p, p->st_tot.p);
            HCHECK(hipGetLastError());
            double *dst = g2;
            if (!on_device) {
                p->g2stage.alloc(nst * p->stride);
                dst = p->g2stage.p;
            }
            PairQArgs a{};
            a.h = p->h.p;
            a.Kp = (int)p->Kp;
            a.CP = (int)p->CP;
            a.nwv = (int)p->nwv;
            a.kint = p->kint.p;
            a.L = (int)p->L;
            a.Lf0 = (int)p->Lf[0]; a.Lf1 = (int)p->Lf[1]; a.Lf2 = (int)p->Lf[2];
            a.T = p->T.p;
            a.Rk = c->gpcell.p;
            a.npair = (int)c->npairk;
            a.Vk = p->Vk.p;
            a.nwin = p->nwin.p;
            a.wband = p->wband.p;
            a.rowmap = p->rowmap.p;
            a.rowoff = p->rowoff.p;
            a.st_om = p->st_om.p;
            a.st_lp = c->polar ? p->st_lp.p : nullptr;
            a.st_slot = p->st_slot.p;
            a.nb = (int)nb;
            a.out = dst;
            a.stride = p->stride;
            const int CB = (int)(p->CP / 16);
            const size_t lds = std::max((size_t)(2 * kPairQKC * p->CP + p->L) * sizeof(double2) +
                                            (size_t)(p->Kp / 2) * sizeof(int4),
                                        (size_t)64 * (p->CP + 1) * sizeof(double2));
            dim3 grid((unsigned)((p->nwv + 127) / 128), (unsigned)nst);
            HCHECK(hipEventRecord(c->ev_a, s));
            a.Nel0 = (int)p->Nel[0];
            a.Nel1 = (int)p->Nel[1];
            a.Nel2 = (int)p->Nel[2];
            a.Lmul = phase_lmul(p->L);
            const bool packed = p->L <= 255 && p->Nel[0] <= 255 && p->Nel[1] <= 255 && p->Nel[2] <= 255;
            auto launch = [&](auto pk) {
                constexpr bool PK = decltype(pk)::value;
                if (p->general) {
                    // chunks of states whose Y fits 512 MB
                    constexpr int CBg = 2;
                    const int64_t per = p->nwv * p->CP * (int64_t)sizeof(double2);
                    const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(nst, ((int64_t)1 << 29) / per));
                    p->Y.alloc(chunk * p->nwv * p->CP);
                    const size_t ldsg =
                        (size_t)(2 * 32 * 16 * CBg + p->L) * sizeof(double2) + (size_t)p->Kp * sizeof(int4);
                    for (int64_t t0 = 0; t0 < nst; t0 += chunk) {
                        const int64_t nc = std::min(chunk, nst - t0);
                        dim3 gy((unsigned)((p->nwv + 127) / 128), (unsigned)nc, (unsigned)(p->CP / (16 * CBg)));
                        hipLaunchKernelGGL((k_ibzq_y<CBg, PK>), gy, dim3(256), ldsg, s, a, (int)t0, p->Y.p);
                        dim3 ge((unsigned)((p->nwv + 63) / 64), (unsigned)nc);
                        hipLaunchKernelGGL(k_ibzq_epi, ge, dim3(256), 0, s, a, (int)t0, p->Y.p);
                    }
                    return;
                }
                switch (CB) {
                    case 1: hipLaunchKernelGGL((k_pair_q_fp64<1, PK>), grid, dim3(256), lds, s, a); break;
                    case 2: hipLaunchKernelGGL((k_pair_q_fp64<2, PK>), grid, dim3(256), lds, s, a); break;
                    case 3: hipLaunchKernelGGL((k_pair_q_fp64<3, PK>), grid, dim3(256), lds, s, a); break;
                    case 4: hipLaunchKernelGGL((k_pair_q_fp64<4, PK>), grid, dim3(256), lds, s, a); break;
                    case 5: hipLaunchKernelGGL((k_pair_q_fp64<5, PK>), grid, dim3(256), lds, s, a); break;
                    case 6: hipLaunchKernelGGL((k_pair_q_fp64<6, PK>), grid, dim3(256), lds, s, a); break;
                    default: throw Error{ELPH_EINVAL, "unsupported column count"};
                }
            };
            if (packed) launch(std::true_type{});
            else launch(std::false_type{});
            HCHECK(hipGetLastError());
            HCHECK(hipEventRecord(c->ev_b, s));
            std::vector<int64_t> tot_h(nst);
            HCHECK(hipMemcpyAsync(tot_h.data(), p->st_tot.p, nst * sizeof(int64_t), hipMemcpyDeviceToHost, s));
            HCHECK(hipStreamSynchronize(s));
            float ms = 0;
            HCHECK(hipEventElapsedTime(&ms, c->ev_a, c->ev_b));
            c->t_dominant_ms += ms;
            c->n_dominant += 1;
            for (int64_t t = 0; t < nst; ++t) {
                counts[sts[t] - first] = tot_h[t];
                if (tot_h[t] > p->stride) throw Error{ELPH_EINVAL, "ibzq count exceeds the row stride"};
            }
            if (!on_device) {
                for (int64_t t = 0; t < nst; ++t)
                    HCHECK(hipMemcpyAsync(g2 + (sts[t] - first) * p->stride, p->g2stage.p + t * p->stride,
                                          tot_h[t] * sizeof(double), hipMemcpyDeviceToHost, s));
                HCHECK(hipStreamSynchronize(s));
            }
            istate = j;
        }
        HCHECK(hipEventRecord(c->ev_d, s));
        HCHECK(hipEventSynchronize(c->ev_d));
        float ms = 0;
        HCHECK(hipEventElapsedTime(&ms, c->ev_c, c->ev_d));
        c->t_all_ms = ms;
    } catch (const Error &e) {
        return ret_error(c, e);
    }
    return ELPH_OK;
}

}  // extern "C"
