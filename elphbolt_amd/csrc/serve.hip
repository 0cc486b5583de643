// Request serving of the electron-phonon vertex at arbitrary k- and q-lists.
//
// A request names a resident material (an elph_ctx whose Wannier tensors stay
// in HBM) and gives a k-list and a q-list in crystal coordinates; the answer
// is |g_mn,s(k,q)|^2 (FP32, eV^2) for every k of the list, every q of the list
// and all bands m, n and branches s, with the reference's g2 'ph' convention
// (src/wannier.f90:903-1011): U_k = el_wann(k), U_k+q = el_wann(k+q), e_q =
// ph_wann(q), g(k,Rp) = gkRp(k) (src/wannier.f90:1073-1132).  Layout:
//   out[(((ik * nq + iq) * nb + m) * nb + n) * nbr + s].
//
// Execution (one host scheduler thread per server):
//   * admission: the request's resident data (lists, phonons, phase matrix,
//     output) is carved from one pre-allocated HBM pool, first come first
//     served while it fits; nothing is allocated or freed on the device while
//     serving (hipFree would synchronise the device);
//   * set-up of an admitted request (phonons at its q-list, the phase matrix
//     e^{i 2pi q.Rp} with the argument reduced in FP64, its int8 digits) runs
//     on a set-up stream;
//   * work units of up to KT k-points of one material, drawn from one or more
//     admitted requests (continuous batching: the Re contraction of a unit is
//     one GEMM for all of them), run on slot streams, several in flight:
//       el_wann(k) -> gkRp (FP64 or int8 digits) -> rotate by U_k
//       per request segment: el_wann(k+q) -> Y = Phi h (FP64 or int8 matrix
//       cores) -> epilogue (phonon and band rotations, |g|^2, FP32 store);
//   * a request completes when its last segment's unit has completed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "ctx.h"

using namespace elph;

namespace {

constexpr int kKT = 64;                // k-points per unit (the gkRp column tile)
constexpr int64_t kPairMax = 1 << 16;  // (k, q) pairs per unit
constexpr int64_t kQMax = 4096;        // q rows per segment (multiple of 64)
constexpr int64_t kElChunk = 4096;     // points per el_wann launch
constexpr int64_t kGkMax = 1ll << 31;  // bytes of g(k,Rp) per unit
constexpr int64_t kYMax = 1ll << 29;   // bytes of Y per state chunk
constexpr int64_t kFlagRing = 1 << 16; // pinned status words (requests in flight < this)

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

int64_t rup(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

__global__ void k_zero_int(int *p) { *p = 0; }

// kq(p) = k(k0 + i) + q(q0 + j), p = i * nqs + j; vectors are (n,3) column-major.
__global__ void k_srv_kq(int nks, int nqs, const double *__restrict__ kv, int ldk, int k0, const double *__restrict__ qv,
                         int ldq, int q0, double *__restrict__ out, int ldo) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nks * nqs) return;
    const int i = p / nqs, j = p % nqs;
#pragma unroll
    for (int d = 0; d < 3; ++d) out[p + (size_t)ldo * d] = kv[k0 + i + (size_t)ldk * d] + qv[q0 + j + (size_t)ldq * d];
}

// Phonon post-processing of src/wannier.f90:740-780 on the device: omega =
// sign(sqrt|w2|, w2) Ryd2eV; modes 1..3 zero at q = 0; (-0.005, 0) -> 0 and a
// flag for omega < -0.005 eV.  u = e / sqrt(M) in place of the eigenvectors.
__global__ void k_ph_post(int nq, int nbr, const double *__restrict__ qv, int ldq, const double *__restrict__ W,
                          double2 *__restrict__ V, const double *__restrict__ invsqrtm, double *__restrict__ om,
                          int *__restrict__ flag) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    const bool gamma = qv[q] == 0.0 && qv[q + ldq] == 0.0 && qv[q + 2 * (size_t)ldq] == 0.0;
    int bad = 0;
    for (int s = 0; s < nbr; ++s) {
        const double o2 = W[(size_t)q * nbr + s];
        const double a = sqrt(fabs(o2));
        double w = (signbit(o2) ? -a : a) * kRyd2eV;
        if (gamma && s < 3) w = 0.0;
        if (w < -0.005) bad = 1;
        else if (w < 0.0 && w > -0.005) w = 0.0;
        om[(size_t)q * nbr + s] = w;
        for (int p = 0; p < nbr; ++p) {
            double2 &v = V[((size_t)q * nbr + s) * nbr + p];
            v.x *= invsqrtm[p];
            v.y *= invsqrtm[p];
        }
    }
    if (bad) *flag = 1;  // vector store; every writer stores the same value
}

// PhiT[K2][q] (real, row stride ldp) of the cell pairs (R_p, -R_p): cos (K2 =
// 2p) and sin (K2 = 2p + 1) of 2 pi (q . R_p - nint(q . R_p)); 0 past the
// pairs and for columns q >= nq.  With the folded h of k_rotate_all,
// sum_K2 PhiT[K2][q] h[K2] = sum_R e^{i 2 pi q . R} h(R).
__global__ void k_srv_phase(int nq, int ldp, int Kp, int npair, const double *__restrict__ qv, int ldq,
                            const int *__restrict__ pcell, double *__restrict__ phi) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)ldp * (Kp / 2)) return;
    const int p = (int)(t / ldp), q = (int)(t % ldp);
    double c = 0.0, sn = 0.0;
    if (q < nq && p < npair) {
        const double x = qv[q] * (double)pcell[3 * p] + qv[q + ldq] * (double)pcell[3 * p + 1] +
                         qv[q + 2 * (size_t)ldq] * (double)pcell[3 * p + 2];
        sincos(kTwoPi * (x - rint(x)), &sn, &c);
    }
    phi[(size_t)(2 * p) * ldp + q] = c;
    phi[(size_t)(2 * p + 1) * ldp + q] = sn;
}

// int8 digits of Phi in the A-operand order of the 16x16x64 int8 MFMA:
// adig[((qb * nkc + kc) * kNS + d][lane], lane = i + 16 g holding rows
// qb*16 + i and K2 = kc*64 + 16 g .. +15; rows >= nq are 0.
__global__ void k_srv_phase_digits(int nq, int nqb, int nkc, int ldp, const double *__restrict__ phi,
                                   v4i *__restrict__ adig) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nqb * nkc * 64) return;
    const int lane = (int)(t % 64);
    const int kc = (int)((t / 64) % nkc);
    const int qb = (int)(t / (64 * nkc));
    const int q = qb * 16 + (lane & 15), g = lane >> 4;
    unsigned u[kNS][4] = {};
#pragma unroll
    for (int b = 0; b < 16; ++b) {
        const double x = q < nq ? phi[(size_t)(kc * 64 + g * 16 + b) * ldp + q] : 0.0;
        signed char d[kNS];
        split_digits<kNS>(x, d);
#pragma unroll
        for (int e = 0; e < kNS; ++e) put_byte(u[e], b, d[e]);
    }
    v4i *o = adig + ((size_t)qb * nkc + kc) * kNS * 64 + lane;
#pragma unroll
    for (int e = 0; e < kNS; ++e) o[e * 64] = as_v4i(u[e]);
}

// Y[st][r][col] = sum_K2 PhiT[K2][q0 + r] h[st0 + st][K2][col] (FP64 matrix
// cores; real phase operand, two products per K2).
// Block: 4 waves x 32 rows; grid (rows / 128, states, column tiles of 16*CB).
// Each K chunk of both operands (128 phase rows, K-major, and the 16*CB
// columns of h) is staged in LDS; the next chunk is loaded into registers
// while the current one computes.
template <int CB>
__global__ __launch_bounds__(256) void k_srv_y_fp64(const double2 *__restrict__ h, int Kp, int CPT, int st0,
                                                    const double *__restrict__ phiT, int ldp, int q0, int nqs,
                                                    double2 *__restrict__ Y) {
    constexpr int KC = 16, RB = 2, CP = 16 * CB, ROWS = 128;
    constexpr int NA = KC * ROWS / 256, NB = (KC * CP + 255) / 256;
    extern __shared__ double2 smem[];
    double *As = reinterpret_cast<double *>(smem);
    double2 *Bs = smem + KC * ROWS / 2;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int li = lane & 15, lk = lane >> 4;
    const int st = blockIdx.y, col0 = blockIdx.z * CP;
    const int row0 = blockIdx.x * ROWS;
    const double2 *B = h + (size_t)(st0 + st) * Kp * CPT + col0;
    double pa[NA];
    dv2 pb[NB];
    const dv2 *Bv = reinterpret_cast<const dv2 *>(B);
    dv2 *Bsv = reinterpret_cast<dv2 *>(Bs);
    auto load = [&](int ch) {
#pragma unroll
        for (int j = 0; j < NA; ++j) {
            const int i = tid + 256 * j, k = i / ROWS, r = i % ROWS;
            pa[j] = (row0 + r < nqs) ? phiT[(size_t)(ch * KC + k) * ldp + q0 + row0 + r] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            const int i = tid + 256 * j;
            if ((KC * CP) % 256 == 0 || i < KC * CP) pb[j] = Bv[(size_t)(ch * KC + i / CP) * CPT + i % CP];
        }
    };
    auto store = [&]() {
#pragma unroll
        for (int j = 0; j < NA; ++j) As[tid + 256 * j] = pa[j];
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            const int i = tid + 256 * j;
            if ((KC * CP) % 256 == 0 || i < KC * CP) Bsv[i] = pb[j];
        }
    };
    d4 cr[RB][CB], ci[RB][CB];
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
        for (int c = 0; c < CB; ++c) { cr[rb][c] = d4{0, 0, 0, 0}; ci[rb][c] = d4{0, 0, 0, 0}; }
    load(0);
    store();
    __syncthreads();
    const int nchunks = Kp / KC;
    for (int ch = 0; ch < nchunks; ++ch) {
        const bool more = ch + 1 < nchunks;
        if (more) load(ch + 1);
#pragma unroll
        for (int ks = 0; ks < KC / 4; ++ks) {
            const int kl = ks * 4 + lk;
            double av[RB];
#pragma unroll
            for (int rb = 0; rb < RB; ++rb) av[rb] = As[kl * ROWS + wave * 32 + rb * 16 + li];
#pragma unroll
            for (int c = 0; c < CB; ++c) {
                const double2 bv = Bs[kl * CP + c * 16 + li];
#pragma unroll
                for (int rb = 0; rb < RB; ++rb) {
                    cr[rb][c] = mfma_f64(av[rb], bv.x, cr[rb][c]);
                    ci[rb][c] = mfma_f64(av[rb], bv.y, ci[rb][c]);
                }
            }
        }
        __syncthreads();
        if (more) {
            store();
            __syncthreads();
        }
    }
    const int row_base = row0 + wave * 32;
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
        for (int c = 0; c < CB; ++c)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = row_base + rb * 16 + lk + 4 * r;
                if (row < nqs)
                    Y[((size_t)st * nqs + row) * CPT + col0 + c * 16 + li] = make_double2(cr[rb][c][r], ci[rb][c][r]);
            }
}

// h(R) = sum_mp U_k(m,mp) gk[k](np, mp, sp, R) / phwsdeg(R) (c = np + nb sp),
// folded over the cell pairs (R_p, -R_p) of the Rp list for all bands m:
//   h[(k * nb + m)][2p][c] = h(R_p) + h(-R_p),  h[(k * nb + m)][2p + 1][c] = i (h(R_p) - h(-R_p))
// (zero past the pairs or for c >= nb nbr; -R_p absent: h(-R_p) = 0).  One
// thread per (k, p, c) reads its two g columns once and writes all nb bands.
template <int NB>
__global__ __launch_bounds__(256) void k_rotate_all(int nb, int nbr, int npair, int Kp, int CPT,
                                                    const double2 *__restrict__ gk, int64_t kstride,
                                                    const double2 *__restrict__ Uk, const int *__restrict__ degq,
                                                    const int2 *__restrict__ pairs, double2 *__restrict__ h) {
    __shared__ double2 U[NB * NB];
    const int k = blockIdx.y;
    for (int i = threadIdx.x; i < nb * nb; i += blockDim.x) U[i] = Uk[(size_t)k * nb * nb + i];
    __syncthreads();
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)(Kp / 2) * CPT) return;
    const int p = (int)(t / CPT), c = (int)(t % CPT);
    const bool live = p < npair && c < nb * nbr;
    const int np = c % nb, sp = c / nb;
    const int2 pr = live ? pairs[p] : make_int2(0, -1);
    // h(R_p) for every band m
    double2 hi[NB];
    {
        double2 g[NB];
        if (live) {
            const double2 *src = gk + (size_t)k * kstride + ((size_t)(sp + nbr * pr.x)) * nb * nb + np;
            const double inv = 1.0 / (double)degq[pr.x];
#pragma unroll
            for (int mp = 0; mp < NB; ++mp) {
                g[mp] = mp < nb ? src[(size_t)mp * nb] : make_double2(0.0, 0.0);
                g[mp].x *= inv;
                g[mp].y *= inv;
            }
        }
#pragma unroll
        for (int m = 0; m < NB; ++m) {
            double2 acc = make_double2(0.0, 0.0);
            if (live && m < nb)
#pragma unroll
                for (int mp = 0; mp < NB; ++mp)
                    if (mp < nb) cfma(acc, U[m * nb + mp], g[mp]);
            hi[m] = acc;
        }
    }
    double2 g[NB];
    const bool partner = live && pr.y >= 0;
    if (partner) {
        const double2 *src = gk + (size_t)k * kstride + ((size_t)(sp + nbr * pr.y)) * nb * nb + np;
        const double inv = 1.0 / (double)degq[pr.y];
#pragma unroll
        for (int mp = 0; mp < NB; ++mp) {
            g[mp] = mp < nb ? src[(size_t)mp * nb] : make_double2(0.0, 0.0);
            g[mp].x *= inv;
            g[mp].y *= inv;
        }
    }
    double2 *dst = h + ((size_t)k * nb * Kp + 2 * p) * CPT + c;
#pragma unroll
    for (int m = 0; m < NB; ++m) {
        if (m < nb) {
            double2 hj = make_double2(0.0, 0.0);
            if (partner)
#pragma unroll
                for (int mp = 0; mp < NB; ++mp)
                    if (mp < nb) cfma(hj, U[m * nb + mp], g[mp]);
            dst[(size_t)m * Kp * CPT] = make_double2(hi[m].x + hj.x, hi[m].y + hj.y);
            dst[(size_t)m * Kp * CPT + CPT] = make_double2(hj.y - hi[m].y, hi[m].x - hj.x);
        }
    }
}

// Mixed-precision form of k_srv_y_fp64: kNS int8 digits of the real Phi
// (adig) and of h / column scale (bsl, [st][kc][cb][Re, Im][digit][lane]);
// the digit products with i + j < kNS accumulate exactly in int32 per level
// i + j and are recombined in FP64 with the weights 1 / (127^2 254^l).
// Block: 4 waves x 16 rows; grid (rows / 64, states, column tiles of 16*CB).
template <int CB>
__global__ __launch_bounds__(256, 2) void k_srv_y_i8(const v4i *__restrict__ bsl, int nkc, int CBT, int st0,
                                                     const double *__restrict__ bscale, int CPT,
                                                     const v4i *__restrict__ adig, int qb0, int nqs,
                                                     double2 *__restrict__ Y) {
    constexpr int NS = kNS;
    constexpr int BCH = CB * 2 * NS * 64;
    constexpr int NPF = (BCH + 255) / 256;
    extern __shared__ v4i smem8[];
    v4i *buf0 = smem8, *buf1 = smem8 + BCH;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int li = lane & 15, lg = lane >> 4;
    const int st = blockIdx.y, cb0 = blockIdx.z * CB;
    const int row_base = blockIdx.x * 64 + wave * 16;
    const int qb = qb0 + row_base / 16;
    const v4i *ast = adig + (size_t)qb * nkc * NS * 64 + lane;
    const v4i *bst = bsl + (size_t)(st0 + st) * nkc * CBT * 2 * NS * 64 + (size_t)cb0 * 2 * NS * 64;
    const size_t bkc = (size_t)CBT * 2 * NS * 64;  // v4i per K chunk of one state
    v4i acc[CB][NS][2];
#pragma unroll
    for (int c = 0; c < CB; ++c)
#pragma unroll
        for (int l = 0; l < NS; ++l) { acc[c][l][0] = v4i{0, 0, 0, 0}; acc[c][l][1] = v4i{0, 0, 0, 0}; }
    for (int i = tid; i < BCH; i += 256) buf0[i] = bst[i];
    v4i A[NS];
#pragma unroll
    for (int d = 0; d < NS; ++d) A[d] = ast[d * 64];
    __syncthreads();
    for (int kc = 0; kc < nkc; ++kc) {
        const int kn = min(kc + 1, nkc - 1);
        const v4i *src = bst + (size_t)kn * bkc;
        v4i pf[NPF];
#pragma unroll
        for (int j = 0; j < NPF; ++j) {
            const int i = tid + 256 * j;
            if (BCH % 256 == 0 || i < BCH) pf[j] = src[i];
        }
        v4i nA[NS];
        const v4i *an = ast + (size_t)kn * NS * 64;
#pragma unroll
        for (int d = 0; d < NS; ++d) nA[d] = an[d * 64];
        const v4i *B = (kc & 1) ? buf1 : buf0;
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
        }
#pragma unroll
        for (int d = 0; d < NS; ++d) A[d] = nA[d];
        v4i *dst = (kc & 1) ? buf0 : buf1;
#pragma unroll
        for (int j = 0; j < NPF; ++j) {
            const int i = tid + 256 * j;
            if (BCH % 256 == 0 || i < BCH) dst[i] = pf[j];
        }
        __syncthreads();
    }
    double wl[NS];
    wl[0] = 1.0 / (127.0 * 127.0);
#pragma unroll
    for (int l = 1; l < NS; ++l) wl[l] = wl[l - 1] / 254.0;
    const double *sc = bscale + (size_t)(st0 + st) * CPT + cb0 * 16;
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
            if (row < nqs) Y[((size_t)st * nqs + row) * CPT + (cb0 + c) * 16 + li] = make_double2(re * s, im * s);
        }
    }
}

struct EpiArgs {
    const double2 *Y;      // [st - st0][r][CPT]
    int CPT, nqs, st0, ns;  // states st0 .. st0 + ns of the segment
    int nb, nbr;
    int k_unit0, k_req0, q0, nq;
    const double2 *Ukq;    // [(kseg * nqs + r)][n][np]
    const double2 *Uk;     // unit [k][m][mp]
    const double2 *u;      // request [q][s][sp], e / sqrt(M)
    const double *om;      // request [q][s] (eV)
    const double2 *lp;     // request [q][s] polar prefactor, or null
    float *out;
    int G;                 // q rows handled together by one block
};

// Epilogue of the states (k, m) of one k and G q rows per block:
//   Z(n,sp) = sum_np conj(U_k+q(n,np)) Y(np + nb sp)
//   g(n,s)  = sum_sp u_q(s,sp) Z(n,sp) [+ lp(q,s) sum_i conj(U_k+q(n,i)) U_k(m,i)]
//   |g|^2   = 1/2 |g|^2 / omega * Ryd2eV^3 Ryd2amu (0 when omega == 0).
// The block loads U_k+q and u_q once and loops over the bands m of its k that
// lie in the launch's state range [st0, st0 + ns) (states st = k nb + m of the
// segment; Y holds them from st0 on).
__global__ __launch_bounds__(256) void k_srv_epi(EpiArgs a) {
    extern __shared__ double2 sm[];
    const int nb = a.nb, nbr = a.nbr, nbnr = nb * nbr;
    const int G = a.G, TPQ = 256 / G;
    const int tid = threadIdx.x, sub = tid / TPQ, tt = tid % TPQ;
    // grid (k of the launch, q-row groups): consecutive blocks are the k of
    // one q-row group, which share u_q while it is in L2
    const int kseg = a.st0 / nb + blockIdx.x;
    const int m_lo = max(0, a.st0 - kseg * nb), m_hi = min(nb, a.st0 + a.ns - kseg * nb);
    const int r = blockIdx.y * G + sub;
    const int per = 3 * nbnr + nb * nb + nbr * nbr;
    double2 *Ys = sm + sub * per, *Zs = Ys + nbnr, *Us = Zs + nbnr, *us = Us + nb * nb, *un = us + nbr * nbr;
    const bool live = sub < G && r < a.nqs;
    const int q = a.q0 + r;
    if (live) {
        const double2 *Ur = a.Ukq + ((size_t)kseg * a.nqs + r) * nb * nb;
        const double2 *ur = a.u + (size_t)q * nbr * nbr;
        for (int i = tt; i < nb * nb; i += TPQ) Us[i] = Ur[i];
        for (int i = tt; i < nbr * nbr; i += TPQ) us[i] = ur[i];
    }
    for (int m = m_lo; m < m_hi; ++m) {
        const int st = kseg * nb + m;
        if (live) {
            const double2 *Yr = a.Y + ((size_t)(st - a.st0) * a.nqs + r) * a.CPT;
            for (int i = tt; i < nbnr; i += TPQ) Ys[i] = Yr[i];
        }
        // Ys (and, before the first band, Us and us) complete; Zs and un of
        // the previous band no longer read
        __syncthreads();
        if (live) {
            for (int i = tt; i < nbnr; i += TPQ) {
                const int n = i % nb, sp = i / nb;
                double2 z = make_double2(0.0, 0.0);
                for (int np = 0; np < nb; ++np) {
                    const double2 t = cmulc(Us[n * nb + np], Ys[np + nb * sp]);
                    z.x += t.x;
                    z.y += t.y;
                }
                Zs[n + nb * sp] = z;
            }
            if (a.lp) {
                const double2 *Ukm = a.Uk + ((size_t)(a.k_unit0 + kseg) * nb + m) * nb;
                for (int n = tt; n < nb; n += TPQ) {
                    double2 z = make_double2(0.0, 0.0);
                    for (int i = 0; i < nb; ++i) {
                        const double2 t = cmulc(Us[n * nb + i], Ukm[i]);
                        z.x += t.x;
                        z.y += t.y;
                    }
                    un[n] = z;
                }
            }
        }
        // Zs and un complete; Ys no longer read
        __syncthreads();
        if (live) {
            float *o = a.out + ((((size_t)(a.k_req0 + kseg) * a.nq + q) * nb + m) * nb) * nbr;
            for (int i = tt; i < nbnr; i += TPQ) {
                const int n = i / nbr, s = i % nbr;
                double2 g = make_double2(0.0, 0.0);
                for (int sp = 0; sp < nbr; ++sp) cfma(g, us[s * nbr + sp], Zs[n + nb * sp]);
                if (a.lp) cfma(g, a.lp[(size_t)q * nbr + s], un[n]);
                const double w = a.om[(size_t)q * nbr + s];
                o[i] = (w == 0.0) ? 0.0f : (float)(0.5 * (g.x * g.x + g.y * g.y) / w * kG2Unit);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

// First-fit sub-allocator over one device allocation (256-byte granules).
struct Pool {
    char *base = nullptr;
    int64_t size = 0, used = 0;
    std::map<int64_t, int64_t> free_;  // offset -> bytes
    void init(int64_t bytes) {
        bytes = rup(bytes, 256);
        // memory freed just before (e.g. by another server) is returned by the
        // driver asynchronously: retry for up to 5 s
        for (int attempt = 0;; ++attempt) {
            if (hipMalloc(&base, bytes) == hipSuccess) break;
            base = nullptr;
            (void)hipGetLastError();
            if (attempt == 50)
                throw Error{ELPH_ENOMEM, "server pool: hipMalloc of " + std::to_string(bytes) + " bytes failed"};
            std::this_thread::sleep_for(std::chrono::milliseconds(100));
        }
        size = bytes;
        free_.clear();
        free_[0] = bytes;
    }
    int64_t alloc(int64_t bytes) {
        bytes = rup(std::max<int64_t>(bytes, 1), 256);
        for (auto it = free_.begin(); it != free_.end(); ++it)
            if (it->second >= bytes) {
                const int64_t off = it->first, rest = it->second - bytes;
                free_.erase(it);
                if (rest > 0) free_[off + bytes] = rest;
                used += bytes;
                return off;
            }
        return -1;
    }
    void release(int64_t off, int64_t bytes) {
        bytes = rup(std::max<int64_t>(bytes, 1), 256);
        used -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    ~Pool() {
        if (base) (void)hipFree(base);
    }
};

struct Material {
    elph_ctx *c = nullptr;
    // Rp cells (rcells_q(1:nwsg)) in pairs (R_p, -R_p): Kp = 2 npair cos/sin
    // rows rounded up to 64; pcell[p] = R_p
    int64_t npair = 0;
    DBuf<int2> pairs;
    DBuf<int> pcell;
    int64_t nb = 0, nbr = 0, nb2 = 0, Kp = 0, nkc = 0, CB = 1, CPT = 16, KT = 1;
    int64_t gk_k = 0;  // complex per k of g(k,Rp)
    DBuf<double> mscale, invsqrtm;
};

struct Request {
    int64_t id = 0;
    int mat = 0;
    int64_t nk = 0, nq = 0, nq64 = 0;
    std::vector<double> kv, qv;  // (n,3) column-major
    double t_submit = 0, t_done = 0;
    int status = ELPH_OK;
    bool admitted = false, setup_done = false, done = false, released = false;
    int64_t off = -1, bytes = 0;
    // device views into the pool
    double *d_kv = nullptr, *d_qv = nullptr, *d_W = nullptr, *d_om = nullptr;
    double2 *d_D = nullptr, *d_V = nullptr, *d_Lq = nullptr, *d_lp = nullptr;
    double *d_phi = nullptr;  // real cell-pair phases [K2][nq64]
    v4i *d_adig = nullptr;
    float *d_out = nullptr;
    int *d_flag = nullptr;
    hipEvent_t ev_setup = nullptr;
    int *h_flag = nullptr;  // pinned (the server's ring)
    // work cursor: q chunks outer, k inner
    int64_t cur_q0 = 0, cur_k = 0;
    int64_t outstanding = 0;
    bool issued_all = false;
};

struct Segment {
    int64_t req;
    int64_t k_unit0, nks, k_req0, q0, nqs;
};

struct Slot {
    hipStream_t st = nullptr;
    hipEvent_t ev = nullptr;
    bool busy = false;
    int mat = -1;
    std::vector<Segment> segs;
    // workspace (sized for the largest material)
    DBuf<double> kv, W, kq;
    DBuf<double2> phi, H, Uk, Ukq, gk, h, Y, phiT;
    DBuf<v4i> bphi, bsl;
    DBuf<double> bscale;
    double *h_kv = nullptr;  // pinned staging of the unit's k list
};

}  // namespace

struct elph_server {
    int device = 0;
    int32_t precision = ELPH_FP64;
    Pool pool;
    std::vector<std::unique_ptr<Material>> mats;
    std::vector<std::unique_ptr<Slot>> slots;
    hipStream_t setup_st = nullptr, copy_st = nullptr;
    int *h_flags = nullptr;        // pinned, kFlagRing entries
    std::vector<hipEvent_t> free_events;  // set-up events for reuse (created by the scheduler)
    std::unordered_map<int64_t, std::unique_ptr<Request>> reqs;
    std::deque<int64_t> pending;   // submitted, not admitted (FIFO)
    std::vector<int64_t> active;   // admitted, not done (arrival order)
    int64_t next_id = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::thread worker;
    bool stop = false, started = false;
    std::string err;
    // statistics
    int64_t n_done = 0, n_failed = 0, vertices = 0, units = 0;
    double t_first_submit = -1, t_last_done = 0;
    int64_t max_slots_busy = 0;
};

namespace {

void size_slots(elph_server *s) {
    // workspace of every slot for the largest registered material
    int64_t kv = 0, W = 0, kq = 0, phi = 0, H = 0, Uk = 0, Ukq = 0, gk = 0, hh = 0, phiT = 0, bphi = 0, bsl = 0,
            bscale = 0;
    for (auto &mp : s->mats) {
        const Material &m = *mp;
        elph_ctx *c = m.c;
        const int64_t pts = std::max<int64_t>(kElChunk, m.KT);
        kv = std::max(kv, 3 * m.KT);
        W = std::max(W, pts * m.nb);
        kq = std::max(kq, 3 * kPairMax);
        phi = std::max(phi, pts * c->nwsk);
        H = std::max(H, pts * m.nb2);
        Uk = std::max(Uk, m.KT * m.nb2);
        Ukq = std::max(Ukq, kPairMax * m.nb2);
        gk = std::max(gk, m.KT * m.gk_k);
        hh = std::max(hh, m.KT * m.nb * m.Kp * m.CPT);
        phiT = std::max(phiT, gk_phi_count(c));
        if (s->precision == ELPH_MIXED) {
            bphi = std::max(bphi, gk_bphi_count(c));
            bsl = std::max(bsl, m.KT * m.nb * m.nkc * (m.CPT / 16) * 2 * kNS * 64);
            bscale = std::max(bscale, m.KT * m.nb * m.CPT);
        }
    }
    for (auto &sp : s->slots) {
        Slot &sl = *sp;
        sl.kv.alloc(kv);
        sl.W.alloc(W);
        sl.kq.alloc(kq);
        sl.phi.alloc(phi);
        sl.H.alloc(H);
        sl.Uk.alloc(Uk);
        sl.Ukq.alloc(Ukq);
        sl.gk.alloc(gk);
        sl.h.alloc(hh);
        sl.Y.alloc(kYMax / sizeof(double2));
        sl.phiT.alloc(phiT);
        if (s->precision == ELPH_MIXED) {
            sl.bphi.alloc(bphi);
            sl.bsl.alloc(bsl);
            sl.bscale.alloc(bscale);
        }
    }
}

// Device bytes of a request's resident data in the pool.
struct ReqLayout {
    int64_t kv, qv, D, V, W, om, phi, adig, Lq, lp, out, flag, total;
};

ReqLayout req_layout(const elph_server *s, const Material &m, int64_t nk, int64_t nq) {
    ReqLayout L{};
    int64_t off = 0;
    auto take = [&](int64_t bytes) {
        const int64_t o = off;
        off += rup(std::max<int64_t>(bytes, 1), 256);
        return o;
    };
    const int64_t nq64 = rup(nq, 64), nbr2 = m.nbr * m.nbr;
    L.kv = take(nk * 3 * 8);
    L.qv = take(nq * 3 * 8);
    L.D = take(nq * nbr2 * 16);
    L.V = take(nq * nbr2 * 16);
    L.W = take(nq * m.nbr * 8);
    L.om = take(nq * m.nbr * 8);
    L.phi = take(nq64 * m.Kp * 8);
    L.adig = s->precision == ELPH_MIXED ? take(nq64 * m.Kp * kNS) : -1;
    L.Lq = m.c->polar ? take(nq * m.nbr * 16) : -1;
    L.lp = m.c->polar ? take(nq * m.nbr * 16) : -1;
    L.out = take(nk * nq * m.nb2 * m.nbr * 4);
    L.flag = take(4);
    L.total = off;
    return L;
}

// Admission (FIFO, head of line blocks): returns true when a request was admitted.
bool admit(elph_server *s) {
    if (s->pending.empty()) return false;
    const int64_t id = s->pending.front();
    Request &r = *s->reqs.at(id);
    Material &m = *s->mats[r.mat];
    elph_ctx *c = m.c;
    const ReqLayout L = req_layout(s, m, r.nk, r.nq);
    const int64_t off = s->pool.alloc(L.total);
    if (off < 0) return false;
    s->pending.pop_front();
    s->active.push_back(id);
    r.off = off;
    r.bytes = L.total;
    char *b = s->pool.base + off;
    r.d_kv = reinterpret_cast<double *>(b + L.kv);
    r.d_qv = reinterpret_cast<double *>(b + L.qv);
    r.d_D = reinterpret_cast<double2 *>(b + L.D);
    r.d_V = reinterpret_cast<double2 *>(b + L.V);
    r.d_W = reinterpret_cast<double *>(b + L.W);
    r.d_om = reinterpret_cast<double *>(b + L.om);
    r.d_phi = reinterpret_cast<double *>(b + L.phi);
    r.d_adig = L.adig >= 0 ? reinterpret_cast<v4i *>(b + L.adig) : nullptr;
    r.d_Lq = L.Lq >= 0 ? reinterpret_cast<double2 *>(b + L.Lq) : nullptr;
    r.d_lp = L.lp >= 0 ? reinterpret_cast<double2 *>(b + L.lp) : nullptr;
    r.d_out = reinterpret_cast<float *>(b + L.out);
    r.d_flag = reinterpret_cast<int *>(b + L.flag);
    r.nq64 = rup(r.nq, 64);
    r.admitted = true;
    if (!s->free_events.empty()) {
        r.ev_setup = s->free_events.back();
        s->free_events.pop_back();
    } else {
        HCHECK(hipEventCreateWithFlags(&r.ev_setup, hipEventDisableTiming));
    }
    // set-up on the set-up stream: phonons at the q-list, phases, digits
    hipStream_t st = s->setup_st;
    const int64_t nq = r.nq, nbr = m.nbr;
    HCHECK(hipMemcpyAsync(r.d_kv, r.kv.data(), r.nk * 3 * 8, hipMemcpyHostToDevice, st));
    HCHECK(hipMemcpyAsync(r.d_qv, r.qv.data(), nq * 3 * 8, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, st, r.d_flag);
    launch_dynmat(c, st, nq, r.d_qv, r.d_D);
    launch_jacobi(st, nq, (int)nbr, r.d_D, m.mscale.p, r.d_W, r.d_V);
    hipLaunchKernelGGL(k_ph_post, dim3((unsigned)((nq + 63) / 64)), dim3(64), 0, st, (int)nq, (int)nbr, r.d_qv, (int)nq,
                       r.d_W, r.d_V, m.invsqrtm.p, r.d_om, r.d_flag);
    if (c->polar) {
        polar_launch(c, st, nq, r.d_qv, nullptr, nullptr, r.d_Lq);
        launch_polar_lp(st, (int)(nq * nbr), (int)nbr, (int)nbr, nullptr, r.d_Lq, r.d_V, r.d_lp);
    }
    const int64_t np = r.nq64 * m.Kp;
    hipLaunchKernelGGL(k_srv_phase, dim3((unsigned)((np / 2 + 255) / 256)), dim3(256), 0, st, (int)nq, (int)r.nq64,
                       (int)m.Kp, (int)m.npair, r.d_qv, (int)nq, m.pcell.p, r.d_phi);
    if (s->precision == ELPH_MIXED) {
        const int64_t nt = (r.nq64 / 16) * m.nkc * 64;
        hipLaunchKernelGGL(k_srv_phase_digits, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, st, (int)nq,
                           (int)(r.nq64 / 16), (int)m.nkc, (int)r.nq64, r.d_phi, r.d_adig);
    }
    HCHECK(hipMemcpyAsync(r.h_flag, r.d_flag, sizeof(int), hipMemcpyDeviceToHost, st));
    HCHECK(hipGetLastError());
    HCHECK(hipEventRecord(r.ev_setup, st));
    return true;
}

// Next unit: the oldest active request with work left fixes the material;
// k-points are taken from it and then from later requests of that material.
bool form_unit(elph_server *s, int &mat, std::vector<Segment> &segs) {
    segs.clear();
    mat = -1;
    int64_t nk_unit = 0;
    for (int64_t id : s->active) {
        Request &r = *s->reqs.at(id);
        if (r.issued_all || r.done) continue;
        if (!r.setup_done) {
            if (hipEventQuery(r.ev_setup) != hipSuccess) continue;
            r.setup_done = true;
            if (*r.h_flag) {
                r.status = ELPH_ENEGPHONON;
                r.issued_all = true;
                continue;
            }
        }
        if (mat < 0) mat = r.mat;
        if (r.mat != mat) continue;
        const Material &m = *s->mats[mat];
        while (!r.issued_all && nk_unit < m.KT) {
            const int64_t nqs = std::min(kQMax, r.nq - r.cur_q0);
            int64_t pairs_used = 0;
            for (auto &sg : segs) pairs_used += sg.nks * sg.nqs;
            const int64_t kmax = std::min<int64_t>(m.KT - nk_unit, (kPairMax - pairs_used) / nqs);
            if (kmax <= 0) break;
            const int64_t nks = std::min(kmax, r.nk - r.cur_k);
            segs.push_back({id, nk_unit, nks, r.cur_k, r.cur_q0, nqs});
            nk_unit += nks;
            r.outstanding += 1;
            r.cur_k += nks;
            if (r.cur_k == r.nk) {
                r.cur_k = 0;
                r.cur_q0 += nqs;
                if (r.cur_q0 >= r.nq) r.issued_all = true;
            }
        }
        if (nk_unit >= m.KT) break;
    }
    return !segs.empty();
}

template <int CB>
void launch_y_fp64(hipStream_t st, const Material &m, const double2 *h, int st0, int nst, const double *phi, int ldp,
                   int q0, int nqs, double2 *Y) {
    dim3 grid((unsigned)((nqs + 127) / 128), (unsigned)nst, (unsigned)(m.CPT / (16 * CB)));
    const size_t lds = (size_t)16 * 128 * sizeof(double) + (size_t)16 * 16 * CB * sizeof(double2);
    hipLaunchKernelGGL(k_srv_y_fp64<CB>, grid, dim3(256), lds, st, h, (int)m.Kp, (int)m.CPT, st0, phi, ldp, q0, nqs,
                       Y);
}

template <int NB>
void launch_rotate_all(hipStream_t st, const Material &m, int nk, const double2 *gk, const double2 *Uk, double2 *h) {
    const int64_t per = (m.Kp / 2) * m.CPT;
    dim3 grid((unsigned)((per + 255) / 256), (unsigned)nk);
    hipLaunchKernelGGL(k_rotate_all<NB>, grid, dim3(256), 0, st, (int)m.nb, (int)m.nbr, (int)m.npair, (int)m.Kp,
                       (int)m.CPT, gk, m.gk_k, Uk, m.c->degq.p, m.pairs.p, h);
}

template <int CB>
void launch_y_i8(hipStream_t st, const Material &m, const v4i *bsl, int st0, int nst, const double *bscale,
                 const v4i *adig, int qb0, int nqs, double2 *Y) {
    dim3 grid((unsigned)((nqs + 63) / 64), (unsigned)nst, (unsigned)(m.CPT / (16 * CB)));
    const size_t lds = (size_t)2 * CB * 2 * kNS * 64 * sizeof(v4i);
    hipLaunchKernelGGL(k_srv_y_i8<CB>, grid, dim3(256), lds, st, bsl, (int)m.nkc, (int)(m.CPT / 16), st0, bscale,
                       (int)m.CPT, adig, qb0, nqs, Y);
}

void launch_unit(elph_server *s, Slot &sl, int mat, std::vector<Segment> &segs) {
    const Material &m = *s->mats[mat];
    elph_ctx *c = m.c;
    hipStream_t st = sl.st;
    const bool mixed = s->precision == ELPH_MIXED;
    const int64_t nb = m.nb, nbr = m.nbr, nb2 = m.nb2;
    int64_t nk = 0;
    for (auto &sg : segs) nk = std::max(nk, sg.k_unit0 + sg.nks);
    // the unit's k list, (nk,3) column-major, staged through pinned memory
    for (auto &sg : segs) {
        const Request &r = *s->reqs.at(sg.req);
        for (int64_t i = 0; i < sg.nks; ++i)
            for (int d = 0; d < 3; ++d) sl.h_kv[sg.k_unit0 + i + nk * d] = r.kv[sg.k_req0 + i + r.nk * d];
    }
    HCHECK(hipMemcpyAsync(sl.kv.p, sl.h_kv, nk * 3 * 8, hipMemcpyHostToDevice, st));
    // U_k, g(k,Rp), h = U_k g / phwsdeg
    launch_build_H(c, st, nk, sl.kv.p, nk, sl.phi.p, sl.H.p);
    launch_jacobi(st, nk, (int)nb, sl.H.p, nullptr, sl.W.p, sl.Uk.p);
    launch_gkRp(c, st, nk, sl.kv.p, sl.gk.p, m.gk_k, mixed, sl.phiT.p, sl.bphi.p);
    const int nst = (int)(nk * nb);
    if (nb <= 8) launch_rotate_all<8>(st, m, (int)nk, sl.gk.p, sl.Uk.p, sl.h.p);
    else if (nb <= 16) launch_rotate_all<16>(st, m, (int)nk, sl.gk.p, sl.Uk.p, sl.h.p);
    else if (nb <= 24) launch_rotate_all<24>(st, m, (int)nk, sl.gk.p, sl.Uk.p, sl.h.p);
    else launch_rotate_all<32>(st, m, (int)nk, sl.gk.p, sl.Uk.p, sl.h.p);
    if (mixed) {
        launch_colscale(st, nst, (int)m.Kp, (int)m.CPT, sl.h.p, sl.bscale.p);
        launch_slice_b_ri(st, nst, (int)m.nkc, (int)(m.CPT / 16), (int)m.Kp, (int)m.CPT, sl.h.p, sl.bscale.p,
                          sl.bsl.p);
    }
    for (auto &sg : segs) {
        const Request &r = *s->reqs.at(sg.req);
        const int64_t npairs = sg.nks * sg.nqs;
        hipLaunchKernelGGL(k_srv_kq, dim3((unsigned)((npairs + 255) / 256)), dim3(256), 0, st, (int)sg.nks, (int)sg.nqs,
                           sl.kv.p, (int)nk, (int)sg.k_unit0, r.d_qv, (int)r.nq, (int)sg.q0, sl.kq.p, (int)npairs);
        for (int64_t p0 = 0; p0 < npairs; p0 += kElChunk) {
            const int64_t n = std::min(kElChunk, npairs - p0);
            launch_build_H(c, st, n, sl.kq.p + p0, npairs, sl.phi.p, sl.H.p);
            launch_jacobi(st, n, (int)nb, sl.H.p, nullptr, sl.W.p, sl.Ukq.p + p0 * nb2);
        }
        // states of the segment: (k_unit0 .. k_unit0 + nks) x m, in chunks of Y
        const int64_t st_seg0 = sg.k_unit0 * nb, nst_seg = sg.nks * nb;
        const int64_t per_state = sg.nqs * m.CPT * (int64_t)sizeof(double2);
        const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(nst_seg, kYMax / per_state));
        for (int64_t s0 = 0; s0 < nst_seg; s0 += chunk) {
            const int ns = (int)std::min(chunk, nst_seg - s0);
            if (mixed) {
                const int qb0 = (int)(sg.q0 / 16);
                switch (m.CB) {
                    case 1: launch_y_i8<1>(st, m, sl.bsl.p, (int)(st_seg0 + s0), ns, sl.bscale.p, r.d_adig, qb0, (int)sg.nqs, sl.Y.p); break;
                    case 2: launch_y_i8<2>(st, m, sl.bsl.p, (int)(st_seg0 + s0), ns, sl.bscale.p, r.d_adig, qb0, (int)sg.nqs, sl.Y.p); break;
                    default: launch_y_i8<3>(st, m, sl.bsl.p, (int)(st_seg0 + s0), ns, sl.bscale.p, r.d_adig, qb0, (int)sg.nqs, sl.Y.p); break;
                }
            } else {
                switch (m.CB) {
                    case 1: launch_y_fp64<1>(st, m, sl.h.p, (int)(st_seg0 + s0), ns, r.d_phi, (int)r.nq64, (int)sg.q0, (int)sg.nqs, sl.Y.p); break;
                    case 2: launch_y_fp64<2>(st, m, sl.h.p, (int)(st_seg0 + s0), ns, r.d_phi, (int)r.nq64, (int)sg.q0, (int)sg.nqs, sl.Y.p); break;
                    case 3: launch_y_fp64<3>(st, m, sl.h.p, (int)(st_seg0 + s0), ns, r.d_phi, (int)r.nq64, (int)sg.q0, (int)sg.nqs, sl.Y.p); break;
                    default: launch_y_fp64<4>(st, m, sl.h.p, (int)(st_seg0 + s0), ns, r.d_phi, (int)r.nq64, (int)sg.q0, (int)sg.nqs, sl.Y.p); break;
                }
            }
            EpiArgs a{};
            a.Y = sl.Y.p;
            a.CPT = (int)m.CPT;
            a.nqs = (int)sg.nqs;
            a.st0 = (int)s0;
            a.ns = ns;
            a.nb = (int)nb;
            a.nbr = (int)nbr;
            a.k_unit0 = (int)sg.k_unit0;
            a.k_req0 = (int)sg.k_req0;
            a.q0 = (int)sg.q0;
            a.nq = (int)r.nq;
            a.Ukq = sl.Ukq.p;
            a.Uk = sl.Uk.p;
            a.u = r.d_V;
            a.om = r.d_om;
            a.lp = r.d_lp;
            a.out = r.d_out;
            a.G = (int)std::max<int64_t>(1, std::min<int64_t>(8, 256 / (nb * nbr)));
            const size_t lds = (size_t)a.G * (3 * nb * nbr + nb * nb + nbr * nbr) * sizeof(double2);
            dim3 grid((unsigned)((s0 + ns - 1) / nb - s0 / nb + 1), (unsigned)((sg.nqs + a.G - 1) / a.G));
            hipLaunchKernelGGL(k_srv_epi, grid, dim3(256), lds, st, a);
        }
        HCHECK(hipGetLastError());
    }
    HCHECK(hipEventRecord(sl.ev, st));
    sl.busy = true;
    sl.mat = mat;
    sl.segs = segs;
}

// (all scheduler helpers run with s->mu held)
void finish_request(elph_server *s, Request &r) {
    r.done = true;
    r.t_done = now_ms();
    s->active.erase(std::find(s->active.begin(), s->active.end(), r.id));
    if (r.status == ELPH_OK) {
        s->n_done += 1;
        s->vertices += r.nk * r.nq * s->mats[r.mat]->nb2 * s->mats[r.mat]->nbr;
    } else {
        s->n_failed += 1;
    }
    s->t_last_done = std::max(s->t_last_done, r.t_done);
    s->cv_done.notify_all();
}

void finish_slot(elph_server *s, Slot &sl) {
    sl.busy = false;
    s->units += 1;
    for (auto &sg : sl.segs) {
        Request &r = *s->reqs.at(sg.req);
        r.outstanding -= 1;
        if (r.issued_all && r.outstanding == 0 && !r.done) finish_request(s, r);
    }
    sl.segs.clear();
}

void run_loop(elph_server *s) {
    std::unique_lock<std::mutex> lk(s->mu);
    try {
        HCHECK(hipSetDevice(s->device));
        for (;;) {
            bool progressed = false;
            int busy = 0;
            for (auto &sp : s->slots)
                if (sp->busy) {
                    if (hipEventQuery(sp->ev) == hipSuccess) {
                        finish_slot(s, *sp);
                        progressed = true;
                    } else {
                        busy += 1;
                    }
                }
            while (admit(s)) progressed = true;
            // requests failed at set-up with nothing in flight
            for (int64_t id : std::vector<int64_t>(s->active)) {
                Request &r = *s->reqs.at(id);
                if (!r.done && r.setup_done && r.issued_all && r.outstanding == 0) finish_request(s, r);
            }
            for (auto &sp : s->slots) {
                if (sp->busy) continue;
                int mat;
                std::vector<Segment> segs;
                if (!form_unit(s, mat, segs)) break;
                launch_unit(s, *sp, mat, segs);
                busy += 1;
                progressed = true;
            }
            s->max_slots_busy = std::max<int64_t>(s->max_slots_busy, busy);
            if (s->stop) {
                // requests not admitted yet are refused; admitted ones finish
                for (int64_t id : s->pending) {
                    Request &r = *s->reqs.at(id);
                    r.status = ELPH_EINVAL;
                    r.done = true;
                    r.t_done = now_ms();
                }
                s->pending.clear();
                s->cv_done.notify_all();
                if (s->active.empty() && busy == 0) break;
            }
            if (!progressed) s->cv_work.wait_for(lk, std::chrono::microseconds(busy ? 20 : 500));
        }
    } catch (const Error &e) {
        s->err = e.msg;
        for (auto &kv : s->reqs)
            if (!kv.second->done) {
                kv.second->status = e.code;
                kv.second->done = true;
                kv.second->t_done = now_ms();
            }
        s->active.clear();
        s->pending.clear();
        s->cv_done.notify_all();
    }
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int elph_server_create(elph_server **out, int device, int32_t precision, int64_t pool_bytes, int32_t nslots) {
    if (!out || nslots < 1 || nslots > 16 || pool_bytes <= 0) return ELPH_EINVAL;
    if (precision != ELPH_FP64 && precision != ELPH_MIXED) return ELPH_EINVAL;
    *out = nullptr;
    auto s = std::make_unique<elph_server>();
    try {
        HCHECK(hipSetDevice(device));
        s->device = device;
        s->precision = precision;
        s->pool.init(pool_bytes);
        HCHECK(hipStreamCreateWithFlags(&s->setup_st, hipStreamNonBlocking));
        HCHECK(hipStreamCreateWithFlags(&s->copy_st, hipStreamNonBlocking));
        for (int i = 0; i < nslots; ++i) {
            auto sl = std::make_unique<Slot>();
            HCHECK(hipStreamCreateWithFlags(&sl->st, hipStreamNonBlocking));
            HCHECK(hipEventCreateWithFlags(&sl->ev, hipEventDisableTiming));
            HCHECK(hipHostMalloc((void **)&sl->h_kv, 3 * kKT * sizeof(double), hipHostMallocDefault));
            s->slots.push_back(std::move(sl));
        }
        HCHECK(hipHostMalloc((void **)&s->h_flags, kFlagRing * sizeof(int), hipHostMallocDefault));
    } catch (const Error &e) {
        return e.code;
    }
    *out = s.release();
    return ELPH_OK;
}

int elph_server_add_material(elph_server *s, elph_ctx *ctx) {
    if (!s || !ctx) return ELPH_EINVAL;
    if (s->started) {
        s->err = "materials must be added before the first request";
        return ELPH_EINVAL;
    }
    try {
        if (ctx->device != s->device) throw Error{ELPH_EINVAL, "context and server are on different devices"};
        if (ctx->nwsg > ctx->nwsq) throw Error{ELPH_EINVAL, "g2 'ph' sums nwsg cells of rcells_q: needs nwsg <= nwsq"};
        if (ctx->nb > 32 || ctx->nbr > 32) throw Error{ELPH_EINVAL, "numwannbands and numbranches must be <= 32"};
        HCHECK(hipSetDevice(s->device));
        auto m = std::make_unique<Material>();
        m->c = ctx;
        m->nb = ctx->nb;
        m->nbr = ctx->nbr;
        m->nb2 = ctx->nb * ctx->nb;
        {
            const std::vector<int2> pr = cell_pairs(ctx->rcells_q.data(), ctx->nwsq, ctx->nwsg);
            m->npair = (int64_t)pr.size();
            std::vector<int> pc(3 * m->npair);
            for (int64_t q = 0; q < m->npair; ++q)
                for (int d = 0; d < 3; ++d) pc[3 * q + d] = (int)ctx->rcells_q[pr[q].x + ctx->nwsq * d];
            m->pairs.upload(pr, ctx->stream);
            m->pcell.upload(pc, ctx->stream);
        }
        m->Kp = rup(2 * m->npair, 64);
        m->nkc = m->Kp / 64;
        const int64_t cp16 = rup(m->nb * m->nbr, 16) / 16;
        m->CB = std::min<int64_t>(3, cp16);
        m->CPT = rup(m->nb * m->nbr, 16 * m->CB);
        m->gk_k = m->nb2 * m->nbr * ctx->nwsg;
        const int64_t per_k = std::max(m->gk_k, m->nb * m->Kp * m->CPT) * (int64_t)sizeof(double2);
        m->KT = std::max<int64_t>(1, std::min<int64_t>(kKT, kGkMax / per_k));
        m->mscale.upload(mass_scale(ctx), ctx->stream);
        std::vector<double> ism(m->nbr);
        for (int64_t p = 0; p < m->nbr; ++p) ism[p] = 1.0 / std::sqrt(ctx->masses[p / 3]);
        m->invsqrtm.upload(ism, ctx->stream);
        if (s->precision == ELPH_MIXED) ensure_gwann_digits(ctx);
        HCHECK(hipStreamSynchronize(ctx->stream));
        s->mats.push_back(std::move(m));
        size_slots(s);
        return (int)s->mats.size() - 1;
    } catch (const Error &e) {
        s->err = e.msg;
        return -e.code;
    }
}

int64_t elph_server_submit(elph_server *s, int32_t material, int64_t nk, const double *kvecs, int64_t nq,
                           const double *qvecs) {
    if (!s || nk < 1 || nq < 1 || !kvecs || !qvecs) return -ELPH_EINVAL;
    if (material < 0 || material >= (int32_t)s->mats.size()) return -ELPH_EINVAL;
    try {
        const Material &m = *s->mats[material];
        if (req_layout(s, m, nk, nq).total > s->pool.size) {
            s->err = "request needs more device memory than the server pool holds";
            return -ELPH_ENOMEM;
        }
        if (nk * nq * m.nb2 * m.nbr > ((int64_t)1 << 40)) return -ELPH_EINVAL;
        auto r = std::make_unique<Request>();
        r->mat = material;
        r->nk = nk;
        r->nq = nq;
        r->kv.assign(kvecs, kvecs + 3 * nk);
        r->qv.assign(qvecs, qvecs + 3 * nq);
        // no HIP call on the caller's thread: device objects of a request are
        // made by the scheduler thread
        std::lock_guard<std::mutex> lk(s->mu);
        r->id = s->next_id++;
        r->h_flag = s->h_flags + (r->id % kFlagRing);
        *r->h_flag = 0;
        r->t_submit = now_ms();
        if (s->t_first_submit < 0) s->t_first_submit = r->t_submit;
        const int64_t id = r->id;
        s->reqs[id] = std::move(r);
        s->pending.push_back(id);
        if (!s->started) {
            s->started = true;
            s->worker = std::thread(run_loop, s);
        }
        s->cv_work.notify_one();
        return id;
    } catch (const Error &e) {
        s->err = e.msg;
        return -e.code;
    }
}

int elph_server_wait(elph_server *s, int64_t id, double *latency_ms) {
    if (!s) return ELPH_EINVAL;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return ELPH_EINVAL;
    Request &r = *it->second;
    s->cv_done.wait(lk, [&] { return r.done; });
    if (latency_ms) *latency_ms = r.t_done - r.t_submit;
    return r.status;
}

int elph_server_test(elph_server *s, int64_t id) {
    if (!s) return -ELPH_EINVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return -ELPH_EINVAL;
    return it->second->done ? 1 : 0;
}

int elph_server_fetch(elph_server *s, int64_t id, float *out) {
    if (!s || !out) return ELPH_EINVAL;
    const float *src = nullptr;
    int64_t n = 0;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->reqs.find(id);
        if (it == s->reqs.end() || !it->second->done || it->second->released) return ELPH_EINVAL;
        Request &r = *it->second;
        if (r.status != ELPH_OK) return r.status;
        src = r.d_out;
        const Material &m = *s->mats[r.mat];
        n = r.nk * r.nq * m.nb2 * m.nbr;
    }
    try {
        HCHECK(hipSetDevice(s->device));
        HCHECK(hipMemcpyAsync(out, src, n * sizeof(float), hipMemcpyDeviceToHost, s->copy_st));
        HCHECK(hipStreamSynchronize(s->copy_st));
    } catch (const Error &e) {
        s->err = e.msg;
        return e.code;
    }
    return ELPH_OK;
}

int elph_server_result_device(elph_server *s, int64_t id, const float **dptr, int64_t *count) {
    if (!s || !dptr) return ELPH_EINVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end() || !it->second->done || it->second->released) return ELPH_EINVAL;
    Request &r = *it->second;
    if (r.status != ELPH_OK) return r.status;
    *dptr = r.d_out;
    if (count) *count = r.nk * r.nq * s->mats[r.mat]->nb2 * s->mats[r.mat]->nbr;
    return ELPH_OK;
}

int elph_server_release(elph_server *s, int64_t id) {
    if (!s) return ELPH_EINVAL;
    std::unique_ptr<Request> r;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->reqs.find(id);
        if (it == s->reqs.end() || !it->second->done) return ELPH_EINVAL;
        r = std::move(it->second);
        s->reqs.erase(it);
        if (r->off >= 0) s->pool.release(r->off, r->bytes);
        if (r->ev_setup) s->free_events.push_back(r->ev_setup);
        s->cv_work.notify_one();
    }
    return ELPH_OK;
}

int elph_server_stats(elph_server *s, elph_server_stats_t *st) {
    if (!s || !st) return ELPH_EINVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    st->completed = s->n_done;
    st->failed = s->n_failed;
    st->pending = (int64_t)s->pending.size();
    st->active = (int64_t)s->active.size();
    st->vertices = s->vertices;
    st->units = s->units;
    st->pool_bytes = s->pool.size;
    st->pool_used = s->pool.used;
    st->max_slots_busy = s->max_slots_busy;
    st->busy_span_ms = s->t_first_submit < 0 ? 0.0 : s->t_last_done - s->t_first_submit;
    return ELPH_OK;
}

const char *elph_server_error(const elph_server *s) { return s ? s->err.c_str() : "null server"; }

int elph_server_destroy(elph_server *s) {
    if (!s) return ELPH_EINVAL;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv_work.notify_all();
    }
    if (s->worker.joinable()) s->worker.join();
    (void)hipSetDevice(s->device);
    for (auto &sp : s->slots) {
        if (sp->st) (void)hipStreamSynchronize(sp->st);
    }
    if (s->setup_st) (void)hipStreamSynchronize(s->setup_st);
    for (auto &kv : s->reqs)
        if (kv.second->ev_setup) (void)hipEventDestroy(kv.second->ev_setup);
    for (auto e : s->free_events) (void)hipEventDestroy(e);
    if (s->h_flags) (void)hipHostFree(s->h_flags);
    for (auto &sp : s->slots) {
        if (sp->h_kv) (void)hipHostFree(sp->h_kv);
        if (sp->ev) (void)hipEventDestroy(sp->ev);
        if (sp->st) (void)hipStreamDestroy(sp->st);
    }
    if (s->setup_st) (void)hipStreamDestroy(s->setup_st);
    if (s->copy_st) (void)hipStreamDestroy(s->copy_st);
    hipError_t fe = hipSuccess;
    if (s->pool.base) {
        fe = hipFree(s->pool.base);
        s->pool.base = nullptr;
    }
    delete s;
    if (fe != hipSuccess) {
        std::fprintf(stderr, "elph_server_destroy: hipFree of the pool: %s\n", hipGetErrorString(fe));
        return ELPH_EDEVICE;
    }
    return ELPH_OK;
}

}  // extern "C"
