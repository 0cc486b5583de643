// Internal declarations shared by the library's translation units: device
// buffers, the context (HBM-resident Wannier tensors) and stream-parameterised
// launchers of kernels that more than one driver uses.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "common.h"
#include "kernels.h"

#define HCHECK(call)                                                                    \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            throw Error{ELPH_EDEVICE, std::string(#call) + ": " + hipGetErrorString(e_)}; \
        }                                                                               \
    } while (0)

namespace elph {

constexpr double kTwoPi = 2.0 * 3.14159265358979323846;

// ---------------------------------------------------------------------------
// Device buffers
// ---------------------------------------------------------------------------
template <typename T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    bool owned = true;  // false: a view of memory owned elsewhere (borrow)
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    ~DBuf() { release(); }
    void release() {
        if (p && owned) (void)hipFree(p);
        p = nullptr;
        n = 0;
        owned = true;
    }
    // Uses count elements at ptr (not freed by this buffer); a later alloc or
    // upload of at most count elements writes there without allocating.
    void borrow(T *ptr, size_t count) {
        release();
        p = ptr;
        n = count;
        owned = false;
    }
    void alloc(size_t count) {
        if (count <= n && p) return;
        if (!owned) throw Error{ELPH_ENOMEM, "a borrowed device buffer is too small"};
        release();
        if (count == 0) count = 1;
        hipError_t e = hipMalloc(&p, count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            throw Error{ELPH_ENOMEM, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed"};
        }
        n = count;
    }
    void upload(const T *h, size_t count, hipStream_t s) {
        alloc(count);
        HCHECK(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void upload(const std::vector<T> &v, hipStream_t s) { upload(v.data(), v.size(), s); }
};

}  // namespace elph

using elph::DBuf;
using elph::v4i;

struct elph_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev_a = nullptr, ev_b = nullptr, ev_c = nullptr, ev_d = nullptr;
    // event pairs bracketing the dominant kernel launches of a run (created
    // with the context; read when full and at the end of the run, so the
    // host does not wait after each launch)
    std::vector<hipEvent_t> tev_start, tev_stop;
    int tev_used = 0;
    int64_t nb = 0, nbr = 0, nwsk = 0, nwsq = 0, nwsg = 0, natoms = 0;
    int64_t coarse_qmesh[3] = {0, 0, 0};
    double lattvecs[9], reclattvecs[9], volume = 0;
    int polar = 0;
    double epsilon[9];
    std::vector<double> basis, masses, born;
    std::vector<int64_t> rcells_k, rcells_q, rcells_g, elwsdeg, phwsdeg, gwsdeg;
    // HBM-resident tensors
    DBuf<double2> gwann;   // (nb*nb, nwsk, nbr*nwsg)
    DBuf<double2> HwT;     // [Re][i + nb j]
    DBuf<double2> DphT;    // [iuc][i + nbr j]
    DBuf<int> rk, rq, rg;  // [cell][3]
    DBuf<int> degk, degq, degg;
    DBuf<double> rcart_k, rcart_q;  // [cell][3] Cartesian (nm)
    // int8 digits of gwann for the mixed-precision Re contraction (built on
    // first use): [slice][row block][K chunk][re/im][digit][lane]
    DBuf<v4i> gdig;
    DBuf<double> gscale;
    int64_t gnrb = 0, gnkc = 0;
    // Re cells in pairs (R_p, -R_p) for the cell-pair form of the Re
    // contraction: gpair[p] = (index of R_p, index of -R_p or -1), gpcell[p] =
    // R_p, ginvdeg[cell] = 1 / elwsdeg; K2 = 2 npairk cos/sin rows, padded to
    // gk2p (a multiple of 64)
    DBuf<int2> gpair;
    DBuf<int> gpcell;
    DBuf<double> ginvdeg;
    int64_t npairk = 0, gk2p = 0;
    // the same over the Rp cells of the g set (rcells_g) for the cell-pair
    // form of gReq: rgpair, rgpcell, rgpwt = 1 / gwsdeg of R_p and of -R_p
    // (0 when missing) per pair
    DBuf<int2> rgpair;
    DBuf<int> rgpcell;
    DBuf<double2> rgpwt;
    int64_t npairg = 0;
    // D(q) interpolation terms (wrapped cells of src/wannier.f90:631-650)
    DBuf<int> term_off, term_iuc, term_T, term_div;
    // polar crystals: reciprocal vectors of the (2N+1)^3 box (1/bohr), Born
    // charges, atom positions (bohr) and the q-independent part of D_L
    DBuf<double> pol_G, pol_Z, pol_tau;
    DBuf<double2> pol_D0;
    int64_t pol_nG = 0;
    double pol_alpha = 0, pol_fac = 0;
    std::string err;
    // device address of the library's file-scope scratch array on this device
    // (the allocation-free entry points carve their buffers from it)
    unsigned char *noalloc_scratch = nullptr;
    int64_t lds_max = 0;  // dynamic LDS a workgroup may allocate on this device (bytes)
    int64_t ncu = 0;      // compute units of the device (persistent launches)
    double t_dominant_ms = 0, t_all_ms = 0;
    int64_t n_dominant = 0;
};


namespace elph {

int ret_error(elph_ctx *ctx, const Error &e);
// Throws ELPH_EINVAL when a kernel's dynamic LDS (`bytes`) exceeds what a
// workgroup may allocate on the context's device; `what` names the kernel and
// the table that sets the size.
void require_lds(const elph_ctx *c, size_t bytes, const char *what);
int64_t lcm64(int64_t a, int64_t b);
// Pairs (i, j) of opposite cells R_j = -R_i among the first n columns of the
// 3 x ld column-major cell list, in order of first appearance; j = -1 when -R_i
// is not in the list (R_i = 0).  Every cell is in exactly one pair.
std::vector<int2> cell_pairs(const int64_t *cells, int64_t ld, int64_t n);

// Launchers (all enqueue on `st`; none synchronises or allocates).
// gkRp for nk device k vectors ((nk,3) column-major) into out[k] (nb*nb, nbr*nwsg);
// phiT: gk_phi_count complex scratch; bphi (mixed only): gk_bphi_count v4i scratch.
void launch_gkRp(elph_ctx *c, hipStream_t st, int64_t nk, const double *d_kvecs, double2 *out, int64_t out_kstride,
                 bool mixed, double2 *phiT, v4i *bphi);
void ensure_gwann_digits(elph_ctx *c);  // int8 digits of gwann, once per context; returns when they are built
int64_t gk_phi_count(const elph_ctx *c);
int64_t gk_bphi_count(const elph_ctx *c);
// h(Rp) = sum_mp U(kb,m,mp) gk[kb](np,mp,sp,Rp) / phwsdeg(Rp), folded over the
// cell pairs (R_p, -R_p): h[st][2p][CP] = h(R_p) + h(-R_p), h[st][2p+1][CP] =
// i (h(R_p) - h(-R_p)); pairs[p] = (index of R_p, index of -R_p or -1), Kp even.
void launch_rotate_h(hipStream_t st, int nst, int nb, int nbr, int npair, int Kp, int CP, const double2 *gk,
                     int64_t gk_kstride, const int *st_kb, const int *st_m, const double2 *Uk, const int *degq,
                     const int2 *pairs, double2 *h);
void launch_colscale(hipStream_t st, int nst, int Kp, int CP, const double2 *h, double *scale);
// int8 B digits [st][kc][cb][Re, Im, -Im][digit][lane] (complex phase operand)
void launch_slice_b(hipStream_t st, int nst, int nkc, int CB, int Kp, int CP, const double2 *h, const double *scale,
                    v4i *bsl);
// int8 B digits [st][kc][cb][Re, Im][digit][lane] (real phase operand, cell pairs)
void launch_slice_b_ri(hipStream_t st, int nst, int nkc, int CB, int Kp, int CP, const double2 *h,
                       const double *scale, v4i *bsl);
// Batched zheev('V','U') (optionally on A .* scale): W ascending, V[mat][col j][row]
void launch_jacobi(hipStream_t st, int64_t nmat, int n, const double2 *A, const double *scale, double *W, double2 *V);
// Short-range D(q) (plus the polar D_L for polar contexts) at nq device q vectors
void launch_dynmat(elph_ctx *c, hipStream_t st, int64_t nq, const double *d_qvecs, double2 *D);
// Mass normalisation factors Ryd2amu / sqrt(M_i M_j) of the context, (nbr, nbr)
std::vector<double> mass_scale(const elph_ctx *c);
// Polar terms (see k_polar): D_L into D, dD_L into dD, long-range vector into Lq
void polar_launch(elph_ctx *c, hipStream_t st, int64_t nq, const double *d_qvecs, double2 *D, double2 *dD,
                  double2 *Lq);
// H(k) at n device k vectors (k_d of point v at d_kvecs[v + vstride d]) into
// H[k][i + nb j]; phi: n * nwsk complex scratch
void launch_build_H(elph_ctx *c, hipStream_t st, int64_t n, const double *d_kvecs, int64_t vstride, double2 *phi,
                    double2 *H);
// lp[i] = sum_p Lq[r][p] u[i][p], r = ridx[i] (or i / div)
void launch_polar_lp(hipStream_t st, int n, int nbr, int div, const int *ridx, const double2 *Lq, const double2 *u,
                     double2 *lp);

// Scratch of the consumers of an ibzk plan's g2 rows (transition
// probabilities and rates), owned by the plan.
struct ConsumerScratch {
    DBuf<int2> vmap;     // [nprocs] (k' row r, band n << 8 | branch s) of each process of a row
    DBuf<int4> st_k;     // per in-window state: k (electron mesh ints)
    DBuf<double> st_e;   // per in-window state: electron energy
    DBuf<int> st_row;    // per in-window state: its row among the call's states
    DBuf<int> st_idx;    // per state of the call: index among the in-window ones, -1 outside
    DBuf<double> partial;
    DBuf<double> bose;   // Bose factors of the phonon modes
};

// The k'-row tables of an ibzk plan, as its consumers read them.  A g2 row is
// ordered k' ascending (row r), then in-window band rank, then branch s: the
// process of rank j at row r is at offs[r] + j nbr + s, band wband[r nb + j].
struct IbzkRows {
    elph_ctx *ctx;
    int nb, nbr, nwv, mesh_ref;
    int Nel[3], Nph[3];
    int64_t nirr, nprocs;
    const int *kp_int, *wband;
    const int64_t *offs;
    const double *om, *omf;      // phonon energies [nq][s] on the phonon mesh / every electron-mesh q
    const int *kint_irr;         // host [nirr][3]
    const double *ens_irred;     // host (nirr, nb)
    double enref, fsthick;
    ConsumerScratch *scr;
};
// Builds the vmap table on first use (synchronous).
IbzkRows ibzk_rows(elph_ibzk_plan *p);

// Timing of dominant launches: timed_begin records the start event of the
// next pair on `st`, timed_end its stop event; timing_flush waits for the
// recorded pairs and adds their times to t_dominant_ms / n_dominant.
void timed_begin(elph_ctx *c, hipStream_t st);
void timed_end(elph_ctx *c, hipStream_t st);
void timing_flush(elph_ctx *c);

}  // namespace elph
