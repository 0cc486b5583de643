/* CPU oracle (test infrastructure only): a literal FP64 restatement of the
 * reference's electron-phonon vertex path.  Every function cites the
 * reference lines it follows; loop orders and floating-point expressions are
 * kept as written there.  Only tests/, smoke() and the CPU-baseline leg of
 * bench.py load this library, and only as the checker.
 *
 * Pins (see tests/): the reference's own unit-test vectors for the index
 * algebra (test/test_misc.f90, test/test_vector.f90), the reference's
 * published symmetry counts (48 ops / 413 IBZ points on 24^3 for Si,
 * examples/Si/output_gcc.txt), scipy's LAPACK zheev for the eigensolver,
 * closed forms and numpy brute force for g2 / gkRp / gReq.
 */
#include "oracle.h"

#include <math.h>
#include <quadmath.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static const double QE = 1.602176634e-19;
static const double ME = 9.1093837015e-31;
static const double AMU = 1.66053906660e-27;
static const double HBAR = 1.05457172647e-22;
#define PI_ (4.0 * atan(1.0))
#define TWOPI_ (2.0 * PI_)
#define RYD2EV (0.5 * 27.2116)
#define RYD2AMU (2.0 * ME / AMU)
#define BOHR2NM 0.052917721092
#define RYD2RADTHZ (0.5 * 27.2116 * QE / HBAR)

/* column-major helpers */
#define IX2(i, j, n1) ((i) + (n1) * (j))
#define IX3(i, j, k, n1, n2) ((i) + (n1) * ((j) + (n2) * (k)))

static int64_t fnint(double x) { return (int64_t)llround(x); } /* Fortran nint */
static int64_t fmodulo(int64_t a, int64_t p) { int64_t r = a % p; return (r != 0 && ((r < 0) != (p < 0))) ? r + p : r; }

/* misc.f90:588-594 expi */
void orc_expi(double x, double out[2]) { out[0] = cos(x); out[1] = sin(x); }
static zdouble cexpi(double x) { return cos(x) + I * sin(x); }

/* misc.f90:596-610 add_and_fold (.umklapp.) */
void orc_add_and_fold(const double q1[3], const double q2[3], double q3[3]) {
    for (int c = 0; c < 3; ++c) {
        q3[c] = q1[c] + q2[c];
        if (q3[c] >= 1.0) q3[c] = q3[c] - 1.0;
        if (q3[c] < 0.0) q3[c] = q3[c] + 1.0;
    }
}

/* misc.f90:906-924 mux_vector */
int64_t orc_mux_vector(const int64_t v[3], const int64_t mesh[3], int64_t base) {
    if (base == 0) return (v[2] * mesh[1] + v[1]) * mesh[0] + v[0] + 1;
    return ((v[2] - 1) * mesh[1] + (v[1] - 1)) * mesh[0] + v[0];
}

/* misc.f90:926-942 demux_vector (int_div = Fortran / and mod, truncating) */
void orc_demux_vector(int64_t i, int64_t v[3], const int64_t mesh[3], int64_t base) {
    int64_t aux = (i - 1) / mesh[0];
    v[0] = (i - 1) % mesh[0];
    v[2] = aux / mesh[1];
    v[1] = aux % mesh[1];
    if (base == 1) { v[0] += 1; v[1] += 1; v[2] += 1; }
}

/* misc.f90:965-975 mux_state */
int64_t orc_mux_state(int64_t nbands, int64_t iband, int64_t ik) { return (ik - 1) * nbands + iband; }

/* misc.f90:977-990 demux_state */
void orc_demux_state(int64_t m, int64_t nbands, int64_t *iband, int64_t *ik) {
    *iband = fmodulo(m - 1, nbands) + 1;
    *ik = (m - 1) / nbands + 1;
}

/* misc.f90:738-763 binsearch; returns 1-based index or -1 */
int64_t orc_binsearch(const int64_t *array, int64_t n, int64_t e) {
    int64_t a = 1, b = n, m = (b + a) / 2;
    int64_t mid = array[m - 1];
    while (mid != e) {
        if (e > mid) a = m + 1;
        else if (e < mid) b = m - 1;
        if (a > b) return -1;
        m = (b + a) / 2;
        mid = array[m - 1];
    }
    return m;
}

/* misc.f90:506-540 distribute_points, for image this_image of num_images */
void orc_distribute_points(int64_t npts, int64_t num_images, int64_t this_image, int64_t *chunk,
                           int64_t *istart, int64_t *iend, int64_t *num_active) {
    int64_t na = npts < num_images ? npts : num_images;
    *num_active = na;
    int64_t smallest = npts / na, rem = fmodulo(npts, na);
    if (this_image <= rem) {
        *chunk = smallest + 1;
        *istart = (this_image - 1) * (*chunk) + 1;
        *iend = *istart + *chunk - 1;
    } else if (this_image <= na) {
        int64_t off = (smallest + 1) * rem;
        *chunk = smallest;
        *istart = off + (this_image - rem - 1) * (*chunk) + 1;
        *iend = *istart + *chunk - 1;
    } else {
        *chunk = 0; *istart = 0; *iend = 0;
    }
}

static void matvec3(const double m[9], const double v[3], double out[3]) {
    for (int r = 0; r < 3; ++r) out[r] = m[IX2(r, 0, 3)] * v[0] + m[IX2(r, 1, 3)] * v[1] + m[IX2(r, 2, 3)] * v[2];
}

/* vector.f90:38-55 create */
void orc_vec_create(int64_t ind, const int64_t grid[3], const double prim[9], orc_vec *v) {
    v->muxed = ind;
    orc_demux_vector(ind, v->ints, grid, 0);
    for (int c = 0; c < 3; ++c) v->frac[c] = (double)v->ints[c] / (double)grid[c];
    matvec3(prim, v->frac, v->cart);
}

/* vector.f90:68-84 vector_allreps_add */
void orc_vec_add(const orc_vec *v1, const orc_vec *v2, const int64_t grid[3], const double prim[9], orc_vec *v3) {
    orc_add_and_fold(v1->frac, v2->frac, v3->frac);
    matvec3(prim, v3->frac, v3->cart);
    for (int c = 0; c < 3; ++c) v3->ints[c] = fnint(v3->frac[c] * (double)grid[c]);
    v3->muxed = orc_mux_vector(v3->ints, grid, 0);
}

/* vector.f90:86-102 vector_allreps_sub */
void orc_vec_sub(const orc_vec *v1, const orc_vec *v2, const int64_t grid[3], const double prim[9], orc_vec *v3) {
    double neg[3] = {-v2->frac[0], -v2->frac[1], -v2->frac[2]};
    orc_add_and_fold(v1->frac, neg, v3->frac);
    matvec3(prim, v3->frac, v3->cart);
    for (int c = 0; c < 3; ++c) v3->ints[c] = fnint(v3->frac[c] * (double)grid[c]);
    v3->muxed = orc_mux_vector(v3->ints, grid, 0);
}

/* vector.f90:104-120 vector_allreps_change_grid */
void orc_vec_change_grid(const orc_vec *vin, const int64_t grid[3], orc_vec *vout) {
    for (int c = 0; c < 3; ++c) { vout->frac[c] = vin->frac[c]; vout->cart[c] = vin->cart[c]; }
    for (int c = 0; c < 3; ++c) vout->ints[c] = fnint(vin->frac[c] * (double)grid[c]);
    vout->muxed = orc_mux_vector(vout->ints, grid, 0);
}

/* crystal.f90:285-293 reciprocal lattice vectors and cell volume */
void orc_reclattvecs(const double L[9], double rec[9], double *volume) {
    for (int i = 0; i < 3; ++i) {
        int j = (i + 1) % 3, k = (j + 1) % 3;
        const double *a = &L[3 * j], *b = &L[3 * k];
        rec[3 * i + 0] = a[1] * b[2] - a[2] * b[1];
        rec[3 * i + 1] = a[2] * b[0] - a[0] * b[2];
        rec[3 * i + 2] = a[0] * b[1] - a[1] * b[0];
    }
    double vol = fabs(L[0] * rec[0] + L[1] * rec[1] + L[2] * rec[2]);
    double vbz = TWOPI_ / vol;
    for (int i = 0; i < 9; ++i) rec[i] *= vbz;
    *volume = vol;
}

/* ---------------------------------------------------------------------------
 * Hermitian eigensolver with zheev('V','U') semantics (only the upper triangle
 * of `a` is read; eigenvalues ascending; orthonormal eigenvectors returned in
 * the columns of `a`).  Cyclic two-sided Jacobi: textbook, not LAPACK's QR path,
 * so eigenvector phases differ from LAPACK's; tests compare eigenvalues and
 * eigenprojectors against scipy.linalg.lapack.zheev.
 * ------------------------------------------------------------------------- */
int orc_zheev(int64_t n, zdouble *a, int64_t lda, double *w) {
    zdouble *A = malloc(sizeof(zdouble) * n * n);
    zdouble *V = malloc(sizeof(zdouble) * n * n);
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < n; ++i) {
            if (i < j) A[IX2(i, j, n)] = a[IX2(i, j, lda)];
            else if (i == j) A[IX2(i, i, n)] = creal(a[IX2(i, i, lda)]);
            else A[IX2(i, j, n)] = conj(a[IX2(j, i, lda)]);
            V[IX2(i, j, n)] = (i == j) ? 1.0 : 0.0;
        }
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, diag = 0.0;
        for (int64_t j = 0; j < n; ++j)
            for (int64_t i = 0; i < n; ++i) {
                double m = cabs(A[IX2(i, j, n)]);
                if (i != j) off += m * m; else diag += m * m;
            }
        if (off <= 1e-32 * (diag + off) || off == 0.0) break;
        for (int64_t p = 0; p < n - 1; ++p)
            for (int64_t q = p + 1; q < n; ++q) {
                zdouble c = A[IX2(p, q, n)];
                double ac = cabs(c);
                if (ac == 0.0) continue;
                double app = creal(A[IX2(p, p, n)]), aqq = creal(A[IX2(q, q, n)]);
                zdouble ph = c / ac; /* e^{i phi} */
                double theta = (aqq - app) / (2.0 * ac);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double cs = 1.0 / sqrt(t * t + 1.0), sn = t * cs;
                /* U = diag(1, conj(ph)) * [[cs, sn], [-sn, cs]] acting on columns p, q */
                zdouble u_pp = cs, u_pq = sn, u_qp = -sn * conj(ph), u_qq = cs * conj(ph);
                for (int64_t k = 0; k < n; ++k) { /* A <- A U (columns) */
                    zdouble akp = A[IX2(k, p, n)], akq = A[IX2(k, q, n)];
                    A[IX2(k, p, n)] = akp * u_pp + akq * u_qp;
                    A[IX2(k, q, n)] = akp * u_pq + akq * u_qq;
                    zdouble vkp = V[IX2(k, p, n)], vkq = V[IX2(k, q, n)];
                    V[IX2(k, p, n)] = vkp * u_pp + vkq * u_qp;
                    V[IX2(k, q, n)] = vkp * u_pq + vkq * u_qq;
                }
                for (int64_t k = 0; k < n; ++k) { /* A <- U^H A (rows) */
                    zdouble apk = A[IX2(p, k, n)], aqk = A[IX2(q, k, n)];
                    A[IX2(p, k, n)] = conj(u_pp) * apk + conj(u_qp) * aqk;
                    A[IX2(q, k, n)] = conj(u_pq) * apk + conj(u_qq) * aqk;
                }
                A[IX2(p, q, n)] = 0.0; A[IX2(q, p, n)] = 0.0;
                A[IX2(p, p, n)] = creal(A[IX2(p, p, n)]);
                A[IX2(q, q, n)] = creal(A[IX2(q, q, n)]);
            }
    }
    /* sort ascending (stable selection) */
    int64_t *ord = malloc(sizeof(int64_t) * n);
    double *ev = malloc(sizeof(double) * n);
    for (int64_t i = 0; i < n; ++i) { ord[i] = i; ev[i] = creal(A[IX2(i, i, n)]); }
    for (int64_t i = 0; i < n; ++i)
        for (int64_t j = i + 1; j < n; ++j)
            if (ev[ord[j]] < ev[ord[i]]) { int64_t t = ord[i]; ord[i] = ord[j]; ord[j] = t; }
    for (int64_t j = 0; j < n; ++j) {
        w[j] = ev[ord[j]];
        for (int64_t i = 0; i < n; ++i) a[IX2(i, j, lda)] = V[IX2(i, ord[j], n)];
    }
    free(ord); free(ev); free(A); free(V);
    return 0;
}

/* wannier.f90:430-524 el_wann.  kvecs (nk,3), energies (nk,nb), velocities
 * (nk,nb,3) or NULL, evecs (nk,nb,nb) or NULL, scissor (nb) or NULL. */
int orc_el_wann(const orc_wannier *w, int64_t nk, const double *kvecs, double *energies, double *velocities,
                zdouble *evecs, const double *scissor) {
    const int64_t nb = w->nb, nws = w->nwsk;
    if (velocities && !evecs) return 1; /* "velocity is present but not eigenvecs" */
    zdouble *H = malloc(sizeof(zdouble) * nb * nb);
    zdouble *dH = malloc(sizeof(zdouble) * 3 * nb * nb);
    double *en = malloc(sizeof(double) * nb);
    for (int64_t ik = 0; ik < nk; ++ik) {
        memset(H, 0, sizeof(zdouble) * nb * nb);
        memset(dH, 0, sizeof(zdouble) * 3 * nb * nb);
        for (int64_t iuc = 0; iuc < nws; ++iuc) {
            double dot = 0.0;
            for (int c = 0; c < 3; ++c) dot += kvecs[ik + nk * c] * (double)w->rcells_k[iuc + nws * c];
            zdouble caux = cexpi(TWOPI_ * dot) / (double)w->elwsdeg[iuc];
            for (int64_t jb = 0; jb < nb; ++jb)
                for (int64_t ib = 0; ib < nb; ++ib)
                    H[IX2(ib, jb, nb)] += caux * w->Hwann[IX3(iuc, ib, jb, nws, nb)];
            if (velocities) {
                double R[3] = {(double)w->rcells_k[iuc], (double)w->rcells_k[iuc + nws], (double)w->rcells_k[iuc + 2 * nws]};
                double rcart[3];
                matvec3(w->lattvecs, R, rcart);
                for (int64_t jb = 0; jb < nb; ++jb)
                    for (int64_t ib = 0; ib < nb; ++ib)
                        for (int ip = 0; ip < 3; ++ip)
                            dH[IX3(ip, ib, jb, 3, nb)] += I * rcart[ip] * caux * w->Hwann[IX3(iuc, ib, jb, nws, nb)];
            }
        }
        /* force Hermiticity (481-487): lower triangle gets the same value */
        for (int64_t ib = 0; ib < nb; ++ib)
            for (int64_t jb = ib + 1; jb < nb; ++jb) {
                H[IX2(ib, jb, nb)] = (H[IX2(ib, jb, nb)] + conj(H[IX2(jb, ib, nb)])) * 0.5;
                H[IX2(jb, ib, nb)] = H[IX2(ib, jb, nb)];
            }
        orc_zheev(nb, H, nb, en);
        for (int64_t ib = 0; ib < nb; ++ib) energies[ik + nk * ib] = en[ib];
        if (evecs) /* evecs(ik,:,:) = transpose(H) */
            for (int64_t ib = 0; ib < nb; ++ib)
                for (int64_t jb = 0; jb < nb; ++jb) evecs[IX3(ik, ib, jb, nk, nb)] = H[IX2(jb, ib, nb)];
        if (velocities) {
            for (int64_t ib = 0; ib < nb; ++ib)
                for (int ip = 0; ip < 3; ++ip) {
                    /* real(dot_product(v, matmul(dH, v))), v = evecs(ik,ib,:) */
                    zdouble acc = 0.0;
                    for (int64_t r = 0; r < nb; ++r) {
                        zdouble mv = 0.0;
                        for (int64_t c = 0; c < nb; ++c) mv += dH[IX3(ip, r, c, 3, nb)] * evecs[IX3(ik, ib, c, nk, nb)];
                        acc += conj(evecs[IX3(ik, ib, r, nk, nb)]) * mv;
                    }
                    velocities[IX3(ik, ib, ip, nk, nb)] = creal(acc);
                }
        }
        for (int64_t ib = 0; ib < nb; ++ib) {
            energies[ik + nk * ib] *= RYD2EV;
            if (scissor) energies[ik + nk * ib] += scissor[ib];
        }
        if (velocities)
            for (int64_t ib = 0; ib < nb; ++ib)
                for (int ip = 0; ip < 3; ++ip) velocities[IX3(ik, ib, ip, nk, nb)] *= RYD2RADTHZ;
    }
    free(H); free(dH); free(en);
    return 0;
}

/* wannier.f90:532-582 ws_wrapping_vectors.  Returns the number of wrappers
 * (written as (3,n) int32 up to maxw). */
int orc_ws_wrapping_vectors(const double lvec[9], const double vl[3], double eps, int32_t *wrappers, int32_t maxw) {
    const int sc = 3;
    int iv0[3];
    double wl[3], wc[3];
    for (int c = 0; c < 3; ++c) { iv0[c] = (int)vl[c]; wl[c] = vl[c] - (double)iv0[c]; }
    int n = 0;
    double d = 1.7976931348623157e308;
    for (int k = -sc; k <= sc; ++k)
        for (int j = -sc; j <= sc; ++j)
            for (int i = -sc; i <= sc; ++i) {
                double v[3] = {wl[0] + i, wl[1] + j, wl[2] + k};
                matvec3(lvec, v, wc);
                double nrm = sqrt(wc[0] * wc[0] + wc[1] * wc[1] + wc[2] * wc[2]);
                if (fabs(nrm - d) < eps) n = n + 1;
                else if (nrm < d) { n = 1; d = nrm; }
            }
    n = 0;
    for (int k = -sc; k <= sc; ++k)
        for (int j = -sc; j <= sc; ++j)
            for (int i = -sc; i <= sc; ++i) {
                double v[3] = {wl[0] + i, wl[1] + j, wl[2] + k};
                matvec3(lvec, v, wc);
                double nrm = sqrt(wc[0] * wc[0] + wc[1] * wc[1] + wc[2] * wc[2]);
                if (fabs(nrm - d) < eps) {
                    if (n < maxw) {
                        wrappers[3 * n + 0] = i - iv0[0];
                        wrappers[3 * n + 1] = j - iv0[1];
                        wrappers[3 * n + 2] = k - iv0[2];
                    }
                    n = n + 1;
                }
            }
    return n;
}

/* wannier.f90:784-901 dyn_nonanalytic.  q Cartesian (1/bohr); dyn_l (nbr,nbr);
 * ddyn_l (nbr,nbr,3) or NULL. */
void orc_dyn_nonanalytic(const orc_wannier *w, const double q[3], zdouble *dyn_l, zdouble *ddyn_l) {
    const int64_t nat = w->natoms, nbr = w->nbr;
    double rec[9], vol;
    orc_reclattvecs(w->lattvecs, rec, &vol);
    double a1 = sqrt(w->lattvecs[0] * w->lattvecs[0] + w->lattvecs[1] * w->lattvecs[1] + w->lattvecs[2] * w->lattvecs[2]);
    double tpiba = TWOPI_ / a1 * BOHR2NM;
    double gmax = 14.0, alph = tpiba * tpiba;
    double fac = 8.0 * PI_ / (vol / (BOHR2NM * BOHR2NM * BOHR2NM));
    double *bc = malloc(sizeof(double) * 3 * nat);
    for (int64_t a = 0; a < nat; ++a) matvec3(w->lattvecs, &w->basis[3 * a], &bc[3 * a]);
    memset(dyn_l, 0, sizeof(zdouble) * nbr * nbr);
    if (ddyn_l) memset(ddyn_l, 0, sizeof(zdouble) * nbr * nbr * 3);
    const double *eps = w->epsilon, *born = w->born;
#define BORN(i, j, a) born[(i) + 3 * ((j) + 3 * (a))]
    for (int64_t m1 = -w->coarse_qmesh[0]; m1 <= w->coarse_qmesh[0]; ++m1)
        for (int64_t m2 = -w->coarse_qmesh[1]; m2 <= w->coarse_qmesh[1]; ++m2)
            for (int64_t m3 = -w->coarse_qmesh[2]; m3 <= w->coarse_qmesh[2]; ++m3) {
                double g[3];
                for (int c = 0; c < 3; ++c) g[c] = (m1 * rec[c] + m2 * rec[3 + c] + m3 * rec[6 + c]) * BOHR2NM;
                double eg[3];
                matvec3(eps, g, eg);
                double qeq = g[0] * eg[0] + g[1] * eg[1] + g[2] * eg[2];
                if (qeq > 0.0 && qeq / alph / 4.0 < gmax) {
                    double facqd = exp(-qeq / alph / 4.0) / qeq;
                    for (int64_t ia = 0; ia < nat; ++ia) {
                        double zig[3];
                        for (int c = 0; c < 3; ++c) zig[c] = g[0] * BORN(0, c, ia) + g[1] * BORN(1, c, ia) + g[2] * BORN(2, c, ia);
                        zdouble fnat[3] = {0, 0, 0};
                        for (int64_t ja = 0; ja < nat; ++ja) {
                            double rr[3];
                            for (int c = 0; c < 3; ++c) rr[c] = (bc[3 * ia + c] - bc[3 * ja + c]) / BOHR2NM;
                            double arg = g[0] * rr[0] + g[1] * rr[1] + g[2] * rr[2];
                            double zjg[3];
                            for (int c = 0; c < 3; ++c) zjg[c] = g[0] * BORN(0, c, ja) + g[1] * BORN(1, c, ja) + g[2] * BORN(2, c, ja);
                            for (int c = 0; c < 3; ++c) fnat[c] += zjg[c] * cexpi(arg);
                        }
                        for (int ip = 0; ip < 3; ++ip)
                            for (int jp = 0; jp < 3; ++jp) {
                                int64_t idim = ia * 3 + ip, jdim = ia * 3 + jp;
                                dyn_l[IX2(idim, jdim, nbr)] -= facqd * zig[ip] * fnat[jp];
                            }
                    }
                }
                for (int c = 0; c < 3; ++c) g[c] = g[c] + q[c];
                matvec3(eps, g, eg);
                qeq = g[0] * eg[0] + g[1] * eg[1] + g[2] * eg[2];
                if (qeq > 0.0 && qeq / alph / 4.0 < gmax) {
                    double facqd = exp(-qeq / alph / 4.0) / qeq;
                    double dgeg[3];
                    for (int r = 0; r < 3; ++r)
                        dgeg[r] = (eps[r] + eps[3 * r]) * g[0] + (eps[r + 3] + eps[3 * r + 1]) * g[1] + (eps[r + 6] + eps[3 * r + 2]) * g[2];
                    double dgeg_coeff = 0.25 / alph + 1.0 / qeq;
                    for (int64_t ia = 0; ia < nat; ++ia) {
                        double zig[3];
                        for (int c = 0; c < 3; ++c) zig[c] = g[0] * BORN(0, c, ia) + g[1] * BORN(1, c, ia) + g[2] * BORN(2, c, ia);
                        for (int64_t ja = 0; ja < nat; ++ja) {
                            double rr[3];
                            for (int c = 0; c < 3; ++c) rr[c] = (bc[3 * ia + c] - bc[3 * ja + c]) / BOHR2NM;
                            double zjg[3];
                            for (int c = 0; c < 3; ++c) zjg[c] = g[0] * BORN(0, c, ja) + g[1] * BORN(1, c, ja) + g[2] * BORN(2, c, ja);
                            double arg = g[0] * rr[0] + g[1] * rr[1] + g[2] * rr[2];
                            zdouble facq = facqd * cexpi(arg);
                            for (int ip = 0; ip < 3; ++ip)
                                for (int jp = 0; jp < 3; ++jp) {
                                    int64_t idim = ia * 3 + ip, jdim = ja * 3 + jp;
                                    dyn_l[IX2(idim, jdim, nbr)] += facq * zig[ip] * zjg[jp];
                                    if (ddyn_l)
                                        for (int c = 0; c < 3; ++c)
                                            ddyn_l[IX3(idim, jdim, c, nbr, nbr)] +=
                                                facq * (zjg[jp] * BORN(c, ip, ia) + zig[ip] * BORN(c, jp, ja) +
                                                        zig[ip] * zjg[jp] * (I * rr[c] - dgeg_coeff * dgeg[c]));
                                }
                        }
                    }
                }
            }
#undef BORN
    for (int64_t i = 0; i < nbr * nbr; ++i) dyn_l[i] *= fac;
    if (ddyn_l)
        for (int64_t i = 0; i < nbr * nbr * 3; ++i) ddyn_l[i] *= fac;
    free(bc);
}

/* wannier.f90:585-782 ph_wann.  qvecs (nq,3); energies (nq,nbr);
 * evecs (nq,nbr,nbr) or NULL; velocities (nq,nbr,3) or NULL.
 * Returns 2 when an energy below -0.005 eV is found (the reference stops). */
int orc_ph_wann(const orc_wannier *w, int64_t nq, const double *qvecs, double *energies, zdouble *evecs,
                double *velocities) {
    const int64_t nbr = w->nbr, nat = w->natoms, nws = w->nwsq;
    if (velocities && !evecs) return 1;
    int status = 0;
    /* The WS wrappers depend on (iuc, iat, jat) only; the reference recomputes
     * them per q (631-650), precomputing gives the identical sets. */
    const int MAXW = 64;
    int32_t *wrap = malloc(sizeof(int32_t) * 3 * MAXW * nws * nat * nat);
    int32_t *nwrap = malloc(sizeof(int32_t) * nws * nat * nat);
    for (int64_t iuc = 0; iuc < nws; ++iuc)
        for (int64_t ja = 0; ja < nat; ++ja)
            for (int64_t ia = 0; ia < nat; ++ia) {
                double dist[3];
                for (int c = 0; c < 3; ++c) {
                    dist[c] = (double)w->rcells_q[iuc + nws * c] - w->basis[3 * ia + c] + w->basis[3 * ja + c];
                    dist[c] = dist[c] / (double)w->coarse_qmesh[c];
                }
                int64_t slot = iuc + nws * (ia + nat * ja);
                nwrap[slot] = orc_ws_wrapping_vectors(w->lattvecs, dist, 1.0e-6, &wrap[3 * MAXW * slot], MAXW);
                if (nwrap[slot] > MAXW) nwrap[slot] = MAXW;
            }
    double rec[9], vol;
    orc_reclattvecs(w->lattvecs, rec, &vol);
    zdouble *dyn = malloc(sizeof(zdouble) * nbr * nbr), *dyn_l = malloc(sizeof(zdouble) * nbr * nbr);
    zdouble *cmat = malloc(sizeof(zdouble) * nbr * nbr);
    zdouble *ddyn = malloc(sizeof(zdouble) * nbr * nbr * 3), *ddyn_l = malloc(sizeof(zdouble) * nbr * nbr * 3);
    double *om2 = malloc(sizeof(double) * nbr);
    for (int64_t iq = 0; iq < nq; ++iq) {
        double q[3] = {qvecs[iq], qvecs[iq + nq], qvecs[iq + 2 * nq]};
        memset(dyn, 0, sizeof(zdouble) * nbr * nbr);
        memset(ddyn, 0, sizeof(zdouble) * nbr * nbr * 3);
        for (int64_t iuc = 0; iuc < nws; ++iuc) {
            memset(cmat, 0, sizeof(zdouble) * nbr * nbr);
            for (int64_t ja = 0; ja < nat; ++ja)
                for (int64_t ia = 0; ia < nat; ++ia) {
                    int64_t slot = iuc + nws * (ia + nat * ja);
                    int nw = nwrap[slot];
                    for (int iw = 0; iw < nw; ++iw) {
                        double dot = 0.0;
                        for (int c = 0; c < 3; ++c)
                            dot += q[c] * (double)(w->rcells_q[iuc + nws * c] +
                                                   (int64_t)wrap[3 * MAXW * slot + 3 * iw + c] * w->coarse_qmesh[c]);
                        zdouble caux = cexpi(TWOPI_ * dot);
                        for (int jj = 0; jj < 3; ++jj)
                            for (int ii = 0; ii < 3; ++ii)
                                cmat[IX2(ia * 3 + ii, ja * 3 + jj, nbr)] += caux / (double)(w->phwsdeg[iuc] * nw);
                    }
                }
            for (int64_t j = 0; j < nbr; ++j)
                for (int64_t i = 0; i < nbr; ++i)
                    dyn[IX2(i, j, nbr)] += cmat[IX2(i, j, nbr)] * w->Dphwann[IX3(iuc, i, j, nws, nbr)];
            if (velocities) {
                double R[3] = {(double)w->rcells_q[iuc], (double)w->rcells_q[iuc + nws], (double)w->rcells_q[iuc + 2 * nws]};
                double rcart[3];
                matvec3(w->lattvecs, R, rcart);
                for (int ip = 0; ip < 3; ++ip)
                    for (int64_t j = 0; j < nbr; ++j)
                        for (int64_t i = 0; i < nbr; ++i)
                            ddyn[IX3(i, j, ip, nbr, nbr)] += I * rcart[ip] * cmat[IX2(i, j, nbr)] * w->Dphwann[IX3(iuc, i, j, nws, nbr)];
            }
        }
        if (w->polar) {
            double qc[3];
            matvec3(rec, q, qc);
            for (int c = 0; c < 3; ++c) qc[c] *= BOHR2NM;
            orc_dyn_nonanalytic(w, qc, dyn_l, velocities ? ddyn_l : NULL);
            if (velocities)
                for (int64_t i = 0; i < nbr * nbr * 3; ++i) ddyn[i] += ddyn_l[i] * BOHR2NM;
            for (int64_t i = 0; i < nbr * nbr; ++i) dyn[i] += dyn_l[i];
        }
        for (int64_t ib = 0; ib < nbr; ++ib)
            for (int64_t jb = ib + 1; jb < nbr; ++jb) {
                dyn[IX2(ib, jb, nbr)] = (dyn[IX2(ib, jb, nbr)] + conj(dyn[IX2(jb, ib, nbr)])) * 0.5;
                dyn[IX2(jb, ib, nbr)] = dyn[IX2(ib, jb, nbr)];
            }
        for (int64_t na = 0; na < nat; ++na)
            for (int64_t nb2 = 0; nb2 < nat; ++nb2) {
                double massnorm = 1.0 / sqrt(w->masses[na] * w->masses[nb2]) * RYD2AMU;
                for (int jj = 0; jj < 3; ++jj)
                    for (int ii = 0; ii < 3; ++ii) {
                        dyn[IX2(3 * na + ii, 3 * nb2 + jj, nbr)] *= massnorm;
                        if (velocities)
                            for (int c = 0; c < 3; ++c) ddyn[IX3(3 * na + ii, 3 * nb2 + jj, c, nbr, nbr)] *= massnorm;
                    }
            }
        orc_zheev(nbr, dyn, nbr, om2);
        for (int64_t s = 0; s < nbr; ++s) {
            double a = sqrt(fabs(om2[s]));
            energies[iq + nq * s] = (om2[s] >= 0.0 && !signbit(om2[s])) ? a : -a; /* sign(a, om2) */
        }
        if (evecs)
            for (int64_t s = 0; s < nbr; ++s)
                for (int64_t p = 0; p < nbr; ++p) evecs[IX3(iq, s, p, nq, nbr)] = dyn[IX2(p, s, nbr)];
        if (velocities) {
            for (int64_t ib = 0; ib < nbr; ++ib) {
                for (int ip = 0; ip < 3; ++ip) {
                    zdouble acc = 0.0;
                    for (int64_t r = 0; r < nbr; ++r) {
                        zdouble mv = 0.0;
                        for (int64_t c = 0; c < nbr; ++c) mv += ddyn[IX3(r, c, ip, nbr, nbr)] * dyn[IX2(c, ib, nbr)];
                        acc += conj(dyn[IX2(r, ib, nbr)]) * mv;
                    }
                    velocities[IX3(iq, ib, ip, nq, nbr)] = creal(acc);
                }
                for (int ip = 0; ip < 3; ++ip)
                    velocities[IX3(iq, ib, ip, nq, nbr)] = velocities[IX3(iq, ib, ip, nq, nbr)] / (2.0 * energies[iq + nq * ib]);
            }
        }
        for (int64_t s = 0; s < nbr; ++s) energies[iq + nq * s] *= RYD2EV;
        if (velocities)
            for (int64_t i = 0; i < nbr * 3; ++i) velocities[iq + nq * i] *= RYD2RADTHZ;
        if (q[0] == 0 && q[1] == 0 && q[2] == 0) {
            for (int64_t s = 0; s < 3 && s < nbr; ++s) energies[iq + nq * s] = 0.0;
            if (velocities)
                for (int64_t i = 0; i < nbr * 3; ++i) velocities[iq + nq * i] = 0.0;
        }
        for (int64_t s = 0; s < nbr; ++s) {
            double e = energies[iq + nq * s];
            if (e < -0.005) status = 2;
            else if (e < 0 && e > -0.005) energies[iq + nq * s] = 0.0;
        }
    }
    free(wrap); free(nwrap); free(dyn); free(dyn_l); free(cmat); free(ddyn); free(ddyn_l); free(om2);
    return status;
}

/* wannier.f90:1013-1071 long_range_prefac.  q Cartesian (1/bohr). */
void orc_long_range_prefac(const orc_wannier *w, const double q[3], const zdouble *uqs, zdouble *glprefac) {
    const int64_t nat = w->natoms;
    double rec[9], vol;
    orc_reclattvecs(w->lattvecs, rec, &vol);
    double a1 = sqrt(w->lattvecs[0] * w->lattvecs[0] + w->lattvecs[1] * w->lattvecs[1] + w->lattvecs[2] * w->lattvecs[2]);
    double tpiba = TWOPI_ / a1 * BOHR2NM;
    double gmax = 14.0, alph = tpiba * tpiba;
    zdouble fac = 8.0 * PI_ / (vol / (BOHR2NM * BOHR2NM * BOHR2NM)) * I;
    zdouble acc = 0.0;
    const double *eps = w->epsilon, *born = w->born;
    for (int64_t m1 = -w->coarse_qmesh[0]; m1 <= w->coarse_qmesh[0]; ++m1)
        for (int64_t m2 = -w->coarse_qmesh[1]; m2 <= w->coarse_qmesh[1]; ++m2)
            for (int64_t m3 = -w->coarse_qmesh[2]; m3 <= w->coarse_qmesh[2]; ++m3) {
                double g[3], eg[3];
                for (int c = 0; c < 3; ++c) g[c] = (m1 * rec[c] + m2 * rec[3 + c] + m3 * rec[6 + c]) * BOHR2NM + q[c];
                matvec3(eps, g, eg);
                double qeq = g[0] * eg[0] + g[1] * eg[1] + g[2] * eg[2];
                if (qeq > 0.0 && qeq / alph / 4.0 < gmax) {
                    double facqd = exp(-qeq / alph / 4.0) / qeq;
                    for (int64_t na = 0; na < nat; ++na) {
                        double bc[3];
                        matvec3(w->lattvecs, &w->basis[3 * na], bc);
                        double arg = -(g[0] * bc[0] + g[1] * bc[1] + g[2] * bc[2]) / BOHR2NM;
                        zdouble facq = facqd * cexpi(arg);
                        for (int ip = 0; ip < 3; ++ip) {
                            const double *b = &born[3 * (ip + 3 * na)];
                            double zaq = g[0] * b[0] + g[1] * b[1] + g[2] * b[2];
                            acc += facq * zaq * uqs[3 * na + ip];
                        }
                    }
                }
            }
    *glprefac = acc * fac;
}

/* wannier.f90:903-1011 g2.  Eigenvectors are read with a stride so callers
 * can pass row slices of (nk,nb,nb) arrays as the reference does.
 * gmixed is (nb,nb,nbr,nws) with nws = nwsg ('ph') or nwsk ('el'). */
double orc_g2(const orc_wannier *w, const double kvec[3], const double qvec[3], const zdouble *el_evec_k,
              int64_t stride_k, const zdouble *el_evec_kp, int64_t stride_kp, const zdouble *ph_evec_q,
              int64_t stride_q, double ph_en, const zdouble *gmixed, int wannspace_ph) {
    const int64_t nb = w->nb, nbr = w->nbr;
    const double g2unitfactor = RYD2EV * RYD2EV * RYD2EV * RYD2AMU;
    zdouble u[64];
    for (int64_t ip = 0; ip < nbr; ++ip) {
        int64_t mtype = ip / 3;
        u[ip] = ph_evec_q[ip * stride_q] / sqrt(w->masses[mtype]);
    }
    if (ph_en == 0) return 0.0;
    int64_t nws = wannspace_ph ? w->nwsg : w->nwsk;
    zdouble *A = calloc(nbr * nws, sizeof(zdouble));
    zdouble *B = calloc(nws, sizeof(zdouble));
    zdouble ov[64 * 64];
    zdouble gbloch = 0.0;
    for (int64_t np = 0; np < nb; ++np)
        for (int64_t mp = 0; mp < nb; ++mp) ov[IX2(mp, np, nb)] = conj(el_evec_kp[np * stride_kp]) * el_evec_k[mp * stride_k];
    for (int64_t iws = 0; iws < nws; ++iws)
        for (int64_t sp = 0; sp < nbr; ++sp) {
            zdouble caux = 0.0;
            for (int64_t np = 0; np < nb; ++np)
                for (int64_t mp = 0; mp < nb; ++mp)
                    caux += ov[IX2(mp, np, nb)] * gmixed[np + nb * (mp + nb * (sp + nbr * iws))];
            A[sp + nbr * iws] += caux;
        }
    for (int64_t iws = 0; iws < nws; ++iws) {
        zdouble acc = 0.0; /* dot_product(conjg(u), A(:,iws)) = sum u*A */
        for (int64_t sp = 0; sp < nbr; ++sp) acc += u[sp] * A[sp + nbr * iws];
        B[iws] += acc;
    }
    for (int64_t iws = 0; iws < nws; ++iws) {
        zdouble caux;
        if (wannspace_ph) {
            double dot = 0.0;
            for (int c = 0; c < 3; ++c) dot += qvec[c] * (double)w->rcells_q[iws + w->nwsq * c];
            caux = cexpi(TWOPI_ * dot) / (double)w->phwsdeg[iws];
        } else {
            double dot = 0.0;
            for (int c = 0; c < 3; ++c) dot += kvec[c] * (double)w->rcells_k[iws + w->nwsk * c];
            caux = cexpi(TWOPI_ * dot) / (double)w->elwsdeg[iws];
        }
        gbloch += caux * B[iws];
    }
    if (w->polar) {
        zdouble unm = 0.0; /* dot_product(el_evec_kp, el_evec_k) */
        for (int64_t i = 0; i < nb; ++i) unm += conj(el_evec_kp[i * stride_kp]) * el_evec_k[i * stride_k];
        double rec[9], vol, qc[3];
        orc_reclattvecs(w->lattvecs, rec, &vol);
        matvec3(rec, qvec, qc);
        for (int c = 0; c < 3; ++c) qc[c] *= BOHR2NM;
        zdouble glp;
        orc_long_range_prefac(w, qc, u, &glp);
        gbloch = gbloch + glp * unm;
    }
    free(A); free(B);
    return 0.5 * creal(gbloch * conj(gbloch)) / ph_en * g2unitfactor;
}

/* wannier.f90:1073-1132 gkRp: gmixed(np,mp,sp,Rp) = sum_Re phase(Re) gwann(np,mp,Re,sp,Rp).
 * Output (nb,nb,nbr,nwsq); entries beyond nwsg stay zero. */
void orc_gkRp(const orc_wannier *w, const double kvec[3], zdouble *gmixed) {
    const int64_t nb = w->nb, nbr = w->nbr, nwsk = w->nwsk, nwsg = w->nwsg;
    const int64_t nb2 = nb * nb;
    memset(gmixed, 0, sizeof(zdouble) * nb2 * nbr * w->nwsq);
    zdouble *phase = malloc(sizeof(zdouble) * nwsk);
    for (int64_t iuc = 0; iuc < nwsk; ++iuc) {
        double dot = 0.0;
        for (int c = 0; c < 3; ++c) dot += kvec[c] * (double)w->rcells_k[iuc + nwsk * c];
        phase[iuc] = cexpi(TWOPI_ * dot);
    }
    for (int64_t iuc = 0; iuc < nwsk; ++iuc) phase[iuc] = phase[iuc] / (double)w->elwsdeg[iuc];
    for (int64_t iuc = 0; iuc < nwsk; ++iuc) {
        zdouble caux = phase[iuc];
        for (int64_t rp = 0; rp < nwsg; ++rp)
            for (int64_t sp = 0; sp < nbr; ++sp) {
                const zdouble *src = &w->gwann[nb2 * (iuc + nwsk * (sp + nbr * rp))];
                zdouble *dst = &gmixed[nb2 * (sp + nbr * rp)];
                for (int64_t i = 0; i < nb2; ++i) dst[i] = dst[i] + caux * src[i];
            }
    }
    free(phase);
}

/* wannier.f90:1134-1188 gReq: gmixed(np,mp,s,Re) = sum_Rp phase(q,Rp) gwann(np,mp,Re,s,Rp).
 * Output (nb,nb,nbr,nwsk). */
void orc_gReq(const orc_wannier *w, const double qvec[3], zdouble *gmixed) {
    const int64_t nb = w->nb, nbr = w->nbr, nwsk = w->nwsk, nwsg = w->nwsg;
    const int64_t nb2 = nb * nb;
    memset(gmixed, 0, sizeof(zdouble) * nb2 * nbr * nwsk);
    for (int64_t iuc = 0; iuc < nwsg; ++iuc) {
        double dot = 0.0;
        for (int c = 0; c < 3; ++c) dot += qvec[c] * (double)w->rcells_g[iuc + nwsg * c];
        zdouble phase = cexpi(TWOPI_ * dot) / (double)w->gwsdeg[iuc];
        for (int64_t s = 0; s < nbr; ++s)
            for (int64_t re = 0; re < nwsk; ++re) {
                const zdouble *src = &w->gwann[nb2 * (re + nwsk * (s + nbr * iuc))];
                zdouble *dst = &gmixed[nb2 * (s + nbr * re)];
                for (int64_t i = 0; i < nb2; ++i) dst[i] = dst[i] + phase * src[i];
            }
    }
}

/* ---------------------------------------------------------------------------
 * symmetry.f90 (without spglib): point-group search by brute force.
 * ------------------------------------------------------------------------- */

/* All integer matrices R (entries -1..1, fractional coordinates, x' = R x + t)
 * that preserve the lattice metric and map the atomic basis onto itself with
 * species, the operation set spglib returns for spg_get_dataset
 * (symmetry.f90:97).  Returns nsymm; rot_out is (3,3,nsymm) column-major R. */
int orc_find_rotations(const double L[9], const double *basis, const int64_t *types, int64_t natoms,
                       double symprec, int32_t *rot_out) {
    double G[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0;
            for (int c = 0; c < 3; ++c) s += L[3 * i + c] * L[3 * j + c];
            G[IX2(i, j, 3)] = s;
        }
    double gscale = (G[0] + G[4] + G[8]) / 3.0;
    int nsymm = 0;
    int e[9];
    for (int code = 0; code < 19683; ++code) {
        int c = code;
        for (int i = 0; i < 9; ++i) { e[i] = c % 3 - 1; c /= 3; }
        /* R column-major: R(r,cidx) = e[r + 3*cidx] */
        int det = e[0] * (e[4] * e[8] - e[7] * e[5]) - e[3] * (e[1] * e[8] - e[7] * e[2]) + e[6] * (e[1] * e[5] - e[4] * e[2]);
        if (det != 1 && det != -1) continue;
        int ok = 1;
        for (int i = 0; i < 3 && ok; ++i)
            for (int j = 0; j < 3 && ok; ++j) {
                double s = 0; /* (R^T G R)(i,j) */
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) s += e[a + 3 * i] * G[IX2(a, b, 3)] * e[b + 3 * j];
                if (fabs(s - G[IX2(i, j, 3)]) > 1e-8 * gscale) ok = 0;
            }
        if (!ok) continue;
        /* find a translation t = tau_j - R tau_0 mapping the basis onto itself */
        int found = 0;
        for (int64_t j0 = 0; j0 < natoms && !found; ++j0) {
            if (types[j0] != types[0]) continue;
            double Rt0[3], t[3];
            for (int r = 0; r < 3; ++r) Rt0[r] = e[r] * basis[0] + e[r + 3] * basis[1] + e[r + 6] * basis[2];
            for (int r = 0; r < 3; ++r) t[r] = basis[3 * j0 + r] - Rt0[r];
            int all = 1;
            for (int64_t a = 0; a < natoms && all; ++a) {
                double x[3];
                for (int r = 0; r < 3; ++r)
                    x[r] = e[r] * basis[3 * a] + e[r + 3] * basis[3 * a + 1] + e[r + 6] * basis[3 * a + 2] + t[r];
                int hit = 0;
                for (int64_t b = 0; b < natoms && !hit; ++b) {
                    if (types[b] != types[a]) continue;
                    double d[3], dc[3];
                    for (int r = 0; r < 3; ++r) { d[r] = x[r] - basis[3 * b + r]; d[r] -= nearbyint(d[r]); }
                    matvec3(L, d, dc);
                    if (sqrt(dc[0] * dc[0] + dc[1] * dc[1] + dc[2] * dc[2]) < symprec) hit = 1;
                }
                if (!hit) all = 0;
            }
            if (all) found = 1;
        }
        if (!found) continue;
        for (int i = 0; i < 9; ++i) rot_out[9 * nsymm + i] = e[i];
        nsymm++;
        if (nsymm >= 48) break;
    }
    return nsymm;
}

static void inv3(const double a[9], double r[9]) {
    double det = a[0] * (a[4] * a[8] - a[7] * a[5]) - a[3] * (a[1] * a[8] - a[7] * a[2]) + a[6] * (a[1] * a[5] - a[4] * a[2]);
    r[0] = (a[4] * a[8] - a[5] * a[7]) / det; r[3] = (a[6] * a[5] - a[3] * a[8]) / det; r[6] = (a[3] * a[7] - a[6] * a[4]) / det;
    r[1] = (a[7] * a[2] - a[1] * a[8]) / det; r[4] = (a[0] * a[8] - a[6] * a[2]) / det; r[7] = (a[6] * a[1] - a[0] * a[7]) / det;
    r[2] = (a[1] * a[5] - a[4] * a[2]) / det; r[5] = (a[3] * a[2] - a[0] * a[5]) / det; r[8] = (a[0] * a[4] - a[3] * a[1]) / det;
}

/* symmetry.f90:134-211: qrotations = G R G^-1 (G = L^T L), doubled by time
 * reversal, then rotations incompatible with the mesh and duplicates of the
 * real-space rotations are dropped.  Returns nsymm_rot. */
int64_t orc_qrotations(const double L[9], const int32_t *rot, int64_t nsymm, const int64_t mesh[3], double *qrot_out,
                       double *crot_out) {
    double G[9], Gi[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0;
            for (int c = 0; c < 3; ++c) s += L[3 * i + c] * L[3 * j + c];
            G[IX2(i, j, 3)] = s;
        }
    inv3(G, Gi);
    int64_t n2 = 2 * nsymm;
    double *qr = malloc(sizeof(double) * 9 * n2);
    int32_t *rr = malloc(sizeof(int32_t) * 9 * n2);
    for (int64_t s = 0; s < nsymm; ++s) {
        double GR[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double v = 0;
                for (int a = 0; a < 3; ++a) v += G[IX2(i, a, 3)] * rot[9 * s + IX2(a, j, 3)];
                GR[IX2(i, j, 3)] = v;
            }
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double v = 0;
                for (int a = 0; a < 3; ++a) v += GR[IX2(i, a, 3)] * Gi[IX2(a, j, 3)];
                qr[9 * s + IX2(i, j, 3)] = v;
                qr[9 * (s + nsymm) + IX2(i, j, 3)] = -v;
            }
        for (int i = 0; i < 9; ++i) { rr[9 * s + i] = rot[9 * s + i]; rr[9 * (s + nsymm) + i] = -rot[9 * s + i]; }
    }
    int64_t nq = mesh[0] * mesh[1] * mesh[2];
    int64_t *emap = malloc(sizeof(int64_t) * n2 * nq);
    orc_find_equiv_map(n2, emap, mesh, qr, NULL, nq);
    char *valid = malloc(n2);
    for (int64_t s = 0; s < n2; ++s) {
        valid[s] = 1;
        for (int64_t i = 0; i < nq; ++i)
            if (emap[s + n2 * i] == -1) { valid[s] = 0; break; }
    }
    for (int64_t s = 1; s < n2; ++s)
        for (int64_t t = 0; t < s; ++t) {
            if (!valid[t]) continue;
            if (memcmp(&rr[9 * s], &rr[9 * t], 9 * sizeof(int32_t)) == 0) { valid[s] = 0; break; }
        }
    int64_t k = 0;
    for (int64_t s = 0; s < n2; ++s)
        if (valid[s]) { memcpy(&qrot_out[9 * k], &qr[9 * s], 9 * sizeof(double)); k++; }
    if (crot_out) {
        /* symmetry.f90:138, 159, 191-205: crotations in the same filtered order */
        double *c = malloc(sizeof(double) * 9 * nsymm);
        orc_cartesian_rotations(L, rot, nsymm, c);
        int64_t kc = 0;
        for (int64_t s = 0; s < n2; ++s) {
            if (!valid[s]) continue;
            for (int e = 0; e < 9; ++e) crot_out[9 * kc + e] = s < nsymm ? c[9 * s + e] : -c[9 * (s - nsymm) + e];
            kc++;
        }
        free(c);
    }
    free(qr); free(rr); free(emap); free(valid);
    return k;
}

/* symmetry.f90:214-301 find_star + find_equiv_map.  equiv_map (nsymm_rot, nlist). */
void orc_find_equiv_map(int64_t nsymm_rot, int64_t *equiv_map, const int64_t mesh[3], const double *qrot,
                        const int64_t *indexlist, int64_t nlist) {
    for (int64_t i = 0; i < nlist; ++i) {
        int64_t ijk[3];
        orc_demux_vector(indexlist ? indexlist[i] : i + 1, ijk, mesh, 0);
        double f[3] = {(double)ijk[0] / (double)mesh[0], (double)ijk[1] / (double)mesh[1], (double)ijk[2] / (double)mesh[2]};
        for (int64_t s = 0; s < nsymm_rot; ++s) {
            double v[3], d[3];
            matvec3(&qrot[9 * s], f, v);
            int64_t iv[3];
            for (int c = 0; c < 3; ++c) { v[c] = (double)mesh[c] * v[c]; iv[c] = fnint(v[c]); d[c] = fabs(v[c] - (double)iv[c]); }
            if (sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) >= 1e-2) equiv_map[s + nsymm_rot * i] = -1;
            else {
                int64_t m[3] = {fmodulo(iv[0], mesh[0]), fmodulo(iv[1], mesh[1]), fmodulo(iv[2], mesh[2])};
                equiv_map[s + nsymm_rot * i] = orc_mux_vector(m, mesh, 0);
            }
        }
    }
}

/* symmetry.f90:303-451 find_irred_wedge (serial form of the running-list scan).
 * indexlist may be NULL (full mesh).  Returns nwv_irred or -1 on failure.
 * ibz2fbz is (nsymm_rot, nlist, 2) holding (rotation, image) pairs (1-based). */
int64_t orc_find_irred_wedge(const int64_t mesh[3], int64_t nsymm_rot, const double *qrot, const int64_t *indexlist,
                             int64_t nlist, int64_t *indexlist_irred, int64_t *nequiv, int64_t *ibz2fbz) {
    int64_t *emap = malloc(sizeof(int64_t) * nsymm_rot * nlist);
    orc_find_equiv_map(nsymm_rot, emap, mesh, qrot, indexlist, nlist);
    int64_t nmesh = mesh[0] * mesh[1] * mesh[2];
    char *in_running = calloc(nmesh + 1, 1); /* membership of the running list */
    int64_t *imagelist = malloc(sizeof(int64_t) * nsymm_rot);
    int64_t nirr = 0, counter = 0;
    for (int64_t i = 0; i < nlist; ++i) {
        int64_t imux = indexlist ? indexlist[i] : i + 1;
        if (in_running[imux]) continue;
        nirr++;
        indexlist_irred[nirr - 1] = imux;
        nequiv[nirr - 1] = 0;
        for (int64_t s = 0; s < nsymm_rot; ++s) {
            int64_t image = emap[s + nsymm_rot * i];
            int present = 0;
            for (int64_t l = 0; l < nequiv[nirr - 1]; ++l)
                if (imagelist[l] == image) { present = 1; break; }
            if (!present) {
                int64_t aux = ++nequiv[nirr - 1];
                imagelist[aux - 1] = image;
                if (image >= 1 && image <= nmesh) in_running[image] = 1;
                if (ibz2fbz) {
                    ibz2fbz[(aux - 1) + nsymm_rot * ((nirr - 1) + nlist * 0)] = s + 1;
                    ibz2fbz[(aux - 1) + nsymm_rot * ((nirr - 1) + nlist * 1)] = image;
                }
            }
        }
        counter += nequiv[nirr - 1];
    }
    free(emap); free(in_running); free(imagelist);
    if (counter != nlist) return -1; /* "Could not find irreducible wedge." */
    return nirr;
}

/* electron.f90:321-595 (steps 1-10 that feed the vertex) + apply_energy_window
 * (597-636). */
int orc_calculate_electrons(const orc_wannier *w, const double *qrot, int64_t nsr, const int64_t wvmesh[3],
                            double enref, double fsthick, orc_electrons *el) {
    const int64_t nb = w->nb;
    memset(el, 0, sizeof(*el));
    for (int c = 0; c < 3; ++c) el->wvmesh[c] = wvmesh[c];
    el->enref = enref; el->fsthick = fsthick;
    int64_t nwv = wvmesh[0] * wvmesh[1] * wvmesh[2];
    /* 2. IBZ of the full mesh */
    int64_t *irr = malloc(sizeof(int64_t) * nwv), *neq = malloc(sizeof(int64_t) * nwv);
    int64_t *map = malloc(sizeof(int64_t) * nsr * nwv * 2);
    int64_t nirr = orc_find_irred_wedge(wvmesh, nsr, qrot, NULL, nwv, irr, neq, map);
    if (nirr < 0) return 3;
    /* 3. IBZ energies */
    double *kir = malloc(sizeof(double) * nirr * 3);
    for (int64_t i = 0; i < nirr; ++i) {
        int64_t ijk[3];
        orc_demux_vector(irr[i], ijk, wvmesh, 0);
        for (int c = 0; c < 3; ++c) kir[i + nirr * c] = (double)ijk[c] / (double)wvmesh[c];
    }
    double *eir = malloc(sizeof(double) * nirr * nb);
    orc_el_wann(w, nirr, kir, eir, NULL, NULL, NULL);
    /* 4. FBZ energies copied from the IBZ */
    double *efbz = malloc(sizeof(double) * nwv * nb);
    for (int64_t i = 0; i < nirr; ++i)
        for (int64_t l = 0; l < neq[i]; ++l) {
            int64_t il = map[l + nsr * (i + nwv * 1)];
            for (int64_t b = 0; b < nb; ++b) efbz[(il - 1) + nwv * b] = eir[i + nirr * b];
        }
    /* 5. window (any band within fsthick), 6. sorted (already ascending) */
    int64_t *idx = malloc(sizeof(int64_t) * nwv);
    int64_t cnt = 0;
    for (int64_t ik = 0; ik < nwv; ++ik) {
        int in = 0;
        for (int64_t b = 0; b < nb; ++b)
            if (fabs(efbz[ik + nwv * b] - enref) <= fsthick) in = 1;
        if (in) idx[cnt++] = ik + 1;
    }
    if (cnt == 0) return 4; /* "No states found within Fermi window." */
    el->nwv = cnt;
    el->indexlist = idx;
    el->ens = malloc(sizeof(double) * cnt * nb);
    for (int64_t i = 0; i < cnt; ++i)
        for (int64_t b = 0; b < nb; ++b) el->ens[i + cnt * b] = efbz[(idx[i] - 1) + nwv * b];
    /* 7. FBZ-block eigenvectors from direct calculations */
    double *kw = malloc(sizeof(double) * cnt * 3);
    for (int64_t i = 0; i < cnt; ++i) {
        int64_t ijk[3];
        orc_demux_vector(idx[i], ijk, wvmesh, 0);
        for (int c = 0; c < 3; ++c) kw[i + cnt * c] = (double)ijk[c] / (double)wvmesh[c];
    }
    el->evecs = malloc(sizeof(zdouble) * cnt * nb * nb);
    double *etmp = malloc(sizeof(double) * cnt * nb);
    double *vtmp = malloc(sizeof(double) * cnt * nb * 3);
    orc_el_wann(w, cnt, kw, etmp, vtmp, el->evecs, NULL);
    /* 8. IBZ of the blocks */
    int64_t *irr2 = malloc(sizeof(int64_t) * cnt), *neq2 = malloc(sizeof(int64_t) * cnt);
    int64_t *map2 = malloc(sizeof(int64_t) * nsr * cnt * 2);
    int64_t nirr2 = orc_find_irred_wedge(wvmesh, nsr, qrot, idx, cnt, irr2, neq2, map2);
    if (nirr2 < 0) return 3;
    el->nwv_irred = nirr2;
    el->indexlist_irred = irr2;
    /* 9. IBZ-block energies and eigenvectors */
    double *kir2 = malloc(sizeof(double) * nirr2 * 3);
    for (int64_t i = 0; i < nirr2; ++i) {
        int64_t ijk[3];
        orc_demux_vector(irr2[i], ijk, wvmesh, 0);
        for (int c = 0; c < 3; ++c) kir2[i + nirr2 * c] = (double)ijk[c] / (double)wvmesh[c];
    }
    el->ens_irred = malloc(sizeof(double) * nirr2 * nb);
    el->evecs_irred = malloc(sizeof(zdouble) * nirr2 * nb * nb);
    double *vir2 = malloc(sizeof(double) * nirr2 * nb * 3);
    orc_el_wann(w, nirr2, kir2, el->ens_irred, vir2, el->evecs_irred, NULL);
    /* 10. number of in-window FBZ-block states */
    el->nstates_inwindow = 0;
    for (int64_t i = 0; i < cnt; ++i)
        for (int64_t b = 0; b < nb; ++b)
            if (fabs(el->ens[i + cnt * b] - enref) <= fsthick) el->nstates_inwindow++;
    /* 11. FBZ-block energies re-copied from the IBZ blocks (580-585 of the
     * later loop; identical values to step 4 for a symmetric window) */
    for (int64_t i = 0; i < nirr2; ++i)
        for (int64_t l = 0; l < neq2[i]; ++l) {
            int64_t il = map2[l + nsr * (i + cnt * 1)];
            int64_t aux = orc_binsearch(idx, cnt, il);
            if (aux > 0)
                for (int64_t b = 0; b < nb; ++b) el->ens[(aux - 1) + cnt * b] = el->ens_irred[i + nirr2 * b];
        }
    free(irr); free(neq); free(map); free(kir); free(eir); free(efbz); free(kw); free(etmp); free(vtmp);
    free(neq2); free(map2); free(kir2); free(vir2);
    return 0;
}

void orc_free_electrons(orc_electrons *el) {
    free(el->indexlist_irred); free(el->ens_irred); free(el->evecs_irred);
    free(el->indexlist); free(el->ens); free(el->evecs);
    memset(el, 0, sizeof(*el));
}

/* phonon.f90:159-310 calculate_phonons, Wannier branch: FBZ ph_wann, IBZ, and
 * energies copied from the IBZ representative. */
int orc_calculate_phonons(const orc_wannier *w, const double *qrot, int64_t nsr, const int64_t wvmesh[3], orc_phonons *ph) {
    const int64_t nbr = w->nbr;
    memset(ph, 0, sizeof(*ph));
    for (int c = 0; c < 3; ++c) ph->wvmesh[c] = wvmesh[c];
    int64_t nq = wvmesh[0] * wvmesh[1] * wvmesh[2];
    ph->nwv = nq;
    double *qv = malloc(sizeof(double) * nq * 3);
    for (int64_t i = 0; i < nq; ++i) {
        int64_t ijk[3];
        orc_demux_vector(i + 1, ijk, wvmesh, 0);
        for (int c = 0; c < 3; ++c) qv[i + nq * c] = (double)ijk[c] / (double)wvmesh[c];
    }
    ph->ens = malloc(sizeof(double) * nq * nbr);
    ph->evecs = malloc(sizeof(zdouble) * nq * nbr * nbr);
    int st = orc_ph_wann(w, nq, qv, ph->ens, ph->evecs, NULL);
    free(qv);
    if (st) return st;
    int64_t *irr = malloc(sizeof(int64_t) * nq), *neq = malloc(sizeof(int64_t) * nq);
    int64_t *map = malloc(sizeof(int64_t) * nsr * nq * 2);
    int64_t nirr = orc_find_irred_wedge(wvmesh, nsr, qrot, NULL, nq, irr, neq, map);
    if (nirr < 0) return 3;
    for (int64_t i = 0; i < nirr; ++i) {
        int64_t ii = irr[i];
        for (int64_t l = 0; l < neq[i]; ++l) {
            int64_t il = map[l + nsr * (i + nq * 1)];
            for (int64_t s = 0; s < nbr; ++s) ph->ens[(il - 1) + nq * s] = ph->ens[(ii - 1) + nq * s];
        }
    }
    ph->nwv_irred = nirr;
    ph->indexlist_irred = irr;
    free(neq); free(map);
    return 0;
}

void orc_free_phonons(orc_phonons *ph) {
    free(ph->ens); free(ph->evecs); free(ph->indexlist_irred);
    memset(ph, 0, sizeof(*ph));
}

/* interactions.f90:2095-2400 calculate_eph_interaction_ibzk, key 'g'.
 * The gkRp of every needed IBZ k is formed in memory (the reference writes and
 * re-reads it as gkRp.ik<ik>, 1120-1131 / 2173-2182). */
int orc_ibzk_g2(const orc_wannier *w, const orc_electrons *el, const orc_phonons *ph, int64_t mesh_ref,
                int64_t istate_first, int64_t istate_last, int64_t *out_count, double *out_g2, int nthreads) {
    return orc_ibzk_g2_sample(w, el, ph, mesh_ref, istate_first, istate_last, out_count, out_g2, nthreads, -1);
}

/* Same driver restricted to the final wave vectors ikp_first..ikp_first+nkp-1
 * (1-based) of the window; out_count[t] counts the vertices of that range and
 * row t holds them from its first entry on.  The full driver's row has them at
 * nbr times the number of in-window (k', n) with k' < ikp_first. */
int orc_ibzk_g2_window(const orc_wannier *w, const orc_electrons *el, const orc_phonons *ph, int64_t mesh_ref,
                       int64_t istate_first, int64_t istate_last, int64_t *out_count, double *out_g2, int nthreads,
                       int64_t ikp_first, int64_t ikp_count) {
    const int64_t nb = w->nb, nbr = w->nbr, nirr = el->nwv_irred;
    const int64_t nwv = el->nwv;
    if (ikp_first < 1) ikp_first = 1;
    const int64_t ikp_end = (ikp_count >= 0 && ikp_first - 1 + ikp_count < nwv) ? ikp_first - 1 + ikp_count : nwv;
    const int64_t nprocs = el->nstates_inwindow * nbr;
    const int64_t gsz = nb * nb * nbr * w->nwsq;
    double rec[9], vol;
    orc_reclattvecs(w->lattvecs, rec, &vol);
    int64_t nst = istate_last - istate_first + 1;
    int status = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
    {
        zdouble *gk = malloc(sizeof(zdouble) * gsz);
        double phe[64];
        zdouble phv[64 * 64];
        int64_t cur_ik = -1;
#pragma omp for schedule(dynamic, 1)
        for (int64_t t = 0; t < nst; ++t) {
            int64_t istate = istate_first + t, m, ik;
            orc_demux_state(istate, nb, &m, &ik);
            double en_el = el->ens_irred[(ik - 1) + nirr * (m - 1)];
            if (fabs(en_el - el->enref) > el->fsthick) { out_count[t] = -1; continue; }
            int64_t ik_fbz = el->indexlist_irred[ik - 1];
            orc_vec kv, kpv, qv, qc;
            orc_vec_create(ik_fbz, el->wvmesh, rec, &kv);
            if (cur_ik != ik) { orc_gkRp(w, kv.frac, gk); cur_ik = ik; }
            int64_t count = 0;
            double *row = &out_g2[t * nprocs];
            for (int64_t ikp = ikp_first; ikp <= ikp_end; ++ikp) {
                orc_vec_create(el->indexlist[ikp - 1], el->wvmesh, rec, &kpv);
                orc_vec_sub(&kpv, &kv, el->wvmesh, rec, &qv);
                int need_fine = 0;
                for (int c = 0; c < 3; ++c)
                    if (qv.ints[c] % mesh_ref != 0) need_fine = 1;
                int64_t iqc = 0;
                if (need_fine) {
                    double ql[3] = {qv.frac[0], qv.frac[1], qv.frac[2]};
                    if (orc_ph_wann(w, 1, ql, phe, phv, NULL) == 2) status = 2;
                } else {
                    orc_vec_change_grid(&qv, ph->wvmesh, &qc);
                    iqc = qc.muxed;
                }
                for (int64_t n = 1; n <= nb; ++n) {
                    if (fabs(el->ens[(ikp - 1) + nwv * (n - 1)] - el->enref) > el->fsthick) continue;
                    for (int64_t s = 1; s <= nbr; ++s) {
                        count++;
                        if (count > nprocs) { status = 5; continue; }
                        const zdouble *uk = &el->evecs_irred[(ik - 1) + nirr * (m - 1)];
                        const zdouble *ukp = &el->evecs[(ikp - 1) + nwv * (n - 1)];
                        if (need_fine) {
                            row[count - 1] = orc_g2(w, kv.frac, qv.frac, uk, nirr * nb, ukp, nwv * nb,
                                                    &phv[s - 1], nbr, phe[s - 1], gk, 1);
                        } else {
                            const int64_t nq = ph->nwv;
                            row[count - 1] = orc_g2(w, kv.frac, qc.frac, uk, nirr * nb, ukp, nwv * nb,
                                                    &ph->evecs[(iqc - 1) + nq * (s - 1)], nq * nbr,
                                                    ph->ens[(iqc - 1) + nq * (s - 1)], gk, 1);
                        }
                    }
                }
            }
            out_count[t] = count;
        }
        free(gk);
    }
    return status;
}

/* The first ikp_limit final wave vectors of each state when ikp_limit >= 0 (a
 * bounded sample for CPU timing); the whole window otherwise. */
int orc_ibzk_g2_sample(const orc_wannier *w, const orc_electrons *el, const orc_phonons *ph, int64_t mesh_ref,
                       int64_t istate_first, int64_t istate_last, int64_t *out_count, double *out_g2, int nthreads,
                       int64_t ikp_limit) {
    return orc_ibzk_g2_window(w, el, ph, mesh_ref, istate_first, istate_last, out_count, out_g2, nthreads, 1,
                              ikp_limit);
}

/* orc_g2 over a list of n vertices, OpenMP over the list (each call is the
 * reference's g2 as restated above).  Vertex i reads kvecs/qvecs row i
 * ((n,3) column-major), el_evec_k/el_evec_kp (n,nb), ph_evec_q (n,nbr),
 * ph_en (n) and gmixed block gidx[i] of gsz complex entries. */
void orc_g2_list(const orc_wannier *w, int64_t n, const double *kvecs, const double *qvecs,
                 const zdouble *el_evec_k, const zdouble *el_evec_kp, const zdouble *ph_evec_q,
                 const double *ph_en, const zdouble *gmixed, int64_t gsz, const int64_t *gidx, int wannspace_ph,
                 double *out, int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t i = 0; i < n; ++i) {
        double k[3] = {kvecs[i], kvecs[i + n], kvecs[i + 2 * n]};
        double q[3] = {qvecs[i], qvecs[i + n], qvecs[i + 2 * n]};
        out[i] = orc_g2(w, k, q, &el_evec_k[i], n, &el_evec_kp[i], n, &ph_evec_q[i], n, ph_en[i],
                        &gmixed[gidx[i] * gsz], wannspace_ph);
    }
}

/* symmetry.f90:131-137: tmp1 = transpose(lattvecs), tmp2 = transpose(lattvecs
 * . rotations_orig(:,:,i)), dgesv(tmp1, tmp2), crotations_orig(:,:,i) =
 * transpose(tmp2).  rot (3,3,nsymm) are the operations on fractional
 * coordinates (rotations_orig), column-major. */
static void solve3(const double A_in[9], double B[9]) {
    /* B <- A^-1 B by Gaussian elimination with partial pivoting (dgesv) */
    double A[9];
    memcpy(A, A_in, sizeof(A));
    for (int c = 0; c < 3; ++c) {
        int piv = c;
        for (int r = c + 1; r < 3; ++r)
            if (fabs(A[r + 3 * c]) > fabs(A[piv + 3 * c])) piv = r;
        if (piv != c)
            for (int j = 0; j < 3; ++j) {
                double t = A[c + 3 * j]; A[c + 3 * j] = A[piv + 3 * j]; A[piv + 3 * j] = t;
                t = B[c + 3 * j]; B[c + 3 * j] = B[piv + 3 * j]; B[piv + 3 * j] = t;
            }
        for (int r = c + 1; r < 3; ++r) {
            double f = A[r + 3 * c] / A[c + 3 * c];
            for (int j = 0; j < 3; ++j) { A[r + 3 * j] -= f * A[c + 3 * j]; B[r + 3 * j] -= f * B[c + 3 * j]; }
        }
    }
    for (int c = 2; c >= 0; --c)
        for (int j = 0; j < 3; ++j) {
            double v = B[c + 3 * j];
            for (int k = c + 1; k < 3; ++k) v -= A[c + 3 * k] * B[k + 3 * j];
            B[c + 3 * j] = v / A[c + 3 * c];
        }
}

void orc_cartesian_rotations(const double L[9], const int32_t *rot, int64_t nsymm, double *crot_out) {
    for (int64_t s = 0; s < nsymm; ++s) {
        double tmp1[9], LW[9], tmp2[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                tmp1[i + 3 * j] = L[j + 3 * i];
                double v = 0;
                for (int a = 0; a < 3; ++a) v += L[i + 3 * a] * rot[9 * s + a + 3 * j];
                LW[i + 3 * j] = v;
            }
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) tmp2[i + 3 * j] = LW[j + 3 * i];
        solve3(tmp1, tmp2);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) crot_out[9 * s + i + 3 * j] = tmp2[j + 3 * i];
    }
}

/* electron.f90:462-476: symmetrizers(:,:,i) = sum of crotations_orig(:,:,jj)
 * over jj = 1..nsymm with equiv_map(jj, i) == indexlist(i), divided by the
 * count when it exceeds 1; equiv_map from find_equiv_map over the list with
 * the (filtered) qrotations. */
void orc_symmetrizers(const int64_t mesh[3], int64_t nsymm_rot, const double *qrot, int64_t nsymm,
                      const double *crot_orig, int64_t nk, const int64_t *indexlist, double *out) {
    int64_t *emap = malloc(sizeof(int64_t) * nsymm_rot * nk);
    orc_find_equiv_map(nsymm_rot, emap, mesh, qrot, indexlist, nk);
    for (int64_t i = 0; i < nk; ++i) {
        int64_t ii = indexlist ? indexlist[i] : i + 1;
        double *S = &out[9 * i];
        for (int e = 0; e < 9; ++e) S[e] = 0.0;
        int64_t kk = 0;
        for (int64_t jj = 0; jj < nsymm; ++jj)
            if (emap[jj + nsymm_rot * i] == ii) {
                for (int e = 0; e < 9; ++e) S[e] += crot_orig[9 * jj + e];
                kk++;
            }
        if (kk > 1)
            for (int e = 0; e < 9; ++e) S[e] /= (double)kk;
    }
    free(emap);
}

/* interactions.f90:1673-1930 calculate_eph_interaction_ibzq, key 'g'.
 * Phonon states are (branch s, IBZ q of the phonon mesh). */
int orc_ibzq_g2(const orc_wannier *w, const orc_electrons *el, const orc_phonons *ph, int64_t istate_first,
                int64_t istate_last, int64_t *out_count, double *out_g2, int nthreads) {
    const int64_t nb = w->nb, nbr = w->nbr, nwv = el->nwv, nq = ph->nwv;
    const int64_t stride = el->nstates_inwindow * nb;
    const int64_t gsz = nb * nb * nbr * w->nwsk;
    double rec[9], vol;
    orc_reclattvecs(w->lattvecs, rec, &vol);
    int64_t nst = istate_last - istate_first + 1;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
    {
        zdouble *gq = malloc(sizeof(zdouble) * gsz);
        int64_t cur_iq = -1;
#pragma omp for schedule(dynamic, 1)
        for (int64_t t = 0; t < nst; ++t) {
            int64_t istate = istate_first + t, s, iq;
            orc_demux_state(istate, nbr, &s, &iq);
            int64_t iq_fbz = ph->indexlist_irred[iq - 1];
            double en_ph = ph->ens[(iq_fbz - 1) + nq * (s - 1)];
            (void)en_ph;
            orc_vec qv, kv, kpv;
            orc_vec_create(iq_fbz, ph->wvmesh, rec, &qv);
            if (cur_iq != iq) { orc_gReq(w, qv.frac, gq); cur_iq = iq; }
            int64_t count = 0;
            double *row = &out_g2[t * stride];
            for (int64_t ik = 1; ik <= nwv; ++ik) {
                orc_vec_create(el->indexlist[ik - 1], el->wvmesh, rec, &kv);
                orc_vec_add(&kv, &qv, el->wvmesh, rec, &kpv);
                int64_t ikpw = orc_binsearch(el->indexlist, nwv, kpv.muxed);
                if (ikpw < 0) continue;
                for (int64_t m = 1; m <= nb; ++m) {
                    if (fabs(el->ens[(ik - 1) + nwv * (m - 1)] - el->enref) > el->fsthick) continue;
                    for (int64_t n = 1; n <= nb; ++n) {
                        if (fabs(el->ens[(ikpw - 1) + nwv * (n - 1)] - el->enref) > el->fsthick) continue;
                        count++;
                        row[count - 1] = orc_g2(w, kv.frac, qv.frac, &el->evecs[(ik - 1) + nwv * (m - 1)], nwv * nb,
                                                &el->evecs[(ikpw - 1) + nwv * (n - 1)], nwv * nb,
                                                &ph->evecs[(iq_fbz - 1) + nq * (s - 1)], nq * nbr,
                                                ph->ens[(iq_fbz - 1) + nq * (s - 1)], gq, 0);
                    }
                }
            }
            out_count[t] = count;
        }
        free(gq);
    }
    return 0;
}

/* ======================================================================
 * Consumers of g2: tetrahedron delta function, transition probabilities
 * X+/- and the electron RTA scattering rates.
 * ====================================================================== */

/* delta.f90:494-605 form_tetrahedra_3d.  tetra (6 nk, 4), tetracount (nk),
 * tetramap (2, nk, 24), column-major, 1-based indices; with blocks != 0 the
 * wave vectors are indexlist(1:nk) (muxed, sorted) and a vertex outside the
 * list is stored as -1 (binsearch's miss value). */
void orc_form_tetrahedra_3d(int64_t nk, const int64_t mesh[3], int blocks, const int64_t *indexlist, int64_t *tetra,
                            int64_t *tetracount, int64_t *tetramap) {
    static const int64_t labels[6][4] = {{1, 2, 3, 6}, {1, 3, 5, 6}, {3, 5, 6, 7},
                                         {3, 6, 7, 8}, {3, 4, 6, 8}, {2, 3, 4, 6}};
    const int64_t ntet = 6 * nk;
    memset(tetra, 0, sizeof(int64_t) * ntet * 4);
    memset(tetracount, 0, sizeof(int64_t) * nk);
    memset(tetramap, 0, sizeof(int64_t) * 2 * nk * 24);
    int64_t count = 1;
    for (int64_t ik = 1; ik <= nk; ++ik) {
        int64_t ijk[3];
        orc_demux_vector(blocks ? indexlist[ik - 1] : ik, ijk, mesh, 1);
        const int64_t i = ijk[0], j = ijk[1], k = ijk[2];
        const int64_t ip1 = i == mesh[0] ? 1 : i + 1;
        const int64_t jp1 = j == mesh[1] ? 1 : j + 1;
        const int64_t kp1 = k == mesh[2] ? 1 : k + 1;
        const int64_t sc[8][3] = {{i, j, k},     {ip1, j, k},     {i, jp1, k},     {ip1, jp1, k},
                                  {i, j, kp1},   {ip1, j, kp1},   {i, jp1, kp1},   {ip1, jp1, kp1}};
        for (int tk = 0; tk < 6; ++tk) {
            for (int tl = 1; tl <= 4; ++tl) {
                const int64_t *v = sc[labels[tk][tl - 1] - 1];
                const int64_t aux = orc_mux_vector(v, mesh, 1);
                int64_t tmp = aux;
                if (blocks) tmp = orc_binsearch(indexlist, nk, aux);
                tetra[(count - 1) + ntet * (tl - 1)] = tmp;
                if (tmp > 0) {
                    const int64_t c = ++tetracount[tmp - 1];
                    tetramap[0 + 2 * ((tmp - 1) + nk * (c - 1))] = count;
                    tetramap[1 + 2 * ((tmp - 1) + nk * (c - 1))] = tl;
                }
            }
            count++;
        }
    }
}

/* delta.f90:607-646 fill_tetrahedra_3d: tetra_evals (ntet, nb, 4), the band
 * energies evals (nk, nb) at the four vertices (0 for a vertex outside the
 * list), sorted ascending over the vertices. */
void orc_fill_tetrahedra_3d(int64_t ntet, const int64_t *tetra, int64_t nk, int64_t nb, const double *evals,
                            double *tetra_evals) {
    memset(tetra_evals, 0, sizeof(double) * ntet * nb * 4);
    for (int64_t it = 0; it < ntet; ++it)
        for (int iv = 0; iv < 4; ++iv) {
            const int64_t aux = tetra[it + ntet * iv];
            if (aux > 0)
                for (int64_t ib = 0; ib < nb; ++ib) tetra_evals[it + ntet * (ib + nb * iv)] = evals[(aux - 1) + nk * ib];
        }
    for (int64_t it = 0; it < ntet; ++it)
        for (int64_t ib = 0; ib < nb; ++ib) {
            double *e[4];
            for (int iv = 0; iv < 4; ++iv) e[iv] = &tetra_evals[it + ntet * (ib + nb * iv)];
            /* misc.f90:716-736 sort_real (exchange sort, ascending) */
            for (int a = 0; a < 4; ++a) {
                double aux = *e[a];
                for (int b = a + 1; b < 4; ++b)
                    if (aux > *e[b]) {
                        const double t = *e[b];
                        *e[b] = aux;
                        *e[a] = t;
                        aux = t;
                    }
            }
        }
}

/* delta.f90:61-225 delta_fn_tetra, the vertex weight of the linear
 * tetrahedron method for sample energy e at wave vector ik (1-based, of the
 * list the tetrahedra were formed on) and band ib.  The case selector iv is
 * the wave vector's position in the tetrahedron as formed (tetramap(2,:,:)),
 * the energies e1 <= e2 <= e3 <= e4 are the sorted vertex energies: both as
 * the reference has them.  numtetra = 6 product(mesh). */
double orc_delta_fn_tetra(double e, int64_t ik, int64_t ib, const int64_t mesh[3], int64_t nk, int64_t ntet,
                          int64_t nb, const int64_t *tetramap, const int64_t *tetracount, const double *tetra_evals) {
    double tmp = 0.0, d = 0.0;
    const int64_t numtetra = mesh[0] * mesh[1] * mesh[2] * 6;
    const int64_t num = tetracount[ik - 1];
    for (int64_t itk = 1; itk <= num; ++itk) {
        const int64_t it = tetramap[0 + 2 * ((ik - 1) + nk * (itk - 1))];
        const int64_t iv = tetramap[1 + 2 * ((ik - 1) + nk * (itk - 1))];
        const double e1 = tetra_evals[(it - 1) + ntet * ((ib - 1) + nb * 0)];
        const double e2 = tetra_evals[(it - 1) + ntet * ((ib - 1) + nb * 1)];
        const double e3 = tetra_evals[(it - 1) + ntet * ((ib - 1) + nb * 2)];
        const double e4 = tetra_evals[(it - 1) + ntet * ((ib - 1) + nb * 3)];
        const double e1e = e1 - e, e2e = e2 - e, e3e = e3 - e, e4e = e4 - e;
        const double e21 = e2 - e1, e31 = e3 - e1, e41 = e4 - e1, e32 = e3 - e2, e42 = e4 - e2, e43 = e4 - e3;
        const int c1 = e1 <= e && e <= e2, c2 = e2 <= e && e <= e3, c3 = e3 <= e && e <= e4;
        if (e < e1 || e > e4) continue;
        switch (iv) {
        case 1:
            if (c1) {
                tmp = (e2e / e21 + e3e / e31 + e4e / e41) * (e1e * e1e) / e41 / e31 / e21;
                if (e1 == e2) tmp = 0.0;
            } else if (c2) {
                tmp = -0.5 * (e3e / (e31 * e31) * (e3e * e2e / e42 / e32 + e4e * e1e / e41 / e42 + e3e * e1e / e32 / e41) +
                              e4e / (e41 * e41) * (e4e * e1e / e42 / e31 + e4e * e2e / e42 / e32 + e3e * e1e / e31 / e32));
                if (e2 == e3)
                    tmp = -0.5 * (e4e * e1e / e41 / e42 + e1e / e41 +
                                  e4e / (e41 * e41) * (e4e * e1e / e42 / e31 + e4e / e42 + e1e / e31));
            } else if (c3) {
                tmp = (e4e * e4e * e4e) / (e41 * e41) / e42 / e43;
                if (e3 == e4) tmp = (e4e * e4e) / (e41 * e41) / e42;
            }
            break;
        case 2:
            if (c1) {
                tmp = -(e1e * e1e * e1e) / (e21 * e21) / e31 / e41;
                if (e1 == e2) tmp = 0.0;
            } else if (c2) {
                tmp = -0.5 * (e3e / (e32 * e32) * (e3e * e2e / e42 / e31 + e4e * e2e / e42 / e41 + e3e * e1e / e31 / e41) +
                              e4e / (e42 * e42) * (e3e * e2e / e32 / e31 + e4e * e1e / e41 / e31 + e4e * e2e / e32 / e41));
                if (e2 == e3) tmp = -0.5 * (0.0 + e4e / e42 / e41 + 0.0 + e4e / (e42 * e42) * (0.0 + e4e * e1e / e41 / e31 + 1.0));
            } else if (c3) {
                tmp = (e4e * e4e * e4e) / e41 / (e42 * e42) / e43;
                if (e3 == e4) tmp = 0.0;
            }
            break;
        case 3:
            if (c1) {
                tmp = -(e1e * e1e * e1e) / e21 / (e31 * e31) / e41;
                if (e1 == e2) tmp = 0.0;
            } else if (c2) {
                tmp = 0.5 * (e2e / (e32 * e32) * (e3e * e2e / e42 / e31 + e4e * e2e / e42 / e41 + e3e * e1e / e31 / e41) +
                             e1e / (e31 * e31) * (e3e * e2e / e42 / e32 + e4e * e1e / e41 / e42 + e3e * e1e / e32 / e41));
                if (e2 == e3)
                    tmp = 0.5 * (0.0 + e4e / e42 / e41 + e1e / e31 / e41 +
                                 e1e / (e31 * e31) * (0.0 + e4e * e1e / e41 / e42 + e1e / e41));
            } else if (c3) {
                tmp = (e4e * e4e * e4e) / e41 / e42 / (e43 * e43);
                if (e3 == e4) tmp = 0.0;
            }
            break;
        case 4:
            if (c1) {
                tmp = -(e1e * e1e * e1e) / e21 / e31 / (e41 * e41);
                if (e1 == e2) tmp = 0.0;
            } else if (c2) {
                tmp = 0.5 * (e2e / (e42 * e42) * (e3e * e2e / e32 / e31 + e4e * e1e / e41 / e31 + e4e * e2e / e32 / e41) +
                             e1e / (e41 * e41) * (e4e * e1e / e42 / e31 + e4e * e2e / e42 / e32 + e3e * e1e / e31 / e32));
                if (e2 == e3) tmp = 0.5 * (0.0 + e1e / (e41 * e41) * (e4e * e1e / e42 / e31 + e4e / e42 + e1e / e31));
            } else if (c3) {
                tmp = -(e3e / e43 + e2e / e42 + e1e / e41) * (e4e * e4e) / e41 / e42 / e43;
                if (e3 == e4) tmp = 0.0;
            }
            break;
        }
        if (e1 == e2 && e1 == e3 && e1 == e4 && e == e1) tmp = 0.25;
        d += tmp;
    }
    if (d < 1.0e-12) d = 0.0;
    return d / (double)numtetra;
}

/* misc.f90:992-1019 Bose (with the quad-precision expm1) and Fermi; kB from
 * params.f90:39. */
static const double KB_EV = 1.380649e-23 / 1.602176634e-19;
double orc_bose(double e, double T) {
    const __float128 x = (__float128)(e / KB_EV / T);
    const double em1 = (double)(expq(x) - (__float128)1.0);
    return 1.0 / em1;
}
double orc_fermi(double e, double chempot, double T) { return 1.0 / (exp((e - chempot) / KB_EV / T) + 1.0); }

/* interactions.f90:2190-2218, 2228-2251, 2278-2332, 2351-2354
 * calculate_eph_interaction_ibzk, key 'X': for IBZ states
 * istate_first..istate_last with in-window counts (count -1 = outside the
 * window: rows untouched), the transition probabilities (THz)
 *   X+ = 2pi/hbar g2 (n_B + f(e + w)) delta(e + w; k', n),
 *   X- = 2pi/hbar g2 (1 + n_B - f(e - w)) delta(e - w; k', n),
 * zero for phonon energies below 0.5 meV, and the final electron and
 * phonon state of each process (fine-mesh phonons negative).  g2 rows have
 * stride nprocs = nstates_inwindow nbr, as orc_ibzk_g2 writes them. */
/* delta_fn_ptr of get_delta_fn_pointer (delta.f90:47-59): tetrahedra
 * (nvert 4) or triangles (nvert 3). */
static double orc_delta_simplex(int nvert, double e, int64_t ik, int64_t ib, const int64_t mesh[3], int64_t nk,
                                int64_t nb, const int64_t *map, const int64_t *count, const double *evals) {
    if (nvert == 3) return orc_delta_fn_triang(e, ik, ib, mesh, nk, 2 * nk, nb, map, count, evals);
    return orc_delta_fn_tetra(e, ik, ib, mesh, nk, 6 * nk, nb, map, count, evals);
}

int orc_ibzk_X(const orc_wannier *w, const orc_electrons *el, const orc_phonons *ph, int64_t mesh_ref,
               int64_t istate_first, int64_t istate_last, const int64_t *counts, const double *g2, double T,
               double chempot, int nvert, const int64_t *tetramap, const int64_t *tetracount,
               const double *tetra_evals, double *Xplus, double *Xminus, int64_t *istate_el, int64_t *istate_ph,
               int nthreads) {
    const int64_t nb = w->nb, nbr = w->nbr, nirr = el->nwv_irred, nwv = el->nwv;
    const int64_t nprocs = el->nstates_inwindow * nbr;
    const double cnst = TWOPI_ / (HBAR / QE);
    double rec[9], vol;
    orc_reclattvecs(w->lattvecs, rec, &vol);
    const int64_t nst = istate_last - istate_first + 1;
    int status = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
    {
        double phe[64];
        zdouble phv[64 * 64];
#pragma omp for schedule(dynamic, 1)
        for (int64_t t = 0; t < nst; ++t) {
            if (counts[t] < 0) continue;
            const int64_t istate = istate_first + t;
            int64_t m, ik;
            orc_demux_state(istate, nb, &m, &ik);
            const double en_el = el->ens_irred[(ik - 1) + nirr * (m - 1)];
            orc_vec kv, kpv, qv, qc;
            orc_vec_create(el->indexlist_irred[ik - 1], el->wvmesh, rec, &kv);
            const double *g = &g2[t * nprocs];
            double *xp = &Xplus[t * nprocs], *xm = &Xminus[t * nprocs];
            int64_t *se = &istate_el[t * nprocs], *sp = &istate_ph[t * nprocs];
            int64_t count = 0;
            for (int64_t ikp = 1; ikp <= nwv; ++ikp) {
                orc_vec_create(el->indexlist[ikp - 1], el->wvmesh, rec, &kpv);
                orc_vec_sub(&kpv, &kv, el->wvmesh, rec, &qv);
                int need_fine = 0;
                for (int c = 0; c < 3; ++c)
                    if (qv.ints[c] % mesh_ref != 0) need_fine = 1;
                int64_t iqc = 0;
                if (need_fine) {
                    double ql[3] = {qv.frac[0], qv.frac[1], qv.frac[2]};
                    if (orc_ph_wann(w, 1, ql, phe, phv, NULL) == 2) status = 2;
                } else {
                    orc_vec_change_grid(&qv, ph->wvmesh, &qc);
                    iqc = qc.muxed;
                }
                for (int64_t n = 1; n <= nb; ++n) {
                    if (fabs(el->ens[(ikp - 1) + nwv * (n - 1)] - el->enref) > el->fsthick) continue;
                    for (int64_t s = 1; s <= nbr; ++s) {
                        count++;
                        if (count > nprocs) { status = 5; continue; }
                        const double en_ph = need_fine ? phe[s - 1] : ph->ens[(iqc - 1) + ph->nwv * (s - 1)];
                        const double bosefac = en_ph != 0.0 ? orc_bose(en_ph, T) : 0.0;
                        const double fp = orc_fermi(en_el + en_ph, chempot, T);
                        const double fm = orc_fermi(en_el - en_ph, chempot, T);
                        double delta = orc_delta_simplex(nvert, en_el + en_ph, ikp, n, el->wvmesh, nwv, nb, tetramap,
                                                         tetracount, tetra_evals);
                        double x = 0.0;
                        if (en_ph >= 0.5e-3) x = g[count - 1] * (bosefac + fp) * delta;
                        xp[count - 1] = cnst * x;
                        delta = orc_delta_simplex(nvert, en_el - en_ph, ikp, n, el->wvmesh, nwv, nb, tetramap,
                                                  tetracount, tetra_evals);
                        x = 0.0;
                        if (en_ph >= 0.5e-3) x = g[count - 1] * (1.0 + bosefac - fm) * delta;
                        xm[count - 1] = cnst * x;
                        se[count - 1] = orc_mux_state(nb, n, ikp);
                        sp[count - 1] = need_fine ? -orc_mux_state(nbr, s, qv.muxed) : orc_mux_state(nbr, s, iqc);
                    }
                }
            }
            if (count != counts[t]) status = 6;
        }
    }
    return status;
}

/* interactions.f90:2957-3063 calculate_el_rta_rates, e-ph channel: the
 * scattering rate of each in-window IBZ state is sum X+ + sum X- over its
 * processes (THz), each sum in process order.  rates has one entry per
 * state of istate_first..istate_last (0 outside the window). */
void orc_el_rta_rates(int64_t nst, const int64_t *counts, int64_t nprocs, const double *Xplus, const double *Xminus,
                      double *rates) {
    for (int64_t t = 0; t < nst; ++t) {
        double r = 0.0;
        for (int64_t i = 0; i < counts[t]; ++i) r = r + Xplus[t * nprocs + i];
        for (int64_t i = 0; i < counts[t]; ++i) r = r + Xminus[t * nprocs + i];
        rates[t] = r;
    }
}

/* bte.f90:912-1011 calculate_field_term, electrons (species 'el'): field 'E'
 * (A = qe, pow = 0) or 'T' (A = 1/T, pow = 1).  ibz2fbz_map (maxeq, nk_ibz, 2)
 * column-major, (:, :, 2) the muxed FBZ index, located in indexlist (nk) by
 * binsearch; rates_ibz (nk_ibz, nb); field_term (nk, nb, 3) is overwritten
 * (zero where no IBZ image lands or the rate is 0). */
void orc_field_term(char field, const int64_t *nequiv, int64_t nk_ibz, int64_t maxeq, const int64_t *ibz2fbz_map,
                    double T, double chempot, int64_t nk, int64_t nb, const double *ens, const double *vels,
                    const double *rates_ibz, const int64_t *indexlist, double *field_term) {
    const double A = field == 'T' ? 1.0 / T : QE;
    const int pow_ = field == 'T' ? 1 : 0;
    memset(field_term, 0, sizeof(double) * nk * nb * 3);
    for (int64_t i = 0; i < nk_ibz; ++i)
        for (int64_t ieq = 0; ieq < nequiv[i]; ++ieq) {
            const int64_t ik = orc_binsearch(indexlist, nk, ibz2fbz_map[ieq + maxeq * (i + nk_ibz * 1)]);
            for (int64_t ib = 0; ib < nb; ++ib) {
                const double r = rates_ibz[i + nk_ibz * ib];
                if (r == 0.0) continue;
                const double e = ens[(ik - 1) + nk * ib];
                const double ep = pow_ == 1 ? (e - chempot) : 1.0;
                for (int d = 0; d < 3; ++d)
                    field_term[(ik - 1) + nk * (ib + nb * d)] = A * vels[(ik - 1) + nk * (ib + nb * d)] * ep / r;
            }
        }
}

/* symmetry.f90:496-513 symmetrize_3x3_tensor: t = sum_r C_r t C_r^T / nrot. */
static void orc_symmetrize_3x3(double t[9], const double *crot, int64_t nrot) {
    double aux[9] = {0};
    for (int64_t r = 0; r < nrot; ++r) {
        const double *C = &crot[9 * r];
        double m[9];
        for (int i = 0; i < 3; ++i) /* m = t C^T */
            for (int j = 0; j < 3; ++j) {
                double s = 0.0;
                for (int k = 0; k < 3; ++k) s += t[i + 3 * k] * C[j + 3 * k];
                m[i + 3 * j] = s;
            }
        for (int i = 0; i < 3; ++i) /* C m */
            for (int j = 0; j < 3; ++j) {
                double s = 0.0;
                for (int k = 0; k < 3; ++k) s += C[i + 3 * k] * m[k + 3 * j];
                aux[i + 3 * j] += s;
            }
    }
    for (int i = 0; i < 9; ++i) t[i] = aux[i] / (double)nrot;
}

/* bz_sums.f90:484-611 calculate_transport_coeff for electrons ('el'), field
 * 'E' or 'T': trans_coeff_hc and trans_coeff_cc (nb, 3, 3) column-major,
 * symmetrized with the nrot Cartesian rotations crot (3,3,nrot) when nrot >
 * 0.  The literal 1.0e21 of the source is a default-kind (single precision)
 * real, so its value is the float nearest 1e21. */
void orc_transport_coeff(char field, double T, int64_t deg, double chempot, int64_t nk, int64_t nb, const double *ens,
                         const double *vels, double volume, const int64_t mesh[3], const double *response,
                         int64_t nrot, const double *crot, double *hc, double *cc) {
    const double fac = (double)1.0e21f / KB_EV / T / volume / (double)(mesh[0] * mesh[1] * mesh[2]);
    double A_hc, A_cc;
    int pow_hc, pow_cc;
    if (field == 'T') {
        A_cc = -deg * QE * fac;
        pow_cc = 0;
        A_hc = deg * QE * fac;
        pow_hc = 1;
    } else {
        A_cc = deg * fac;
        pow_cc = 0;
        A_hc = -A_cc;
        pow_hc = 1;
    }
    memset(hc, 0, sizeof(double) * nb * 9);
    memset(cc, 0, sizeof(double) * nb * 9);
    for (int64_t ik = 0; ik < nk; ++ik)
        for (int64_t ib = 0; ib < nb; ++ib) {
            const double e = ens[ik + nk * ib];
            double dist = orc_fermi(e, chempot, T);
            dist = dist * (1.0 - dist);
            const double p_hc = pow_hc == 1 ? (e - chempot) : 1.0, p_cc = pow_cc == 1 ? (e - chempot) : 1.0;
            for (int j = 0; j < 3; ++j) {
                const double v = vels[ik + nk * (ib + nb * j)];
                for (int l = 0; l < 3; ++l) {
                    const double r = response[ik + nk * (ib + nb * l)];
                    hc[ib + nb * (j + 3 * l)] += p_hc * dist * v * r;
                    if (A_cc != 0.0) cc[ib + nb * (j + 3 * l)] += p_cc * dist * v * r;
                }
            }
        }
    for (int64_t i = 0; i < nb * 9; ++i) {
        hc[i] = A_hc * hc[i];
        if (A_cc != 0.0) cc[i] = A_cc * cc[i];
    }
    if (nrot > 0)
        for (int64_t ib = 0; ib < nb; ++ib) {
            double t[9], u[9];
            for (int i = 0; i < 9; ++i) { t[i] = hc[ib + nb * i]; u[i] = cc[ib + nb * i]; }
            orc_symmetrize_3x3(t, crot, nrot);
            if (A_cc != 0.0) orc_symmetrize_3x3(u, crot, nrot);
            for (int i = 0; i < 9; ++i) { hc[ib + nb * i] = t[i]; cc[ib + nb * i] = u[i]; }
        }
}

/* bte.f90:241-244 (and 262-265): field_term(ik, :, :) =
 * transpose(matmul(symmetrizers(:, :, ik), transpose(field_term(ik, :, :)))),
 * i.e. F(ik, ib, :) <- S_ik F(ik, ib, :); symmetrizers (3, 3, nk). */
void orc_symmetrize_field_term(int64_t nk, int64_t nb, const double *symmetrizers, double *field_term) {
    for (int64_t ik = 0; ik < nk; ++ik) {
        const double *S = &symmetrizers[9 * ik];
        for (int64_t ib = 0; ib < nb; ++ib) {
            double f[3], g[3];
            for (int d = 0; d < 3; ++d) f[d] = field_term[ik + nk * (ib + nb * d)];
            for (int i = 0; i < 3; ++i) g[i] = S[i] * f[0] + S[i + 3] * f[1] + S[i + 6] * f[2];
            for (int d = 0; d < 3; ++d) field_term[ik + nk * (ib + nb * d)] = g[d];
        }
    }
}

/* delta.f90:648-737 form_triangles: two triangles per subcell of each z
 * layer.  triang (2 nk, 3), triangcount (nk), triangmap (2, nk, 6), 1-based;
 * with blocks != 0 a vertex off indexlist is stored as -1. */
void orc_form_triangles(int64_t nk, const int64_t mesh[3], int blocks, const int64_t *indexlist, int64_t *triang,
                        int64_t *triangcount, int64_t *triangmap) {
    static const int64_t labels[2][3] = {{1, 2, 3}, {1, 3, 4}};
    const int64_t ntri = 2 * nk;
    memset(triang, 0, sizeof(int64_t) * ntri * 3);
    memset(triangcount, 0, sizeof(int64_t) * nk);
    memset(triangmap, 0, sizeof(int64_t) * 2 * nk * 6);
    int64_t count = 1;
    for (int64_t ik = 1; ik <= nk; ++ik) {
        int64_t ijk[3];
        orc_demux_vector(blocks ? indexlist[ik - 1] : ik, ijk, mesh, 1);
        const int64_t i = ijk[0], j = ijk[1], k = ijk[2];
        const int64_t ip1 = i == mesh[0] ? 1 : i + 1;
        const int64_t jp1 = j == mesh[1] ? 1 : j + 1;
        const int64_t sc[4][3] = {{i, j, k}, {ip1, j, k}, {i, jp1, k}, {ip1, jp1, k}};
        for (int tk = 0; tk < 2; ++tk) {
            for (int tl = 1; tl <= 3; ++tl) {
                const int64_t *v = sc[labels[tk][tl - 1] - 1];
                const int64_t aux = orc_mux_vector(v, mesh, 1);
                int64_t tmp = aux;
                if (blocks) tmp = orc_binsearch(indexlist, nk, aux);
                triang[(count - 1) + ntri * (tl - 1)] = tmp;
                if (tmp > 0) {
                    const int64_t c = ++triangcount[tmp - 1];
                    triangmap[0 + 2 * ((tmp - 1) + nk * (c - 1))] = count;
                    triangmap[1 + 2 * ((tmp - 1) + nk * (c - 1))] = tl;
                }
            }
            count++;
        }
    }
}

/* delta.f90:739-788 fill_triangles: triang_evals (ntri, nb, 3), 0 off the
 * list, sorted ascending over the vertices. */
void orc_fill_triangles(int64_t ntri, const int64_t *triang, int64_t nk, int64_t nb, const double *evals,
                        double *triang_evals) {
    memset(triang_evals, 0, sizeof(double) * ntri * nb * 3);
    for (int64_t it = 0; it < ntri; ++it)
        for (int iv = 0; iv < 3; ++iv) {
            const int64_t aux = triang[it + ntri * iv];
            if (aux > 0)
                for (int64_t ib = 0; ib < nb; ++ib) triang_evals[it + ntri * (ib + nb * iv)] = evals[(aux - 1) + nk * ib];
        }
    for (int64_t it = 0; it < ntri; ++it)
        for (int64_t ib = 0; ib < nb; ++ib) {
            double *e[3];
            for (int iv = 0; iv < 3; ++iv) e[iv] = &triang_evals[it + ntri * (ib + nb * iv)];
            for (int a = 0; a < 3; ++a) { /* misc.f90:716-736 sort_real */
                double aux = *e[a];
                for (int b = a + 1; b < 3; ++b)
                    if (aux > *e[b]) {
                        const double t = *e[b];
                        *e[b] = aux;
                        *e[a] = t;
                        aux = t;
                    }
            }
        }
}

/* delta.f90:790-882 delta_fn_triang (triangle method of Kurganskii et al.):
 * the weight of the wave vector's vertex label iv with sorted energies
 * e1 <= e2 <= e3, summed over its triangles, zero below 1e-12, divided by
 * numtriangs = 2 product(mesh). */
double orc_delta_fn_triang(double e, int64_t ik, int64_t ib, const int64_t mesh[3], int64_t nk, int64_t ntri,
                           int64_t nb, const int64_t *triangmap, const int64_t *triangcount,
                           const double *triang_evals) {
    double d = 0.0;
    const int64_t numtriangs = mesh[0] * mesh[1] * mesh[2] * 2;
    const int64_t num = triangcount[ik - 1];
    for (int64_t itk = 1; itk <= num; ++itk) {
        const int64_t it = triangmap[0 + 2 * ((ik - 1) + nk * (itk - 1))];
        const int64_t iv = triangmap[1 + 2 * ((ik - 1) + nk * (itk - 1))];
        const double e1 = triang_evals[(it - 1) + ntri * ((ib - 1) + nb * 0)];
        const double e2 = triang_evals[(it - 1) + ntri * ((ib - 1) + nb * 1)];
        const double e3 = triang_evals[(it - 1) + ntri * ((ib - 1) + nb * 2)];
        const int c1 = e <= e1, c2 = e1 < e && e <= e2, c3 = e2 < e && e <= e3, c4 = e3 < e;
        double tmp = 0.0;
        if (c1 || c4) continue;
        const double E12 = (e - e2) / (e1 - e2), E21 = (e - e1) / (e2 - e1), E13 = (e - e3) / (e1 - e3);
        const double E31 = (e - e1) / (e3 - e1), E23 = (e - e3) / (e2 - e3), E32 = (e - e2) / (e3 - e2);
        switch (iv) {
        case 1:
            if (c2) tmp = E21 * (E12 + E13) / (e3 - e1);
            else if (c3) tmp = E23 * E13 / (e3 - e1);
            break;
        case 2:
            if (c2) tmp = E21 * E21 / (e3 - e1);
            else if (c3) tmp = E23 * E23 / (e3 - e1);
            break;
        case 3:
            if (c2) tmp = E21 * E31 / (e3 - e1);
            else if (c3) tmp = E23 * (E31 + E32) / (e3 - e1);
            break;
        }
        d += tmp;
    }
    if (d < 1.0e-12) d = 0.0;
    return d / (double)numtriangs;
}
