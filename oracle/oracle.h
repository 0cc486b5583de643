/* CPU oracle: a literal FP64 restatement of elphbolt's electron-phonon vertex
 * path, used only as the checker in tests, smoke() and the CPU-baseline leg of
 * bench.py.  It is not part of the product and is never linked into it.
 *
 * All arrays are column-major (Fortran order) with the shapes given in the
 * reference; indices that the reference returns 1-based stay 1-based.
 */
#ifndef ELPH_ORACLE_H
#define ELPH_ORACLE_H
#include <stdint.h>
#include <complex.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef double _Complex zdouble;

typedef struct {
    int64_t nb, nbr, nwsk, nwsq, nwsg, natoms;
    int64_t coarse_qmesh[3];
    const int64_t *rcells_k, *rcells_q, *rcells_g; /* (nws,3) */
    const int64_t *elwsdeg, *phwsdeg, *gwsdeg;    /* (nws) */
    const zdouble *Hwann;   /* (nwsk,nb,nb) */
    const zdouble *Dphwann; /* (nwsq,nbr,nbr) */
    const zdouble *gwann;   /* (nb,nb,nwsk,nbr,nwsg) */
    double lattvecs[9];     /* (3,3) nm, columns */
    const double *basis;    /* (3,natoms) fractional */
    const double *masses;   /* (natoms) amu, per atom */
    int32_t polar;
    double epsilon[9];      /* (3,3) */
    const double *born;     /* (3,3,natoms) */
} orc_wannier;

typedef struct {
    double frac[3], cart[3];
    int64_t ints[3];
    int64_t muxed;
} orc_vec;

/* ---- misc.f90 / vector.f90 ---- */
void orc_expi(double x, double out[2]);
void orc_add_and_fold(const double q1[3], const double q2[3], double q3[3]);
int64_t orc_mux_vector(const int64_t v[3], const int64_t mesh[3], int64_t base);
void orc_demux_vector(int64_t i, int64_t v[3], const int64_t mesh[3], int64_t base);
int64_t orc_mux_state(int64_t nbands, int64_t iband, int64_t ik);
void orc_demux_state(int64_t m, int64_t nbands, int64_t *iband, int64_t *ik);
int64_t orc_binsearch(const int64_t *array, int64_t n, int64_t e);
void orc_distribute_points(int64_t npts, int64_t num_images, int64_t this_image, int64_t *chunk,
                           int64_t *istart, int64_t *iend, int64_t *num_active);
void orc_vec_create(int64_t ind, const int64_t grid[3], const double prim[9], orc_vec *v);
void orc_vec_add(const orc_vec *v1, const orc_vec *v2, const int64_t grid[3], const double prim[9], orc_vec *v3);
void orc_vec_sub(const orc_vec *v1, const orc_vec *v2, const int64_t grid[3], const double prim[9], orc_vec *v3);
void orc_vec_change_grid(const orc_vec *vin, const int64_t grid[3], orc_vec *vout);
void orc_reclattvecs(const double lattvecs[9], double rec[9], double *volume);

/* ---- LAPACK zheev('V','U') semantics: ascending eigenvalues, eigenvectors in
 * columns; only the upper triangle of a is read. ---- */
int orc_zheev(int64_t n, zdouble *a, int64_t lda, double *w);

/* ---- wannier.f90 ---- */
int orc_el_wann(const orc_wannier *w, int64_t nk, const double *kvecs, double *energies,
                double *velocities, zdouble *evecs, const double *scissor);
int orc_ws_wrapping_vectors(const double lvec[9], const double vl[3], double eps, int32_t *wrappers,
                            int32_t maxw);
int orc_ph_wann(const orc_wannier *w, int64_t nq, const double *qvecs, double *energies, zdouble *evecs,
                double *velocities);
void orc_dyn_nonanalytic(const orc_wannier *w, const double q[3], zdouble *dyn_l, zdouble *ddyn_l);
double orc_g2(const orc_wannier *w, const double kvec[3], const double qvec[3], const zdouble *el_evec_k,
              int64_t stride_k, const zdouble *el_evec_kp, int64_t stride_kp, const zdouble *ph_evec_q,
              int64_t stride_q, double ph_en, const zdouble *gmixed, int wannspace_ph);
void orc_long_range_prefac(const orc_wannier *w, const double q[3], const zdouble *uqs, zdouble *glprefac);
void orc_gkRp(const orc_wannier *w, const double kvec[3], zdouble *gmixed);
void orc_gReq(const orc_wannier *w, const double qvec[3], zdouble *gmixed);

/* ---- symmetry.f90 ---- */
int orc_find_rotations(const double lattvecs[9], const double *basis, const int64_t *types, int64_t natoms,
                       double symprec, int32_t *rot_out /* (3,3,48) */);
/* qrot_out (3,3,2*nsymm); crot_out (the Cartesian operations in the same
 * filtered order) may be NULL */
int64_t orc_qrotations(const double lattvecs[9], const int32_t *rot, int64_t nsymm, const int64_t mesh[3],
                       double *qrot_out, double *crot_out);
void orc_find_equiv_map(int64_t nsymm_rot, int64_t *equiv_map, const int64_t mesh[3], const double *qrot,
                        const int64_t *indexlist, int64_t nlist);
int64_t orc_find_irred_wedge(const int64_t mesh[3], int64_t nsymm_rot, const double *qrot,
                             const int64_t *indexlist, int64_t nlist, int64_t *indexlist_irred,
                             int64_t *nequiv, int64_t *ibz2fbz /* (nsymm_rot, nlist, 2) */);

void orc_cartesian_rotations(const double L[9], const int32_t *rot, int64_t nsymm, double *crot_out);
void orc_symmetrizers(const int64_t mesh[3], int64_t nsymm_rot, const double *qrot, int64_t nsymm,
                      const double *crot_orig, int64_t nk, const int64_t *indexlist, double *out);

/* ---- electron.f90 calculate_electrons (vertex-relevant subset) ---- */
typedef struct {
    int64_t wvmesh[3];
    double enref, fsthick;
    int64_t nwv_irred, nwv, nstates_inwindow;
    int64_t *indexlist_irred;  /* (nwv_irred) */
    double *ens_irred;         /* (nwv_irred, nb) */
    zdouble *evecs_irred;      /* (nwv_irred, nb, nb) */
    int64_t *indexlist;        /* (nwv) */
    double *ens;               /* (nwv, nb) */
    zdouble *evecs;            /* (nwv, nb, nb) */
} orc_electrons;
int orc_calculate_electrons(const orc_wannier *w, const double *qrot, int64_t nsymm_rot,
                            const int64_t wvmesh[3], double enref, double fsthick, orc_electrons *el);
void orc_free_electrons(orc_electrons *el);

typedef struct {
    int64_t wvmesh[3];
    int64_t nwv;
    double *ens;     /* (nwv, nbr) */
    zdouble *evecs;  /* (nwv, nbr, nbr) */
    int64_t nwv_irred;
    int64_t *indexlist_irred;
} orc_phonons;
int orc_calculate_phonons(const orc_wannier *w, const double *qrot, int64_t nsymm_rot,
                          const int64_t wvmesh[3], orc_phonons *ph);
void orc_free_phonons(orc_phonons *ph);

/* ---- interactions.f90 ---- */
/* calculate_eph_interaction_ibzk, key 'g', for IBZ states istate_first..istate_last
 * (1-based, inclusive).  out_count[i] = count for state istate_first+i or -1 when the
 * state is outside the window; out_g2 has row stride nprocs = nstates_inwindow*nbr. */
int orc_ibzk_g2(const orc_wannier *w, const orc_electrons *el, const orc_phonons *ph, int64_t mesh_ref,
                int64_t istate_first, int64_t istate_last, int64_t *out_count, double *out_g2,
                int nthreads);
int orc_ibzk_g2_sample(const orc_wannier *w, const orc_electrons *el, const orc_phonons *ph, int64_t mesh_ref,
                       int64_t istate_first, int64_t istate_last, int64_t *out_count, double *out_g2, int nthreads,
                       int64_t ikp_limit);
int orc_ibzk_g2_window(const orc_wannier *w, const orc_electrons *el, const orc_phonons *ph, int64_t mesh_ref,
                       int64_t istate_first, int64_t istate_last, int64_t *out_count, double *out_g2, int nthreads,
                       int64_t ikp_first, int64_t ikp_count);
void orc_g2_list(const orc_wannier *w, int64_t n, const double *kvecs, const double *qvecs,
                 const zdouble *el_evec_k, const zdouble *el_evec_kp, const zdouble *ph_evec_q,
                 const double *ph_en, const zdouble *gmixed, int64_t gsz, const int64_t *gidx, int wannspace_ph,
                 double *out, int nthreads);
/* calculate_eph_interaction_ibzq, key 'g'; out row stride nprocs = nstates_inwindow*nb. */
int orc_ibzq_g2(const orc_wannier *w, const orc_electrons *el, const orc_phonons *ph,
                int64_t istate_first, int64_t istate_last, int64_t *out_count, double *out_g2,
                int nthreads);

/* ---- consumers of g2: delta.f90, interactions.f90 key 'X', RTA rates ---- */
void orc_form_tetrahedra_3d(int64_t nk, const int64_t mesh[3], int blocks, const int64_t *indexlist, int64_t *tetra,
                            int64_t *tetracount, int64_t *tetramap);
void orc_fill_tetrahedra_3d(int64_t ntet, const int64_t *tetra, int64_t nk, int64_t nb, const double *evals,
                            double *tetra_evals);
double orc_delta_fn_tetra(double e, int64_t ik, int64_t ib, const int64_t mesh[3], int64_t nk, int64_t ntet,
                          int64_t nb, const int64_t *tetramap, const int64_t *tetracount, const double *tetra_evals);
double orc_bose(double e, double T);
double orc_fermi(double e, double chempot, double T);
void orc_form_triangles(int64_t nk, const int64_t mesh[3], int blocks, const int64_t *indexlist, int64_t *triang,
                        int64_t *triangcount, int64_t *triangmap);
void orc_fill_triangles(int64_t ntri, const int64_t *triang, int64_t nk, int64_t nb, const double *evals,
                        double *triang_evals);
double orc_delta_fn_triang(double e, int64_t ik, int64_t ib, const int64_t mesh[3], int64_t nk, int64_t ntri,
                           int64_t nb, const int64_t *triangmap, const int64_t *triangcount,
                           const double *triang_evals);
/* nvert 4: tetrahedra tables, 3: triangle tables (the simplex_* arrays). */
int orc_ibzk_X(const orc_wannier *w, const orc_electrons *el, const orc_phonons *ph, int64_t mesh_ref,
               int64_t istate_first, int64_t istate_last, const int64_t *counts, const double *g2, double T,
               double chempot, int nvert, const int64_t *tetramap, const int64_t *tetracount,
               const double *tetra_evals, double *Xplus, double *Xminus, int64_t *istate_el, int64_t *istate_ph,
               int nthreads);
void orc_el_rta_rates(int64_t nst, const int64_t *counts, int64_t nprocs, const double *Xplus, const double *Xminus,
                      double *rates);

void orc_field_term(char field, const int64_t *nequiv, int64_t nk_ibz, int64_t maxeq, const int64_t *ibz2fbz_map,
                    double T, double chempot, int64_t nk, int64_t nb, const double *ens, const double *vels,
                    const double *rates_ibz, const int64_t *indexlist, double *field_term);
void orc_transport_coeff(char field, double T, int64_t deg, double chempot, int64_t nk, int64_t nb, const double *ens,
                         const double *vels, double volume, const int64_t mesh[3], const double *response,
                         int64_t nrot, const double *crot, double *hc, double *cc);

void orc_symmetrize_field_term(int64_t nk, int64_t nb, const double *symmetrizers, double *field_term);

#ifdef __cplusplus
}
#endif
#endif
