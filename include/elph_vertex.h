/* elph_vertex: MI355X (gfx950) electron-phonon vertex library, C ABI.
 *
 * Drop-in for the electron-phonon vertex path of elphbolt (Fortran, coarray).
 * Every entry point takes plain pointers and sizes; arrays use the reference's
 * column-major (Fortran) layouts; complex128 arrays are interleaved (re, im)
 * doubles, i.e. `complex(r64)` memory.  Status 0 = success; any other value is
 * an error whose text `elph_last_error()` returns (the Fortran wrapper passes
 * it to `exit_with_message`, src/misc.f90:194-203).
 *
 * One context owns one GPU; its Wannier-space tensors stay resident in HBM
 * until elph_ctx_destroy().  A context is not thread-safe (one coarray image
 * drives it, as `resource%gpu_manager` does for the OpenACC path,
 * src/resource.f90:75-77).
 */
#ifndef ELPH_VERTEX_H
#define ELPH_VERTEX_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct elph_ctx elph_ctx;
typedef struct elph_ibzk_plan elph_ibzk_plan;
typedef struct elph_tetra elph_tetra;
typedef struct elph_el_transport elph_el_transport;
typedef struct elph_ibzq_plan elph_ibzq_plan;

enum { ELPH_OK = 0, ELPH_EINVAL = 1, ELPH_EDEVICE = 2, ELPH_EIO = 3, ELPH_ENEGPHONON = 4, ELPH_EWINDOW = 5,
       ELPH_ENOMEM = 6, ELPH_EWEDGE = 7 };

/* Arithmetic of the two Bloch-phase contractions. */
enum { ELPH_FP64 = 0,  /* FP64 MFMA, FP64 everywhere */
       ELPH_MIXED = 1  /* phase contraction as sliced low-precision MFMA products with exact
                          integer accumulation, FP64 recombination; eigensolves and k.R
                          reduction stay FP64 */ };

/* Sizes of `type Wannier` (src/wannier.f90:46-93) and the crystal data the
 * path reads (src/crystal.f90). */
typedef struct {
    int64_t numwannbands;   /* nb  */
    int64_t numbranches;    /* nbr = 3*natoms */
    int64_t nwsk, nwsq, nwsg;
    int64_t natoms;
    int64_t coarse_qmesh[3];
    double lattvecs[9];     /* crys%lattvecs (3,3) nm, columns */
    int32_t polar;          /* crys%polar */
    double epsilon[9];      /* crys%epsilon (3,3) */
} elph_dims;

/* Replaces `wann%read` / `read_EPW_Wannier` (src/wannier.f90:289-428) when the
 * caller already holds the arrays.  rcells_* are (nws,3) int64; *wsdeg (nws);
 * Hwann (nwsk,nb,nb) c128 Ry; Dphwann (nwsq,nbr,nbr) c128; gwann
 * (nb,nb,nwsk,nbr,nwsg) c128 (the un-sharded `gwann`, src/wannier.f90:41);
 * basis (3,natoms) fractional; masses (natoms) amu, per atom (i.e.
 * crys%masses(crys%atomtypes(:))); born (3,3,natoms) or NULL when !polar. */
int elph_ctx_create(elph_ctx **ctx, int device, const elph_dims *dims, const int64_t *rcells_k,
                    const int64_t *rcells_q, const int64_t *rcells_g, const int64_t *elwsdeg,
                    const int64_t *phwsdeg, const int64_t *gwsdeg, const double *Hwann, const double *Dphwann,
                    const double *gwann, const double *basis, const double *masses, const double *born);

/* Replaces `read_EPW_Wannier` (src/wannier.f90:289-428): reads epwdata.fmt,
 * epmatwp1, rcells_{k,q,g}, wsdeg_{k,q,g} from `dir`.  The crystal data comes
 * from the caller (`crys`, src/crystal.f90). */
int elph_ctx_create_epw(elph_ctx **ctx, int device, const char *dir, int64_t natoms, const double *lattvecs,
                        const double *basis, const double *masses, const int64_t *coarse_qmesh, int32_t polar,
                        const double *epsilon, const double *born);

/* The reader of elph_ctx_create_epw on its own, host only (no device): the
 * sizes (nb, nwsk, nbr, nwsq, nwsg) of the EPW files in `dir`, then the arrays
 * in the layouts of elph_ctx_create (any output may be NULL).  The text files
 * are read as the reference's list-directed `read(u, *)` statements read them
 * (src/wannier.f90:313-340, 379-427). */
int elph_epw_read_dims(const char *dir, int64_t dims[5]);
int elph_epw_read(const char *dir, const int64_t dims[5], int64_t *rcells_k, int64_t *rcells_q, int64_t *rcells_g,
                  int64_t *elwsdeg, int64_t *phwsdeg, int64_t *gwsdeg, double *Hwann, double *Dphwann,
                  double *gwann);

int elph_ctx_destroy(elph_ctx *ctx);
const char *elph_last_error(const elph_ctx *ctx);
/* Sizes read back from a context (nb, nbr, nwsk, nwsq, nwsg, natoms). */
int elph_ctx_dims(const elph_ctx *ctx, int64_t out[6]);

/* Replaces `Wannier%el_wann` (src/wannier.f90:430-524).  kvecs (nk,3) crystal
 * coords; energies (nk,nb) eV; velocities (nk,nb,3) or NULL; evecs (nk,nb,nb)
 * c128 or NULL (evecs(ik,ib,:) = eigenvector ib); scissor (nb) or NULL.
 * velocities without evecs is an error, as in the reference (454-455). */
int elph_el_wann(elph_ctx *ctx, int64_t nk, const double *kvecs, double *energies, double *velocities,
                 double *evecs, const double *scissor);

/* Replaces `Wannier%ph_wann` (src/wannier.f90:585-782).  qvecs (nq,3);
 * energies (nq,nbr) eV; evecs (nq,nbr,nbr) c128 or NULL; velocities (nq,nbr,3)
 * or NULL.  Returns ELPH_ENEGPHONON when an energy below -0.005 eV occurs. */
int elph_ph_wann(elph_ctx *ctx, int64_t nq, const double *qvecs, double *energies, double *evecs,
                 double *velocities);

/* Replaces `Wannier%gkRp` (src/wannier.f90:1073-1132) for nk wave vectors at
 * once.  out: nk consecutive (nb,nb,nbr,nwsq) c128 blocks, the byte layout of
 * the reference's gkRp.ik<ik> files (entries beyond nwsg are zero). */
int elph_gkRp(elph_ctx *ctx, int64_t nk, const double *kvecs, double *out);

/* Replaces `Wannier%gReq` (src/wannier.f90:1134-1188) for nq wave vectors.
 * out: nq consecutive (nb,nb,nbr,nwsk) c128 blocks (gReq.iq<iq> layout). */
int elph_gReq(elph_ctx *ctx, int64_t nq, const double *qvecs, double *out);

/* Replaces `Wannier%g2` (src/wannier.f90:903-1011) for a list of n vertices.
 * Vertex i uses kvecs/qvecs row i ((n,3)), el_evec_k/el_evec_kp (n,nb) c128,
 * ph_evec_q (n,nbr) c128, ph_en (n) eV and the mixed-representation block
 * gmixed[gidx[i]] of shape (nb,nb,nbr,nws) with nws = nwsg for wannspace 'p'
 * ('ph') or nwsk for 'e' ('el').  out (n) eV^2. */
int elph_g2(elph_ctx *ctx, int64_t n, const double *kvecs, const double *qvecs, const double *el_evec_k,
            const double *el_evec_kp, const double *ph_evec_q, const double *ph_en, const double *gmixed,
            int64_t ngmixed, const int64_t *gidx, char wannspace, double *out);

/* Electron and phonon data consumed by the vertex drivers: the fields of
 * `type electron` / `type phonon` that calculate_eph_interaction_ibzk/ibzq
 * read (src/interactions.f90:2095-2400, 1673-1930).  All host pointers. */
typedef struct {
    int64_t wvmesh[3];
    double enref, fsthick;
    int64_t nwv_irred;              /* IBZ blocks */
    const int64_t *indexlist_irred; /* (nwv_irred) muxed FBZ index, 1-based */
    const double *ens_irred;        /* (nwv_irred,nb) eV */
    const double *evecs_irred;      /* (nwv_irred,nb,nb) c128 */
    int64_t nwv;                    /* FBZ blocks */
    const int64_t *indexlist;       /* (nwv) sorted, 1-based */
    const double *ens;              /* (nwv,nb) */
    const double *evecs;            /* (nwv,nb,nb) c128 */
    int64_t nstates_inwindow;
} elph_electrons;

typedef struct {
    int64_t wvmesh[3];
    int64_t nwv;                    /* full phonon mesh */
    const double *ens;              /* (nwv,nbr) eV */
    const double *evecs;            /* (nwv,nbr,nbr) c128 */
    int64_t nwv_irred;
    const int64_t *indexlist_irred; /* (nwv_irred) */
} elph_phonons;

/* calculate_eph_interaction_ibzk, key 'g' (src/interactions.f90:2095-2400),
 * split into plan (upload + window/offset tables, HBM-resident) and run.
 * run() evaluates IBZ states istate_first..istate_last (1-based, state =
 * (ik-1)*nb + m as demux_state, src/misc.f90:977-990): counts[i] = count of
 * state istate_first+i, or -1 when the state is outside the window (the
 * reference writes no file then); g2 row i (length nprocs =
 * nstates_inwindow*nbr) holds the vertices in the reference's order (k'
 * ascending, in-window n ascending, s ascending).  g2 is device memory when
 * g2_on_device != 0, host memory otherwise.  mesh_ref > 1 (fine phonons off
 * the phonon mesh, 2241-2251) interpolates those phonons with ph_wann on the
 * device. */
int elph_ibzk_plan_create(elph_ctx *ctx, const elph_electrons *el, const elph_phonons *ph, int64_t mesh_ref,
                          int32_t precision, elph_ibzk_plan **plan);
int elph_ibzk_plan_run(elph_ibzk_plan *plan, int64_t istate_first, int64_t istate_last, int64_t *counts,
                       double *g2, int32_t g2_on_device);
int64_t elph_ibzk_plan_nprocs(const elph_ibzk_plan *plan);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc): elph_ibzk_plan_run with the g2 rows in device memory (row i =
 * state istate_first+i, stride nprocs), allocating nothing.  Its group tables
 * and the scratch of its kernels come from one 128 MiB array per GPU that
 * the library holds from its first use on that GPU until the process ends;
 * the IBZ k of a call are processed in groups sized to fit it (ELPH_ENOMEM
 * when a single IBZ k needs more).  Every call on a GPU shares that array:
 * calls that overlap on one GPU (other threads, other contexts) race.  Create
 * the plan (which allocates) beforehand. */
int elph_ibzk_plan_run_noalloc(elph_ibzk_plan *plan, int64_t istate_first, int64_t istate_last, int64_t *counts,
                               double *g2_device);
int elph_ibzk_plan_destroy(elph_ibzk_plan *plan);
/* One-shot form (plan + run + destroy). */
int elph_ibzk_g2(elph_ctx *ctx, const elph_electrons *el, const elph_phonons *ph, int64_t mesh_ref,
                 int32_t precision, int64_t istate_first, int64_t istate_last, int64_t *counts, double *g2);

/* calculate_eph_interaction_ibzq, key 'g' (src/interactions.f90:1673-1930).
 * Phonon state = (iq-1)*nbr + s over IBZ q.  g2 row stride =
 * nstates_inwindow*nb (the reference sizes its buffer nstates_inwindow*nbr,
 * 1724, which a window with more than nbr band pairs per k would overrun).
 * The ibzq contractions run in FP64 for either precision value. */
int elph_ibzq_plan_create(elph_ctx *ctx, const elph_electrons *el, const elph_phonons *ph, int32_t precision,
                          elph_ibzq_plan **plan);
int elph_ibzq_plan_run(elph_ibzq_plan *plan, int64_t istate_first, int64_t istate_last, int64_t *counts,
                       double *g2, int32_t g2_on_device);
int64_t elph_ibzq_plan_stride(const elph_ibzq_plan *plan);
int elph_ibzq_plan_destroy(elph_ibzq_plan *plan);

/* ---- Consumers of the g2 rows ----
 * form_tetrahedra_3d + fill_tetrahedra_3d (src/delta.f90:494-646) over a wave
 * vector list: indexlist (nk) muxed 1-based FBZ indices, strictly ascending,
 * as electron.f90:582-587 passes el%indexlist (blocks = .true.), or NULL for
 * the full mesh (nk = prod(mesh)); ens (nk, nb) column-major eV.  Vertices
 * off the list carry energy 0, as in the reference.  Tables live on the
 * context's device. */
int elph_tetra_create(elph_ctx *ctx, int64_t nk, const int64_t *mesh, const int64_t *indexlist, int64_t nb,
                      const double *ens, elph_tetra **out);
/* The same for the triangle method, form_triangles + fill_triangles
 * (src/delta.f90:648-788, used by the reference when tetrahedra = .false.):
 * the delta function of the returned object is delta_fn_triang (790-882). */
int elph_triang_create(elph_ctx *ctx, int64_t nk, const int64_t *mesh, const int64_t *indexlist, int64_t nb,
                       const double *ens, elph_tetra **out);
/* Host copies in the reference's layouts, any pointer may be NULL: tetra
 * (6 nk, 4) (-1 = vertex off the list), tetracount (nk), tetramap (2, nk, 24),
 * tetra_evals (6 nk, nb, 4) sorted over the vertices; for triangles (2 nk, 3),
 * (nk), (2, nk, 6), (2 nk, nb, 3). */
int elph_tetra_tables(const elph_tetra *t, int64_t *tetra, int64_t *tetracount, int64_t *tetramap,
                      double *tetra_evals);
/* delta_fn_tetra (src/delta.f90:61-225) or delta_fn_triang (790-882)
 * evaluated on the device for n host samples (energy e[i], 1-based list
 * index ik[i], 1-based band ib[i]). */
int elph_tetra_delta(elph_tetra *t, int64_t n, const double *e, const int64_t *ik, const int64_t *ib, double *out);
int elph_tetra_destroy(elph_tetra *t);

/* calculate_eph_interaction_ibzk, key 'X' (src/interactions.f90:2278-2354),
 * on the device rows g2 (stride nprocs) that elph_ibzk_plan_run wrote for the
 * same istate_first..istate_last, the tetrahedra (or triangles) formed on
 * the plan's electron list: X+ and X- (THz, rows of stride nprocs; rows of states
 * outside the window are left untouched), the final electron state
 * mux_state(nb, n, ik') and phonon state mux_state(nbr, s, iq) (negative for
 * a phonon off the phonon mesh), and the e-ph RTA rate of each state, sum X+
 * + sum X- (calculate_el_rta_rates, 2957-3063; 0 outside the window).  All
 * output pointers are device memory and may be NULL (X+ and X- together, the
 * two state maps together); T in K, chempot in eV.  At most 65535 states per
 * call. */
int elph_ibzk_plan_X(elph_ibzk_plan *plan, const elph_tetra *tet, double T, double chempot, int64_t istate_first,
                     int64_t istate_last, const double *g2, double *xplus, double *xminus, int64_t *istate_el,
                     int64_t *istate_ph, double *rates);

/* Electron RTA transport (dragless_ebte_RTA, src/bte.f90:200-277) on the
 * device.  create() takes the static inputs, host memory, reference layouts:
 * ens (nk, nb) eV and vels (nk, nb, 3) km/s on the list indexlist (nk,
 * ascending muxed), the IBZ wedge nequiv (nk_ibz) and ibz2fbz_map (maxeq,
 * nk_ibz, 2), the per-k symmetrizers (3, 3, nk) applied to the field terms
 * (NULL: none), the Cartesian rotations (3, 3, nrot) that symmetrize each
 * band's tensor (nrot = 0: none), the primitive cell volume (nm^3), the
 * mesh and the spin degeneracy. */
int elph_el_transport_create(elph_ctx *ctx, int64_t nk, int64_t nb, const double *ens, const double *vels,
                             const int64_t *indexlist, int64_t nk_ibz, int64_t maxeq, const int64_t *nequiv,
                             const int64_t *ibz2fbz_map, const double *symmetrizers, int64_t nrot,
                             const double *crotations, double volume, const int64_t *mesh, int64_t spindeg,
                             elph_el_transport **out);
/* From the IBZ e-ph rates (device memory, THz, state (ik-1)*nb + m as
 * elph_ibzk_plan_X writes them for istate 1..nk_ibz*nb): the field terms of
 * calculate_field_term (src/bte.f90:912-1011) for gradT and E, RTA
 * responses, and calculate_transport_coeff (src/bz_sums.f90:484-611).
 * Outputs host (nb, 3, 3) column-major, any may be NULL: kappa0 (W/m/K),
 * sigmaS (A/m/K), sigma (1/Ohm/m), alphabyT (A/m/K). */
int elph_el_transport_rta(elph_el_transport *t, const double *rates_ibz, double T, double chempot, double *kappa0,
                          double *sigmaS, double *sigma, double *alphabyT);
int elph_el_transport_destroy(elph_el_transport *t);

/* Writes gk2.istate<i> (kind 'k': int64 count + nprocs doubles, the whole
 * buffer as src/interactions.f90:2337-2349 does) or gq2.istate<i> (kind 'q':
 * int64 count + count doubles, 1877-1889) for states with counts >= 0. */
int elph_write_g2_files(const char *dir, char kind, int64_t istate_first, int64_t nstates, const int64_t *counts,
                        const double *g2, int64_t row_stride);
/* The resume side (`read_gk2` / `read_gq2`, app/elphbolt.f90:185,212 and
 * src/numerics.f90:197-198; the consumers' reads, src/interactions.f90:
 * 2194-2217 and src/eliashberg.f90:176-189): reads gk2/gq2.istate<i> of
 * states istate_first.. into host rows of stride row_stride, the first
 * counts[t] values of each; counts[t] = -1 where no file exists (a state the
 * driver skipped).  ELPH_EIO on a short or corrupt file. */
int elph_read_g2_files(const char *dir, char kind, int64_t istate_first, int64_t nstates, int64_t *counts,
                       double *g2, int64_t row_stride);
/* The same files written while the device computes the next states (the
 * reference writes each state's file as soon as it has it,
 * src/interactions.f90:2337-2349, 1877-1889).  submit() takes device rows
 * (stride row_stride) of up to max_states states, queues their copy to pinned
 * host memory behind the work already on the context's stream, makes later
 * work on that stream wait for the copy -- so the caller may run the next
 * states into the same device buffer at once -- and returns; a host thread
 * writes the files.  Two staging buffers alternate (submit waits while both
 * hold batches not yet on disk).  wait() returns when every submitted state
 * is on disk, with the first error (ELPH_EIO when a file could not be
 * written).  No counterpart in the reference as a separate object. */
typedef struct elph_g2_writer elph_g2_writer;
int elph_g2_writer_create(elph_ctx *ctx, const char *dir, char kind, int64_t row_stride, int64_t max_states,
                          elph_g2_writer **writer);
int elph_g2_writer_submit(elph_g2_writer *writer, int64_t istate_first, int64_t nstates, const int64_t *counts,
                          const double *g2_device);
int elph_g2_writer_wait(elph_g2_writer *writer);
const char *elph_g2_writer_error(const elph_g2_writer *writer);
int elph_g2_writer_destroy(elph_g2_writer *writer);
/* Writes Xplus.istate<i> and Xminus.istate<i> as calculate_eph_interaction_ibzk
 * key 'X' does (src/interactions.f90:2356-2377) and read_transition_probs_e
 * reads them (3065-3091): int64 count, count doubles X, count int64
 * istate_el, count int64 istate_ph; states with counts >= 0; host rows of
 * stride row_stride (copy the device rows of elph_ibzk_plan_X first). */
int elph_write_X_files(const char *dir, int64_t istate_first, int64_t nstates, const int64_t *counts,
                       const double *xplus, const double *xminus, const int64_t *istate_el, const int64_t *istate_ph,
                       int64_t row_stride);

/* Setup the vertex path consumes (host + device): symmetry operations
 * (src/symmetry.f90:66-212, found by a lattice/basis search in place of
 * spglib), the irreducible wedge (303-451), calculate_electrons
 * (src/electron.f90:321-595) and calculate_phonons (src/phonon.f90:159-310).
 * Outputs are library-owned until elph_state_free. */
int elph_symmetry(const double *lattvecs, int64_t natoms, const double *basis, const int64_t *atomtypes,
                  double symprec, const int64_t *mesh, int64_t *nsymm, int64_t *nsymm_rot,
                  double *qrotations /* (3,3,96) */);
/* The Cartesian operations of elph_symmetry (src/symmetry.f90:131-137,
 * 154-159, 191-205): crotations_orig (3,3,nsymm) = L W L^-1 of each point
 * operation W (`sym%crotations_orig`), crotations (3,3,nsymm_rot) in the order
 * of elph_symmetry's qrotations (`sym%crotations`, which the transport tensors
 * are symmetrized with).  Either output may be NULL. */
int elph_symmetry_cartesian(const double *lattvecs, int64_t natoms, const double *basis, const int64_t *atomtypes,
                            double symprec, const int64_t *mesh, int64_t *nsymm, int64_t *nsymm_rot,
                            double *crotations_orig, double *crotations);
/* The symmetrizers of wave-vector dependent vectors (src/electron.f90:
 * 462-476): for each point of indexlist (nk muxed 1-based, NULL = the whole
 * mesh), the mean of crotations_orig(:,:,jj), jj = 1..nsymm, over the
 * operations whose qrotations(:,:,jj) map the point onto itself; out
 * (3,3,nk).  This is `el%symmetrizers`, which the field terms are multiplied
 * by (src/bte.f90:242-246, 262-266; elph_el_transport_create). */
int elph_symmetrizers(const int64_t *mesh, int64_t nsymm, const double *qrotations, const double *crotations_orig,
                      int64_t nk, const int64_t *indexlist, double *symmetrizers);
int elph_irred_wedge(const int64_t *mesh, int64_t nsymm_rot, const double *qrotations, int64_t nlist,
                     const int64_t *indexlist /* NULL = full mesh */, int64_t *nirr,
                     int64_t *indexlist_irred /* (nlist) */, int64_t *nequiv /* (nlist) */,
                     int64_t *ibz2fbz /* (nsymm_rot,nlist,2) or NULL */);

typedef struct elph_state elph_state;
int elph_calculate_electrons(elph_ctx *ctx, int64_t nsymm_rot, const double *qrotations, const int64_t *wvmesh,
                             double enref, double fsthick, elph_state **state, elph_electrons *view);
int elph_calculate_phonons(elph_ctx *ctx, int64_t nsymm_rot, const double *qrotations, const int64_t *wvmesh,
                           elph_state **state, elph_phonons *view);
int elph_state_free(elph_state *state);

/* Device memory of the context's GPU for callers without their own allocator
 * (e.g. to keep run() outputs in HBM: pass the pointer with g2_on_device = 1).
 * No counterpart in the reference (its outputs go to files). */
int elph_device_alloc(elph_ctx *ctx, int64_t bytes, void **dptr);
int elph_device_free(elph_ctx *ctx, void *dptr);
int elph_device_to_host(elph_ctx *ctx, void *host, const void *dptr, int64_t bytes);
int elph_host_to_device(elph_ctx *ctx, void *dptr, const void *host, int64_t bytes);
/* Waits for all work of the context's device (the reference's `sync all`
 * boundary after a driver, src/interactions.f90:2387). */
int elph_ctx_synchronize(elph_ctx *ctx);

/* Device-to-device copy on the context's stream (synchronous).  No counterpart
 * in the reference; used to calibrate hardware memory counters. */
int elph_device_copy(elph_ctx *ctx, void *dst, const void *src, int64_t bytes);
/* Free and total device memory of a GPU (bytes).  No counterpart in the reference. */
int elph_device_meminfo(int device, int64_t *free_bytes, int64_t *total_bytes);

/* Fills the context's Wannier vertex in HBM with the synthetic law of the
 * benchmark workloads (|g| = 0.1 eV e^{-|Re|/2 A - |Rp|/2 A}, phase uniform in
 * [0, 2 pi) from a counter-based hash of (seed, element)); pass gwann = NULL to
 * elph_ctx_create to skip the host copy.  No counterpart in the reference. */
int elph_synth_vertex(elph_ctx *ctx, uint64_t seed);

/* Checks the fp64 MFMA operand layout the kernels assume (returns max abs
 * error of a 16x16x4 product against the host; the device must be present). */
int elph_selftest_mfma(int device, double *max_err);

/* Device time of the last run()'s dominant kernel launches in ms, measured with
 * HIP events on the stream the kernels ran on: [0] sum over launches, [1]
 * number of launches, [2] sum of all kernels of run(). */
int elph_last_timing(const elph_ctx *ctx, double out[3]);

/* ------------------------------------------------------------------------
 * Batched multi-material driver (no counterpart in the reference; it runs the
 * work of several calculate_eph_interaction_ibzk('g') calls at once).  Each
 * added plan contributes the in-window IBZ states first..last of its material,
 * |g|^2 rows written to g2_device (device memory, row stride = the plan's
 * nprocs; one row per in-window state, in istate order) and counts filled as
 * elph_ibzk_plan_run does (-1 for out-of-window states).  prepare() picks the
 * pair-stage tile shape of each size class (nb*nbr and Kp buckets) by timing
 * the candidates on the class's first group when autotune != 0, assigns
 * plans to streams and captures the whole step into one hipGraph; run()
 * replays it (ms = device time of the replay).  Plans are tuned in place:
 * their later runs use the chosen shapes.
 * ------------------------------------------------------------------------ */
typedef struct elph_batch elph_batch;
typedef struct {
    int32_t precision, path;  /* path 0 fused kernel, 1 general column-tiled */
    int32_t cb, rb;           /* column blocks of 16, row blocks of 16 per wave */
    int32_t candidates;       /* shapes timed */
    int64_t cols, kp;         /* nb*nbr and Kp of the class's tuning material */
    int64_t members;
    double ms;                /* pair stage of the tuning group with the chosen shape */
} elph_batch_class_t;
int elph_batch_create(elph_batch **out, int device, int32_t nstreams);
int elph_batch_add(elph_batch *b, elph_ibzk_plan *plan, int64_t first, int64_t last, double *g2_device,
                   int64_t *counts);
int elph_batch_prepare(elph_batch *b, int32_t autotune);
int elph_batch_run(elph_batch *b, double *ms);
int64_t elph_batch_classes(const elph_batch *b, elph_batch_class_t *out, int64_t max);
const char *elph_batch_error(const elph_batch *b);
int elph_batch_destroy(elph_batch *b);

/* ------------------------------------------------------------------------
 * Serving (no counterpart in the reference; not called by the Fortran side).
 * A server keeps the Wannier tensors of its materials (contexts) resident and
 * answers requests for |g|^2 at arbitrary k- and q-lists: for every k of the
 * list, every q of the list and all m, n, s (the g2 'ph' form of
 * src/wannier.f90:903-1011 with U_k, U_k+q from el_wann, e_q from ph_wann and
 * g(k,Rp) from gkRp), as FP32 eV^2 in the order
 *   out[(((ik * nq + iq) * nb + m) * nb + n) * nbr + s].
 * Requests are admitted first come first served while their device data
 * (lists, phonons, phases, output) fits in the pool given at creation, and
 * are executed by a scheduler thread as work units of up to 64 k-points of
 * one material taken from one or more requests, on nslots streams at once.
 * precision: ELPH_FP64 or ELPH_MIXED (int8 signed-digit phase contractions).
 * All calls are thread safe.  Materials are added before the first submit and
 * the contexts must outlive the server.
 * ------------------------------------------------------------------------ */
typedef struct elph_server elph_server;
typedef struct {
    int64_t completed, failed, pending, active;
    int64_t vertices;        /* |g|^2 values of completed requests */
    int64_t units;           /* work units executed */
    int64_t pool_bytes, pool_used;
    int64_t max_slots_busy;
    double busy_span_ms;     /* first submit to last completion */
} elph_server_stats_t;
int elph_server_create(elph_server **out, int device, int32_t precision, int64_t pool_bytes, int32_t nslots);
/* returns the material id (>= 0) or -status */
int elph_server_add_material(elph_server *s, elph_ctx *ctx);
/* kvecs (nk,3), qvecs (nq,3) column-major crystal coordinates; returns the
 * request id (>= 0) or -status (ELPH_ENOMEM: larger than the pool) */
int64_t elph_server_submit(elph_server *s, int32_t material, int64_t nk, const double *kvecs, int64_t nq,
                           const double *qvecs);
/* blocks until the request is done; returns its status (ELPH_ENEGPHONON when
 * ph_wann at its q-list found omega < -0.005 eV); latency = completion - submit */
int elph_server_wait(elph_server *s, int64_t id, double *latency_ms);
int elph_server_test(elph_server *s, int64_t id);  /* 1 done, 0 not yet, <0 error */
int elph_server_fetch(elph_server *s, int64_t id, float *out);  /* copy the result to host */
int elph_server_result_device(elph_server *s, int64_t id, const float **dptr, int64_t *count);
int elph_server_release(elph_server *s, int64_t id);  /* frees the request's device data */
int elph_server_stats(elph_server *s, elph_server_stats_t *stats);
const char *elph_server_error(const elph_server *s);
int elph_server_destroy(elph_server *s);

#ifdef __cplusplus
}
#endif
#endif
