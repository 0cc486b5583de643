! elph_vertex_mod: the Fortran side of the C ABI in elph_vertex.h, as
! iso_c_binding interfaces for elphbolt (Fortran 2018).  One coarray image
! owns the GPU and calls these; the others wait at the `sync all` that follows
! each driver (the resource%gpu_manager pattern, src/resource.f90:75-77).
!
! Conventions: every status is integer(c_int), 0 = success; on failure the
! caller passes the text of elph_last_error(ctx) to exit_with_message
! (src/misc.f90:194-203).  Arrays are the reference's own column-major arrays
! (complex(r64) as interleaved doubles, declared here as real(c_double)).
! Optional outputs are type(c_ptr), value: pass c_loc(x) or c_null_ptr.
! Strings end with c_null_char.
!
! Replaced reference symbols (signatures): Wannier%el_wann
! (src/wannier.f90:430-440), Wannier%ph_wann (585-594), Wannier%g2 (903-922),
! Wannier%gkRp / gReq (1073-1076, 1134-1137), read_EPW_Wannier (289-295),
! calculate_eph_interaction_ibzk / ibzq (src/interactions.f90:2130-2133,
! 1706-1709), the gk2/gq2 writers (2337-2349, 1877-1889).
module elph_vertex_mod
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: ELPH_OK = 0, ELPH_EINVAL = 1, ELPH_EDEVICE = 2, ELPH_EIO = 3, &
       ELPH_ENEGPHONON = 4, ELPH_EWINDOW = 5, ELPH_ENOMEM = 6, ELPH_EWEDGE = 7
  integer(c_int32_t), parameter :: ELPH_FP64 = 0, ELPH_MIXED = 1

  type, bind(C) :: elph_dims
     integer(c_int64_t) :: numwannbands, numbranches, nwsk, nwsq, nwsg, natoms
     integer(c_int64_t) :: coarse_qmesh(3)
     real(c_double)     :: lattvecs(3, 3)
     integer(c_int32_t) :: polar
     real(c_double)     :: epsilon(3, 3)
  end type elph_dims

  type, bind(C) :: elph_electrons
     integer(c_int64_t) :: wvmesh(3)
     real(c_double)     :: enref, fsthick
     integer(c_int64_t) :: nwv_irred
     type(c_ptr)        :: indexlist_irred, ens_irred, evecs_irred
     integer(c_int64_t) :: nwv
     type(c_ptr)        :: indexlist, ens, evecs
     integer(c_int64_t) :: nstates_inwindow
  end type elph_electrons

  type, bind(C) :: elph_phonons
     integer(c_int64_t) :: wvmesh(3)
     integer(c_int64_t) :: nwv
     type(c_ptr)        :: ens, evecs
     integer(c_int64_t) :: nwv_irred
     type(c_ptr)        :: indexlist_irred
  end type elph_phonons

  type, bind(C) :: elph_batch_class_t
     integer(c_int32_t) :: precision, path, cb, rb, candidates
     integer(c_int64_t) :: cols, kp, members
     real(c_double)     :: ms
  end type elph_batch_class_t

  type, bind(C) :: elph_server_stats_t
     integer(c_int64_t) :: completed, failed, pending, active, vertices, units, pool_bytes, pool_used, &
          max_slots_busy
     real(c_double)     :: busy_span_ms
  end type elph_server_stats_t

  interface
     ! ---- context: read_EPW_Wannier / type Wannier ----
     integer(c_int) function elph_ctx_create(ctx, device, dims, rcells_k, rcells_q, rcells_g, elwsdeg, phwsdeg, &
          gwsdeg, Hwann, Dphwann, gwann, basis, masses, born) bind(C, name="elph_ctx_create")
       import
       type(c_ptr) :: ctx
       integer(c_int), value :: device
       type(elph_dims) :: dims
       integer(c_int64_t) :: rcells_k(*), rcells_q(*), rcells_g(*), elwsdeg(*), phwsdeg(*), gwsdeg(*)
       real(c_double) :: Hwann(*), Dphwann(*), basis(*), masses(*)
       type(c_ptr), value :: gwann, born
     end function elph_ctx_create

     integer(c_int) function elph_ctx_create_epw(ctx, device, dir, natoms, lattvecs, basis, masses, coarse_qmesh, &
          polar, epsilon, born) bind(C, name="elph_ctx_create_epw")
       import
       type(c_ptr) :: ctx
       integer(c_int), value :: device
       character(kind=c_char) :: dir(*)
       integer(c_int64_t), value :: natoms
       real(c_double) :: lattvecs(3, 3), basis(3, *), masses(*)
       integer(c_int64_t) :: coarse_qmesh(3)
       integer(c_int32_t), value :: polar
       type(c_ptr), value :: epsilon, born
     end function elph_ctx_create_epw

     integer(c_int) function elph_epw_read_dims(dir, dims) bind(C, name="elph_epw_read_dims")
       import
       character(kind=c_char) :: dir(*)
       integer(c_int64_t) :: dims(5)
     end function elph_epw_read_dims

     integer(c_int) function elph_epw_read(dir, dims, rcells_k, rcells_q, rcells_g, elwsdeg, phwsdeg, gwsdeg, &
          Hwann, Dphwann, gwann) bind(C, name="elph_epw_read")
       import
       character(kind=c_char) :: dir(*)
       integer(c_int64_t) :: dims(5)
       type(c_ptr), value :: rcells_k, rcells_q, rcells_g, elwsdeg, phwsdeg, gwsdeg, Hwann, Dphwann, gwann
     end function elph_epw_read

     integer(c_int) function elph_ctx_destroy(ctx) bind(C, name="elph_ctx_destroy")
       import
       type(c_ptr), value :: ctx
     end function elph_ctx_destroy

     type(c_ptr) function elph_last_error(ctx) bind(C, name="elph_last_error")
       import
       type(c_ptr), value :: ctx
     end function elph_last_error

     integer(c_int) function elph_ctx_dims(ctx, out) bind(C, name="elph_ctx_dims")
       import
       type(c_ptr), value :: ctx
       integer(c_int64_t) :: out(6)
     end function elph_ctx_dims

     ! ---- Wannier interpolation ----
     integer(c_int) function elph_el_wann(ctx, nk, kvecs, energies, velocities, evecs, scissor) &
          bind(C, name="elph_el_wann")
       import
       type(c_ptr), value :: ctx
       integer(c_int64_t), value :: nk
       real(c_double) :: kvecs(nk, 3), energies(nk, *)
       type(c_ptr), value :: velocities, evecs, scissor
     end function elph_el_wann

     integer(c_int) function elph_ph_wann(ctx, nq, qvecs, energies, evecs, velocities) bind(C, name="elph_ph_wann")
       import
       type(c_ptr), value :: ctx
       integer(c_int64_t), value :: nq
       real(c_double) :: qvecs(nq, 3), energies(nq, *)
       type(c_ptr), value :: evecs, velocities
     end function elph_ph_wann

     integer(c_int) function elph_gkRp(ctx, nk, kvecs, out) bind(C, name="elph_gkRp")
       import
       type(c_ptr), value :: ctx
       integer(c_int64_t), value :: nk
       real(c_double) :: kvecs(nk, 3), out(*)
     end function elph_gkRp

     integer(c_int) function elph_gReq(ctx, nq, qvecs, out) bind(C, name="elph_gReq")
       import
       type(c_ptr), value :: ctx
       integer(c_int64_t), value :: nq
       real(c_double) :: qvecs(nq, 3), out(*)
     end function elph_gReq

     integer(c_int) function elph_g2(ctx, n, kvecs, qvecs, el_evec_k, el_evec_kp, ph_evec_q, ph_en, gmixed, &
          ngmixed, gidx, wannspace, out) bind(C, name="elph_g2")
       import
       type(c_ptr), value :: ctx
       integer(c_int64_t), value :: n, ngmixed
       real(c_double) :: kvecs(n, 3), qvecs(n, 3), el_evec_k(*), el_evec_kp(*), ph_evec_q(*), ph_en(n), &
            gmixed(*), out(n)
       integer(c_int64_t) :: gidx(n)
       character(kind=c_char), value :: wannspace
     end function elph_g2

     ! ---- drivers: calculate_eph_interaction_ibzk / ibzq, key 'g' ----
     integer(c_int) function elph_ibzk_plan_create(ctx, el, ph, mesh_ref, precision, plan) &
          bind(C, name="elph_ibzk_plan_create")
       import
       type(c_ptr), value :: ctx
       type(elph_electrons) :: el
       type(elph_phonons) :: ph
       integer(c_int64_t), value :: mesh_ref
       integer(c_int32_t), value :: precision
       type(c_ptr) :: plan
     end function elph_ibzk_plan_create

     integer(c_int) function elph_ibzk_plan_run(plan, first, last, counts, g2, on_device) &
          bind(C, name="elph_ibzk_plan_run")
       import
       type(c_ptr), value :: plan, g2
       integer(c_int64_t), value :: first, last
       integer(c_int64_t) :: counts(*)
       integer(c_int32_t), value :: on_device
     end function elph_ibzk_plan_run

     integer(c_int) function elph_ibzk_plan_run_noalloc(plan, first, last, counts, g2_device) &
          bind(C, name="elph_ibzk_plan_run_noalloc")
       import
       type(c_ptr), value :: plan, g2_device
       integer(c_int64_t), value :: first, last
       integer(c_int64_t) :: counts(*)
     end function elph_ibzk_plan_run_noalloc

     integer(c_int64_t) function elph_ibzk_plan_nprocs(plan) bind(C, name="elph_ibzk_plan_nprocs")
       import
       type(c_ptr), value :: plan
     end function elph_ibzk_plan_nprocs

     integer(c_int) function elph_ibzk_plan_destroy(plan) bind(C, name="elph_ibzk_plan_destroy")
       import
       type(c_ptr), value :: plan
     end function elph_ibzk_plan_destroy

     integer(c_int) function elph_ibzk_g2(ctx, el, ph, mesh_ref, precision, first, last, counts, g2) &
          bind(C, name="elph_ibzk_g2")
       import
       type(c_ptr), value :: ctx
       type(elph_electrons) :: el
       type(elph_phonons) :: ph
       integer(c_int64_t), value :: mesh_ref, first, last
       integer(c_int32_t), value :: precision
       integer(c_int64_t) :: counts(*)
       real(c_double) :: g2(*)
     end function elph_ibzk_g2

     integer(c_int) function elph_ibzq_plan_create(ctx, el, ph, precision, plan) bind(C, name="elph_ibzq_plan_create")
       import
       type(c_ptr), value :: ctx
       type(elph_electrons) :: el
       type(elph_phonons) :: ph
       integer(c_int32_t), value :: precision
       type(c_ptr) :: plan
     end function elph_ibzq_plan_create

     integer(c_int) function elph_ibzq_plan_run(plan, first, last, counts, g2, on_device) &
          bind(C, name="elph_ibzq_plan_run")
       import
       type(c_ptr), value :: plan, g2
       integer(c_int64_t), value :: first, last
       integer(c_int64_t) :: counts(*)
       integer(c_int32_t), value :: on_device
     end function elph_ibzq_plan_run

     integer(c_int64_t) function elph_ibzq_plan_stride(plan) bind(C, name="elph_ibzq_plan_stride")
       import
       type(c_ptr), value :: plan
     end function elph_ibzq_plan_stride

     integer(c_int) function elph_ibzq_plan_destroy(plan) bind(C, name="elph_ibzq_plan_destroy")
       import
       type(c_ptr), value :: plan
     end function elph_ibzq_plan_destroy

     ! ---- consumers: tetrahedra / triangles, key 'X', RTA transport ----
     integer(c_int) function elph_tetra_create(ctx, nk, mesh, indexlist, nb, ens, tet) bind(C, name="elph_tetra_create")
       import
       type(c_ptr), value :: ctx, indexlist
       integer(c_int64_t), value :: nk, nb
       integer(c_int64_t) :: mesh(3)
       real(c_double) :: ens(nk, nb)
       type(c_ptr) :: tet
     end function elph_tetra_create

     integer(c_int) function elph_triang_create(ctx, nk, mesh, indexlist, nb, ens, tet) &
          bind(C, name="elph_triang_create")
       import
       type(c_ptr), value :: ctx, indexlist
       integer(c_int64_t), value :: nk, nb
       integer(c_int64_t) :: mesh(3)
       real(c_double) :: ens(nk, nb)
       type(c_ptr) :: tet
     end function elph_triang_create

     integer(c_int) function elph_tetra_tables(t, tetra, tetracount, tetramap, tetra_evals) &
          bind(C, name="elph_tetra_tables")
       import
       type(c_ptr), value :: t, tetra, tetracount, tetramap, tetra_evals
     end function elph_tetra_tables

     integer(c_int) function elph_tetra_delta(t, n, e, ik, ib, out) bind(C, name="elph_tetra_delta")
       import
       type(c_ptr), value :: t
       integer(c_int64_t), value :: n
       real(c_double) :: e(n), out(n)
       integer(c_int64_t) :: ik(n), ib(n)
     end function elph_tetra_delta

     integer(c_int) function elph_tetra_destroy(t) bind(C, name="elph_tetra_destroy")
       import
       type(c_ptr), value :: t
     end function elph_tetra_destroy

     integer(c_int) function elph_ibzk_plan_X(plan, tet, T, chempot, first, last, g2, xplus, xminus, istate_el, &
          istate_ph, rates) bind(C, name="elph_ibzk_plan_X")
       import
       type(c_ptr), value :: plan, tet, g2, xplus, xminus, istate_el, istate_ph, rates
       real(c_double), value :: T, chempot
       integer(c_int64_t), value :: first, last
     end function elph_ibzk_plan_X

     integer(c_int) function elph_el_transport_create(ctx, nk, nb, ens, vels, indexlist, nk_ibz, maxeq, nequiv, &
          ibz2fbz_map, symmetrizers, nrot, crotations, volume, mesh, spindeg, tr) &
          bind(C, name="elph_el_transport_create")
       import
       type(c_ptr), value :: ctx, symmetrizers, crotations
       integer(c_int64_t), value :: nk, nb, nk_ibz, maxeq, nrot, spindeg
       real(c_double) :: ens(nk, nb), vels(nk, nb, 3)
       real(c_double), value :: volume
       integer(c_int64_t) :: indexlist(nk), nequiv(nk_ibz), ibz2fbz_map(maxeq, nk_ibz, 2), mesh(3)
       type(c_ptr) :: tr
     end function elph_el_transport_create

     integer(c_int) function elph_el_transport_rta(tr, rates, T, chempot, kappa0, sigmaS, sigma, alphabyT) &
          bind(C, name="elph_el_transport_rta")
       import
       type(c_ptr), value :: tr, rates, kappa0, sigmaS, sigma, alphabyT
       real(c_double), value :: T, chempot
     end function elph_el_transport_rta

     integer(c_int) function elph_el_transport_destroy(tr) bind(C, name="elph_el_transport_destroy")
       import
       type(c_ptr), value :: tr
     end function elph_el_transport_destroy

     ! ---- on-disk g2 / X layouts ----
     integer(c_int) function elph_write_g2_files(dir, kind, first, nstates, counts, g2, stride) &
          bind(C, name="elph_write_g2_files")
       import
       character(kind=c_char) :: dir(*)
       character(kind=c_char), value :: kind
       integer(c_int64_t), value :: first, nstates, stride
       integer(c_int64_t) :: counts(*)
       real(c_double) :: g2(*)
     end function elph_write_g2_files

     integer(c_int) function elph_g2_writer_create(ctx, dir, kind, stride, max_states, writer) &
          bind(C, name="elph_g2_writer_create")
       import
       type(c_ptr), value :: ctx
       character(kind=c_char) :: dir(*)
       character(kind=c_char), value :: kind
       integer(c_int64_t), value :: stride, max_states
       type(c_ptr) :: writer
     end function elph_g2_writer_create

     integer(c_int) function elph_g2_writer_submit(writer, first, nstates, counts, g2_device) &
          bind(C, name="elph_g2_writer_submit")
       import
       type(c_ptr), value :: writer, g2_device
       integer(c_int64_t), value :: first, nstates
       integer(c_int64_t) :: counts(*)
     end function elph_g2_writer_submit

     integer(c_int) function elph_g2_writer_wait(writer) bind(C, name="elph_g2_writer_wait")
       import
       type(c_ptr), value :: writer
     end function elph_g2_writer_wait

     type(c_ptr) function elph_g2_writer_error(writer) bind(C, name="elph_g2_writer_error")
       import
       type(c_ptr), value :: writer
     end function elph_g2_writer_error

     integer(c_int) function elph_g2_writer_destroy(writer) bind(C, name="elph_g2_writer_destroy")
       import
       type(c_ptr), value :: writer
     end function elph_g2_writer_destroy

     integer(c_int) function elph_read_g2_files(dir, kind, first, nstates, counts, g2, stride) &
          bind(C, name="elph_read_g2_files")
       import
       character(kind=c_char) :: dir(*)
       character(kind=c_char), value :: kind
       integer(c_int64_t), value :: first, nstates, stride
       integer(c_int64_t) :: counts(*)
       real(c_double) :: g2(*)
     end function elph_read_g2_files

     integer(c_int) function elph_write_X_files(dir, first, nstates, counts, xplus, xminus, istate_el, istate_ph, &
          stride) bind(C, name="elph_write_X_files")
       import
       character(kind=c_char) :: dir(*)
       integer(c_int64_t), value :: first, nstates, stride
       integer(c_int64_t) :: counts(*), istate_el(*), istate_ph(*)
       real(c_double) :: xplus(*), xminus(*)
     end function elph_write_X_files

     ! ---- setup: symmetry, IBZ, calculate_electrons / calculate_phonons ----
     integer(c_int) function elph_symmetry(lattvecs, natoms, basis, atomtypes, symprec, mesh, nsymm, nsymm_rot, &
          qrotations) bind(C, name="elph_symmetry")
       import
       real(c_double) :: lattvecs(3, 3), basis(3, *), qrotations(3, 3, 96)
       integer(c_int64_t), value :: natoms
       integer(c_int64_t) :: atomtypes(*), mesh(3), nsymm, nsymm_rot
       real(c_double), value :: symprec
     end function elph_symmetry

     integer(c_int) function elph_symmetry_cartesian(lattvecs, natoms, basis, atomtypes, symprec, mesh, nsymm, &
          nsymm_rot, crotations_orig, crotations) bind(C, name="elph_symmetry_cartesian")
       import
       real(c_double) :: lattvecs(3, 3), basis(3, *)
       integer(c_int64_t), value :: natoms
       integer(c_int64_t) :: atomtypes(*), mesh(3), nsymm, nsymm_rot
       real(c_double), value :: symprec
       type(c_ptr), value :: crotations_orig, crotations
     end function elph_symmetry_cartesian

     integer(c_int) function elph_symmetrizers(mesh, nsymm, qrotations, crotations_orig, nk, indexlist, &
          symmetrizers) bind(C, name="elph_symmetrizers")
       import
       integer(c_int64_t) :: mesh(3)
       integer(c_int64_t), value :: nsymm, nk
       real(c_double) :: qrotations(3, 3, *), crotations_orig(3, 3, *), symmetrizers(3, 3, nk)
       type(c_ptr), value :: indexlist
     end function elph_symmetrizers

     integer(c_int) function elph_irred_wedge(mesh, nsymm_rot, qrotations, nlist, indexlist, nirr, indexlist_irred, &
          nequiv, ibz2fbz) bind(C, name="elph_irred_wedge")
       import
       integer(c_int64_t) :: mesh(3), nirr, indexlist_irred(*), nequiv(*)
       integer(c_int64_t), value :: nsymm_rot, nlist
       real(c_double) :: qrotations(3, 3, *)
       type(c_ptr), value :: indexlist, ibz2fbz
     end function elph_irred_wedge

     integer(c_int) function elph_calculate_electrons(ctx, nsymm_rot, qrotations, wvmesh, enref, fsthick, state, &
          view) bind(C, name="elph_calculate_electrons")
       import
       type(c_ptr), value :: ctx
       integer(c_int64_t), value :: nsymm_rot
       real(c_double) :: qrotations(3, 3, *)
       integer(c_int64_t) :: wvmesh(3)
       real(c_double), value :: enref, fsthick
       type(c_ptr) :: state
       type(elph_electrons) :: view
     end function elph_calculate_electrons

     integer(c_int) function elph_calculate_phonons(ctx, nsymm_rot, qrotations, wvmesh, state, view) &
          bind(C, name="elph_calculate_phonons")
       import
       type(c_ptr), value :: ctx
       integer(c_int64_t), value :: nsymm_rot
       real(c_double) :: qrotations(3, 3, *)
       integer(c_int64_t) :: wvmesh(3)
       type(c_ptr) :: state
       type(elph_phonons) :: view
     end function elph_calculate_phonons

     integer(c_int) function elph_state_free(state) bind(C, name="elph_state_free")
       import
       type(c_ptr), value :: state
     end function elph_state_free

     ! ---- device memory and timing (no counterpart in the reference) ----
     integer(c_int) function elph_device_alloc(ctx, bytes, dptr) bind(C, name="elph_device_alloc")
       import
       type(c_ptr), value :: ctx
       integer(c_int64_t), value :: bytes
       type(c_ptr) :: dptr
     end function elph_device_alloc

     integer(c_int) function elph_device_free(ctx, dptr) bind(C, name="elph_device_free")
       import
       type(c_ptr), value :: ctx, dptr
     end function elph_device_free

     integer(c_int) function elph_device_to_host(ctx, host, dptr, bytes) bind(C, name="elph_device_to_host")
       import
       type(c_ptr), value :: ctx, host, dptr
       integer(c_int64_t), value :: bytes
     end function elph_device_to_host

     integer(c_int) function elph_host_to_device(ctx, dptr, host, bytes) bind(C, name="elph_host_to_device")
       import
       type(c_ptr), value :: ctx, dptr, host
       integer(c_int64_t), value :: bytes
     end function elph_host_to_device

     integer(c_int) function elph_ctx_synchronize(ctx) bind(C, name="elph_ctx_synchronize")
       import
       type(c_ptr), value :: ctx
     end function elph_ctx_synchronize

     integer(c_int) function elph_device_copy(ctx, dst, src, bytes) bind(C, name="elph_device_copy")
       import
       type(c_ptr), value :: ctx, dst, src
       integer(c_int64_t), value :: bytes
     end function elph_device_copy

     integer(c_int) function elph_device_meminfo(device, free_bytes, total_bytes) bind(C, name="elph_device_meminfo")
       import
       integer(c_int), value :: device
       integer(c_int64_t) :: free_bytes, total_bytes
     end function elph_device_meminfo

     integer(c_int) function elph_synth_vertex(ctx, seed) bind(C, name="elph_synth_vertex")
       import
       type(c_ptr), value :: ctx
       integer(c_int64_t), value :: seed   ! uint64_t bits
     end function elph_synth_vertex

     integer(c_int) function elph_selftest_mfma(device, max_err) bind(C, name="elph_selftest_mfma")
       import
       integer(c_int), value :: device
       real(c_double) :: max_err
     end function elph_selftest_mfma

     integer(c_int) function elph_last_timing(ctx, out) bind(C, name="elph_last_timing")
       import
       type(c_ptr), value :: ctx
       real(c_double) :: out(3)
     end function elph_last_timing

     ! ---- batched multi-material driver ----
     integer(c_int) function elph_batch_create(b, device, nstreams) bind(C, name="elph_batch_create")
       import
       type(c_ptr) :: b
       integer(c_int), value :: device
       integer(c_int32_t), value :: nstreams
     end function elph_batch_create

     integer(c_int) function elph_batch_add(b, plan, first, last, g2_device, counts) bind(C, name="elph_batch_add")
       import
       type(c_ptr), value :: b, plan, g2_device
       integer(c_int64_t), value :: first, last
       integer(c_int64_t) :: counts(*)
     end function elph_batch_add

     integer(c_int) function elph_batch_prepare(b, autotune) bind(C, name="elph_batch_prepare")
       import
       type(c_ptr), value :: b
       integer(c_int32_t), value :: autotune
     end function elph_batch_prepare

     integer(c_int) function elph_batch_run(b, ms) bind(C, name="elph_batch_run")
       import
       type(c_ptr), value :: b
       real(c_double) :: ms
     end function elph_batch_run

     integer(c_int64_t) function elph_batch_classes(b, out, nmax) bind(C, name="elph_batch_classes")
       import
       type(c_ptr), value :: b
       type(elph_batch_class_t) :: out(*)
       integer(c_int64_t), value :: nmax
     end function elph_batch_classes

     type(c_ptr) function elph_batch_error(b) bind(C, name="elph_batch_error")
       import
       type(c_ptr), value :: b
     end function elph_batch_error

     integer(c_int) function elph_batch_destroy(b) bind(C, name="elph_batch_destroy")
       import
       type(c_ptr), value :: b
     end function elph_batch_destroy

     ! ---- serving ----
     integer(c_int) function elph_server_create(s, device, precision, pool_bytes, nslots) &
          bind(C, name="elph_server_create")
       import
       type(c_ptr) :: s
       integer(c_int), value :: device
       integer(c_int32_t), value :: precision, nslots
       integer(c_int64_t), value :: pool_bytes
     end function elph_server_create

     integer(c_int) function elph_server_add_material(s, ctx) bind(C, name="elph_server_add_material")
       import
       type(c_ptr), value :: s, ctx
     end function elph_server_add_material

     integer(c_int64_t) function elph_server_submit(s, material, nk, kvecs, nq, qvecs) bind(C, name="elph_server_submit")
       import
       type(c_ptr), value :: s
       integer(c_int32_t), value :: material
       integer(c_int64_t), value :: nk, nq
       real(c_double) :: kvecs(nk, 3), qvecs(nq, 3)
     end function elph_server_submit

     integer(c_int) function elph_server_wait(s, id, latency_ms) bind(C, name="elph_server_wait")
       import
       type(c_ptr), value :: s
       integer(c_int64_t), value :: id
       real(c_double) :: latency_ms
     end function elph_server_wait

     integer(c_int) function elph_server_test(s, id) bind(C, name="elph_server_test")
       import
       type(c_ptr), value :: s
       integer(c_int64_t), value :: id
     end function elph_server_test

     integer(c_int) function elph_server_fetch(s, id, out) bind(C, name="elph_server_fetch")
       import
       type(c_ptr), value :: s
       integer(c_int64_t), value :: id
       real(c_float) :: out(*)
     end function elph_server_fetch

     integer(c_int) function elph_server_result_device(s, id, dptr, count) bind(C, name="elph_server_result_device")
       import
       type(c_ptr), value :: s
       integer(c_int64_t), value :: id
       type(c_ptr) :: dptr
       integer(c_int64_t) :: count
     end function elph_server_result_device

     integer(c_int) function elph_server_release(s, id) bind(C, name="elph_server_release")
       import
       type(c_ptr), value :: s
       integer(c_int64_t), value :: id
     end function elph_server_release

     integer(c_int) function elph_server_stats(s, stats) bind(C, name="elph_server_stats")
       import
       type(c_ptr), value :: s
       type(elph_server_stats_t) :: stats
     end function elph_server_stats

     type(c_ptr) function elph_server_error(s) bind(C, name="elph_server_error")
       import
       type(c_ptr), value :: s
     end function elph_server_error

     integer(c_int) function elph_server_destroy(s) bind(C, name="elph_server_destroy")
       import
       type(c_ptr), value :: s
     end function elph_server_destroy
  end interface
end module elph_vertex_mod
