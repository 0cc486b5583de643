! A plain (non-coarray) Fortran program that drives the library through
! include/elph_vertex_mod.f90 the way elphbolt's image 1 would
! (app/elphbolt.f90:184-230): test infrastructure for the binding.
!
!   elph_driver host   <epwdir> <g2dir>   symmetry + IBZ counts, the EPW host
!                                         reader, g2 file write/read (no GPU)
!   elph_driver device <epwdir> <g2dir>   context from EPW files, setup,
!                                         ibzk 'g' in FP64, gk2 files
! <epwdir>/crystal.txt holds (list-directed): natoms; lattvecs(3,3);
! basis(3,natoms); masses(natoms); coarse_qmesh(3); wvmesh(3); enref fsthick.
program elph_driver
  use iso_c_binding
  use elph_vertex_mod
  implicit none
  character(len=512) :: mode, epwdir, g2dir
  call get_command_argument(1, mode)
  call get_command_argument(2, epwdir)
  call get_command_argument(3, g2dir)
  if (trim(mode) == "host") then
     call host_mode()
  else if (trim(mode) == "device") then
     call device_mode()
  else
     print '(A)', "usage: elph_driver host|device <epwdir> <g2dir>"
     stop 2
  end if

contains

  subroutine check(st, what)
    integer(c_int), intent(in) :: st
    character(len=*), intent(in) :: what
    if (st /= ELPH_OK) then
       print '(A,A,A,I0)', "FAIL ", what, " status ", st
       stop 1
    end if
  end subroutine check

  subroutine host_mode()
    real(c_double) :: lat(3, 3), basis(3, 2), qrot(3, 3, 96), g2(4, 3), back(4, 3)
    integer(c_int64_t) :: types(2), mesh(3), nsymm, nsymm_rot, nirr, dims(5)
    integer(c_int64_t), allocatable :: irr(:), neq(:)
    integer(c_int64_t), allocatable, target :: rk(:, :), deg(:)
    complex(c_double), allocatable, target :: hw(:, :, :)
    integer(c_int64_t) :: counts(3), cback(3)
    integer :: i, j
    ! fcc a = 5.43 A (examples/Si/input/input.nml orientation), two equal atoms
    lat = 0.2715d0 * reshape([-1d0, 0d0, 1d0, 0d0, 1d0, 1d0, -1d0, 1d0, 0d0], [3, 3])
    basis = reshape([0d0, 0d0, 0d0, 0.25d0, 0.25d0, 0.25d0], [3, 2])
    types = [1, 1]
    mesh = [24, 24, 24]
    call check(elph_symmetry(lat, 2_c_int64_t, basis, types, 1d-5, mesh, nsymm, nsymm_rot, qrot), "symmetry")
    allocate(irr(product(mesh)), neq(product(mesh)))
    call check(elph_irred_wedge(mesh, nsymm_rot, qrot, product(mesh), c_null_ptr, nirr, irr, neq, c_null_ptr), &
         "irred_wedge")
    print '(A,I0,A,I0,A,I0)', "symmetry nsymm ", nsymm, " nirr ", nirr, " nequiv_sum ", sum(neq(1:nirr))
    ! EPW host reader
    call check(elph_epw_read_dims(trim(epwdir)//c_null_char, dims), "epw_read_dims")
    allocate(rk(dims(2), 3), deg(dims(2)), hw(dims(2), dims(1), dims(1)))
    call check(elph_epw_read(trim(epwdir)//c_null_char, dims, c_loc(rk), c_null_ptr, c_null_ptr, c_loc(deg), &
         c_null_ptr, c_null_ptr, c_loc(hw), c_null_ptr, c_null_ptr), "epw_read")
    print '(A,5(1X,I0))', "epw dims", dims
    print '(A,I0,A,I0,A,ES24.16E3)', "epw rcells_k_sum ", sum(abs(rk)), " wsdeg_sum ", sum(deg), &
         " hwann_abs_sum ", sum(abs(hw))
    ! g2 files: write, read back
    do j = 1, 3
       do i = 1, 4
          g2(i, j) = 10d0 * j + i + 0.5d0
       end do
    end do
    counts = [4, -1, 2]
    call check(elph_write_g2_files(trim(g2dir)//c_null_char, 'k', 7_c_int64_t, 3_c_int64_t, counts, g2, &
         4_c_int64_t), "write_g2_files")
    back = 0d0
    call check(elph_read_g2_files(trim(g2dir)//c_null_char, 'k', 7_c_int64_t, 3_c_int64_t, cback, back, &
         4_c_int64_t), "read_g2_files")
    if (any(cback /= counts) .or. any(back(:, 1) /= g2(:, 1)) .or. any(back(1:2, 3) /= g2(1:2, 3))) then
       print '(A)', "FAIL g2 round trip"
       stop 1
    end if
    print '(A,3(1X,I0))', "g2files counts", cback
    print '(A)', "host ok"
  end subroutine host_mode

  subroutine device_mode()
    integer(c_int64_t) :: natoms, cq(3), wv(3), nsymm, nsymm_rot, nst, nprocs, ninw, ntot, dims6(6)
    real(c_double) :: lat(3, 3), enref, fsthick, qrot(3, 3, 96)
    real(c_double), allocatable :: basis(:, :), masses(:)
    integer(c_int64_t), allocatable :: types(:), counts(:)
    real(c_double), allocatable :: g2(:)
    type(c_ptr) :: ctx, est, pst
    type(elph_electrons) :: el
    type(elph_phonons) :: ph
    integer :: u, i, j
    open(newunit=u, file=trim(epwdir)//"/crystal.txt", status="old")
    read(u, *) natoms
    allocate(basis(3, natoms), masses(natoms), types(natoms))
    read(u, *) lat
    read(u, *) basis
    read(u, *) masses
    read(u, *) cq
    read(u, *) wv
    read(u, *) enref, fsthick
    close(u)
    ! atom types from equal masses (the reference's crys%atomtypes)
    do i = 1, int(natoms)
       types(i) = i
       do j = 1, i - 1
          if (masses(j) == masses(i)) then
             types(i) = types(j)
             exit
          end if
       end do
    end do
    ctx = c_null_ptr
    call check(elph_ctx_create_epw(ctx, 0_c_int, trim(epwdir)//c_null_char, natoms, lat, basis, masses, cq, &
         0_c_int32_t, c_null_ptr, c_null_ptr), "ctx_create_epw")
    call check(elph_ctx_dims(ctx, dims6), "ctx_dims")
    call check(elph_symmetry(lat, natoms, basis, types, 1d-5, wv, nsymm, nsymm_rot, qrot), "symmetry")
    call check(elph_calculate_electrons(ctx, nsymm_rot, qrot, wv, enref, fsthick, est, el), "calculate_electrons")
    call check(elph_calculate_phonons(ctx, nsymm_rot, qrot, wv, pst, ph), "calculate_phonons")
    nst = el%nwv_irred * dims6(1)
    nprocs = el%nstates_inwindow * dims6(2)
    allocate(counts(nst), g2(nst * nprocs))
    call check(elph_ibzk_g2(ctx, el, ph, 1_c_int64_t, ELPH_FP64, 1_c_int64_t, nst, counts, g2), "ibzk_g2")
    call check(elph_write_g2_files(trim(g2dir)//c_null_char, 'k', 1_c_int64_t, nst, counts, g2, nprocs), &
         "write_g2_files")
    ninw = count(counts >= 0)
    ntot = sum(counts, mask=counts >= 0)
    print '(A,I0,A,I0,A,I0,A,I0)', "device nsymm ", nsymm, " nirr ", el%nwv_irred, " states ", nst, &
         " in_window ", ninw
    print '(A,I0,A,I0)', "device nprocs ", nprocs, " vertices ", ntot
    call check(elph_state_free(est), "state_free")
    call check(elph_state_free(pst), "state_free")
    call check(elph_ctx_destroy(ctx), "ctx_destroy")
    print '(A)', "device ok"
  end subroutine device_mode
end program elph_driver
