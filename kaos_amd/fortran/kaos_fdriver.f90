! kaos_fdriver.f90 -- a Fortran driver of the hot path through the
! iso_c_binding interface: the call sequence of ComputationalTimeStepLoop
! (_8_KineticSolver.f90:54-239) with the push, counts, distribution function,
! moments, fields and boundary conditions replaced by one kaos_step call per
! time step, and the master-step outputs read back the way the reference
! stores them per master step.
!
! Usage: kaos_fdriver INPUT OUTPUT [EXPORTDIR [SEED]]
!   EXPORTDIR: DataExport2Sub's files of every master step (kaos_export_master)
!   SEED: initial ions generated by the library (kaos_cell_counts from the
!         density, kaos_seed_particles with that seed) instead of q0/vpar0/vperp0
! INPUT (stream, native endianness):
!   int32  nq nvpar nvperp grav eamb eambsign filter_pt lbcond ubcond ndatfac precision nsteps
!   int64  n_inj_lb n_inj_ub nions capacity seed
!   real64 p cphi sphi h ms qs nsnormfac Ts_lb Ts_ub
!   real64 qGL qGH qGC hqC dqC d3xC rGC cthGC ellGC Te        (nq each)
!   real64 vparGL vparGH vparGC dvpar                          (nvpar each)
!   real64 vperpGL vperpGH vperpGC dvperp                      (nvperp each)
!   real64 d3v                                                 (nvpar*nvperp, Vpar fastest)
!   real64 q0 vpar0 vperp0                                     (nions each)
!   with EXPORTDIR: real64 thetaGC TsGC densityGC (nq each), phi, ns0
! OUTPUT (stream), one record per master step (and step 1):
!   int64 n, int64 n_alive, int64 NqReNorm(nq), real64 M(6*nq), real64 EAmag(nq), real64 esc(4)
program kaos_fdriver
  use, intrinsic :: iso_c_binding
  use kaos_hip
  implicit none
  character(len=1024) :: fin, fout
  integer(c_int32_t) :: hdr(12)
  integer(c_int64_t) :: h64(5)
  real(c_double) :: hd(9)
  real(c_double), allocatable, target :: cg(:, :), vpg(:, :), vrg(:, :), d3v(:), q0(:), v0(:), w0(:)
  real(c_double), allocatable, target :: M(:), EA(:), th(:), ts(:), dens(:)
  real(c_double) :: phi_ns0(2)
  character(len=1024) :: xdir, sarg
  character(len=1, kind=c_char), allocatable, target :: cdir(:)
  integer(c_int64_t), allocatable, target :: nsfa(:)
  type(kaos_export_info) :: xinfo
  integer(c_int64_t) :: nnx, iseed
  logical :: do_export, do_seed
  integer :: i
  real(c_double), target :: esc(4)
  integer(c_int64_t), allocatable, target :: nqc(:)
  type(kaos_cfg) :: cfg
  type(kaos_grid) :: g
  type(kaos_step_out), target :: so
  type(c_ptr) :: ctx
  integer :: u, uo, nq, nvpar, nvperp
  integer(c_int) :: rc
  integer(c_int64_t) :: n, nsteps

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  do_export = command_argument_count() >= 3
  do_seed = command_argument_count() >= 4
  if (do_export) call get_command_argument(3, xdir)
  if (do_seed) then
    call get_command_argument(4, sarg)
    read (sarg, *) iseed
  end if
  open (newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read (u) hdr, h64, hd
  nq = hdr(1); nvpar = hdr(2); nvperp = hdr(3); nsteps = hdr(12)
  allocate (cg(nq, 10), vpg(nvpar, 4), vrg(nvperp, 4), d3v(nvpar*nvperp))
  allocate (q0(h64(3)), v0(h64(3)), w0(h64(3)), M(6*nq), EA(nq), nqc(nq))
  read (u) cg, vpg, vrg, d3v, q0, v0, w0
  if (do_export) then
    allocate (th(nq), ts(nq), dens(nq), nsfa(nq))
    read (u) th, ts, dens, phi_ns0
  end if
  close (u)

  cfg%ntubes = 1; cfg%precision = hdr(11); cfg%ndatfac = hdr(10); cfg%device = 0; cfg%seed = h64(5)
  cfg%reorder_every = 0; cfg%flags = 0
  g%nq = hdr(1); g%nvpar = hdr(2); g%nvperp = hdr(3); g%grav = hdr(4); g%eamb = hdr(5)
  g%eambsign = hdr(6); g%filter_pt = hdr(7); g%lbcond = hdr(8); g%ubcond = hdr(9); g%filter_pt_m1par = 0
  g%n_inj_lb = h64(1); g%n_inj_ub = h64(2)
  g%p = hd(1); g%cphi = hd(2); g%sphi = hd(3); g%h = hd(4); g%ms = hd(5); g%qs = hd(6)
  g%nsnormfac = hd(7); g%Ts_lb = hd(8); g%Ts_ub = hd(9)
  g%qGL = c_loc(cg(1, 1)); g%qGH = c_loc(cg(1, 2)); g%qGC = c_loc(cg(1, 3)); g%hqC = c_loc(cg(1, 4))
  g%dqC = c_loc(cg(1, 5)); g%d3xC = c_loc(cg(1, 6)); g%rGC = c_loc(cg(1, 7)); g%cthGC = c_loc(cg(1, 8))
  g%ellGC = c_loc(cg(1, 9)); g%Te = c_loc(cg(1, 10))
  g%vparGL = c_loc(vpg(1, 1)); g%vparGH = c_loc(vpg(1, 2)); g%vparGC = c_loc(vpg(1, 3)); g%dvpar = c_loc(vpg(1, 4))
  g%vperpGL = c_loc(vrg(1, 1)); g%vperpGH = c_loc(vrg(1, 2)); g%vperpGC = c_loc(vrg(1, 3))
  g%dvperp = c_loc(vrg(1, 4)); g%d3v = c_loc(d3v(1))

  rc = kaos_init(cfg, ctx)
  if (rc /= KAOS_OK) then
    print *, 'ERROR: kaos_init ', rc, trim(kaos_error_message(c_null_ptr))
    stop 1
  end if
  rc = kaos_set_grid(ctx, 0_c_int32_t, g, h64(4))
  call check('kaos_set_grid')
  if (do_seed) then
    ! NsFARRp from the density (DensityProfileA1Sub), then DensityProfileBSub +
    ! VelocityDistributionSub + InitialConditionsSub on the device
    rc = kaos_cell_counts(int(nq, c_int32_t), dens, cg(:, 6), hd(7), nsfa)
    call check('kaos_cell_counts')
    rc = kaos_seed_particles(ctx, 0_c_int32_t, nsfa, ts, iseed)
    call check('kaos_seed_particles')
  else
    rc = kaos_upload_particles(ctx, 0_c_int32_t, h64(3), q0, v0, w0)
    call check('kaos_upload_particles')
  end if
  if (do_export) then
    allocate (cdir(len_trim(xdir) + 1))
    do i = 1, len_trim(xdir)
      cdir(i) = xdir(i:i)
    end do
    cdir(len_trim(xdir) + 1) = c_null_char
    xinfo%dir = c_loc(cdir(1))
    xinfo%species = 1_c_int32_t; xinfo%flux_tube = 1_c_int32_t
    xinfo%ns0 = phi_ns0(2); xinfo%phi = phi_ns0(1)
    xinfo%theta = c_loc(th(1)); xinfo%Ts = c_loc(ts(1))
  end if

  open (newunit=uo, file=trim(fout), access='stream', form='unformatted', status='replace')
  do n = 1, nsteps
    rc = kaos_step(ctx, n, c_null_ptr, 0_c_int64_t, c_loc(so))
    call check('kaos_step')
    if (so%master /= 0) then
      rc = kaos_download_master(ctx, 0_c_int32_t, c_loc(nqc), c_null_ptr, c_null_ptr, c_loc(M), &
                                c_loc(EA), c_null_ptr, c_loc(esc))
      call check('kaos_download_master')
      write (uo) n, so%n_alive, nqc, M, EA, esc
      if (do_export) then
        rc = kaos_export_master(ctx, 0_c_int32_t, xinfo, nnx)
        call check('kaos_export_master')
      end if
    end if
  end do
  close (uo)
  rc = kaos_finalize(ctx)
  print '(a,i0,a)', 'kaos_fdriver: ', nsteps, ' steps done'

contains

  subroutine check(what)
    character(len=*), intent(in) :: what
    if (rc /= KAOS_OK) then
      print *, 'ERROR: ', what, ' returned ', rc, ': ', trim(kaos_error_message(ctx))
      stop 1
    end if
  end subroutine check

end program kaos_fdriver
