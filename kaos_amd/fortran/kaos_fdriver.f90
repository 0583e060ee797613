! kaos_fdriver.f90 -- a Fortran driver of the hot path through the
! iso_c_binding interface: the call sequence of ComputationalTimeStepLoop
! (_8_KineticSolver.f90:54-239) with the push, counts, distribution function,
! moments, fields and boundary conditions replaced by one kaos_step call per
! time step, and the master-step outputs read back the way the reference
! stores them per master step.
!
! Usage: kaos_fdriver INPUT OUTPUT [EXPORTDIR [SEED [GRID3D]]]
!   EXPORTDIR: DataExport2Sub's files of every master step (kaos_export_master)
!   SEED: initial ions generated by the library (kaos_cell_counts from the
!         density, kaos_seed_particles with that seed) instead of q0/vpar0/vperp0
!   GRID3D: the 3-D velocity grid (ION2VPERPflag = 1; KAOS_FLAG_VPERP3D), stream:
!         int32 nvperp1 nvperp2; real64 Vperp1 GL GH GC dV (nvperp1 each),
!         Vperp2 GL GH GC dV (nvperp2 each), d3v3 (nvperp1*nvperp2*nvpar, Vpar
!         fastest); the records then end with the eight 3-D moments M8(8*nq), and
!         EXPORTDIR also receives the spin-up files (kaos_export_spinup)
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
! Environment: KAOS_HANDOVER_STEP=k hands the run to a fresh context after
!   step k: the ions through kaos_download_particles/kaos_download_vperp and
!   kaos_upload_state, the last master step's moments through
!   kaos_upload_moments, then kaos_resume(k); with KAOS_CHECKPOINT=path the
!   fresh context instead reads a restart file (kaos_save_state/kaos_load_state).
! OUTPUT (stream), one record per master step (and step 1):
!   int64 n, int64 n_alive, int64 NqReNorm(nq), real64 M(6*nq), real64 EAmag(nq), real64 esc(4)
!   [, real64 M8(8*nq) with GRID3D]
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
  type(kaos_spinup_info) :: sinfo
  logical :: do_3d
  integer(c_int32_t) :: n12(2)
  real(c_double), allocatable, target :: v1g(:, :), v2g(:, :), d3v3(:), M8(:)
  character(len=1024) :: g3file
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
  integer(c_int64_t) :: n, nsteps, hstep
  character(len=1024) :: ckpt, envv
  integer :: elen, est

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  do_export = command_argument_count() >= 3
  do_seed = command_argument_count() >= 4
  do_3d = command_argument_count() >= 5
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
  if (do_3d) then
    call get_command_argument(5, g3file)
    open (newunit=u, file=trim(g3file), access='stream', form='unformatted', status='old')
    read (u) n12
    allocate (v1g(n12(1), 4), v2g(n12(2), 4), d3v3(n12(1)*n12(2)*nvpar), M8(8*nq))
    read (u) v1g, v2g, d3v3
    close (u)
  end if

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
  if (do_3d) then
    cfg%flags = ior(cfg%flags, KAOS_FLAG_VPERP3D)
    g%nvperp1 = n12(1); g%nvperp2 = n12(2)
    g%vperp1GL = c_loc(v1g(1, 1)); g%vperp1GH = c_loc(v1g(1, 2)); g%vperp1GC = c_loc(v1g(1, 3))
    g%dvperp1 = c_loc(v1g(1, 4))
    g%vperp2GL = c_loc(v2g(1, 1)); g%vperp2GH = c_loc(v2g(1, 2)); g%vperp2GC = c_loc(v2g(1, 3))
    g%dvperp2 = c_loc(v2g(1, 4))
    g%d3v3 = c_loc(d3v3(1))
  end if

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
    ! spin-up files into the same directory (the reference's Densitydatadir)
    sinfo%dir = c_loc(cdir(1))
    sinfo%species = 1_c_int32_t; sinfo%flux_tube = 1_c_int32_t
    sinfo%nq_lb = 1_c_int64_t; sinfo%nq_ub = int(nq, c_int64_t)
    sinfo%nsnorm_clb = real(h64(1), c_double); sinfo%nsnorm_cub = real(h64(2), c_double)
    sinfo%lb_replenish = 0_c_int64_t; sinfo%ub_replenish = 0_c_int64_t
  end if

  hstep = -1
  ckpt = ' '
  call get_environment_variable('KAOS_HANDOVER_STEP', envv, elen, est)
  if (est == 0 .and. elen > 0) read (envv, *) hstep
  call get_environment_variable('KAOS_CHECKPOINT', ckpt, elen, est)
  if (est /= 0) ckpt = ' '

  open (newunit=uo, file=trim(fout), access='stream', form='unformatted', status='replace')
  do n = 1, nsteps
    rc = kaos_step(ctx, n, c_null_ptr, 0_c_int64_t, c_loc(so))
    call check('kaos_step')
    if (so%master /= 0) then
      rc = kaos_download_master(ctx, 0_c_int32_t, c_loc(nqc), c_null_ptr, c_null_ptr, c_loc(M), &
                                c_loc(EA), c_null_ptr, c_loc(esc))
      call check('kaos_download_master')
      if (do_3d) then
        rc = kaos_download_master3d(ctx, 0_c_int32_t, c_null_ptr, c_loc(M8))
        call check('kaos_download_master3d')
        write (uo) n, so%n_alive, nqc, M, EA, esc, M8
      else
        write (uo) n, so%n_alive, nqc, M, EA, esc
      end if
      if (do_export) then
        rc = kaos_export_master(ctx, 0_c_int32_t, xinfo, nnx)
        call check('kaos_export_master')
        if (do_3d) then
          rc = kaos_export_spinup(ctx, 0_c_int32_t, sinfo)
          call check('kaos_export_spinup')
        end if
      end if
    end if
    if (n == hstep) call handover()
  end do
  close (uo)
  rc = kaos_finalize(ctx)
  print '(a,i0,a)', 'kaos_fdriver: ', nsteps, ' steps done'

contains

  ! Hand the run to a fresh context after step hstep (see the header).
  subroutine handover()
    type(c_ptr) :: ctx2
    integer(c_int64_t) :: nlive, nget
    real(c_double), allocatable, target :: px(:, :), pv1(:), pv2(:), Mh(:)
    integer(c_int64_t), allocatable, target :: pidx(:, :)
    character(len=1, kind=c_char), allocatable, target :: cpath(:)
    integer :: k
    rc = kaos_init(cfg, ctx2)
    if (rc /= KAOS_OK) then
      print *, 'ERROR: kaos_init (handover) ', rc
      stop 1
    end if
    if (len_trim(ckpt) > 0) then
      allocate (cpath(len_trim(ckpt) + 1))
      do k = 1, len_trim(ckpt)
        cpath(k) = ckpt(k:k)
      end do
      cpath(len_trim(ckpt) + 1) = c_null_char
      rc = kaos_save_state(ctx, c_loc(cpath(1)))
      call check('kaos_save_state')
      rc = kaos_finalize(ctx)
      ctx = ctx2
      rc = kaos_set_grid(ctx, 0_c_int32_t, g, h64(4))
      call check('kaos_set_grid (handover)')
      rc = kaos_load_state(ctx, c_loc(cpath(1)))
      call check('kaos_load_state')
      return
    end if
    rc = kaos_count(ctx, 0_c_int32_t, nlive)
    call check('kaos_count')
    allocate (px(max(nlive, 1_c_int64_t), 10), pv1(max(nlive, 1_c_int64_t)), pv2(max(nlive, 1_c_int64_t)))
    allocate (pidx(max(nlive, 1_c_int64_t), 3), Mh(6*nq))
    rc = kaos_download_particles(ctx, 0_c_int32_t, nlive, nget, c_loc(px(1, 1)), c_loc(px(1, 2)), &
         c_loc(px(1, 3)), c_loc(px(1, 4)), c_loc(px(1, 5)), c_loc(px(1, 6)), c_loc(px(1, 7)), &
         c_loc(px(1, 8)), c_loc(px(1, 9)), c_loc(px(1, 10)), c_loc(pidx(1, 1)), c_loc(pidx(1, 2)), &
         c_loc(pidx(1, 3)))
    call check('kaos_download_particles')
    rc = kaos_download_vperp(ctx, 0_c_int32_t, nlive, nget, c_loc(pv1(1)), c_loc(pv2(1)))
    call check('kaos_download_vperp')
    rc = kaos_download_master(ctx, 0_c_int32_t, c_null_ptr, c_null_ptr, c_null_ptr, c_loc(Mh(1)), &
                              c_null_ptr, c_null_ptr, c_null_ptr)
    call check('kaos_download_master (handover)')
    rc = kaos_finalize(ctx)
    ctx = ctx2
    rc = kaos_set_grid(ctx, 0_c_int32_t, g, h64(4))
    call check('kaos_set_grid (handover)')
    rc = kaos_upload_state(ctx, 0_c_int32_t, nget, c_loc(px(1, 1)), c_loc(px(1, 2)), c_loc(px(1, 3)), &
         c_loc(px(1, 4)), c_loc(px(1, 5)), c_loc(px(1, 6)), c_loc(pv1(1)), c_loc(pv2(1)), c_loc(px(1, 7)), &
         c_loc(px(1, 9)), c_loc(px(1, 10)), c_loc(px(1, 8)), c_loc(pidx(1, 1)), c_loc(pidx(1, 2)), &
         c_loc(pidx(1, 3)))
    call check('kaos_upload_state')
    rc = kaos_upload_moments(ctx, 0_c_int32_t, Mh)
    call check('kaos_upload_moments')
    rc = kaos_resume(ctx, hstep)
    call check('kaos_resume')
  end subroutine handover

  subroutine check(what)
    character(len=*), intent(in) :: what
    if (rc /= KAOS_OK) then
      print *, 'ERROR: ', what, ' returned ', rc, ': ', trim(kaos_error_message(ctx))
      stop 1
    end if
  end subroutine check

end program kaos_fdriver
