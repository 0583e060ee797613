! kaos_hip_iface.f90 -- iso_c_binding interface of libkaos_hip.so for a Fortran
! driver (the layouts and entry points of include/kaos_hip.h, one for one).
!
! The reference's hot-path subroutines take no arguments and share module
! globals (_KineticMainParams.f90:33-554); a driver that keeps the reference's
! call sequence (_8_KineticSolver.f90:54-239) calls these functions in place of
! KineticSolverA/BSub's push loop, ParticleCountsSub, DistributionFunctionsSub,
! the moment subroutines, AmbipolarEfieldSub, GravfieldSub and
! BoundaryConditionsSub, passing c_loc of its own arrays (valid only during the
! call).  Index conventions: config cells 1..nq (0 = LB escape, -1 = UB), the
! 2-D counts as [Q][Vperp][Vpar] with Vpar fastest.
module kaos_hip
  use, intrinsic :: iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: KAOS_OK = 0, KAOS_E_ARG = 1, KAOS_E_CAPACITY = 2, &
                                       KAOS_E_DEVICE = 3, KAOS_E_STATE = 4, KAOS_E_NAN = 5, &
                                       KAOS_E_IO = 6
  integer(c_int32_t), parameter, public :: KAOS_FP64 = 0, KAOS_MIXED = 1
  integer(c_int32_t), parameter, public :: KAOS_FLAG_NO_GRAPH = 1, KAOS_FLAG_NO_AUTOTUNE = 2, &
                                            KAOS_FLAG_VPERP3D = 4

  type, bind(c), public :: kaos_cfg
    integer(c_int32_t) :: ntubes, precision, ndatfac, device
    integer(c_int64_t) :: seed                   ! uint64_t on the C side
    integer(c_int32_t) :: reorder_every = 0      ! 0: library default; < 0: never
    integer(c_int32_t) :: flags = 0              ! KAOS_FLAG_*
  end type kaos_cfg

  type, bind(c), public :: kaos_grid
    integer(c_int32_t) :: nq, nvpar, nvperp, grav, eamb, eambsign, filter_pt, lbcond, ubcond, filter_pt_m1par
    integer(c_int64_t) :: n_inj_lb, n_inj_ub
    real(c_double) :: p, cphi, sphi, h, ms, qs, nsnormfac, Ts_lb, Ts_ub
    type(c_ptr) :: qGL, qGH, qGC, hqC, dqC, d3xC, rGC, cthGC, ellGC, Te
    type(c_ptr) :: vparGL, vparGH, vparGC, dvpar, vperpGL, vperpGH, vperpGC, dvperp, d3v
    ! ICR wave heating (ICRflagT; stochastic): off unless icr = 1
    integer(c_int32_t) :: icr = 0, pad1 = 0
    real(c_double) :: icr_lambda_perp = 0, icr_eta_lh = 0, icr_xi_perp1 = 0, icr_xi_perp2 = 0
    real(c_double) :: icr_s0 = 0, icr_f0 = 0, icr_chi_perp1 = 0, icr_chi_perp2 = 0
    ! 3-D velocity grid (ION2VPERPflag = 1; read when the context has KAOS_FLAG_VPERP3D)
    integer(c_int32_t) :: nvperp1 = 0, nvperp2 = 0
    type(c_ptr) :: vperp1GL = c_null_ptr, vperp1GH = c_null_ptr, vperp1GC = c_null_ptr, dvperp1 = c_null_ptr
    type(c_ptr) :: vperp2GL = c_null_ptr, vperp2GH = c_null_ptr, vperp2GC = c_null_ptr, dvperp2 = c_null_ptr
    type(c_ptr) :: d3v3 = c_null_ptr
    ! ion-neutral charge exchange (QEXCHANGEflagT = 1; needs the 3-D grid)
    integer(c_int32_t) :: cx = 0, cx_filter = 0, cx_filter_pt = 11, cx_pad = 0
    real(c_double) :: t_neut = 0, m_neut = 0, cx_sigma_scale = 1
    type(c_ptr) :: ns_neut = c_null_ptr
    integer(c_int64_t) :: ena_cap = 0
    ! ENA velocity grid (|Vp|, Vq, Vphi cells; per config cell and bin hVp hVq hVphi and d33vCT)
    integer(c_int32_t) :: nvp = 0, nvq = 0, nvphi = 0, ena_pad = 0
    type(c_ptr) :: vpGL = c_null_ptr, vpGH = c_null_ptr, vpGC = c_null_ptr, dvp = c_null_ptr
    type(c_ptr) :: vqGL = c_null_ptr, vqGH = c_null_ptr, vqGC = c_null_ptr, dvq = c_null_ptr
    type(c_ptr) :: vphiGL = c_null_ptr, vphiGH = c_null_ptr, vphiGC = c_null_ptr, dvphi = c_null_ptr
    type(c_ptr) :: ena_hh = c_null_ptr, ena_d33v = c_null_ptr
  end type kaos_grid

  ! dir: c_loc of a NUL-terminated character array (trim(dataexportdir)//c_null_char)
  type, bind(c), public :: kaos_export_info
    type(c_ptr) :: dir
    integer(c_int32_t) :: species, flux_tube
    real(c_double) :: ns0, phi
    type(c_ptr) :: theta, Ts
  end type kaos_export_info

  ! spin-up (restart) files: dir = c_loc of a NUL-terminated Densitydatadir
  type, bind(c), public :: kaos_spinup_info
    type(c_ptr) :: dir
    integer(c_int32_t) :: species, flux_tube
    integer(c_int64_t) :: nq_lb, nq_ub
    real(c_double) :: nsnorm_clb, nsnorm_cub
    integer(c_int64_t) :: lb_replenish, ub_replenish
  end type kaos_spinup_info

  ! scalars of a previous run's spin-up files (kaos_read_spinup)
  type, bind(c), public :: kaos_spinup_state
    integer(c_int64_t) :: nq_lb, nq_ub
    real(c_double) :: nsnorm_clb, nsnorm_cub
    integer(c_int64_t) :: lb_replenish, ub_replenish, nq
  end type kaos_spinup_state

  ! grid set-up inputs (ConfigGridGeneratorSub, ConfigGridGeneratorSubB, VelGridGeneratorSub)
  type, bind(c), public :: kaos_gridgen_in
    integer(c_int64_t) :: nq_ub
    real(c_double) :: qGA, qGB, L, phi, ms, ti, te, ns0, nsnormfac
    integer(c_int32_t) :: vperp3d = 1, symvpar = 1, cx = 0, device = 0
    integer(c_int64_t) :: nvpar_gp, nvperp_gp
    real(c_double) :: vpar_sigma_fac, vperp_sigma_fac
    real(c_double) :: t_neut = 0d0, zns0_neut = 0d0, ns0_neut = 0d0
  end type kaos_gridgen_in

  type, bind(c), public :: kaos_gridgen_dims
    integer(c_int32_t) :: nq, nvpar, nvperp, nvperp1, nvperp2, nvp, nvq, nvphi
  end type kaos_gridgen_dims

  ! grid set-up outputs: c_loc of caller arrays (c_null_ptr: not wanted)
  type, bind(c), public :: kaos_gridgen_out
    type(c_ptr) :: qGL = c_null_ptr, qGH = c_null_ptr, qGC = c_null_ptr, hqC = c_null_ptr, dqC = c_null_ptr, &
                   d3xC = c_null_ptr, rGC = c_null_ptr, cthGC = c_null_ptr, ellGC = c_null_ptr, Ts = c_null_ptr, &
                   Te = c_null_ptr, density = c_null_ptr, counts = c_null_ptr, ns_neut = c_null_ptr
    real(c_double) :: nsnorm_clb = 0d0, nsnorm_cub = 0d0
    type(c_ptr) :: vparGL = c_null_ptr, vparGH = c_null_ptr, vparGC = c_null_ptr, dvpar = c_null_ptr
    type(c_ptr) :: vperpGL = c_null_ptr, vperpGH = c_null_ptr, vperpGC = c_null_ptr, dvperp = c_null_ptr, &
                   d3v = c_null_ptr
    type(c_ptr) :: vperp1GL = c_null_ptr, vperp1GH = c_null_ptr, vperp1GC = c_null_ptr, dvperp1 = c_null_ptr
    type(c_ptr) :: vperp2GL = c_null_ptr, vperp2GH = c_null_ptr, vperp2GC = c_null_ptr, dvperp2 = c_null_ptr, &
                   d3v3 = c_null_ptr
    type(c_ptr) :: vpGL = c_null_ptr, vpGH = c_null_ptr, vpGC = c_null_ptr, dvp = c_null_ptr
    type(c_ptr) :: vqGL = c_null_ptr, vqGH = c_null_ptr, vqGC = c_null_ptr, dvq = c_null_ptr
    type(c_ptr) :: vphiGL = c_null_ptr, vphiGH = c_null_ptr, vphiGC = c_null_ptr, dvphi = c_null_ptr
    type(c_ptr) :: ena_hh = c_null_ptr, ena_d33v = c_null_ptr
  end type kaos_gridgen_out

  integer(c_int32_t), parameter, public :: KAOS_FIELD_COMPUTED = 0, KAOS_FIELD_INPUT_FIRST = 1, &
                                            KAOS_FIELD_INPUT_FIXED = 2

  type, bind(c), public :: kaos_step_out
    integer(c_int64_t) :: n_alive, n_injected
    integer(c_int32_t) :: master, status
  end type kaos_step_out

  public :: kaos_init, kaos_set_grid, kaos_upload_particles, kaos_upload_vperp, kaos_cell_counts, &
            kaos_gridgen_dims_of, kaos_generate_grid, kaos_seed_particles, kaos_step, kaos_run, &
            kaos_download_master, kaos_download_master3d, kaos_export_master, kaos_export_spinup, &
            kaos_read_spinup, kaos_set_field, kaos_download_cx, kaos_download_enas, &
            kaos_download_ena_master, kaos_download_ena_moments, &
            kaos_count, kaos_download_particles, kaos_download_vperp, kaos_upload_state, &
            kaos_upload_moments, kaos_resume, kaos_save_state, kaos_load_state, kaos_set_launch, &
            kaos_get_launch, kaos_set_launch_block, kaos_get_launch_block, kaos_profile, &
            kaos_profile_read, kaos_last_error, kaos_finalize, kaos_error_message

  interface
    integer(c_int) function kaos_init(cfg, ctx) bind(c, name='kaos_init')
      import :: c_int, c_ptr, kaos_cfg
      type(kaos_cfg), intent(in) :: cfg
      type(c_ptr), intent(out) :: ctx
    end function kaos_init

    integer(c_int) function kaos_set_grid(ctx, tube, g, capacity) bind(c, name='kaos_set_grid')
      import :: c_int, c_ptr, c_int32_t, c_int64_t, kaos_grid
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: tube
      type(kaos_grid), intent(in) :: g
      integer(c_int64_t), value :: capacity
    end function kaos_set_grid

    integer(c_int) function kaos_upload_particles(ctx, tube, n, q0, vpar0, vperp0) &
        bind(c, name='kaos_upload_particles')
      import :: c_int, c_ptr, c_int32_t, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: tube
      integer(c_int64_t), value :: n
      real(c_double), intent(in) :: q0(*), vpar0(*), vperp0(*)
    end function kaos_upload_particles

    integer(c_int) function kaos_upload_vperp(ctx, tube, n, vperp1, vperp2) bind(c, name='kaos_upload_vperp')
      import :: c_int, c_ptr, c_int32_t, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: tube
      integer(c_int64_t), value :: n
      real(c_double), intent(in) :: vperp1(*), vperp2(*)
    end function kaos_upload_vperp

    ! NsFARRp of every cell (DensityProfileA1Sub), host arithmetic only
    integer(c_int) function kaos_cell_counts(nq, density, d3x, nsnormfac, counts) bind(c, name='kaos_cell_counts')
      import :: c_int, c_int32_t, c_int64_t, c_double
      integer(c_int32_t), value :: nq
      real(c_double), intent(in) :: density(*), d3x(*)
      real(c_double), value :: nsnormfac
      integer(c_int64_t), intent(out) :: counts(*)
    end function kaos_cell_counts

    ! grid sizes of a set-up (host arithmetic only)
    integer(c_int) function kaos_gridgen_dims_of(gin, dims) bind(c, name='kaos_gridgen_dims_of')
      import :: c_int, kaos_gridgen_in, kaos_gridgen_dims
      type(kaos_gridgen_in), intent(in) :: gin
      type(kaos_gridgen_dims), intent(out) :: dims
    end function kaos_gridgen_dims_of

    ! ConfigGridGeneratorSub + ConfigGridGeneratorSubB + VelGridGeneratorSub on the device
    integer(c_int) function kaos_generate_grid(gin, gout) bind(c, name='kaos_generate_grid')
      import :: c_int, kaos_gridgen_in, kaos_gridgen_out
      type(kaos_gridgen_in), intent(in) :: gin
      type(kaos_gridgen_out), intent(inout) :: gout
    end function kaos_generate_grid

    ! DensityProfileBSub + VelocityDistributionSub + InitialConditionsSub on the device
    integer(c_int) function kaos_seed_particles(ctx, tube, counts, temperature, seed) &
        bind(c, name='kaos_seed_particles')
      import :: c_int, c_ptr, c_int32_t, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: tube
      integer(c_int64_t), intent(in) :: counts(*)
      real(c_double), intent(in) :: temperature(*)
      integer(c_int64_t), value :: seed           ! uint64_t on the C side
    end function kaos_seed_particles

    integer(c_int) function kaos_step(ctx, n, uniforms, n_uniforms, out) bind(c, name='kaos_step')
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: n
      type(c_ptr), value :: uniforms             ! c_null_ptr: seeded device stream
      integer(c_int64_t), value :: n_uniforms
      type(c_ptr), value :: out                  ! c_loc(kaos_step_out) or c_null_ptr
    end function kaos_step

    integer(c_int) function kaos_run(ctx, n_first, nsteps, particle_steps) bind(c, name='kaos_run')
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: n_first, nsteps
      integer(c_int64_t), intent(out) :: particle_steps
    end function kaos_run

    integer(c_int) function kaos_download_master(ctx, tube, nq_counts, nph_counts, NqRT, M, EA, EG, esc) &
        bind(c, name='kaos_download_master')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: tube
      type(c_ptr), value :: nq_counts, nph_counts, NqRT, M, EA, EG, esc
    end function kaos_download_master

    ! 3-D counts [Q][V1][V2][Vpar] and the eight 3-D moments of the last master step
    integer(c_int) function kaos_download_master3d(ctx, tube, n3_counts, M) &
        bind(c, name='kaos_download_master3d')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: tube
      type(c_ptr), value :: n3_counts, M
    end function kaos_download_master3d

    ! DataExport2Sub's spin-up files (SPINUPflag = 1) of the last master step
    integer(c_int) function kaos_export_spinup(ctx, tube, info) bind(c, name='kaos_export_spinup')
      import :: c_int, c_ptr, c_int32_t, kaos_spinup_info
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: tube
      type(kaos_spinup_info), intent(in) :: info
    end function kaos_export_spinup

    ! charge exchange: sigmaIonNeutRT, nuIonNeutRT, nuIonNeutPoiT of the last master step
    integer(c_int) function kaos_download_cx(ctx, tube, sigma, nu, poi) bind(c, name='kaos_download_cx')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: tube
      type(c_ptr), value :: sigma, nu, poi
    end function kaos_download_cx

    ! the tube's ENAs in id order
    integer(c_int) function kaos_download_enas(ctx, tube, n_cap, n_out, x, y, z, Vx, Vy, Vz, id, qind) &
        bind(c, name='kaos_download_enas')
      import :: c_int, c_ptr, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: tube
      integer(c_int64_t), value :: n_cap
      integer(c_int64_t), intent(out) :: n_out
      type(c_ptr), value :: x, y, z, Vx, Vy, Vz, id, qind
    end function kaos_download_enas

    ! NqReNormENAT and the LB/UB ENA escapes (counts, fluxes) of the last master step
    integer(c_int) function kaos_download_ena_master(ctx, tube, nq_ena, esc) bind(c, name='kaos_download_ena_master')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: tube
      type(c_ptr), value :: nq_ena, esc
    end function kaos_download_ena_master

    ! ENA phase-space counts [Q][Vp][Vq][Vphi] and the eight ENA moments of the last master step
    integer(c_int) function kaos_download_ena_moments(ctx, tube, counts, M) bind(c, name='kaos_download_ena_moments')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: tube
      type(c_ptr), value :: counts, M
    end function kaos_download_ena_moments

    ! the spin-up files of a previous run (SPINUPflag = 0 restart); dir as in kaos_spinup_info
    integer(c_int) function kaos_read_spinup(dir, species, flux_tube, nq_cap, st, density, temperature, &
                                             ea_inertial, ea_pressure, ea_mag) bind(c, name='kaos_read_spinup')
      import :: c_int, c_ptr, c_int32_t, c_int64_t, kaos_spinup_state
      type(c_ptr), value :: dir
      integer(c_int32_t), value :: species, flux_tube
      integer(c_int64_t), value :: nq_cap
      type(kaos_spinup_state), intent(out) :: st
      type(c_ptr), value :: density, temperature, ea_inertial, ea_pressure, ea_mag
    end function kaos_read_spinup

    ! input ambipolar field (EAmagInputT) of a restart and its mode (KAOS_FIELD_*)
    integer(c_int) function kaos_set_field(ctx, tube, EA, mode) bind(c, name='kaos_set_field')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: tube
      type(c_ptr), value :: EA
      integer(c_int32_t), value :: mode
    end function kaos_set_field

    ! DataExport2Sub's files of the last master step
    integer(c_int) function kaos_export_master(ctx, tube, info, nn) bind(c, name='kaos_export_master')
      import :: c_int, c_ptr, c_int32_t, c_int64_t, kaos_export_info
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: tube
      type(kaos_export_info), intent(in) :: info
      integer(c_int64_t), intent(out) :: nn
    end function kaos_export_master

    integer(c_int) function kaos_count(ctx, tube, n) bind(c, name='kaos_count')
      import :: c_int, c_ptr, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: tube
      integer(c_int64_t), intent(out) :: n
    end function kaos_count

    integer(c_int) function kaos_download_particles(ctx, tube, n_cap, n_out, x, y, z, Vx, Vy, Vz, &
        vperp, qk4, aea, ag, qind, vparind, vperpind) bind(c, name='kaos_download_particles')
      import :: c_int, c_ptr, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: tube
      integer(c_int64_t), value :: n_cap
      integer(c_int64_t), intent(out) :: n_out
      type(c_ptr), value :: x, y, z, Vx, Vy, Vz, vperp, qk4, aea, ag, qind, vparind, vperpind
    end function kaos_download_particles

    ! Vperp1, Vperp2 of the live ions, in the order of kaos_download_particles
    integer(c_int) function kaos_download_vperp(ctx, tube, n_cap, n_out, vperp1, vperp2) &
        bind(c, name='kaos_download_vperp')
      import :: c_int, c_ptr, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: tube
      integer(c_int64_t), value :: n_cap
      integer(c_int64_t), intent(out) :: n_out
      type(c_ptr), value :: vperp1, vperp2
    end function kaos_download_vperp

    ! The driver's per-particle arrays handed to a tube (integer(kind=dp) indices)
    integer(c_int) function kaos_upload_state(ctx, tube, n, x, y, z, Vx, Vy, Vz, vperp1, vperp2, vperp, &
        aea, ag, qk4, qind, vparind, vperpind) bind(c, name='kaos_upload_state')
      import :: c_int, c_ptr, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: tube
      integer(c_int64_t), value :: n
      type(c_ptr), value :: x, y, z, Vx, Vy, Vz, vperp1, vperp2, vperp, aea, ag, qk4, qind, vparind, vperpind
    end function kaos_upload_state

    ! M*phRT(nn, :) of the last master step (6*nq), read by the next master step's fields
    integer(c_int) function kaos_upload_moments(ctx, tube, M) bind(c, name='kaos_upload_moments')
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: tube
      real(c_double), intent(in) :: M(*)
    end function kaos_upload_moments

    ! continue another run after its step n_last
    integer(c_int) function kaos_resume(ctx, n_last) bind(c, name='kaos_resume')
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: n_last
    end function kaos_resume

    ! restart files (path: NUL-terminated)
    integer(c_int) function kaos_save_state(ctx, path) bind(c, name='kaos_save_state')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, path
    end function kaos_save_state

    integer(c_int) function kaos_load_state(ctx, path) bind(c, name='kaos_load_state')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, path
    end function kaos_load_state

    integer(c_int) function kaos_set_launch(ctx, tiles_per_group, hist_words, persistent) &
        bind(c, name='kaos_set_launch')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: tiles_per_group, hist_words, persistent
    end function kaos_set_launch

    integer(c_int) function kaos_get_launch(ctx, tiles_per_group, hist_words, persistent, tuning) &
        bind(c, name='kaos_get_launch')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(out) :: tiles_per_group, hist_words, persistent, tuning
    end function kaos_get_launch

    integer(c_int) function kaos_set_launch_block(ctx, block) bind(c, name='kaos_set_launch_block')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: block
    end function kaos_set_launch_block

    integer(c_int) function kaos_get_launch_block(ctx, block) bind(c, name='kaos_get_launch_block')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(out) :: block
    end function kaos_get_launch_block

    integer(c_int) function kaos_profile(ctx, enable) bind(c, name='kaos_profile')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: enable
    end function kaos_profile

    integer(c_int) function kaos_profile_read(ctx, push_ms, push_launches, run_ms) &
        bind(c, name='kaos_profile_read')
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: push_ms, run_ms
      integer(c_int64_t), intent(out) :: push_launches
    end function kaos_profile_read

    type(c_ptr) function kaos_last_error(ctx) bind(c, name='kaos_last_error')
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function kaos_last_error

    integer(c_int) function kaos_finalize(ctx) bind(c, name='kaos_finalize')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function kaos_finalize
  end interface

contains

  ! Text of kaos_last_error(ctx) as a Fortran string (the reference reports
  ! errors by printing 'ERROR: ...' and continuing; a driver can do the same).
  function kaos_error_message(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=256) :: msg
    character(kind=c_char), pointer :: s(:)
    integer :: i
    msg = ' '
    call c_f_pointer(kaos_last_error(ctx), s, [256])
    do i = 1, 256
      if (s(i) == c_null_char) exit
      msg(i:i) = s(i)
    end do
  end function kaos_error_message

end module kaos_hip
