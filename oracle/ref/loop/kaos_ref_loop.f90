! kaos_ref_loop.f90 -- TEST INFRASTRUCTURE ONLY.  Runs the reference's own
! program sequence (_KineticMain.f90: SimParameterizationSub ..
! InitialConditionsSub, then KineticSolverSub with its ComputationalTimeStepLoop,
! _8_KineticSolver.f90:54-239), every subroutine compiled unmodified from the
! reference sources, with the harness's ModelInput and GaussianRNG modules.
! KAOS_REF_NT (steps) and KAOS_REF_NSNORMFAC (macro-ion weight) set the run
! length and the ion count.  Between the set-up and the loop it writes the
! inputs the reference generated,
! so that the CPU oracle and the GPU library can be run on the same inputs:
!   kaos_ref_grid.bin (rank 0, stream): int64 nq, nvpar, nvperp1, nvperp2,
!     ndatfac, NNt, Nt, ranks; real64 h, L, phi, ms, qs, nsnormfac,
!     TsGT(1, NqLB), nsnormCLBGT(1); real64 (nq each, master-step row 1)
!     qGLGT, qGHGT, qGCGT, hqCGT, dqCGT, d3xCGT, rGCGT, thetaGCGT, ellGCGT,
!     TeGT, TsGT; the 3-D velocity cells of config cell 1 (V2PerpCellT,
!     ION2VPERPflag = 1): real64 Vperp1GLT, Vperp1GHT, Vperp1GCT, dVperp1GT
!     (nvperp1 each), the same for Vperp2 (nvperp2 each), VparGLT, VparGHT,
!     VparGCGT, dVparGT (nvpar each), d3vCT (nvperp1 x nvperp2 x nvpar, Vpar
!     fastest);
!   kaos_ref_ions_<rank>.bin: int64 NsT; real64 q0T, Vpar0T, Vperp10T,
!     Vperp20T, Vperp0T, x0T, y0T, z0T, Vx0T, Vy0T, Vz0T (NsT each).
! After the loop every rank writes the state its last step left:
!   kaos_ref_state_<rank>.bin: int64 NsT; real64 x, y, z, Vx, Vy, Vz, xN, yN,
!     zN, VxN, VyN, VzN, AEAmag, AGmag, AEAmagN, AGmagN, Vperp1, Vperp2
!     (NsT each, zeros for arrays the loop left unallocated); int64 Qindk1,
!     Vparindk1 (NsT each).
! It times KineticSolverSub with MPI_Wtime between barriers and prints
!   kaos_ref_loop seconds <max over ranks> ranks <n> ions0 <sum> ions_end <sum>
!   steps <Nt> ndatfac <ndatfacGT(1)>
program kaos_ref_loop
  use KineticMainParams
  use SimParameterization
  use DensityProfileA
  use DensityProfileB
  use DipolePolynomialSolver
  use VelocityDistribution
  use InitialConditions
  use KineticSolver
  implicit none
  real(kind= dp) :: t0, t1, dtl, dtmax
  integer(kind= dp) :: n0, n0sum, nend, nendsum
  integer(4) :: nranks
  character(len= 64) :: envv
  integer :: elen, est

  call MPI_INIT(ierr)
  call MPI_COMM_RANK(MPI_COMM_WORLD, rank, ierr)
  call MPI_COMM_SIZE(MPI_COMM_WORLD, nranks, ierr)
  ! run length and macro-ion weight of this run (time step kept at 11 s)
  call get_environment_variable('KAOS_REF_NT', envv, elen, est)
  if (est == 0 .and. elen > 0) then
    read (envv, *) Nt
    B = real(Nt, dp)*11d0
  end if
  call get_environment_variable('KAOS_REF_NSNORMFAC', envv, elen, est)
  if (est == 0 .and. elen > 0) read (envv, *) nsnormfacIC

  call SimParameterizationSub
  call DensityProfileASub
  call DensityProfileBSub
  call DipolePolynomialSolverSub
  call VelocityDistributionSub
  call InitialConditionsSub

  call write_inputs()
  n0 = SpecieT(1)%FluxTubeT(1)%NsT(1)

  call MPI_BARRIER(MPI_COMM_WORLD, ierr)
  t0 = MPI_WTIME()
  call KineticSolverSub
  t1 = MPI_WTIME()
  call MPI_BARRIER(MPI_COMM_WORLD, ierr)
  dtl = t1 - t0
  nend = SpecieT(1)%FluxTubeT(1)%NsT(1)
  call write_state()
  call MPI_REDUCE(dtl, dtmax, 1, MPI_DOUBLE_PRECISION, MPI_MAX, 0, MPI_COMM_WORLD, ierr)
  call MPI_REDUCE(n0, n0sum, 1, MPI_INTEGER8, MPI_SUM, 0, MPI_COMM_WORLD, ierr)
  call MPI_REDUCE(nend, nendsum, 1, MPI_INTEGER8, MPI_SUM, 0, MPI_COMM_WORLD, ierr)
  if (rank == 0) then
    write (*, '(a, es16.8, a, i0, a, i0, a, i0, a, i0, a, i0)') 'kaos_ref_loop seconds ', dtmax, &
      ' ranks ', nranks, ' ions0 ', n0sum, ' ions_end ', nendsum, ' steps ', SpecieT(1)%FluxTubeT(1)%NtT(1), &
      ' ndatfac ', SpecieT(1)%FluxTubeT(1)%ndatfacGT(1)
  end if
  call MPI_FINALIZE(ierr)

contains

  subroutine put(u, a, ns)
    integer, intent(in) :: u
    real(kind= dp), allocatable, intent(in) :: a(:)
    integer(kind= dp), intent(in) :: ns
    real(kind= dp), allocatable :: z0(:)
    if (allocated(a)) then
      if (size(a) >= ns) then
        write (u) a(1:ns)
        return
      end if
    end if
    allocate (z0(ns))
    z0 = 0d0
    write (u) z0
  end subroutine put

  subroutine puti(u, a, ns)
    integer, intent(in) :: u
    integer(kind= dp), allocatable, intent(in) :: a(:)
    integer(kind= dp), intent(in) :: ns
    integer(kind= dp), allocatable :: z0(:)
    if (allocated(a)) then
      if (size(a) >= ns) then
        write (u) a(1:ns)
        return
      end if
    end if
    allocate (z0(ns))
    z0 = 0
    write (u) z0
  end subroutine puti

  subroutine write_state()
    integer :: u
    integer(kind= dp) :: ns
    character(len= 64) :: name
    ns = SpecieT(1)%FluxTubeT(1)%NsT(1)
    write (name, '(a, i0, a)') 'kaos_ref_state_', rank, '.bin'
    open (newunit= u, file= trim(name), access= 'stream', form= 'unformatted', status= 'replace')
    write (u) ns
    call put(u, x, ns); call put(u, y, ns); call put(u, z, ns)
    call put(u, Vx, ns); call put(u, Vy, ns); call put(u, Vz, ns)
    call put(u, xN, ns); call put(u, yN, ns); call put(u, zN, ns)
    call put(u, VxN, ns); call put(u, VyN, ns); call put(u, VzN, ns)
    call put(u, AEAmag, ns); call put(u, AGmag, ns); call put(u, AEAmagN, ns); call put(u, AGmagN, ns)
    call put(u, Vperp1, ns); call put(u, Vperp2, ns)
    call puti(u, Qindk1, ns); call puti(u, Vparindk1, ns)
    close (u)
  end subroutine write_state

  subroutine write_inputs()
    integer :: u, lb, ub, nvp, nv1, nv2, ia, ib, ic
    integer(kind= dp) :: hdr(8)
    real(kind= dp) :: sc(8)
    character(len= 64) :: name
    lb = int(SpecieT(1)%FluxTubeT(1)%NqLBT(1))
    ub = int(SpecieT(1)%FluxTubeT(1)%NqUBT(1))
    nvp = int(SpecieT(1)%FluxTubeT(1)%QCellT(1)%NVparGT(1))
    nv1 = int(SpecieT(1)%FluxTubeT(1)%QCellT(1)%NVperp1GT(1))
    nv2 = int(SpecieT(1)%FluxTubeT(1)%QCellT(1)%NVperp2GT(1))
    if (rank == 0) then
      call MPI_COMM_SIZE(MPI_COMM_WORLD, u, ierr)
      hdr = [int(ub - lb + 1, dp), int(nvp, dp), int(nv1, dp), int(nv2, dp), &
             SpecieT(1)%FluxTubeT(1)%ndatfacGT(1), SpecieT(1)%FluxTubeT(1)%NNtT(1), &
             SpecieT(1)%FluxTubeT(1)%NtT(1), int(u, dp)]
      sc = [SpecieT(1)%hT, SpecieT(1)%FluxTubeT(1)%pGCGT(1, lb), SpecieT(1)%FluxTubeT(1)%phiGCGT(1, lb), &
            SpecieT(1)%msT(1), SpecieT(1)%qsT(1), SpecieT(1)%FluxTubeT(1)%nsnormfacT(1), &
            SpecieT(1)%FluxTubeT(1)%TsGT(1, lb), SpecieT(1)%FluxTubeT(1)%nsnormCLBGT(1)]
      open (newunit= u, file= 'kaos_ref_grid.bin', access= 'stream', form= 'unformatted', status= 'replace')
      write (u) hdr, sc
      write (u) SpecieT(1)%FluxTubeT(1)%qGLGT(1, lb:ub), SpecieT(1)%FluxTubeT(1)%qGHGT(1, lb:ub), &
        SpecieT(1)%FluxTubeT(1)%qGCGT(1, lb:ub), SpecieT(1)%FluxTubeT(1)%hqCGT(1, lb:ub), &
        SpecieT(1)%FluxTubeT(1)%dqCGT(1, lb:ub), SpecieT(1)%FluxTubeT(1)%d3xCGT(1, lb:ub), &
        SpecieT(1)%FluxTubeT(1)%rGCGT(1, lb:ub), SpecieT(1)%FluxTubeT(1)%thetaGCGT(1, lb:ub), &
        SpecieT(1)%FluxTubeT(1)%ellGCGT(1, lb:ub), SpecieT(1)%FluxTubeT(1)%TeGT(1, lb:ub), &
        SpecieT(1)%FluxTubeT(1)%TsGT(1, lb:ub)
      associate (V => SpecieT(1)%FluxTubeT(1)%QCellT(1)%V2PerpCellT)
        write (u) (V(ia, 1, 1)%Vperp1GLT(1), ia = 1, nv1), (V(ia, 1, 1)%Vperp1GHT(1), ia = 1, nv1), &
          (V(ia, 1, 1)%Vperp1GCT(1), ia = 1, nv1), (V(ia, 1, 1)%dVperp1GT(1), ia = 1, nv1)
        write (u) (V(1, ib, 1)%Vperp2GLT(1), ib = 1, nv2), (V(1, ib, 1)%Vperp2GHT(1), ib = 1, nv2), &
          (V(1, ib, 1)%Vperp2GCT(1), ib = 1, nv2), (V(1, ib, 1)%dVperp2GT(1), ib = 1, nv2)
        write (u) (V(1, 1, ic)%VparGLT(1), ic = 1, nvp), (V(1, 1, ic)%VparGHT(1), ic = 1, nvp), &
          (V(1, 1, ic)%VparGCGT(1), ic = 1, nvp), (V(1, 1, ic)%dVparGT(1), ic = 1, nvp)
        write (u) (((V(ia, ib, ic)%d3vCT(1), ic = 1, nvp), ib = 1, nv2), ia = 1, nv1)
      end associate
      close (u)
    end if
    write (name, '(a, i0, a)') 'kaos_ref_ions_', rank, '.bin'
    open (newunit= u, file= trim(name), access= 'stream', form= 'unformatted', status= 'replace')
    write (u) SpecieT(1)%FluxTubeT(1)%NsT(1)
    write (u) SpecieT(1)%FluxTubeT(1)%q0T, SpecieT(1)%FluxTubeT(1)%Vpar0T, SpecieT(1)%FluxTubeT(1)%Vperp10T, &
      SpecieT(1)%FluxTubeT(1)%Vperp20T, SpecieT(1)%FluxTubeT(1)%Vperp0T, SpecieT(1)%FluxTubeT(1)%x0T, &
      SpecieT(1)%FluxTubeT(1)%y0T, SpecieT(1)%FluxTubeT(1)%z0T, SpecieT(1)%FluxTubeT(1)%Vx0T, &
      SpecieT(1)%FluxTubeT(1)%Vy0T, SpecieT(1)%FluxTubeT(1)%Vz0T
    close (u)
  end subroutine write_inputs

end program kaos_ref_loop
