! kaos_ref_push.f90 -- TEST INFRASTRUCTURE ONLY.  Drives the reference's own
! RK4 push (KineticRK4UpdateSub, _8_1_KineticRK4Update.f90:29-377, with
! KineticUpdateA/B/C, RK4DipolePolynomialSolverSub and WaveHeatingSub, compiled
! unmodified from the reference sources) over a set of ions read from a file,
! and writes the updated state, so the oracle's push can be checked against
! the reference arithmetic.
!
! Input (stream, little-endian): int64 nq, nvpar, nvperp, ns; int64 flags(8) =
! (grav, eamb, eambsign, mirror, qexchange, n, ndatfac, icr); real64 sc(8) =
! (p, phi, h, ms, qs, -, -, -); real64 qGL(nq), qGH(nq); vparGL, vparGH (nvpar);
! vperpGL, vperpGH (nvperp); per ion: x, y, z, Vx, Vy, Vz, Vperp, AEAmag,
! AGmag (ns each); int64 Qindk1, Vparindk1, Vperpindk1 (ns each, previous).
! Output: real64 xN, yN, zN, VxN, VyN, VzN, Vpar, qk4 (ns each); int64 Qindk1,
! Vparindk1, Vperpindk1 (ns each).
program kaos_ref_push
  use KineticMainParams
  use KineticRK4Update
  implicit none
  integer(kind= dp) :: hnq, hnvpar, hnvperp, hns, fl(8), ia, ib
  real(kind= dp) :: sc(8)
  real(kind= dp), allocatable :: qL(:), qH(:), vpL(:), vpH(:), vrL(:), vrH(:)
  character(len= 1024) :: fin, fout
  integer :: u
  integer(kind= 8) :: clk0, clk1, clkrate

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit= u, file= trim(fin), access= 'stream', form= 'unformatted', status= 'old')
  read(u) hnq, hnvpar, hnvperp, hns
  read(u) fl
  read(u) sc
  allocate(qL(hnq), qH(hnq), vpL(hnvpar), vpH(hnvpar), vrL(hnvperp), vrH(hnvperp))
  read(u) qL, qH, vpL, vpH, vrL, vrH

  allocate(SpecieT(1))
  allocate(SpecieT(1)%FluxTubeT(1))
  SpecieT(1)%hT= sc(3)
  SpecieT(1)%msT(1)= sc(4)
  SpecieT(1)%qsT(1)= sc(5)
  SpecieT(1)%NfT(1)= 1
  SpecieT(1)%FluxTubeT(1)%NqLBT(1)= 1
  SpecieT(1)%FluxTubeT(1)%NqUBT(1)= hnq
  SpecieT(1)%FluxTubeT(1)%NsT(1)= hns
  SpecieT(1)%FluxTubeT(1)%GRAVflagT(1)= fl(1)
  SpecieT(1)%FluxTubeT(1)%EAMBflagT(1)= fl(2)
  SpecieT(1)%FluxTubeT(1)%EAMBSIGNflagT(1)= fl(3)
  SpecieT(1)%FluxTubeT(1)%MIRRORflagT(1)= fl(4)
  SpecieT(1)%FluxTubeT(1)%QEXCHANGEflagT(1)= fl(5)
  SpecieT(1)%FluxTubeT(1)%ICRflagT(1)= fl(8)
  SpecieT(1)%FluxTubeT(1)%EPARflagT(1)= 0
  SpecieT(1)%FluxTubeT(1)%ION2VPERPflagT(1)= 0
  allocate(SpecieT(1)%FluxTubeT(1)%ndatfacGT(1))
  SpecieT(1)%FluxTubeT(1)%ndatfacGT(1)= fl(7)
  allocate(SpecieT(1)%FluxTubeT(1)%qGLGT(1, hnq), SpecieT(1)%FluxTubeT(1)%qGHGT(1, hnq))
  allocate(SpecieT(1)%FluxTubeT(1)%pGCGT(1, hnq), SpecieT(1)%FluxTubeT(1)%phiGCGT(1, hnq))
  SpecieT(1)%FluxTubeT(1)%qGLGT(1, :)= qL
  SpecieT(1)%FluxTubeT(1)%qGHGT(1, :)= qH
  SpecieT(1)%FluxTubeT(1)%pGCGT(1, :)= sc(1)
  SpecieT(1)%FluxTubeT(1)%phiGCGT(1, :)= sc(2)
  allocate(SpecieT(1)%FluxTubeT(1)%QCellT(1))
  SpecieT(1)%FluxTubeT(1)%QCellT(1)%NVparGT(1)= hnvpar
  SpecieT(1)%FluxTubeT(1)%QCellT(1)%NVperpGT(1)= hnvperp
  allocate(SpecieT(1)%FluxTubeT(1)%QCellT(1)%VCellT(hnvperp, hnvpar))
  do ia= 1, hnvperp
    do ib= 1, hnvpar
      SpecieT(1)%FluxTubeT(1)%QCellT(1)%VCellT(ia, ib)%VparGLT(1)= vpL(ib)
      SpecieT(1)%FluxTubeT(1)%QCellT(1)%VCellT(ia, ib)%VparGHT(1)= vpH(ib)
      SpecieT(1)%FluxTubeT(1)%QCellT(1)%VCellT(ia, ib)%VperpGLT(1)= vrL(ia)
      SpecieT(1)%FluxTubeT(1)%QCellT(1)%VCellT(ia, ib)%VperpGHT(1)= vrH(ia)
    end do
  end do

  allocate(x(hns), y(hns), z(hns), Vx(hns), Vy(hns), Vz(hns), Vperp1(hns), Vperp2(hns), Vperp(hns), &
    Vpar(hns), Vp(hns), Vphi(hns), AEAmag(hns), AGmag(hns), AEPmag(hns), xN(hns), yN(hns), zN(hns), &
    VxN(hns), VyN(hns), VzN(hns), Vperp1N(hns), Vperp2N(hns), VperpN(hns), VparConvSign(hns), &
    Qindk1(hns), Vparindk1(hns), Vperpindk1(hns), Vperp1indk1(hns), Vperp2indk1(hns), &
    Vpindk1(hns), Vqindk1(hns), Vphiindk1(hns), rk4(hns), thetak4(hns), phik4(hns), qk4(hns), &
    pk4(hns), ellk4(hns), ENAflag(hns))
  read(u) x, y, z, Vx, Vy, Vz, Vperp, AEAmag, AGmag
  read(u) Qindk1, Vparindk1, Vperpindk1
  close(u)
  Vperp1= Vperp
  Vperp2= 0d0
  AEPmag= 0d0
  ENAflag= .false.
  i= (0d0, 1d0)
  s= 1
  f= 1
  nnind= 1
  rank= 0
  n= fl(6)
  Time(1)= 0d0
  call system_clock(clk0, clkrate)
  do j= 1, hns
    call KineticRK4UpdateSub
  end do
  call system_clock(clk1)
  ! wall time of the push loop alone (for the CPU baseline)
  write(*, '(a, es14.6)') 'push_seconds ', real(clk1 - clk0, 8)/real(clkrate, 8)
  open(newunit= u, file= trim(fout), access= 'stream', form= 'unformatted', status= 'replace')
  write(u) xN, yN, zN, VxN, VyN, VzN, Vpar, qk4
  write(u) Qindk1, Vparindk1, Vperpindk1
  close(u)
end program kaos_ref_push
