! kaos_ref_input.f90 -- TEST INFRASTRUCTURE ONLY.  The harness's own input
! module for the reference's step loop: it takes the place of the reference's
! _ModelInput.f90 (same module name and the same parameter names, which every
! reference module reads through `use ModelInput`) with the values of a
! configuration sized like the CPU-reference workload: one O+ flux tube on L = 4
! from q = 0.779 (1.05 RE) to q = 0.0556 (3 RE), 64 config cells, a 2-D
! (Vperp, Vpar) velocity grid, gravity and the self-consistent ambipolar field,
! lower-boundary injection, no heating, no charge exchange, no convection.  The
! time step makes the reference's own master-step cadence (the lower-boundary
! cell's thermal transit time in steps, _0_ConfigGridGenerator.f90:664-666)
! come out at 3, the minimum it allows (_0_1_1_ConfigGridGeneratorB.f90:29).
! Physical constants are the reference's values.  Output paths are relative:
! the harness runs in a scratch directory.
#ifndef KAOS_REF_EAMB
#define KAOS_REF_EAMB 1
#endif
module ModelInput
  implicit none
  integer, parameter :: dp = kind(1.d0)
  complex(kind= dp) :: i

  ! run mode and I/O (spin-up: the grid and density come from these inputs)
  integer(kind= dp), parameter :: SPINUPflag = 1
  character(*), parameter :: Densitydatadir = './'
  character(*), parameter :: dataexportdir = './kaos_ref_out/'

  ! physical constants
  real(kind= dp), parameter :: pi = 4.d0*datan(1.d0)
  real(kind= dp), parameter :: kB = 138d-25
  real(kind= dp), parameter :: epsilon0 = 8.854d-12
  real(kind= dp), parameter :: mu0 = 1.257d-6
  real(kind= dp), parameter :: m = 8.06d15
  real(kind= dp), parameter :: GG = 667d-13
  real(kind= dp), parameter :: ME = 598d22
  real(kind= dp), parameter :: RE = 638d4
  real(kind= dp), parameter :: melec = 9.109d-31

  ! time: Nt steps of (B - A)/Nt = 11 s (variables: the harness's main program
  ! sets Nt and B from KAOS_REF_NT before the set-up runs)
  real(kind= dp), parameter :: A = 0d0
  real(kind= dp) :: B = 1000d0*11d0
  integer(kind= dp) :: Nt = 1000

  ! ion flags
  integer(kind= dp), parameter :: IONNOISEflag = 0
  integer(kind= dp), parameter :: FLUIDIONEXPORTflag = 1
  integer(kind= dp), parameter :: SYMVPARflag = 1
  integer(kind= dp), parameter :: ION2VPERPflag = 1

  ! boundary conditions
  integer(kind= dp), parameter :: LBCONDITIONflag = 1
  integer(kind= dp), parameter :: LBREPLENISHflag = 0
  integer(kind= dp), parameter :: UBCONDITIONflag = 0
  integer(kind= dp), parameter :: UBREPLENISHflag = 0
  integer(kind= dp), parameter :: STATICINJECTIONflag = 0

  ! forces
  integer(kind= dp), parameter :: ICRflag = 0
  integer(kind= dp), parameter :: ICRCOHERENCEflag = 0
  integer(kind= dp), parameter :: MIRRORflag = 0
  integer(kind= dp), parameter :: GRAVflag = 1
  ! (KAOS_REF_EAMB: 1, or 0 for the build without the ambipolar field)
  integer(kind= dp), parameter :: EAMBflag = KAOS_REF_EAMB
  integer(kind= dp), parameter :: EAMBSELFCONSISTflag = 1
  integer(kind= dp), parameter :: EAMBSIGNflag = 1
  integer(kind= dp), parameter :: EAINERTIALflag = 0
  integer(kind= dp), parameter :: EAPRESSUREflag = 1
  integer(kind= dp), parameter :: EPARflag = 0
  integer(kind= dp), parameter :: CONVECTIONflag = 0
  integer(kind= dp), parameter :: DYNAMICGRIDflag = 0

  ! moments
  integer(kind= dp), parameter :: MOMENTFILTERflag = 1
  integer(kind= dp), parameter :: PHASEIONDISTRIBflag = 1
  integer(kind= dp), parameter :: PHASEDENSITYIONMOMENTflag = 1
  integer(kind= dp), parameter :: PHASEVELPERPIONMOMENTflag = 1
  integer(kind= dp), parameter :: PHASEVELPARIONMOMENTflag = 1
  integer(kind= dp), parameter :: PHASEENERGYIONMOMENTflag = 1
  integer(kind= dp), parameter :: PHASEENERGYPERPIONMOMENTflag = 1
  integer(kind= dp), parameter :: PHASEENERGYPARIONMOMENTflag = 1
  integer(kind= dp), parameter :: PHASEENERGYPERPIONMOMENTCENTERflag = 0
  integer(kind= dp), parameter :: PHASEENERGYPARIONMOMENTCENTERflag = 1
  integer(kind= dp), parameter :: FLUIDIONREFflag = 0

  ! neutrals (off)
  integer(kind= dp), parameter :: QEXCHANGEflag = 0
  integer(kind= dp), parameter :: ENANOISEflag = 0

  ! grid
  integer(kind= dp), parameter :: Stot = 1
  integer(kind= dp), parameter :: Nf = 1
  real(kind= dp), parameter :: LshellIC = 4d0
  real(kind= dp), parameter :: phiLshellIC = pi/2d0
  real(kind= dp), parameter :: qGAIC = 0.7790d0
  real(kind= dp), parameter :: qGBIC = 0.0556d0
  real(kind= dp), parameter :: mO = (16d0)*(1.66054d-27)
  real(kind= dp), parameter :: qO = 1.602d-19
  integer(kind= dp), parameter :: dq = 1
  integer(kind= dp), parameter :: ddVperp1 = 1
  integer(kind= dp), parameter :: ddVperp2 = 1
  integer(kind= dp), parameter :: ddVperp = 1
  integer(kind= dp), parameter :: ddVpar = 1
  integer(kind= dp), parameter :: ddVp = 1
  integer(kind= dp), parameter :: ddVq = 1
  integer(kind= dp), parameter :: ddVphi = 1
  real(kind= dp), parameter :: NVparGpF = 14d0
  real(kind= dp), parameter :: NVperpGpF = 14d0
  real(kind= dp), parameter :: VperpsigmaFacIC = 7.4d0
  real(kind= dp), parameter :: VparsigmaFacIC = 7.4d0
  real(kind= dp), parameter :: TiIC = 3482.6d0
  real(kind= dp), parameter :: TeIC = 1d4
  real(kind= dp), parameter :: TNeutIC = 848d0

  ! initial ions
  integer(kind= dp), parameter :: NqLB = 1
  integer(kind= dp), parameter :: NqUB = 64
  real(kind= dp) :: nsnormfacIC = 7.7d16    ! KAOS_REF_NSNORMFAC (about 1e5 ions at 7.7e16)
  real(kind= dp), parameter :: dNTe = 0d0
  real(kind= dp), parameter :: dNTeEND = 3d10
  real(kind= dp), parameter :: ns0IC = 1d8

  ! neutral oxygen (unused without charge exchange)
  real(kind= dp), parameter :: zns0NeutIC = RE + 719d3
  real(kind= dp), parameter :: ns0NeutIC = 1.423d11

  ! wave heating (unused with ICRflag = 0)
  real(kind= dp), parameter :: EtaLHp = 0.125d0
  real(kind= dp), parameter :: XiPerp1p = 0.5d0
  real(kind= dp), parameter :: XiPerp2p = 0.5d0
  real(kind= dp), parameter :: lambdaPerpp = 0d0
  real(kind= dp), parameter :: f0p = 6.5d0
  real(kind= dp), parameter :: S0p = 5d-5
  real(kind= dp), parameter :: ChiPerp1p = 2.1d0
  real(kind= dp), parameter :: ChiPerp2p = 2.1d0

  ! parallel potential (unused with EPARflag = 0)
  integer(kind= dp), parameter :: Eparlim = 1
  real(kind= dp), parameter :: PhiPar0p = 5d-6
  real(kind= dp), parameter :: dPhiPar0p = 1d0

  ! limits and moment filter windows
  real(kind= dp), parameter :: IonNoiseLimitNph = 2500d0
  real(kind= dp), parameter :: ENANoiseLimitNph = 5d0
  integer(kind= dp), parameter :: M0MAfilterPt = 11
  integer(kind= dp), parameter :: M1Perp1MAfilterPt = 11
  integer(kind= dp), parameter :: M1Perp2MAfilterPt = 11
  integer(kind= dp), parameter :: M1ParMAfilterPt = 11
  integer(kind= dp), parameter :: M2ParMAfilterPt = 11
end module ModelInput
