! kaos_ref_gauss.f90 -- TEST INFRASTRUCTURE ONLY.  The harness's GaussianRNG
! module (in place of _3_1_GaussianRNG.f90, same module and subroutine names):
! the same Box-Muller formula on two uniforms, taken from the logged, seeded
! stream of kaos_ref_rng.c instead of random_number.
module GaussianRNG
  use KineticMainParams
  use, intrinsic :: iso_c_binding, only: c_int, c_double
  implicit none
  interface
    subroutine kaos_ref_uniforms(r, u1, u2) bind(c, name='kaos_ref_uniforms')
      import :: c_int, c_double
      integer(c_int), value :: r
      real(c_double), intent(out) :: u1, u2
    end subroutine kaos_ref_uniforms
  end interface
contains
  subroutine GaussianRNGSub
    real(c_double) :: a, b
    call kaos_ref_uniforms(int(rank, c_int), a, b)
    UniformRN1(1) = a
    UniformRN2(1) = b
    GaussianRN(1) = sqrt(-2d0*log(UniformRN1(1)))*cos(2d0*pi*UniformRN2(1))
  end subroutine GaussianRNGSub
end module GaussianRNG
