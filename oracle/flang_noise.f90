! flang_noise.f90 - the U[0,1) array the reference's LEQ_BICGS0 draws for
! Rtilde (model/leq_bicgs.f:298: call random_number(Rtilde(:))), regenerated
! with the same compiler runtime.
!
! TEST INFRASTRUCTURE ONLY.  The reference harness (ref_harness.f90, op 1)
! seeds the runtime's generator with seed + 7919*i for each element i of the
! seed array before calling LEQ_BICGS; the intrinsic random_number then
! fills the whole DIMENSION_3 array in one call.  This routine repeats both
! steps, so a test can hand the HIP solver and the oracle the exact noise the
! reference used at any size, instead of shipping the array itself.  It is
! plain Fortran of this project's own; only the intrinsics come from the
! compiler's runtime (the same one that built the reference harness).
subroutine oracle_flang_noise(n, seed, out) bind(C, name="oracle_flang_noise")
  use iso_c_binding, only: c_int, c_double
  implicit none
  integer(c_int), value :: n, seed
  real(c_double), intent(out) :: out(n)
  integer :: nseed, i
  integer, allocatable :: seedarr(:)
  double precision, allocatable :: u(:)
  call random_seed(size=nseed)
  allocate(seedarr(nseed), u(n))
  do i = 1, nseed
     seedarr(i) = seed + 7919 * i
  end do
  call random_seed(put=seedarr)
  call random_number(u)
  out = u
  deallocate(seedarr, u)
end subroutine oracle_flang_noise
