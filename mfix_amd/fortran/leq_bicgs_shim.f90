! Fortran drop-in for MFIX's LEQ_BICGS (model/leq_bicgs.f:23-90).
!
! SUBROUTINE LEQ_BICGS keeps the reference's external (non-module) name and
! dummy argument list, so SOLVE_LIN_EQ (model/solve_lin_eq.f:151-152) links
! against it unchanged; the arrays are taken as assumed-size, which accepts
! the contiguous slices A_M(:,:,M) / B_M(:,M) SOLVE_LIN_EQ passes without a
! copy.  The solve runs on the GPU through the C ABI of libmfix_leq.so
! (include/mfix_leq.h).  MFIX_LEQ_GRID_INIT replaces the module state the
! reference routine reads (compar/geometry/indices, set by GRIDMAP_INIT and
! SET_INCREMENTS) and must be called once after the grid is known.
!
! Built inside the MFIX tree with -DMFIX_LEQSOL_ITER_TOT (preprocessed), the
! drop-in also adds each solve's count to leqsol's ITER_TOT(VNO), as
! LEQ_BICGS0 does (model/leq_bicgs.f:458, 491), so REPORT_SOLVER_STATS
! (model/leqsol_mod.f:50-84) prints the GPU solves' iterations.  Built on its
! own, the library keeps the counts (mfix_leq_iter_tot_f).
module mfix_leq_c
  use iso_c_binding
  implicit none
  integer(c_int), parameter :: MFIX_LEQ_EBADPC = -101
  integer(c_int), parameter :: MFIX_LEQ_EDGTSV = -106

  type, bind(c) :: mfix_leq_grid_t
     integer(c_int) :: imax, jmax, kmax, no_k
     integer(c_int) :: cyclic_x, cyclic_y, cyclic_z
     integer(c_int) :: is_serial, minimize_dotproducts, icheck_bicgs
  end type mfix_leq_grid_t

  interface
     integer(c_int) function c_grid_init(g) bind(c, name='mfix_leq_grid_init')
       import :: c_int, mfix_leq_grid_t
       type(mfix_leq_grid_t), intent(in) :: g
     end function c_grid_init

     integer(c_int) function c_bicgs(vname, vname_len, vno, var, a_m, b_m, cmethod, cmethod_len, tol, pc, &
                                     itmax, ier) bind(c, name='mfix_leq_bicgs')
       import :: c_int, c_double, c_char
       character(kind=c_char), dimension(*), intent(in) :: vname, cmethod, pc
       integer(c_int), value :: vname_len, vno, cmethod_len, itmax
       real(c_double), dimension(*), intent(inout) :: var, a_m, b_m
       real(c_double), value :: tol
       integer(c_int), intent(out) :: ier
     end function c_bicgs

     type(c_ptr) function c_error_string(code) bind(c, name='mfix_leq_error_string')
       import :: c_int, c_ptr
       integer(c_int), value :: code
     end function c_error_string

     integer(c_int) function c_iter_tot(vno) bind(c, name='mfix_leq_iter_tot')
       import :: c_int
       integer(c_int), value :: vno
     end function c_iter_tot
  end interface
end module mfix_leq_c

! Grid descriptor (the serial layout of the reference: one ghost layer,
! DIMENSION_3 = (IMAX+2)(JMAX+2)(KMAX+2), J fastest).
subroutine mfix_leq_grid_init_f(imax, jmax, kmax, no_k, cyclic_x, cyclic_y, cyclic_z, is_serial, &
                                minimize_dotproducts, icheck_bicgs, ierr)
  use mfix_leq_c
  implicit none
  integer, intent(in) :: imax, jmax, kmax, icheck_bicgs
  logical, intent(in) :: no_k, cyclic_x, cyclic_y, cyclic_z, is_serial, minimize_dotproducts
  integer, intent(out) :: ierr
  type(mfix_leq_grid_t) :: g
  g%imax = imax; g%jmax = jmax; g%kmax = kmax
  g%no_k = merge(1, 0, no_k)
  g%cyclic_x = merge(1, 0, cyclic_x); g%cyclic_y = merge(1, 0, cyclic_y); g%cyclic_z = merge(1, 0, cyclic_z)
  g%is_serial = merge(1, 0, is_serial)
  g%minimize_dotproducts = merge(1, 0, minimize_dotproducts)
  g%icheck_bicgs = icheck_bicgs
  ierr = c_grid_init(g)
end subroutine mfix_leq_grid_init_f

! Drop-in LEQ_BICGS: model/leq_bicgs.f:23-69 argument list.
subroutine leq_bicgs(vname, vno, var, a_m, b_m, cmethod, tol, pc, itmax, ier)
  use mfix_leq_c
#ifdef MFIX_LEQSOL_ITER_TOT
  use leqsol, only: iter_tot
#endif
  implicit none
  character(len=*), intent(in) :: vname
  integer, intent(in) :: vno
  double precision, intent(inout) :: var(*)
  double precision, intent(inout) :: a_m(*)
  double precision, intent(inout) :: b_m(*)
  character(len=*), intent(in) :: cmethod
  double precision, intent(in) :: tol
  character(len=4), intent(in) :: pc
  integer, intent(in) :: itmax
  integer, intent(inout) :: ier
  integer(c_int) :: rc, cier
#ifdef MFIX_LEQSOL_ITER_TOT
  integer(c_int) :: before
  before = c_iter_tot(vno)
#endif
  rc = c_bicgs(vname, len(vname, kind=c_int), vno, var, a_m, b_m, cmethod, len(cmethod, kind=c_int), tol, pc, &
               itmax, cier)
  if (rc == MFIX_LEQ_EBADPC) then
     ! the reference logs this and calls mfix_exit (leq_bicgs.f:83-87)
     write (*, *) 'preconditioner option not found - check mfix.dat and readme'
     error stop 1
  else if (rc == MFIX_LEQ_EDGTSV) then
     ! a JK/IJ line sweep failed: fatal in the reference (leqsol_mod.f:991-996)
     write (*, *) 'DGTSV failed in a line sweep of ', vname
     error stop 1
  else if (rc /= 0) then
     write (*, *) 'LEQ_BICGS: GPU solver error ', rc
     error stop 1
  end if
  ier = cier
#ifdef MFIX_LEQSOL_ITER_TOT
  iter_tot(vno) = iter_tot(vno) + (c_iter_tot(vno) - before)
#endif
end subroutine leq_bicgs

! ITER_TOT(VNO) as kept by the reference (leqsol_mod.f:17).
integer function mfix_leq_iter_tot_f(vno)
  use mfix_leq_c
  implicit none
  integer, intent(in) :: vno
  mfix_leq_iter_tot_f = c_iter_tot(vno)
end function mfix_leq_iter_tot_f
