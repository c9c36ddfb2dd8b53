! Fortran drop-in layer of sprgpu.
!
! SPRFC1, SPRFC2, SPRFC3, SPRFS1, SPRFS2, SPRBS1, SPRBS2 and SPRSOL with the
! argument lists of the reference routines (src/SPR/sprfc1.f:5-7,
! sprfc2.f:5-7, sprfc3.f:5-7, sprfs1.f:5-7, sprfs2.f:5-7, sprbs1.f:5-6,
! sprbs2.f:5-6, sprsol.f:5-6), forwarding through ISO_C_BINDING to the C ABI
! declared in include/sprgpu.h.  A program that links this library ahead of
! the reference SPR library keeps its call sequence unchanged: the symbolic
! phase and SPRADM come from the reference, the numeric factor and solves
! run on the GPU.  Integers are INTEGER*8, as in the reference's default
! build (src/SPR/CMakeLists.txt:12-29).
module sprgpu_c
  use iso_c_binding, only: c_int64_t, c_double, c_funptr, c_null_funptr
  implicit none
  interface
    subroutine c_sprfc1(mspar, mtrees, msifa, sm, tol, iwork, rwork, lpu, ierr, mv, mm) &
        bind(C, name='sprgpu_sprfc1')
      import :: c_int64_t, c_double, c_funptr
      integer(c_int64_t) :: mspar(*), mtrees(*), msifa(*), iwork(*), lpu, ierr
      real(c_double) :: sm(*), tol(*), rwork(*)
      type(c_funptr), value :: mv, mm
    end subroutine
    subroutine c_sprfc3(mspar, mtrees, msifa, sm, tol, iwork, rwork, dcor, lpu, ierr, mv, mm) &
        bind(C, name='sprgpu_sprfc3')
      import :: c_int64_t, c_double, c_funptr
      integer(c_int64_t) :: mspar(*), mtrees(*), msifa(*), iwork(*), lpu, ierr
      real(c_double) :: sm(*), tol(*), rwork(*), dcor(*)
      type(c_funptr), value :: mv, mm
    end subroutine
    subroutine c_sprfs1(mspar, mtrees, msifa, sm, b, ldb, nrhs, iwork, rwork, lpu, ierr) &
        bind(C, name='sprgpu_sprfs1')
      import :: c_int64_t, c_double
      integer(c_int64_t) :: mspar(*), mtrees(*), msifa(*), iwork(*), ldb, nrhs, lpu, ierr
      real(c_double) :: sm(*), b(*), rwork(*)
    end subroutine
    subroutine c_sprbs1(mspar, mtrees, msifa, sm, b, ldb, nrhs, rwork, lpu, ierr) &
        bind(C, name='sprgpu_sprbs1')
      import :: c_int64_t, c_double
      integer(c_int64_t) :: mspar(*), mtrees(*), msifa(*), ldb, nrhs, lpu, ierr
      real(c_double) :: sm(*), b(*), rwork(*)
    end subroutine
    subroutine c_sprfc2(mspar, mtrees, msifa, sm, tol, iwork, rwork, lpu, ierr, mv, mm) &
        bind(C, name='sprgpu_sprfc2')
      import :: c_int64_t, c_double, c_funptr
      integer(c_int64_t) :: mspar(*), mtrees(*), msifa(*), iwork(*), lpu, ierr
      real(c_double) :: sm(*), tol(*), rwork(*)
      type(c_funptr), value :: mv, mm
    end subroutine
    subroutine c_sprfs2(mspar, mtrees, msifa, sm, b, ldb, nrhs, iwork, rwork, lpu, ierr) &
        bind(C, name='sprgpu_sprfs2')
      import :: c_int64_t, c_double
      integer(c_int64_t) :: mspar(*), mtrees(*), msifa(*), iwork(*), ldb, nrhs, lpu, ierr
      real(c_double) :: sm(*), b(*), rwork(*)
    end subroutine
    subroutine c_sprbs2(mspar, mtrees, msifa, sm, b, ldb, nrhs, rwork, lpu, ierr) &
        bind(C, name='sprgpu_sprbs2')
      import :: c_int64_t, c_double
      integer(c_int64_t) :: mspar(*), mtrees(*), msifa(*), ldb, nrhs, lpu, ierr
      real(c_double) :: sm(*), b(*), rwork(*)
    end subroutine
    subroutine c_release(sm) bind(C, name='sprgpu_release')
      import :: c_double
      real(c_double) :: sm(*)
    end subroutine
  end interface
contains
  ! Message on unit LPU in the spirit of SPRER1 (silent if LPU < 0).
  subroutine sprgpu_message(name, lpu, ierr)
    character(len=*), intent(in) :: name
    integer(c_int64_t), intent(in) :: lpu, ierr
    if (lpu < 0 .or. ierr == 0) return
    if (ierr > 0) then
      write(lpu, '(/1X,A,A,A,I4)') '*** WARNING from ', name, ' (GPU): IERR =', ierr
    else
      write(lpu, '(/1X,A,A,A,I4)') '*** ERROR from ', name, ' (GPU): IERR =', ierr
    end if
  end subroutine
end module sprgpu_c

subroutine SPRFC1(MSPAR, MTREES, MSIFA, SM, TOL, IWORK, RWORK, LPU, IERR, MVMULT, MMMULT)
  use sprgpu_c
  implicit none
  integer(c_int64_t) :: MSPAR(*), MTREES(*), MSIFA(*), IWORK(*), LPU, IERR
  real(c_double) :: SM(*), TOL(3), RWORK(*)
  external :: MVMULT, MMMULT
  call c_sprfc1(MSPAR, MTREES, MSIFA, SM, TOL, IWORK, RWORK, LPU, IERR, &
                c_null_funptr, c_null_funptr)
  call sprgpu_message('SPRFC1', LPU, IERR)
end subroutine SPRFC1

subroutine SPRFC3(MSPAR, MTREES, MSIFA, SM, TOL, IWORK, RWORK, DCOR, LPU, IERR, MVMULT, MMMULT)
  use sprgpu_c
  implicit none
  integer(c_int64_t) :: MSPAR(*), MTREES(*), MSIFA(*), IWORK(*), LPU, IERR
  real(c_double) :: SM(*), TOL(4), RWORK(*), DCOR(*)
  external :: MVMULT, MMMULT
  call c_sprfc3(MSPAR, MTREES, MSIFA, SM, TOL, IWORK, RWORK, DCOR, LPU, IERR, &
                c_null_funptr, c_null_funptr)
  call sprgpu_message('SPRFC3', LPU, IERR)
end subroutine SPRFC3

subroutine SPRFS1(MSPAR, MTREES, MSIFA, SM, B, LDB, NRHS, IWORK, RWORK, LPU, IERR)
  use sprgpu_c
  implicit none
  integer(c_int64_t) :: MSPAR(*), MTREES(*), MSIFA(*), IWORK(*), LDB, NRHS, LPU, IERR
  real(c_double) :: SM(*), B(*), RWORK(*)
  call c_sprfs1(MSPAR, MTREES, MSIFA, SM, B, LDB, NRHS, IWORK, RWORK, LPU, IERR)
  call sprgpu_message('SPRFS1', LPU, IERR)
end subroutine SPRFS1

subroutine SPRBS1(MSPAR, MTREES, MSIFA, SM, B, LDB, NRHS, RWORK, LPU, IERR)
  use sprgpu_c
  implicit none
  integer(c_int64_t) :: MSPAR(*), MTREES(*), MSIFA(*), LDB, NRHS, LPU, IERR
  real(c_double) :: SM(*), B(*), RWORK(*)
  call c_sprbs1(MSPAR, MTREES, MSIFA, SM, B, LDB, NRHS, RWORK, LPU, IERR)
  call sprgpu_message('SPRBS1', LPU, IERR)
end subroutine SPRBS1

subroutine SPRFC2(MSPAR, MTREES, MSIFA, SM, TOL, IWORK, RWORK, LPU, IERR, MVMULT, MMMULT)
  use sprgpu_c
  implicit none
  integer(c_int64_t) :: MSPAR(*), MTREES(*), MSIFA(*), IWORK(*), LPU, IERR
  real(c_double) :: SM(*), TOL(3), RWORK(*)
  external :: MVMULT, MMMULT
  call c_sprfc2(MSPAR, MTREES, MSIFA, SM, TOL, IWORK, RWORK, LPU, IERR, &
                c_null_funptr, c_null_funptr)
  call sprgpu_message('SPRFC2', LPU, IERR)
end subroutine SPRFC2

subroutine SPRFS2(MSPAR, MTREES, MSIFA, SM, B, LDB, NRHS, IWORK, RWORK, LPU, IERR)
  use sprgpu_c
  implicit none
  integer(c_int64_t) :: MSPAR(*), MTREES(*), MSIFA(*), IWORK(*), LDB, NRHS, LPU, IERR
  real(c_double) :: SM(*), B(*), RWORK(*)
  call c_sprfs2(MSPAR, MTREES, MSIFA, SM, B, LDB, NRHS, IWORK, RWORK, LPU, IERR)
  call sprgpu_message('SPRFS2', LPU, IERR)
end subroutine SPRFS2

subroutine SPRBS2(MSPAR, MTREES, MSIFA, SM, B, LDB, NRHS, RWORK, LPU, IERR)
  use sprgpu_c
  implicit none
  integer(c_int64_t) :: MSPAR(*), MTREES(*), MSIFA(*), LDB, NRHS, LPU, IERR
  real(c_double) :: SM(*), B(*), RWORK(*)
  call c_sprbs2(MSPAR, MTREES, MSIFA, SM, B, LDB, NRHS, RWORK, LPU, IERR)
  call sprgpu_message('SPRBS2', LPU, IERR)
end subroutine SPRBS2

! sprsol.f:29-50 with the GPU factor and solves
subroutine SPRSOL(IOP, MSPAR, MTREES, MSIFA, SM, B, LDB, NRHS, TOL, IWORK, RWORK, LPU, IERR)
  use sprgpu_c
  implicit none
  integer(c_int64_t) :: IOP, MSPAR(*), MTREES(*), MSIFA(*), IWORK(*), LDB, NRHS, LPU, IERR
  real(c_double) :: SM(*), B(*), TOL(3), RWORK(*)
  IERR = 0
  if (IOP <= 3) then
    call c_sprfc1(MSPAR, MTREES, MSIFA, SM, TOL, IWORK, RWORK, LPU, IERR, &
                  c_null_funptr, c_null_funptr)
    call sprgpu_message('SPRFC1', LPU, IERR)
  end if
  if (IOP >= 2 .and. IOP <= 5 .and. IERR >= 0) then
    call c_sprfs1(MSPAR, MTREES, MSIFA, SM, B, LDB, NRHS, IWORK, RWORK, LPU, IERR)
    call sprgpu_message('SPRFS1', LPU, IERR)
  end if
  if (IOP >= 3 .and. IOP /= 5 .and. IERR >= 0) then
    call c_sprbs1(MSPAR, MTREES, MSIFA, SM, B, LDB, NRHS, RWORK, LPU, IERR)
    call sprgpu_message('SPRBS1', LPU, IERR)
  end if
end subroutine SPRSOL

! Releases the device factor the drop-in layer keeps for SM (no counterpart in
! the reference, which keeps nothing between calls): for a caller that is
! done with SM, or that will reuse its memory for another model.
subroutine SPRGPU_RELEASE(SM)
  use sprgpu_c
  implicit none
  real(c_double) :: SM(*)
  call c_release(SM)
end subroutine SPRGPU_RELEASE
