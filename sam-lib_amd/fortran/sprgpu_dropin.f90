! Fortran drop-in layer of sprgpu.
!
! SPRFC1, SPRFC2, SPRFC3, SPRFC4, SPRFS1, SPRFS2, SPRBS1, SPRBS2 and SPRSOL with the
! argument lists of the reference routines (src/SPR/sprfc1.f:5-7,
! sprfc2.f:5-7, sprfc3.f:5-7, sprfc4.f:5-7, sprfs1.f:5-7, sprfs2.f:5-7, sprbs1.f:5-6,
! sprbs2.f:5-6, sprsol.f:5-6), forwarding through ISO_C_BINDING to the C ABI
! declared in include/sprgpu.h.  A program that links this library ahead of
! the reference SPR library keeps its call sequence unchanged: the symbolic
! phase and SPRADM come from the reference, the numeric factor and solves
! run on the GPU.  Integers are INTEGER*8, as in the reference's default
! build (src/SPR/CMakeLists.txt:12-29).
module sprgpu_c
  use iso_c_binding, only: c_int64_t, c_double, c_funptr, c_null_funptr, c_ptr, c_char, &
                           c_size_t, c_f_pointer
  implicit none
  interface
    function c_lierr() bind(C, name='sprgpu_dropin_lierr')
      import :: c_int64_t
      integer(c_int64_t) :: c_lierr
    end function
    function c_last_error() bind(C, name='sprgpu_last_error')
      import :: c_ptr
      type(c_ptr) :: c_last_error
    end function
    function c_strlen(p) bind(C, name='strlen')
      import :: c_ptr, c_size_t
      type(c_ptr), value :: p
      integer(c_size_t) :: c_strlen
    end function
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
    subroutine c_sprfc4(mspar, mtrees, msifa, sm, tol, iwork, rwork, dcor, lpu, ierr, mv, mm) &
        bind(C, name='sprgpu_sprfc4')
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
  ! Messages on unit LPU as the reference drivers write them (silent if
  ! LPU < 0): SPRER1, the reference's own message routine (src/SPR/sprer1.f,
  ! linked from the host library with the symbolic phase), with the codes
  ! the reference passes (sprfc1.f:143,234,257,264; sprfc2.f:136,227,250;
  ! sprfc3.f:142,233,256,265; sprfs1.f:140,229; sprbs1.f:135,214), plus the
  ! pivot diagnostic of IERR -3 and the SPRFC3 lists in the same layout as
  ! the reference's, and one message of this layer for a HIP error (IERR -5).
  subroutine factor_messages(name, mspar, tol, dcor, neq, lpu, ierr)
    character(len=*), intent(in) :: name
    integer(c_int64_t), intent(in) :: mspar(*), neq, lpu, ierr
    real(c_double), intent(in) :: tol(*), dcor(2, *)
    integer(c_int64_t) :: lierr, i
    external :: SPRER1
    if (ierr == -1) call SPRER1(30_c_int64_t, name, 0_c_int64_t, 0_c_int64_t, 0_c_int64_t, &
                                lpu, ierr)
    if (ierr == -2) call SPRER1(36_c_int64_t, name, 0_c_int64_t, 0_c_int64_t, 0_c_int64_t, &
                                lpu, ierr)
    if (ierr == -5) call hip_message(name, lpu)
    if (ierr == -3) then
      lierr = c_lierr()
      if (lierr == -31 .or. lierr == -33) then
        ! SPRFC3, SPRFC4: SUPEQT counts the replaced pivots, MSPAR(27) = -SUPEQT
        call SPRER1(40_c_int64_t, name, mspar(26), -mspar(27), lierr, lpu, ierr)
        if (lpu >= 0) then
          write(lpu, 700) tol(1), -mspar(27)
          do i = 1, neq
            if (dcor(2, i) /= 0.0d0) write(lpu, 710) i, dcor(1, i), dcor(2, i)
          end do
          write(lpu, '(A)') ''
        end if
      else
        call SPRER1(40_c_int64_t, name, mspar(26), mspar(27), lierr, lpu, ierr)
        if (lpu >= 0 .and. lierr == -30) write(lpu, 600) tol(1), tol(2), tol(3), &
                                                    abs(tol(2) / tol(3))
        if (lpu >= 0 .and. lierr == -32 .and. name == 'SPRFC2') write(lpu, 601) tol(1), &
                                                    tol(2), tol(3), tol(2) / abs(tol(3))
      end if
    else if (ierr == 3) then
      call SPRER1(60_c_int64_t, name, 0_c_int64_t, mspar(27), mspar(26), lpu, ierr)
      if (lpu >= 0 .and. neq > 0) then  ! SPRFC3: its negative pivots
        do i = 1, neq
          if (dcor(1, i) < 0.0d0) write(lpu, 720) i, dcor(1, i)
        end do
        write(lpu, '(A)') ''
      end if
    end if
600 format(5x, 'ZERO FACTOR THRESHOLD,   TOL(1) = ', 1pe12.5, &
           / 5x, 'DIAGONAL ENTRY (PIVOT),   DIAG  = ', 1pe12.5, &
           / 5x, 'PREVIOUS PIVOT,          ADIAG  = ', 1pe12.5, &
           / 5x, 'DIAGONAL DECAY, ABS(DIAG/ADIAG) = ', 1pe12.5, ' < TOL(1)')
601 format(5x, 'ZERO FACTOR THRESHOLD,   TOL(1) = ', 1pe12.5, &
           / 5x, 'DIAGONAL ENTRY (PIVOT),   DIAG  = ', 1pe12.5, &
           / 5x, 'PREVIOUS PIVOT,          ADIAG  = ', 1pe12.5, &
           / 5x, 'DIAGONAL DECAY, DIAG/ABS(ADIAG) = ', 1pe12.5, ' < TOL(1)')
700 format(5x, 'ZERO FACTOR THRESHOLD,   TOL(1) = ', 1pe12.5, &
           / 5x, 'NUMBER OF SINGULARITIES FOUND   = ', i8, &
           // 5x, 'EQUATION   DIAGONAL ENTRY  INSERTED NEW ENTRY')
710 format(i10, 1p2e20.8)
720 format(5x, 'EQUATION #', i8, ' :  DIAG =', 1pe16.8)
  end subroutine

  subroutine solve_messages(name, lpu, ierr)
    character(len=*), intent(in) :: name
    integer(c_int64_t), intent(in) :: lpu, ierr
    external :: SPRER1
    if (ierr == -1) call SPRER1(30_c_int64_t, name, 0_c_int64_t, 0_c_int64_t, 0_c_int64_t, &
                                lpu, ierr)
    if (ierr == -2) call SPRER1(37_c_int64_t, name, 0_c_int64_t, 0_c_int64_t, 0_c_int64_t, &
                                lpu, ierr)
    if (ierr == -5) call hip_message(name, lpu)
  end subroutine

  ! IERR -5 (a HIP runtime error, no counterpart in the reference): the
  ! library's reason (sprgpu_last_error)
  subroutine hip_message(name, lpu)
    character(len=*), intent(in) :: name
    integer(c_int64_t), intent(in) :: lpu
    type(c_ptr) :: p
    character(kind=c_char), pointer :: txt(:)
    integer :: n, i
    character(len=256) :: line
    if (lpu < 0) return
    p = c_last_error()
    n = int(min(c_strlen(p), 256_c_size_t))
    call c_f_pointer(p, txt, [max(n, 1)])
    line = ' '
    do i = 1, n
      line(i:i) = txt(i)
    end do
    write(lpu, '(///A,A/,4X,A,A)') ' *** ERROR RETURN FROM  S A M / SPR  LIBRARY ROUTINE : ', &
      name, ' GPU RUNTIME ERROR, IERR = -5: ', line(1:max(n, 1))
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
  call factor_messages('SPRFC1', MSPAR, TOL, TOL, 0_c_int64_t, LPU, IERR)
end subroutine SPRFC1

subroutine SPRFC3(MSPAR, MTREES, MSIFA, SM, TOL, IWORK, RWORK, DCOR, LPU, IERR, MVMULT, MMMULT)
  use sprgpu_c
  implicit none
  integer(c_int64_t) :: MSPAR(*), MTREES(*), MSIFA(*), IWORK(*), LPU, IERR
  real(c_double) :: SM(*), TOL(4), RWORK(*), DCOR(*)
  external :: MVMULT, MMMULT
  call c_sprfc3(MSPAR, MTREES, MSIFA, SM, TOL, IWORK, RWORK, DCOR, LPU, IERR, &
                c_null_funptr, c_null_funptr)
  call factor_messages('SPRFC3', MSPAR, TOL, DCOR, MSPAR(8), LPU, IERR)
end subroutine SPRFC3

subroutine SPRFC4(MSPAR, MTREES, MSIFA, SM, TOL, IWORK, RWORK, DCOR, LPU, IERR, MVMULT, MMMULT)
  use sprgpu_c
  implicit none
  integer(c_int64_t) :: MSPAR(*), MTREES(*), MSIFA(*), IWORK(*), LPU, IERR
  real(c_double) :: SM(*), TOL(4), RWORK(*), DCOR(*)
  external :: MVMULT, MMMULT
  call c_sprfc4(MSPAR, MTREES, MSIFA, SM, TOL, IWORK, RWORK, DCOR, LPU, IERR, &
                c_null_funptr, c_null_funptr)
  call factor_messages('SPRFC4', MSPAR, TOL, DCOR, MSPAR(8), LPU, IERR)
end subroutine SPRFC4

subroutine SPRFS1(MSPAR, MTREES, MSIFA, SM, B, LDB, NRHS, IWORK, RWORK, LPU, IERR)
  use sprgpu_c
  implicit none
  integer(c_int64_t) :: MSPAR(*), MTREES(*), MSIFA(*), IWORK(*), LDB, NRHS, LPU, IERR
  real(c_double) :: SM(*), B(*), RWORK(*)
  call c_sprfs1(MSPAR, MTREES, MSIFA, SM, B, LDB, NRHS, IWORK, RWORK, LPU, IERR)
  call solve_messages('SPRFS1', LPU, IERR)
end subroutine SPRFS1

subroutine SPRBS1(MSPAR, MTREES, MSIFA, SM, B, LDB, NRHS, RWORK, LPU, IERR)
  use sprgpu_c
  implicit none
  integer(c_int64_t) :: MSPAR(*), MTREES(*), MSIFA(*), LDB, NRHS, LPU, IERR
  real(c_double) :: SM(*), B(*), RWORK(*)
  call c_sprbs1(MSPAR, MTREES, MSIFA, SM, B, LDB, NRHS, RWORK, LPU, IERR)
  call solve_messages('SPRBS1', LPU, IERR)
end subroutine SPRBS1

subroutine SPRFC2(MSPAR, MTREES, MSIFA, SM, TOL, IWORK, RWORK, LPU, IERR, MVMULT, MMMULT)
  use sprgpu_c
  implicit none
  integer(c_int64_t) :: MSPAR(*), MTREES(*), MSIFA(*), IWORK(*), LPU, IERR
  real(c_double) :: SM(*), TOL(3), RWORK(*)
  external :: MVMULT, MMMULT
  call c_sprfc2(MSPAR, MTREES, MSIFA, SM, TOL, IWORK, RWORK, LPU, IERR, &
                c_null_funptr, c_null_funptr)
  call factor_messages('SPRFC2', MSPAR, TOL, TOL, 0_c_int64_t, LPU, IERR)
end subroutine SPRFC2

subroutine SPRFS2(MSPAR, MTREES, MSIFA, SM, B, LDB, NRHS, IWORK, RWORK, LPU, IERR)
  use sprgpu_c
  implicit none
  integer(c_int64_t) :: MSPAR(*), MTREES(*), MSIFA(*), IWORK(*), LDB, NRHS, LPU, IERR
  real(c_double) :: SM(*), B(*), RWORK(*)
  call c_sprfs2(MSPAR, MTREES, MSIFA, SM, B, LDB, NRHS, IWORK, RWORK, LPU, IERR)
  call solve_messages('SPRFS2', LPU, IERR)
end subroutine SPRFS2

subroutine SPRBS2(MSPAR, MTREES, MSIFA, SM, B, LDB, NRHS, RWORK, LPU, IERR)
  use sprgpu_c
  implicit none
  integer(c_int64_t) :: MSPAR(*), MTREES(*), MSIFA(*), LDB, NRHS, LPU, IERR
  real(c_double) :: SM(*), B(*), RWORK(*)
  call c_sprbs2(MSPAR, MTREES, MSIFA, SM, B, LDB, NRHS, RWORK, LPU, IERR)
  call solve_messages('SPRBS2', LPU, IERR)
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
    call factor_messages('SPRFC1', MSPAR, TOL, TOL, 0_c_int64_t, LPU, IERR)
  end if
  if (IOP >= 2 .and. IOP <= 5 .and. IERR >= 0) then
    call c_sprfs1(MSPAR, MTREES, MSIFA, SM, B, LDB, NRHS, IWORK, RWORK, LPU, IERR)
    call solve_messages('SPRFS1', LPU, IERR)
  end if
  if (IOP >= 3 .and. IOP /= 5 .and. IERR >= 0) then
    call c_sprbs1(MSPAR, MTREES, MSIFA, SM, B, LDB, NRHS, RWORK, LPU, IERR)
    call solve_messages('SPRBS1', LPU, IERR)
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
