! A host program in the shape of an SPR application, for the drop-in layer.
!
! It does what a Fortran program built on the reference SPR library does
! (src/SPR/sprprp.f ... sprsol.f), with every integer INTEGER*8 as in the
! reference's default build:
!   1. reads a finite-element model (SAM arrays MADOF, MSC, MPMNPC, MMNPC,
!      the element matrices and the nodal loads) from a stream file;
!   2. runs the reference's symbolic phase SPRPRP, SPRSAS, SPRRNM, SPRTRS,
!      SPRSMB, SPRPMP and calls SPRADM once per element (from
!      libspr_host.so: the reference routines compiled unchanged);
!   3. runs the numeric path through the drop-in library (libsprgpu_f.so,
!      the same names and argument lists as the reference): SPRSOL(IOP=3);
!      SPRFC1 followed by SPRFS1 and SPRBS1 as separate calls; SPRFC3 with
!      TOL(1) from the input and its DCOR; SPRGPU_RELEASE for each SM.
!      The MVMULT/MMMULT arguments of SPRFC1/SPRFC3 (the reference's per-
!      supernode kernel plugins, LMVY18/LMMY18 in SPRSOL) are accepted and
!      never called by the drop-in: the ones passed here stop the program
!      if they ever are;
!   4. writes MEQN, the solutions, MSPAR, TOL, IERR and DCOR to a stream file.
!
!   usage: spr_driver <model file> <output file>
!
! Input (stream, 8-byte integers and reals, in this order): NANOD, NEL,
! NDOF, NMMNPC, NEM, NRHS, TOL1; MADOF(NANOD+1), MSC(NDOF), MPMNPC(NEL+1),
! MMNPC(NMMNPC), NEDOF(NEL), EMOFF(NEL) (0-based offsets of each
! column-major element matrix in EM), EM(NEM), F(NDOF,NRHS).
program spr_driver
  implicit none
  integer, parameter :: nspar = 60, lpu = 6
  integer :: nanod, nel, ndof, nmmnpc, nem, nrhs, neq, i, k, iel, ierr
  integer :: nmsica, ntrees, nmsifa, niwork, mfront, maxsup, nzerol, nsm, ierr1, ierr2, ierr3
  double precision :: tol1
  integer, allocatable :: madof(:), msc(:), mpmnpc(:), mmnpc(:), nedof(:), emoff(:)
  integer, allocatable :: mpar(:), mspar(:), meqn(:), msica(:), mtrees(:), msifa(:), mdofnc(:)
  integer, allocatable :: iwork(:), minex(:), ms1(:), ms2(:), ms3(:)
  integer :: mpmceq(1), mmceq(1)
  double precision, allocatable :: em(:), f(:,:), sm(:), sv(:), b(:,:), x1(:,:), x2(:,:)
  double precision, allocatable :: sm1(:), sm2(:), sm3(:), rwork(:), dcor(:)
  double precision :: ttcc(1), rinfo(4), tol(3), tol3(4)
  character(len=512) :: fin, fout
  external :: SPRPRP, SPRSAS, SPRRNM, SPRTRS, SPRSMB, SPRPMP, SPRADM
  external :: SPRSOL, SPRFC1, SPRFS1, SPRBS1, SPRFC3, SPRGPU_RELEASE
  external :: HOST_MVMULT, HOST_MMMULT

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(unit=10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) nanod, nel, ndof, nmmnpc, nem, nrhs, tol1
  allocate(madof(nanod+1), msc(ndof), mpmnpc(nel+1), mmnpc(nmmnpc), nedof(nel), emoff(nel))
  allocate(em(nem), f(ndof, nrhs))
  read(10) madof, msc, mpmnpc, mmnpc, nedof, emoff, em, f
  close(10)

  ! ---- symbolic phase (reference routines) --------------------------------
  allocate(mpar(50), mspar(nspar), meqn(ndof), minex(nanod))
  mpar = 0
  mpar(1) = nanod
  mpar(2) = nel
  mpar(3) = ndof
  mspar = 0
  mpmceq = 1
  mmceq = 0
  do i = 1, nanod
    minex(i) = i
  end do
  allocate(iwork(2*ndof + nanod))
  call SPRPRP(madof, minex, mpmnpc, mmnpc, mpmceq, mmceq, msc, nspar, lpu, mpar, mspar, &
              meqn, iwork, ierr)
  call check('SPRPRP', ierr)
  nmsica = mspar(35)
  ntrees = mspar(36)
  allocate(msica(nmsica))
  call SPRSAS(mpar, mpmnpc, mmnpc, madof, msc, mpmceq, mmceq, meqn, mspar, msica, iwork, &
              nspar, lpu, ierr)
  call check('SPRSAS', ierr)
  deallocate(iwork)
  allocate(iwork(mspar(37)))
  call SPRRNM(mspar, msica, iwork, lpu, ierr)
  call check('SPRRNM', ierr)
  deallocate(iwork)
  allocate(mtrees(ntrees))
  niwork = mspar(38)
  allocate(iwork(niwork))
  call SPRTRS(mspar, msica, mtrees, meqn, nspar, nmsica, ntrees, ndof, niwork, iwork, rinfo, &
              lpu, ierr)
  call check('SPRTRS', ierr)
  deallocate(iwork)
  nmsifa = mspar(3)
  allocate(msifa(nmsifa))
  allocate(iwork(3*mspar(6) + 2*mspar(8) + 1))
  call SPRSMB(mspar, msica, mtrees, msifa, meqn, iwork, lpu, ierr)
  call check('SPRSMB', ierr)
  deallocate(iwork)
  neq = mspar(8)
  allocate(mdofnc(2*neq + 1))
  call SPRPMP(msica, mtrees, meqn, mpar, mspar, mdofnc)

  ! ---- element assembly (reference SPRADM, once per element) ---------------
  nzerol = mspar(16)
  nsm = neq + nzerol
  allocate(sm(nsm), sv(neq))
  sm = 0.0d0
  sv = 0.0d0
  allocate(iwork(mspar(39) + 16))
  do iel = 1, nel
    call SPRADM(em(emoff(iel)+1), ttcc, mpar, mspar, madof, meqn, mpmnpc, mmnpc, mpmceq, &
                mmceq, msica, mtrees, msifa, mdofnc, sm, sv, iwork, iel, nedof(iel), lpu, &
                0_8, ierr)
    call check('SPRADM', ierr)
  end do
  deallocate(iwork)

  ! right-hand sides in equation order
  allocate(b(neq, nrhs))
  b = 0.0d0
  do k = 1, nrhs
    do i = 1, ndof
      if (meqn(i) > 0) b(meqn(i), k) = f(i, k)
    end do
  end do

  ! ---- numeric phase through the drop-in names ---------------------------
  mfront = mspar(14)
  maxsup = mspar(13)
  allocate(iwork(max(mfront + 1, maxsup) + 1), rwork(max(mspar(17), mfront) + 1))
  ! 1. SPRSOL, factor + forward + backward
  allocate(sm1(nsm), x1(neq, nrhs), ms1(nspar))
  sm1 = sm
  x1 = b
  ms1 = mspar
  tol = 0.0d0
  call SPRSOL(3_8, ms1, mtrees, msifa, sm1, x1, neq, nrhs, tol, iwork, rwork, lpu, ierr1)
  call SPRGPU_RELEASE(sm1)
  ! 2. SPRFC1, then SPRFS1 and SPRBS1 as separate calls
  allocate(sm2(nsm), x2(neq, nrhs), ms2(nspar))
  sm2 = sm
  x2 = b
  ms2 = mspar
  tol = 0.0d0
  call SPRFC1(ms2, mtrees, msifa, sm2, tol, iwork, rwork, lpu, ierr2, HOST_MVMULT, HOST_MMMULT)
  if (ierr2 >= 0) call SPRFS1(ms2, mtrees, msifa, sm2, x2, neq, nrhs, iwork, rwork, lpu, ierr2)
  if (ierr2 >= 0) call SPRBS1(ms2, mtrees, msifa, sm2, x2, neq, nrhs, rwork, lpu, ierr2)
  call SPRGPU_RELEASE(sm2)
  ! 3. SPRFC3 with the given TOL(1), TOL(4) = 1
  allocate(sm3(nsm), ms3(nspar), dcor(2*neq))
  sm3 = sm
  ms3 = mspar
  dcor = 0.0d0
  tol3 = (/ tol1, 0.0d0, 0.0d0, 1.0d0 /)
  call SPRFC3(ms3, mtrees, msifa, sm3, tol3, iwork, rwork, dcor, lpu, ierr3, HOST_MVMULT, HOST_MMMULT)
  call SPRGPU_RELEASE(sm3)

  open(unit=11, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(11) neq, nrhs, meqn
  write(11) ierr1, ms1, x1
  write(11) ierr2, ms2, x2, tol
  write(11) ierr3, ms3, tol3, dcor, sm3(1:neq)
  close(11)
  write(*, '(A,I8,A,3I4)') ' spr_driver: NEQ =', neq, ', IERR =', ierr1, ierr2, ierr3

contains

  subroutine check(name, code)
    character(len=*), intent(in) :: name
    integer, intent(in) :: code
    if (code < 0) then
      write(*, '(A,A,A,I6)') ' spr_driver: ', name, ' failed, IERR =', code
      stop 2
    end if
  end subroutine check

end program spr_driver

! MVMULT / MMMULT plugins handed to SPRFC1 and SPRFC3 (lmvy18.f:5, lmmy18.f:5
! signatures); the GPU replaces the whole factorization, so neither may run.
subroutine HOST_MVMULT(m, n, y, d, apnt, a)
  implicit none
  integer :: m, n, apnt(*)
  double precision :: y(*), d(*), a(*)
  write(*, '(A)') ' spr_driver: MVMULT was called'
  stop 3
end subroutine HOST_MVMULT

subroutine HOST_MMMULT(m, n, q, xpnt, x, d, y, ldy)
  implicit none
  integer :: m, n, q, xpnt(*), ldy
  double precision :: x(*), d(*), y(*)
  write(*, '(A)') ' spr_driver: MMMULT was called'
  stop 3
end subroutine HOST_MMMULT
