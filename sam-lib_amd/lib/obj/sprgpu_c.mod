﻿!mod$ v1 sum:b5d721015e6e24cd
!need$ 0bde2ac47243ead2 i iso_c_binding
module sprgpu_c
use,intrinsic::iso_c_binding,only:c_int64_t
use,intrinsic::iso_c_binding,only:c_double
use,intrinsic::iso_c_binding,only:c_funptr
use,intrinsic::iso_c_binding,only:c_null_funptr
interface
subroutine c_sprfc1(mspar,mtrees,msifa,sm,tol,iwork,rwork,lpu,ierr,mv,mm) bind(c,name="sprgpu_sprfc1")
import::c_funptr
integer(8)::mspar(1_8:*)
integer(8)::mtrees(1_8:*)
integer(8)::msifa(1_8:*)
real(8)::sm(1_8:*)
real(8)::tol(1_8:*)
integer(8)::iwork(1_8:*)
real(8)::rwork(1_8:*)
integer(8)::lpu
integer(8)::ierr
type(c_funptr),value::mv
type(c_funptr),value::mm
end
end interface
interface
subroutine c_sprfc3(mspar,mtrees,msifa,sm,tol,iwork,rwork,dcor,lpu,ierr,mv,mm) bind(c,name="sprgpu_sprfc3")
import::c_funptr
integer(8)::mspar(1_8:*)
integer(8)::mtrees(1_8:*)
integer(8)::msifa(1_8:*)
real(8)::sm(1_8:*)
real(8)::tol(1_8:*)
integer(8)::iwork(1_8:*)
real(8)::rwork(1_8:*)
real(8)::dcor(1_8:*)
integer(8)::lpu
integer(8)::ierr
type(c_funptr),value::mv
type(c_funptr),value::mm
end
end interface
interface
subroutine c_sprfs1(mspar,mtrees,msifa,sm,b,ldb,nrhs,iwork,rwork,lpu,ierr) bind(c,name="sprgpu_sprfs1")
integer(8)::mspar(1_8:*)
integer(8)::mtrees(1_8:*)
integer(8)::msifa(1_8:*)
real(8)::sm(1_8:*)
real(8)::b(1_8:*)
integer(8)::ldb
integer(8)::nrhs
integer(8)::iwork(1_8:*)
real(8)::rwork(1_8:*)
integer(8)::lpu
integer(8)::ierr
end
end interface
interface
subroutine c_sprbs1(mspar,mtrees,msifa,sm,b,ldb,nrhs,rwork,lpu,ierr) bind(c,name="sprgpu_sprbs1")
integer(8)::mspar(1_8:*)
integer(8)::mtrees(1_8:*)
integer(8)::msifa(1_8:*)
real(8)::sm(1_8:*)
real(8)::b(1_8:*)
integer(8)::ldb
integer(8)::nrhs
real(8)::rwork(1_8:*)
integer(8)::lpu
integer(8)::ierr
end
end interface
contains
subroutine sprgpu_message(name,lpu,ierr)
character(*,1),intent(in)::name
integer(8),intent(in)::lpu
integer(8),intent(in)::ierr
end
end
