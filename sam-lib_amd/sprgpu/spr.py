"""Host mirror of the reference SPR numeric interface, running on the GPU.

Functions take the same arrays as the Fortran routines they replace (numpy,
INTEGER*8 / float64, 1-based content, SM and B modified in place) and return
IERR, so that code written against the reference call sequence reads the
same:

    sprfc1(mspar, mtrees, msifa, sm, tol)            # SPRFC1, sprfc1.f:5-7
    sprfc3(mspar, mtrees, msifa, sm, tol, dcor)      # SPRFC3, sprfc3.f:5-7
    sprfc4(mspar, mtrees, msifa, sm, tol, dcor)      # SPRFC4, sprfc4.f:5-7 (CC')
    sprfs1(mspar, mtrees, msifa, sm, b)              # SPRFS1, sprfs1.f:5-7
    sprbs1(mspar, mtrees, msifa, sm, b)              # SPRBS1, sprbs1.f:5-6
    sprsol(iop, mspar, mtrees, msifa, sm, b, tol)    # SPRSOL, sprsol.f:5-6

B is a Fortran-ordered (LDB, NRHS) float64 array (or a 1-D vector).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import lib

I8 = np.int64


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _ref(v):
    return C.byref(C.c_int64(int(v)))


def _check_int(*arrs):
    for a in arrs:
        if a.dtype != I8 or not a.flags.c_contiguous:
            raise TypeError("MSPAR/MTREES/MSIFA must be contiguous int64 arrays")


def _check_real(*arrs):
    for a in arrs:
        if a.dtype != np.float64:
            raise TypeError("SM/B/TOL must be float64 arrays")


def _b_view(b):
    if b.ndim == 1:
        return b, b.shape[0], 1
    if not b.flags.f_contiguous:
        raise TypeError("B must be Fortran-ordered (LDB, NRHS)")
    return b, b.shape[0], b.shape[1]


def sprfc1(mspar, mtrees, msifa, sm, tol, lpu: int = -1) -> int:
    _check_int(mspar, mtrees, msifa)
    _check_real(sm, tol)
    ierr = C.c_int64(0)
    lib().sprgpu_sprfc1(_p(mspar), _p(mtrees), _p(msifa), _p(sm), _p(tol), None, None,
                        _ref(lpu), C.byref(ierr), None, None)
    return int(ierr.value)


def sprfc3(mspar, mtrees, msifa, sm, tol, dcor, lpu: int = -1) -> int:
    _check_int(mspar, mtrees, msifa)
    _check_real(sm, tol, dcor)
    if tol.size < 4:
        raise ValueError("SPRFC3 needs TOL(4)")
    ierr = C.c_int64(0)
    lib().sprgpu_sprfc3(_p(mspar), _p(mtrees), _p(msifa), _p(sm), _p(tol), None, None,
                        _p(dcor), _ref(lpu), C.byref(ierr), None, None)
    return int(ierr.value)


def sprfc4(mspar, mtrees, msifa, sm, tol, dcor, lpu: int = -1) -> int:
    _check_int(mspar, mtrees, msifa)
    _check_real(sm, tol, dcor)
    if tol.size < 4:
        raise ValueError("SPRFC4 needs TOL(4)")
    ierr = C.c_int64(0)
    lib().sprgpu_sprfc4(_p(mspar), _p(mtrees), _p(msifa), _p(sm), _p(tol), None, None,
                        _p(dcor), _ref(lpu), C.byref(ierr), None, None)
    return int(ierr.value)


def sprfs1(mspar, mtrees, msifa, sm, b, lpu: int = -1) -> int:
    _check_int(mspar, mtrees, msifa)
    _check_real(sm, b)
    b, ldb, nrhs = _b_view(b)
    ierr = C.c_int64(0)
    lib().sprgpu_sprfs1(_p(mspar), _p(mtrees), _p(msifa), _p(sm), _p(b), _ref(ldb), _ref(nrhs),
                        None, None, _ref(lpu), C.byref(ierr))
    return int(ierr.value)


def sprbs1(mspar, mtrees, msifa, sm, b, lpu: int = -1) -> int:
    _check_int(mspar, mtrees, msifa)
    _check_real(sm, b)
    b, ldb, nrhs = _b_view(b)
    ierr = C.c_int64(0)
    lib().sprgpu_sprbs1(_p(mspar), _p(mtrees), _p(msifa), _p(sm), _p(b), _ref(ldb), _ref(nrhs),
                        None, _ref(lpu), C.byref(ierr))
    return int(ierr.value)


def sprfc2(mspar, mtrees, msifa, sm, tol, lpu: int = -1) -> int:
    _check_int(mspar, mtrees, msifa)
    _check_real(sm, tol)
    ierr = C.c_int64(0)
    lib().sprgpu_sprfc2(_p(mspar), _p(mtrees), _p(msifa), _p(sm), _p(tol), None, None,
                        _ref(lpu), C.byref(ierr), None, None)
    return int(ierr.value)


def sprfs2(mspar, mtrees, msifa, sm, b, lpu: int = -1) -> int:
    _check_int(mspar, mtrees, msifa)
    _check_real(sm, b)
    b, ldb, nrhs = _b_view(b)
    ierr = C.c_int64(0)
    lib().sprgpu_sprfs2(_p(mspar), _p(mtrees), _p(msifa), _p(sm), _p(b), _ref(ldb), _ref(nrhs),
                        None, None, _ref(lpu), C.byref(ierr))
    return int(ierr.value)


def sprbs2(mspar, mtrees, msifa, sm, b, lpu: int = -1) -> int:
    _check_int(mspar, mtrees, msifa)
    _check_real(sm, b)
    b, ldb, nrhs = _b_view(b)
    ierr = C.c_int64(0)
    lib().sprgpu_sprbs2(_p(mspar), _p(mtrees), _p(msifa), _p(sm), _p(b), _ref(ldb), _ref(nrhs),
                        None, _ref(lpu), C.byref(ierr))
    return int(ierr.value)


def sprsol(iop, mspar, mtrees, msifa, sm, b, tol, lpu: int = -1) -> int:
    _check_int(mspar, mtrees, msifa)
    _check_real(sm, tol)
    if b is None:
        b = np.zeros((int(mspar[7]), 1), order="F")
    b, ldb, nrhs = _b_view(b)
    ierr = C.c_int64(0)
    lib().sprgpu_sprsol(_ref(iop), _p(mspar), _p(mtrees), _p(msifa), _p(sm), _p(b), _ref(ldb),
                        _ref(nrhs), _p(tol), None, None, _ref(lpu), C.byref(ierr))
    return int(ierr.value)


def release(sm) -> None:
    """Frees the device factor kept for SM by the drop-in entry points."""
    lib().sprgpu_release(_p(sm))
