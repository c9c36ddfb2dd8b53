"""ctypes access to the reference SPR library built by ``oracle/Makefile``.

TEST INFRASTRUCTURE ONLY.  ``oracle/_ref/libspr_ref.so`` is the reference
Fortran (src/SPR/*.f) compiled unchanged with flang -fdefault-integer-8.  It
is the checker and the CPU baseline: SPRSOL / SPRFC1 / SPRFS1 / SPRBS1 run on
the CPU, timed with CLOCK_MONOTONIC.  Problem set-up (the reference symbolic
phase and SPRADM assembly) is the product's host side, ``sprgpu.host``; it
is re-exported here so that the tests read as one module.

Every array is INTEGER*8 or DOUBLE PRECISION and 1-based in content; every
scalar is passed by reference, as Fortran 77 does.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "sam-lib_amd"))
from sprgpu.host import (SprError, SprSystem, assemble, equations_to_dofs,  # noqa: E402,F401
                         factor_flops, loads_to_equations, symbolic)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_ref", "libspr_ref.so")
I8 = np.int64
NSPAR = 60
LPU = 6

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} missing: run `make -C oracle`")
        _lib = C.CDLL(LIB_PATH)
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _i(v: int):
    return C.byref(C.c_int64(int(v)))


def sol(iop: int, sysm: SprSystem, sm: np.ndarray, b: np.ndarray | None, tol=None,
        mspar: np.ndarray | None = None):
    """Reference SPRSOL (sprsol.f:5) with timing.  Returns (ierr, seconds, tol).

    ``sm`` and ``b`` (Fortran-ordered (neq, nrhs)) are modified in place, as in
    the reference.
    """
    L = lib()
    mspar = sysm.mspar if mspar is None else mspar
    neq = int(mspar[7])
    mfront = int(mspar[13])
    niwork = max(mfront + 1, int(mspar[12])) + 1
    nrwork = max(int(mspar[16]), mfront) + 1
    iwork = np.zeros(niwork, dtype=I8)
    rwork = np.zeros(nrwork)
    tolv = np.zeros(3) if tol is None else np.asarray(tol, dtype=np.float64).copy()
    if b is None:
        b = np.zeros((neq, 1), order="F")
    assert b.flags.f_contiguous and b.dtype == np.float64
    nrhs = b.shape[1] if b.ndim == 2 else 1
    ldb = b.shape[0]  # < NEQ: the partial solve of MSPAR(57) = 1
    secs = C.c_double(0.0)
    L.spr_sol_timed.restype = C.c_int64
    ierr = L.spr_sol_timed(C.c_int64(iop), _p(mspar), _p(sysm.mtrees), _p(sysm.msifa), _p(sm),
                           _p(b), C.c_int64(ldb), C.c_int64(nrhs), _p(tolv), _p(iwork),
                           _p(rwork), C.c_int64(LPU), C.byref(secs))
    return int(ierr), secs.value, tolv




def prm(a, b, mspar, mtrees, msifa, iflag=1, ksa=1, c=None):
    """Reference SPRPRM (sprprm.f:5-7): C = +-A B or C = C +- A B.  Returns
    (ierr, c, wa); wa is the reference's work array on exit (the last
    column's +-A b, or the result for IFLAG = 88)."""
    b = np.asfortranarray(b, dtype=np.float64)
    if b.ndim == 1:
        b = b[:, None]
    m, n = b.shape
    c = np.zeros((m, n), order="F") if c is None else c
    wa = np.zeros(m)
    ierr = C.c_int64(0)
    lib().sprprm_(_p(np.ascontiguousarray(a, dtype=np.float64)), _p(b), _p(c), _p(wa),
                  _p(mspar), _p(mtrees), _p(msifa), _i(m), _i(n), _i(ksa), _i(iflag), _i(-1),
                  C.byref(ierr))
    return int(ierr.value), c, wa


def tmm(u, b, mspar, mtrees, msifa, iflag=1, ksu=1):
    """Reference SPRTMM (sprtmm.f:5-7).  Returns (ierr, c, wa)."""
    b = np.asfortranarray(b, dtype=np.float64)
    if b.ndim == 1:
        b = b[:, None]
    m, n = b.shape
    c = np.zeros((m, n), order="F")
    wa = np.zeros(m)
    ierr = C.c_int64(0)
    lib().sprtmm_(_p(np.ascontiguousarray(u, dtype=np.float64)), _p(b), _p(c), _p(wa),
                  _p(mspar), _p(mtrees), _p(msifa), _i(m), _i(n), _i(ksu), _i(iflag), _i(-1),
                  C.byref(ierr))
    return int(ierr.value), c, wa


def fc2(sysm: SprSystem, sm: np.ndarray, tol=(0.0, 0.0, 0.0), mspar: np.ndarray | None = None):
    """Reference SPRFC2 (sprfc2.f:5-7), the CC' factorization, with CMVY18 /
    CMMY18.  ``sm`` and ``mspar`` are modified in place.  Returns (ierr, tol)."""
    mspar = sysm.mspar if mspar is None else mspar
    mfront = int(mspar[13])
    iwork = np.zeros(max(mfront + 1, int(mspar[12])) + 1, dtype=I8)
    rwork = np.zeros(max(int(mspar[16]), mfront) + 1)
    tolv = np.asarray(tol, dtype=np.float64).copy()
    f = lib().spr_fc2
    f.restype = C.c_int64
    ierr = f(_p(mspar), _p(sysm.mtrees), _p(sysm.msifa), _p(sm), _p(tolv), _p(iwork), _p(rwork),
             C.c_int64(-1))
    return int(ierr), tolv


def fc4(sysm: SprSystem, sm: np.ndarray, tol=(0.0, 0.0, 0.0, 0.0),
        mspar: np.ndarray | None = None):
    """Reference SPRFC4 (sprfc4.f:5-7), the CC' factorization with
    replacement, with CMVY18 / CMMY18.  ``sm`` and ``mspar`` are modified in
    place.  Returns (ierr, tol, dcor (neq, 2))."""
    mspar = sysm.mspar if mspar is None else mspar
    mfront = int(mspar[13])
    iwork = np.zeros(mfront + 2, dtype=I8)
    rwork = np.zeros(max(int(mspar[16]), mfront) + 1)
    tolv = np.zeros(4)
    tolv[:len(tol)] = tol
    dcor = np.zeros(2 * int(mspar[7]))
    f = lib().spr_fc4
    f.restype = C.c_int64
    ierr = f(_p(mspar), _p(sysm.mtrees), _p(sysm.msifa), _p(sm), _p(tolv), _p(iwork), _p(rwork),
             _p(dcor), C.c_int64(-1))
    return int(ierr), tolv, dcor.reshape(-1, 2)


def fs2_bs2(which: str, sysm: SprSystem, sm: np.ndarray, b: np.ndarray,
            mspar: np.ndarray | None = None):
    """Reference SPRFS2 (C y = b, sprfs2.f:5) or SPRBS2 (C^T x = y, sprbs2.f:5)
    in place on Fortran-ordered b (neq, nrhs).  Returns ierr."""
    mspar = sysm.mspar if mspar is None else mspar
    neq, mfront = int(mspar[7]), int(mspar[13])
    assert b.flags.f_contiguous and b.dtype == np.float64
    nrhs = b.shape[1] if b.ndim == 2 else 1
    rwork = np.zeros(mfront + 1)
    ierr = C.c_int64(0)
    if which == "fs":
        iwork = np.zeros(max(int(mspar[12]), mfront) + 1, dtype=I8)
        lib().sprfs2_(_p(mspar), _p(sysm.mtrees), _p(sysm.msifa), _p(sm), _p(b), _i(neq),
                      _i(nrhs), _p(iwork), _p(rwork), _i(-1), C.byref(ierr))
    else:
        lib().sprbs2_(_p(mspar), _p(sysm.mtrees), _p(sysm.msifa), _p(sm), _p(b), _i(neq),
                      _i(nrhs), _p(rwork), _i(-1), C.byref(ierr))
    return int(ierr.value)
