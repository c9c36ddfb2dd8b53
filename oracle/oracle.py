"""ctypes wrapper for the C restatement ``oracle/liboracle.so``.

TEST INFRASTRUCTURE ONLY: the checker that the HIP library is compared with.
The functions take the same arrays as the reference routines they restate
(see spr_oracle.h for the file:line of each).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
I8 = np.int64

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} missing: run `make -C oracle`")
        _lib = C.CDLL(LIB_PATH)
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _i(v):
    return C.byref(C.c_int64(int(v)))


def sprsol(iop: int, mspar, mtrees, msifa, sm, b=None, tol=None):
    """Restated SPRSOL; sm and b are modified in place. Returns (ierr, tol)."""
    neq = int(mspar[7])
    tolv = np.zeros(3) if tol is None else np.asarray(tol, dtype=np.float64).copy()
    if b is None:
        b = np.zeros((neq, 1), order="F")
    assert b.flags.f_contiguous and b.dtype == np.float64
    nrhs = b.shape[1] if b.ndim == 2 else 1
    ldb = b.shape[0]  # < NEQ: the partial solve of MSPAR(57) = 1
    ierr = C.c_int64(0)
    lib().oracle_sprsol(_i(iop), _p(mspar), _p(mtrees), _p(msifa), _p(sm), _p(b),
                        _i(ldb), _i(nrhs), _p(tolv), None, None, _i(-1), C.byref(ierr))
    return int(ierr.value), tolv


def sprfc3(mspar, mtrees, msifa, sm, tol):
    """Restated SPRFC3 (singular-tolerant factor). Returns (ierr, tol, dcor(2,neq))."""
    neq = int(mspar[7])
    tolv = np.asarray(tol, dtype=np.float64).copy()
    assert tolv.size >= 4
    dcor = np.zeros(2 * neq)
    ierr = C.c_int64(0)
    lib().oracle_sprfc3(_p(mspar), _p(mtrees), _p(msifa), _p(sm), _p(tolv), None, None,
                        _p(dcor), _i(-1), C.byref(ierr))
    return int(ierr.value), tolv, dcor.reshape(neq, 2).T


def sprprm(a, b, mspar, mtrees, msifa, iflag=1, c=None, ksa=1):
    """Restated SPRPRM: C = +-A*B (|iflag|=1) or C +-= A*B (|iflag|=2); A is
    the SPR-layout lower triangle (ksa > 0) or a diagonal vector (ksa <= 0)."""
    b = np.asfortranarray(b, dtype=np.float64)
    if b.ndim == 1:
        b = b[:, None]
    m, n = b.shape
    c = np.zeros((m, n), order="F") if c is None else c
    wa = np.zeros(m)
    ierr = C.c_int64(0)
    lib().oracle_sprprm(_p(a), _p(b), _p(c), _p(wa), _p(mspar), _p(mtrees), _p(msifa),
                        _i(m), _i(n), _i(ksa), _i(iflag), _i(-1), C.byref(ierr))
    return int(ierr.value), c


def assemble_free(mspar, mtrees, msifa, meen_ptr, meen, em, em_off, escale=None, sm=None):
    """Restated SPRADM element loop (free dofs): SM(NEQ+1:) += EM_e per element.

    Returns (rc, sm); sm is a zero SM(NEQ+NZEROL) unless one is given (it is
    then updated in place)."""
    neq, nzerol = int(mspar[7]), int(mspar[15])
    sm = np.zeros(neq + nzerol) if sm is None else sm
    meen_ptr = np.ascontiguousarray(meen_ptr, dtype=I8)
    meen = np.ascontiguousarray(meen, dtype=I8)
    em = np.ascontiguousarray(em, dtype=np.float64)
    em_off = np.ascontiguousarray(em_off, dtype=I8)
    sc = None if escale is None else np.ascontiguousarray(escale, dtype=np.float64)
    f = lib().oracle_assemble_free
    f.restype = C.c_int64
    rc = f(_p(mspar), _p(mtrees), _p(msifa), C.c_int64(meen_ptr.size - 1), _p(meen_ptr),
           _p(meen), _p(em), _p(em_off), _p(sc) if sc is not None else None, _p(sm))
    return int(rc), sm


def assemble_ceq(mspar, mtrees, msifa, meen_ptr, meen, em, em_off, escale, mpmceq, mmceq,
                 ttcc, meqn, nsv, sm=None, sv=None):
    """Restated SPRADM element loop with constraint equations (SPRAD2 complete):
    SM(NEQ+NZEROL) for nsv >= 0 and SV(NEQ*|nsv|) for nsv != 0.  Returns
    (rc, sm, sv); zero arrays unless given (then updated in place)."""
    neq, nzerol = int(mspar[7]), int(mspar[15])
    sm = np.zeros(neq + nzerol) if sm is None else sm
    sv = np.zeros(max(neq * max(abs(nsv), 1), 1)) if sv is None else sv
    a = [np.ascontiguousarray(x, dtype=I8) for x in (meen_ptr, meen, em_off, mpmceq, mmceq, meqn)]
    em = np.ascontiguousarray(em, dtype=np.float64)
    tt = np.ascontiguousarray(ttcc, dtype=np.float64)
    sc = None if escale is None else np.ascontiguousarray(escale, dtype=np.float64)
    f = lib().oracle_assemble_ceq
    f.restype = C.c_int64
    rc = f(_p(mspar), _p(mtrees), _p(msifa), C.c_int64(a[0].size - 1), _p(a[0]), _p(a[1]),
           _p(em), _p(a[2]), _p(sc) if sc is not None else None, _p(a[3]), _p(a[4]), _p(tt),
           _p(a[5]), C.c_int64(nsv), _p(sm), _p(sv))
    return int(rc), sm, sv


def sprtmm(u, b, mspar, mtrees, msifa, iflag=1, ksu=1):
    """Restated SPRTMM: IFLAG 1 C = L^T B, 2 C = L B (L the stored lower
    triangle of U); 11/22 with one column: the result in WA.  Returns
    (ierr, c, wa)."""
    b = np.asfortranarray(b, dtype=np.float64)
    if b.ndim == 1:
        b = b[:, None]
    m, n = b.shape
    c = np.zeros((m, n), order="F")
    wa = np.zeros(m)
    ierr = C.c_int64(0)
    lib().oracle_sprtmm(_p(np.ascontiguousarray(u, dtype=np.float64)), _p(b), _p(c), _p(wa),
                        _p(mspar), _p(mtrees), _p(msifa), _i(m), _i(n), _i(ksu), _i(iflag),
                        _i(-1), C.byref(ierr))
    return int(ierr.value), c, wa
