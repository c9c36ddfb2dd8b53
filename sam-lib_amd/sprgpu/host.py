"""Host side of the drop-in: the reference symbolic phase and assembly.

The GPU library replaces SPR's numeric factor and solves; the host
application keeps running the reference's own symbolic phase (SPRPRP ->
SPRSAS -> SPRRNM -> SPRTRS -> SPRSMB -> SPRPMP) and element assembly
(SPRADM), and hands their arrays (MSPAR, MTREES, MSIFA, SM) to the drop-in
entry points.  ``lib/libspr_host.so`` is those reference routines compiled
unchanged from src/SPR (see the package Makefile) plus a C loop over
SPRADM; this module is its ctypes binding, used to set up problems the way a
host application would.

Every array is INTEGER*8 or DOUBLE PRECISION and 1-based in content; every
scalar is passed by reference, as Fortran 77 does.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(HERE), "lib", "libspr_host.so")
I8 = np.int64
NSPAR = 60
LPU = 6

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} missing: run `make -C sam-lib_amd`")
        _lib = C.CDLL(LIB_PATH)
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _i(v: int):
    return C.byref(C.c_int64(int(v)))


class SprError(RuntimeError):
    pass


@dataclass
class SprSystem:
    """Output of the reference symbolic phase and assembly for one model."""

    mpar: np.ndarray
    mspar: np.ndarray
    meqn: np.ndarray
    msica: np.ndarray
    mtrees: np.ndarray
    msifa: np.ndarray
    mdofnc: np.ndarray
    rinfo: np.ndarray
    sm: np.ndarray | None = None

    @property
    def neq(self) -> int:
        return int(self.mspar[7])

    @property
    def nsuper(self) -> int:
        return int(self.mspar[10])

    @property
    def nzerol(self) -> int:
        return int(self.mspar[15])

    @property
    def mfront(self) -> int:
        return int(self.mspar[13])

    def xsuper(self) -> np.ndarray:
        p = int(self.mspar[44]) - 1
        return self.mtrees[p:p + self.nsuper + 1]

    def xsibl(self) -> np.ndarray:
        p = int(self.mspar[44]) - 1 + self.nsuper + 1
        return self.mtrees[p:p + self.nsuper + 1]

    def sibl(self) -> np.ndarray:
        p = int(self.mspar[44]) - 1 + 2 * (self.nsuper + 1)
        return self.mtrees[p:p + self.nsuper + 1]

    def xlindx(self) -> np.ndarray:
        p = int(self.mspar[45]) - 1
        return self.mtrees[p:p + self.nsuper + 1]

    def lindx(self) -> np.ndarray:
        p = int(self.mspar[46]) - 1
        return self.msifa[p:p + int(self.mspar[14])]

    def xlnz(self) -> np.ndarray:
        p = int(self.mspar[47]) - 1
        return self.msifa[p:p + self.nsuper + 1]


def symbolic(model, nspar: int = NSPAR, cache_kb: int = 0, nperm: int = 0) -> SprSystem:
    """Run the reference symbolic phase (sprprp.f:5 ... sprpmp.f:5) on a SamModel.

    nperm = 2 reserves the PERI2E/PERE2I permutation arrays at the start of
    MSIFA (MSPAR(58) = 2 before SPRSMB, sprsmb.f:181, sprpmp.f:47); fill them
    with set_permutation (SPRABC) afterwards."""
    L = lib()
    mpar = model.mpar()
    ndof, nanod, nel = model.ndof, model.nanod, model.nel
    madof = np.ascontiguousarray(model.madof, dtype=I8)
    msc = np.ascontiguousarray(model.msc, dtype=I8).copy()
    mpmnpc = np.ascontiguousarray(model.mpmnpc, dtype=I8)
    mmnpc = np.ascontiguousarray(model.mmnpc, dtype=I8)
    minex = np.arange(1, nanod + 1, dtype=I8)
    mpmceq, mmceq, _ = _ceq(model)
    mspar = np.zeros(nspar, dtype=I8)
    meqn = np.zeros(ndof, dtype=I8)
    ierr = C.c_int64(0)

    mwork = np.zeros(2 * ndof + nanod, dtype=I8)
    L.sprprp_(_p(madof), _p(minex), _p(mpmnpc), _p(mmnpc), _p(mpmceq), _p(mmceq), _p(msc),
              _i(nspar), _i(LPU), _p(mpar), _p(mspar), _p(meqn), _p(mwork), C.byref(ierr))
    if ierr.value < 0:
        raise SprError(f"SPRPRP IERR={ierr.value}")
    if cache_kb:
        mspar[8] = cache_kb  # MSPAR(9): cache size for panelling (sprsmb.f:137)
    if nperm:
        mspar[57] = nperm    # MSPAR(58): permutation arrays first in MSIFA
    nmsica = int(mspar[34])
    ntrees = int(mspar[35])
    msica = np.zeros(nmsica, dtype=I8)
    iwork = np.zeros(2 * ndof + nanod, dtype=I8)
    L.sprsas_(_p(mpar), _p(mpmnpc), _p(mmnpc), _p(madof), _p(msc), _p(mpmceq), _p(mmceq),
              _p(meqn), _p(mspar), _p(msica), _p(iwork), _i(nspar), _i(LPU), C.byref(ierr))
    if ierr.value < 0:
        raise SprError(f"SPRSAS IERR={ierr.value}")
    iwork = np.zeros(int(mspar[36]), dtype=I8)
    L.sprrnm_(_p(mspar), _p(msica), _p(iwork), _i(LPU), C.byref(ierr))
    if ierr.value < 0:
        raise SprError(f"SPRRNM IERR={ierr.value}")
    mtrees = np.zeros(ntrees, dtype=I8)
    niwork = int(mspar[37])
    iwork = np.zeros(niwork, dtype=I8)
    rinfo = np.zeros(4)
    L.sprtrs_(_p(mspar), _p(msica), _p(mtrees), _p(meqn), _i(nspar), _i(nmsica), _i(ntrees),
              _i(ndof), _i(niwork), _p(iwork), _p(rinfo), _i(LPU), C.byref(ierr))
    if ierr.value < 0:
        raise SprError(f"SPRTRS IERR={ierr.value}")
    if nperm:
        mspar[2] += nperm * int(mspar[7])  # MSIFA also holds the permutations
    msifa = np.zeros(int(mspar[2]), dtype=I8)
    iwork = np.zeros(3 * int(mspar[5]) + 2 * int(mspar[7]) + 1, dtype=I8)
    L.sprsmb_(_p(mspar), _p(msica), _p(mtrees), _p(msifa), _p(meqn), _p(iwork), _i(LPU),
              C.byref(ierr))
    if ierr.value < 0:
        raise SprError(f"SPRSMB IERR={ierr.value}")
    neq = int(mspar[7])
    mdofnc = np.zeros(2 * neq + 1, dtype=I8)
    L.sprpmp_(_p(msica), _p(mtrees), _p(meqn), _p(mpar), _p(mspar), _p(mdofnc))
    return SprSystem(mpar, mspar, meqn, msica, mtrees, msifa, mdofnc, rinfo)


def set_permutation(sysm: SprSystem, meqn_ext: np.ndarray) -> None:
    """SPRABC (sprabc.f:5-66): PERI2E and PERE2I between the SPR equation
    order (sysm.meqn) and the caller's external order meqn_ext (1-based
    external equation of each dof, 0 where the dof has none), placed first
    in MSIFA; sets MSPAR(58) = 2.  The system must come from
    symbolic(..., nperm=2)."""
    if int(sysm.mspar[46]) < 1 + 2 * sysm.neq:
        raise SprError("MSIFA has no room for the permutations: use symbolic(..., nperm=2)")
    meqn_ext = np.ascontiguousarray(meqn_ext, dtype=I8)
    ierr = C.c_int64(0)
    lib().sprabc_(_p(sysm.mspar), _p(sysm.msifa), _p(sysm.meqn), _p(meqn_ext),
                  _i(sysm.meqn.size), _i(LPU), C.byref(ierr))
    if ierr.value != 0:
        raise SprError(f"SPRABC IERR={ierr.value}")


def _ceq(model):
    """(MPMCEQ, MMCEQ, TTCC) of a model, placeholders when it has none."""
    f = getattr(model, "ceq_arrays", None)
    if f is None:
        return np.ones(1, dtype=I8), np.zeros(1, dtype=I8), np.zeros(1)
    return f()


def assemble(model, sysm: SprSystem) -> np.ndarray:
    """Assemble all element matrices with the reference SPRADM (spradm.f:5).

    Returns SM(NEQ+NZEROL): SM(1:NEQ) is untouched scratch (zero) and
    SM(NEQ+1:) holds the lower triangle of A in the supernodal LNZ layout.
    Constraint equations of the model enter through SPRAD2's constrained part
    (spradm.f:861-1002).
    """
    return assemble_rhs(model, sysm, 0)[0]


def assemble_rhs(model, sysm: SprSystem, nsv: int, sm: np.ndarray | None = None,
                 sv: np.ndarray | None = None):
    """SPRADM over all elements with NSV = nsv (spradm.f:5, SPRAD2 at
    spradm.f:743-1002): nsv >= 0 adds to SM, nsv != 0 subtracts the
    prescribed (and, for |nsv| = 1, dependent) dofs' columns, weighted by
    their c0, from SV(NEQ, |nsv|).  Returns (sm, sv); zero arrays unless given
    (then updated in place)."""
    L = lib()
    neq, nzerol = sysm.neq, sysm.nzerol
    if sm is None:
        sm = np.zeros(neq + nzerol)
    if sv is None:
        sv = np.zeros(max(neq * max(abs(nsv), 1), 1))
    iwork = np.zeros(int(sysm.mspar[38]) + 16, dtype=I8)
    mpmceq, mmceq, ttcc = _ceq(model)
    nedof = np.ascontiguousarray(model.nedof, dtype=I8)
    em = np.ascontiguousarray(model.em, dtype=np.float64)
    em_off = np.ascontiguousarray(model.em_off, dtype=I8)
    scale = None if model.escale is None else np.ascontiguousarray(model.escale, dtype=np.float64)
    madof = np.ascontiguousarray(model.madof, dtype=I8)
    mpmnpc = np.ascontiguousarray(model.mpmnpc, dtype=I8)
    mmnpc = np.ascontiguousarray(model.mmnpc, dtype=I8)
    bad = C.c_int64(0)
    L.spr_assemble_all.restype = C.c_int64
    ierr = L.spr_assemble_all(
        C.c_int64(model.nel), _p(nedof), _p(em), _p(em_off),
        _p(scale) if scale is not None else None, _p(ttcc), _p(sysm.mpar), _p(sysm.mspar),
        _p(madof), _p(sysm.meqn), _p(mpmnpc), _p(mmnpc), _p(mpmceq), _p(mmceq),
        _p(sysm.msica), _p(sysm.mtrees), _p(sysm.msifa), _p(sysm.mdofnc), _p(sm), _p(sv),
        C.c_int64(nsv), _p(iwork), C.c_int64(LPU), C.byref(bad))
    if ierr != 0:
        raise SprError(f"SPRADM IERR={ierr} at element {bad.value}")
    return sm, sv


def loads_to_equations(sysm: SprSystem, f_dof: np.ndarray) -> np.ndarray:
    """Map natural-dof loads (ndof, nrhs) to the SPR equation order (neq, nrhs)."""
    f_dof = np.asarray(f_dof, dtype=np.float64)
    if f_dof.ndim == 1:
        f_dof = f_dof[:, None]
    b = np.zeros((sysm.neq, f_dof.shape[1]), order="F")
    free = sysm.meqn > 0
    b[sysm.meqn[free] - 1, :] = f_dof[free, :]
    return b


def equations_to_dofs(sysm: SprSystem, x_eq: np.ndarray) -> np.ndarray:
    """Map an SPR-ordered solution (neq[, nrhs]) back to natural dofs (zeros where fixed)."""
    x_eq = np.asarray(x_eq)
    shape = (sysm.meqn.size,) + x_eq.shape[1:]
    u = np.zeros(shape)
    free = sysm.meqn > 0
    u[free] = x_eq[sysm.meqn[free] - 1]
    return u


def factor_flops(sysm: SprSystem) -> float:
    """RINFO(3): the reference's factor flop count (sprtrs.f:1961-1969)."""
    return float(sysm.rinfo[2])


def lnz_positions(sysm: SprSystem, rows: np.ndarray, cols: np.ndarray) -> np.ndarray:
    """0-based SM indices of the lower-triangle entries (rows[k], cols[k]).

    rows/cols are 0-based SPR equations with rows >= cols.  The entry of
    column c lies in the supernode holding c, in the packed trapezoid that
    starts at XLNZ: column k of a supernode of order F starts at
    OFFSET(F,k) = k*F - k*(k-1)/2 and holds the rows LINDX(k..F-1)
    (sprsmb.f:512-575, sprfc1.f:201-203); SM(1:NEQ) precedes LNZ.
    """
    rows = np.asarray(rows, dtype=np.int64)
    cols = np.asarray(cols, dtype=np.int64)
    if np.any(rows < cols):
        raise ValueError("lnz_positions takes lower-triangle entries (row >= col)")
    xsuper = sysm.xsuper() - 1
    xlindx = sysm.xlindx() - 1
    xlnz = sysm.xlnz() - 1
    lindx = sysm.lindx() - 1
    s = np.searchsorted(xsuper, cols, side="right") - 1
    k = cols - xsuper[s]
    F = xlindx[s + 1] - xlindx[s]
    p = np.empty_like(rows)
    for j in np.unique(s):  # rows are searched in the supernode's sorted list
        m = s == j
        seg = lindx[xlindx[j]:xlindx[j + 1]]
        p[m] = np.searchsorted(seg, rows[m])
        if np.any(seg[np.minimum(p[m], seg.size - 1)] != rows[m]):
            raise ValueError("entry outside the symbolic pattern")
    return sysm.neq + xlnz[s] + k * F - k * (k - 1) // 2 + (p - k)


def element_equations(model, sysm: SprSystem) -> np.ndarray:
    """(nel, nedof) 0-based equations of each element's dofs, -1 where a dof
    has none (fixed).  All elements must have the same dof count."""
    nedof = int(model.nedof[0])
    if np.any(model.nedof != nedof):
        raise ValueError("element_equations needs a uniform element dof count")
    eq = np.empty((model.nel, nedof), dtype=np.int64)
    for e in range(model.nel):
        eq[e] = sysm.meqn[model.element_dofs(e)] - 1
    eq[eq < 0] = -1
    return eq


def element_meen(model, sysm: SprSystem):
    """(meen_ptr, meen): element e's dofs have the 1-based SPR equations
    meen[meen_ptr[e]:meen_ptr[e+1]] in element-matrix order, 0 where a dof has
    none.  This is the MEEN array that SPREEQ builds for each SPRADM call
    (spradm.f:169-221, MEEN(I) = MEQN(IDOF))."""
    nedof = np.asarray(model.nedof, dtype=I8)
    meen_ptr = np.zeros(model.nel + 1, dtype=I8)
    np.cumsum(nedof, out=meen_ptr[1:])
    if np.all(nedof == nedof[0]) and model.nel:
        nn = int(model.mpmnpc[1] - model.mpmnpc[0])
        if np.all(np.diff(model.mpmnpc) == nn):
            # uniform elements: every node has nedof/nn dofs, in MADOF order
            nodes = (model.mmnpc - 1).reshape(model.nel, nn)
            per = int(nedof[0]) // nn
            first = model.madof[nodes] - 1
            if np.all(model.madof[nodes + 1] - 1 - first == per):
                dofs = (first[:, :, None] + np.arange(per)).reshape(model.nel, -1)
                return meen_ptr, np.ascontiguousarray(sysm.meqn[dofs].ravel(), dtype=I8)
    meen = np.concatenate([sysm.meqn[model.element_dofs(e)] for e in range(model.nel)])
    return meen_ptr, np.ascontiguousarray(meen, dtype=I8)


@dataclass
class MergedSystem:
    """Several independent SPR systems as one forest (block-diagonal A).

    The arrays follow the SPR layout the drop-in consumes (sprfc1.f:161-203):
    supernodes, equations, LINDX entries and LNZ positions of system k are
    shifted by the totals of systems 0..k-1, and each system's roots stay
    roots.  The GPU plan batches all supernodes of one tree height, so every
    launch covers all systems.
    """

    mspar: np.ndarray
    mtrees: np.ndarray
    msifa: np.ndarray
    neq_off: np.ndarray     # (k+1,) equation offsets
    nzl_off: np.ndarray     # (k+1,) LNZ offsets
    neq: int
    nzerol: int

    def merge_sm(self, sms) -> np.ndarray:
        """SM of the forest: all diagonals, then all LNZ parts, in system order."""
        k = len(sms)
        out = np.empty(self.neq + self.nzerol)
        for i in range(k):
            n = self.neq_off[i + 1] - self.neq_off[i]
            out[self.neq_off[i]:self.neq_off[i + 1]] = sms[i][:n]
            out[self.neq + self.nzl_off[i]:self.neq + self.nzl_off[i + 1]] = sms[i][n:]
        return out

    def split(self, v: np.ndarray):
        """Per-system slices of a vector in forest equation order."""
        return [v[self.neq_off[i]:self.neq_off[i + 1]] for i in range(len(self.neq_off) - 1)]


def merge_systems(systems) -> MergedSystem:
    """Merges SprSystems into one forest-structured system (see MergedSystem)."""
    ns_tot = sum(s.nsuper for s in systems)
    neq_off = np.zeros(len(systems) + 1, dtype=I8)
    nzl_off = np.zeros(len(systems) + 1, dtype=I8)
    nof_tot = 0
    maxsup = mfront = 0
    xsuper, xsibl, sibl, xlindx, supsup, lindx, xlnz = [], [], [], [], [], [], []
    ns0 = nsib0 = nof0 = 0
    for i, s in enumerate(systems):
        neq_off[i + 1] = neq_off[i] + s.neq
        nzl_off[i + 1] = nzl_off[i] + s.nzerol
        ns = s.nsuper
        nofsub = int(s.mspar[14])
        maxsup = max(maxsup, int(s.mspar[12]))
        mfront = max(mfront, int(s.mspar[13]))
        xs = s.xsuper()
        xsuper.append(xs[:-1] + neq_off[i])
        xsb = s.xsibl()
        nch = int(xsb[ns] - xsb[0])
        xsibl.append(xsb[:-1] - xsb[0] + 1 + nsib0)
        sibl.append(s.sibl()[xsb[0] - 1:xsb[0] - 1 + nch] + ns0)
        xl = s.xlindx()
        xlindx.append(xl[:-1] - xl[0] + 1 + nof0)
        p = int(s.mspar[45]) - 1 + ns + 1
        supsup.append(s.mtrees[p:p + ns])
        lindx.append(s.lindx() + neq_off[i])
        xlnz.append(s.xlnz()[:-1] + nzl_off[i])
        ns0 += ns
        nsib0 += nch
        nof0 += nofsub
        nof_tot += nofsub
    neq, nzerol = int(neq_off[-1]), int(nzl_off[-1])
    xsuper = np.concatenate(xsuper + [np.array([neq + 1], dtype=I8)])
    xsibl = np.concatenate(xsibl + [np.array([nsib0 + 1], dtype=I8)])
    sibl = np.concatenate(sibl + [np.zeros(ns_tot + 1 - nsib0, dtype=I8)])
    xlindx = np.concatenate(xlindx + [np.array([nof_tot + 1], dtype=I8)])
    supsup = np.concatenate(supsup)
    lindx = np.concatenate(lindx)
    xlnz = np.concatenate(xlnz + [np.array([nzerol + 1], dtype=I8)])
    mtrees = np.concatenate([xsuper, xsibl, sibl, xlindx, supsup])
    msifa = np.concatenate([lindx, xlnz])
    mspar = np.zeros(NSPAR, dtype=I8)
    mspar[0] = 4                      # symbolic phase complete
    mspar[7] = neq
    mspar[10] = ns_tot
    mspar[12] = maxsup
    mspar[13] = mfront
    mspar[14] = nof_tot
    mspar[15] = nzerol
    mspar[44] = 1                              # XSUPER
    mspar[45] = 3 * (ns_tot + 1) + 1           # XLINDX
    mspar[46] = 1                              # LINDX
    mspar[47] = nof_tot + 1                    # XLNZ
    return MergedSystem(mspar, mtrees, msifa, neq_off, nzl_off, neq, nzerol)
