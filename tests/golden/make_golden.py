"""Writes the golden fixtures in this folder from the reference SPR build.

Run from the repository root where /root/reference is present:

    make -C oracle ref && python tests/golden/make_golden.py

Each fixture holds, for one small model, the reference's own symbolic output
(MSPAR, MTREES, MSIFA, MEQN, RINFO), the assembled SM from SPRADM, the
right-hand sides, and the reference's numeric results (factored SM, solution,
TOL, MSPAR after the call, IERR, and DCOR for SPRFC3 cases).  All arrays are
data produced by running the reference; no reference source is stored.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "sam-lib_amd"), os.path.join(ROOT, "oracle")]

import refspr  # noqa: E402
from sprgpu import fem  # noqa: E402


def shifted(model, s):
    """Same mesh, element matrices K_e - s*I (assembled: K - s*diag(multiplicity))."""
    nel = model.nel
    out = []
    for e in range(nel):
        k = model.element_matrix(e) - s * np.eye(int(model.nedof[e]))
        out.append(k.ravel(order="F"))
    model.em = np.concatenate(out)
    model.em_off = (np.arange(nel) * int(model.nedof[0]) ** 2).astype(np.int64)
    model.escale = None
    return model


def unclamped(model):
    model.msc[:] = 1
    return model


def run_case(name, model, nrhs=2, tol=(0.0, 0.0, 0.0), mode=1, zero_diag=None, cache_kb=0):
    sysm = refspr.symbolic(model, cache_kb=cache_kb)
    sm = refspr.assemble(model, sysm)
    if zero_diag is not None:
        # zero the assembled diagonal of equation zero_diag (1-based)
        xs, xl, xlnz = sysm.xsuper(), sysm.xlindx(), sysm.xlnz()
        js = int(np.searchsorted(xs, zero_diag, side="right")) - 1
        c = zero_diag - xs[js]
        F = xl[js + 1] - xl[js]
        sm[sysm.neq + xlnz[js] - 1 + c * F - c * (c - 1) // 2] = 0.0
    rng = np.random.default_rng(0)
    b = refspr.loads_to_equations(sysm, fem.uniform_loads(model.ndof, rng, nrhs))
    out = dict(mspar_in=sysm.mspar.copy(), mtrees=sysm.mtrees, msifa=sysm.msifa,
               meqn=sysm.meqn, rinfo=sysm.rinfo, sm_in=sm.copy(), b=b.copy(),
               tol_in=np.array(tol, dtype=float), mode=np.int64(mode))
    ms = sysm.mspar.copy()
    smf = sm.copy()
    x = b.copy(order="F")
    if mode == 3:
        ierr, tolv, dcor = refspr_sprfc3(sysm, ms, smf, tol)
        out["dcor"] = dcor
        if ierr >= 0 or ierr == -3:
            ierr2, _, _ = refspr.sol(4, sysm, smf, x, tol=tolv[:3], mspar=ms)
            out["ierr_solve"] = np.int64(ierr2)
    else:
        ierr, _, tolv = refspr.sol(3, sysm, smf, x, tol=tol, mspar=ms)
    out.update(sm_out=smf, x=x, tol_out=tolv, mspar_out=ms, ierr=np.int64(ierr))
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
    print(f"{name}: neq={sysm.neq} nsuper={sysm.nsuper} ierr={ierr} numneg={ms[25]}")


def refspr_sprfc3(sysm, ms, sm, tol):
    """Reference SPRFC3 (sprfc3.f:5) through ctypes."""
    import ctypes as C

    L = refspr.lib()
    neq = int(ms[7])
    tolv = np.array(tol, dtype=float)
    dcor = np.zeros(2 * neq)
    iwork = np.zeros(int(ms[13]) + 2, dtype=np.int64)
    rwork = np.zeros(int(ms[16]) + 1)
    ierr = C.c_int64(0)
    lpu = C.c_int64(-1)
    p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    L.sprfc3_(p(ms), p(sysm.mtrees), p(sysm.msifa), p(sm), p(tolv), p(iwork), p(rwork),
              p(dcor), C.byref(lpu), C.byref(ierr), L.lmvy18_, L.lmmy18_)
    return int(ierr.value), tolv, dcor


def main():
    run_case("hex8_3_lognormal", fem.hex8_grid(3, e_sigma=0.5, e_seed=1))
    run_case("hex8_4_panels", fem.hex8_grid(4), cache_kb=2)
    run_case("quad4_6", fem.quad4_grid(6))
    run_case("frame_25", fem.beam_frame(25, np.random.default_rng(3)))
    run_case("hex8_3_indefinite", shifted(fem.hex8_grid(3), 0.08))
    run_case("hex8_2_zero_diag", fem.hex8_grid(2), zero_diag=5)
    run_case("hex8_2_singular", unclamped(fem.hex8_grid(2)), tol=(1e-8, 0.0, 0.0))
    run_case("hex8_2_sprfc3", unclamped(fem.hex8_grid(2)), tol=(1e-8, 0.0, 0.0, 1.0), mode=3)


if __name__ == "__main__":
    main()
