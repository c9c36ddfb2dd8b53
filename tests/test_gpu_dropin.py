"""GPU parity of the drop-in entry points against the oracle (C restatement).

Every case runs the reference symbolic phase and SPRADM (the caller's set-up),
then factors and solves on the GPU through the C ABI and compares with the
oracle on the same inputs.  Floating-point results agree to within rounding:
the GPU eliminates in a different order (level-batched, blocked, MFMA), so
factor entries and solutions are compared with relative tolerances; integer
outputs (IERR, NUMNEG, SUPEQT, MSPAR) must match exactly.
"""
import numpy as np
import pytest

import oracle
from conftest import make_system

pytestmark = pytest.mark.gpu

# relative tolerance of FP64 solutions against the oracle (factor computed in
# a different elimination order; the bound is kappa * eps with margin)
SOL_RTOL = 1e-9
TOL_RTOL = 1e-9


def _oracle_solve(sysm, sm, b, tol1=0.0):
    sm_o = sm.copy()
    b_o = np.asfortranarray(b.copy())
    ms = sysm.mspar.copy()
    ierr, tol = oracle.sprsol(3, ms, sysm.mtrees, sysm.msifa, sm_o, b_o, tol=[tol1, 0, 0])
    return ierr, tol, sm_o, b_o, ms


CASES = [("hex8", 2, {}), ("hex8", 4, {"e_sigma": 0.5, "e_seed": 1}),
         ("hex8", 8, {"e_sigma": 0.5, "e_seed": 1}), ("hex8", 12, {}),
         ("quad4", 30, {}), ("frame", 60, {"seed": 3})]


@pytest.mark.parametrize("kind,n,kw", CASES, ids=[f"{c[0]}{c[1]}" for c in CASES])
def test_sprsol_matches_oracle(kind, n, kw):
    from sprgpu import spr

    model, sysm, sm = make_system(kind, n, **kw)
    print(f"case {kind} n={n} neq={sysm.neq} nsuper={sysm.nsuper}", flush=True)
    rng = np.random.default_rng(0)
    b = rng.uniform(-1, 1, (sysm.neq, 3))
    ierr_o, tol_o, sm_o, x_o, ms_o = _oracle_solve(sysm, sm, b)

    sm_g = sm.copy()
    x_g = np.asfortranarray(b.copy())
    ms_g = sysm.mspar.copy()
    tol_g = np.zeros(3)
    ierr_g = spr.sprsol(3, ms_g, sysm.mtrees, sysm.msifa, sm_g, x_g, tol_g)
    spr.release(sm_g)
    assert ierr_g == ierr_o
    assert np.array_equal(ms_g, ms_o)
    np.testing.assert_allclose(tol_g[1:], tol_o[1:], rtol=TOL_RTOL)
    rel = np.linalg.norm(x_g - x_o) / np.linalg.norm(x_o)
    assert rel < SOL_RTOL, rel
    # the factor written back into SM agrees with the oracle's
    d_rel = np.max(np.abs(sm_g[:sysm.neq] - sm_o[:sysm.neq]) / np.abs(sm_o[:sysm.neq]))
    assert d_rel < 1e-9, d_rel
    l_err = np.linalg.norm(sm_g[sysm.neq:] - sm_o[sysm.neq:]) / np.linalg.norm(sm_o[sysm.neq:])
    assert l_err < 1e-9, l_err


def test_factor_then_separate_solves():
    """SPRFC1 then SPRFS1 and SPRBS1 as separate calls (device factor reused)."""
    from sprgpu import spr

    model, sysm, sm = make_system("hex8", 6, e_sigma=0.5, e_seed=1)
    b = np.asfortranarray(np.random.default_rng(1).normal(size=(sysm.neq, 2)))
    _, _, _, x_o, _ = _oracle_solve(sysm, sm, b)
    ms = sysm.mspar.copy()
    tol = np.zeros(3)
    assert spr.sprfc1(ms, sysm.mtrees, sysm.msifa, sm, tol) == 0
    assert ms[0] == 5
    x = b.copy(order="F")
    assert spr.sprfs1(ms, sysm.mtrees, sysm.msifa, sm, x) == 0
    assert ms[0] == 7
    assert spr.sprbs1(ms, sysm.mtrees, sysm.msifa, sm, x) == 0
    spr.release(sm)
    assert np.linalg.norm(x - x_o) / np.linalg.norm(x_o) < SOL_RTOL
