"""The partial solve (MSPAR(57) = 1 with LDB < NEQ; sprfs1.f:200-201,
sprbs1.f:189) on the GPU: B holds the leading LDB equations, the others are
zero in the sweeps and their results are dropped (SPRFB1/SPRFB2).

Checked against the oracle, whose partial SPRFS1/SPRBS1 equal the reference
built from source bit for bit (tests/test_oracle.py), for the forward and
backward sweeps alone and together, with LDB inside a supernode, at a
supernode boundary and one short of NEQ; through the handle
(sprgpu_solve_partial), the drop-in SPRSOL / SPRFS1 / SPRBS1, and with PERI2E
when the leading external equations are the leading internal ones (else -2).
"""
import numpy as np
import pytest

import oracle
from conftest import make_system
from test_oracle import _partial_cases, dense_equation_matrix

pytestmark = pytest.mark.gpu

RTOL = 1e-10


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("kind,n", [("hex8", 6), ("quad4", 20)])
def test_partial_solve_handle_matches_oracle(kind, n):
    import torch
    from sprgpu.device import SprFactor

    model, sysm, sm = make_system(kind, n)
    fac = sm.copy()
    assert oracle.sprsol(1, sysm.mspar.copy(), sysm.mtrees, sysm.msifa, fac)[0] == 0
    A = dense_equation_matrix(model, sysm)
    h = SprFactor(sysm.mspar, sysm.mtrees, sysm.msifa, "fp64")
    try:
        assert h.factor(torch.from_numpy(sm).cuda())["ierr"] == 0
        rng = np.random.default_rng(5)
        mspar = sysm.mspar.copy()
        mspar[56] = 1
        for ldb in _partial_cases(sysm):
            b = np.asfortranarray(rng.uniform(-1, 1, (ldb, 3)))
            print(f"{kind} {n}: neq {sysm.neq} ldb {ldb}", flush=True)
            for which, iop in ((1, 5), (2, 6), (3, 4)):
                xo = b.copy(order="F")
                assert oracle.sprsol(iop, mspar.copy(), sysm.mtrees, sysm.msifa, fac.copy(),
                                     xo)[0] == 0
                xg = torch.from_numpy(np.ascontiguousarray(b.T)).cuda()
                h.solve_partial(xg, which)
                assert _rel(xg.cpu().numpy().T, xo) < RTOL, (ldb, which)
            x1 = np.linalg.solve(A[:ldb, :ldb], b)
            assert _rel(xg.cpu().numpy().T, x1) < 1e-9
        # LDB >= NEQ is the ordinary solve
        b = rng.uniform(-1, 1, (2, sysm.neq))
        x = torch.from_numpy(b.copy()).cuda()
        y = torch.from_numpy(b.copy()).cuda()
        h.solve_partial(x)
        h.solve(y)
        assert torch.equal(x, y)
    finally:
        h.close()


def test_partial_solve_dropin_matches_reference():
    import refspr
    from sprgpu import spr

    model, sysm, sm = make_system("hex8", 7)
    rng = np.random.default_rng(9)
    for ldb in _partial_cases(sysm)[:2]:
        mspar = sysm.mspar.copy()
        mspar[56] = 1
        b = np.asfortranarray(rng.uniform(-1, 1, (ldb, 2)))
        xr, smr = b.copy(order="F"), sm.copy()
        assert refspr.sol(3, sysm, smr, xr, mspar=mspar.copy())[0] == 0
        xg, smg, msg = b.copy(order="F"), sm.copy(), mspar.copy()
        assert spr.sprsol(3, msg, sysm.mtrees, sysm.msifa, smg, xg, np.zeros(3)) == 0
        assert _rel(xg, xr) < 1e-9, ldb
        # the two sweeps as separate calls
        x2 = b.copy(order="F")
        assert spr.sprfs1(msg, sysm.mtrees, sysm.msifa, smg, x2) == 0
        assert spr.sprbs1(msg, sysm.mtrees, sysm.msifa, smg, x2) == 0
        assert _rel(x2, xr) < 1e-9
        # MSPAR(57) = 0: LDB < NEQ is an error, as in the reference
        ms0 = msg.copy()
        ms0[56] = 0
        assert spr.sprfs1(ms0, sysm.mtrees, sysm.msifa, smg, b.copy(order="F")) == -2
        spr.release(smg)


def test_partial_solve_with_peri2e():
    import torch
    from sprgpu import fem, host
    from sprgpu._lib import SprGpuError
    from sprgpu.device import SprFactor

    model = fem.hex8_grid(6, e_sigma=0.5, e_seed=1)
    sysm = host.symbolic(model, nperm=2)
    ldb = sysm.neq - 40
    # external numbers: a permutation inside the leading LDB and inside the rest
    rng = np.random.default_rng(3)
    order = np.concatenate([rng.permutation(ldb), ldb + rng.permutation(sysm.neq - ldb)])
    free = sysm.meqn > 0
    ext = np.zeros_like(sysm.meqn)
    ext[free] = order[sysm.meqn[free] - 1] + 1
    host.set_permutation(sysm, ext)
    sm = host.assemble(model, sysm)
    fac = sm.copy()
    assert oracle.sprsol(1, sysm.mspar.copy(), sysm.mtrees, sysm.msifa, fac)[0] == 0
    mspar = sysm.mspar.copy()
    mspar[56] = 1
    b = np.asfortranarray(rng.uniform(-1, 1, (ldb, 2)))
    xo = b.copy(order="F")
    assert oracle.sprsol(4, mspar.copy(), sysm.mtrees, sysm.msifa, fac.copy(), xo)[0] == 0
    h = SprFactor(sysm.mspar, sysm.mtrees, sysm.msifa, "fp64")
    try:
        assert h.factor(torch.from_numpy(sm).cuda())["ierr"] == 0
        xg = torch.from_numpy(np.ascontiguousarray(b.T)).cuda()
        h.solve_partial(xg, 3)
        assert _rel(xg.cpu().numpy().T, xo) < RTOL
        # a shorter LDB whose external equations are not the leading internal ones
        with pytest.raises(SprGpuError, match="first LDB external"):
            h.solve_partial(torch.from_numpy(np.ascontiguousarray(b[:ldb // 2].T)).cuda(), 3)
    finally:
        h.close()
