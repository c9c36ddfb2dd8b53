"""The PERI2E-permuted variants (MSPAR(58) >= 1): the caller numbers its
vectors in an external equation order, and MSIFA(1:NEQ) = PERI2E (filled by
SPRABC, sprabc.f:5-66) maps every internal equation to it.

CPU: the oracle's restatement of SPRFS4 / SPRBS5 (with SPRFB3 / SPRFB4,
sprfs1.f:396-528, :636-926, sprbs1.f:371-495), SPRPRR (sprprm.f:327-397) and
SPRTMR (sprtmm.f:345-480) equals the reference built from source bit for
bit, and the permuted results are the unpermuted ones moved by PERI2E.
GPU: the handle's solve, refinement, SpMV, SPRPRM and SPRTMM and the drop-in
SPRSOL on a permuted system against the oracle and the reference.
"""
import numpy as np
import pytest

import oracle
from conftest import gpu_available


def perm_system(kind="hex8", n=5, seed=7):
    """A system whose symbolic phase reserved PERI2E/PERE2I (MSPAR(58) = 2)
    and whose external order is a random numbering of the free dofs."""
    from sprgpu import fem, host

    model = fem.hex8_grid(n, e_sigma=0.5, e_seed=1) if kind == "hex8" else fem.quad4_grid(n)
    sysm = host.symbolic(model, nperm=2)
    free = sysm.meqn > 0
    ext = np.zeros_like(sysm.meqn)
    ext[free] = np.random.default_rng(seed).permutation(sysm.neq) + 1
    host.set_permutation(sysm, ext)
    sm = host.assemble(model, sysm)
    peri2e = sysm.msifa[:sysm.neq] - 1
    return model, sysm, sm, ext, peri2e


def unpermuted(mspar):
    ms = mspar.copy()
    ms[57] = 0  # same arrays, the permutation switched off
    return ms


def test_set_permutation_fills_peri2e_and_pere2i():
    _, sysm, _, ext, peri2e = perm_system()
    assert sysm.mspar[57] == 2 and sysm.mspar[46] == 1 + 2 * sysm.neq
    pere2i = sysm.msifa[sysm.neq:2 * sysm.neq] - 1
    assert np.array_equal(np.sort(peri2e), np.arange(sysm.neq))
    assert np.array_equal(pere2i[peri2e], np.arange(sysm.neq))
    free = sysm.meqn > 0
    assert np.array_equal(peri2e[sysm.meqn[free] - 1], ext[free] - 1)


def test_oracle_permuted_solves_equal_the_reference():
    import refspr

    _, sysm, sm, _, peri2e = perm_system()
    b = np.asfortranarray(np.random.default_rng(1).uniform(-1, 1, (sysm.neq, 3)))
    sm_r, x_r, ms_r = sm.copy(), b.copy(order="F"), sysm.mspar.copy()
    ierr_r, _, _ = refspr.sol(3, sysm, sm_r, x_r, mspar=ms_r)
    sm_o, x_o, ms_o = sm.copy(), b.copy(order="F"), sysm.mspar.copy()
    ierr_o, _ = oracle.sprsol(3, ms_o, sysm.mtrees, sysm.msifa, sm_o, x_o)
    assert ierr_r == ierr_o == 0
    assert np.array_equal(ms_r, ms_o)
    assert np.array_equal(sm_r.view(np.int64), sm_o.view(np.int64))
    assert np.array_equal(x_r.view(np.int64), x_o.view(np.int64))
    # the permuted solution is the internal-order one moved by PERI2E
    x_i = np.asfortranarray(b[peri2e])
    ierr, _ = oracle.sprsol(3, unpermuted(sysm.mspar), sysm.mtrees, sysm.msifa, sm.copy(), x_i)
    assert ierr == 0
    assert np.array_equal(x_o[peri2e], x_i)


@pytest.mark.parametrize("iflag", [1, -2])
def test_oracle_sprprr_equals_the_reference(iflag):
    import refspr

    _, sysm, sm, _, peri2e = perm_system(seed=3)
    lnz = sm[sysm.neq:]
    rng = np.random.default_rng(2)
    b = np.asfortranarray(rng.normal(size=(sysm.neq, 2)))
    c0 = np.asfortranarray(rng.normal(size=(sysm.neq, 2)))
    ierr_r, c_r, _ = refspr.prm(lnz, b, sysm.mspar.copy(), sysm.mtrees, sysm.msifa, iflag=iflag) \
        if iflag == 1 else _ref_prm_acc(refspr, lnz, b, c0, sysm, iflag)
    ierr_o, c_o = oracle.sprprm(lnz, b, sysm.mspar.copy(), sysm.mtrees, sysm.msifa, iflag=iflag,
                                c=c0.copy(order="F"))
    assert ierr_r == ierr_o == 0
    assert np.array_equal(c_r.view(np.int64), c_o.view(np.int64))
    # moved by PERI2E from the internal-order product
    _, c_i = oracle.sprprm(lnz, np.asfortranarray(b[peri2e]), unpermuted(sysm.mspar),
                           sysm.mtrees, sysm.msifa, iflag=iflag,
                           c=np.asfortranarray(c0[peri2e]))
    assert np.array_equal(c_o[peri2e], c_i)
    # a diagonal A with PERI2E is the reference's IERR -3
    ierr_d, _ = oracle.sprprm(np.ones(sysm.neq), b, sysm.mspar.copy(), sysm.mtrees, sysm.msifa,
                              ksa=0)
    assert ierr_d == -3


def _ref_prm_acc(refspr, lnz, b, c0, sysm, iflag):
    return refspr.prm(lnz, b, sysm.mspar.copy(), sysm.mtrees, sysm.msifa, iflag=iflag,
                      c=c0.copy(order="F"))


@pytest.mark.parametrize("iflag", [1, 2])
def test_oracle_sprtmr_equals_the_reference(iflag):
    import refspr

    _, sysm, sm, _, peri2e = perm_system(seed=5)
    f = sm.copy()
    ms = sysm.mspar.copy()
    assert oracle.sprsol(1, ms, sysm.mtrees, sysm.msifa, f)[0] == 0  # U = the factor
    b = np.asfortranarray(np.random.default_rng(4).normal(size=(sysm.neq, 2)))
    ierr_r, c_r, _ = refspr.tmm(f[sysm.neq:], b, sysm.mspar.copy(), sysm.mtrees, sysm.msifa,
                                iflag=iflag)
    ierr_o, c_o, _ = oracle.sprtmm(f[sysm.neq:], b, sysm.mspar.copy(), sysm.mtrees, sysm.msifa,
                                   iflag=iflag)
    assert ierr_r == ierr_o == 0
    assert np.array_equal(c_r.view(np.int64), c_o.view(np.int64))
    _, c_i, _ = oracle.sprtmm(f[sysm.neq:], np.asfortranarray(b[peri2e]), unpermuted(sysm.mspar),
                              sysm.mtrees, sysm.msifa, iflag=iflag)
    assert np.array_equal(c_o[peri2e], c_i)


# --------------------------------------------------------------------------
# GPU
# --------------------------------------------------------------------------


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs an MI355X")
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_handle_with_peri2e(precision):
    """The handle takes B, X, C and WA in the external order."""
    import torch
    from sprgpu.device import SprFactor

    _, sysm, sm, _, peri2e = perm_system(n=10)
    print(f"PERI2E hex8 10^3: neq={sysm.neq}, {precision}", flush=True)
    b = np.random.default_rng(6).uniform(-1, 1, (sysm.neq, 3))
    x_o = np.asfortranarray(b.copy())
    assert oracle.sprsol(3, sysm.mspar.copy(), sysm.mtrees, sysm.msifa, sm.copy(), x_o)[0] == 0
    h = SprFactor(sysm.mspar, sysm.mtrees, sysm.msifa, precision)
    try:
        h.set_pattern(sm)
        sm_dev = torch.from_numpy(sm).cuda()
        assert h.factor(sm_dev)["ierr"] == 0
        bd = torch.from_numpy(np.ascontiguousarray(b.T)).cuda()
        if precision == "fp64":
            x = bd.clone()
            h.solve(x)
        else:
            x = torch.empty_like(bd)
            ir = h.solve_ir(bd, x, rtol=1e-12, maxit=30)
            assert ir["converged"], ir
        rel = np.linalg.norm(x.cpu().numpy().T - x_o) / np.linalg.norm(x_o)
        print(f"solution vs oracle {rel:.2e}", flush=True)
        assert rel < 1e-9
        # A x in the external order gives b back
        y = torch.empty_like(x)
        h.spmv(x, y)
        assert np.linalg.norm(y.cpu().numpy().T - b) / np.linalg.norm(b) < 1e-10
        # SPRPRR and SPRTMR: bit for bit the oracle's (itself the reference's)
        a = sm_dev[sysm.neq:]
        xb = np.asfortranarray(np.random.default_rng(8).normal(size=(sysm.neq, 2)))
        c0 = np.asfortranarray(np.random.default_rng(9).normal(size=(sysm.neq, 2)))
        _, c_o = oracle.sprprm(sm[sysm.neq:], xb, sysm.mspar.copy(), sysm.mtrees, sysm.msifa,
                               iflag=-2, c=c0.copy(order="F"))
        cg = torch.from_numpy(np.ascontiguousarray(c0.T)).cuda()
        h.sprprm(a, torch.from_numpy(np.ascontiguousarray(xb.T)).cuda(), cg, iflag=-2)
        assert np.array_equal(cg.cpu().numpy().T, c_o)
        for iflag in (1, 2):
            _, t_o, _ = oracle.sprtmm(sm[sysm.neq:], xb, sysm.mspar.copy(), sysm.mtrees,
                                      sysm.msifa, iflag=iflag)
            tg = h.sprtmm(a, torch.from_numpy(np.ascontiguousarray(xb.T)).cuda(), iflag=iflag)
            assert np.array_equal(tg.cpu().numpy().T, t_o), iflag
    finally:
        h.close()


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs an MI355X")
def test_dropin_sprsol_with_peri2e_matches_the_reference():
    import refspr
    from sprgpu import spr

    _, sysm, sm, _, _ = perm_system(n=8, seed=11)
    b = np.asfortranarray(np.random.default_rng(12).uniform(-1, 1, (sysm.neq, 2)))
    sm_r, x_r, ms_r = sm.copy(), b.copy(order="F"), sysm.mspar.copy()
    assert refspr.sol(3, sysm, sm_r, x_r, mspar=ms_r)[0] == 0
    sm_g, x_g, ms_g = sm.copy(), b.copy(order="F"), sysm.mspar.copy()
    ierr = spr.sprsol(3, ms_g, sysm.mtrees, sysm.msifa, sm_g, x_g, np.zeros(3))
    # separate SPRFS1 / SPRBS1 calls with LDB > NEQ
    x3 = np.zeros((sysm.neq + 5, 2), order="F")
    x3[:sysm.neq] = b
    assert spr.sprfs1(ms_g, sysm.mtrees, sysm.msifa, sm_g, x3) == 0
    assert spr.sprbs1(ms_g, sysm.mtrees, sysm.msifa, sm_g, x3) == 0
    spr.release(sm_g)
    assert ierr == 0 and np.array_equal(ms_g, ms_r)
    for x in (x_g, x3[:sysm.neq]):
        rel = np.linalg.norm(x - x_r) / np.linalg.norm(x_r)
        assert rel < 1e-9, rel
    assert np.all(x3[sysm.neq:] == 0)
