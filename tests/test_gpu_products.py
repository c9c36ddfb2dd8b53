"""GPU SPR-layout product (sprgpu_sprprm) against the oracle's SPRPRM.

The oracle's SPRPRM is pinned bit for bit to the reference's
(tests/test_oracle.py::test_sprprm_bitexact_vs_reference).  The device product
sums each row in the reference's order, so the bar is equality of the values
(the sign of an exactly zero sum is not compared).  Covered: IFLAG +-1, +-2
and 88, a diagonal A (KSA <= 0), several right-hand sides and WA, hex8 with
per-element moduli, quad4, frames, a second matrix on the assembly map's
pattern (another parameter set, as a mass or geometric matrix would be), and
the check that rejects an A with non-zeros off the pattern.
"""
import numpy as np
import pytest

import oracle
from conftest import make_system

pytestmark = pytest.mark.gpu


def _torch():
    import torch

    return torch


def _dev(a):
    return _torch().from_numpy(np.ascontiguousarray(a)).cuda()


def _oracle_prm(sysm, a, x, iflag, ksa=1, c0=None):
    c = np.asfortranarray(np.zeros_like(x) if c0 is None else c0.copy())
    ierr, c = oracle.sprprm(a, np.asfortranarray(x), sysm.mspar.copy(), sysm.mtrees, sysm.msifa,
                            iflag=iflag, c=c, ksa=ksa)
    assert ierr == 0
    return c


@pytest.mark.parametrize("kind,n,kw", [("hex8", 10, dict(e_sigma=0.5, e_seed=1)),
                                       ("quad4", 30, {}), ("frame", 300, dict(seed=3))])
def test_sprprm_matches_oracle(kind, n, kw):
    from sprgpu.device import SprFactor

    torch = _torch()
    model, sysm, sm = make_system(kind, n, **kw)
    h = SprFactor(sysm.mspar, sysm.mtrees, sysm.msifa, "fp64")
    h.set_pattern(sm)
    rng = np.random.default_rng(3)
    x = rng.normal(size=(sysm.neq, 3))
    a = sm[sysm.neq:]
    a_dev, x_dev = _dev(a), _dev(x.T)
    for iflag in (1, -1, 2, -2):
        c0 = rng.normal(size=(sysm.neq, 3))
        c_dev = _dev(c0.T)
        wa = torch.empty(sysm.neq, dtype=torch.float64, device="cuda")
        h.sprprm(a_dev, x_dev, c_dev, iflag=iflag, wa_dev=wa)
        want = _oracle_prm(sysm, a, x, iflag, c0=c0)
        got = c_dev.cpu().numpy().T
        assert np.array_equal(got, want), (iflag, np.abs(got - want).max())
        # WA holds the last column's +-A b
        last = _oracle_prm(sysm, a, x[:, -1:], 1 if iflag > 0 else -1)[:, 0]
        assert np.array_equal(wa.cpu().numpy(), last)
    # IFLAG 88: one column, result in WA only
    wa = torch.empty(sysm.neq, dtype=torch.float64, device="cuda")
    h.sprprm(a_dev, x_dev[:1], None, iflag=88, wa_dev=wa)
    assert np.array_equal(wa.cpu().numpy(), _oracle_prm(sysm, a, x[:, :1], 1)[:, 0])
    # a diagonal A (KSA <= 0)
    dvec = rng.normal(size=sysm.neq)
    c_dev = torch.zeros(3, sysm.neq, dtype=torch.float64, device="cuda")
    h.sprprm(_dev(dvec), x_dev, c_dev, iflag=-1, ksa=0)
    assert np.array_equal(c_dev.cpu().numpy().T, _oracle_prm(sysm, dvec, x, -1, ksa=0))
    h.close()


def test_sprprm_on_the_assembly_pattern_and_the_pattern_check():
    """The assembly map's pattern serves any matrix assembled on the same
    connectivity; an A with a non-zero off the pattern is rejected."""
    from sprgpu import host
    from sprgpu._lib import SprGpuError
    from sprgpu.device import SprFactor

    torch = _torch()
    model, sysm, sm = make_system("hex8", 8, e_sigma=0.5, e_seed=1)
    h = SprFactor(sysm.mspar, sysm.mtrees, sysm.msifa, "fp64")
    mp, me = host.element_meen(model, sysm)
    h.set_assembly(mp, me)
    # a second matrix on the same connectivity: other per-element factors
    scale2 = np.random.default_rng(8).uniform(0.5, 2.0, model.nel)
    rc, sm2 = oracle.assemble_free(sysm.mspar, sysm.mtrees, sysm.msifa, mp, me, model.em,
                                   model.em_off, scale2)
    assert rc == 0
    a2 = sm2[sysm.neq:]
    x = np.random.default_rng(9).normal(size=(sysm.neq, 2))
    c_dev = torch.empty(2, sysm.neq, dtype=torch.float64, device="cuda")
    h.sprprm(_dev(a2), _dev(x.T), c_dev, iflag=1)
    assert np.array_equal(c_dev.cpu().numpy().T, _oracle_prm(sysm, a2, x, 1))
    # a fill slot (zero in every assembled matrix) made non-zero
    fill = np.nonzero(a2 == 0.0)[0]
    assert fill.size > 0
    a3 = a2.copy()
    a3[fill[fill.size // 2]] = 1.0
    with pytest.raises(SprGpuError, match="outside the pattern"):
        h.sprprm(_dev(a3), _dev(x.T), c_dev, iflag=1)
    # with a pattern taken from that matrix it is accepted
    sm3 = np.concatenate([np.zeros(sysm.neq), a3])
    h.set_pattern(sm3)
    h.sprprm(_dev(a3), _dev(x.T), c_dev, iflag=1)
    assert np.array_equal(c_dev.cpu().numpy().T, _oracle_prm(sysm, a3, x, 1))
    with pytest.raises(SprGpuError, match="IFLAG"):
        h.sprprm(_dev(a3), _dev(x.T), c_dev, iflag=3)
    h.close()


@pytest.mark.parametrize("kind,n,kw", [("hex8", 10, dict(e_sigma=0.5, e_seed=1)),
                                       ("quad4", 30, {}), ("frame", 300, dict(seed=3))])
def test_sprtmm_matches_oracle(kind, n, kw):
    """SPRTMM with the stored lower triangle of the assembled matrix and of
    the factor (fill included): bit for bit equal to the oracle's SPRTMM,
    which tests/test_oracle.py pins to the reference."""
    from sprgpu.device import SprFactor

    torch = _torch()
    model, sysm, sm = make_system(kind, n, **kw)
    fac = sm.copy()
    assert oracle.sprsol(1, sysm.mspar.copy(), sysm.mtrees, sysm.msifa, fac)[0] == 0
    h = SprFactor(sysm.mspar, sysm.mtrees, sysm.msifa, "fp64")
    rng = np.random.default_rng(4)
    x = rng.normal(size=(sysm.neq, 3))
    x_dev = _dev(x.T)

    def bits(a):
        return np.ascontiguousarray(a).view(np.int64)

    for u in (sm[sysm.neq:], fac[sysm.neq:], rng.normal(size=sysm.neq)):
        ksu = 1 if u.size == sysm.nzerol else 0
        u_dev = _dev(u)
        for iflag in (1, 2):
            wa = torch.empty(sysm.neq, dtype=torch.float64, device="cuda")
            c = h.sprtmm(u_dev, x_dev, iflag=iflag, ksu=ksu, wa_dev=wa)
            ierr, c_o, wa_o = oracle.sprtmm(u, x, sysm.mspar.copy(), sysm.mtrees, sysm.msifa,
                                            iflag, ksu)
            assert ierr == 0
            assert np.array_equal(bits(c.cpu().numpy().T), bits(c_o)), (ksu, iflag)
            assert np.array_equal(bits(wa.cpu().numpy()), bits(wa_o))
        for iflag in (11, 22):
            wa = torch.empty(sysm.neq, dtype=torch.float64, device="cuda")
            h.sprtmm(u_dev, x_dev[:1], None, iflag=iflag, ksu=ksu, wa_dev=wa)
            ierr, _, wa_o = oracle.sprtmm(u, x[:, :1], sysm.mspar.copy(), sysm.mtrees,
                                          sysm.msifa, iflag, ksu)
            assert ierr == 0 and np.array_equal(bits(wa.cpu().numpy()), bits(wa_o))
    h.close()


@pytest.mark.parametrize("kind,n,kw", [("hex8", 10, dict(e_sigma=0.5, e_seed=1)),
                                       ("quad4", 30, {})])
def test_sprtra_is_bt_a_b(kind, n, kw):
    """SPRTRA (sprtra.f:5-56): C = B' A B.  The reference cannot be built here
    (it calls the host BLAS DDOT, absent from the image), so C is held to the
    mathematics: B' W with W = A B from the oracle's SPRPRM (bit-exact to the
    reference's), to 1e-13 of |B'||A B|, exactly symmetric; WA is the last
    column's A b bit for bit (the reference's SPRPRM IFLAG 88); a diagonal A
    (KSA <= 0) gives B' diag(a) B."""
    from sprgpu.device import SprFactor

    model, sysm, sm = make_system(kind, n, **kw)
    h = SprFactor(sysm.mspar, sysm.mtrees, sysm.msifa, "fp64")
    try:
        h.set_pattern(sm)
        rng = np.random.default_rng(4)
        b = rng.normal(size=(sysm.neq, 6))
        a = sm[sysm.neq:]
        c, wa = h.sprtra(_dev(a), _dev(b.T), ksa=1)
        c = c.cpu().numpy()
        w = _oracle_prm(sysm, a, b, 1)
        want = b.T @ w
        scale = np.abs(b).T @ np.abs(w)
        assert np.all(np.abs(c - want) <= 1e-13 * scale), np.abs(c - want).max()
        assert np.array_equal(c, c.T)
        assert np.array_equal(wa.cpu().numpy(), w[:, -1])
        dvec = rng.normal(size=sysm.neq)
        cd, _ = h.sprtra(_dev(dvec), _dev(b.T), ksa=0)
        want = b.T @ (dvec[:, None] * b)
        scale = np.abs(b).T @ (np.abs(dvec)[:, None] * np.abs(b))
        assert np.all(np.abs(cd.cpu().numpy() - want) <= 1e-13 * scale)
    finally:
        h.close()
