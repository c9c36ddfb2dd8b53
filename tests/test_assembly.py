"""Element assembly: the oracle's restatement of SPRADM's free-dof loop.

The GPU assembly is checked against this restatement (tests/test_gpu_assembly.py),
so it is pinned here first:
* bit for bit against the reference SPRADM (spradm.f:5-166, compiled from its
  sources into lib/libspr_host.so and called once per element), on hex8 with
  per-element scales, quad4, beam frames, and on non-symmetric element matrices,
  which fixes that SPRAD2 reads EM(I,J) with I <= J only (spradm.f:818-859);
* bit for bit against a dense assembly A[eq, eq] += K_e in element order
  (mathematics, no SPR code), gathered at the LNZ slots;
* planted mistakes (a transposed element matrix, the reverse element order)
  must fail the reference comparison, so the comparison pins the index
  orientation and the summation order.
"""
import dataclasses

import numpy as np
import pytest

import oracle
from conftest import make_system
from test_oracle import dense_equation_matrix

CASES = [("hex8", 6, dict(e_sigma=0.5, e_seed=1)), ("quad4", 12, {}), ("frame", 60, dict(seed=3))]


def _bits(a):
    return np.ascontiguousarray(a).view(np.int64)


def _oracle(model, sysm, **kw):
    from sprgpu import host

    mp, me = host.element_meen(model, sysm)
    return oracle.assemble_free(sysm.mspar, sysm.mtrees, sysm.msifa, mp, me, model.em,
                                model.em_off, model.escale, **kw)


@pytest.mark.parametrize("kind,n,kw", CASES)
def test_oracle_assembly_equals_reference_spradm(kind, n, kw):
    model, sysm, sm = make_system(kind, n, **kw)
    rc, sm_o = _oracle(model, sysm)
    assert rc == 0
    assert np.array_equal(_bits(sm_o), _bits(sm))


def test_oracle_reads_the_upper_triangle_like_spradm():
    """Non-symmetric element matrices: SPRAD2 takes EM(I,J), I <= J."""
    from sprgpu import host

    model, sysm, _ = make_system("hex8", 4, e_sigma=0.5, e_seed=1)
    rng = np.random.default_rng(5)
    nd = int(model.nedof[0])
    em = rng.normal(size=model.nel * nd * nd)
    m = dataclasses.replace(model, em=em, em_off=(nd * nd * np.arange(model.nel)).astype(np.int64))
    ref = host.assemble(m, sysm)
    rc, sm_o = _oracle(m, sysm)
    assert rc == 0 and np.array_equal(_bits(sm_o), _bits(ref))
    # planted mistake: EM(J,I) instead of EM(I,J)
    emt = em.reshape(model.nel, nd, nd).transpose(0, 2, 1).ravel()
    rc, sm_t = _oracle(dataclasses.replace(m, em=emt), sysm)
    assert rc == 0 and not np.array_equal(sm_t, ref)


def test_oracle_summation_order_is_pinned():
    """The reverse element order gives other bits somewhere: bit equality with
    the reference fixes the order of the sums, element by element."""
    from sprgpu import host

    model, sysm, sm = make_system("hex8", 6, e_sigma=0.5, e_seed=1)
    mp, me = host.element_meen(model, sysm)
    nd = np.diff(mp)
    order = np.arange(model.nel)[::-1]
    me_r = np.concatenate([me[mp[e]:mp[e + 1]] for e in order])
    mp_r = np.concatenate([[0], np.cumsum(nd[order])])
    rc, sm_r = oracle.assemble_free(sysm.mspar, sysm.mtrees, sysm.msifa, mp_r, me_r, model.em,
                                    model.em_off[order], model.escale[order])
    assert rc == 0
    assert np.max(np.abs(sm_r - sm)) <= 1e-14 * np.max(np.abs(sm))
    assert not np.array_equal(_bits(sm_r), _bits(sm))


@pytest.mark.parametrize("kind,n,kw", CASES[:2])
def test_oracle_assembly_equals_dense_assembly(kind, n, kw):
    from sprgpu import host

    model, sysm, _ = make_system(kind, n, **kw)
    rc, sm_o = _oracle(model, sysm)
    A = dense_equation_matrix(model, sysm)
    r, c = np.nonzero(np.tril(np.ones_like(A, dtype=bool)) & (A != 0))
    pos = host.lnz_positions(sysm, r, c)
    assert np.array_equal(_bits(sm_o[pos]), _bits(A[r, c]))
    rest = np.ones(sm_o.size, dtype=bool)
    rest[pos] = False
    assert not np.any(sm_o[rest])


def test_oracle_rejects_an_element_outside_the_pattern():
    from sprgpu import host

    model, sysm, _ = make_system("hex8", 6)
    mp, me = host.element_meen(model, sysm)
    me = me.copy()
    e = 3
    seg = me[mp[e]:mp[e + 1]]
    # couple element e's first equation with an equation its pattern lacks
    i0 = int(seg[seg > 0][0]) - 1
    for q in range(sysm.neq):
        try:
            host.lnz_positions(sysm, np.array([max(i0, q)]), np.array([min(i0, q)]))
        except ValueError:
            break
    else:
        pytest.skip("the pattern is dense")
    me[mp[e] + int(np.argmax(seg > 0)) + 1] = q + 1
    rc, _ = oracle.assemble_free(sysm.mspar, sysm.mtrees, sysm.msifa, mp, me, model.em,
                                 model.em_off, model.escale)
    assert rc == -(e + 1)


def test_element_meen_matches_element_dofs():
    from sprgpu import host

    for kind, n, kw in CASES:
        model, sysm, _ = make_system(kind, n, **kw)
        mp, me = host.element_meen(model, sysm)
        for e in range(model.nel):
            assert np.array_equal(me[mp[e]:mp[e + 1]], sysm.meqn[model.element_dofs(e)])
