"""Element assembly: the oracle's restatement of SPRADM (free dofs, and the
constrained dofs of SAM's constraint equations with the SV right-hand side).

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
  orientation and the summation order;
* with constraint equations (spradm.f:861-1002): SM and SV bit for bit
  against the reference SPRADM for NSV = 0, 1, -1 and NPDOF, and against the
  transformation they stand for (SM = T'KT, SV = -T'K u0).  Doing the free
  terms of all elements before the constrained ones (a plausible slip) gives
  other bits: the per-element order is pinned.
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


# ---------------------------------------------------------------------------
# constraint equations: SPRAD2's constrained part (spradm.f:861-1002)
# ---------------------------------------------------------------------------
def _oracle_ceq(model, sysm, nsv, **kw):
    from sprgpu import host

    mp, me = host.element_meen(model, sysm)
    mpmceq, mmceq, ttcc = model.ceq_arrays()
    return oracle.assemble_ceq(sysm.mspar, sysm.mtrees, sysm.msifa, mp, me, model.em,
                               model.em_off, model.escale, mpmceq, mmceq, ttcc, sysm.meqn,
                               nsv, **kw)


def _npdof(model):
    return int(np.count_nonzero(np.diff(model.mpmceq) == 1))


@pytest.mark.parametrize("n", [3, 5])
def test_oracle_constrained_assembly_equals_reference_spradm(n):
    """SM and SV bit for bit against the reference SPRADM for NSV = 0, 1, -1
    and NPDOF (one vector per prescribed dof)."""
    from sprgpu import host

    model, sysm, sm = make_system("hex8c", n)
    assert np.any(host.element_meen(model, sysm)[1] < 0)
    for nsv in (0, 1, -1, _npdof(model)):
        sm_r, sv_r = host.assemble_rhs(model, sysm, nsv)
        rc, sm_o, sv_o = _oracle_ceq(model, sysm, nsv)
        assert rc == 0
        assert np.array_equal(_bits(sm_o), _bits(sm_r)), nsv
        assert np.array_equal(_bits(sv_o), _bits(sv_r)), nsv
        assert (np.count_nonzero(sv_r) > 0) == (nsv != 0)
        assert (np.count_nonzero(sm_r) > 0) == (nsv >= 0)
    assert np.array_equal(_bits(host.assemble_rhs(model, sysm, 0)[0]), _bits(sm))


def _transformation(model, sysm):
    """T (NDOF x NEQ) and u0 with u = T q + u0: free dofs map to their
    equation, a slave to its masters' equations with the TTCC weights, and
    every constrained dof carries its c0 in u0."""
    T = np.zeros((model.ndof, sysm.neq))
    u0 = np.zeros(model.ndof)
    free = sysm.meqn > 0
    T[np.nonzero(free)[0], sysm.meqn[free] - 1] = 1.0
    mp, mm, tt = model.mpmceq - 1, model.mmceq - 1, model.ttcc
    for i in range(model.nceq):
        s = mm[mp[i]]
        u0[s] = tt[mp[i]]
        for ip in range(mp[i] + 1, mp[i + 1]):
            T[s, sysm.meqn[mm[ip]] - 1] += tt[ip]
    return T, u0


def test_oracle_constrained_assembly_is_tt_k_t():
    """Mathematics: SM = T'KT (lower triangle), SV = -T'K u0 for NSV = 1, and
    the NPDOF vectors are -T'K e_p c0_p of the prescribed dofs in equation
    order; K the natural-dof stiffness, T and u0 the constraint map.  The model
    has slaves whose master sits in the same element (the doubled diagonal
    term), slaves sharing masters (master pairs on one slot) and a master with
    a zero coefficient."""
    from test_oracle import _dense_lower

    model, sysm, _ = make_system("hex8c", 4)
    K = np.zeros((model.ndof, model.ndof))
    for e in range(model.nel):
        d = model.element_dofs(e)
        K[np.ix_(d, d)] += model.element_matrix(e)
    T, u0 = _transformation(model, sysm)
    A = T.T @ K @ T
    rc, sm, sv1 = _oracle_ceq(model, sysm, 1)
    assert rc == 0
    L = _dense_lower(sysm, sm[sysm.neq:])
    tol = 1e-13 * np.max(np.abs(A))
    assert np.max(np.abs(L - np.tril(A))) <= tol
    assert np.max(np.abs(sv1 + T.T @ K @ u0)) <= 1e-13 * np.max(np.abs(K @ u0))
    npd = _npdof(model)
    rc, _, svp = _oracle_ceq(model, sysm, -npd)
    assert rc == 0
    svp = svp.reshape(npd, sysm.neq).T
    presc = [i for i in range(model.nceq) if model.mpmceq[i + 1] - model.mpmceq[i] == 1]
    for k, i in enumerate(presc):
        s = model.mmceq[model.mpmceq[i] - 1] - 1
        ref = -T.T @ K[:, s] * model.ttcc[model.mpmceq[i] - 1]
        assert np.max(np.abs(svp[:, k] - ref)) <= 1e-13 * max(np.max(np.abs(K[:, s])), 1.0)


def test_oracle_constrained_assembly_rejects_too_few_vectors():
    model, sysm, _ = make_system("hex8c", 3)
    rc, _, _ = _oracle_ceq(model, sysm, 2)
    assert rc == -(model.nel + 1)
