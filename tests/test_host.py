"""Host side of the drop-in (sprgpu.host): SPR addressing against the reference assembly."""
import numpy as np
import pytest

from conftest import make_system


@pytest.mark.parametrize("kind,n,kw", [("hex8", 3, {"e_sigma": 0.5, "e_seed": 4}),
                                       ("quad4", 6, {}), ("frame", 25, {"seed": 3})])
def test_lnz_positions_reproduce_spradm(kind, n, kw):
    """Scattering every element matrix through lnz_positions, in element order,
    gives SPRADM's SM bit for bit (spradm.f:5; the reference adds elements in
    the same order), so the addressing of every stored entry is right."""
    from sprgpu import host

    model, sysm, sm = make_system(kind, n, **kw)
    own = np.zeros_like(sm)
    for e in range(model.nel):
        q = sysm.meqn[model.element_dofs(e)] - 1
        k = model.element_matrix(e)
        ii, jj = np.nonzero((q[:, None] >= 0) & (q[None, :] >= 0) & (q[:, None] >= q[None, :]))
        np.add.at(own, host.lnz_positions(sysm, q[ii], q[jj]), k[ii, jj])
    assert np.array_equal(own[sysm.neq:], sm[sysm.neq:])


def test_lnz_positions_reject_entries_outside_the_pattern():
    from sprgpu import host

    _, sysm, _ = make_system("hex8", 3)
    with pytest.raises(ValueError):
        host.lnz_positions(sysm, np.array([0]), np.array([1]))  # upper triangle
    # the first and the last equation of a 3^3 hex8 grid are not coupled
    with pytest.raises(ValueError):
        host.lnz_positions(sysm, np.array([sysm.neq - 1]), np.array([0]))


def test_element_equations_mark_fixed_dofs():
    from sprgpu import host

    model, sysm, _ = make_system("hex8", 3)
    eq = host.element_equations(model, sysm)
    assert eq.shape == (model.nel, 24)
    # element 0 touches the clamped face z=0 with its first four nodes
    assert np.all(eq[0, :12] == -1) and np.all(eq[0, 12:] >= 0)
    assert eq.max() == sysm.neq - 1


def test_merged_forest_solves_each_system():
    """The forest of several systems (host.merge_systems) gives each system's
    own solution bit for bit with the oracle, and plans without error."""
    import ctypes as C

    import oracle
    from sprgpu import _lib, fem, host

    rng = np.random.default_rng(3)
    models = [fem.beam_frame(int(n), rng) for n in (30, 45, 25)] + [fem.hex8_grid(2)]
    systems = [host.symbolic(m) for m in models]
    sms = [host.assemble(m, s) for m, s in zip(models, systems)]
    merged = host.merge_systems(systems)
    out = np.zeros(10)
    assert _lib.lib().sprgpu_plan_summary(*(a.ctypes.data_as(C.c_void_p) for a in
                                            (merged.mspar, merged.mtrees, merged.msifa, out))) == 0
    assert out[0] == merged.neq == sum(s.neq for s in systems)
    b = np.random.default_rng(0).normal(size=(merged.neq, 1))
    xm = np.asfortranarray(b.copy())
    ierr, _ = oracle.sprsol(3, merged.mspar.copy(), merged.mtrees, merged.msifa,
                            merged.merge_sm(sms), xm, tol=[0, 0, 0])
    assert ierr == 0
    for k, (s, sm, x) in enumerate(zip(systems, sms, merged.split(xm[:, 0]))):
        bi = np.asfortranarray(b[merged.neq_off[k]:merged.neq_off[k + 1]].copy())
        e, _ = oracle.sprsol(3, s.mspar.copy(), s.mtrees, s.msifa, sm.copy(), bi, tol=[0, 0, 0])
        assert e == 0 and np.array_equal(bi[:, 0], x)
