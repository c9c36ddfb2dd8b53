"""GPU element assembly (sprgpu_set_assembly / sprgpu_assemble) against the
reference SPRADM loop and the oracle.

The bar is bit equality (compared as 64-bit patterns): the device sums every
LNZ slot's element entries in SPRADM's order.  Covered: hex8 with per-element
scales of one base matrix, hex8 20^3 (the C1 mesh), quad4, beam frames with
one matrix per element, consecutive element matrices (no offsets), adding to
an existing SM, the residual pattern that the map sets (FP32 factor + FP64
refinement on the device-assembled SM, without sprgpu_set_pattern), and the
error paths.  With constraint equations (sprgpu_set_assembly_ceq): SM and SV
against the reference SPRADM for NSV = 0, 1, -1, NPDOF, new TTCC values on
one map, accumulation, the solution of the constrained problem, and the
error paths.
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
    torch = _torch()
    return None if a is None else torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _bits(a):
    return np.ascontiguousarray(a).view(np.int64)


def _setup(kind, n, kw, precision="fp64"):
    from sprgpu import host
    from sprgpu.device import SprFactor

    model, sysm, sm = make_system(kind, n, **kw)
    mp, me = host.element_meen(model, sysm)
    h = SprFactor(sysm.mspar, sysm.mtrees, sysm.msifa, precision)
    h.set_assembly(mp, me)
    return model, sysm, sm, h, mp, me


@pytest.mark.parametrize("kind,n,kw", [("hex8", 10, dict(e_sigma=0.5, e_seed=1)),
                                       ("hex8", 20, {}),
                                       ("quad4", 40, {}),
                                       ("frame", 400, dict(seed=3))])
def test_device_assembly_equals_spradm(kind, n, kw):
    torch = _torch()
    model, sysm, sm, h, _, _ = _setup(kind, n, kw)
    print(f"{kind} {n}: neq {sysm.neq} nzerol {sysm.nzerol} nel {model.nel}", flush=True)
    out = torch.full((sysm.neq + sysm.nzerol,), float("nan"), dtype=torch.float64, device="cuda")
    h.assemble(_dev(model.em), out, _dev(model.em_off), _dev(model.escale))
    torch.cuda.synchronize()
    assert np.array_equal(_bits(out.cpu().numpy()), _bits(sm))
    h.close()


def test_consecutive_matrices_and_accumulate():
    torch = _torch()
    model, sysm, sm, h, mp, me = _setup("hex8", 8, dict(e_sigma=0.5, e_seed=1))
    nd = int(model.nedof[0])
    # one matrix per element, one after the other (no offsets), unscaled
    em = np.concatenate([model.element_matrix(e).ravel(order="F") for e in range(model.nel)])
    out = torch.empty(sysm.neq + sysm.nzerol, dtype=torch.float64, device="cuda")
    h.assemble(_dev(em), out)
    rc, sm_o = oracle.assemble_free(sysm.mspar, sysm.mtrees, sysm.msifa, mp, me, em,
                                    nd * nd * np.arange(model.nel), None)
    assert rc == 0
    torch.cuda.synchronize()
    assert np.array_equal(_bits(out.cpu().numpy()), _bits(sm_o))
    # added to an existing SM, entry by entry in the same order
    sm0 = np.random.default_rng(4).normal(size=sysm.neq + sysm.nzerol)
    acc = _dev(sm0)
    h.assemble(_dev(model.em), acc, _dev(model.em_off), _dev(model.escale), accumulate=True)
    rc, sm_a = oracle.assemble_free(sysm.mspar, sysm.mtrees, sysm.msifa, mp, me, model.em,
                                    model.em_off, model.escale, sm=sm0.copy())
    assert rc == 0
    torch.cuda.synchronize()
    assert np.array_equal(_bits(acc.cpu().numpy()), _bits(sm_a))
    h.close()


def test_assembly_pattern_serves_the_refinement():
    """set_assembly sets the residual pattern: the FP32 factor of the device
    SM refines to 1e-10 and agrees with the oracle's FP64 solution."""
    torch = _torch()
    model, sysm, sm, h, _, _ = _setup("hex8", 12, dict(e_sigma=0.5, e_seed=1), "fp32")
    out = torch.empty(sysm.neq + sysm.nzerol, dtype=torch.float64, device="cuda")
    h.assemble(_dev(model.em), out, _dev(model.em_off), _dev(model.escale))
    assert h.factor(out)["ierr"] == 0
    b = np.random.default_rng(2).uniform(-1, 1, (sysm.neq, 2))
    x = torch.empty(2, sysm.neq, dtype=torch.float64, device="cuda")
    ir = h.solve_ir(_dev(b.T), x, rtol=1e-10, maxit=30)
    print("IR:", ir, flush=True)
    assert ir["converged"] and ir["relres"] <= 1e-10
    x_o = np.asfortranarray(b.copy())
    ierr, _ = oracle.sprsol(3, sysm.mspar.copy(), sysm.mtrees, sysm.msifa, sm.copy(), x_o)
    assert ierr == 0
    assert np.linalg.norm(x.cpu().numpy().T - x_o) / np.linalg.norm(x_o) < 1e-8
    h.close()


def test_assembly_error_paths():
    from sprgpu import host
    from sprgpu._lib import SprGpuError
    from sprgpu.device import SprFactor

    torch = _torch()
    model, sysm, _ = make_system("hex8", 6)
    mp, me = host.element_meen(model, sysm)
    h = SprFactor(sysm.mspar, sysm.mtrees, sysm.msifa, "fp64")
    out = torch.empty(sysm.neq + sysm.nzerol, dtype=torch.float64, device="cuda")
    with pytest.raises(SprGpuError, match="set_assembly first"):
        h.assemble(_dev(model.em), out, _dev(model.em_off))
    bad = me.copy()
    bad[5] = -1
    with pytest.raises(SprGpuError, match="constrained dof"):
        h.set_assembly(mp, bad)
    bad = me.copy()
    bad[5] = sysm.neq + 1
    with pytest.raises(SprGpuError, match="beyond NEQ"):
        h.set_assembly(mp, bad)
    # element 3 couples two equations that the symbolic pattern does not hold
    # (the oracle rejects the same connectivity)
    e, seg = 3, me[mp[3]:mp[4]]
    i0 = int(seg[seg > 0][0]) - 1
    for q in range(sysm.neq):
        try:
            host.lnz_positions(sysm, np.array([max(i0, q)]), np.array([min(i0, q)]))
        except ValueError:
            break
    bad = me.copy()
    bad[mp[e] + int(np.argmax(seg > 0)) + 1] = q + 1
    rc, _ = oracle.assemble_free(sysm.mspar, sysm.mtrees, sysm.msifa, mp, bad, model.em,
                                 model.em_off, model.escale)
    assert rc == -(e + 1)
    with pytest.raises(SprGpuError, match="outside the symbolic pattern"):
        h.set_assembly(mp, bad)
    # a valid map after the failures still assembles correctly
    h.set_assembly(mp, me)
    h.assemble(_dev(model.em), out, _dev(model.em_off), _dev(model.escale))
    rc, sm_o = oracle.assemble_free(sysm.mspar, sysm.mtrees, sysm.msifa, mp, me, model.em,
                                    model.em_off, model.escale)
    torch.cuda.synchronize()
    assert np.array_equal(_bits(out.cpu().numpy()), _bits(sm_o))
    h.close()


# ---------------------------------------------------------------------------
# constraint equations (sprgpu_set_assembly_ceq / sprgpu_assemble_ceq)
# ---------------------------------------------------------------------------
def _ceq_setup(n, nsv):
    from sprgpu import host
    from sprgpu.device import SprFactor

    model, sysm, _ = make_system("hex8c", n)
    mp, me = host.element_meen(model, sysm)
    h = SprFactor(sysm.mspar, sysm.mtrees, sysm.msifa, "fp64")
    mpmceq, mmceq, ttcc = model.ceq_arrays()
    h.set_assembly_ceq(mp, me, mpmceq, mmceq, sysm.meqn, nsv)
    return model, sysm, h, mp, me


def _npdof(model):
    return int(np.count_nonzero(np.diff(model.mpmceq) == 1))


def _run_ceq(h, model, sysm, nsv, ttcc, sm0=None, sv0=None):
    torch = _torch()
    nvec = max(abs(nsv), 1)
    sm = None if nsv < 0 else (_dev(sm0) if sm0 is not None else torch.full(
        (sysm.neq + sysm.nzerol,), float("nan"), dtype=torch.float64, device="cuda"))
    sv = None if nsv == 0 else (_dev(sv0) if sv0 is not None else torch.full(
        (nvec * sysm.neq,), float("nan"), dtype=torch.float64, device="cuda"))
    h.assemble_ceq(_dev(model.em), _dev(ttcc), sm, sv, _dev(model.em_off), _dev(model.escale),
                   accumulate=sm0 is not None or sv0 is not None)
    torch.cuda.synchronize()
    return (None if sm is None else sm.cpu().numpy()), (None if sv is None else sv.cpu().numpy())


@pytest.mark.parametrize("n", [4, 12])
def test_device_constrained_assembly_equals_spradm(n):
    """SM and SV bit for bit against the reference SPRADM with constraint
    equations, for NSV = 0, 1, -1 and NPDOF."""
    from sprgpu import host

    model, sysm, _ = make_system("hex8c", n)
    for nsv in (0, 1, -1, _npdof(model)):
        print(f"hex8c {n}: nsv {nsv} neq {sysm.neq} nceq {model.nceq}", flush=True)
        _, _, h, _, _ = _ceq_setup(n, nsv)
        sm, sv = _run_ceq(h, model, sysm, nsv, model.ttcc)
        sm_r, sv_r = host.assemble_rhs(model, sysm, nsv)
        if nsv >= 0:
            assert np.array_equal(_bits(sm), _bits(sm_r)), nsv
        if nsv != 0:
            assert np.array_equal(_bits(sv), _bits(sv_r)), nsv
        h.close()


def test_constrained_assembly_reads_ttcc_at_each_call():
    """One map, new TTCC values (C0 changed, a zero coefficient made non-zero,
    a coupling coefficient changed): each assembly equals SPRADM's with that
    TTCC; and accumulation into given SM/SV."""
    import dataclasses

    from sprgpu import host

    model, sysm, h, _, _ = _ceq_setup(6, 1)
    rng = np.random.default_rng(7)
    for step in range(3):
        tt = model.ttcc.copy()
        if step > 0:
            tt = tt * rng.uniform(0.5, 1.5, tt.size)
            tt[tt == 0] = rng.uniform(-0.1, 0.1, int(np.count_nonzero(tt == 0)))
        m2 = dataclasses.replace(model, ttcc=tt)
        sm, sv = _run_ceq(h, m2, sysm, 1, tt)
        sm_r, sv_r = host.assemble_rhs(m2, sysm, 1)
        assert np.array_equal(_bits(sm), _bits(sm_r)), step
        assert np.array_equal(_bits(sv), _bits(sv_r)), step
    sm0 = rng.normal(size=sysm.neq + sysm.nzerol)
    sv0 = rng.normal(size=sysm.neq)
    sm, sv = _run_ceq(h, model, sysm, 1, model.ttcc, sm0, sv0)
    sm_r, sv_r = host.assemble_rhs(model, sysm, 1, sm=sm0.copy(), sv=sv0.copy())
    assert np.array_equal(_bits(sm), _bits(sm_r))
    assert np.array_equal(_bits(sv), _bits(sv_r))
    h.close()


def test_constrained_assembly_solves_the_constrained_problem():
    """Device SM and SV, factored and solved on the device with loads F:
    the solution q gives u = T q + u0 with T'(K u - F) = 0 (equilibrium of
    the constrained model), to 1e-10."""
    from test_assembly import _transformation

    torch = _torch()
    model, sysm, h, _, _ = _ceq_setup(6, 1)
    sm, sv = _run_ceq(h, model, sysm, 1, model.ttcc)
    rng = np.random.default_rng(3)
    F = rng.uniform(-1, 1, model.ndof)
    T, u0 = _transformation(model, sysm)
    b = T.T @ F + sv
    assert h.factor(_dev(sm))["ierr"] == 0
    x = _dev(b.copy())
    h.solve(x)
    torch.cuda.synchronize()
    q = x.cpu().numpy()
    K = np.zeros((model.ndof, model.ndof))
    for e in range(model.nel):
        d = model.element_dofs(e)
        K[np.ix_(d, d)] += model.element_matrix(e)
    u = T @ q + u0
    r = T.T @ (K @ u - F)
    assert np.linalg.norm(r) <= 1e-10 * np.linalg.norm(T.T @ F)
    h.close()


def test_constrained_assembly_error_paths():
    from sprgpu import host
    from sprgpu._lib import SprGpuError
    from sprgpu.device import SprFactor

    torch = _torch()
    model, sysm, _ = make_system("hex8c", 4)
    mp, me = host.element_meen(model, sysm)
    mpmceq, mmceq, ttcc = model.ceq_arrays()
    h = SprFactor(sysm.mspar, sysm.mtrees, sysm.msifa, "fp64")
    with pytest.raises(SprGpuError, match="sprgpu_set_assembly_ceq"):
        h.set_assembly(mp, me)
    bad = mpmceq.copy()
    bad[0] = 2
    with pytest.raises(SprGpuError, match="MPMCEQ"):
        h.set_assembly_ceq(mp, me, bad, mmceq, sysm.meqn)
    bad = mmceq.copy()
    slave = mmceq[mpmceq[0] - 1]
    bad[mpmceq[0]] = slave  # a master that is a constrained dof
    with pytest.raises(SprGpuError, match="no equation"):
        h.set_assembly_ceq(mp, me, mpmceq, bad, sysm.meqn)
    with pytest.raises(SprGpuError, match="beyond"):
        h.set_assembly_ceq(mp, me, mpmceq, mmceq, sysm.meqn, 2)
    h.set_assembly_ceq(mp, me, mpmceq, mmceq, sysm.meqn, 1)
    out = torch.empty(sysm.neq + sysm.nzerol, dtype=torch.float64, device="cuda")
    with pytest.raises(SprGpuError, match="TTCC is required"):
        h.assemble(_dev(model.em), out, _dev(model.em_off), _dev(model.escale))
    with pytest.raises(SprGpuError, match="SV is required"):
        h.assemble_ceq(_dev(model.em), _dev(ttcc), out, None, _dev(model.em_off),
                       _dev(model.escale))
    h.close()


# ---------------------------------------------------------------------------
# element matrices on the device (sprgpu_hex8_stiffness)
# ---------------------------------------------------------------------------
def test_device_hex8_stiffness_and_assembly():
    """Distorted hex8 6^3 with log-normal moduli: the device element matrices
    against the isoparametric formula (fem.hex8_stiffness_at, itself pinned
    by tests/test_fem.py) to 1e-13 of each matrix's largest entry; assembled
    on the device they give the host SPRADM loop's SM of those matrices to
    1e-13; a node number out of range gives that element a NaN matrix and
    leaves the others alone."""
    import dataclasses

    from sprgpu import fem, host

    torch = _torch()
    n, nu = 6, 0.28
    model, sysm, _, h, mp, me = _setup("hex8", n, dict(e_sigma=0.5, e_seed=1))
    xyz = fem.hex8_coordinates(n, distort=0.2, seed=5)
    E = model.escale
    conn = model.mmnpc.reshape(model.nel, 8) - 1
    em_ref = np.concatenate([fem.hex8_stiffness_at(xyz[conn[e]], E[e], nu).ravel(order="F")
                             for e in range(model.nel)])
    em = h.hex8_stiffness(_dev(xyz), _dev(model.mmnpc), _dev(E), nu)
    torch.cuda.synchronize()
    got = em.cpu().numpy().reshape(model.nel, 576)
    ref = em_ref.reshape(model.nel, 576)
    err = np.max(np.abs(got - ref), axis=1) / np.max(np.abs(ref), axis=1)
    print(f"hex8 {n}^3 distorted: max element error {err.max():.2e}", flush=True)
    assert err.max() <= 1e-13
    out = torch.empty(sysm.neq + sysm.nzerol, dtype=torch.float64, device="cuda")
    h.assemble(em, out)
    m2 = dataclasses.replace(model, em=em_ref, em_off=(576 * np.arange(model.nel)).astype(np.int64),
                             escale=None)
    sm_ref = host.assemble(m2, sysm)
    torch.cuda.synchronize()
    sm = out.cpu().numpy()
    assert np.max(np.abs(sm - sm_ref)) <= 1e-13 * np.max(np.abs(sm_ref))
    bad = model.mmnpc.copy()
    bad[8 * 5 + 3] = model.nanod + 1
    em2 = h.hex8_stiffness(_dev(xyz), _dev(bad), _dev(E), nu)
    torch.cuda.synchronize()
    g2 = em2.cpu().numpy().reshape(model.nel, 576)
    assert np.all(np.isnan(g2[5])) and np.array_equal(np.delete(g2, 5, 0), np.delete(got, 5, 0))
    h.close()
