"""GPU element assembly (sprgpu_set_assembly / sprgpu_assemble) against the
reference SPRADM loop and the oracle.

The bar is bit equality (compared as 64-bit patterns): the device sums every
LNZ slot's element entries in SPRADM's order.  Covered: hex8 with per-element
scales of one base matrix, hex8 20^3 (the C1 mesh), quad4, beam frames with
one matrix per element, consecutive element matrices (no offsets), adding to
an existing SM, the residual pattern that the map sets (FP32 factor + FP64
refinement on the device-assembled SM, without sprgpu_set_pattern), and the
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
