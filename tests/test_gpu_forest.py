"""The persistent forest schedule: structures whose fronts all have F <= 64
(batches of small systems merged into one forest, small meshes) are factored,
and solved one right-hand side at a time, by one launch over all supernodes.

Checked against the level schedule on the same handle (sprgpu_set_plan_params
forest = 0): the factor in SM layout, d, DCOR and the factor outputs are the
same bits (the per-entry operations are the level schedule's small-front
kernel's); solutions agree to rounding (the forward step is folded into the
elimination, which sums a row's update vector in another order).  Solutions
are also held to a backward-error bar per system and against the oracle."""
import ctypes as C

import numpy as np
import pytest

from conftest import backward_error, gpu_available, make_system

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs an MI355X")]


def _say(msg):
    print(msg, flush=True)


def _frames(nsys, seed, lo=80, hi=300):
    from sprgpu import fem, host

    rng = np.random.default_rng(seed)
    models = [fem.beam_frame(int(n), rng) for n in rng.integers(lo, hi, nsys)]
    systems = [host.symbolic(m) for m in models]
    sms = [host.assemble(m, s) for m, s in zip(models, systems)]
    merged = host.merge_systems(systems)
    return systems, sms, merged


def _systems():
    """(name, mspar, mtrees, msifa, sm) of structures with fronts F <= 64."""
    out = []
    systems, sms, merged = _frames(16, 11)
    out.append(("frames16", merged.mspar, merged.mtrees, merged.msifa, merged.merge_sm(sms)))
    for kind, n in (("hex8", 3), ("quad4", 10)):
        _, s, sm = make_system(kind, n)
        out.append((f"{kind}_{n}", s.mspar, s.mtrees, s.msifa, sm))
    return out


def _both(mspar, mtrees, msifa, precision):
    from sprgpu.device import SprFactor

    hf = SprFactor(mspar, mtrees, msifa, precision)
    hl = SprFactor(mspar, mtrees, msifa, precision)
    hl.set_plan_params(forest=0)
    assert hf.plan_params()["forest"] == 1, "structure not on the forest schedule"
    assert hl.plan_params()["forest"] == 0
    return hf, hl


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_forest_factor_equals_level_schedule(precision):
    import torch

    dev = torch.device("cuda:0")
    for name, mspar, mtrees, msifa, sm in _systems():
        _say(f"forest factor {precision}: {name}, neq {int(mspar[7])}, nsuper {int(mspar[10])}")
        hf, hl = _both(mspar, mtrees, msifa, precision)
        try:
            sm_dev = torch.from_numpy(sm).to(dev)
            fi_f = hf.factor(sm_dev)
            fi_l = hl.factor(sm_dev)
            assert fi_f == fi_l, (fi_f, fi_l)
            assert fi_f["ierr"] == 0
            lf, ll = hf.get_factor(), hl.get_factor()
            assert np.array_equal(lf, ll), name
            b = np.random.default_rng(5).uniform(-1, 1, sm.size - int(mspar[15]))
            b_dev = torch.from_numpy(b).to(dev).view(1, -1)
            xf, xl = b_dev.clone(), b_dev.clone()
            hf.solve(xf)
            hl.solve(xl)
            rel = float(torch.linalg.norm(xf - xl) / torch.linalg.norm(xl))
            _say(f"  solution difference forest vs level {rel:.2e}")
            assert rel < (1e-12 if precision == "fp64" else 1e-5), rel
            # forward, then backward as two calls: the same as one call
            x2 = b_dev.clone()
            hf.solve(x2, which=1)
            hf.solve(x2, which=2)
            assert torch.equal(x2, xf)
        finally:
            hf.close()
            hl.close()


def test_forest_solutions_per_frame_backward_error():
    """FP64 factor and solve of a forest of frames in one launch (the graph
    path), each frame's solution held to a backward-error bar and compared
    with the oracle's."""
    import oracle
    import torch
    from sprgpu import _lib
    from sprgpu.device import SprFactor

    systems, sms, merged = _frames(24, 13, 80, 900)
    sm = merged.merge_sm(sms)
    b = np.random.default_rng(6).uniform(-1, 1, merged.neq)
    dev = torch.device("cuda:0")
    sm_dev = torch.from_numpy(sm).to(dev)
    x_dev = torch.from_numpy(b).to(dev).view(1, -1).clone()
    h = SprFactor(merged.mspar, merged.mtrees, merged.msifa, "fp64")
    try:
        assert h.plan_params()["forest"] == 1
        _lib.check(_lib.lib().sprgpu_factor_solve_graph(h._h, C.c_void_p(sm_dev.data_ptr()),
                                                        C.c_void_p(x_dev.data_ptr()),
                                                        merged.neq, 1), "factor_solve_graph")
        torch.cuda.synchronize()
        assert h.factor_info()["ierr"] == 0
    finally:
        h.close()
    x = x_dev.view(-1).cpu().numpy()
    worst = 0.0
    for k, (s, smk) in enumerate(zip(systems, sms)):
        bk = b[merged.neq_off[k]:merged.neq_off[k + 1]].copy()
        xo = np.asfortranarray(bk[:, None].copy())
        assert oracle.sprsol(3, s.mspar.copy(), s.mtrees, s.msifa, smk.copy(), xo)[0] == 0
        be_g = backward_error(s, smk, merged.split(x)[k], bk)
        be_o = backward_error(s, smk, xo[:, 0], bk)
        worst = max(worst, be_g)
        assert be_g <= 1e-14, (k, be_g, be_o)
    _say(f"24 frames: worst backward error {worst:.2e}")


@pytest.mark.parametrize("fixture", ["hex8_2_zero_diag", "hex8_2_singular", "hex8_3_indefinite"])
def test_forest_failures_and_replacements_equal_level_schedule(fixture):
    """Error paths and SPRFC3 replacements on the forest schedule: IERR,
    NUMNEG, SUPEQT, TOL(2:3), d and DCOR are the level schedule's bits."""
    import os

    import torch

    path = os.path.join(os.path.dirname(__file__), "golden", fixture + ".npz")
    g = np.load(path)
    dev = torch.device("cuda:0")
    sm_dev = torch.from_numpy(np.ascontiguousarray(g["sm_in"])).to(dev)
    for mode, tol in ((1, (1e-12, 0, 0, 0)), (3, (0.6, 0, 0, 1.0))):
        hf, hl = _both(g["mspar_in"], g["mtrees"], g["msifa"], "fp64")
        try:
            fi_f = hf.factor(sm_dev, tol=tol, mode=mode)
            fi_l = hl.factor(sm_dev, tol=tol, mode=mode)
            _say(f"{fixture} mode {mode}: {fi_f}")
            assert fi_f == fi_l, (fi_f, fi_l)
            if mode == 3:
                assert np.array_equal(hf.dcor(), hl.dcor())
            assert np.array_equal(hf.get_factor(), hl.get_factor())
        finally:
            hf.close()
            hl.close()
