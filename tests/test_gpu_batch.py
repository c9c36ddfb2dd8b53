"""Batches of systems as one forest, the captured factor+solve graph and the autotuner."""
import ctypes as C

import numpy as np
import pytest

from conftest import backward_error, gpu_available, make_system

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs an MI355X")]


def _graph_run(h, sm_dev, x_dev, neq):
    from sprgpu import _lib

    _lib.check(_lib.lib().sprgpu_factor_solve_graph(h._h, C.c_void_p(sm_dev.data_ptr()),
                                                    C.c_void_p(x_dev.data_ptr()), neq, 1),
               "factor_solve_graph")


def _check_replays(mspar, mtrees, msifa, sm, b, precision="fp64"):
    """Every graph replay equals the eager factor+solve bit for bit, with
    unrelated eager work and synchronisations between the replays."""
    import torch
    from sprgpu.device import SprFactor

    dev = torch.device("cuda:0")
    sm_dev = torch.from_numpy(sm).to(dev)
    b_dev = torch.from_numpy(b).to(dev).view(1, -1)
    h = SprFactor(mspar, mtrees, msifa, precision)
    try:
        x_eager = b_dev.clone()
        h.factor(sm_dev)
        h.solve(x_eager)
        torch.cuda.synchronize()
        x_graph = torch.empty_like(b_dev)
        other = torch.from_numpy(np.random.default_rng(1).normal(size=b.size)).to(dev).view(1, -1)
        for rep in range(4):
            x_graph.copy_(b_dev)
            _graph_run(h, sm_dev, x_graph, b.size)
            torch.cuda.synchronize()
            assert torch.equal(x_graph, x_eager), f"replay {rep}"
            y = other.clone()  # eager work between replays
            h.factor(sm_dev)
            h.solve(y)
            torch.cuda.synchronize()
        assert h.factor_info()["ierr"] == 0
        return x_eager.view(-1).cpu().numpy()
    finally:
        h.close()


def test_graph_replays_match_eager_on_a_frame_forest():
    import oracle
    from sprgpu import fem, host

    rng = np.random.default_rng(3)
    models = [fem.beam_frame(int(n), rng) for n in rng.integers(80, 300, 12)]
    systems = [host.symbolic(m) for m in models]
    sms = [host.assemble(m, s) for m, s in zip(models, systems)]
    merged = host.merge_systems(systems)
    b = rng.uniform(-1, 1, merged.neq)
    x = _check_replays(merged.mspar, merged.mtrees, merged.msifa, merged.merge_sm(sms), b)
    # per frame: a backward-error bar (the frames are ill-conditioned, so the
    # forward difference to the oracle is bounded by kappa times it, not by a
    # fixed number), and the oracle's own solution for comparison
    for k, (s, sm) in enumerate(zip(systems, sms)):
        bk = b[merged.neq_off[k]:merged.neq_off[k + 1]].copy()
        bi = np.asfortranarray(bk[:, None].copy())
        assert oracle.sprsol(3, s.mspar.copy(), s.mtrees, s.msifa, sm.copy(), bi,
                             tol=[0, 0, 0])[0] == 0
        xk = merged.split(x)[k]
        be_g = backward_error(s, sm, xk, bk)
        be_o = backward_error(s, sm, bi[:, 0], bk)
        print(f"frame {k}: neq {s.neq} backward error {be_g:.2e} (oracle {be_o:.2e}), "
              f"difference to the oracle {np.linalg.norm(xk - bi[:, 0]) / np.linalg.norm(bi[:, 0]):.2e}",
              flush=True)
        assert be_g <= 1e-14, (k, be_g, be_o)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_graph_replays_match_eager_with_wide_supernodes(precision):
    """hex8 16^3 has supernodes of over 1000 columns: the dataflow solves'
    ready flags must see a new epoch on every replay."""
    _, sysm, sm = make_system("hex8", 16)
    b = np.random.default_rng(2).uniform(-1, 1, sysm.neq)
    _check_replays(sysm.mspar, sysm.mtrees, sysm.msifa, sm, b, precision)


def test_autotune_keeps_results():
    import torch
    from sprgpu import _lib
    from sprgpu.device import SprFactor

    _, sysm, sm = make_system("hex8", 8)
    dev = torch.device("cuda:0")
    sm_dev = torch.from_numpy(sm).to(dev)
    b = torch.from_numpy(np.random.default_rng(4).uniform(-1, 1, sysm.neq)).to(dev).view(1, -1)
    h = SprFactor(sysm.mspar, sysm.mtrees, sysm.msifa, "fp64")
    try:
        x0 = b.clone()
        h.factor(sm_dev)
        h.solve(x0)
        report = np.zeros(256)
        scratch = b.clone()
        _lib.check(_lib.lib().sprgpu_autotune(h._h, C.c_void_p(sm_dev.data_ptr()),
                                              C.c_void_p(scratch.data_ptr()), sysm.neq, 1,
                                              report.ctypes.data_as(C.c_void_p)), "autotune")
        n = int(report[0])
        assert n >= 2 and np.all(report[6:6 * n + 1:6] > 0)
        params = (C.c_int * 6)()
        _lib.lib().sprgpu_plan_params(h._h, params)
        assert params[0] in (32, 48, 64) and params[1] in (1, 2, 4)
        x1 = b.clone()
        h.factor(sm_dev)
        h.solve(x1)
        rel = float(torch.linalg.norm(x1 - x0) / torch.linalg.norm(x0))
        assert rel < 1e-12, rel
    finally:
        h.close()


def test_graph_recaptured_after_the_solve_buffers_grow():
    """A graph captured with one right-hand side must not replay after an
    eager solve with more right-hand sides reallocated the solve workspace
    (its captured kernels would address freed memory): the next call
    captures again, and its result equals the eager factor + solve."""
    import torch
    from sprgpu.device import SprFactor

    _, sysm, sm = make_system("hex8", 12, e_sigma=0.5, e_seed=1)
    dev = torch.device("cuda:0")
    sm_dev = torch.from_numpy(sm).to(dev)
    b = torch.from_numpy(np.random.default_rng(5).uniform(-1, 1, sysm.neq)).to(dev).view(1, -1)
    h = SprFactor(sysm.mspar, sysm.mtrees, sysm.msifa, "fp64")
    try:
        x_graph = b.clone()
        _graph_run(h, sm_dev, x_graph, sysm.neq)      # captured with 1 RHS
        torch.cuda.synchronize()
        B8 = torch.randn(8, sysm.neq, dtype=torch.float64, device=dev)
        h.factor(sm_dev)
        h.solve(B8)                                   # grows the solve workspace
        torch.cuda.synchronize()
        x_eager = b.clone()
        h.factor(sm_dev)
        h.solve(x_eager)
        torch.cuda.synchronize()
        for rep in range(2):
            x_graph.copy_(b)
            _graph_run(h, sm_dev, x_graph, sysm.neq)
            torch.cuda.synchronize()
            assert torch.equal(x_graph, x_eager), f"replay {rep}"
    finally:
        h.close()
