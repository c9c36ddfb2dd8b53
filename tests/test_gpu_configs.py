"""The five benchmark configurations at their stated sizes, on the GPU.

C1  hex8 20^3, E = 1, nu = 0.3, z = 0 clamped, one U[-1,1] RHS (seed 0):
    drop-in SPRSOL against the oracle (the C restatement of SPRSOL,
    sprsol.f:5-52): IERR, MSPAR, TOL(2:3), the factor in SM and the solution.
C2  hex8 40^3, E log-normal(0, 0.5) seed 1, one U[-1,1] RHS: the FP64 GPU
    factor + solve and the FP32 factor + FP64 refinement, each within 1e-8 of
    the reference's own FP64 solution (tests/golden/hex8_40_reference_solution.npy,
    written by the reference SPRSOL built from source) with relative residual
    at most 1e-10.
C3  the C2 matrix with 128 N(0,1) right-hand sides (seed 2), and quad4 plane
    stress 1000^2 (E = 1, nu = 0.25, x = 0 clamped) with 128 N(0,1)
    right-hand sides: FP32 + refinement of all columns in one batched call,
    every column's residual at most 1e-10 (FP64 SpMV), every column within
    1e-8 of the FP64 solve, and sampled columns of the batched plain solves
    within 1e-11 (relative) of their single-column solves (the batched
    rectangle products are MFMA GEMMs summed in another order).
C4  512 beam frames (nodes U[80,900], seed 3) merged into one forest, FP64
    factor + solve as one captured graph, after the autotuner: per frame the
    normwise backward error ||b - A x|| / (||A|| ||x|| + ||b||) (infinity
    norms, A x and |A| row sums from the reference's own SPRPRM) at most
    1e-14, which bounds the error of an ill-conditioned frame by what its
    condition number allows; the oracle's own backward error on the same
    frame is reported beside it.
C5  compliance adjoint at 40^3 (E seed 4): dL/dE_e of all 64,000 elements
    through the torch.autograd binding, against the FP64 CPU adjoint
    (tests/golden/hex8_40_adjoint_reference.npz, from the reference SPRSOL)
    to 1e-6 and against central differences (h = 1e-6 E_e) on 64 random
    elements to 1e-6.
"""
import ctypes as C
import os
import sys

import numpy as np
import pytest

from conftest import ROOT, backward_error, gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs an MI355X")]

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _say(msg):
    print(msg, flush=True)


def _hex8_40_problem():
    if ROOT not in sys.path:
        sys.path.insert(0, ROOT)
    from bench import build_problem

    return build_problem(40)


# --------------------------------------------------------------------------
# C1
# --------------------------------------------------------------------------
def test_c1_hex8_20_dropin_matches_oracle():
    import oracle
    from sprgpu import fem, host, spr

    model = fem.hex8_grid(20)
    sysm = host.symbolic(model)
    sm = host.assemble(model, sysm)
    b = host.loads_to_equations(sysm, fem.uniform_loads(model.ndof, np.random.default_rng(0)))
    _say(f"C1 hex8 20^3: neq={sysm.neq} nsuper={sysm.nsuper} nzerol={sysm.nzerol}")
    assert sysm.neq == 26460
    sm_o, x_o, ms_o = sm.copy(), np.asfortranarray(b.copy()), sysm.mspar.copy()
    ierr_o, tol_o = oracle.sprsol(3, ms_o, sysm.mtrees, sysm.msifa, sm_o, x_o)
    sm_g, x_g, ms_g = sm.copy(), np.asfortranarray(b.copy()), sysm.mspar.copy()
    tol_g = np.zeros(3)
    ierr_g = spr.sprsol(3, ms_g, sysm.mtrees, sysm.msifa, sm_g, x_g, tol_g)
    spr.release(sm_g)
    assert ierr_g == ierr_o == 0
    assert np.array_equal(ms_g, ms_o)
    np.testing.assert_allclose(tol_g[1:], tol_o[1:], rtol=1e-9)
    rel = np.linalg.norm(x_g - x_o) / np.linalg.norm(x_o)
    d_rel = np.max(np.abs(sm_g[:sysm.neq] / sm_o[:sysm.neq] - 1))
    l_rel = (np.linalg.norm(sm_g[sysm.neq:] - sm_o[sysm.neq:]) /
             np.linalg.norm(sm_o[sysm.neq:]))
    _say(f"C1: solution {rel:.2e}, pivots {d_rel:.2e}, L {l_rel:.2e} against the oracle")
    assert rel < 1e-9 and d_rel < 1e-9 and l_rel < 1e-9


# --------------------------------------------------------------------------
# C2
# --------------------------------------------------------------------------
def test_c2_hex8_40_fp64_and_fp32_ir_match_reference():
    import torch
    from sprgpu.device import SprFactor

    model, sysm, sm, b, _ = _hex8_40_problem()
    assert sysm.neq == 201720
    xr = np.load(os.path.join(GOLDEN, "hex8_40_reference_solution.npy"), allow_pickle=False)
    sm_dev = torch.from_numpy(sm).cuda()
    b_dev = torch.from_numpy(b).cuda().view(1, -1)
    out = {}
    for prec in ("fp64", "fp32"):
        h = SprFactor(sysm.mspar, sysm.mtrees, sysm.msifa, prec)
        try:
            h.set_pattern(sm)
            info = h.factor(sm_dev)
            assert info["ierr"] == 0 and info["numneg"] == 0, info
            x = torch.empty_like(b_dev)
            if prec == "fp64":
                x.copy_(b_dev)
                h.solve(x)
                it = 0
            else:
                ir = h.solve_ir(b_dev, x, rtol=1e-10, maxit=30)
                assert ir["converged"], ir
                it = ir["iterations"]
            r = torch.empty_like(x)
            h.spmv(x, r)
            relres = float(torch.linalg.norm(b_dev - r) / torch.linalg.norm(b_dev))
            xx = x.view(-1).cpu().numpy()
            err = float(np.linalg.norm(xx - xr) / np.linalg.norm(xr))
            out[prec] = (relres, err, it)
            _say(f"C2 {prec}: relres {relres:.2e}, error vs reference {err:.2e}, "
                 f"refinement steps {it}")
        finally:
            h.close()
    for prec, (relres, err, _) in out.items():
        assert relres <= 1e-10, (prec, relres)
        assert err <= 1e-8, (prec, err)


# --------------------------------------------------------------------------
# C3
# --------------------------------------------------------------------------
def _many_rhs_case(kind):
    from sprgpu import fem, host

    if kind == "hex8_40":
        model, sysm, sm, _, _ = _hex8_40_problem()
    else:
        model = fem.quad4_grid(1000)
        sysm = host.symbolic(model)
        sm = host.assemble(model, sysm)
    B = host.loads_to_equations(sysm, fem.normal_loads(model.ndof, np.random.default_rng(2), 128))
    return sysm, sm, np.ascontiguousarray(B.T)


@pytest.mark.parametrize("kind", ["hex8_40", "quad4_1000"])
def test_c3_128_rhs_batched(kind):
    import torch
    from sprgpu.device import SprFactor

    sysm, sm, Bt = _many_rhs_case(kind)
    _say(f"C3 {kind}: neq={sysm.neq} nzerol={sysm.nzerol}, 128 right-hand sides")
    sm_dev = torch.from_numpy(sm).cuda()
    b_dev = torch.from_numpy(Bt).cuda()  # (nrhs, neq) = B(LDB, NRHS)
    bn = b_dev.norm(dim=1)
    sample = (0, 37, 127)
    # FP64: factor + one batched solve; sampled columns agree with their own
    # solve (the batched rectangle products are MFMA GEMMs summed in another
    # order: to rounding, at most 1e-11 relative)
    h = SprFactor(sysm.mspar, sysm.mtrees, sysm.msifa, "fp64")
    try:
        assert h.factor(sm_dev)["ierr"] == 0
        x64 = b_dev.clone()
        h.solve(x64)
        for k in sample:
            xk = b_dev[k:k + 1].clone()
            h.solve(xk)
            d = ((xk[0] - x64[k]).norm() / xk[0].norm()).item()
            assert d <= 1e-11, ("fp64", k, d)
        h.set_pattern(sm)
        r = torch.empty_like(x64)
        h.spmv(x64, r)
        rr64 = ((b_dev - r).norm(dim=1) / bn).max().item()
    finally:
        h.close()
    # FP32 factor + FP64 refinement of all 128 columns in one call
    h = SprFactor(sysm.mspar, sysm.mtrees, sysm.msifa, "fp32")
    try:
        h.set_pattern(sm)
        assert h.factor(sm_dev)["ierr"] == 0
        y = b_dev.clone()
        h.solve(y)
        for k in sample:
            yk = b_dev[k:k + 1].clone()
            h.solve(yk)
            d = ((yk[0] - y[k]).norm() / yk[0].norm()).item()
            assert d <= 1e-11, ("fp32", k, d)
        x32 = torch.empty_like(b_dev)
        ir = h.solve_ir(b_dev, x32, rtol=1e-10, maxit=40)
        r = torch.empty_like(x32)
        h.spmv(x32, r)
        rr = ((b_dev - r).norm(dim=1) / bn)
        diff = ((x32 - x64).norm(dim=1) / x64.norm(dim=1)).max().item()
    finally:
        h.close()
    _say(f"C3 {kind}: fp64 max relres {rr64:.2e}; fp32+IR {ir['iterations']} steps, "
         f"max relres {rr.max().item():.2e}, max column difference to fp64 {diff:.2e}")
    assert ir["converged"]
    assert rr.max().item() <= 1e-10
    assert rr64 <= 1e-10
    assert diff <= 1e-8


# --------------------------------------------------------------------------
# C4
# --------------------------------------------------------------------------
def _frames():
    from sprgpu import fem, host

    rng = np.random.default_rng(3)
    nodes = rng.integers(80, 901, 512)
    models = [fem.beam_frame(int(n), rng) for n in nodes]
    systems = [host.symbolic(m) for m in models]
    sms = [host.assemble(m, s) for m, s in zip(models, systems)]
    bs = [host.loads_to_equations(s, fem.uniform_loads(m.ndof, rng))[:, 0].copy()
          for m, s in zip(models, systems)]
    return systems, sms, bs


def test_c4_512_frames_one_graph():
    import oracle
    import torch
    from sprgpu import _lib, host
    from sprgpu.device import SprFactor

    systems, sms, bs = _frames()
    merged = host.merge_systems(systems)
    sm = merged.merge_sm(sms)
    b = np.concatenate(bs)
    _say(f"C4: 512 frames, NEQ {min(s.neq for s in systems)}..{max(s.neq for s in systems)}, "
         f"total {merged.neq}")
    dev = torch.device("cuda:0")
    sm_dev = torch.from_numpy(sm).to(dev)
    b_dev = torch.from_numpy(b).to(dev).view(1, -1)
    x_dev = b_dev.clone()
    h = SprFactor(merged.mspar, merged.mtrees, merged.msifa, "fp64")
    L = _lib.lib()
    try:
        report = np.zeros(256)
        _lib.check(L.sprgpu_autotune(h._h, C.c_void_p(sm_dev.data_ptr()),
                                     C.c_void_p(x_dev.data_ptr()), merged.neq, 1,
                                     report.ctypes.data_as(C.c_void_p)), "autotune")
        xs = []
        for rep in range(2):
            x_dev.copy_(b_dev)
            _lib.check(L.sprgpu_factor_solve_graph(h._h, C.c_void_p(sm_dev.data_ptr()),
                                                   C.c_void_p(x_dev.data_ptr()), merged.neq, 1),
                       "factor_solve_graph")
            torch.cuda.synchronize()
            xs.append(x_dev.view(-1).cpu().numpy().copy())
        assert h.factor_info()["ierr"] == 0
    finally:
        h.close()
    assert np.array_equal(xs[0], xs[1]), "graph replays differ"
    worst_be, worst_ratio = 0.0, 0.0
    for k, (s, smk, bk) in enumerate(zip(systems, sms, bs)):
        xg = merged.split(xs[1])[k]
        xo = np.asfortranarray(bk[:, None].copy())
        ierr, _ = oracle.sprsol(3, s.mspar.copy(), s.mtrees, s.msifa, smk.copy(), xo)
        assert ierr == 0
        be_g = backward_error(s, smk, xg, bk)
        be_o = backward_error(s, smk, xo[:, 0], bk)
        worst_be = max(worst_be, be_g)
        worst_ratio = max(worst_ratio, be_g / max(be_o, 1e-17))
        assert be_g <= 1e-14, (k, be_g, be_o)
    _say(f"C4: worst backward error {worst_be:.2e} (the oracle's own: ratio at most "
         f"{worst_ratio:.1f})")


# --------------------------------------------------------------------------
# C5
# --------------------------------------------------------------------------
def test_c5_hex8_40_adjoint():
    import torch
    from sprgpu import fem
    from sprgpu.adjoint import ElasticSystem, spr_solve

    n, seed = 40, 4
    model = fem.hex8_grid(n, e_sigma=0.5, e_seed=seed)
    rng = np.random.default_rng(seed)
    rng.normal(0.0, 0.5, model.nel)
    f = rng.uniform(-1.0, 1.0, model.ndof)
    sysobj = ElasticSystem(model, precision="fp32", rtol=1e-10, maxit=30)
    try:
        E = model.escale.copy()
        ft = torch.from_numpy(f).cuda()
        p = torch.from_numpy(E.copy()).cuda().requires_grad_(True)
        u = spr_solve(p, ft, sysobj)
        L = torch.dot(ft, u)
        L.backward()
        g = p.grad.cpu().numpy()
        ref = np.load(os.path.join(GOLDEN, "hex8_40_adjoint_reference.npz"), allow_pickle=False)
        g_rel = float(np.abs(g - ref["grad"]).max() / np.abs(ref["grad"]).max())
        l_rel = abs(float(L) - float(ref["compliance"])) / abs(float(ref["compliance"]))
        f_eq = sysobj.to_eq(ft)
        sysobj.factor(E)
        sel = np.random.default_rng(seed).choice(model.nel, 64, replace=False)
        fd_rel = []
        for e in sel:
            fd = sysobj.central_difference(E, f_eq, int(e), h_rel=1e-6)
            fd_rel.append(abs(fd - g[e]) / abs(g[e]))
        _say(f"C5: gradient vs FP64 CPU adjoint {g_rel:.2e}, compliance {l_rel:.2e}, "
             f"central differences (64 elements) max {max(fd_rel):.2e}")
        assert g.shape == (64000,)
        assert g_rel <= 1e-6 and l_rel <= 1e-9
        assert max(fd_rel) <= 1e-6
    finally:
        sysobj.close()
