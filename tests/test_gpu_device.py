"""GPU tests of the device-resident handle API against the oracle.

Covers the blocked factor path (fronts wider than 64), the blocked solves
(supernodes wider than 128), multiple right-hand sides, the FP32 factor with
FP64 iterative refinement, SPRFC3 pivot replacement and the error paths.
Tolerances: FP64 solutions within 1e-9 relative of the oracle's; FP32+IR
residuals at or below the requested 1e-10.
"""
import os

import numpy as np
import pytest

import oracle
from conftest import make_system

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _torch():
    import torch

    return torch


def _oracle(sysm, sm, b, tol=(0.0, 0.0, 0.0)):
    sm_o, x_o, ms = sm.copy(), np.asfortranarray(b.copy()), sysm.mspar.copy()
    ierr, tolv = oracle.sprsol(3, ms, sysm.mtrees, sysm.msifa, sm_o, x_o, tol=list(tol))
    return ierr, tolv, sm_o, x_o, ms


def _handle(sysm, sm, precision):
    torch = _torch()
    from sprgpu.device import SprFactor

    h = SprFactor(sysm.mspar, sysm.mtrees, sysm.msifa, precision)
    sm_dev = torch.from_numpy(sm).cuda()
    return h, sm_dev


@pytest.mark.parametrize("n", [10, 16])
def test_fp64_factor_and_blocked_solve(n):
    torch = _torch()
    model, sysm, sm = make_system("hex8", n, e_sigma=0.5, e_seed=1)
    xs, xl = sysm.xsuper(), sysm.xlindx()
    print(f"hex8 {n}: neq={sysm.neq} maxS={np.diff(xs).max()} maxF={np.diff(xl).max()}", flush=True)
    b = np.random.default_rng(0).uniform(-1, 1, (sysm.neq, 4))
    ierr_o, tol_o, sm_o, x_o, ms_o = _oracle(sysm, sm, b)
    h, sm_dev = _handle(sysm, sm, "fp64")
    info = h.factor(sm_dev)
    assert info["ierr"] == ierr_o == 0
    assert info["numneg"] == ms_o[25]
    np.testing.assert_allclose([info["tol2"], info["tol3"]], tol_o[1:], rtol=1e-9)
    x = torch.from_numpy(np.ascontiguousarray(b.T)).cuda()
    h.solve(x)
    xg = x.cpu().numpy().T
    rel = np.linalg.norm(xg - x_o) / np.linalg.norm(x_o)
    assert rel < 1e-9, rel
    # forward and backward separately give the same
    x2 = torch.from_numpy(np.ascontiguousarray(b.T)).cuda()
    h.solve(x2, which=1)
    h.solve(x2, which=2)
    assert torch.equal(x2, x)
    # factor in SM layout agrees with the oracle's
    f = h.get_factor()
    assert np.max(np.abs(f[:sysm.neq] / sm_o[:sysm.neq] - 1)) < 1e-9
    assert np.linalg.norm(f[sysm.neq:] - sm_o[sysm.neq:]) / np.linalg.norm(sm_o[sysm.neq:]) < 1e-9
    h.close()


def test_fp32_factor_with_refinement():
    torch = _torch()
    model, sysm, sm = make_system("hex8", 12, e_sigma=0.5, e_seed=1)
    b = np.random.default_rng(2).normal(size=(sysm.neq, 3))
    _, _, _, x_o, _ = _oracle(sysm, sm, b)
    h, sm_dev = _handle(sysm, sm, "fp32")
    h.set_pattern(sm)
    info = h.factor(sm_dev)
    assert info["ierr"] == 0
    bd = torch.from_numpy(np.ascontiguousarray(b.T)).cuda()
    x = torch.empty_like(bd)
    ir = h.solve_ir(bd, x, rtol=1e-10, maxit=30)
    print("IR:", ir, flush=True)
    assert ir["converged"] and ir["relres"] <= 1e-10
    xg = x.cpu().numpy().T
    assert np.linalg.norm(xg - x_o) / np.linalg.norm(x_o) < 1e-8
    # the FP64 SpMV of the handle is A x
    y = torch.empty_like(x)
    h.spmv(torch.from_numpy(np.ascontiguousarray(x_o.T)).cuda(), y)
    assert np.linalg.norm(y.cpu().numpy().T - b) / np.linalg.norm(b) < 1e-12
    h.close()


def test_many_rhs_match_single():
    torch = _torch()
    model, sysm, sm = make_system("quad4", 40)
    h, sm_dev = _handle(sysm, sm, "fp64")
    assert h.factor(sm_dev)["ierr"] == 0
    B = torch.randn(16, sysm.neq, dtype=torch.float64, device="cuda")
    X = B.clone()
    h.solve(X)
    for k in (0, 7, 15):
        xk = B[k:k + 1].clone()
        h.solve(xk)
        assert torch.equal(xk[0], X[k])
    h.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_grouped_rhs_match_single_with_wide_supernodes(precision):
    """Right-hand sides in groups on hex8 12^3, whose top supernodes take the
    wide dataflow path: 37 of them (rectangle launches: a group of 32 and a
    partial group of 5; block chains and narrow supernodes: four groups of 8
    and a partial group of 5), 11 (groups of 8 throughout, the forward
    rectangles in their own launch) and single columns (the forward
    rectangles fused with the chains).  Every column equals its own
    single-RHS solve bit for bit."""
    torch = _torch()
    model, sysm, sm = make_system("hex8", 12, e_sigma=0.5, e_seed=1)
    assert np.diff(sysm.xsuper()).max() > 128
    h, sm_dev = _handle(sysm, sm, precision)
    assert h.factor(sm_dev)["ierr"] == 0
    B = torch.randn(37, sysm.neq, dtype=torch.float64, device="cuda")
    X = B.clone()
    h.solve(X)
    for k in range(37):
        xk = B[k:k + 1].clone()
        h.solve(xk)
        assert torch.equal(xk[0], X[k]), k
    X11 = B[:11].clone()
    h.solve(X11)
    assert torch.equal(X11, X[:11])
    h.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_mfma_rectangles_with_many_rhs(precision):
    """From 64 right-hand sides on, the wide supernodes' rectangle products
    (forward L21 x1, backward L21^T x2) run as FP64 MFMA GEMMs over groups of
    64: 70 columns on hex8 12^3 (a full group and a partial group of 6).  The
    products are summed in another order than the per-column kernels, so a
    column agrees with its own single-RHS solve to rounding (relative
    difference at most 1e-11, measured about 1e-15), and every column
    solves A x = b (relative residual at most 1e-12 in FP64; the FP32 factor's
    columns to the factor's own accuracy, 1e-5)."""
    torch = _torch()
    model, sysm, sm = make_system("hex8", 12, e_sigma=0.5, e_seed=1)
    assert np.diff(sysm.xsuper()).max() > 128
    h, sm_dev = _handle(sysm, sm, precision)
    assert h.factor(sm_dev)["ierr"] == 0
    h.set_pattern(sm)
    B = torch.randn(70, sysm.neq, dtype=torch.float64, device="cuda")
    X = B.clone()
    h.solve(X)
    worst = 0.0
    for k in (0, 1, 17, 63, 64, 69):
        xk = B[k:k + 1].clone()
        h.solve(xk)
        worst = max(worst, ((xk[0] - X[k]).norm() / xk[0].norm()).item())
    r = torch.empty_like(X)
    h.spmv(X, r)
    relres = ((B - r).norm(dim=1) / B.norm(dim=1)).max().item()
    print(f"{precision}: worst column difference {worst:.2e}, max relres {relres:.2e}", flush=True)
    h.close()
    assert worst <= 1e-11
    assert relres <= (1e-12 if precision == "fp64" else 1e-5)


def _golden(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_error_paths_match_reference():
    torch = _torch()
    from sprgpu.device import SprFactor

    for name, want_ierr in (("hex8_2_zero_diag", -3), ("hex8_2_singular", -3)):
        g = _golden(name)
        h = SprFactor(g["mspar_in"], g["mtrees"], g["msifa"], "fp64")
        info = h.factor(torch.from_numpy(g["sm_in"]).cuda(), tol=tuple(g["tol_in"]))
        assert info["ierr"] == want_ierr == int(g["ierr"])
        assert info["supeqt"] == g["mspar_out"][26], name
        assert info["numneg"] == g["mspar_out"][25], name
        h.close()


def test_sprfc3_replacement_matches_reference():
    torch = _torch()
    from sprgpu.device import SprFactor

    g = _golden("hex8_2_sprfc3")
    h = SprFactor(g["mspar_in"], g["mtrees"], g["msifa"], "fp64")
    info = h.factor(torch.from_numpy(g["sm_in"]).cuda(), tol=tuple(g["tol_in"]), mode=3)
    assert info["ierr"] == -3
    assert info["supeqt"] == g["mspar_out"][26]      # -number of replaced pivots
    assert info["numneg"] == g["mspar_out"][25]
    h.close()


@pytest.mark.parametrize("mode", [1, 3])
@pytest.mark.parametrize("block", [1, 2, 0], ids=["outer_blocks_fused", "outer_blocks",
                                                   "panel_steps"])
def test_pivot_rules_inside_blocked_fronts(block, mode):
    """TOL(1) = 0.6 puts about 50 pivots of the wide fronts (the root has
    S = 702: six 128-column outer blocks) under the threshold d/a_jj:
    SPRFC1 stops at the first one (IERR -3, MSPAR(27) = its equation),
    SPRFC3 replaces every one of them; every schedule (outer blocks with the
    rows below in the same launch or in their own, panel steps) must match
    the oracle."""
    torch = _torch()
    from sprgpu.device import SprFactor

    model, sysm, sm = make_system("hex8", 10, e_sigma=0.5, e_seed=1)
    xs = sysm.xsuper()
    assert np.diff(xs).max() > 256
    tol = [0.6, 0.0, 0.0, 1.0]
    ms_o, sm_o = sysm.mspar.copy(), sm.copy()
    if mode == 1:
        ierr_o, tol_o = oracle.sprsol(1, ms_o, sysm.mtrees, sysm.msifa, sm_o, tol=tol[:3])
    else:
        ierr_o, tol_o, dcor_o = oracle.sprfc3(ms_o, sysm.mtrees, sysm.msifa, sm_o, tol)
    h = SprFactor(sysm.mspar, sysm.mtrees, sysm.msifa, "fp64")
    h.set_plan_params(block_factor=block)
    assert h.plan_params()["block_factor"] == block
    info = h.factor(torch.from_numpy(sm).cuda(), tol=tuple(tol), mode=mode)
    print(f"mode {mode} block {block}: gpu {info} oracle ierr {ierr_o} "
          f"mspar26/27 {ms_o[25]} {ms_o[26]}", flush=True)
    assert info["ierr"] == ierr_o == -3
    assert info["supeqt"] == ms_o[26]
    assert info["numneg"] == ms_o[25]
    if mode == 3:
        assert ms_o[26] < -20  # many replacements
        f = h.get_factor()
        assert np.max(np.abs(f[:sysm.neq] / sm_o[:sysm.neq] - 1)) < 1e-9
        assert np.linalg.norm(f[sysm.neq:] - sm_o[sysm.neq:]) / np.linalg.norm(sm_o[sysm.neq:]) < 1e-9
        # solves with the replaced-pivot factor (the wide blocks through their
        # explicit inverses) against the oracle's solve with its own factor
        b = np.random.default_rng(7).uniform(-1, 1, (sysm.neq, 2))
        x_o = np.asfortranarray(b.copy())
        ierr_s, _ = oracle.sprsol(4, ms_o, sysm.mtrees, sysm.msifa, sm_o, x_o)  # FS + BS
        assert ierr_s == 0
        x = torch.from_numpy(np.ascontiguousarray(b.T)).cuda()
        h.solve(x)
        rel = np.linalg.norm(x.cpu().numpy().T - x_o) / np.linalg.norm(x_o)
        print(f"solve with replaced pivots: rel {rel:.2e}", flush=True)
        assert rel < 1e-8, rel
    h.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_outer_blocks_match_panel_steps(precision):
    """The blocked-front schedules compute the same factor: panel steps and
    outer blocks to rounding (the rank-32 updates of a row are summed in a
    different order), the outer blocks in one launch or two bit for bit (the
    same operations; only where the rows below read their operands
    differs)."""
    torch = _torch()
    from sprgpu.device import SprFactor

    model, sysm, sm = make_system("hex8", 12, e_sigma=0.5, e_seed=1)
    sm_dev = torch.from_numpy(sm).cuda()
    fs = []
    for block in (0, 1, 2):
        h = SprFactor(sysm.mspar, sysm.mtrees, sysm.msifa, precision)
        h.set_plan_params(block_factor=block)
        info = h.factor(sm_dev)
        assert info["ierr"] == 0
        fs.append(h.get_factor())
        h.close()
    a, b, c = fs
    assert np.array_equal(b, c)
    tol = 1e-11 if precision == "fp64" else 2e-4
    assert np.max(np.abs(a[:sysm.neq] / b[:sysm.neq] - 1)) < tol
    assert np.linalg.norm(a[sysm.neq:] - b[sysm.neq:]) / np.linalg.norm(b[sysm.neq:]) < tol


def test_spmv_many_rhs_matches_columns():
    """From 16 right-hand sides on, the FP64 product runs equation-major (X
    transposed in, one wavefront per row over all columns, transposed out):
    each column agrees with the one-column product to rounding (1e-14
    relative) and with the oracle's SPRPRM (A x on the SPR layout)."""
    torch = _torch()
    model, sysm, sm = make_system("hex8", 8, e_sigma=0.5, e_seed=1)
    h, sm_dev = _handle(sysm, sm, "fp64")
    h.set_pattern(sm)
    assert h.factor(sm_dev)["ierr"] == 0  # the product takes A from the factorization's input
    X = torch.randn(40, sysm.neq, dtype=torch.float64, device="cuda")
    Y = torch.empty_like(X)
    h.spmv(X, Y)
    worst = 0.0
    for k in (0, 1, 17, 39):
        yk = torch.empty_like(X[k:k + 1])
        h.spmv(X[k:k + 1].clone(), yk)
        worst = max(worst, ((yk[0] - Y[k]).norm() / yk[0].norm()).item())
    _, c_o = oracle.sprprm(sm[sysm.neq:], np.asfortranarray(X.cpu().numpy().T),
                           sysm.mspar.copy(), sysm.mtrees, sysm.msifa, iflag=1)
    rel = np.linalg.norm(Y.cpu().numpy().T - c_o) / np.linalg.norm(c_o)
    print(f"many-RHS product: worst column difference {worst:.2e}, vs SPRPRM {rel:.2e}", flush=True)
    h.close()
    assert worst <= 1e-14 and rel <= 1e-14
