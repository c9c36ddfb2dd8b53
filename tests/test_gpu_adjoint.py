"""Adjoint kernels and the torch.autograd binding (sprgpu.adjoint) on the GPU.

The gradient of the compliance L = f^T u with respect to the element moduli
is checked against (1) the same formula evaluated in numpy from the GPU
solution, (2) central finite-difference quotients (h = 1e-6 relative),
evaluated without cancellation, and (3) Euler's identity for a function
homogeneous of degree -1 in the moduli: sum_e E_e dL/dE_e = -L.
"""
import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs an MI355X")]


def _system(n, precision, seed=4):
    from sprgpu import fem
    from sprgpu.adjoint import ElasticSystem

    model = fem.hex8_grid(n, e_sigma=0.5, e_seed=seed)
    rng = np.random.default_rng(seed)
    rng.normal(0.0, 0.5, model.nel)
    f = rng.uniform(-1.0, 1.0, model.ndof)
    return model, f, ElasticSystem(model, precision=precision, rtol=1e-12, maxit=40)


def test_element_grad_kernel_matches_numpy():
    import torch

    model, f, sysobj = _system(5, "fp64")
    try:
        p = model.escale.copy()
        sysobj.factor(p)
        x = sysobj.solve_eq(sysobj.to_eq(torch.from_numpy(f).cuda()))
        lam = torch.from_numpy(np.random.default_rng(7).normal(size=sysobj.neq)).cuda()
        g = sysobj.element_grad(lam, x).cpu().numpy()
        eq = sysobj.elem_eq.cpu().numpy()
        xh, lh = x.cpu().numpy(), lam.cpu().numpy()
        xe = np.where(eq >= 0, xh[np.maximum(eq, 0)], 0.0)
        le = np.where(eq >= 0, lh[np.maximum(eq, 0)], 0.0)
        k0 = sysobj.k0_host[0]
        ref = -np.einsum("ei,ij,ej->e", le, k0, xe)
        np.testing.assert_allclose(g, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    finally:
        sysobj.close()


def test_pattern_grad_kernel_matches_numpy():
    import torch
    from sprgpu import host

    model, f, sysobj = _system(3, "fp64")
    try:
        sysobj.factor(model.escale.copy())
        rng = np.random.default_rng(11)
        lam = torch.from_numpy(rng.normal(size=sysobj.neq)).cuda()
        x = torch.from_numpy(rng.normal(size=sysobj.neq)).cuda()
        g = sysobj.h.pattern_grad(lam, x).cpu().numpy()
        lh, xh = lam.cpu().numpy(), x.cpu().numpy()
        s = sysobj.sysm
        xsuper, xlindx, lindx = s.xsuper() - 1, s.xlindx() - 1, s.lindx() - 1
        for j in range(s.nsuper):
            rows = lindx[xlindx[j]:xlindx[j + 1]]
            for c in range(xsuper[j], xsuper[j + 1]):
                r = rows[rows >= c]
                pos = host.lnz_positions(s, r, np.full(r.size, c))
                want = np.where(r == c, -lh[c] * xh[c], -(lh[r] * xh[c] + lh[c] * xh[r]))
                # the kernel may contract a*b + c*d into one FMA: a few ulps
                tol = 4e-16 * (np.abs(lh[r] * xh[c]) + np.abs(lh[c] * xh[r]))
                assert np.all(np.abs(g[pos] - want) <= tol), (j, c)
    finally:
        sysobj.close()


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_compliance_gradient(precision):
    import torch
    from sprgpu.adjoint import spr_solve

    model, f, sysobj = _system(8, precision)
    try:
        p = torch.from_numpy(model.escale.copy()).cuda().requires_grad_(True)
        ft = torch.from_numpy(f).cuda()
        u = spr_solve(p, ft, sysobj)
        L = torch.dot(ft, u)
        L.backward()
        g = p.grad.detach().cpu().numpy()
        Lv = float(L.detach())
        # Euler: L is homogeneous of degree -1 in the moduli
        assert abs(float(np.dot(model.escale, g)) + Lv) <= 1e-9 * abs(Lv)
        # central differences on a few elements
        f_eq = sysobj.to_eq(ft)
        for e in np.random.default_rng(5).choice(model.nel, 6, replace=False):
            fd = sysobj.central_difference(model.escale, f_eq, int(e), h_rel=1e-6)
            assert abs(fd - g[e]) <= 1e-6 * abs(g[e]), (e, fd, g[e])
    finally:
        sysobj.close()


def test_load_gradient_is_the_adjoint_solution():
    """For L = g^T u, dL/df = K^-1 g (K symmetric)."""
    import torch
    from sprgpu.adjoint import spr_solve

    model, f, sysobj = _system(6, "fp64")
    try:
        p = torch.from_numpy(model.escale.copy()).cuda()
        ft = torch.from_numpy(f).cuda().requires_grad_(True)
        gw = torch.from_numpy(np.random.default_rng(3).normal(size=model.ndof)).cuda()
        u = spr_solve(p, ft, sysobj)
        torch.dot(gw, u).backward()
        want = sysobj.to_dofs(sysobj.solve_eq(sysobj.to_eq(gw)))
        np.testing.assert_allclose(ft.grad.cpu().numpy(), want.cpu().numpy(), rtol=1e-10,
                                   atol=1e-12 * float(want.abs().max()))
    finally:
        sysobj.close()
