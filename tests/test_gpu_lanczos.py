"""GPU shift-invert Lanczos (sprgpu_lanczos, the task of SPRLAN) against
dense generalized eigenvalues.

The reference SPRLAN calls SAM library routines and BLAS that are not part of
this build, so it cannot run here: parity is pinned by mathematics instead.
The eigenvalues nearest the shift must equal scipy.linalg.eigh(A, B)'s to
1e-9 relative, the eigenvectors must satisfy |A v - lambda B v| <= 1e-8 |A v|
and V' B V = I to 1e-9.  Cases: a diagonal B (KSB <= 0) with shift 0 (the
lowest modes), an interior shift, and a sparse B on the same connectivity
(the unit-modulus stiffness), plus the error for a singular A - SHIFT*B.
"""
import numpy as np
import pytest
import scipy.linalg as sl

from conftest import make_system
from test_oracle import dense_equation_matrix

pytestmark = pytest.mark.gpu


def _torch():
    import torch

    return torch


def _setup():
    from sprgpu.device import SprFactor

    torch = _torch()
    model, sysm, sm = make_system("hex8", 5, e_sigma=0.5, e_seed=1)
    A = dense_equation_matrix(model, sysm)
    h = SprFactor(sysm.mspar, sysm.mtrees, sysm.msifa, "fp64")
    h.set_pattern(sm)
    return model, sysm, sm, A, h, torch.from_numpy(sm).cuda()


def _check(A, B, lam_ref, evl, everr, v, nconv, neval):
    print("evl", evl, "ref", lam_ref, "everr", everr, "nconv", nconv, flush=True)
    assert nconv == neval
    np.testing.assert_allclose(evl, lam_ref, rtol=1e-9)
    V = v.cpu().numpy().T
    R = A @ V - (B @ V) * evl[None, :len(V[0])]
    assert np.all(np.linalg.norm(R, axis=0) <= 1e-8 * np.linalg.norm(A @ V, axis=0))
    np.testing.assert_allclose(V.T @ B @ V, np.eye(V.shape[1]), atol=1e-9)


def test_lowest_modes_with_diagonal_b():
    model, sysm, sm, A, h, sm_dev = _setup()
    bdiag = np.random.default_rng(1).uniform(0.5, 1.5, sysm.neq)
    B = np.diag(bdiag)
    lam = sl.eigh(A, B, eigvals_only=True)
    evl, everr, v, nconv = h.lanczos(sm_dev, _torch().from_numpy(bdiag).cuda(), 0.0, 6, nevec=6,
                                     ksb=0)
    _check(A, B, lam[:6], evl, everr, v, nconv, 6)
    h.close()


def test_interior_shift_with_diagonal_b():
    model, sysm, sm, A, h, sm_dev = _setup()
    bdiag = np.random.default_rng(1).uniform(0.5, 1.5, sysm.neq)
    B = np.diag(bdiag)
    lam = sl.eigh(A, B, eigvals_only=True)
    shift = 0.5 * (lam[20] + lam[21]) + 0.1 * (lam[21] - lam[20])
    near = lam[np.argsort(np.abs(lam - shift))[:4]]
    evl, everr, v, nconv = h.lanczos(sm_dev, _torch().from_numpy(bdiag).cuda(), shift, 4,
                                     nevec=4, ksb=0)
    _check(A, B, near, evl, everr, v, nconv, 4)
    h.close()


def test_sparse_b_on_the_same_connectivity_and_singular_shift():
    from sprgpu._lib import SprGpuError

    torch = _torch()
    model, sysm, sm, A, h, sm_dev = _setup()
    model_b, sysm_b, sm_b = make_system("hex8", 5)   # unit moduli, same mesh
    assert np.array_equal(sysm_b.mspar, sysm.mspar)
    B = dense_equation_matrix(model_b, sysm_b)
    lam = sl.eigh(A, B, eigvals_only=True)
    b_dev = torch.from_numpy(sm_b[sysm.neq:]).cuda()
    evl, everr, v, nconv = h.lanczos(sm_dev, b_dev, 0.0, 5, nevec=3)
    _check(A, B, lam[:5], evl, everr, v, nconv, 5)
    # A - SHIFT*B with SHIFT = A's own scale on a B = A: exactly singular
    with pytest.raises(SprGpuError, match="singular"):
        h.lanczos(sm_dev, sm_dev[sysm.neq:], 1.0, 2)
    h.close()
