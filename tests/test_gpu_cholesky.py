"""GPU CC' form (sprgpu_factor mode 2, the SPRFC2 semantics) against the
reference's SPRFC2 / SPRFS2 / SPRBS2 built from source (oracle/_ref).

tests/test_oracle.py::test_reference_cct_is_ldlt_scaled pins on the reference
itself the relations the GPU form is built on (C = L sqrt(D), SPRFS2 =
sqrt(D) D^-1 L^-1, SPRBS2 = L^-T D^-1/2).  Bars: IERR, SUPEQT and MSPAR(26)
exact; TOL(2:3), the factor in SM layout and the solutions within 1e-9
relative.  Cases: an SPD hex8 mesh (success), TOL(1) = 0.6 (a pivot under the
threshold), an indefinite matrix (the first negative pivot fails, which
SPRFC1 would accept) and a zero diagonal.
"""
import os
import types

import numpy as np
import pytest

from conftest import make_system

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _torch():
    import torch

    return torch


def _ref_fc2(sysm, sm, tol):
    import refspr

    cc, ms = sm.copy(), sysm.mspar.copy()
    ierr, tolv = refspr.fc2(sysm, cc, tol=tol, mspar=ms)
    return ierr, tolv, cc, ms


def _gpu(sysm, sm, tol):
    from sprgpu.device import SprFactor

    torch = _torch()
    h = SprFactor(sysm.mspar, sysm.mtrees, sysm.msifa, "fp64")
    info = h.factor(torch.from_numpy(sm).cuda(), tol=tuple(tol) + (0.0,), mode=2)
    return h, info


def _compare_info(info, ierr, tolv, ms):
    print("gpu", info, "ref", ierr, tolv, ms[25], ms[26], flush=True)
    assert info["ierr"] == ierr
    assert info["supeqt"] == ms[26]
    assert info["numneg"] == ms[25]
    np.testing.assert_allclose([info["tol2"], info["tol3"]], tolv[1:], rtol=1e-9)


def test_cct_factor_and_solves_match_reference():
    import refspr

    torch = _torch()
    model, sysm, sm = make_system("hex8", 10, e_sigma=0.5, e_seed=1)
    ierr, tolv, cc, ms = _ref_fc2(sysm, sm, (0.0, 0.0, 0.0))
    assert ierr == 0
    h, info = _gpu(sysm, sm, (0.0, 0.0, 0.0))
    _compare_info(info, ierr, tolv, ms)
    n = sysm.neq
    f = h.get_factor()
    assert np.max(np.abs(f[:n] / cc[:n] - 1)) < 1e-9
    assert np.linalg.norm(f[n:] - cc[n:]) / np.linalg.norm(cc[n:]) < 1e-9
    b = np.random.default_rng(5).uniform(-1, 1, (n, 3))
    y_r = np.asfortranarray(b.copy())
    assert refspr.fs2_bs2("fs", sysm, cc, y_r, mspar=ms) == 0
    x = torch.from_numpy(np.ascontiguousarray(b.T)).cuda()
    h.solve(x, which=1)
    y = x.cpu().numpy().T
    assert np.linalg.norm(y - y_r) / np.linalg.norm(y_r) < 1e-9
    x_r = y_r.copy(order="F")
    assert refspr.fs2_bs2("bs", sysm, cc, x_r, mspar=ms) == 0
    h.solve(x, which=2)
    assert np.linalg.norm(x.cpu().numpy().T - x_r) / np.linalg.norm(x_r) < 1e-9
    x3 = torch.from_numpy(np.ascontiguousarray(b.T)).cuda()
    h.solve(x3)
    assert np.linalg.norm(x3.cpu().numpy().T - x_r) / np.linalg.norm(x_r) < 1e-9
    h.close()


def test_cct_pivot_threshold_matches_reference():
    model, sysm, sm = make_system("hex8", 10, e_sigma=0.5, e_seed=1)
    tol = (0.6, 0.0, 0.0)
    ierr, tolv, _, ms = _ref_fc2(sysm, sm, tol)
    assert ierr == -3 and ms[26] > 0
    h, info = _gpu(sysm, sm, tol)
    _compare_info(info, ierr, tolv, ms)
    h.close()


@pytest.mark.parametrize("name", ["hex8_3_indefinite", "hex8_2_zero_diag"])
def test_cct_failures_match_reference(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        g = {k: z[k] for k in z.files}
    sysm = types.SimpleNamespace(mspar=g["mspar_in"], mtrees=g["mtrees"], msifa=g["msifa"])
    tol = (0.0, 0.0, 0.0)
    ierr, tolv, _, ms = _ref_fc2(sysm, g["sm_in"], tol)
    assert ierr == -3
    h, info = _gpu(sysm, g["sm_in"], tol)
    _compare_info(info, ierr, tolv, ms)
    h.close()
