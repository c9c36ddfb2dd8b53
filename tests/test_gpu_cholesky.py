"""GPU CC' forms (sprgpu_factor mode 2, the SPRFC2 semantics, and mode 4 /
sprgpu_sprfc4, the SPRFC4 semantics) against the
reference's SPRFC2 / SPRFC4 / SPRFS2 / SPRBS2 built from source (oracle/_ref).

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


# --------------------------------------------------------------------------
# SPRFC4: the CC' form with replacement (mode 4, sprgpu_sprfc4)
# --------------------------------------------------------------------------

def _ref_fc4(sysm, sm, tol):
    import refspr

    cc, ms = sm.copy(), sysm.mspar.copy()
    ierr, tolv, dcor = refspr.fc4(sysm, cc, tol, mspar=ms)
    return ierr, tolv, dcor, cc, ms


def _check_fc4(sysm, sm, tol, dcor_rows=None, sol_tol=1e-9):
    """Handle (mode 4) and drop-in SPRFC4 against the reference's SPRFC4:
    IERR, MSPAR(26:27) exact; TOL(2:3), DCOR and C to 1e-9 (DCOR entries to
    1e-9 of the largest |a_jj|), the SPRFS2/SPRBS2 solutions to sol_tol."""
    import refspr
    from sprgpu import spr

    torch = _torch()
    ierr, tolv, dcor, cc, ms = _ref_fc4(sysm, sm, tol)
    n = int(sysm.mspar[7])
    rows = slice(0, n) if dcor_rows is None else dcor_rows
    scale = np.abs(sm[n:]).max()
    from sprgpu.device import SprFactor

    h = SprFactor(sysm.mspar, sysm.mtrees, sysm.msifa, "fp64")
    try:
        info = h.factor(torch.from_numpy(sm).cuda(), tol=tol, mode=4)
        print("gpu", info, "ref", ierr, ms[25], ms[26], tolv, flush=True)
        assert info["ierr"] == ierr and info["supeqt"] == ms[26] and info["numneg"] == ms[25] == 0
        np.testing.assert_allclose([info["tol2"], info["tol3"]], tolv[1:3], rtol=1e-9)
        np.testing.assert_allclose(h.dcor()[rows], dcor[rows], rtol=1e-9, atol=1e-9 * scale)
        if ierr == 0 or ms[26] < 0:  # every pivot taken: C and the solves
            f = h.get_factor()
            assert np.max(np.abs(f[:n] / cc[:n] - 1)) < 1e-9
            assert np.linalg.norm(f[n:] - cc[n:]) / np.linalg.norm(cc[n:]) < 1e-9
            b = np.random.default_rng(7).uniform(-1, 1, (n, 2))
            y_r = np.asfortranarray(b.copy())
            assert refspr.fs2_bs2("fs", sysm, cc, y_r, mspar=ms.copy()) == 0
            x_r = y_r.copy(order="F")
            assert refspr.fs2_bs2("bs", sysm, cc, x_r, mspar=ms.copy()) == 0
            x = torch.from_numpy(np.ascontiguousarray(b.T)).cuda()
            h.solve(x, which=1)
            assert np.linalg.norm(x.cpu().numpy().T - y_r) / np.linalg.norm(y_r) < sol_tol
            h.solve(x, which=2)
            assert np.linalg.norm(x.cpu().numpy().T - x_r) / np.linalg.norm(x_r) < sol_tol
    finally:
        h.close()
    # the drop-in SPRFC4 (then SPRFS2 / SPRBS2 on its C)
    sg, msg = sm.copy(), sysm.mspar.copy()
    tg = np.zeros(4)
    tg[:len(tol)] = tol
    dg = np.zeros(2 * n)
    ierr_g = spr.sprfc4(msg, sysm.mtrees, sysm.msifa, sg, tg, dg)
    try:
        assert ierr_g == ierr and msg[0] == ms[0] == 6 and msg[25] == ms[25] and msg[26] == ms[26]
        np.testing.assert_allclose(tg[1:3], tolv[1:3], rtol=1e-9)
        np.testing.assert_allclose(dg.reshape(-1, 2)[rows], dcor[rows], rtol=1e-9,
                                   atol=1e-9 * scale)
        if ierr == 0 or ms[26] < 0:
            assert np.linalg.norm(sg[n:] - cc[n:]) / np.linalg.norm(cc[n:]) < 1e-9
            b = np.asfortranarray(np.random.default_rng(8).uniform(-1, 1, (n, 1)))
            y, y_r = b.copy(order="F"), b.copy(order="F")
            assert spr.sprfs2(msg, sysm.mtrees, sysm.msifa, sg, y) == 0 and msg[0] == 8
            assert spr.sprbs2(msg, sysm.mtrees, sysm.msifa, sg, y) == 0
            assert refspr.fs2_bs2("fs", sysm, cc, y_r, mspar=ms.copy()) == 0
            assert refspr.fs2_bs2("bs", sysm, cc, y_r, mspar=ms.copy()) == 0
            assert np.linalg.norm(y - y_r) / np.linalg.norm(y_r) < sol_tol
    finally:
        spr.release(sg)
    return ierr, ms


@pytest.mark.parametrize("tol", [(0.0, 0.0, 0.0, 0.0), (0.6, 0.0, 0.0, 1.0),
                                 (0.6, 0.0, 0.0, 10.0)])
def test_sprfc4_matches_reference(tol):
    model, sysm, sm = make_system("hex8", 10, e_sigma=0.5, e_seed=1)
    ierr, ms = _check_fc4(sysm, sm, tol)
    assert (ierr == 0) == (tol[0] == 0.0) and (tol[0] == 0.0 or ms[26] < 0)


@pytest.mark.parametrize("name,tol,sol_tol", [("hex8_3_indefinite", (1e-12, 0.0, 0.0, 1.0), 1e-9),
                                              ("hex8_2_zero_diag", (1e-12, 0.0, 0.0, 1.0), 1e-9),
                                              ("hex8_2_singular", (1e-6, 0.0, 0.0, 1.0), 1e-7)])
def test_sprfc4_special_cases_match_reference(name, tol, sol_tol):
    """A negative pivot (fails: DCOR compared up to the failing equation, where
    the reference stops), zero diagonals (replaced by one, no error) and
    near-singular pivots (replaced; six pivots of order 1e-6 of their a_jj
    make C C' ill-conditioned, and the solutions, not the factor, then
    differ from the reference's by about 3e-9)."""
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        g = {k: z[k] for k in z.files}
    sysm = types.SimpleNamespace(mspar=g["mspar_in"], mtrees=g["mtrees"], msifa=g["msifa"])
    ierr, tolv, dcor, cc, ms = _ref_fc4(sysm, g["sm_in"], tol)
    rows = slice(0, int(ms[26])) if ms[26] > 0 else None
    _check_fc4(sysm, g["sm_in"], tol, dcor_rows=rows, sol_tol=sol_tol)
