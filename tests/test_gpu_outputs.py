"""Every output of the factor entry points against the reference, on the GPU.

The reference-written fixtures (tests/golden, produced by the reference SPR
built from source, see tests/golden/make_golden.py) and the oracle pin:
* DCOR of SPRFC3 (sprfc3.f:560-575, :1083-1099) on the handle (mode 3) and
  through the SPRFC3 drop-in, with MSPAR(27) = -(replaced pivots), TOL(2:3)
  and the factor written back to SM;
* IERR = 3 and NUMNEG = MSPAR(26) of an indefinite matrix in the LDL' form
  (sprfc1.f:262-267), on the handle and through SPRSOL;
* a structure with cache panels, MSPAR(9) > 0 (SPRSM2, sprsmb.f:590-747; the
  GPU's blocking ignores the panels, so the results agree to rounding);
* the FP32 factorization in mode 3: the replaced pivots and their values;
* the drop-in's device factor cache: a factor computed elsewhere into an SM
  the drop-in has seen before is taken from SM.
Integer outputs must match exactly; floating-point ones to 1e-9 relative
(the GPU eliminates in another order), FP32 ones as stated per test.
"""
import os

import numpy as np
import pytest

import oracle
from conftest import gpu_available, make_system

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs an MI355X")]

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(autouse=True, scope="module")
def _torch_runtime_first():
    import torch

    assert torch.cuda.is_available()


def _golden(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _close(a, b, tol=1e-9):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.linalg.norm(a - b) <= tol * max(np.linalg.norm(b), 1e-300)


def _check_dcor(dcor, ref, tol=1e-9):
    """Same replaced / recorded equations; values to tol (DCOR(2,j), the
    replacement sign(d) TOL(1) max(1,|TOL(4) a_jj|), is exact up to a_jj)."""
    dcor, ref = np.asarray(dcor).reshape(-1, 2), np.asarray(ref).reshape(-1, 2)
    assert np.array_equal(dcor[:, 1] != 0, ref[:, 1] != 0), "replaced pivots differ"
    assert np.array_equal(dcor[:, 0] != 0, ref[:, 0] != 0), "recorded pivots differ"
    np.testing.assert_allclose(dcor[:, 1], ref[:, 1], rtol=1e-14)
    # DCOR(1,j) of a replaced pivot is a near-zero difference: compare on the
    # scale of the diagonal it was tested against
    np.testing.assert_allclose(dcor[:, 0], ref[:, 0], rtol=tol, atol=1e-12)


def test_dcor_on_the_handle_matches_the_reference_fixture():
    import torch
    from sprgpu.device import SprFactor

    g = _golden("hex8_2_sprfc3")
    h = SprFactor(g["mspar_in"], g["mtrees"], g["msifa"], "fp64")
    try:
        info = h.factor(torch.from_numpy(g["sm_in"]).cuda(), tol=tuple(g["tol_in"]), mode=3)
        assert info["ierr"] == int(g["ierr"]) == -3
        assert info["supeqt"] == g["mspar_out"][26] == -6
        assert info["numneg"] == g["mspar_out"][25]
        np.testing.assert_allclose([info["tol2"], info["tol3"]], g["tol_out"][1:3], rtol=1e-9)
        dc = h.dcor()
        _check_dcor(dc, g["dcor"])
        # the matrix is singular (unclamped): columns after the first
        # replaced pivot hold cancellation residues scaled by 1/TOL(1), so
        # the factor is compared before it
        j0 = int(np.nonzero(dc[:, 1])[0][0])
        f = h.get_factor()
        assert _close(f[:j0], g["sm_out"][:j0])
    finally:
        h.close()


def test_sprfc3_dropin_matches_the_reference_fixture():
    from sprgpu import spr

    g = _golden("hex8_2_sprfc3")
    ms, sm = g["mspar_in"].copy(), g["sm_in"].copy()
    tol = g["tol_in"].astype(np.float64).copy()
    dcor = np.zeros(2 * int(ms[7]))
    ierr = spr.sprfc3(ms, g["mtrees"], g["msifa"], sm, tol, dcor)
    try:
        assert ierr == int(g["ierr"]) == -3
        # MSPAR(1) = 5 after the factor (the fixture's 7 is after its solves)
        assert ms[0] == 5 and np.array_equal(ms[1:], g["mspar_out"][1:])
        np.testing.assert_allclose(tol[1:3], g["tol_out"][1:3], rtol=1e-9)
        assert tol[0] == g["tol_out"][0] and tol[3] == g["tol_out"][3]
        _check_dcor(dcor, g["dcor"])
        j0 = int(np.nonzero(dcor.reshape(-1, 2)[:, 1])[0][0])
        assert _close(sm[:j0], g["sm_out"][:j0])
        # the replaced-pivot factor goes on to the solves as the reference's did
        # (their values amplify rounding residues by 1/TOL(1): not compared)
        x = np.asfortranarray(g["b"].copy())
        assert spr.sprfs1(ms, g["mtrees"], g["msifa"], sm, x) == int(g["ierr_solve"]) == 0
        assert spr.sprbs1(ms, g["mtrees"], g["msifa"], sm, x) == 0
        assert ms[0] == g["mspar_out"][0] == 7
        assert np.all(np.isfinite(x))
    finally:
        spr.release(sm)


def test_sprfc3_dropin_many_replacements_matches_oracle():
    """TOL(1) = 0.6 on hex8 10^3: many pivots of the wide fronts are replaced."""
    from sprgpu import spr

    _, sysm, sm = make_system("hex8", 10, e_sigma=0.5, e_seed=1)
    tol = np.array([0.6, 0.0, 0.0, 1.0])
    ms_o, sm_o = sysm.mspar.copy(), sm.copy()
    ierr_o, tol_o, dcor_o = oracle.sprfc3(ms_o, sysm.mtrees, sysm.msifa, sm_o, tol)
    ms, smg, tolg = sysm.mspar.copy(), sm.copy(), tol.copy()
    dcor = np.zeros(2 * sysm.neq)
    ierr = spr.sprfc3(ms, sysm.mtrees, sysm.msifa, smg, tolg, dcor)
    spr.release(smg)
    print(f"SPRFC3 TOL(1)=0.6: gpu ierr {ierr} MSPAR(27) {ms[26]}, oracle {ierr_o} {ms_o[26]}",
          flush=True)
    assert ierr == ierr_o == -3
    assert np.array_equal(ms, ms_o) and ms[26] < -20
    np.testing.assert_allclose(tolg[1:3], tol_o[1:3], rtol=1e-9)
    _check_dcor(dcor, dcor_o.T)  # the oracle returns DCOR as (2, NEQ)
    assert _close(smg[:sysm.neq], sm_o[:sysm.neq]) and _close(smg[sysm.neq:], sm_o[sysm.neq:])


def test_indefinite_ldl_reports_numneg():
    import torch
    from sprgpu import spr
    from sprgpu.device import SprFactor

    g = _golden("hex8_3_indefinite")
    assert int(g["mspar_out"][25]) == 11 and int(g["ierr"]) == 0
    h = SprFactor(g["mspar_in"], g["mtrees"], g["msifa"], "fp64")
    try:
        info = h.factor(torch.from_numpy(g["sm_in"]).cuda())
        assert info["ierr"] == 3  # NUMNEG > 0: the handle's IERR +3 (sprfc1.f:262-265)
        assert info["numneg"] == 11
        np.testing.assert_allclose([info["tol2"], info["tol3"]], g["tol_out"][1:3], rtol=1e-9)
        x = torch.from_numpy(np.ascontiguousarray(g["b"].T)).cuda()
        h.solve(x)
        assert _close(x.cpu().numpy().T, g["x"])
    finally:
        h.close()
    # the drop-in: SPRFC1 reports IERR = 3; SPRSOL goes on to solve (IERR >= 0)
    # and ends with SPRBS1's IERR = 0, as the reference's SPRSOL did
    ms, sm = g["mspar_in"].copy(), g["sm_in"].copy()
    tol = np.zeros(3)
    assert spr.sprfc1(ms, g["mtrees"], g["msifa"], sm, tol) == 3
    assert ms[25] == 11
    spr.release(sm)
    ms, sm = g["mspar_in"].copy(), g["sm_in"].copy()
    x = np.asfortranarray(g["b"].copy())
    ierr = spr.sprsol(3, ms, g["mtrees"], g["msifa"], sm, x, tol)
    spr.release(sm)
    assert ierr == int(g["ierr"]) == 0
    assert np.array_equal(ms, g["mspar_out"])
    assert _close(sm, g["sm_out"]) and _close(x, g["x"])


def test_cache_panels_fixture():
    from sprgpu import spr

    g = _golden("hex8_4_panels")
    assert g["mspar_in"][8] > 0 and g["mspar_out"][21] > 0  # MSPAR(9) KB, MSPAR(22) panels
    ms, sm = g["mspar_in"].copy(), g["sm_in"].copy()
    x = np.asfortranarray(g["b"].copy())
    tol = g["tol_in"].astype(np.float64).copy()
    ierr = spr.sprsol(3, ms, g["mtrees"], g["msifa"], sm, x, tol)
    spr.release(sm)
    assert ierr == int(g["ierr"]) == 0
    assert np.array_equal(ms, g["mspar_out"])
    np.testing.assert_allclose(tol[1:3], g["tol_out"][1:3], rtol=1e-9)
    assert _close(sm, g["sm_out"]) and _close(x, g["x"])


def test_fp32_mode3_replacements():
    """The FP32 factor in SPRFC3 mode on hex8 10^3 with TOL(1) = 0.6: the
    pivots are tested in FP64 against the FP64 a_jj, so the replaced set
    equals the oracle's except for pivots within FP32 rounding of the
    threshold, and the replacement values are exact."""
    import torch
    from sprgpu.device import SprFactor

    _, sysm, sm = make_system("hex8", 10, e_sigma=0.5, e_seed=1)
    tol = [0.6, 0.0, 0.0, 1.0]
    ms_o, sm_o = sysm.mspar.copy(), sm.copy()
    ierr_o, _, dcor_o = oracle.sprfc3(ms_o, sysm.mtrees, sysm.msifa, sm_o, tol)
    h = SprFactor(sysm.mspar, sysm.mtrees, sysm.msifa, "fp32")
    try:
        info = h.factor(torch.from_numpy(sm).cuda(), tol=tuple(tol), mode=3)
        dcor = h.dcor()
        f = h.get_factor()
    finally:
        h.close()
    ref = np.asarray(dcor_o).T  # the oracle returns DCOR as (2, NEQ)
    rep_g, rep_o = dcor[:, 1] != 0, ref[:, 1] != 0
    # a_jj of every equation, for the ratio the pivot test uses
    adiag = np.asarray(sm[sysm.neq:])[sysm_diag_positions(sysm)]
    differ = np.nonzero(rep_g != rep_o)[0]
    # |d| / |a_jj| of the oracle's pivot before any replacement
    ratio = np.abs(np.where(rep_o, ref[:, 0], sm_o[:sysm.neq])) / np.abs(adiag)
    print(f"FP32 mode 3: {rep_g.sum()} replaced (oracle {rep_o.sum()}), "
          f"{differ.size} differ, ratios there {ratio[differ]}", flush=True)
    assert info["ierr"] == ierr_o == -3
    assert differ.size <= max(2, rep_o.sum() // 20)
    assert np.all(np.abs(ratio[differ] - 0.6) < 1e-3)
    both = rep_g & rep_o
    np.testing.assert_allclose(dcor[both, 1], ref[both, 1], rtol=1e-14)


def sysm_diag_positions(sysm):
    from sprgpu import host

    eq = np.arange(sysm.neq)
    return host.lnz_positions(sysm, eq, eq) - sysm.neq


def test_dropin_takes_a_factor_computed_elsewhere_into_a_known_sm():
    """SPRFC1 on the GPU, then the reference factors another matrix into the
    same SM array: the next SPRFS1/SPRBS1 must use the factor now in SM."""
    import refspr
    from sprgpu import spr

    _, sysm, sm = make_system("hex8", 6, e_sigma=0.5, e_seed=1)
    b = np.asfortranarray(np.random.default_rng(3).uniform(-1, 1, (sysm.neq, 2)))
    work = sm.copy()
    ms = sysm.mspar.copy()
    assert spr.sprfc1(ms, sysm.mtrees, sysm.msifa, work, np.zeros(3)) == 0
    try:
        # another matrix on the same pattern, factored by the reference into `work`
        sm2 = 2.0 * sm
        work[:] = sm2
        ms2 = sysm.mspar.copy()
        ierr, _, _ = refspr.sol(1, sysm, work, None, mspar=ms2)
        assert ierr == 0
        x = b.copy(order="F")
        assert spr.sprfs1(ms2, sysm.mtrees, sysm.msifa, work, x) == 0
        assert spr.sprbs1(ms2, sysm.mtrees, sysm.msifa, work, x) == 0
        x_o = b.copy(order="F")
        assert oracle.sprsol(3, sysm.mspar.copy(), sysm.mtrees, sysm.msifa, sm2.copy(), x_o)[0] == 0
        assert _close(x, x_o)
    finally:
        spr.release(work)


def test_dropin_cache_evicts_old_entries():
    """Six SM arrays factored in turn: the cache keeps at most four device
    factors, and every one still solves (an evicted one from SM)."""
    from sprgpu import spr

    _, sysm, sm = make_system("hex8", 4, e_sigma=0.5, e_seed=1)
    b = np.asfortranarray(np.random.default_rng(4).uniform(-1, 1, (sysm.neq, 1)))
    x_o = b.copy(order="F")
    assert oracle.sprsol(3, sysm.mspar.copy(), sysm.mtrees, sysm.msifa, sm.copy(), x_o)[0] == 0
    sms = [sm.copy() for _ in range(6)]
    mss = [sysm.mspar.copy() for _ in range(6)]
    try:
        for s, m in zip(sms, mss):
            assert spr.sprfc1(m, sysm.mtrees, sysm.msifa, s, np.zeros(3)) == 0
        for s, m in zip(sms, mss):
            x = b.copy(order="F")
            assert spr.sprfs1(m, sysm.mtrees, sysm.msifa, s, x) == 0
            assert spr.sprbs1(m, sysm.mtrees, sysm.msifa, s, x) == 0
            assert _close(x, x_o)
    finally:
        for s in sms:
            spr.release(s)
