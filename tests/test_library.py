"""The C-ABI libraries load and export their declared symbols; the host
planner (no GPU needed) produces a consistent schedule."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT, make_system


def _declared_functions():
    text = open(os.path.join(ROOT, "include", "sprgpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sprgpu_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    from sprgpu import _lib

    names = _declared_functions()
    assert len(names) >= 20
    L = C.CDLL(_lib.LIB_PATH)
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_lib.SIGNATURES), "ctypes table out of sync with the header"


def test_fortran_dropin_exports_reference_names():
    from sprgpu import _lib

    out = subprocess.run(["nm", "-D", "--defined-only", _lib.FLIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    for n in ("sprfc1_", "sprfc2_", "sprfc3_", "sprfs1_", "sprfs2_", "sprbs1_", "sprbs2_",
              "sprsol_", "sprgpu_release_"):
        assert re.search(rf"\bT {n}$", out, flags=re.M), n


def _summary(sysm):
    from sprgpu import _lib

    out = np.zeros(10)
    p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    rc = _lib.lib().sprgpu_plan_summary(p(sysm.mspar), p(sysm.mtrees), p(sysm.msifa), p(out))
    assert rc == 0, _lib.last_error()
    return out


@pytest.mark.parametrize("kind,n", [("hex8", 6), ("quad4", 40), ("frame", 300)])
def test_plan_is_consistent(kind, n):
    import refspr

    kw = {"seed": 1} if kind == "frame" else {}
    model, sysm, _ = make_system(kind, n, **kw)
    out = _summary(sysm)
    assert out[0] == sysm.neq and out[1] == sysm.nsuper
    # factor flops equal the reference's RINFO(3)
    assert out[9] == refspr.factor_flops(sysm)
    # solve workspace holds every front once: sum of F = NOFSUB
    assert out[5] == sysm.mspar[14]
    xs, xl = sysm.xsuper(), sysm.xlindx()
    F, S = np.diff(xl), np.diff(xs)
    assert out[6] == np.sum(F <= 64)
    # one diagonal-block launch slot per 128-column outer block of a blocked front
    assert out[7] == np.sum(np.where(F > 64, (S + 127) // 128, 0))


def test_plan_rejects_retained_supernodes():
    from sprgpu import _lib

    model, sysm, _ = make_system("hex8", 3)
    ms = sysm.mspar.copy()
    ms[56] = 1  # MSPAR(57): partial factorization
    out = np.zeros(10)
    p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    assert _lib.lib().sprgpu_plan_summary(p(ms), p(sysm.mtrees), p(sysm.msifa), p(out)) == -2
    assert "MSPAR(57)" in _lib.last_error()


def test_dropin_stage_protocol_without_gpu():
    """MSPAR(1)<0 is rejected before any device work (sprfc1.f:140-145)."""
    from sprgpu import spr

    model, sysm, sm = make_system("hex8", 2)
    ms = sysm.mspar.copy()
    ms[0] = -3
    assert spr.sprfc1(ms, sysm.mtrees, sysm.msifa, sm, np.zeros(3)) == -1
    assert ms[0] == -5
    ms[0] = -3
    b = np.zeros((sysm.neq, 1), order="F")
    assert spr.sprfs1(ms, sysm.mtrees, sysm.msifa, sm, b) == -1
    assert ms[0] == -7
    ms = sysm.mspar.copy()
    ms[3] = 1  # NTREES too small
    assert spr.sprfc1(ms, sysm.mtrees, sysm.msifa, sm, np.zeros(3)) == -2
    assert ms[0] == -2
