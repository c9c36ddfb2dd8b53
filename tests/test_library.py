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


@pytest.mark.parametrize("kind,n", [("hex8", 14), ("quad4", 60), ("frame", 300)])
def test_front_pool_holds_only_live_fronts(kind, n):
    """Fronts live in the pool from their level to their parent's assembly
    (the factored panels are kept apart); the planner checks that the fronts
    alive in each level are disjoint (it fails the plan otherwise), and the
    pool is well below the sum of all fronts (the layout in which every
    front stays allocated): at most the largest front memory alive in any
    level plus 25%."""
    kw = {"seed": 1} if kind == "frame" else {}
    model, sysm, _ = make_system(kind, n, **kw)
    out = _summary(sysm)
    xs, xl = sysm.xsuper(), sysm.xlindx()
    F, S = np.diff(xl), np.diff(xs)
    ld = (F + 15) // 16 * 16
    size = (ld * F + 63) // 64 * 64
    # heights and parents from the row lists (parent: the supernode of the
    # first row below the diagonal block)
    ns = len(F)
    col_node = np.repeat(np.arange(ns), S)
    lindx = sysm.lindx() - 1
    parent = np.array([col_node[lindx[xl[j] - 1 + S[j]]] if F[j] > S[j] else -1
                       for j in range(ns)])
    height = np.zeros(ns, dtype=int)
    for j in range(ns):
        if parent[j] >= 0:
            height[parent[j]] = max(height[parent[j]], height[j] + 1)
    last = np.where(parent >= 0, height[np.maximum(parent, 0)], height + 1)
    alive = np.zeros(height.max() + 2)
    for j in range(ns):
        alive[height[j]:last[j] + 1] += size[j]
    print(f"{kind} {n}: pool {out[4]:.3e}, live max {alive.max():.3e}, all fronts {size.sum():.3e}")
    assert alive.max() <= out[4] <= 1.25 * alive.max()
    assert out[4] < 0.6 * size.sum()


def test_plan_rejects_retained_supernodes():
    """MSPAR(57) = 1 alone is accepted (without retained supernodes it changes
    only the solves' LDB rule); a supernode with SUPSUP > 0 is rejected."""
    from sprgpu import _lib

    model, sysm, _ = make_system("hex8", 3)
    ms = sysm.mspar.copy()
    ms[56] = 1  # MSPAR(57)
    out = np.zeros(10)
    p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    assert _lib.lib().sprgpu_plan_summary(p(ms), p(sysm.mtrees), p(sysm.msifa), p(out)) == 0
    mt = sysm.mtrees.copy()
    mt[int(ms[45]) - 1 + sysm.nsuper + 1 + sysm.nsuper - 1] = sysm.nsuper  # SUPSUP of the root
    assert _lib.lib().sprgpu_plan_summary(p(ms), p(mt), p(sysm.msifa), p(out)) == -2
    assert "retained" in _lib.last_error()


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
