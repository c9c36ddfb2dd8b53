"""Pins the oracle (oracle/spr_oracle.c, the C restatement of the SPR numeric
path) before it is trusted as the checker of the GPU library.

1. Golden fixtures written by the reference itself (tests/golden/*.npz): the
   oracle must reproduce the reference's factor, solution, TOL, MSPAR, IERR
   and DCOR bit for bit, including the error paths (-35 zero diagonal, -30
   pivot below TOL(1), SPRFC3 replacement) and cache panelling (MSPAR(9)).
2. The live reference build (oracle/_ref) on further models, bit for bit.
3. Mathematics that does not depend on the reference: a single dense front
   equals the textbook LDL' obtained from numpy's Cholesky; NUMNEG equals the
   number of negative eigenvalues (Sylvester's law of inertia); the solution
   of A x = A x* recovers x*; SPRPRM equals the dense product.
"""
import glob
import os

import numpy as np
import pytest

import oracle
from conftest import make_system

# the reference's factor/solve fixtures (make_golden.py); the 40^3 adjoint
# and solution files (make_adjoint_reference.py, make_reference_solution.py)
# are checked by the GPU benchmark
GOLDEN = sorted(p for p in glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz"))
                if "reference" not in os.path.basename(p))


def _load(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
def test_oracle_reproduces_reference_fixture(path):
    g = _load(path)
    ms = g["mspar_in"].copy()
    sm = g["sm_in"].copy()
    x = np.asfortranarray(g["b"].copy())
    if int(g["mode"]) == 3:
        ierr, tol, dcor = oracle.sprfc3(ms, g["mtrees"], g["msifa"], sm, g["tol_in"])
        assert np.array_equal(dcor.T.ravel(), g["dcor"])
        if "ierr_solve" in g:
            ierr2, _ = oracle.sprsol(4, ms, g["mtrees"], g["msifa"], sm, x, tol=tol[:3])
            assert ierr2 == int(g["ierr_solve"])
    else:
        ierr, tol = oracle.sprsol(3, ms, g["mtrees"], g["msifa"], sm, x, tol=g["tol_in"])
    assert ierr == int(g["ierr"])
    assert np.array_equal(ms, g["mspar_out"])
    assert np.array_equal(tol, g["tol_out"])
    assert np.array_equal(sm, g["sm_out"])
    assert np.array_equal(x, g["x"])


def test_fixtures_cover_the_error_paths():
    by = {os.path.basename(p)[:-4]: _load(p) for p in GOLDEN}
    # SPRSOL(IOP=3) returns the IERR of its last step: SPRFS1 resets the
    # factor's warning 3 to 0 (sprsol.f:38-47); NUMNEG stays in MSPAR(26)
    assert int(by["hex8_3_indefinite"]["ierr"]) == 0 and by["hex8_3_indefinite"]["mspar_out"][25] > 0
    assert int(by["hex8_2_zero_diag"]["ierr"]) == -3
    assert int(by["hex8_2_singular"]["ierr"]) == -3
    assert int(by["hex8_2_sprfc3"]["mspar_out"][26]) < 0   # -#replaced pivots
    assert int(by["hex8_4_panels"]["mspar_in"][21]) > 0      # NBLOCK > 0: panelled path


def test_live_reference_symbolic_matches_fixture():
    """The reference build used by the tests is the one that wrote the fixtures."""
    import refspr
    from sprgpu import fem

    g = _load(os.path.join(os.path.dirname(__file__), "golden", "hex8_3_lognormal.npz"))
    model = fem.hex8_grid(3, e_sigma=0.5, e_seed=1)
    sysm = refspr.symbolic(model)
    assert np.array_equal(sysm.mspar, g["mspar_in"])
    assert np.array_equal(sysm.mtrees, g["mtrees"])
    assert np.array_equal(sysm.msifa, g["msifa"])
    assert np.array_equal(refspr.assemble(model, sysm), g["sm_in"])


LIVE = [("hex8", 5, {"e_sigma": 0.5, "e_seed": 1}), ("quad4", 12, {}),
        ("frame", 120, {"seed": 5}), ("hex8", 7, {})]


@pytest.mark.parametrize("kind,n,kw", LIVE, ids=[f"{c[0]}{c[1]}" for c in LIVE])
def test_oracle_bitexact_vs_live_reference(kind, n, kw):
    import refspr

    model, sysm, sm = make_system(kind, n, **kw)
    b = np.asfortranarray(np.random.default_rng(7).uniform(-1, 1, (sysm.neq, 2)))
    sm_r, x_r, ms_r = sm.copy(), b.copy(order="F"), sysm.mspar.copy()
    ierr_r, _, tol_r = refspr.sol(3, sysm, sm_r, x_r, mspar=ms_r)
    sm_o, x_o, ms_o = sm.copy(), b.copy(order="F"), sysm.mspar.copy()
    ierr_o, tol_o = oracle.sprsol(3, ms_o, sysm.mtrees, sysm.msifa, sm_o, x_o)
    assert ierr_o == ierr_r == 0
    assert np.array_equal(sm_o, sm_r)
    assert np.array_equal(x_o, x_r)
    assert np.array_equal(tol_o, tol_r)
    assert np.array_equal(ms_o, ms_r)


# ---------------------------------------------------------------------------
# mathematics, independent of the reference
# ---------------------------------------------------------------------------

def dense_equation_matrix(model, sysm):
    """Assembled A in SPR equation order, built from the element matrices."""
    neq = sysm.neq
    A = np.zeros((neq, neq))
    for e in range(model.nel):
        dofs = model.element_dofs(e)
        eq = sysm.meqn[dofs]
        k = model.element_matrix(e)
        keep = eq > 0
        idx = eq[keep] - 1
        A[np.ix_(idx, idx)] += k[np.ix_(keep, keep)]
    return A


def _single_front_model(nn, rng):
    """One element coupling nn one-dof nodes: the factor is a single dense front."""
    from sprgpu.fem import SamModel

    Q = rng.normal(size=(nn, nn))
    K = Q @ Q.T + nn * np.eye(nn)
    return SamModel("dense", nn, 1, nn, np.arange(1, nn + 2, dtype=np.int64),
                    np.ones(nn, dtype=np.int64), np.array([1, nn + 1], dtype=np.int64),
                    np.arange(1, nn + 1, dtype=np.int64), np.array([nn], dtype=np.int64),
                    K.ravel(order="F").copy(), np.zeros(1, dtype=np.int64), None)


def test_single_front_equals_textbook_ldlt():
    import refspr

    model = _single_front_model(40, np.random.default_rng(11))
    sysm = refspr.symbolic(model)
    assert sysm.nsuper == 1
    sm = refspr.assemble(model, sysm)
    A = dense_equation_matrix(model, sysm)
    ms = sysm.mspar.copy()
    ierr, _ = oracle.sprsol(1, ms, sysm.mtrees, sysm.msifa, sm)
    assert ierr == 0
    C = np.linalg.cholesky(A)          # A = C C^T, C = L sqrt(D)
    d_ref = np.diag(C) ** 2
    L_ref = C / np.diag(C)
    n = sysm.neq
    np.testing.assert_allclose(sm[:n], d_ref, rtol=1e-12)
    # packed columns: d on the diagonal, l below
    pos = n
    for j in range(n):
        col = sm[pos:pos + n - j]
        np.testing.assert_allclose(col[0], d_ref[j], rtol=1e-12)
        np.testing.assert_allclose(col[1:], L_ref[j + 1:, j], rtol=1e-10, atol=1e-13)
        pos += n - j


def test_numneg_is_inertia():
    """Sylvester: NUMNEG of LDL' = number of negative eigenvalues of A."""
    import refspr
    from sprgpu import fem

    sys_path = os.path.join(os.path.dirname(__file__), "golden")
    g = _load(os.path.join(sys_path, "hex8_3_indefinite.npz"))
    model = fem.hex8_grid(3)
    sysm = refspr.symbolic(model)
    # the fixture's matrix: element matrices K_e - 0.08 I
    A = dense_equation_matrix(model, sysm) - 0.08 * np.diag(_multiplicity(model, sysm))
    nneg = int(np.sum(np.linalg.eigvalsh(A) < 0))
    assert nneg == int(g["mspar_out"][25]) > 0
    ms = g["mspar_in"].copy()
    sm = g["sm_in"].copy()
    ierr, _ = oracle.sprsol(1, ms, g["mtrees"], g["msifa"], sm)
    assert ierr == 3 and ms[25] == nneg


def _multiplicity(model, sysm):
    m = np.zeros(sysm.neq)
    for e in range(model.nel):
        eq = sysm.meqn[model.element_dofs(e)]
        eq = eq[eq > 0] - 1
        m[eq] += 1
    return m


@pytest.mark.parametrize("kind,n", [("hex8", 4), ("quad4", 10), ("frame", 50)])
def test_manufactured_solution(kind, n):
    kw = {"seed": 2} if kind == "frame" else {}
    model, sysm, sm = make_system(kind, n, **kw)
    A = dense_equation_matrix(model, sysm)
    xs = np.random.default_rng(3).normal(size=(sysm.neq, 2))
    b = np.asfortranarray(A @ xs)
    ms = sysm.mspar.copy()
    ierr, _ = oracle.sprsol(3, ms, sysm.mtrees, sysm.msifa, sm, b)
    assert ierr == 0
    cond = np.linalg.cond(A)
    assert np.linalg.norm(b - xs) / np.linalg.norm(xs) < 50 * cond * np.finfo(float).eps


def test_sprprm_equals_dense_product():
    model, sysm, sm = make_system("hex8", 4, e_sigma=0.5, e_seed=1)
    A = dense_equation_matrix(model, sysm)
    x = np.random.default_rng(5).normal(size=(sysm.neq, 3))
    ierr, c = oracle.sprprm(sm[sysm.neq:], x, sysm.mspar.copy(), sysm.mtrees, sysm.msifa)
    assert ierr == 0
    np.testing.assert_allclose(c, A @ x, rtol=1e-12, atol=1e-12 * np.abs(A).max())
    ierr, c2 = oracle.sprprm(sm[sysm.neq:], x, sysm.mspar.copy(), sysm.mtrees, sysm.msifa,
                             iflag=-1)
    np.testing.assert_array_equal(c2, -c)


@pytest.mark.parametrize("kind,n,kw", [("hex8", 5, dict(e_sigma=0.5, e_seed=1)),
                                       ("quad4", 10, {}), ("frame", 50, dict(seed=3))])
def test_sprprm_bitexact_vs_reference(kind, n, kw):
    """The restated SPRPRM equals the reference's (oracle/_ref) bit for bit for
    every IFLAG and for a diagonal A (KSA <= 0), and fails for a planted
    change of the summation order (the column's own entries first)."""
    import refspr

    model, sysm, sm = make_system(kind, n, **kw)
    rng = np.random.default_rng(6)
    x = rng.normal(size=(sysm.neq, 3))
    a = sm[sysm.neq:]
    for ksa in (1, 0):
        aa = a if ksa > 0 else rng.normal(size=sysm.neq)
        for iflag in (1, -1, 2, -2):
            c0 = rng.normal(size=(sysm.neq, 3))
            ierr_r, c_r, wa_r = refspr.prm(aa, x, sysm.mspar.copy(), sysm.mtrees, sysm.msifa,
                                           iflag=iflag, ksa=ksa, c=np.asfortranarray(c0.copy()))
            c_o = np.asfortranarray(c0.copy())
            ierr_o, c_o = oracle.sprprm(aa, x, sysm.mspar.copy(), sysm.mtrees, sysm.msifa,
                                        iflag=iflag, c=c_o, ksa=ksa)
            assert ierr_r == ierr_o == 0
            assert np.array_equal(c_o.view(np.int64), c_r.view(np.int64)), (ksa, iflag)
    # planted mistake: each row summed with its own column's entries first
    A = dense_equation_matrix(model, sysm)
    ierr_r, c_r, _ = refspr.prm(a, x, sysm.mspar.copy(), sysm.mtrees, sysm.msifa)
    w = np.zeros_like(x)
    for r in range(sysm.neq):
        cols = np.nonzero(A[r])[0]
        for cc in np.concatenate([cols[cols >= r], cols[cols < r]]):
            w[r] += A[r, cc] * x[cc]
    assert not np.array_equal(w, c_r)


@pytest.mark.parametrize("kind,n,kw", [("hex8", 5, dict(e_sigma=0.5, e_seed=1)),
                                       ("quad4", 10, {}), ("frame", 50, dict(seed=3))])
def test_sprtmm_bitexact_vs_reference(kind, n, kw):
    """The restated SPRTMM equals the reference's bit for bit (IFLAG 1, 2, 11,
    22, a diagonal U), on the unfactored matrix and on the factor; it equals
    L^T B and L B of the dense lower triangle; a planted change (the
    diagonal term added last in IFLAG 1) is caught."""
    import refspr

    model, sysm, sm = make_system(kind, n, **kw)
    rng = np.random.default_rng(7)
    x = rng.normal(size=(sysm.neq, 2))
    fac = sm.copy()
    assert oracle.sprsol(1, sysm.mspar.copy(), sysm.mtrees, sysm.msifa, fac)[0] == 0
    for u in (sm[sysm.neq:], fac[sysm.neq:]):
        for ksu, uu in ((1, u), (0, rng.normal(size=sysm.neq))):
            for iflag in (1, 2):
                r = refspr.tmm(uu, x, sysm.mspar.copy(), sysm.mtrees, sysm.msifa, iflag, ksu)
                o = oracle.sprtmm(uu, x, sysm.mspar.copy(), sysm.mtrees, sysm.msifa, iflag, ksu)
                assert r[0] == o[0] == 0
                assert np.array_equal(r[1].view(np.int64), o[1].view(np.int64)), (ksu, iflag)
                assert np.array_equal(r[2].view(np.int64), o[2].view(np.int64))
            for iflag in (11, 22):
                r = refspr.tmm(uu, x[:, :1], sysm.mspar.copy(), sysm.mtrees, sysm.msifa, iflag, ksu)
                o = oracle.sprtmm(uu, x[:, :1], sysm.mspar.copy(), sysm.mtrees, sysm.msifa, iflag,
                                  ksu)
                assert r[0] == o[0] == 0
                assert np.array_equal(r[2].view(np.int64), o[2].view(np.int64))
    # mathematics: the dense lower triangle of the unfactored matrix
    A = dense_equation_matrix(model, sysm)
    Lt = np.tril(A)
    _, c1, _ = oracle.sprtmm(sm[sysm.neq:], x, sysm.mspar.copy(), sysm.mtrees, sysm.msifa, 1)
    _, c2, _ = oracle.sprtmm(sm[sysm.neq:], x, sysm.mspar.copy(), sysm.mtrees, sysm.msifa, 2)
    tol = 1e-12 * np.abs(A).max() * np.abs(x).max()
    np.testing.assert_allclose(c1, Lt.T @ x, rtol=1e-12, atol=tol)
    np.testing.assert_allclose(c2, Lt @ x, rtol=1e-12, atol=tol)
    # planted mistake: column sums with the diagonal term last
    w = np.zeros(sysm.neq)
    for j in range(sysm.neq):
        rows = np.nonzero(Lt[:, j])[0]
        for r_ in np.concatenate([rows[rows > j], rows[rows == j]]):
            w[j] += Lt[r_, j] * x[r_, 0]
    assert not np.array_equal(w, c1[:, 0])


def test_planted_mistake_is_caught():
    """A dropped Schur term in the factor must break the dense-LDL' check."""
    import refspr

    model = _single_front_model(12, np.random.default_rng(2))
    sysm = refspr.symbolic(model)
    sm = refspr.assemble(model, sysm)
    A = dense_equation_matrix(model, sysm)
    ms = sysm.mspar.copy()
    oracle.sprsol(1, ms, sysm.mtrees, sysm.msifa, sm)
    d = sm[:sysm.neq].copy()
    d_wrong = d.copy()
    d_wrong[-1] += 1e-6 * abs(d_wrong[-1])  # what a dropped rank-1 term would leave
    C = np.linalg.cholesky(A)
    assert np.allclose(d, np.diag(C) ** 2, rtol=1e-12)
    assert not np.allclose(d_wrong, np.diag(C) ** 2, rtol=1e-12)


def test_reference_cct_is_ldlt_scaled():
    """The reference's SPRFC2 factor C equals L sqrt(D) from its SPRFC1 factor
    (C C' = L D L'), and SPRFS2 / SPRBS2 are C^-1 and C^-T: the relations the
    GPU's CC' form is built on (checked here on the reference itself)."""
    import refspr

    model, sysm, sm = make_system("hex8", 5, e_sigma=0.5, e_seed=1)
    n = sysm.neq
    cc, ms = sm.copy(), sysm.mspar.copy()
    ierr, tol_c = refspr.fc2(sysm, cc, mspar=ms)
    assert ierr == 0 and ms[0] >= 0
    ld, ms1 = sm.copy(), sysm.mspar.copy()
    ierr1, _, tol_l = refspr.sol(1, sysm, ld, None, mspar=ms1)
    assert ierr1 == 0
    d = ld[:n]
    np.testing.assert_allclose(cc[:n], np.sqrt(d), rtol=1e-13)
    xs, xl = sysm.xsuper() - 1, sysm.xlindx() - 1
    scale = np.empty(sysm.nzerol)
    pos = 0
    for s in range(sysm.nsuper):
        f = xl[s + 1] - xl[s]
        for k in range(xs[s + 1] - xs[s]):
            j = xs[s] + k
            scale[pos:pos + f - k] = np.sqrt(d[j])
            scale[pos] = 1.0 / np.sqrt(d[j])          # diagonal: d -> sqrt(d)
            pos += f - k
    np.testing.assert_allclose(cc[n:], ld[n:] * scale, rtol=1e-11,
                               atol=1e-13 * np.abs(cc[n:]).max())
    np.testing.assert_allclose(tol_c[1:], tol_l[1:], rtol=1e-12)
    b = np.asfortranarray(np.random.default_rng(3).uniform(-1, 1, (n, 2)))
    y = b.copy(order="F")
    assert refspr.fs2_bs2("fs", sysm, cc, y, mspar=ms) == 0
    y1 = b.copy(order="F")
    assert oracle.sprsol(5, ms1.copy(), sysm.mtrees, sysm.msifa, ld, y1)[0] == 0  # forward only
    np.testing.assert_allclose(y, np.sqrt(d)[:, None] * y1, rtol=1e-11, atol=1e-13)
    assert refspr.fs2_bs2("bs", sysm, cc, y, mspar=ms) == 0
    x = b.copy(order="F")
    assert oracle.sprsol(3, sysm.mspar.copy(), sysm.mtrees, sysm.msifa, sm.copy(), x)[0] == 0
    np.testing.assert_allclose(y, x, rtol=1e-10, atol=1e-12 * np.abs(x).max())


def _dense_lower(sysm, lnz):
    """The packed supernodal lower triangle (SM(NEQ+1:) layout) as a dense matrix."""
    n = sysm.neq
    xs, xl = sysm.xsuper() - 1, sysm.xlindx() - 1
    rows_all = sysm.lindx() - 1
    Ld = np.zeros((n, n))
    pos = 0
    for s in range(sysm.nsuper):
        rows = rows_all[xl[s]:xl[s + 1]]
        for k in range(xs[s + 1] - xs[s]):
            Ld[rows[k:], xs[s] + k] = lnz[pos:pos + len(rows) - k]
            pos += len(rows) - k
    return Ld


def test_reference_sprfc4_rules():
    """The reference's SPRFC4 (sprfc4.f), the checker of the GPU's mode 4, on
    the rules the GPU restates: with TOL(1) = 0 it is SPRFC2 bit for bit;
    with replacements its C satisfies C C' = A + sum_j (DCOR(2,j) -
    DCOR(1,j)) e_j e_j' (a replaced pivot is a modified a_jj), the replaced
    values are TOL(1) max(1, |TOL(4) a_jj|) and MSPAR(27) = -count; on an
    indefinite matrix it fails at the first LDL' pivot below -TOL(1)|a_jj|
    (from the reference SPRFC1), with TOL(2:3) = (d, a_jj), DCOR = (d, d)."""
    import types

    import refspr

    model, sysm, sm = make_system("hex8", 4, e_sigma=0.5, e_seed=1)
    n = sysm.neq
    c4, ms4 = sm.copy(), sysm.mspar.copy()
    ierr, tol4, dcor = refspr.fc4(sysm, c4, (0.0, 0.0, 0.0, 0.0), mspar=ms4)
    c2, ms2 = sm.copy(), sysm.mspar.copy()
    assert ierr == 0 and refspr.fc2(sysm, c2, mspar=ms2)[0] == 0
    assert np.array_equal(c4, c2) and ms4[0] == ms2[0] == 6 and not dcor.any()
    # replacements
    A = dense_equation_matrix(model, sysm)
    tol = (0.6, 0.0, 0.0, 1.7)
    c4, ms4 = sm.copy(), sysm.mspar.copy()
    ierr, tolv, dcor = refspr.fc4(sysm, c4, tol, mspar=ms4)
    rep = dcor[:, 1] != 0
    assert ierr == -3 and rep.sum() > 0 and ms4[26] == -rep.sum() and ms4[25] == 0
    a = np.diag(A)
    np.testing.assert_array_equal(dcor[rep, 1], 0.6 * np.maximum(1.0, np.abs(1.7 * a[rep])))
    assert np.all(np.abs(dcor[rep, 0]) <= 0.6 * np.abs(a[rep]))
    Cd = _dense_lower(sysm, c4[n:])
    np.testing.assert_allclose(np.diag(Cd), c4[:n], rtol=0)
    Acorr = A + np.diag(np.where(rep, dcor[:, 1] - dcor[:, 0], 0.0))
    err = np.abs(Cd @ Cd.T - Acorr).max() / np.abs(A).max()
    assert err < 1e-12, err
    # a planted mistake (one correction left out) breaks the identity
    Aw = Acorr.copy()
    j = int(np.flatnonzero(rep)[0])
    Aw[j, j] -= dcor[j, 1] - dcor[j, 0]
    assert np.abs(Cd @ Cd.T - Aw).max() / np.abs(A).max() > 1e-6
    # the first negative pivot fails
    g = dict(np.load(os.path.join(os.path.dirname(__file__), "golden", "hex8_3_indefinite.npz"),
                     allow_pickle=False))
    s = types.SimpleNamespace(mspar=g["mspar_in"], mtrees=g["mtrees"], msifa=g["msifa"])
    c4, ms4 = g["sm_in"].copy(), g["mspar_in"].copy()
    ierr, tolv, dcor = refspr.fc4(s, c4, (1e-12, 0.0, 0.0, 1.0), mspar=ms4)
    ld, ms1 = g["sm_in"].copy(), g["mspar_in"].copy()
    ierr1, _ = oracle.sprsol(1, ms1, g["mtrees"], g["msifa"], ld)
    assert ierr1 == 3  # negative pivots, counted (NUMNEG)
    neq = int(g["mspar_in"][7])
    d = ld[:neq]
    adiag = _assembled_diagonal(s, g["sm_in"])
    first = int(np.flatnonzero(d < -1e-12 * np.abs(adiag))[0])
    assert ierr == -3 and ms4[26] == first + 1 and ms4[25] == 0
    np.testing.assert_allclose(tolv[1:3], [d[first], adiag[first]], rtol=1e-12)
    np.testing.assert_allclose(dcor[first], [d[first], d[first]], rtol=1e-12)


def _assembled_diagonal(sysm, sm):
    """a_jj of the assembled SM: the first entry of each packed column."""
    mspar = sysm.mspar
    ns, neq = int(mspar[10]), int(mspar[7])
    p45, p46, p48 = int(mspar[44]) - 1, int(mspar[45]) - 1, int(mspar[47]) - 1
    xsuper = sysm.mtrees[p45:p45 + ns + 1] - 1
    xlindx = sysm.mtrees[p46:p46 + ns + 1] - 1
    xlnz = sysm.msifa[p48:p48 + ns + 1] - 1
    a = np.empty(neq)
    for s in range(ns):
        f = xlindx[s + 1] - xlindx[s]
        pos = xlnz[s]
        for k in range(xsuper[s + 1] - xsuper[s]):
            a[xsuper[s] + k] = sm[neq + pos]
            pos += f - k
    return a


def _partial_cases(sysm):
    """LDB values for the partial solve: inside a supernode, at a supernode
    boundary, and one equation short."""
    xs = sysm.xsuper() - 1
    mid = int(xs[-3]) + 1 if xs[-2] - xs[-3] > 1 else int(xs[-3])
    return sorted({sysm.neq - 1, int(xs[-2]), mid, sysm.neq // 2})


@pytest.mark.parametrize("kind,n", [("hex8", 5), ("quad4", 14)])
def test_partial_solve_oracle_equals_reference(kind, n):
    """MSPAR(57) = 1 with LDB < NEQ (sprfs1.f:200-201, sprbs1.f:189): the
    restated SPRFS1/SPRBS1 equal the reference's bit for bit for IOP 4 (both
    sweeps), 5 (forward) and 6 (backward); IOP 4 solves the leading block,
    x1 = A11^-1 b1 (mathematics, dense); with MSPAR(57) = 0 LDB < NEQ is
    IERR = -2 in both."""
    import refspr

    model, sysm, sm = make_system(kind, n)
    A = dense_equation_matrix(model, sysm)
    fac = sm.copy()
    assert oracle.sprsol(1, sysm.mspar.copy(), sysm.mtrees, sysm.msifa, fac)[0] == 0
    rng = np.random.default_rng(11)
    for ldb in _partial_cases(sysm):
        mspar = sysm.mspar.copy()
        mspar[56] = 1
        b = np.asfortranarray(rng.uniform(-1, 1, (ldb, 2)))
        for iop in (4, 5, 6):
            xo, xr = b.copy(order="F"), b.copy(order="F")
            ierr_o, _ = oracle.sprsol(iop, mspar.copy(), sysm.mtrees, sysm.msifa, fac.copy(), xo)
            ierr_r, _, _ = refspr.sol(iop, sysm, fac.copy(), xr, mspar=mspar.copy())
            assert ierr_o == 0 and ierr_r == 0
            assert np.array_equal(xo.view(np.int64), xr.view(np.int64)), (ldb, iop)
            if iop == 4:
                x1 = np.linalg.solve(A[:ldb, :ldb], b)
                assert np.max(np.abs(xo - x1)) <= 1e-10 * np.max(np.abs(x1)), ldb
        xo, xr = b.copy(order="F"), b.copy(order="F")
        assert oracle.sprsol(4, sysm.mspar.copy(), sysm.mtrees, sysm.msifa, fac.copy(), xo)[0] == -2
        assert refspr.sol(4, sysm, fac.copy(), xr, mspar=sysm.mspar.copy())[0] == -2
