"""The Fortran boundary, executed: sam-lib_amd/lib/spr_driver is a Fortran
program (INTEGER*8, as the reference's default build) that runs the
reference symbolic phase and SPRADM, then calls SPRSOL, SPRFC1 + SPRFS1 +
SPRBS1 and SPRFC3 by their reference names (sprsol.f:5-6, sprfc1.f:5-7,
sprfs1.f:5-7, sprbs1.f:5-6, sprfc3.f:5-7) from libsprgpu_f.so, the
ISO_C_BINDING layer over the C ABI.  Its outputs are compared with the
oracle (the C restatement) on the same model: IERR, MSPAR, TOL, DCOR
exactly or to 1e-9, the solutions to 1e-9.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle
from conftest import ROOT, gpu_available

DRIVER = os.path.join(ROOT, "sam-lib_amd", "lib", "spr_driver")
NSPAR = 60


def _write_model(path, model, f, tol1):
    """The driver's input file (see spr_driver.f90): SAM arrays, the element
    matrices with their moduli applied, loads in natural dof order."""
    I8 = np.int64
    em = []
    off = []
    pos = 0
    for e in range(model.nel):
        nd = int(model.nedof[e])
        m = model.em[model.em_off[e]:model.em_off[e] + nd * nd]
        if model.escale is not None:
            m = model.escale[e] * m
        em.append(m)
        off.append(pos)
        pos += nd * nd
    em = np.concatenate(em)
    f = np.asarray(f, dtype=np.float64).reshape(model.ndof, -1)
    with open(path, "wb") as fh:
        np.array([model.nanod, model.nel, model.ndof, model.mmnpc.size, em.size, f.shape[1]],
                 dtype=I8).tofile(fh)
        np.array([tol1], dtype=np.float64).tofile(fh)
        for a in (model.madof, model.msc, model.mpmnpc, model.mmnpc, model.nedof,
                  np.array(off)):
            np.asarray(a, dtype=I8).tofile(fh)
        em.tofile(fh)
        np.asfortranarray(f).ravel(order="F").tofile(fh)


def _read_out(path, ndof):
    raw = open(path, "rb").read()
    pos = 0

    def take(dtype, n):
        nonlocal pos
        a = np.frombuffer(raw, dtype=dtype, count=n, offset=pos)
        pos += a.nbytes
        return a.copy()

    neq, nrhs = take(np.int64, 2)
    out = {"neq": int(neq), "meqn": take(np.int64, ndof)}
    for k in (1, 2):
        out[f"ierr{k}"] = int(take(np.int64, 1)[0])
        out[f"ms{k}"] = take(np.int64, NSPAR)
        out[f"x{k}"] = take(np.float64, neq * nrhs).reshape(nrhs, neq).T
    out["tol2"] = take(np.float64, 3)
    out["ierr3"] = int(take(np.int64, 1)[0])
    out["ms3"] = take(np.int64, NSPAR)
    out["tol3"] = take(np.float64, 4)
    out["dcor"] = take(np.float64, 2 * neq).reshape(neq, 2)
    out["d3"] = take(np.float64, neq)
    assert pos == len(raw)
    return out


def _run(tmp_path, n=6, tol1=0.6):
    from sprgpu import fem

    model = fem.hex8_grid(n, e_sigma=0.5, e_seed=1)
    f = fem.uniform_loads(model.ndof, np.random.default_rng(0), 2)
    fin, fout = str(tmp_path / "model.bin"), str(tmp_path / "out.bin")
    _write_model(fin, model, f, tol1)
    print(f"spr_driver on hex8 {n}^3 (TOL(1) = {tol1} for SPRFC3)", flush=True)
    r = subprocess.run([DRIVER, fin, fout], capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0, r.returncode
    out = _read_out(fout, model.ndof)
    out["stdout"] = r.stdout
    return model, f, out


def test_driver_runs_the_reference_symbolic_phase(tmp_path):
    """Without a GPU too: the driver's MEQN is the reference symbolic phase's,
    and the drop-in calls return (IERR 0, or -5 where no GPU is present)."""
    from sprgpu import host

    if not os.path.exists(DRIVER):
        pytest.fail(f"{DRIVER} is not built")
    model, _, out = _run(tmp_path)
    sysm = host.symbolic(model)
    assert out["neq"] == sysm.neq
    assert np.array_equal(out["meqn"], sysm.meqn)
    assert out["ierr1"] in (0, -5) and out["ierr2"] in (0, -5)
    if out["ierr1"] == -5:  # the drop-in's message for a HIP error, on unit LPU
        assert "SPRFC1" in out["stdout"] and "GPU RUNTIME ERROR, IERR = -5" in out["stdout"]


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs an MI355X")
def test_driver_matches_the_oracle(tmp_path):
    from sprgpu import host

    model, f, out = _run(tmp_path)
    sysm = host.symbolic(model)
    sm = host.assemble(model, sysm)
    b = host.loads_to_equations(sysm, f)
    # SPRSOL(IOP=3)
    sm_o, x_o, ms_o = sm.copy(), np.asfortranarray(b.copy()), sysm.mspar.copy()
    ierr_o, tol_o = oracle.sprsol(3, ms_o, sysm.mtrees, sysm.msifa, sm_o, x_o)
    assert out["ierr1"] == ierr_o == 0
    assert np.array_equal(out["ms1"], ms_o)
    for k in (1, 2):
        rel = np.linalg.norm(out[f"x{k}"] - x_o) / np.linalg.norm(x_o)
        print(f"x{k}: {rel:.2e} against the oracle", flush=True)
        assert rel < 1e-9
    # SPRFC1 + SPRFS1 + SPRBS1: the same MSPAR (stage 7) and TOL
    assert out["ierr2"] == 0 and np.array_equal(out["ms2"], ms_o)
    np.testing.assert_allclose(out["tol2"][1:], tol_o[1:], rtol=1e-9)
    # SPRFC3
    ms3, sm3 = sysm.mspar.copy(), sm.copy()
    ierr3, tol3, dcor3 = oracle.sprfc3(ms3, sysm.mtrees, sysm.msifa, sm3, [0.6, 0.0, 0.0, 1.0])
    print(f"SPRFC3: driver IERR {out['ierr3']} MSPAR(27) {out['ms3'][26]}; oracle {ierr3} "
          f"{ms3[26]}", flush=True)
    assert out["ierr3"] == ierr3
    assert np.array_equal(out["ms3"], ms3)
    np.testing.assert_allclose(out["tol3"][1:3], tol3[1:3], rtol=1e-9)
    ref = np.asarray(dcor3).T  # the oracle returns DCOR as (2, NEQ)
    assert np.array_equal(out["dcor"] != 0, ref != 0)
    np.testing.assert_allclose(out["dcor"], ref, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(out["d3"], sm3[:sysm.neq], rtol=1e-9)
    # the messages on unit LPU: SPRER1's, with the list of replaced pivots
    txt = out["stdout"]
    if ierr3 == -3:
        assert "ERROR RETURN FROM  S A M / SPR  LIBRARY ROUTINE : SPRFC3" in txt
        assert "PIVOT VALUE REDUCED TO SIZE LESS THAN" in txt
        assert f"NUMBER OF SINGULARITIES FOUND   = {-int(ms3[26]):8d}" in txt
        listed = [int(l.split()[0]) for l in txt.splitlines()
                  if len(l.split()) == 3 and l.split()[0].isdigit()]
        assert listed == [int(i) + 1 for i in np.nonzero(ref[:, 1])[0]]
