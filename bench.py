#!/usr/bin/env python
"""Factor + solve benchmark: 40^3 hex8 elasticity, FP64 vs FP32 factor + FP64 refinement.

One step = numeric factorization of the assembled matrix (resident in HBM)
plus the solve of one right-hand side:
  * chosen form: FP32 factor on MFMA, then solve + FP64 iterative refinement
    (residual b - A x as an FP64 SpMV) until ||b - A x|| / ||b|| <= 1e-10;
  * plain form: FP64 factor + FP64 forward/backward solve.
value = factor flops (the reference's RINFO(3) count) / seconds of one
chosen-form step, in TFLOP/s.

Set-up outside the timed region, as the host application does it: the mesh,
the reference SPR symbolic phase and SPRADM assembly (sprgpu.host, the
reference's own routines compiled into lib/libspr_host.so), the upload of SM
and b to HBM, and the residual pattern.  The CPU baseline (the only use of
oracle/) is the reference's SPRFC1 + SPRFS1 + SPRBS1 on the 20^3 case, on
one host core.

After the timed steps, on the same 40^3 matrix and outside `value`: the
forward + backward sweeps alone per precision against the HBM roofline, the
device element assembly (bit equality with the host SPRADM loop's SM and its
time), the SPR-layout products (SPRPRM: A x of the FP64 solution against b;
SPRTMM: L^T x + L x - D x against A x) and the shift-invert Lanczos leg (the
10 lowest modes with a lumped mass), each with its own JSON object.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "sam-lib_amd"))

METRIC = "factor+solve s & TFLOP/s, 40^3 hex8: FP64 vs FP32+IR at 1e-10 residual"
PEAK = {"fp64": 78.6, "fp32": 157.3}  # MI355X dense MFMA TFLOP/s
GEMM_F32_TRAFFIC_40 = 470.5e6  # measured HBM bytes per k_gemm_x32<1> launch, 40^3
HBM_PEAK = 8000.0  # GB/s


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_problem(n):
    from sprgpu import fem, host

    model = fem.hex8_grid(n, e_sigma=0.5, e_seed=1)
    rng = np.random.default_rng(1)
    rng.normal(0.0, 0.5, model.nel)           # the moduli drawn by hex8_grid
    f = rng.uniform(-1.0, 1.0, model.ndof)    # then the load vector
    t = time.perf_counter()
    sysm = host.symbolic(model)
    t_sym = time.perf_counter() - t
    t = time.perf_counter()
    sm = host.assemble(model, sysm)
    t_asm = time.perf_counter() - t
    b = host.loads_to_equations(sysm, f)[:, 0].copy()
    log(f"set-up {n}^3: neq={sysm.neq} nsuper={sysm.nsuper} nzerol={sysm.nzerol} "
        f"mfront={sysm.mfront} flops={host.factor_flops(sysm):.4e} (symbolic {t_sym:.1f}s, "
        f"host SPRADM loop {t_asm:.1f}s)")
    return model, sysm, sm, b, t_asm


def bench_product(sysm, h, sm_dev, x, b_dev, reps):
    """SPRPRM on the device (sprgpu_sprprm) with the assembled A: A x for the
    FP64 solution must give b back to the solve's residual (a size-independent
    check of the full product), and the time per call from HIP events (the
    check that A's non-zeros lie on the pattern included, and apart).
    Algorithmic bytes of the row sums: per stored entry of the full symmetric
    pattern its column index, slot and value and the x entry (28 B), plus the
    row pointers and C (16 B per row); the check reads LNZ once (8 B per slot)."""
    import torch

    a = sm_dev[sysm.neq:]
    c = torch.empty_like(x)
    h.sprprm(a, x, c, iflag=1)
    torch.cuda.synchronize()
    rel = float(torch.linalg.norm(c - b_dev) / torch.linalg.norm(b_dev))
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        h.sprprm(a, x, c, iflag=1)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    log(f"device SPRPRM {ms:.3f} ms per call (pattern check included), |Ax - b|/|b| = {rel:.2e}")
    # SPRTMM with the stored lower triangle L of the same A (fill included):
    # L^T x + L x - diag(A) x must equal A x
    from sprgpu import host

    diag_pos = host.lnz_positions(sysm, np.arange(sysm.neq), np.arange(sysm.neq)) - sysm.neq
    dvec = a[torch.from_numpy(diag_pos).cuda()]
    out = {}
    for iflag in (1, 2):
        t = h.sprtmm(a, x, iflag=iflag)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(reps):
            h.sprtmm(a, x, t, iflag=iflag)
        e1.record()
        torch.cuda.synchronize()
        out[iflag] = (t, e0.elapsed_time(e1) / reps)
    d = h.sprtmm(dvec, x, iflag=1, ksu=0)
    sym = out[1][0] + out[2][0] - d
    rel_tmm = float(torch.linalg.norm(sym - c) / torch.linalg.norm(c))
    log(f"device SPRTMM L^T x {out[1][1]:.3f} ms, L x {out[2][1]:.3f} ms; "
        f"|L^T x + L x - D x - A x|/|A x| = {rel_tmm:.2e}")
    return {"ms_per_call": ms, "relres_of_fp64_solution": rel,
            "sprtmm_upper_ms": out[1][1], "sprtmm_lower_ms": out[2][1],
            "sprtmm_vs_sprprm_rel": rel_tmm, "lnz_bytes": 8 * sysm.nzerol}


def bench_solves(sysm, handles, b_dev, reps):
    """The forward + backward sweeps alone (the factor in HBM), one RHS, per
    precision: time per pair from HIP events on the handle's stream, and the
    HBM roofline: algorithmic bytes per sweep = LNZ once (NZEROL entries of
    the factor's type) + the right-hand side read and written in FP64
    (2 NEQ 8 B), two sweeps per pair, against 8 TB/s."""
    import torch

    out = {}
    for name, h, esz in handles:
        x = torch.empty_like(b_dev)
        per = {}
        for which, label in ((3, "pair"), (1, "forward"), (2, "backward")):
            x.copy_(b_dev)
            h.solve(x, which=which)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                x.copy_(b_dev)
                h.solve(x, which=which)
            e1.record()
            torch.cuda.synchronize()
            per[label] = e0.elapsed_time(e1) / reps
        sweep = sysm.nzerol * esz + 2 * sysm.neq * 8
        ach = 2 * sweep / (per["pair"] * 1e-3) / 1e9
        out[name] = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK, "unit": "GB/s",
                     "frac": ach / HBM_PEAK, "traffic": None,
                     "kernel": "forward + backward sweeps (k_fwd_*, k_bwd_*), 1 RHS",
                     "algorithmic_bytes_per_pair": 2 * sweep, "pair_ms": per["pair"],
                     "forward_ms": per["forward"], "backward_ms": per["backward"]}
        log(f"solves {name}: pair {per['pair']:.3f} ms (fwd {per['forward']:.3f}, bwd "
            f"{per['backward']:.3f}), {ach:.0f} GB/s = {ach / HBM_PEAK:.3f} of HBM")
    return out


def bench_lanczos(model, sysm, h, sm_dev, neval=10):
    """The NEVAL lowest modes of K q = lambda M q with a lumped mass (1/8 of
    each adjacent unit-density element per free dof), shift 0, by
    sprgpu_lanczos on the FP64 handle: the time (factor of K included), the
    converged count, and |K v - lambda M v| / |K v| of the first mode from
    the device SPRPRM products."""
    import torch
    from sprgpu import host

    mp, me = host.element_meen(model, sysm)
    mass = np.zeros(sysm.neq)
    np.add.at(mass, me[me > 0] - 1, 1.0 / 8.0)
    m_dev = torch.from_numpy(mass).cuda()
    torch.cuda.synchronize()
    t = time.perf_counter()
    evl, everr, v, nconv = h.lanczos(sm_dev, m_dev, 0.0, neval, nevec=1, ksb=0)
    torch.cuda.synchronize()
    secs = time.perf_counter() - t
    kv = h.sprprm(sm_dev[sysm.neq:], v)
    mv = h.sprprm(m_dev, v, ksa=0)
    res = float(torch.linalg.norm(kv - evl[0] * mv) / torch.linalg.norm(kv))
    log(f"Lanczos: {neval} lowest modes in {secs:.3f} s, {nconv} converged, "
        f"lambda_1 {evl[0]:.6e}, residual {res:.2e}")
    return {"seconds": secs, "neval": neval, "nconv": nconv, "eigenvalues": evl.tolist(),
            "first_mode_residual": res}


def bench_assembly(model, sysm, sm, h, reps, t_host):
    """Element assembly of the same SM on the device (sprgpu_set_assembly once,
    then sprgpu_assemble): bit equality with the host SPRADM loop's SM, and the
    time per assembly from HIP events on the handle's stream (the clearing of
    SM included).  Algorithmic bytes: the cleared SM, per LNZ slot with element
    entries its position, range and result (24 B), per element entry its map
    word, element offset, scale and matrix value (32 B).  Also the element
    matrices themselves made on the device (sprgpu_hex8_stiffness) and
    assembled: their time, and the SM's difference from the host's."""
    import torch
    from sprgpu import host

    t = time.perf_counter()
    mp, me = host.element_meen(model, sysm)
    h.set_assembly(mp, me)
    t_map = time.perf_counter() - t
    em = torch.from_numpy(model.em).cuda()
    off = torch.from_numpy(model.em_off).cuda()
    sc = torch.from_numpy(model.escale).cuda()
    out = torch.empty(sysm.neq + sysm.nzerol, dtype=torch.float64, device="cuda")
    h.assemble(em, out, off, sc)
    torch.cuda.synchronize()
    same = bool(torch.equal(out.view(torch.int64).cpu(), torch.from_numpy(sm).view(torch.int64)))
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        h.assemble(em, out, off, sc)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    # the element matrices themselves on the device (sprgpu_hex8_stiffness):
    # unit cubes with the moduli as E, assembled again; the SM equals the one
    # from base * scale to rounding (the element formula's own order)
    from sprgpu import fem
    xyz = torch.from_numpy(fem.hex8_coordinates(int(model.meta["n"]))).cuda()
    conn = torch.from_numpy(model.mmnpc).cuda()
    em_d = h.hex8_stiffness(xyz, conn, sc, float(model.meta["nu"]))
    h.assemble(em_d, out)
    torch.cuda.synchronize()
    el_rel = float(torch.max(torch.abs(out - torch.from_numpy(sm).cuda())) /
                   float(np.max(np.abs(sm))))
    e0.record()
    for _ in range(reps):
        h.hex8_stiffness(xyz, conn, sc, float(model.meta["nu"]), em_d)
    e1.record()
    torch.cuda.synchronize()
    el_ms = e0.elapsed_time(e1) / reps
    el_bytes = model.nel * (576 * 8 + 8 * 8 + 8 + 8 * 24)  # EM written; nodes, E, coordinates read
    del out, em_d
    log(f"device element matrices {el_ms:.3f} ms ({el_bytes / (el_ms * 1e-3) / 1e9:.0f} GB/s), "
        f"SM from them vs host: {el_rel:.1e}")
    nf = np.add.reduceat(me > 0, mp[:-1]).astype(np.int64)
    ncontrib = int(np.sum(nf * (nf + 1) // 2))
    nslots = int(np.count_nonzero(sm[sysm.neq:]))
    nbytes = 8 * (sysm.neq + sysm.nzerol) + 24 * nslots + 32 * ncontrib
    log(f"device assembly {ms:.3f} ms (bit-equal to SPRADM: {same}; map {t_map:.1f}s; "
        f"host SPRADM loop {t_host:.1f}s)")
    return {"device_ms": ms, "bit_equal_to_host_spradm": same, "map_setup_s": t_map,
            "host_spradm_loop_s": t_host, "element_entries": ncontrib, "lnz_slots": nslots,
            "algorithmic_bytes": nbytes, "achieved_GBps": nbytes / (ms * 1e-3) / 1e9,
            "frac_hbm": nbytes / (ms * 1e-3) / 8e12,
            "element_matrices": {"device_ms": el_ms, "algorithmic_bytes": el_bytes,
                                 "achieved_GBps": el_bytes / (el_ms * 1e-3) / 1e9,
                                 "frac_hbm": el_bytes / (el_ms * 1e-3) / 8e12,
                                 "sm_rel_diff_vs_host": el_rel}}


def cpu_baseline():
    """Reference SPRFC1/SPRFS1/SPRBS1 on the 20^3 case (E=1, nu=0.3, loads seed 0), 1 core."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import refspr  # the reference build, oracle/_ref: this leg only
    from sprgpu import fem

    model = fem.hex8_grid(20)
    sysm = refspr.symbolic(model)
    sm = refspr.assemble(model, sysm)
    b = refspr.loads_to_equations(sysm, fem.uniform_loads(model.ndof, np.random.default_rng(0)))
    b = np.asfortranarray(b)
    ierr, tf, _ = refspr.sol(1, sysm, sm, None)
    ierr2, ts, _ = refspr.sol(4, sysm, sm, b)
    assert ierr == 0 and ierr2 == 0
    flops = refspr.factor_flops(sysm)
    return {"value": flops / (tf + ts) / 1e12, "unit": "TFLOP/s", "cores": 1,
            "kind": "reference",
            "sample": (f"20^3 hex8 (NEQ {sysm.neq}, RINFO(3) {flops:.3e}): full reference "
                       f"SPRFC1 {tf:.2f}s + SPRFS1/SPRBS1 {ts:.3f}s, flang -O2, one thread"),
            "factor_s": tf, "solve_s": ts, "cpu_model": _cpu_model(),
            "nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0))}


def _cpu_model():
    """The host CPU's model name (/proc/cpuinfo), for the baseline's record."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=40, help="elements per edge of the hex8 cube")
    ap.add_argument("--rtol", type=float, default=1e-10)
    ap.add_argument("--maxit", type=int, default=30)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-profile", action="store_true", help="skip the per-kernel timing pass")
    ap.add_argument("--config", default="factor_solve",
                    choices=["factor_solve", "adjoint", "frames", "rhs128", "quad4", "hbm_copy"],
                    help="factor_solve: the headline (40^3, FP64 vs FP32+IR); adjoint: "
                         "dL/dE of the compliance at 40^3 (E seed 4); frames: 512 beam "
                         "frames, FP64 factor+solve in one graph; rhs128: 40^3 with 128 "
                         "N(0,1) RHS; quad4: 1000^2 plane stress with 128 RHS; hbm_copy: a "
                         "1 GiB device copy (calibrates rocprof FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--fd", type=int, default=64, help="adjoint: elements checked by finite differences")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the path's multi-GPU "
              "sharding is a design note, not a measured configuration.", file=sys.stderr)
        sys.exit(2)
    if args.config == "adjoint":
        return bench_adjoint(args)
    if args.config == "frames":
        return bench_frames(args)
    if args.config in ("rhs128", "quad4"):
        return bench_many_rhs(args)
    if args.config == "hbm_copy":
        import torch

        n = 1 << 27  # 1 GiB of float64
        a = torch.ones(n, dtype=torch.float64, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(args.steps):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"config": "hbm_copy", "bytes_read_per_copy": 8 * n,
                          "bytes_written_per_copy": 8 * n, "copies": args.steps}), flush=True)
        return

    import torch
    from sprgpu import host
    from sprgpu.device import SprFactor

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda:0")
    model, sysm, sm, b, t_host_asm = build_problem(args.n)
    flops = host.factor_flops(sysm)
    t = time.perf_counter()
    sm_dev = torch.from_numpy(sm).to(dev)
    b_dev = torch.from_numpy(b).to(dev).view(1, -1)
    h64 = SprFactor(sysm.mspar, sysm.mtrees, sysm.msifa, "fp64")
    h32 = SprFactor(sysm.mspar, sysm.mtrees, sysm.msifa, "fp32")
    h64.set_pattern(sm)
    h32.set_pattern(sm)
    torch.cuda.synchronize()
    log(f"device set-up {time.perf_counter() - t:.1f}s; device bytes fp64 {h64.device_bytes/1e9:.1f} GB "
        f"fp32 {h32.device_bytes/1e9:.1f} GB")

    x32 = torch.empty_like(b_dev)
    x64 = torch.empty_like(b_dev)
    tol = (0.0, 0.0, 0.0, 0.0)

    def step32():
        h32.factor(sm_dev, tol, info=False)
        return h32.solve_ir(b_dev, x32, rtol=args.rtol, maxit=args.maxit)

    def step64():
        h64.factor(sm_dev, tol, info=False)
        x64.copy_(b_dev)
        h64.solve(x64)

    def timed(fn, k):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = None
        for _ in range(k):
            out = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / k, out

    for _ in range(args.warmup):
        step64()
        step32()
    torch.cuda.synchronize()
    log("warm-up done")
    t32, ir = timed(step32, args.steps)
    t64, _ = timed(step64, args.steps)
    log(f"timed: fp32+IR {t32*1e3:.1f} ms/step ({ir}), fp64 {t64*1e3:.1f} ms/step")
    # split of each form into factor and solve (separately timed, outside the headline)
    tf32, _ = timed(lambda: h32.factor(sm_dev, tol, info=False), args.steps)
    tf64, _ = timed(lambda: h64.factor(sm_dev, tol, info=False), args.steps)
    info64 = h64.factor_info()
    info32 = h32.factor_info()
    assert info64["ierr"] == 0 and info32["ierr"] == 0, (info64, info32)

    # accuracy
    def relres(x):
        r = torch.empty_like(b_dev)
        h64.spmv(x, r)
        return float(torch.linalg.norm(b_dev - r) / torch.linalg.norm(b_dev))

    x64.copy_(b_dev)
    h64.solve(x64)
    rr64 = relres(x64)
    ir = h32.solve_ir(b_dev, x32, rtol=args.rtol, maxit=args.maxit)
    rr32 = relres(x32)
    err = {}
    ref_path = os.path.join(ROOT, "tests", "golden", f"hex8_{args.n}_reference_solution.npy")
    if os.path.exists(ref_path):
        xr = np.load(ref_path, allow_pickle=False)
        for name, x in (("fp64", x64), ("fp32_ir", x32)):
            xx = x.view(-1).cpu().numpy()
            err[name] = float(np.linalg.norm(xx - xr) / np.linalg.norm(xr))
    xd = (x32 - x64).view(-1)
    err_32_vs_64 = float(torch.linalg.norm(xd) / torch.linalg.norm(x64))

    result = {
        "metric": METRIC,
        "value": flops / t32 / 1e12,
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t32 * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 factor, f64 residual/update",
        "data": "synthetic",
        "config": {"workload": (f"hex8 {args.n}^3 linear elasticity, E log-normal(0,0.5) seed 1, "
                                f"nu 0.3, z=0 clamped, 1 RHS U[-1,1]"),
                   "neq": sysm.neq, "nsuper": sysm.nsuper, "nzerol": sysm.nzerol,
                   "factor_flops": flops, "rtol": args.rtol},
        "fp32_ir": {"step_s": t32, "factor_s": tf32, "solve_ir_s": t32 - tf32,
                    "iterations": int(ir["iterations"]), "relres": rr32,
                    "tflops_factor": flops / tf32 / 1e12,
                    "rel_err_vs_reference": err.get("fp32_ir"),
                    "rel_diff_vs_fp64_gpu": err_32_vs_64},
        "fp64": {"step_s": t64, "factor_s": tf64, "solve_s": t64 - tf64, "relres": rr64,
                 "tflops": flops / t64 / 1e12, "tflops_factor": flops / tf64 / 1e12,
                 "rel_err_vs_reference": err.get("fp64")},
        "speedup_fp32ir_vs_fp64": t64 / t32,
        "factor_info": {"numneg": info64["numneg"], "tol2": info64["tol2"],
                        "tol3": info64["tol3"]},
        "roofline": None,
        "cpu_baseline": None,
    }

    # the sweeps alone against the HBM roofline
    sol = bench_solves(sysm, (("fp64", h64, 8), ("fp32", h32, 4)), b_dev, max(args.steps, 10))
    result["roofline_solve_fp64"] = sol["fp64"]
    result["roofline_solve_fp32"] = sol["fp32"]

    # element assembly on the device (after the timed steps: it replaces the
    # FP64 handle's residual pattern by the map's, which that handle no longer uses)
    result["assembly"] = bench_assembly(model, sysm, sm, h64, max(args.steps, 5), t_host_asm)
    result["sprprm"] = bench_product(sysm, h64, sm_dev, x64, b_dev, max(args.steps, 5))
    result["lanczos"] = bench_lanczos(model, sysm, h64, sm_dev)

    if not args.no_profile:
        import ctypes as C
        from sprgpu import _lib

        out = np.zeros(27)
        tolv = np.zeros(4)
        classes = ["init", "extend_add", "small_lds", "diag_block", "rows_below", "mfma_update",
                   "other", "mfma_inner", "mfma_update_128"]
        per = {}
        for name, hh in (("fp32", h32), ("fp64", h64)):
            _lib.check(_lib.lib().sprgpu_profile_factor(
                hh._h, C.c_void_p(sm_dev.data_ptr()), tolv.ctypes.data_as(C.c_void_p), 1,
                out.ctypes.data_as(C.c_void_p)), "profile")
            per[name] = {c: {"ms": out[3 * i], "launches": int(out[3 * i + 1]),
                             "flops": out[3 * i + 2]} for i, c in enumerate(classes)}
        def mfma_roofline(prec, cls, kernel):
            g = per[prec][cls]
            ach = g["flops"] / (g["ms"] * 1e-3) / 1e12 if g["ms"] > 0 else 0.0
            return {"bound": "mfma", "achieved": ach, "peak": PEAK[prec], "unit": "TFLOP/s",
                    "frac": ach / PEAK[prec], "traffic": None, "kernel": kernel,
                    "launches": g["launches"], "avg_launch_ms": g["ms"] / max(g["launches"], 1)}

        result["roofline"] = mfma_roofline(
            "fp32", "mfma_update_128",
            "k_gemm_x32<1> (rank-256 and Schur update launches of at least 4096 64-tiles, "
            "128-row tiles, one factorization)")
        # HBM bytes per launch of k_gemm_x32<1> on this workload, from
        # rocprofv3 FETCH_SIZE and WRITE_SIZE runs scaled by a 1 GiB copy
        # (profiles/k_gemm_f32_hex8_40_hbm_traffic.txt); None at other sizes
        if args.n == 40:
            result["roofline"]["traffic"] = GEMM_F32_TRAFFIC_40
        result["roofline_update64_fp32"] = mfma_roofline(
            "fp32", "mfma_update", "k_gemm<float,64,1> (the smaller rank-256 and Schur update launches)")
        result["roofline_inner_fp32"] = mfma_roofline(
            "fp32", "mfma_inner", "k_gemm<float,64,0> (rank-128 updates inside a 256-column group)")
        result["roofline_fp64"] = mfma_roofline(
            "fp64", "mfma_update", "k_gemm<double,64,1> (rank-256 and Schur updates)")
        result["roofline_inner_fp64"] = mfma_roofline(
            "fp64", "mfma_inner", "k_gemm<double,64,0> (rank-128 updates inside a 256-column group)")
        result["kernel_classes_ms"] = {k: {c: round(v[c]["ms"], 3) for c in classes}
                                       for k, v in per.items()}
        log(f"profile: {result['kernel_classes_ms']}")

    if not args.no_cpu:
        log("cpu baseline (reference, 20^3, one core) ...")
        result["cpu_baseline"] = cpu_baseline()
        log(f"cpu baseline {result['cpu_baseline']['value']*1e3:.2f} GFLOP/s")

    print(json.dumps(result), flush=True)


def bench_adjoint(args):
    """Compliance L = f^T u on hex8 n^3, E log-normal(0,0.5) seed 4, loads U[-1,1]
    drawn after the moduli.  Forward: FP32 factor + FP64 refinement (rtol);
    backward: lambda from the same factor, then dL/dE_e on the device.  Timed
    per pass; the forward pass includes the device assembly of K(E).  Checked against cancellation-free
    central differences (h = 1e-6 E_e) on --fd random elements (seed 4) and,
    when tests/golden holds it, the reference's FP64 CPU adjoint."""
    import torch
    from sprgpu import fem
    from sprgpu.adjoint import ElasticSystem, spr_solve

    n, seed = args.n, 4
    model = fem.hex8_grid(n, e_sigma=0.5, e_seed=seed)
    rng = np.random.default_rng(seed)
    rng.normal(0.0, 0.5, model.nel)
    f = rng.uniform(-1.0, 1.0, model.ndof)
    t = time.perf_counter()
    sysobj = ElasticSystem(model, precision="fp32", rtol=args.rtol, maxit=args.maxit)
    E = model.escale.copy()
    sysobj.assemble(E)
    log(f"adjoint set-up {n}^3: neq={sysobj.neq} ({time.perf_counter() - t:.1f}s)")
    ft = torch.from_numpy(f).cuda()
    f_eq = sysobj.to_eq(ft)

    def forward():
        sysobj.factor(E)  # device assembly + factor
        return sysobj.solve_eq(f_eq)

    def backward(x_eq):
        lam = sysobj.solve_eq(f_eq)  # dL/du = f
        return sysobj.element_grad(lam, x_eq)

    x_eq = forward()
    g = backward(x_eq)
    torch.cuda.synchronize()
    tf = tb = 0.0
    for _ in range(args.steps):
        t0 = time.perf_counter()
        x_eq = forward()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        g = backward(x_eq)
        torch.cuda.synchronize()
        tf += t1 - t0
        tb += time.perf_counter() - t1
    tf /= args.steps
    tb /= args.steps
    L = float(torch.dot(f_eq, x_eq))
    gh = g.cpu().numpy()
    euler = abs(float(np.dot(E, gh)) + L) / abs(L)
    # the autograd binding gives the same gradient
    p = torch.from_numpy(E.copy()).cuda().requires_grad_(True)
    u = spr_solve(p, ft, sysobj)
    torch.dot(ft, u).backward()
    autograd_diff = float((p.grad - g).abs().max() / g.abs().max())
    log(f"forward {tf*1e3:.1f} ms, backward {tb*1e3:.1f} ms, L={L:.10e}, euler {euler:.2e}")
    # finite differences
    sel = np.random.default_rng(seed).choice(model.nel, args.fd, replace=False)
    t = time.perf_counter()
    fd_rel = []
    for e in sel:
        fd = sysobj.central_difference(E, f_eq, int(e), h_rel=1e-6)
        fd_rel.append(abs(fd - gh[e]) / abs(gh[e]))
    log(f"finite differences on {len(sel)} elements: max rel {max(fd_rel):.2e} "
        f"({time.perf_counter() - t:.1f}s)")
    ref = {}
    path = os.path.join(ROOT, "tests", "golden", f"hex8_{n}_adjoint_reference.npz")
    if os.path.exists(path):
        d = np.load(path, allow_pickle=False)
        ref["grad_max_rel"] = float(np.abs(gh - d["grad"]).max() / np.abs(d["grad"]).max())
        ref["compliance_rel"] = abs(L - float(d["compliance"])) / abs(float(d["compliance"]))
        xr = d["x"]
        ref["x_rel"] = float(np.linalg.norm(x_eq.cpu().numpy() - xr) / np.linalg.norm(xr))
    result = {
        "metric": f"adjoint dL/dE_e of the compliance, {n}^3 hex8 (E seed 4), FP32 factor + FP64 IR",
        "value": tf + tb, "unit": "s (forward + backward)", "n_gpus": 1, "steps": args.steps,
        "higher_is_better": False, "data": "synthetic",
        "config": {"workload": f"hex8 {n}^3, E log-normal(0,0.5) seed 4, loads U[-1,1]",
                   "neq": sysobj.neq, "elements": model.nel, "rtol": args.rtol},
        "forward_s": tf, "backward_s": tb, "compliance": L,
        "euler_identity_rel": euler, "autograd_vs_kernel_rel": autograd_diff,
        "fd": {"elements": int(len(sel)), "h_rel": 1e-6, "max_rel_diff": float(max(fd_rel)),
               "median_rel_diff": float(np.median(fd_rel))},
        "vs_fp64_cpu_adjoint": ref or None,
    }
    sysobj.close()
    print(json.dumps(result), flush=True)


def bench_many_rhs(args):
    """128 right-hand sides, columns i.i.d. N(0,1) from seed 2 (natural dof
    order), on the 40^3 hex8 matrix (rhs128) or on quad4 plane stress 1000^2
    (E=1, nu=0.25, thickness 1, left edge clamped; quad4).  FP32 factor +
    FP64 refinement of every column to rtol in one batched solve launch per
    kernel, beside the plain FP64 factor + solve."""
    import torch
    from sprgpu import fem, host
    from sprgpu.device import SprFactor

    nrhs = 128
    t = time.perf_counter()
    if args.config == "rhs128":
        model, sysm, sm, _, _ = build_problem(args.n)
        label = f"hex8 {args.n}^3, E log-normal(0,0.5) seed 1"
    else:
        n = 1000 if args.n == 40 else args.n
        model = fem.quad4_grid(n)
        sysm = host.symbolic(model)
        sm = host.assemble(model, sysm)
        label = f"quad4 plane stress {n}^2, E=1, nu=0.25, t=1, x=0 clamped"
    rng = np.random.default_rng(2)
    B = host.loads_to_equations(sysm, fem.normal_loads(model.ndof, rng, nrhs))
    flops = host.factor_flops(sysm)
    log(f"{label}: neq={sysm.neq} nzerol={sysm.nzerol} flops={flops:.3e} "
        f"({time.perf_counter() - t:.1f}s)")
    dev = torch.device("cuda:0")
    sm_dev = torch.from_numpy(sm).to(dev)
    b_dev = torch.from_numpy(np.ascontiguousarray(B.T)).to(dev)  # (nrhs, neq) = B(LDB, NRHS)
    h32 = SprFactor(sysm.mspar, sysm.mtrees, sysm.msifa, "fp32")
    h32.set_pattern(sm)
    x = torch.empty_like(b_dev)
    tol = (0.0, 0.0, 0.0, 0.0)

    def timed(fn, k):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = None
        for _ in range(k):
            out = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / k, out

    def step32():
        h32.factor(sm_dev, tol, info=False)
        return h32.solve_ir(b_dev, x, rtol=args.rtol, maxit=args.maxit)

    step32()
    t32, ir = timed(step32, args.steps)
    tf32, _ = timed(lambda: h32.factor(sm_dev, tol, info=False), args.steps)
    r = torch.empty_like(b_dev)
    h32.spmv(x, r)
    relres = ((b_dev - r).norm(dim=1) / b_dev.norm(dim=1)).max().item()
    log(f"fp32+IR: {t32*1e3:.1f} ms per step (factor {tf32*1e3:.1f}), {ir}, max relres {relres:.2e}")
    h32.close()
    h64 = SprFactor(sysm.mspar, sysm.mtrees, sysm.msifa, "fp64")
    x64 = torch.empty_like(b_dev)

    def step64():
        h64.factor(sm_dev, tol, info=False)
        x64.copy_(b_dev)
        h64.solve(x64)

    step64()
    t64, _ = timed(step64, args.steps)
    tf64, _ = timed(lambda: h64.factor(sm_dev, tol, info=False), args.steps)
    h64.set_pattern(sm)
    h64.spmv(x64, r)
    relres64 = ((b_dev - r).norm(dim=1) / b_dev.norm(dim=1)).max().item()
    h64.close()
    log(f"fp64: {t64*1e3:.1f} ms per step (factor {tf64*1e3:.1f}), max relres {relres64:.2e}")
    solve32 = t32 - tf32
    result = {
        "metric": f"factor + {nrhs}-RHS solve, {label}: FP32+IR vs FP64",
        "value": nrhs / solve32, "unit": "RHS/s (FP32+IR solve phase)", "n_gpus": 1,
        "steps": args.steps, "higher_is_better": True, "dtype": "f32 factor, f64 IR",
        "data": "synthetic",
        "config": {"workload": f"{label}, {nrhs} RHS N(0,1) seed 2", "neq": sysm.neq,
                   "nzerol": sysm.nzerol, "factor_flops": flops, "rtol": args.rtol},
        "fp32_ir": {"step_s": t32, "factor_s": tf32, "solve_ir_s": solve32,
                    "iterations": int(ir["iterations"]), "max_relres": relres},
        "fp64": {"step_s": t64, "factor_s": tf64, "solve_s": t64 - tf64, "max_relres": relres64},
    }
    print(json.dumps(result), flush=True)


def _frame_ref_solve(item):
    """Reference SPRSOL (factor + solve) of one frame: (seconds, x)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import refspr  # the reference build, oracle/_ref: CPU-baseline leg only

    sysm, sm, b = item
    x = np.asfortranarray(b[:, None].copy())
    ierr, secs, _ = refspr.sol(3, sysm, sm.copy(), x)
    assert ierr == 0
    return secs, x[:, 0]


def bench_frames(args):
    """512 random 3D Euler-Bernoulli frames (node counts U[80,900], seed 3, one
    node clamped each, one U[-1,1] RHS each) merged into one forest: FP64
    factor + solve of all of them as one captured hipGraph, with the schedule
    parameters chosen by sprgpu_autotune.  Checked against the reference's
    SPRSOL per frame, which is also the CPU baseline (one core, and all host
    cores with one frame per process)."""
    import ctypes as C
    from multiprocessing import Pool

    import torch
    from sprgpu import _lib, fem, host
    from sprgpu.device import SprFactor

    nsys = 512
    rng = np.random.default_rng(3)
    nodes = rng.integers(80, 901, nsys)
    t = time.perf_counter()
    models = [fem.beam_frame(int(n), rng) for n in nodes]
    systems = [host.symbolic(m) for m in models]
    sms = [host.assemble(m, s) for m, s in zip(models, systems)]
    bs = [host.loads_to_equations(s, fem.uniform_loads(m.ndof, rng))[:, 0].copy()
          for m, s in zip(models, systems)]
    merged = host.merge_systems(systems)
    sm = merged.merge_sm(sms)
    b = np.concatenate(bs)
    log(f"frames set-up: {nsys} systems, NEQ {min(s.neq for s in systems)}..{max(s.neq for s in systems)}, "
        f"total {merged.neq}, NZEROL {merged.nzerol} ({time.perf_counter() - t:.1f}s)")
    dev = torch.device("cuda:0")
    sm_dev = torch.from_numpy(sm).to(dev)
    b_dev = torch.from_numpy(b).to(dev).view(1, -1)
    x_dev = torch.empty_like(b_dev)
    h = SprFactor(merged.mspar, merged.mtrees, merged.msifa, "fp64")
    L = _lib.lib()
    report = np.zeros(256)
    x_dev.copy_(b_dev)
    t = time.perf_counter()
    _lib.check(L.sprgpu_autotune(h._h, C.c_void_p(sm_dev.data_ptr()), C.c_void_p(x_dev.data_ptr()),
                                 merged.neq, 1, report.ctypes.data_as(C.c_void_p)), "autotune")
    torch.cuda.synchronize()
    params = (C.c_int * 6)()
    L.sprgpu_plan_params(h._h, params)
    ncand = int(report[0])
    cands = report[1:1 + 6 * ncand].reshape(ncand, 6)
    level = cands[cands[:, 4] == 0]
    log(f"autotune: {ncand} candidates in {time.perf_counter() - t:.1f}s, chose forest={params[5]} "
        f"small_f={params[0]} waves={params[1]} narrow={params[2]}; ms best {cands[:, 5].min():.3f} "
        f"worst {cands[:, 5].max():.3f}, best level schedule {level[:, 5].min():.3f}")

    def run():
        x_dev.copy_(b_dev)
        _lib.check(L.sprgpu_factor_solve_graph(h._h, C.c_void_p(sm_dev.data_ptr()),
                                               C.c_void_p(x_dev.data_ptr()), merged.neq, 1),
                   "factor_solve_graph")

    for _ in range(max(args.warmup, 1)):
        run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        run()
    torch.cuda.synchronize()
    tg = (time.perf_counter() - t0) / args.steps
    info = h.factor_info()
    assert info["ierr"] == 0, info
    xg = x_dev.view(-1).cpu().numpy()
    # eager sequence for comparison (same kernels, launched one by one)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        x_dev.copy_(b_dev)
        h.factor(sm_dev, info=False)
        h.solve(x_dev)
    torch.cuda.synchronize()
    te = (time.perf_counter() - t0) / args.steps
    assert torch.equal(x_dev.view(-1).cpu(), torch.from_numpy(xg)), "graph and eager differ"
    log(f"graph {tg*1e3:.2f} ms per batch, eager {te*1e3:.2f} ms")
    result = {
        "metric": "512 beam frames (500-5400 dofs each): FP64 factor+solve of the batch",
        "value": tg, "unit": "s per batch", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "higher_is_better": False, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "512 3D Euler-Bernoulli frames, nodes U[80,900], seed 3, "
                               "one U[-1,1] RHS each, merged into one forest",
                   "neq_total": merged.neq, "nzerol_total": merged.nzerol,
                   "systems": nsys},
        "graph_s": tg, "eager_s": te,
        "autotune": {"candidates": ncand, "chosen": {"forest": params[5], "small_f": params[0],
                                                     "waves": params[1],
                                                     "narrow_threads": params[2]},
                     "best_ms": float(cands[:, 5].min()), "worst_ms": float(cands[:, 5].max()),
                     "best_level_schedule_ms": float(level[:, 5].min())},
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        items = list(zip(systems, sms, bs))
        t0 = time.perf_counter()
        outs = [_frame_ref_solve(it) for it in items]
        t1 = time.perf_counter() - t0
        ncores = min(16, os.cpu_count() or 1)
        t0 = time.perf_counter()
        with Pool(ncores) as pool:
            pool.map(_frame_ref_solve, items, chunksize=8)
        tp = time.perf_counter() - t0
        xr = np.concatenate([o[1] for o in outs])
        errs = [np.linalg.norm(a - r) / np.linalg.norm(r)
                for a, r in zip(merged.split(xg), merged.split(xr))]
        result["max_rel_err_vs_reference"] = float(max(errs))
        result["cpu_baseline"] = {
            "value": t1, "unit": "s per batch", "cores": 1, "kind": "reference",
            "sample": f"all {nsys} frames, reference SPRSOL(IOP=3) one after another on one core "
                      f"(sum of per-call {sum(o[0] for o in outs):.2f}s)",
            "all_cores": {"value": tp, "cores": ncores}}
        log(f"reference: {t1:.2f}s on one core, {tp:.2f}s on {ncores} processes; "
            f"max rel err {max(errs):.2e}")
    h.close()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
