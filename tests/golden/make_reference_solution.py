"""Writes hex8_<n>_reference_solution.npy: the reference SPR's FP64 solution of
the benchmark problem (hex8 n^3, E log-normal(0,0.5) seed 1, loads U[-1,1]
drawn after the moduli from the same generator), in SPR equation order.

The solution comes from running the reference (SPRSOL, IOP=3) built from
/root/reference; at n=40 this takes several minutes on one core.

    python tests/golden/make_reference_solution.py 40
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "sam-lib_amd"), os.path.join(ROOT, "oracle")]

import refspr  # noqa: E402
from bench import build_problem  # noqa: E402


def main(n):
    model, sysm, sm, b, _ = build_problem(n)
    x = np.asfortranarray(b[:, None].copy())
    t = time.perf_counter()
    ierr, secs, tol = refspr.sol(3, sysm, sm, x)
    print(f"reference SPRSOL n={n}: ierr={ierr} {secs:.1f}s tol={tol}", flush=True)
    assert ierr == 0
    np.save(os.path.join(HERE, f"hex8_{n}_reference_solution.npy"), x[:, 0])
    print(f"done in {time.perf_counter() - t:.1f}s")


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 40)
