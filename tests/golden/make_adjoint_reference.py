"""Writes hex8_<n>_adjoint_reference.npz: the FP64 CPU adjoint of the
compliance L = f^T u for hex8 n^3 elasticity with element moduli E log-normal
(0, 0.5) from seed 4 and loads U[-1,1] drawn after the moduli from the same
generator (natural dof order).

u comes from the reference SPRSOL (IOP=3) built from /root/reference, on one
core (about 20 minutes at n=40).  For compliance the adjoint variable is u
itself (K symmetric, dL/du = f), so dL/dE_e = -u_e^T K0 u_e with K0 the
unit-modulus element stiffness; that product is taken here in numpy.

    python tests/golden/make_adjoint_reference.py 40
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "sam-lib_amd"), os.path.join(ROOT, "oracle")]

import refspr  # noqa: E402
from sprgpu import fem  # noqa: E402

SEED = 4


def problem(n):
    model = fem.hex8_grid(n, e_sigma=0.5, e_seed=SEED)
    rng = np.random.default_rng(SEED)
    rng.normal(0.0, 0.5, model.nel)          # the moduli drawn by hex8_grid
    f = rng.uniform(-1.0, 1.0, model.ndof)   # then the loads
    return model, f


def element_dofs(model):
    nodes = model.mmnpc.reshape(model.nel, -1) - 1
    return (3 * nodes[:, :, None] + np.arange(3)[None, None, :]).reshape(model.nel, -1)


def main(n):
    model, f = problem(n)
    sysm = refspr.symbolic(model)
    sm = refspr.assemble(model, sysm)
    b = refspr.loads_to_equations(sysm, f)
    x = np.asfortranarray(b.copy())
    t = time.perf_counter()
    ierr, secs, tol = refspr.sol(3, sysm, sm, x)
    print(f"reference SPRSOL n={n}: ierr={ierr} {secs:.1f}s", flush=True)
    assert ierr == 0
    u = refspr.equations_to_dofs(sysm, x[:, 0])
    compliance = float(f @ u)
    k0 = model.em[:24 * 24].reshape(24, 24, order="F")
    ue = u[element_dofs(model)]
    grad = -np.einsum("ei,ij,ej->e", ue, k0, ue)
    np.savez_compressed(os.path.join(HERE, f"hex8_{n}_adjoint_reference.npz"),
                        x=x[:, 0], compliance=compliance, grad=grad)
    print(f"done in {time.perf_counter() - t:.1f}s  L={compliance:.10e}", flush=True)


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 40)
