"""sprgpu_solve_noalloc, the solve with its work space in the library's
static device scratch, against sprgpu_solve through the C ABI: the same
results bit for bit (the same kernels on other buffers), on two calls in a
row with the second smaller, with the permuted (PERI2E) layout, and the
refusal of a call whose work space does not fit."""
import ctypes as C

import numpy as np
import pytest

from conftest import gpu_available, make_system

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs an MI355X")]


def _noalloc(h, x, which=3):
    from sprgpu import _lib

    return int(_lib.lib().sprgpu_solve_noalloc(h._h, C.c_void_p(x.data_ptr()), h.neq,
                                               x.shape[0], which))


def _solve_both(h, B, which):
    import torch

    x_ref = B.clone()
    h.solve(x_ref, which=which)
    x = B.clone()
    rc = _noalloc(h, x, which)
    torch.cuda.synchronize()
    assert rc == 0, rc
    return x, x_ref


def test_noalloc_equals_solve_on_two_calls_the_second_smaller():
    import torch
    from sprgpu.device import SprFactor

    # hex8 12^3 has wide supernodes (the dataflow block chains, the ready
    # flags and partial sums); the second call is a smaller system with
    # fewer right-hand sides, on the same scratch
    cases = [("hex8", 12, 5, "fp64"), ("hex8", 7, 1, "fp32")]
    for kind, n, nrhs, prec in cases:
        _, sysm, sm = make_system(kind, n, e_sigma=0.5, e_seed=1)
        print(f"noalloc: {kind} {n}, neq {sysm.neq}, {nrhs} RHS, {prec}", flush=True)
        h = SprFactor(sysm.mspar, sysm.mtrees, sysm.msifa, prec)
        try:
            assert h.factor(torch.from_numpy(sm).cuda())["ierr"] == 0
            B = torch.from_numpy(np.random.default_rng(n).uniform(-1, 1, (nrhs, sysm.neq))).cuda()
            for which in (3, 1, 2):
                x, x_ref = _solve_both(h, B, which)
                assert torch.equal(x, x_ref), (kind, n, which)
        finally:
            h.close()


def test_noalloc_with_peri2e_and_a_refused_size():
    import torch
    from sprgpu import _lib
    from sprgpu.device import SprFactor
    from test_peri2e import perm_system

    _, sysm, sm, _, _ = perm_system(n=6)
    h = SprFactor(sysm.mspar, sysm.mtrees, sysm.msifa, "fp64")
    try:
        assert h.factor(torch.from_numpy(sm).cuda())["ierr"] == 0
        B = torch.from_numpy(np.random.default_rng(3).uniform(-1, 1, (2, sysm.neq))).cuda()
        x, x_ref = _solve_both(h, B, 3)
        assert torch.equal(x, x_ref)
        # a call beyond the scratch is refused before anything runs
        big = torch.zeros((20000, sysm.neq), dtype=torch.float64, device="cuda")
        assert _noalloc(h, big) == -2
        msg = _lib.last_error()
        print(msg, flush=True)
        assert "scratch holds" in msg
        assert torch.count_nonzero(big).item() == 0
    finally:
        h.close()
