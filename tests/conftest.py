import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "sam-lib_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


_CACHE = {}


def make_system(kind: str, n: int, **kw):
    """(model, SprSystem, assembled SM) from the reference symbolic phase; cached."""
    import refspr
    from sprgpu import fem

    key = (kind, n, tuple(sorted(kw.items())))
    if key not in _CACHE:
        if kind == "hex8":
            model = fem.hex8_grid(n, **kw)
        elif kind == "hex8c":
            model = fem.hex8_constrained(n, **kw)
        elif kind == "quad4":
            model = fem.quad4_grid(n, **kw)
        elif kind == "frame":
            seed = kw.pop("seed", 0)
            model = fem.beam_frame(n, np.random.default_rng(seed), **kw)
        else:
            raise ValueError(kind)
        sysm = refspr.symbolic(model)
        sm = refspr.assemble(model, sysm)
        _CACHE[key] = (model, sysm, sm)
    model, sysm, sm = _CACHE[key]
    return model, sysm, sm.copy()


@pytest.fixture
def system():
    return make_system


def gpu_available() -> bool:
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


def backward_error(sysm, sm, x, b):
    """Normwise backward error ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf)
    of a solution x (SPR equation order) of A x = b, A given by SM's LNZ part.
    A x and the row sums of |A| come from the reference SPRPRM built from
    source (oracle/_ref), the checker's own product."""
    import refspr

    lnz = sm[sysm.neq:]
    _, ax, _ = refspr.prm(lnz, x, sysm.mspar, sysm.mtrees, sysm.msifa)
    _, rows, _ = refspr.prm(np.abs(lnz), np.ones(sysm.neq), sysm.mspar, sysm.mtrees,
                            sysm.msifa)
    r = b - ax[:, 0]
    return float(np.abs(r).max() / (rows.max() * np.abs(x).max() + np.abs(b).max()))
