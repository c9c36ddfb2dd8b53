"""ctypes binding of libsprgpu.so (the C ABI declared in include/sprgpu.h).

The library must have been built (``make -C sam-lib_amd``); there is no
fallback: every numeric call below runs the HIP kernels or raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(HERE)
LIB_DIR = os.path.join(PKG_ROOT, "lib")
LIB_PATH = os.path.join(LIB_DIR, "libsprgpu.so")
FLIB_PATH = os.path.join(LIB_DIR, "libsprgpu_f.so")

I64 = C.c_int64
P = C.c_void_p


class FactorInfo(C.Structure):
    _fields_ = [("ierr", I64), ("numneg", I64), ("supeqt", I64), ("tol2", C.c_double),
                ("tol3", C.c_double), ("nzero_diag", I64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class IrInfo(C.Structure):
    _fields_ = [("iterations", I64), ("relres", C.c_double), ("converged", I64)]


# name -> (restype, argtypes); every function of include/sprgpu.h
SIGNATURES = {
    "sprgpu_sprfc1": (None, [P] * 9 + [P, P]),
    "sprgpu_sprfc3": (None, [P] * 10 + [P, P]),
    "sprgpu_sprfc4": (None, [P] * 10 + [P, P]),
    "sprgpu_sprfs1": (None, [P] * 11),
    "sprgpu_sprbs1": (None, [P] * 10),
    "sprgpu_sprfc2": (None, [P] * 9 + [P, P]),
    "sprgpu_sprfs2": (None, [P] * 11),
    "sprgpu_sprbs2": (None, [P] * 10),
    "sprgpu_sprsol": (None, [P] * 13),
    "sprgpu_release": (None, [P]),
    "sprgpu_dropin_lierr": (I64, []),
    "sprgpu_create": (C.c_int, [P, P, P, C.c_int, C.POINTER(P)]),
    "sprgpu_destroy": (C.c_int, [P]),
    "sprgpu_set_stream": (C.c_int, [P, P]),
    "sprgpu_set_pattern": (C.c_int, [P, P]),
    "sprgpu_set_assembly": (C.c_int, [P, I64, P, P]),
    "sprgpu_assemble": (C.c_int, [P, P, P, P, P, C.c_int]),
    "sprgpu_set_assembly_ceq": (C.c_int, [P, I64, P, P, I64, P, P, I64, P, I64]),
    "sprgpu_assemble_ceq": (C.c_int, [P, P, P, P, P, P, P, C.c_int]),
    "sprgpu_hex8_stiffness": (C.c_int, [P, I64, I64, P, P, P, C.c_double, P]),
    "sprgpu_factor": (C.c_int, [P, P, P, C.c_int, C.POINTER(FactorInfo)]),
    "sprgpu_factor_info_get": (C.c_int, [P, C.POINTER(FactorInfo)]),
    "sprgpu_solve": (C.c_int, [P, P, I64, I64, C.c_int]),
    "sprgpu_solve_noalloc": (C.c_int, [P, P, I64, I64, C.c_int]),
    "sprgpu_solve_partial": (C.c_int, [P, P, I64, I64, C.c_int]),
    "sprgpu_spmv": (C.c_int, [P, P, P, I64, I64]),
    "sprgpu_sprprm": (C.c_int, [P, P, P, P, P, I64, I64, I64, I64]),
    "sprgpu_sprtra": (C.c_int, [P, P, P, P, P, I64, I64, I64]),
    "sprgpu_sprtmm": (C.c_int, [P, P, P, P, P, I64, I64, I64, I64]),
    "sprgpu_lanczos": (C.c_int, [P, P, P, C.c_int, C.c_double, I64, I64, I64, C.c_double, P, P,
                                  P, P]),
    "sprgpu_solve_ir": (C.c_int, [P, P, P, I64, I64, C.c_double, C.c_int, C.POINTER(IrInfo)]),
    "sprgpu_get_factor": (C.c_int, [P, P]),
    "sprgpu_get_dcor": (C.c_int, [P, P]),
    "sprgpu_neq": (I64, [P]),
    "sprgpu_nzerol": (I64, [P]),
    "sprgpu_device_bytes": (I64, [P]),
    "sprgpu_plan_summary": (C.c_int, [P, P, P, P]),
    "sprgpu_profile_factor": (C.c_int, [P, P, P, C.c_int, P]),
    "sprgpu_element_grad": (C.c_int, [P, I64, I64, P, P, P, P, P, P]),
    "sprgpu_pattern_grad": (C.c_int, [P, P, P, P]),
    "sprgpu_factor_solve_graph": (C.c_int, [P, P, P, I64, I64]),
    "sprgpu_autotune": (C.c_int, [P, P, P, I64, I64, P]),
    "sprgpu_plan_params": (C.c_int, [P, P]),
    "sprgpu_set_plan_params": (C.c_int, [P, P]),
    "sprgpu_last_error": (C.c_char_p, []),
}

_lib = None


def lib() -> C.CDLL:
    """Loads libsprgpu.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C sam-lib_amd`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def last_error() -> str:
    return lib().sprgpu_last_error().decode(errors="replace")


class SprGpuError(RuntimeError):
    pass


def check(rc: int, what: str) -> None:
    if rc != 0:
        raise SprGpuError(f"{what} failed ({rc}): {last_error()}")
