"""Device-resident factor handle (include/sprgpu.h, section 2).

Inputs and outputs are torch tensors on the GPU; torch only provides the
device memory and the stream.  Right-hand sides are (nrhs, neq) contiguous
float64 tensors, i.e. column-major (neq, nrhs) blocks with LDB = neq.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import FactorInfo, IrInfo, check, lib

FP64, FP32 = 0, 1


def _ptr(t) -> C.c_void_p:
    return C.c_void_p(t.data_ptr())


def _np(a):
    return a.ctypes.data_as(C.c_void_p)


class SprFactor:
    """A symbolic structure plus the GPU state to factor and solve with it."""

    def __init__(self, mspar, mtrees, msifa, precision: str = "fp64", stream=None):
        import torch

        if not torch.cuda.is_available():
            raise RuntimeError("SprFactor needs a GPU")
        self._torch = torch
        self.mspar = np.ascontiguousarray(mspar, dtype=np.int64)
        self.mtrees = np.ascontiguousarray(mtrees, dtype=np.int64)
        self.msifa = np.ascontiguousarray(msifa, dtype=np.int64)
        prec = {"fp64": FP64, "fp32": FP32}[precision]
        self.precision = precision
        h = C.c_void_p()
        check(lib().sprgpu_create(_np(self.mspar), _np(self.mtrees), _np(self.msifa), prec,
                                  C.byref(h)), "sprgpu_create")
        self._h = h
        self.neq = int(lib().sprgpu_neq(h))
        self.nzerol = int(lib().sprgpu_nzerol(h))
        self.set_stream(stream)

    def set_stream(self, stream=None):
        torch = self._torch
        s = stream if stream is not None else torch.cuda.current_stream()
        self._stream = s
        check(lib().sprgpu_set_stream(self._h, C.c_void_p(s.cuda_stream)), "sprgpu_set_stream")

    def close(self):
        if getattr(self, "_h", None):
            lib().sprgpu_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def device_bytes(self) -> int:
        return int(lib().sprgpu_device_bytes(self._h))

    _PARAM_NAMES = ("small_f", "small_waves", "narrow_threads", "big_tile", "block_factor",
                    "forest")

    def plan_params(self) -> dict:
        """Schedule parameters in use (sprgpu_plan_params)."""
        out = (C.c_int * len(self._PARAM_NAMES))()
        check(lib().sprgpu_plan_params(self._h, out), "sprgpu_plan_params")
        return dict(zip(self._PARAM_NAMES, list(out)))

    def set_plan_params(self, **kw):
        """Re-plan with some schedule parameters changed (sprgpu_set_plan_params)."""
        cur = self.plan_params()
        unknown = set(kw) - set(self._PARAM_NAMES)
        if unknown:
            raise ValueError(f"unknown plan parameters {sorted(unknown)}")
        cur.update(kw)
        arr = (C.c_int * len(self._PARAM_NAMES))(*[int(cur[k]) for k in self._PARAM_NAMES])
        check(lib().sprgpu_set_plan_params(self._h, arr), "sprgpu_set_plan_params")

    def set_pattern(self, sm_host: np.ndarray):
        sm_host = np.ascontiguousarray(sm_host, dtype=np.float64)
        check(lib().sprgpu_set_pattern(self._h, _np(sm_host)), "sprgpu_set_pattern")

    def set_assembly(self, meen_ptr: np.ndarray, meen: np.ndarray):
        """Element connectivity for ``assemble``: 1-based equations (SPREEQ's
        MEEN, 0 = none) of element e's dofs at meen[meen_ptr[e]:meen_ptr[e+1]].
        Also sets the residual pattern."""
        meen_ptr = np.ascontiguousarray(meen_ptr, dtype=np.int64)
        meen = np.ascontiguousarray(meen, dtype=np.int64)
        check(lib().sprgpu_set_assembly(self._h, meen_ptr.size - 1, _np(meen_ptr), _np(meen)),
              "sprgpu_set_assembly")

    def assemble(self, em_dev, sm_dev, em_off_dev=None, escale_dev=None,
                 accumulate: bool = False):
        """SM (device, NEQ+NZEROL float64) from the element matrices on the
        device, bit for bit what a host loop of SPRADM gives."""
        opt = (lambda t: None if t is None else _ptr(t))  # optional arrays: NULL
        check(lib().sprgpu_assemble(self._h, _ptr(em_dev), opt(em_off_dev), opt(escale_dev),
                                    _ptr(sm_dev), int(bool(accumulate))), "sprgpu_assemble")
        return sm_dev

    def hex8_stiffness(self, xyz_dev, mmnpc_dev, emod_dev=None, nu: float = 0.3, em_dev=None):
        """Element matrices of hex8 elements on the device: xyz_dev (nanod, 3)
        float64, mmnpc_dev (nel*8,) int64 1-based nodes, emod_dev (nel,) or
        None.  Returns em_dev (nel*576,) float64, the column-major 24x24
        matrices one after the other."""
        torch = self._torch
        nel = mmnpc_dev.numel() // 8
        if em_dev is None:
            em_dev = torch.empty(nel * 576, dtype=torch.float64, device=xyz_dev.device)
        opt = (lambda t: None if t is None else _ptr(t))
        check(lib().sprgpu_hex8_stiffness(self._h, nel, xyz_dev.shape[0], _ptr(xyz_dev),
                                          _ptr(mmnpc_dev), opt(emod_dev), float(nu),
                                          _ptr(em_dev)), "sprgpu_hex8_stiffness")
        return em_dev

    def set_assembly_ceq(self, meen_ptr: np.ndarray, meen: np.ndarray, mpmceq: np.ndarray,
                         mmceq: np.ndarray, meqn: np.ndarray, nsv: int = 0):
        """``set_assembly`` with constraint equations: meen may hold -ICEQ
        (SPREEQ's MEEN), MPMCEQ/MMCEQ/MEQN are SAM's arrays (1-based), nsv is
        SPRADM's NSV (0 SM, 1 SM+SV, -1 SV, |nsv|>1 one SV per prescribed dof)."""
        a = [np.ascontiguousarray(x, dtype=np.int64) for x in (meen_ptr, meen, mpmceq, mmceq, meqn)]
        check(lib().sprgpu_set_assembly_ceq(self._h, a[0].size - 1, _np(a[0]), _np(a[1]),
                                            a[2].size - 1, _np(a[2]), _np(a[3]), a[4].size,
                                            _np(a[4]), int(nsv)),
              "sprgpu_set_assembly_ceq")

    def assemble_ceq(self, em_dev, ttcc_dev, sm_dev=None, sv_dev=None, em_off_dev=None,
                     escale_dev=None, accumulate: bool = False):
        """SM and/or SV (device float64) with the constraint equations of
        ``set_assembly_ceq``; TTCC (device) is read at this call."""
        opt = (lambda t: None if t is None else _ptr(t))
        check(lib().sprgpu_assemble_ceq(self._h, _ptr(em_dev), opt(em_off_dev), opt(escale_dev),
                                        opt(ttcc_dev), opt(sm_dev), opt(sv_dev),
                                        int(bool(accumulate))), "sprgpu_assemble_ceq")
        return sm_dev, sv_dev

    def factor(self, sm_dev, tol=(0.0, 0.0, 0.0, 0.0), mode: int = 1, info: bool = True):
        tolv = np.zeros(4)
        tolv[:len(tol)] = tol
        if info:
            fi = FactorInfo()
            check(lib().sprgpu_factor(self._h, _ptr(sm_dev), _np(tolv), mode, C.byref(fi)),
                  "sprgpu_factor")
            return fi.as_dict()
        check(lib().sprgpu_factor(self._h, _ptr(sm_dev), _np(tolv), mode, None), "sprgpu_factor")
        return None

    def dcor(self) -> np.ndarray:
        """DCOR(2,NEQ) of the last mode-3 factorization, as (neq, 2) (sprgpu_get_dcor)."""
        out = np.zeros(2 * self.neq)
        check(lib().sprgpu_get_dcor(self._h, _np(out)), "sprgpu_get_dcor")
        return out.reshape(self.neq, 2)

    def factor_info(self) -> dict:
        fi = FactorInfo()
        check(lib().sprgpu_factor_info_get(self._h, C.byref(fi)), "sprgpu_factor_info_get")
        return fi.as_dict()

    def solve(self, b_dev, which: int = 3):
        """In-place forward (1), backward (2) or both (3) on b_dev (nrhs, neq)."""
        b2 = b_dev if b_dev.dim() == 2 else b_dev.view(1, -1)
        check(lib().sprgpu_solve(self._h, _ptr(b2), self.neq, b2.shape[0], which), "sprgpu_solve")
        return b_dev

    def solve_partial(self, b_dev, which: int = 3):
        """The partial solve of MSPAR(57) = 1: b_dev (nrhs, ldb) holds the
        leading ldb < NEQ equations only (the others are zero in the sweeps
        and dropped); which 3 gives A11^-1 b1."""
        b2 = b_dev if b_dev.dim() == 2 else b_dev.view(1, -1)
        check(lib().sprgpu_solve_partial(self._h, _ptr(b2), b2.shape[1], b2.shape[0], which),
              "sprgpu_solve_partial")
        return b_dev

    def spmv(self, x_dev, y_dev):
        x2 = x_dev if x_dev.dim() == 2 else x_dev.view(1, -1)
        y2 = y_dev if y_dev.dim() == 2 else y_dev.view(1, -1)
        check(lib().sprgpu_spmv(self._h, _ptr(x2), _ptr(y2), self.neq, x2.shape[0]),
              "sprgpu_spmv")
        return y_dev

    def sprprm(self, a_dev, b_dev, c_dev=None, iflag: int = 1, ksa: int = 1, wa_dev=None):
        """SPRPRM on the device: C = +-A B (iflag +-1), C +-= A B (+-2), or
        WA = A B (88, one column).  a_dev: LNZ (NZEROL) or a diagonal (NEQ,
        ksa <= 0); b_dev, c_dev: (nrhs, neq) float64."""
        torch = self._torch
        b2 = b_dev if b_dev.dim() == 2 else b_dev.view(1, -1)
        if c_dev is None and iflag != 88:
            c_dev = torch.zeros_like(b2)
        c2 = None if c_dev is None else (c_dev if c_dev.dim() == 2 else c_dev.view(1, -1))
        opt = (lambda t: None if t is None else _ptr(t))
        check(lib().sprgpu_sprprm(self._h, _ptr(a_dev), _ptr(b2), opt(c2), opt(wa_dev),
                                  b2.shape[1], b2.shape[0], int(ksa), int(iflag)),
              "sprgpu_sprprm")
        return c_dev if iflag != 88 else wa_dev

    def sprtra(self, a_dev, b_dev, ksa: int = 1):
        """SPRTRA on the device: C = B^T A B (n x n) for b_dev (n, m) float64 (the
        n columns of B(M, N)), A as in sprprm.  Returns (C as an (n, n) tensor,
        WA = A b of the last column)."""
        torch = self._torch
        b2 = b_dev if b_dev.dim() == 2 else b_dev.view(1, -1)
        n, m = b2.shape
        c = torch.empty(n, n, dtype=torch.float64, device=b2.device)
        wa = torch.empty(m, dtype=torch.float64, device=b2.device)
        check(lib().sprgpu_sprtra(self._h, _ptr(a_dev), _ptr(b2), _ptr(c), _ptr(wa), m, n,
                                  int(ksa)), "sprgpu_sprtra")
        return c.T, wa

    def sprtmm(self, u_dev, b_dev, c_dev=None, iflag: int = 1, ksu: int = 1, wa_dev=None):
        """SPRTMM on the device: C = L^T B (iflag 1) or L B (2) with L the stored
        lower triangle of u_dev (LNZ, or a diagonal when ksu <= 0); 11/22 with
        one column give the result in wa_dev only."""
        torch = self._torch
        b2 = b_dev if b_dev.dim() == 2 else b_dev.view(1, -1)
        if c_dev is None and iflag in (1, 2):
            c_dev = torch.empty_like(b2)
        c2 = None if c_dev is None else (c_dev if c_dev.dim() == 2 else c_dev.view(1, -1))
        opt = (lambda t: None if t is None else _ptr(t))
        check(lib().sprgpu_sprtmm(self._h, _ptr(u_dev), _ptr(b2), opt(c2), opt(wa_dev),
                                  b2.shape[1], b2.shape[0], int(ksu), int(iflag)),
              "sprgpu_sprtmm")
        return c_dev if iflag in (1, 2) else wa_dev

    def lanczos(self, a_sm_dev, b_dev, shift: float, neval: int, nevec: int = 0,
                maxlan: int | None = None, tol: float = 1e-10, ksb: int = 1):
        """Eigenvalues of (A - lambda B) q = 0 nearest to shift (sprgpu_lanczos).
        Returns (evl, everr, v (neq x nevec tensor or None), nconv)."""
        torch = self._torch
        maxlan = maxlan or min(self.neq, max(6 * neval, neval + 40))
        evl, everr = np.zeros(neval), np.zeros(neval)
        v = torch.empty(max(nevec, 1), self.neq, dtype=torch.float64, device=a_sm_dev.device)
        nconv = C.c_int64(0)
        check(lib().sprgpu_lanczos(self._h, _ptr(a_sm_dev), _ptr(b_dev), int(ksb), float(shift),
                                   int(neval), int(nevec), int(maxlan), float(tol), _np(evl),
                                   _np(everr), _ptr(v), C.byref(nconv)), "sprgpu_lanczos")
        return evl, everr, (v[:nevec] if nevec else None), int(nconv.value)

    def solve_ir(self, b_dev, x_dev, rtol: float = 1e-10, maxit: int = 20) -> dict:
        b2 = b_dev if b_dev.dim() == 2 else b_dev.view(1, -1)
        x2 = x_dev if x_dev.dim() == 2 else x_dev.view(1, -1)
        info = IrInfo()
        check(lib().sprgpu_solve_ir(self._h, _ptr(b2), _ptr(x2), self.neq, b2.shape[0],
                                    float(rtol), int(maxit), C.byref(info)), "sprgpu_solve_ir")
        return {"iterations": info.iterations, "relres": info.relres,
                "converged": bool(info.converged)}

    def get_factor(self, out=None):
        """The factor in SM layout (numpy on the host, or a device tensor)."""
        if out is None:
            out = np.zeros(self.neq + self.nzerol)
        if isinstance(out, np.ndarray):
            check(lib().sprgpu_get_factor(self._h, _np(out)), "sprgpu_get_factor")
        else:
            check(lib().sprgpu_get_factor(self._h, _ptr(out)), "sprgpu_get_factor")
        return out

    def element_grad(self, elem_eq_dev, k0_dev, k0_index_dev, lam_dev, x_dev, out=None):
        """grad[e] = -lambda_e^T K0_e x_e on the device (sprgpu_element_grad).

        elem_eq_dev: (nel, nedof) int64 equations (-1: no equation); k0_dev:
        (nmat, nedof, nedof) float64 column-major matrices; k0_index_dev: (nel,)
        int64 or None (one shared matrix)."""
        torch = self._torch
        nel, nedof = elem_eq_dev.shape
        if out is None:
            out = torch.empty(nel, dtype=torch.float64, device=lam_dev.device)
        kidx = C.c_void_p(0) if k0_index_dev is None else _ptr(k0_index_dev)
        check(lib().sprgpu_element_grad(self._h, nel, nedof, _ptr(elem_eq_dev), _ptr(k0_dev), kidx,
                                        _ptr(lam_dev), _ptr(x_dev), _ptr(out)),
              "sprgpu_element_grad")
        return out

    def pattern_grad(self, lam_dev, x_dev, out=None):
        """dL/dA on the stored lower triangle in SM layout (sprgpu_pattern_grad)."""
        torch = self._torch
        if out is None:
            out = torch.zeros(self.neq + self.nzerol, dtype=torch.float64, device=lam_dev.device)
        check(lib().sprgpu_pattern_grad(self._h, _ptr(lam_dev), _ptr(x_dev), _ptr(out)),
              "sprgpu_pattern_grad")
        return out
