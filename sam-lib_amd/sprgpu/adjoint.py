"""Adjoint of K(p) u = f behind torch.autograd.Function.

K(p) = sum_e p_e K0_e is assembled from element matrices that are linear in
one parameter each (a Young's modulus, for instance).  The forward pass is
the drop-in's factor and solve (FP32 factor + FP64 iterative refinement, or
plain FP64).  The backward pass reuses the same factor: lambda = K^-1 dL/du
(K is symmetric), then dL/dp_e = -lambda_e^T K0_e u_e on the device
(sprgpu_element_grad) and dL/df = lambda.  The reference has no adjoint.

Assembly runs on the device (sprgpu_assemble): the unit-parameter element
matrices stay in HBM and SM for the current parameters is summed there, bit
for bit what the reference SPRADM loop gives (EM_e = p_e * K0_e, then the
element sums in SPRADM's order), so a forward pass moves no matrix over PCIe.
"""
from __future__ import annotations

import numpy as np

from . import host
from .device import SprFactor


class ElasticSystem:
    """A model with K = sum_e p_e K0_e, its symbolic structure and a factor handle.

    ``model`` is a :class:`sprgpu.fem.SamModel` whose ``em``/``em_off`` hold
    the unit-parameter element matrices (its ``escale`` is ignored: the
    parameters are passed to :meth:`solve`).
    """

    def __init__(self, model, precision: str = "fp32", rtol: float = 1e-10, maxit: int = 30):
        import torch

        self.torch = torch
        self.model = model
        self.rtol, self.maxit = rtol, maxit
        self.dev = torch.device("cuda:0")
        self.sysm = host.symbolic(model)
        self.neq = self.sysm.neq
        self.h = SprFactor(self.sysm.mspar, self.sysm.mtrees, self.sysm.msifa, precision)
        eq = host.element_equations(model, self.sysm)
        self.nedof = eq.shape[1]
        self.elem_eq = torch.from_numpy(eq).to(self.dev)
        offs, inv = np.unique(np.asarray(model.em_off), return_inverse=True)
        n2 = self.nedof * self.nedof
        k0 = np.stack([model.em[o:o + n2] for o in offs])
        self.k0_host = k0.reshape(-1, self.nedof, self.nedof).transpose(0, 2, 1)  # row-major view
        self.k0 = torch.from_numpy(k0).to(self.dev)
        self.k0_index = None if len(offs) == 1 else torch.from_numpy(inv.astype(np.int64)).to(self.dev)
        free = np.nonzero(self.sysm.meqn > 0)[0]
        self.free_dofs = torch.from_numpy(free).to(self.dev)
        self.free_eqs = torch.from_numpy(self.sysm.meqn[free] - 1).to(self.dev)
        self.p_factored = None  # parameters of the factor currently held by the handle
        # element assembly map (also the residual pattern) and the device inputs
        meen_ptr, meen = host.element_meen(model, self.sysm)
        self.h.set_assembly(meen_ptr, meen)
        self.em_dev = torch.from_numpy(np.ascontiguousarray(model.em)).to(self.dev)
        self.em_off_dev = torch.from_numpy(np.asarray(model.em_off, dtype=np.int64)).to(self.dev)
        self.sm = None

    # -- vector maps between natural dof order and SPR equation order ------
    def to_eq(self, v_dof):
        b = self.torch.zeros(self.neq, dtype=self.torch.float64, device=self.dev)
        b[self.free_eqs] = v_dof[self.free_dofs].to(self.torch.float64)
        return b

    def to_dofs(self, v_eq):
        u = self.torch.zeros(self.model.ndof, dtype=self.torch.float64, device=self.dev)
        u[self.free_dofs] = v_eq[self.free_eqs]
        return u

    # -- factor and solves ----------------------------------------------------
    def assemble(self, p):
        """SM for parameters p (numpy or a device tensor), assembled on the device."""
        torch = self.torch
        p_dev = torch.as_tensor(p, dtype=torch.float64).to(self.dev)
        if self.sm is None:
            self.sm = torch.empty(self.sysm.neq + self.sysm.nzerol, dtype=torch.float64,
                                  device=self.dev)
        self.h.assemble(self.em_dev, self.sm, self.em_off_dev, p_dev)
        return self.sm

    def factor(self, p: np.ndarray, sm_dev=None):
        sm_dev = self.assemble(p) if sm_dev is None else sm_dev
        info = self.h.factor(sm_dev)
        if info["ierr"] < 0:
            raise RuntimeError(f"factorization failed: {info}")
        self.p_factored = np.array(p, dtype=np.float64, copy=True)
        return info

    def solve_eq(self, b_eq):
        x = self.torch.empty_like(b_eq)
        info = self.h.solve_ir(b_eq, x, rtol=self.rtol, maxit=self.maxit)
        if not info["converged"]:
            raise RuntimeError(f"refinement did not converge: {info}")
        return x

    def element_grad(self, lam_eq, x_eq):
        return self.h.element_grad(self.elem_eq, self.k0, self.k0_index, lam_eq, x_eq)

    def close(self):
        self.h.close()

    # -- finite differences ---------------------------------------------------
    def element_sm_update(self, e: int, scale: float):
        """(positions, values): scale * K0_e on the stored lower triangle of SM."""
        q = self.elem_eq[e].cpu().numpy()
        k = self.k0_host[0 if self.k0_index is None else int(self.k0_index[e])]
        ii, jj = np.nonzero((q[:, None] >= 0) & (q[None, :] >= 0) & (q[:, None] >= q[None, :]))
        pos = host.lnz_positions(self.sysm, q[ii], q[jj])
        return pos, scale * k[ii, jj]

    def central_difference(self, p: np.ndarray, f_eq, e: int, h_rel: float = 1e-6) -> float:
        """Central difference quotient [L(p+h e_e) - L(p-h e_e)] / 2h of L = f^T u.

        Evaluated without subtractive cancellation: u- solves K(p-h e_e) u- = f;
        the difference d = u+ - u- solves K(p+h e_e) d = -2h K0_e u- exactly, so
        the quotient is f^T d / 2h.  (At 40^3, a per-element gradient is about
        L/64000; subtracting two computed values of L would leave a relative
        error near eps*64000/h = 7e-6 even with L exact to FP64.)  The two
        perturbed matrices are made from the assembled SM by adding +-h K0_e
        on the device.
        """
        torch = self.torch
        base = self.sm if self.sm is not None else self.assemble(p)
        h = h_rel * float(p[e])
        pos, val = self.element_sm_update(e, h)
        pos_t = torch.from_numpy(pos).to(self.dev)
        val_t = torch.from_numpy(val).to(self.dev)
        sm_m = base.clone()
        sm_m.index_add_(0, pos_t, -val_t)
        self.h.factor(sm_m)
        u_m = self.solve_eq(f_eq)
        sm_p = base.clone()
        sm_p.index_add_(0, pos_t, val_t)
        self.h.factor(sm_p)
        self.p_factored = None
        # rhs = -2h K0_e u-_e scattered to equations
        q = self.elem_eq[e]
        k = torch.from_numpy(np.ascontiguousarray(
            self.k0_host[0 if self.k0_index is None else int(self.k0_index[e])])).to(self.dev)
        ok = q >= 0
        ue = torch.where(ok, u_m[q.clamp(min=0)], torch.zeros_like(u_m[:1]))
        r = -2.0 * h * (k @ ue)
        rhs = torch.zeros_like(f_eq)
        rhs.index_add_(0, q[ok], r[ok])
        d = self.solve_eq(rhs)
        return float(torch.dot(f_eq, d)) / (2.0 * h)


def _solve_function():
    import torch

    class SprSolve(torch.autograd.Function):
        """u = K(p)^-1 f, natural dof order; differentiable in p and f."""

        @staticmethod
        def forward(ctx, p, f, system):
            p_np = p.detach().cpu().numpy().astype(np.float64)
            system.factor(p_np)
            x_eq = system.solve_eq(system.to_eq(f.detach()))
            ctx.system = system
            ctx.p_np = p_np
            ctx.save_for_backward(x_eq)
            return system.to_dofs(x_eq)

        @staticmethod
        def backward(ctx, grad_u):
            system = ctx.system
            (x_eq,) = ctx.saved_tensors
            if system.p_factored is None or not np.array_equal(system.p_factored, ctx.p_np):
                system.factor(ctx.p_np)  # the handle was refactored since the forward pass
            lam_eq = system.solve_eq(system.to_eq(grad_u.contiguous()))
            grad_p = system.element_grad(lam_eq, x_eq) if ctx.needs_input_grad[0] else None
            grad_f = system.to_dofs(lam_eq) if ctx.needs_input_grad[1] else None
            return grad_p, grad_f, None

    return SprSolve


_SprSolve = None


def spr_solve(p, f, system: ElasticSystem):
    """Differentiable u = K(p)^-1 f (torch tensors on the GPU, natural dof order)."""
    global _SprSolve
    if _SprSolve is None:
        _SprSolve = _solve_function()
    return _SprSolve.apply(p, f, system)
