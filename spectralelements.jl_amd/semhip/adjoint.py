"""Adjoint linear solves: ``linsolve(a, problem, solver)`` of adjoint.jl:5-25
as a torch.autograd.Function (the counterpart of the reference's
``Zygote.@adjoint``), with the variable-coefficient diffusion problem of
semPS.jl:157-196 as the structured case whose parameter pullback runs on the
GPU.

Conventions (local arrays of shape (nr*Ex, ns*Ey)):

* ``problem(a)`` returns ``(lhs, rhs)``; ``solver(lhs, rhs, adj)`` returns the
  solution (adjoint.jl:5-8). The forward pass is exactly that.
* The backward pass solves ``lam = solver(lhs, ubar, True)`` and returns the
  pullback of ``resi(u, a) = rhs(a) - lhs(a)(u)`` at ``lam`` (adjoint.jl:10-25).
  ``lhs`` must provide ``param_pullback(u, lam)``, the gradient of
  ``<lam, lhs(a)(u)>`` with respect to ``a``; ``rhs`` may carry a torch
  autograd graph back to ``a`` (then its pullback is taken by torch with the
  mult-weighted pairing ``sum(mult .* lam .* rhs)``).
* ``ubar`` is the derivative with respect to the local copies of u; the
  adjoint source of the assembled system is ``QQ'(M .* ubar)`` (what
  ``diffusion_solver`` builds when ``adj`` is true).
"""
from __future__ import annotations

from typing import Callable, Optional

import torch

from .core import (Diagonal, Mesh, PCGInfo, SEMOperator, adjoint_dnu, combine, gatherScatter,
                   mask, norm_inf, pcg, to_device, vardiff_factors)


class VarDiffusionOperator(SEMOperator):
    """v -> M QQ'[Dr'(nu G11 vr + nu G12 vs) + (nu G12 vr + nu G22 vs) Ds]:
    the 12-argument lapl (lapl.jl:54-68) with nu folded into G
    (semPS.jl:161-165), as one fused device operator."""

    def __init__(self, nu, msh: Mesh, M, precision: str = "f64"):
        self.nu = to_device(nu)
        G = vardiff_factors(self.nu, msh)
        super().__init__(1.0, 0.0, msh, M, precision=precision, G=G)

    def param_pullback(self, u, lam):
        """d <lam, lhs(nu)(u)>_mult / d nu_local = B .* (grad lam . grad u)
        for continuous lam (the GPU kernel of the adjoint gradient)."""
        return adjoint_dnu(u, lam, self.msh)


class VarDiffusion:
    """problem(nu) -> (lhs, rhs) for -div(nu grad u) = f with homogeneous
    Dirichlet data: lhs = VarDiffusionOperator(nu), rhs = QQ'(M .* B .* f)
    (the mass form of semPS.jl:168-170, without dealiasing)."""

    def __init__(self, msh: Mesh, M, f):
        self.msh, self.M = msh, M
        self.rhs = combine(msh, [(1.0, to_device(f), True)], M=M, assemble=True)

    def __call__(self, nu):
        return VarDiffusionOperator(nu, self.msh, self.M), self.rhs


def diffusion_solver(tol: float = 1e-12, maxiter: Optional[int] = None,
                     precond: str = "jacobi", info: Optional[PCGInfo] = None) -> Callable:
    """solver(lhs, rhs, adj) for linsolve: fused PCG to norm(r, Inf) <=
    tol * norm(rhs, Inf). With adj true the cotangent is turned into the
    adjoint source QQ'(M .* rhs) first. precond: 'jacobi' or 'identity'."""

    def solver(lhs: SEMOperator, rhs, adj: bool):
        msh = lhs.msh
        if adj:
            rhs = gatherScatter(mask(to_device(rhs), lhs.M), msh)
        opM = Diagonal(lhs.jacobi()) if precond == "jacobi" else None
        nb = norm_inf(rhs)
        x = pcg(rhs, lhs, opM=opM, mult=msh.mult, tol=tol * nb, maxiter=maxiter, info=info)
        return x

    return solver


class _LinSolve(torch.autograd.Function):
    @staticmethod
    def forward(fctx, a, problem, solver):
        lhs, rhs = problem(a.detach())
        u = solver(lhs, rhs.detach() if isinstance(rhs, torch.Tensor) else rhs, False)
        fctx.problem, fctx.solver, fctx.lhs = problem, solver, lhs
        fctx.save_for_backward(a, u)
        return u

    @staticmethod
    def backward(fctx, ubar):
        a, u = fctx.saved_tensors
        lhs = fctx.lhs
        lam = fctx.solver(lhs, ubar, True)
        # pullback of resi(u, a) = rhs(a) - lhs(a)(u) at lam
        ga = -lhs.param_pullback(u, lam)
        with torch.enable_grad():
            a_ = a.detach().requires_grad_(True)
            _, rhs = fctx.problem(a_)
            if isinstance(rhs, torch.Tensor) and rhs.requires_grad:
                w = lam * lhs.msh.mult
                (gr,) = torch.autograd.grad(rhs, a_, grad_outputs=w, allow_unused=True)
                if gr is not None:
                    ga = ga + gr
        return ga, None, None


def linsolve(a, problem, solver):
    """u = solver(lhs, rhs, false) with (lhs, rhs) = problem(a)
    (adjoint.jl:5-8); differentiable in a through the adjoint solve
    (adjoint.jl:10-25)."""
    return _LinSolve.apply(a, problem, solver)


def nodal_gradient(g, msh: Mesh):
    """QQ' g: the derivative with respect to a continuous nodal field (each
    global node's value shared by its local copies), written to every copy."""
    return gatherScatter(to_device(g), msh)
