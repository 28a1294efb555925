"""The operator interface of the reference's Spectral submodule
(Spectral/Operators.jl:1-209, Spectral/OperatorBasics.jl) over the device
kernels, so that time integrators and linear solvers written against it run
on the GPU path:

    DiagonalOp(d)            diagonal scaling (Operators.jl:37-118)
    TensorProductOp2D(A, B)  (B kron A) u = A U B' (Operators.jl:122-209)
    MatrixOp(M)              1-D mat-vec (Operators.jl:1-35)
    IdentityOp, ZeroOp       (OperatorBasics.jl)
    A + B, A - B, a * A, A * B (fused where the algebra allows, lazy otherwise)
    mul_(v, A, u)            mul!(v, A, u); ldiv_(v, A, u) ldiv!(v, A, u)
    A(u, p, t), A(du, u, p, t)  the SciML call forms (Operators.jl:184-190)

Fields are the (nr, ns) column-major device arrays of semhip (the reference's
Field wraps the same array). An SEMOperator or an MG preconditioner is an
operator of this interface as well (mul_ applies it). Arithmetic runs in the
library: sem_pointwise (products, quotients, inverses), sem_axpby (sums) and
sem_abu (tensor products).
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib as L
from .core import ABu, MG, SEMOperator, _like, _p, ctx, to_device

_F64 = torch.float64


def Field(a) -> torch.Tensor:
    """A field: the column-major device array of a (nr, ns) array."""
    return to_device(np.asarray(a, dtype=np.float64) if not isinstance(a, torch.Tensor) else a)


def _pw(op: int, a: torch.Tensor, b, out: torch.Tensor, alpha: float = 0.0) -> torch.Tensor:
    L.call("sem_pointwise", ctx(), op, a.numel(), float(alpha), _p(a), _p(b), _p(out))
    return out


def _axpby(alpha, x, beta, y, out):
    L.call("sem_axpby", ctx(), x.numel(), float(alpha), _p(x), float(beta), _p(y), _p(out))
    return out


class AbstractOperator:
    """Common behaviour: A * u, A(u, p, t), A(du, u, p, t), adjoint, algebra."""

    def __mul__(self, other):
        if isinstance(other, AbstractOperator):
            return ComposeOp(self, other)
        if isinstance(other, (int, float)):
            return ScaledOp(float(other), self)
        return self.apply(other)

    def __rmul__(self, lam):
        if isinstance(lam, (int, float)):
            return ScaledOp(float(lam), self)
        return NotImplemented

    def __add__(self, other):
        return AffineOp(1.0, self, 1.0, other)

    def __sub__(self, other):
        return AffineOp(1.0, self, -1.0, other)

    def __call__(self, *args):
        if len(args) == 3:          # A(u, p, t) = A * u
            return self * args[0]
        if len(args) == 4:          # A(du, u, p, t) = mul!(du, A, u)
            return mul_(args[0], self, args[1])
        if len(args) == 1:
            return self * args[0]
        raise TypeError("A(u, p, t) or A(du, u, p, t)")

    def apply(self, u):
        """A * u into a new field."""
        u = to_device(u)
        return self.mul(_like(u), u)

    def mul(self, v, u):
        raise NotImplementedError(f"mul! not defined for {type(self).__name__}")

    def ldiv(self, v, u):
        raise NotImplementedError(f"operator inversion not defined for {type(self).__name__}")

    @property
    def T(self):
        return self.adjoint()

    def adjoint(self):
        raise NotImplementedError


class IdentityOp(AbstractOperator):
    def mul(self, v, u):
        v.copy_(u)
        return v

    def ldiv(self, v, u):
        v.copy_(u)
        return v

    def adjoint(self):
        return self


class ZeroOp(AbstractOperator):
    def mul(self, v, u):
        return _axpby(0.0, u, 0.0, u, v)

    def adjoint(self):
        return self


class DiagonalOp(AbstractOperator):
    """v = d .* u; ldiv: v = u ./ d; inv(D) = DiagonalOp(1 ./ d)."""

    def __init__(self, diag):
        self.diag = to_device(diag)

    def mul(self, v, u):
        return _pw(L.SEM_PW_MUL, self.diag, u, v)

    def ldiv(self, v, u):
        return _pw(L.SEM_PW_DIV, u, self.diag, v)

    def inv(self):
        return DiagonalOp(_pw(L.SEM_PW_SCALAR_DIV, self.diag, None, _like(self.diag), 1.0))

    def adjoint(self):
        return self

    # fusion (Operators.jl:80-118)
    def __add__(self, other):
        if isinstance(other, DiagonalOp):
            return DiagonalOp(_axpby(1.0, self.diag, 1.0, other.diag, _like(self.diag)))
        return super().__add__(other)

    def __sub__(self, other):
        if isinstance(other, DiagonalOp):
            return DiagonalOp(_axpby(1.0, self.diag, -1.0, other.diag, _like(self.diag)))
        return super().__sub__(other)

    def __mul__(self, other):
        if isinstance(other, DiagonalOp):
            return DiagonalOp(_pw(L.SEM_PW_MUL, self.diag, other.diag, _like(self.diag)))
        if isinstance(other, (int, float)):
            return DiagonalOp(_axpby(float(other), self.diag, 0.0, self.diag, _like(self.diag)))
        return super().__mul__(other)

    def __rmul__(self, lam):
        return self.__mul__(lam)

    def __truediv__(self, other):
        if isinstance(other, DiagonalOp):
            return DiagonalOp(_pw(L.SEM_PW_DIV, self.diag, other.diag, _like(self.diag)))
        return DiagonalOp(_axpby(1.0 / float(other), self.diag, 0.0, self.diag,
                                 _like(self.diag)))


def _host_mat(A) -> np.ndarray:
    return np.asarray(A.cpu() if isinstance(A, torch.Tensor) else A, dtype=np.float64)


class MatrixOp(AbstractOperator):
    """1-D mat-vec v = M u on a vector (an (n, 1) field), through ABu."""

    def __init__(self, mat):
        self.mat = _host_mat(mat)

    @staticmethod
    def _col(u):
        return to_device(u.reshape(-1, 1) if isinstance(u, torch.Tensor) else
                         np.asarray(u, dtype=np.float64).reshape(-1, 1))

    def apply(self, u):
        return ABu(None, self.mat, self._col(u))

    def mul(self, v, u):
        v.copy_(ABu(None, self.mat, self._col(u)).reshape(v.shape))
        return v

    def ldiv(self, v, u):
        return MatrixOp(np.linalg.inv(self.mat)).mul(v, u)

    def adjoint(self):
        return MatrixOp(self.mat.T)


class TensorProductOp2D(AbstractOperator):
    """(B kron A) u: v = A U B' for a field U of size (size(A,2), size(B,2))."""

    def __init__(self, A, B):
        self.A, self.B = _host_mat(A), _host_mat(B)

    def apply(self, u):
        return ABu(self.B, self.A, to_device(u))

    def mul(self, v, u):
        v.copy_(ABu(self.B, self.A, to_device(u)))
        return v

    def ldiv(self, v, u):
        return TensorProductOp2D(np.linalg.inv(self.A), np.linalg.inv(self.B)).mul(v, u)

    def adjoint(self):
        return TensorProductOp2D(self.A.T, self.B.T)

    def __mul__(self, other):
        # (A1 kron-form)(A2 kron-form) = (A1 A2, B1 B2): exact fusion
        if isinstance(other, TensorProductOp2D):
            return TensorProductOp2D(self.A @ other.A, self.B @ other.B)
        return super().__mul__(other)


class ScaledOp(AbstractOperator):
    def __init__(self, lam, A):
        self.lam, self.A = lam, A

    def mul(self, v, u):
        mul_(v, self.A, u)
        return _axpby(self.lam, v, 0.0, v, v)

    def adjoint(self):
        return ScaledOp(self.lam, self.A.adjoint())


class AffineOp(AbstractOperator):
    """a A + b B, applied lazily (one sum after the two products). (The
    reference fuses A + B of tensor-product operators into the tensor product
    of the sums, Operators.jl:196-203, which is a different operator; the
    lazy sum is the one the notation means.)"""

    def __init__(self, a, A, b, B):
        self.a, self.A, self.b, self.B = float(a), A, float(b), B

    def mul(self, v, u):
        w = _like(v)
        mul_(v, self.A, u)
        mul_(w, self.B, u)
        return _axpby(self.a, v, self.b, w, v)

    def adjoint(self):
        return AffineOp(self.a, self.A.adjoint(), self.b, self.B.adjoint())


class ComposeOp(AbstractOperator):
    """A * B applied as A (B u)."""

    def __init__(self, A, B):
        self.A, self.B = A, B

    def mul(self, v, u):
        w = _like(v)
        mul_(w, self.B, u)
        return mul_(v, self.A, w)

    def adjoint(self):
        return ComposeOp(self.B.adjoint(), self.A.adjoint())


def mul_(v, A, u):
    """mul!(v, A, u): v = A u in place (Operators.jl:27, 65, 174)."""
    if isinstance(A, AbstractOperator):
        return A.mul(v, u)
    if isinstance(A, (SEMOperator, MG)):
        A(to_device(u, dtype=getattr(A, "dtype", _F64)), out=v)
        return v
    raise TypeError(f"mul! not defined for {type(A).__name__}")


def ldiv_(v, A, u):
    """ldiv!(v, A, u): v = A \\ u in place."""
    return A.ldiv(v, u)


def inv(A):
    return A.inv()
