"""The Diffusion equation on the GPU (diffusion.jl:7-137).

Host code keeps the reference's names (``!`` becomes a trailing ``_``):
``Diffusion(bc, msh; Ti, Tf, dt, k)`` with ``opLHS``, ``opPrecond``,
``makeRHS_``, ``solve_``, ``evolve_`` and ``simulate_``. Every field
operation is a call into libsemhip.so: the left-hand side is a fused
``SEMOperator`` (hlmz with the viscosity field and k = bdfB[1], gather-scatter
and mask) that reads ``nu`` in place, so a setVisc! that writes into ``nu``
changes the operator; the solve is the fused PCG; the right-hand side is one
``lapl`` launch for the boundary data plus one ``combine`` launch (mask and
gather-scatter included).

Differences from the reference, all value-preserving or opt-in:

* The field keeps k histories; the reference's Field always keeps 3
  (mesh.jl:187), which equals k for its default k = 3.
* ``precond="jacobi"`` selects the assembled Jacobi diagonal of the step
  operator, rebuilt for every solve because nu may change; the default is
  the reference's identity (diffusion.jl:47-49).
* ``simulate_`` is the reference's ``while time <= Tf`` loop; ``run_`` steps
  by count.
"""
from __future__ import annotations

from typing import Callable, Optional

from .cd import Field, TimeStepper, fixU_
from .core import Diagonal, Mesh, PCGInfo, SEMOperator, combine, identity, lapl, pcg


class Diffusion:
    """Diffusion(bc, msh; Ti=0, Tf=0, dt=0, k=3) (diffusion.jl:7-34): fld
    (u, histories uh, boundary data ub, mask M), nu, f, rhs (device fields),
    tstep, msh. ``tol`` and ``maxiter`` are the pcg! keywords (pcg.jl:64-79;
    absolute tolerance, default 1e-8)."""

    def __init__(self, bc, msh: Mesh, Ti: float = 0.0, Tf: float = 0.0, dt: float = 0.0,
                 k: int = 3, precond: str = "identity", tol: float = 1e-8,
                 maxiter: Optional[int] = None):
        if precond not in ("identity", "jacobi"):
            raise ValueError("precond must be 'identity' or 'jacobi'")
        self.fld = Field(bc, msh, k)
        self.nu = msh.zeros()
        self.f = msh.zeros()
        self.rhs = msh.zeros()
        self.tstep = TimeStepper(Ti, Tf, dt, k)
        self.msh = msh
        self.precond, self.tol, self.maxiter = precond, float(tol), maxiter
        self.info = PCGInfo()
        self.iters = []
        self._ops = {}

    def _op(self) -> SEMOperator:
        b0 = float(self.tstep.bdfB[0])
        op = self._ops.get(b0)
        if op is None:
            op = self._ops[b0] = SEMOperator(self.nu, b0, self.msh, self.fld.M)
        return op

    def opLHS(self, u):
        """mask(gatherScatter(hlmz(u, nu, bdfB[1]))) (diffusion.jl:36-45)."""
        return self._op()(u)

    def opPrecond(self, u):
        """The identity (diffusion.jl:47-49), or d .* u with the inverse
        assembled diagonal for precond='jacobi'."""
        if self.precond == "identity":
            return u
        return self._op().jacobi() * u

    def makeRHS_(self):
        """rhs = QQ' mask(B f - nu .* lapl(ub) - sum_i bdfB[1+i] B uh_i)
        (diffusion.jl:51-65)."""
        b = self.tstep.bdfB
        terms = [(1.0, self.f, True), (-1.0, lapl(self.fld.ub, self.nu, self.msh), False)]
        for i, uh in enumerate(self.fld.uh):
            terms.append((-float(b[1 + i]), uh, True))
        combine(self.msh, terms, M=self.fld.M, assemble=True, out=self.rhs)

    def solve_(self):
        """pcg!(u, rhs, opLHS; opM=opPrecond, mult) then u .+= ub
        (diffusion.jl:67-77)."""
        op = self._op()
        opM = identity if self.precond == "identity" else Diagonal(op.jacobi())
        pcg(self.rhs, op, opM=opM, mult=self.msh.mult, tol=self.tol, maxiter=self.maxiter,
            info=self.info, x=self.fld.u)
        self.iters.append(self.info.iters)
        self.fld.u.add_(self.fld.ub)

    def evolve_(self, setBC: Callable = fixU_, setForcing: Callable = fixU_,
                setVisc: Callable = fixU_):
        """evolve!(dfn, setBC!, setForcing!, setVisc!) (diffusion.jl:81-106):
        histories shift, the stepper advances, ub, f and nu are set at the
        new time, then makeRHS! and solve!."""
        self.fld.updateHist_()
        self.tstep.updateHist_()
        m, t = self.msh, self.tstep.time[0]
        setBC(self.fld.ub, m.x, m.y, t)
        setForcing(self.f, m.x, m.y, t)
        setVisc(self.nu, m.x, m.y, t)
        self.makeRHS_()
        self.solve_()

    def simulate_(self, callback: Callable = lambda dfn: None, setIC: Callable = fixU_,
                  setBC: Callable = fixU_, setForcing: Callable = fixU_,
                  setVisc: Callable = fixU_):
        """simulate!(dfn, callback!, setIC!, setBC!, setForcing!, setVisc!)
        (diffusion.jl:110-137)."""
        m = self.msh
        setIC(self.fld.u, m.x, m.y, self.tstep.time[0])
        callback(self)
        while self.tstep.time[0] <= self.tstep.Tf:
            self.evolve_(setBC, setForcing, setVisc)
            callback(self)
            if self.tstep.time[0] < 1e-12:
                break

    def run_(self, nsteps: int, setBC: Callable = fixU_, setForcing: Callable = fixU_,
             setVisc: Callable = fixU_):
        """nsteps calls of evolve!, counted (not by time)."""
        for _ in range(int(nsteps)):
            self.evolve_(setBC, setForcing, setVisc)
