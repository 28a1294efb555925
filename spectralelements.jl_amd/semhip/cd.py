"""Unsteady convection-diffusion on the GPU: the reference's ConvectionDiffusion
equation (convectionDiffusion.jl:5-179) with its BDFk/EXTk TimeStepper
(time.jl:31-112), and a discrete adjoint through the time loop with
recompute-from-checkpoint (sqrt(N) snapshots).

Host code keeps the reference's names (``!`` becomes a trailing ``_``):
``bdfExtK``, ``TimeStepper.updateHist_``, ``Field``, ``ConvectionDiffusion``
with ``opLHS``, ``opPrecond``, ``makeRHS_``, ``solve_``, ``updateHist_``,
``evolve_``, ``step_`` and ``simulate_``. Every field operation is a call into
libsemhip.so: the step operator is a fused ``SEMOperator`` (hlmz with k =
bdfB[1], gather-scatter and mask), the solve is the fused PCG, the right-hand
side is one ``combine`` launch plus the gather-scatter, and the convective term
is the dealiased ``advect`` kernel.

Differences from the reference, all value-preserving or opt-in:

* ``advect`` of each history is computed once, when it is first needed, and
  shifted with the history (``advH``) (the reference recomputes every history term
  each step, convectionDiffusion.jl:100-104; the values are identical).
* The default preconditioner is the assembled Jacobi diagonal of the step
  operator. ``precond="reference"`` selects u ./ B ./ bdfB[1]
  (convectionDiffusion.jl:87-91), whose per-copy B makes PCG iterates
  discontinuous at shared nodes: that solve forms <p, A p> on the assembled
  A p as pcg.jl:50 does (SEM_PRECOND_DIAG_LOCAL), so it runs the reference's
  recurrence, with host-read scalars each iteration; the adjoint is exact only
  for the continuous preconditioners.
* The PCG stopping test is relative, norm(r, Inf) <= tol * norm(rhs, Inf),
  unless ``abs_tol`` is given (the reference's pcg! default is tol = 1e-8
  absolute, pcg.jl:16).
"""
from __future__ import annotations

import ctypes as C
import math
import weakref
from typing import Callable, List, Optional

import numpy as np
import torch

from . import _lib as L
from .core import (Diagonal, Mesh, PCGInfo, SEMOperator, _p, advect, combine, ctx, dot,
                   dot_acc, empty_field, generateMask, lapl, pcg, to_device)


# ------------------------------------------------------------ time stepper
def bdfExtK(t, k: int = 3):
    """BDFk/EXTk coefficients (a, b) for the time history t = [t_{n+1}, t_n,
    ...] (time.jl:31-53): b = first row of derivMat(unique(t)), a =
    interpMat(t_{n+1}, t_past); zero-padded to k (k+1) while the history
    holds fewer distinct times; steady start (one time) gives a = [1, 0..].
    Computed by the library's host code (sem_bdfextk) once per step."""
    t = np.ascontiguousarray(t, dtype=np.float64).reshape(-1)
    a = np.zeros(int(k))
    b = np.zeros(int(k) + 1)
    L.call("sem_bdfextk", t.size, t.ctypes.data, int(k), a.ctypes.data, b.ctypes.data)
    return a, b


class TimeStepper:
    """TimeStepper(Ti, Tf, dt, k) (time.jl:70-99)."""

    def __init__(self, Ti: float, Tf: float, dt: float, k: int = 3):
        self.k = int(k)
        self.time = [float(Ti)] * (self.k + 1)
        self.bdfA, self.bdfB = bdfExtK(self.time, self.k)
        self.istep = 0
        self.fstep = int(math.ceil(Tf / dt)) if dt != 0 else 0
        self.dt, self.Ti, self.Tf = float(dt), float(Ti), float(Tf)

    def updateHist_(self):
        """updateHist!(tstep) (time.jl:101-112, mesh.jl:217-224)."""
        t = self.time
        for i in range(len(t) - 1, 0, -1):
            t[i] = t[i - 1]
        t[0] = t[1]
        self.istep += 1
        t[0] += self.dt
        self.bdfA, self.bdfB = bdfExtK(t, self.k)

    def state(self):
        return (list(self.time), self.istep, self.bdfA.copy(), self.bdfB.copy())

    def restore(self, st):
        self.time, self.istep = list(st[0]), st[1]
        self.bdfA, self.bdfB = st[2].copy(), st[3].copy()


class Field:
    """Field(bc, msh, k) (mesh.jl:177-195): u, histories uh (k), boundary data
    ub, mask M."""

    def __init__(self, bc, msh: Mesh, k: int = 3):
        self.msh = msh
        self.u = msh.zeros()
        self.uh = [msh.zeros() for _ in range(k)]
        self.ub = msh.zeros()
        self.M = generateMask(bc, msh)
        self._ub_zero = True

    def updateHist_(self):
        """updateHist!(fld) (mesh.jl:199-215): uh[i] <- uh[i-1], uh[1] <- u."""
        last = self.uh[-1]
        for i in range(len(self.uh) - 1, 0, -1):
            self.uh[i] = self.uh[i - 1]
        last.copy_(self.u)
        self.uh[0] = last


def fixU_(u, x, y, t):
    """Leaves u unchanged (mesh.jl:6-8 used as a setter)."""
    return None


# ---------------------------------------------------------------- equation
class ConvectionDiffusion:
    """ConvectionDiffusion(name, fld, vx, vy, tstep, mshD, set0!, set∂!, setF!,
    setν!) (convectionDiffusion.jl:5-56). ``nu`` is a scalar or field
    (the reference fills cdn.ν through setν!); ``f`` an optional forcing field.
    Setters have the reference's signature setter(array, x, y, t) and write in
    place; the default fixU_ leaves the array as it is."""

    def __init__(self, name: str, fld: Field, vx, vy, tstep: TimeStepper, mshD: Mesh,
                 set0: Callable = fixU_, setb: Callable = fixU_, setF: Callable = fixU_,
                 setnu: Optional[Callable] = None, nu=0.0, f=None, precond: str = "jacobi",
                 tol: float = 1e-12, abs_tol: Optional[float] = None,
                 maxiter: Optional[int] = None):
        self.name = name
        self.fld, self.mshV, self.mshD, self.tstep = fld, fld.msh, mshD, tstep
        self.vx, self.vy = to_device(vx), to_device(vy)
        self.nu = nu if isinstance(nu, (int, float)) else to_device(nu)
        self.f = to_device(f) if f is not None else None
        self.rhs = self.mshV.zeros()
        k = len(fld.uh)
        # advect(uh_i) per history (the reference's exH_i is its negative)
        self.advH: List[Optional[torch.Tensor]] = [None] * k
        self.set0, self.setb, self.setF, self.setnu = set0, setb, setF, setnu
        if precond not in ("jacobi", "reference"):
            raise ValueError("precond must be 'jacobi' or 'reference'")
        self.precond, self.tol, self.abs_tol, self.maxiter = precond, tol, abs_tol, maxiter
        self._ops = {}
        self.info = PCGInfo()
        self.iters_total = 0
        self._ts0 = tstep.state()
        self._cd = None          # the library's time loop (sem_cd), built on first use
        self._adv_spare: List[torch.Tensor] = []

    # -- operators of the current step (cached per bdfB[1])
    def _op(self, b0: float):
        e = self._ops.get(b0)
        if e is None:
            op = SEMOperator(self.nu, b0, self.mshV, self.fld.M)
            if self.precond == "jacobi":
                e = (op, Diagonal(op.jacobi(), local=False))
            else:
                e = (op, Diagonal(1.0 / (self.mshV.B * b0), local=True))
            self._ops[b0] = e
        return e

    def opLHS(self, u):
        """mask(gatherScatter(hlmz(u, nu, bdfB[1]))) (convectionDiffusion.jl:76-85)."""
        return self._op(float(self.tstep.bdfB[0]))[0](u)

    def opPrecond(self, u):
        """The step preconditioner applied to u (convectionDiffusion.jl:87-91
        for precond='reference'; the assembled Jacobi diagonal by default)."""
        return self._op(float(self.tstep.bdfB[0]))[1](u)

    def makeRHS_(self):
        """rhs = QQ' mask(B f - nu lapl(ub) - sum b_{1+i} B uh_i + sum a_i exH_i)
        with exH_i = -advect(uh_i) (convectionDiffusion.jl:93-110)."""
        a, b = self.tstep.bdfA, self.tstep.bdfB
        terms = []
        if self.f is not None:
            terms.append((1.0, self.f, True))
        if not self.fld._ub_zero:
            terms.append((-1.0, lapl(self.fld.ub, self.nu, self.mshV), False))
        for i, uh in enumerate(self.fld.uh):
            if self.advH[i] is None and a[i] != 0.0:
                self.advH[i] = advect(uh, self.vx, self.vy, self.mshV, self.mshD)
            if b[1 + i] != 0.0:
                terms.append((-float(b[1 + i]), uh, True))
            if a[i] != 0.0:
                terms.append((-float(a[i]), self.advH[i], False))   # a_i exH_i
        combine(self.mshV, terms, M=self.fld.M, assemble=True, out=self.rhs)

    def solve_(self):
        """pcg!(u, rhs, opLHS; opM, mult) then u .+= ub (convectionDiffusion.jl:112-122)."""
        op, d = self._op(float(self.tstep.bdfB[0]))
        # relative tolerance: norm(rhs, Inf) is formed on the device
        kw = {"tol": self.abs_tol} if self.abs_tol is not None else {"rtol": self.tol}
        x = pcg(self.rhs, op, opM=d, mult=self.mshV.mult, maxiter=self.maxiter,
                info=self.info, x=self.fld.u, **kw)
        self.iters_total += self.info.iters
        if not self.fld._ub_zero:
            x.add_(self.fld.ub)

    def updateHist_(self):
        """updateHist!(cdn) (convectionDiffusion.jl:124-129); the cached
        convective terms shift with the histories."""
        self.fld.updateHist_()
        self.advH = [None] + self.advH[:-1]

    def evolve_(self):
        """evolve!(cdn) (convectionDiffusion.jl:131-144)."""
        t = self.tstep.time[0]
        m = self.mshV
        if self.setb is not fixU_:
            self.setb(self.fld.ub, m.x, m.y, t)
            self.fld._ub_zero = False
        if self.setF is not fixU_:
            if self.f is None:
                self.f = m.zeros()
            self.setF(self.f, m.x, m.y, t)
        if self.setnu is not None:
            raise NotImplementedError("time-dependent nu: rebuild the equation per step")
        self.makeRHS_()
        self.solve_()

    def step_(self):
        """step!(cdn) (convectionDiffusion.jl:148-155)."""
        if self.native():
            self.steps_(1)
            return
        self.updateHist_()
        self.tstep.updateHist_()
        self.evolve_()

    def native(self) -> bool:
        """True when the library's time loop (sem_cd_steps) covers this
        equation: no time-dependent data (default setters, f None, ub zero),
        the Jacobi preconditioner, the relative tolerance, order <= 4."""
        return (self.setb is fixU_ and self.setF is fixU_ and self.setnu is None
                and self.f is None and self.fld._ub_zero and self.precond == "jacobi"
                and self.abs_tol is None and 1 <= len(self.fld.uh) <= 4)

    def steps_(self, nsteps: int, acc: Optional[torch.Tensor] = None, record=None):
        """nsteps calls of step! inside the library (sem_cd_steps): the same
        launches in the same order as step_ issues one by one, with one host
        wait per step (the PCG stop test). acc: a one-element device tensor
        receiving += dt/2 sum(B u^2) after each step; record: a list that
        receives each step's (bdfA, bdfB)."""
        n = int(nsteps)
        if n <= 0:
            return
        if not self.native():
            raise ValueError("steps_ needs an equation the library's time loop covers")
        k = len(self.fld.uh)
        if self._cd is None:
            h = C.c_void_p()
            nu_s, nu_f = ((float(self.nu), None) if isinstance(self.nu, (int, float))
                          else (0.0, self.nu))
            L.call("sem_cd_create", ctx(), self.mshV.h, self.mshD.h, int(self.fld.M._sem_bits),
                   k, nu_s, _p(nu_f), _p(self.vx), _p(self.vy), float(self.tstep.dt),
                   float(self.tol), int(self.maxiter or 0), C.byref(h))
            self._cd = h
            self._cd_fin = weakref.finalize(self, L.load().sem_cd_destroy, h)
        adv, valid = [], []
        for e in self.advH:
            if e is not None:
                adv.append(e)
                valid.append(1)
            else:
                adv.append(self._adv_spare.pop() if self._adv_spare else self.mshV.zeros())
                valid.append(0)
        uh_p = (C.c_void_p * k)(*[t.data_ptr() for t in self.fld.uh])
        adv_p = (C.c_void_p * k)(*[t.data_ptr() for t in adv])
        val = (C.c_int * k)(*valid)
        tt = np.array(self.tstep.time, dtype=np.float64)
        a_out = np.zeros((n, k)) if record is not None else None
        b_out = np.zeros((n, k + 1)) if record is not None else None
        it = C.c_int64()
        st = L.call("sem_cd_steps", ctx(), self._cd, n, _p(self.fld.u), uh_p, adv_p, val,
                    tt.ctypes.data, _p(acc), a_out.ctypes.data if a_out is not None else None,
                    b_out.ctypes.data if b_out is not None else None, C.byref(it))
        uh_of = {t.data_ptr(): t for t in self.fld.uh}
        self.fld.uh = [uh_of[uh_p[i]] for i in range(k)]
        adv_of = {t.data_ptr(): t for t in adv}
        bufs = [adv_of[adv_p[i]] for i in range(k)]
        self.advH = [bufs[i] if val[i] else None for i in range(k)]
        self._adv_spare += [bufs[i] for i in range(k) if not val[i]]
        self.tstep.time = [float(v) for v in tt]
        self.tstep.istep += n
        self.tstep.bdfA, self.tstep.bdfB = bdfExtK(self.tstep.time, self.tstep.k)
        self.iters_total += it.value
        if record is not None:
            record.extend((a_out[s].copy(), b_out[s].copy()) for s in range(n))
        if st == L.SEM_WARN_MAXITER:
            L.maxiter_warning(float("nan"))

    def simulate_(self, callback: Callable = lambda cdn: None):
        """simulate!(cdn, callback!) (convectionDiffusion.jl:159-179)."""
        m = self.mshV
        self.set0(self.fld.u, m.x, m.y, self.tstep.time[0])
        callback(self)
        while self.tstep.time[0] <= self.tstep.Tf:
            self.step_()
            callback(self)
            if self.tstep.time[0] < 1e-12:
                break

    def run_(self, nsteps: int, callback: Optional[Callable] = None):
        """nsteps calls of step!, counted (not by time)."""
        if callback is None and self.native():
            self.steps_(nsteps)
            return
        for _ in range(int(nsteps)):
            self.step_()
            if callback is not None:
                callback(self)

    def reset_(self, u0):
        """u = u0, histories zero, the time stepper back at its start."""
        self.fld.u.copy_(to_device(u0))
        for h in self.fld.uh:
            h.zero_()
        self.advH = [None] * len(self.fld.uh)
        self.tstep.restore(self._ts0)

    # -- state snapshots for checkpointing
    def snapshot(self):
        return (self.fld.u.clone(), [u.clone() for u in self.fld.uh],
                [e.clone() if e is not None else None for e in self.advH], self.tstep.state())

    def restore(self, snap):
        u, uh, advH, ts = snap
        self.fld.u.copy_(u)
        for dst, src in zip(self.fld.uh, uh):
            dst.copy_(src)
        self.advH = [e.clone() if e is not None else None for e in advH]
        self.tstep.restore(ts)


# ----------------------------------------------------------------- adjoint
class CDGradient:
    """Result of ``cd_gradient``: objective J, dJ/du0 per local node, and
    counters (forward steps taken including recomputation, snapshots held)."""

    def __init__(self, J, grad, steps_forward, snapshots, peak_states, iters_fwd, iters_adj):
        self.J, self.grad = J, grad
        self.steps_forward, self.snapshots, self.peak_states = steps_forward, snapshots, peak_states
        self.iters_fwd, self.iters_adj = iters_fwd, iters_adj


def cd_objective_step(cdn: ConvectionDiffusion, acc: Optional[torch.Tensor] = None):
    """dt/2 sum_local B u^2 of the current state; with acc (a one-element
    device tensor) it is added there on the device instead of returned."""
    u = cdn.fld.u
    if acc is not None:
        return dot_acc(acc, u, u, cdn.mshV.B, 0.5 * cdn.tstep.dt)
    return 0.5 * cdn.tstep.dt * dot(u, u, cdn.mshV.B)


def cd_gradient(cdn: ConvectionDiffusion, u0, nsteps: int,
                checkpoints: Optional[int] = None) -> CDGradient:
    """J = sum_{n=1}^{N} dt/2 sum_local B (u^n)^2 and dJ/du0 for the
    equation's stepping from u0 (histories zero, the stepper reset to the
    state it had when the equation was built).

    Forward: one pass accumulating J and keeping a snapshot every c steps
    (c = ceil(sqrt(N)) unless ``checkpoints`` = number of snapshots is given).
    Reverse: for each segment, last first, restore its snapshot, recompute
    and keep its c states, then take the adjoint steps
        lam^n = H_n^{-1} M QQ'(dt B u^n + acc_n),
        acc_{n-1-i} += -b_{1+i} B lam^n - a_i advect'(lam^n),
    with H_n the step operator (symmetric) and advect' the transpose of the
    dealiased convective map; acc_0 is dJ/du0. The equation's solves use its
    own PCG settings; this needs a continuous preconditioner (Jacobi)."""
    if cdn.precond != "jacobi":
        raise ValueError("cd_gradient needs precond='jacobi' (see ConvectionDiffusion)")
    N = int(nsteps)
    if N < 1:
        raise ValueError("nsteps >= 1")
    c = int(math.ceil(math.sqrt(N))) if checkpoints is None else int(math.ceil(N / checkpoints))
    c = max(1, min(c, N))
    mV, M, dt = cdn.mshV, cdn.fld.M, cdn.tstep.dt
    k = len(cdn.fld.uh)
    cdn.reset_(u0)
    cdn.iters_total = 0
    snaps = {}
    coefs = {}
    Jd = torch.zeros(1, dtype=torch.float64, device=cdn.fld.u.device)
    if cdn.native():
        for m0 in range(0, N, c):
            snaps[m0] = cdn.snapshot()
            rec = []
            cdn.steps_(min(c, N - m0), acc=Jd, record=rec)
            for s, ab in enumerate(rec):
                coefs[m0 + 1 + s] = ab
    else:
        for n in range(1, N + 1):
            if (n - 1) % c == 0:
                snaps[n - 1] = cdn.snapshot()
            cdn.step_()
            coefs[n] = (cdn.tstep.bdfA.copy(), cdn.tstep.bdfB.copy())
            cd_objective_step(cdn, Jd)
    J = float(Jd)
    iters_fwd = cdn.iters_total
    steps = N
    acc = {}
    peak = 0
    iters_adj = 0
    info = PCGInfo()
    for m0 in sorted(snaps.keys(), reverse=True):
        m1 = min(m0 + c, N)
        cdn.restore(snaps[m0])
        states = {}
        for n in range(m0 + 1, m1 + 1):
            cdn.step_()
            states[n] = cdn.fld.u.clone()
        steps += m1 - m0
        peak = max(peak, len(states))
        for n in range(m1, m0, -1):
            a, b = coefs[n]
            terms = [(dt, states[n], True)]
            if n in acc:
                terms.append((1.0, acc.pop(n), False))
            mu = combine(mV, terms, M=M, assemble=True)
            op, d = cdn._op(float(b[0]))
            kw = {"tol": cdn.abs_tol} if cdn.abs_tol is not None else {"rtol": cdn.tol}
            lam = pcg(mu, op, opM=d, mult=mV.mult, maxiter=cdn.maxiter, info=info, **kw)
            iters_adj += info.iters
            Ct = None
            for i in range(k):
                mlev = n - 1 - i
                if mlev < 0 or (b[1 + i] == 0.0 and a[i] == 0.0):
                    continue
                if a[i] != 0.0 and Ct is None:
                    Ct = advect(lam, cdn.vx, cdn.vy, mV, cdn.mshD, transpose=True)
                t = []
                if b[1 + i] != 0.0:
                    t.append((-float(b[1 + i]), lam, True))
                if a[i] != 0.0:
                    t.append((-float(a[i]), Ct, False))
                if mlev in acc:
                    t.append((1.0, acc[mlev], False))
                    combine(mV, t, M=None, assemble=False, out=acc[mlev])
                else:
                    acc[mlev] = combine(mV, t, M=None, assemble=False)
            del states[n]
        del snaps[m0]
    grad0 = acc.pop(0) if 0 in acc else mV.zeros()
    return CDGradient(J, grad0, steps, math.ceil(N / c), peak, iters_fwd, iters_adj)


def cd_objective(cdn: ConvectionDiffusion, u0, nsteps: int) -> float:
    """J of ``cd_gradient`` by a plain forward run from u0."""
    cdn.reset_(u0)
    Jd = torch.zeros(1, dtype=torch.float64, device=cdn.fld.u.device)
    if cdn.native():
        cdn.steps_(int(nsteps), acc=Jd)
        return float(Jd)
    for _ in range(int(nsteps)):
        cdn.step_()
        cd_objective_step(cdn, Jd)
    return float(Jd)
