"""The Julia binding (spectralelements.jl_amd/julia/SpectralElementsHIP.jl)
cannot run here (no Julia toolchain). These tests issue, library entry by
library entry and in the same order, the calls its methods make when the
reference's callers run over device fields, and check the results against
the oracle:

* HIPDiffusion's evolve! (diffusion.jl:36-106): updateHist!, the setters
  (host arrays uploaded), makeRHS! (mass, lapl with a field nu, the history
  terms, mask with the host M uploaded once, gatherScatter; each `.=` /
  `.-=` one call plus a copy), solve! as the reference's own pcg! (pcg.jl:
  16-79) over the HIPField algebra with the closures opLHS / opPrecond and
  the host mult (per iteration: norm(r, Inf), two products and a plain sum
  per inner product, axpby and scale for the vector updates), u .+= ub.
* HIPConvectionDiffusion's step! (convectionDiffusion.jl:76-155): the same,
  with exH_i .= -advect(uh_i), the a_i exH_i terms and the preconditioner
  u ./ B ./ bdfB[1] (two pointwise divisions).

Both run the reference's absolute tolerance 1e-8 (pcg.jl:20); per-step
iteration counts agree with the oracle's within one and the states within
1e-7 relative (the inner products sum in another order)."""
import sys

import numpy as np
import pytest

from oracle import sem_oracle as o

pytestmark = pytest.mark.gpu


def _log(msg):
    print(f"[case] {msg}", file=sys.stderr, flush=True)


class JL:
    """The library calls of the binding's HIPField methods (ctypes, same
    entry points and arguments as the Julia ccalls)."""

    def __init__(self, se):
        import torch
        from semhip import _lib as L
        self.se, self.L, self.torch = se, L, torch
        self.ctx = se.ctx()
        self.host = {}

    # device fields and uploads
    def new(self, shape):
        return self.se.empty_field(*shape)

    def upload(self, a):
        return self.se.to_device(np.asarray(a, dtype=np.float64))

    def leaf(self, x):
        """_leaf: a constant host array is uploaded once per object."""
        if isinstance(x, np.ndarray):
            k = id(x)
            if k not in self.host:
                self.host[k] = (x, self.upload(x))
            return self.host[k][1]
        return x

    def p(self, t):
        return self.se.core._p(t)

    # field algebra (the _copy!, _axpby!, _pw!, _dot, norm of the binding)
    def copy_(self, y, x):
        self.L.call("sem_mask", self.ctx, x.numel(), None, self.p(x), self.p(y))
        return y

    def axpby_(self, y, a, x, b, z):
        self.L.call("sem_axpby", self.ctx, y.numel(), float(a), self.p(x), float(b), self.p(z),
                    self.p(y))
        return y

    def pw_(self, y, op, alpha, a, b):
        self.L.call("sem_pointwise", self.ctx, op, y.numel(), float(alpha), self.p(a),
                    self.p(b), self.p(y))
        return y

    def sum(self, a):
        import ctypes as C
        r = C.c_double()
        self.L.call("sem_dot", self.ctx, a.numel(), self.p(a), None, None, C.byref(r))
        return r.value

    def norm_inf(self, a):
        import ctypes as C
        r = C.c_double()
        self.L.call("sem_norm_inf", self.ctx, a.numel(), self.p(a), C.byref(r))
        return r.value

    def zero(self, like):
        return self.torch.zeros_like(like)          # fill!(similar(u), 0): hipMemset

    def copy(self, x):
        return self.copy_(self.new(x.shape), x)

    def add(self, a, b):                            # a + b
        return self.axpby_(self.new(a.shape), 1.0, a, 1.0, b)

    def sub(self, a, b):                            # a - b
        return self.axpby_(self.new(a.shape), 1.0, a, -1.0, b)

    def scale(self, s, a):                          # s * a
        return self.pw_(self.new(a.shape), self.L.SEM_PW_SCALE, s, a, None)

    def neg(self, a):                               # -a
        return self.pw_(self.new(a.shape), self.L.SEM_PW_SCALE, -1.0, a, None)

    def prod3(self, a, b, c):                       # a .* b .* c (n-ary *)
        y = self.new(a.shape)
        self.pw_(y, self.L.SEM_PW_MUL, 0.0, a, self.leaf(b))
        return self.pw_(y, self.L.SEM_PW_MUL, 0.0, y, self.leaf(c))

    # broadcast assignments of the callers
    def assign(self, y, x):                         # y .= x
        return self.copy_(y, x)

    def sub_scaled(self, y, s, x):                  # y .-= s .* x (s a scalar)
        return self.axpby_(y, 1.0, y, -s, x)

    def add_scaled(self, y, s, x):                  # y .+= s .* x
        return self.axpby_(y, 1.0, y, s, x)

    def sub_fieldprod(self, y, f, x):               # y .-= f .* x (f a field)
        t = self.pw_(self.new(y.shape), self.L.SEM_PW_MUL, 0.0, f, x)
        return self.axpby_(y, 1.0, y, -1.0, t)

    # operators (the L3 methods)
    def mass(self, u, m):
        y = self.new(u.shape)
        self.L.call("sem_mass", self.ctx, m.h, self.p(u), self.p(y))
        return y

    def lapl(self, u, m):
        y = self.new(u.shape)
        self.L.call("sem_lapl", self.ctx, m.h, self.p(u), self.p(y))
        return y

    def hlmz(self, u, nu, k, m):                    # nu a field, k a scalar
        y = self.new(u.shape)
        self.L.call("sem_hlmz", self.ctx, m.h, self.p(u), 0.0, self.p(nu), float(k), None,
                    self.p(y))
        return y

    def gs(self, u, m):
        y = self.new(u.shape)
        self.L.call("sem_gather_scatter", self.ctx, m.h, self.p(u), self.p(y))
        return y

    def mask(self, u, M):                           # host M: uploaded once
        y = self.new(u.shape)
        self.L.call("sem_mask", self.ctx, u.numel(), self.p(self.leaf(M)), self.p(u),
                    self.p(y))
        return y

    def advect(self, T, vx, vy, mV, mD):
        y = self.new(T.shape)
        self.L.call("sem_advect", self.ctx, mV.h, mD.h, self.p(T), self.p(vx), self.p(vy),
                    self.p(y))
        return y

    # the reference's pcg (pcg.jl:16-60) over this algebra
    def pcg(self, b, opA, opM, mult, tol=1e-8):
        x = self.zero(b)
        Ax = self.zero(b)
        ra = self.sub(b, Ax)
        hp = rp = u = None
        k = 0
        while self.norm_inf(ra) > tol:
            ha = opM(ra)
            k += 1
            hpp, rpp = hp, rp
            hp, rp = ha, ra
            t = self.sum(self.prod3(rp, hp, mult))
            if k == 1:
                u = self.copy(hp)
            else:
                u = self.add(hp, self.scale(t / self.sum(self.prod3(rpp, hpp, mult)), u))
            Au = opA(u)
            a = t / self.sum(self.prod3(u, Au, mult))
            x = self.add(x, self.scale(a, u))
            ra = self.sub(rp, self.scale(a, Au))
        return x, k


@pytest.fixture(scope="module")
def se():
    import semhip
    return semhip


def _rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


def test_julia_diffusion_evolve_replay(se):
    n, E, dt, k, nsteps = 6, 3, 1e-2, 3, 4
    _log(f"julia replay: HIPDiffusion evolve!, n={n} E={E} {nsteps} steps")
    m = se.Mesh(n, n, E, E, (False, False), se.deform_sinsin(0.1))
    mo = o.Mesh(n, n, E, E, deform=o.deform_sinsin(0.1), dense_gs=False)
    x, y = mo.x, mo.y

    def setIC(a, x, y, t):
        a[...] = np.sin(np.pi * x) * np.sin(np.pi * y)

    def setBC(a, x, y, t):
        a[...] = 0.1 * x * y * t

    def setF(a, x, y, t):
        a[...] = 1.0 + x

    def setNu(a, x, y, t):
        a[...] = 1.0 + 0.1 * x

    uo, iters_o = o.diffusion_simulate(mo, "DDDD", Ti=0.0, Tf=nsteps * dt, dt=dt, k=k,
                                       setIC=setIC, setBC=setBC, setForcing=setF,
                                       setVisc=setNu, nsteps=nsteps)
    jl = JL(se)
    M = o.generateMask("DDDD", mo)          # host mask, as the binding keeps it
    mult = np.array(mo.mult)                # the host msh.mult handed to pcg!
    ts = o.TimeStepper(0.0, nsteps * dt, dt, k=k)
    host = np.zeros_like(x)
    setIC(host, x, y, ts.time[0])
    u = jl.upload(host)
    uh = [jl.upload(np.zeros_like(x)) for _ in range(k)]
    ub, f, nu = (jl.upload(np.zeros_like(x)) for _ in range(3))
    rhs = jl.upload(np.zeros_like(x))
    iters = []
    for _ in range(nsteps):
        # updateHist!(fld), then the host TimeStepper
        for i in range(k - 1, 0, -1):
            jl.assign(uh[i], uh[i - 1])
        jl.assign(uh[0], u)
        ts.update()
        t = ts.time[0]
        for setter, fld in ((setBC, ub), (setF, f), (setNu, nu)):
            a = se.to_host(fld)
            setter(a, x, y, t)
            fld.copy_(jl.upload(a))
        b = ts.bdfB
        # makeRHS!
        jl.assign(rhs, jl.mass(f, m))
        jl.sub_fieldprod(rhs, nu, jl.lapl(ub, m))
        for i in range(k):
            jl.sub_scaled(rhs, float(b[1 + i]), jl.mass(uh[i], m))
        jl.assign(rhs, jl.mask(rhs, M))
        jl.assign(rhs, jl.gs(rhs, m))

        # solve!: pcg!(u, rhs, opL; opM = opP, mult = msh.mult), u .+= ub
        def opL(v):
            lhs = jl.hlmz(v, nu, float(b[0]), m)
            jl.assign(lhs, jl.gs(lhs, m))
            jl.assign(lhs, jl.mask(lhs, M))
            return lhs
        xs, kk = jl.pcg(rhs, opL, lambda v: v, mult)
        jl.assign(u, xs)
        jl.add_scaled(u, 1.0, ub)
        iters.append(kk)
    _log(f"  iterations {iters} oracle {iters_o}")
    assert all(abs(a - b) <= 1 for a, b in zip(iters, iters_o))
    assert _rel(se.to_host(u), uo) < 1e-7


def test_julia_convection_diffusion_step_replay(se):
    n, nd, E, dt, k, nsteps = 6, 9, 3, 1e-2, 2, 4
    _log(f"julia replay: HIPConvectionDiffusion step!, n={n} nd={nd} E={E} {nsteps} steps")
    mV = se.Mesh(n, n, E, E, (False, False), se.deform_sinsin(0.1))
    mD = se.Mesh(nd, nd, E, E, (False, False), se.deform_sinsin(0.1))
    mVo = o.Mesh(n, n, E, E, deform=o.deform_sinsin(0.1), dense_gs=False)
    mDo = o.Mesh(nd, nd, E, E, deform=o.deform_sinsin(0.1), dense_gs=False)
    x, y = mVo.x, mVo.y
    nu0 = 1e-2
    u0 = np.exp(-((x - 0.5) ** 2 + y ** 2) / (2 * 0.3 ** 2))
    vxo, vyo = -y, x
    Mo = o.generateMask("DDDD", mVo)
    # oracle: the same step with the reference's preconditioner and tolerance
    Jr, Js = o.cd_step_matrices(mVo, mDo)
    tso = o.TimeStepper(0.0, nsteps * dt, dt, k=k)
    uo = u0.copy()
    uho = [np.zeros_like(x) for _ in range(k)]
    nuf = nu0 * np.ones_like(x)
    iters_o = []
    for _ in range(nsteps):
        for i in range(k - 1, 0, -1):
            uho[i] = uho[i - 1]
        uho[0] = uo.copy()
        tso.update()
        a, b = tso.bdfA, tso.bdfB
        rhs = np.zeros_like(x)
        for i in range(k):
            exH = -o.advect(uho[i], vxo, vyo, mVo, mDo, Jr, Js)
            rhs = rhs - b[1 + i] * o.mass(uho[i], mVo)
            rhs = rhs + a[i] * exH
        rhs = mVo.gs(o.mask(rhs, Mo))
        info = o.PCGInfo()
        uo = o.pcg(rhs, lambda v: o.mask(mVo.gs(o.hlmz(v, nuf, b[0], mVo)), Mo),
                   opM=lambda r: r / mVo.B / b[0], mult=mVo.mult, tol=1e-8, info=info)
        iters_o.append(info.iters)

    jl = JL(se)
    M = Mo
    mult = np.array(mVo.mult)
    Bhost = np.array(mVo.B)
    ts = o.TimeStepper(0.0, nsteps * dt, dt, k=k)
    u = jl.upload(u0)
    uh = [jl.upload(np.zeros_like(x)) for _ in range(k)]
    exH = [jl.upload(np.zeros_like(x)) for _ in range(k)]
    ub, f = jl.upload(np.zeros_like(x)), jl.upload(np.zeros_like(x))
    nu = jl.upload(nuf)                     # set by setν! (uploaded)
    vx, vy = jl.upload(vxo), jl.upload(vyo)
    rhs = jl.upload(np.zeros_like(x))
    iters = []
    for _ in range(nsteps):
        for i in range(k - 1, 0, -1):
            jl.assign(uh[i], uh[i - 1])
        jl.assign(uh[0], u)
        ts.update()
        a, b = ts.bdfA, ts.bdfB
        jl.assign(rhs, jl.mass(f, mV))
        jl.sub_fieldprod(rhs, nu, jl.lapl(ub, mV))
        for i in range(k):
            jl.assign(exH[i], jl.neg(jl.advect(uh[i], vx, vy, mV, mD)))
            jl.sub_scaled(rhs, float(b[1 + i]), jl.mass(uh[i], mV))
            jl.add_scaled(rhs, float(a[i]), exH[i])
        jl.assign(rhs, jl.mask(rhs, M))
        jl.assign(rhs, jl.gs(rhs, mV))

        def opL(v):
            Au = jl.hlmz(v, nu, float(b[0]), mV)
            jl.assign(Au, jl.gs(Au, mV))
            jl.assign(Au, jl.mask(Au, M))
            return Au

        def opM(v):                         # u ./ B ./ bdfB[1]
            t = jl.pw_(jl.new(v.shape), jl.L.SEM_PW_DIV, 0.0, v, jl.leaf(Bhost))
            return jl.pw_(jl.new(v.shape), jl.L.SEM_PW_DIV_SCALAR, float(b[0]), t, None)
        xs, kk = jl.pcg(rhs, opL, opM, mult)
        jl.assign(u, xs)
        jl.add_scaled(u, 1.0, ub)
        iters.append(kk)
    _log(f"  iterations {iters} oracle {iters_o}")
    assert all(abs(p - q) <= 1 for p, q in zip(iters, iters_o))
    assert _rel(se.to_host(u), uo) < 1e-7
