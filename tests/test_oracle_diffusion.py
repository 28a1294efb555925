"""Pins the oracle's Diffusion time loop (diffusion.jl) by closed forms: the
steady Poisson problem of the annulus example (p2d.jl) against the radial
solution, with homogeneous and with inhomogeneous Dirichlet data carried by
ub (which pins the sign of the -nu lapl(ub) term and the final u += ub); the
unsteady manufactured solution of the d2d example (d2d.jl) at third order in
dt (BDF3 with the stepper's ramp); the reference's while-loop step count; and
a planted mistake (forcing evaluated at the old time) that must fail."""
import math

import numpy as np

from oracle import sem_oracle as o

R0, R1 = 0.5, 1.0
KX = KY = KT = 2.0


def _radial(R, al=0.0, be=0.0):
    """-u'' - u'/r = 1 on r0 < r < r1 with u(r0) = al, u(r1) = be."""
    A = ((be + R1 ** 2 / 4) - (al + R0 ** 2 / 4)) / (math.log(R1) - math.log(R0))
    B = al + R0 ** 2 / 4 - A * math.log(R0)
    return -R ** 2 / 4 + A * np.log(R) + B


def _one(a, x, y, t):
    a[...] = 1.0


def _annulus_mesh():
    return o.Mesh(8, 8, 5, 5, (False, True), o.annulus())


def test_annulus_poisson_matches_radial_solution():
    m = _annulus_mesh()
    R = np.hypot(m.x, m.y)
    u, iters = o.diffusion_simulate(m, "DDNN", setForcing=_one, setVisc=_one, tol=1e-12)
    assert len(iters) == 1
    assert np.abs(u - _radial(R)).max() < 1e-9


def test_annulus_inhomogeneous_dirichlet_through_ub():
    m = _annulus_mesh()
    al, be = 0.3, -0.2

    def setBC(ub, x, y, t):
        ub[...] = al + (be - al) * (np.hypot(x, y) - R0) / (R1 - R0)
    u, _ = o.diffusion_simulate(m, "DDNN", setForcing=_one, setVisc=_one, setBC=setBC,
                                tol=1e-12)
    assert np.abs(u - _radial(np.hypot(m.x, m.y), al, be)).max() < 1e-8


def _utrue(x, y, t):
    return np.sin(KX * np.pi * x) * np.sin(KY * np.pi * y) * np.cos(KT * np.pi * t)


def _setIC(u, x, y, t):
    u[...] = _utrue(x, y, t)


def _setF(f, x, y, t):
    f[...] = (_utrue(x, y, t) * ((KX ** 2 + KY ** 2) * np.pi ** 2)
              - np.sin(KX * np.pi * x) * np.sin(KY * np.pi * y) * np.sin(KT * np.pi * t) * KT * np.pi)


def _d2d_error(dt, n, setF=_setF):
    m = o.Mesh(8, 8, 5, 5)
    u, _ = o.diffusion_simulate(m, "DDDD", Tf=n * dt, dt=dt, setIC=_setIC, setForcing=setF,
                                setVisc=_one, nsteps=n, tol=1e-12)
    return np.abs(u - _utrue(m.x, m.y, n * dt)).max()


def test_d2d_manufactured_solution_third_order():
    e1, e2 = _d2d_error(0.02, 10), _d2d_error(0.01, 20)
    assert e2 < 3e-6
    assert e1 / e2 > 6.0


def test_forcing_at_old_time_is_caught():
    good = _d2d_error(0.01, 20)

    def stale(f, x, y, t):
        _setF(f, x, y, t - 0.01)
    assert _d2d_error(0.01, 20, stale) > 100 * good


def test_while_loop_step_count():
    m = o.Mesh(5, 5, 2, 2)
    _, iters = o.diffusion_simulate(m, "DDDD", setForcing=_one, setVisc=_one)
    assert len(iters) == 1                      # steady: one step, then the break
    dt, Tf = 0.01, 0.05
    t, n = 0.0, 0
    while t <= Tf:
        t += dt
        n += 1
    _, iters = o.diffusion_simulate(m, "DDDD", Tf=Tf, dt=dt, setForcing=_one, setVisc=_one)
    assert len(iters) == n
