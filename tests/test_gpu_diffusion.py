"""GPU parity of the Diffusion equation (diffusion.jl) through the C ABI:
the steady annulus Poisson problem of the p2d example and the unsteady
manufactured solution of the d2d example, each against the oracle's time loop
and the closed form; a viscosity set in place by setVisc! at every step
(which the fused operator must read); the Jacobi-preconditioned option; and
the reference's default absolute tolerance 1e-8."""
import math
import sys

import numpy as np
import pytest

from oracle import sem_oracle as o

pytestmark = pytest.mark.gpu

R0, R1 = 0.5, 1.0
KX = KY = KT = 2.0


def _log(msg):
    print(f"[case] {msg}", file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def se():
    import semhip
    return semhip


def _m(x):
    import torch
    return torch if isinstance(x, torch.Tensor) else np


def _one(a, x, y, t):
    a[...] = 1.0


def _utrue(x, y, t):
    s = _m(x).sin
    return s(KX * np.pi * x) * s(KY * np.pi * y) * math.cos(KT * np.pi * t)


def _setIC(u, x, y, t):
    u[...] = _utrue(x, y, t)


def _setF(f, x, y, t):
    s = _m(x).sin
    f[...] = (_utrue(x, y, t) * ((KX ** 2 + KY ** 2) * np.pi ** 2)
              - s(KX * np.pi * x) * s(KY * np.pi * y) * math.sin(KT * np.pi * t) * KT * np.pi)


def _setVisc(nu, x, y, t):
    nu[...] = 1.0 + 0.5 * _m(x).sin(np.pi * x) ** 2 * (1.0 + t)


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _radial(R, al, be):
    A = ((be + R1 ** 2 / 4) - (al + R0 ** 2 / 4)) / (math.log(R1) - math.log(R0))
    B = al + R0 ** 2 / 4 - A * math.log(R0)
    return -R ** 2 / 4 + A * np.log(R) + B


@pytest.mark.parametrize("al,be", [(0.0, 0.0), (0.3, -0.2)])
def test_annulus_poisson_matches_oracle_and_closed_form(se, al, be):
    _log(f"Diffusion annulus 8x8 pts, 5x5 elements, periodic in s, ub=({al},{be})")
    m = se.Mesh(8, 8, 5, 5, (False, True), se.annulus())
    mo = o.Mesh(8, 8, 5, 5, (False, True), o.annulus(), dense_gs=False)

    def setBC(ub, x, y, t):
        ub[...] = al + (be - al) * (_m(x).hypot(x, y) - R0) / (R1 - R0)
    dfn = se.Diffusion("DDNN", m, tol=1e-12)
    dfn.simulate_(lambda d: None, se.fixU, setBC, _one, _one)
    u = se.to_host(dfn.fld.u)
    uo, it_o = o.diffusion_simulate(mo, "DDNN", setBC=setBC, setForcing=_one, setVisc=_one,
                                    tol=1e-12)
    assert len(dfn.iters) == 1
    assert abs(dfn.iters[0] - it_o[0]) <= 2
    assert _rel(u, uo) < 1e-10
    assert np.abs(u - _radial(np.hypot(mo.x, mo.y), al, be)).max() < 1e-8


def _d2d_pair(se):
    m = se.Mesh(8, 8, 5, 5)
    mo = o.Mesh(8, 8, 5, 5, dense_gs=False)
    return m, mo


def test_d2d_steps_match_oracle_and_manufactured_solution(se):
    _log("Diffusion d2d 8x8 pts, 5x5 elements, dt=0.01, 20 steps, BDF3")
    m, mo = _d2d_pair(se)
    dfn = se.Diffusion("DDDD", m, Tf=0.2, dt=0.01, tol=1e-12)
    _setIC(dfn.fld.u, m.x, m.y, 0.0)
    dfn.run_(20, se.fixU, _setF, _one)
    u = se.to_host(dfn.fld.u)
    uo, it_o = o.diffusion_simulate(mo, "DDDD", Tf=0.2, dt=0.01, setIC=_setIC, setForcing=_setF,
                                    setVisc=_one, nsteps=20, tol=1e-12)
    assert _rel(u, uo) < 1e-10
    assert np.abs(u - _utrue(mo.x, mo.y, 0.2)).max() < 3e-6
    assert max(abs(a - b) for a, b in zip(dfn.iters, it_o)) <= 2


def test_simulate_with_reference_tolerance_and_step_count(se):
    _log("Diffusion d2d simulate_ Tf=0.05 dt=0.01 tol=1e-8 (reference default)")
    m, mo = _d2d_pair(se)
    seen = []
    dfn = se.Diffusion("DDDD", m, Tf=0.05, dt=0.01)
    dfn.simulate_(lambda d: seen.append(d.tstep.istep), _setIC, se.fixU, _setF, _one)
    uo, it_o = o.diffusion_simulate(mo, "DDDD", Tf=0.05, dt=0.01, setIC=_setIC,
                                    setForcing=_setF, setVisc=_one)
    assert seen == list(range(len(it_o) + 1))
    assert np.abs(se.to_host(dfn.fld.u) - uo).max() < 1e-7
    assert dfn.info.res <= 1e-8


@pytest.mark.parametrize("precond", ["identity", "jacobi"])
def test_time_dependent_viscosity_set_in_place(se, precond):
    _log(f"Diffusion deformed 7x7 pts, 4x4 elements, nu(x, t), precond={precond}")
    m = se.Mesh(7, 7, 4, 4, (False, False), se.deform_sinsin(0.1))
    mo = o.Mesh(7, 7, 4, 4, deform=o.deform_sinsin(0.1), dense_gs=False)
    dfn = se.Diffusion("DDDD", m, Tf=0.05, dt=0.01, precond=precond, tol=1e-12)
    _setIC(dfn.fld.u, m.x, m.y, 0.0)
    dfn.run_(5, se.fixU, _setF, _setVisc)
    uo, _ = o.diffusion_simulate(mo, "DDDD", Tf=0.05, dt=0.01, setIC=_setIC, setForcing=_setF,
                                 setVisc=_setVisc, nsteps=5, tol=1e-12)
    assert _rel(se.to_host(dfn.fld.u), uo) < 1e-10
    np.testing.assert_allclose(se.to_host(dfn.nu), 1.0 + 0.5 * np.sin(np.pi * mo.x) ** 2 * 1.05,
                               rtol=1e-14)
