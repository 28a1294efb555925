"""GPU parity of the adjoint rows through the C ABI: the variable-coefficient
operator and its linsolve gradient (torch autograd over the adjoint PCG and
the dJ/dnu kernel) against the oracle and central differences; the
convection-diffusion stepper, its discrete adjoint against the oracle, the
checkpointed reverse sweep against the store-everything sweep (bit-exact:
every kernel is deterministic), and central differences on the cd2d mesh."""
import sys

import numpy as np
import pytest

from oracle import sem_oracle as o

pytestmark = pytest.mark.gpu


def _log(msg):
    print(f"[case] {msg}", file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def se():
    import semhip
    return semhip


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _vd_pair(se, n=7, E=4):
    m = se.Mesh(n, n, E, E, (False, False), se.deform_sinsin(0.1))
    mo = o.Mesh(n, n, E, E, deform=o.deform_sinsin(0.1), dense_gs=False)
    return m, mo


def _nu_host(mo, seed):
    from semhip.workloads import smooth_coefficient
    return mo.mult * mo.gs(smooth_coefficient(mo.x, mo.y, seed))


def test_vardiff_operator_matches_oracle(se):
    _log("vardiff operator 7x7 pts, 4x4 elements")
    m, mo = _vd_pair(se)
    M, Mo = se.generateMask("DDDD", m), o.generateMask("DDDD", mo)
    nu = _nu_host(mo, 41)
    u = np.random.default_rng(1).standard_normal(mo.x.shape)
    G = se.vardiff_factors(se.to_device(nu), m)
    np.testing.assert_allclose(se.to_host(G[0]), nu * mo.G11, rtol=1e-12, atol=1e-12 * abs(mo.G11).max())
    op = se.VarDiffusionOperator(se.to_device(nu), m, M)
    y = se.to_host(op(se.to_device(u)))
    ref = o.vardiff_operator(mo, nu, Mo)(u)
    assert _rel(y, ref) < 1e-12


def _linsolve_grad(se, m, M, nu_h, f_h, ustar, tol):
    import torch
    nu = se.to_device(nu_h).clone().requires_grad_(True)
    problem = se.VarDiffusion(m, M, se.to_device(f_h))
    solver = se.diffusion_solver(tol=tol)
    u = se.linsolve(nu, problem, solver)
    e = u - ustar
    J = 0.5 * torch.sum(m.B * e * e)
    J.backward()
    return float(J.detach()), u.detach(), nu.grad


def test_linsolve_gradient_matches_oracle(se):
    _log("linsolve gradient vs oracle, 7x7 pts, 4x4 elements")
    m, mo = _vd_pair(se)
    M, Mo = se.generateMask("DDDD", m), o.generateMask("DDDD", mo)
    nu, nus = _nu_host(mo, 41), _nu_host(mo, 42)
    f = np.ones_like(mo.x)
    A = o.vardiff_operator(mo, nus, Mo)
    b = mo.gs(Mo * (mo.B * f))
    ustar = o.pcg(b, A, mult=mo.mult, tol=1e-14 * np.abs(b).max())
    Jo, uo, lamo, go = o.vardiff_gradient(mo, nu, f, ustar, Mo, tol_rel=1e-14)
    J, u, g = _linsolve_grad(se, m, M, nu, f, se.to_device(ustar), 1e-14)
    assert abs(J - Jo) <= 1e-10 * abs(Jo)
    assert _rel(se.to_host(u), uo) < 1e-11
    assert _rel(se.to_host(g), go) < 1e-9


def _cd_pair(se, n=6, nd=9, E=3, dt=1e-2, nu=1e-2, k=2, precond="jacobi"):
    mV = se.Mesh(n, n, E, E, (False, False), se.deform_sinsin(0.1))
    mD = se.Mesh(nd, nd, E, E, (False, False), se.deform_sinsin(0.1))
    mVo = o.Mesh(n, n, E, E, deform=o.deform_sinsin(0.1), dense_gs=False)
    mDo = o.Mesh(nd, nd, E, E, deform=o.deform_sinsin(0.1), dense_gs=False)
    fld = se.Field("DDDD", mV, k)
    ts = se.TimeStepper(0.0, 1.0, dt, k)
    cdn = se.ConvectionDiffusion("cd", fld, -mV.y, mV.x, ts, mD, nu=nu, tol=1e-14,
                                 precond=precond)
    return cdn, mVo, mDo


def test_cd_reference_preconditioner_matches_oracle(se):
    """u ./ B ./ b0 (convectionDiffusion.jl:87-91) makes the PCG iterates
    discontinuous; the solve must still follow pcg.jl's recurrence step for
    step (iteration counts equal, states within 1e-11)."""
    _log("cd forward, precond=reference, 6 steps")
    cdn, mVo, mDo = _cd_pair(se, precond="reference")
    u0 = np.exp(-((mVo.x - 0.5) ** 2 + mVo.y ** 2) / (2 * 0.3 ** 2))
    cdn.fld.u.copy_(se.to_device(u0))
    cdn.run_(6)
    Mo = o.generateMask("DDDD", mVo)
    ref = o.cd_forward(mVo, mDo, u0, -mVo.y, mVo.x, 1e-2, 1e-2, 6, Mo, tol_rel=1e-14,
                       precond="reference")
    assert _rel(se.to_host(cdn.fld.u), ref) < 1e-11
    # the continuity check routes such a diagonal by itself as well
    d = se.Diagonal(1.0 / (cdn.mshV.B * 2.0))
    assert d.is_local(cdn.mshV)
    op = se.SEMOperator(1.0, 1.0, cdn.mshV, cdn.fld.M)
    assert not se.Diagonal(op.jacobi()).is_local(cdn.mshV)
    b = op(se.to_device(u0))
    with pytest.raises(se.SemError, match="alias"):
        se.pcg(b, op, opM=d, mult=cdn.mshV.mult, tol=1e-10, x=b)


def test_cd_steps_match_oracle(se):
    _log("cd forward 6x6 pts, 3x3 elements, 6 steps")
    cdn, mVo, mDo = _cd_pair(se)
    u0 = np.exp(-((mVo.x - 0.5) ** 2 + mVo.y ** 2) / (2 * 0.3 ** 2))
    cdn.fld.u.copy_(se.to_device(u0))
    cdn.run_(6)
    Mo = o.generateMask("DDDD", mVo)
    ref = o.cd_forward(mVo, mDo, u0, -mVo.y, mVo.x, 1e-2, 1e-2, 6, Mo, tol_rel=1e-14)
    assert _rel(se.to_host(cdn.fld.u), ref) < 1e-11


def test_cd_library_time_loop_matches_step_by_step(se):
    """sem_cd_steps (the library's step loop) issues the launches of the
    step-by-step path in the same order: states, histories, times, the
    objective and the iteration count are bit-identical, and both match the
    oracle."""
    import torch
    _log("cd library time loop vs step-by-step, 7 steps")
    cdA, mVo, mDo = _cd_pair(se)
    cdB, _, _ = _cd_pair(se)
    assert cdA.native()
    cdB.native = lambda: False          # the step-by-step host path
    u0 = np.exp(-((mVo.x - 0.5) ** 2 + mVo.y ** 2) / (2 * 0.3 ** 2))
    JA = torch.zeros(1, dtype=torch.float64, device="cuda")
    JB = torch.zeros(1, dtype=torch.float64, device="cuda")
    cdA.reset_(u0)
    cdB.reset_(u0)
    rec = []
    cdA.steps_(3, acc=JA, record=rec)
    cdA.steps_(4, acc=JA, record=rec)
    for n in range(7):
        cdB.step_()
        se.cd_objective_step(cdB, JB)
        np.testing.assert_array_equal(rec[n][0], cdB.tstep.bdfA)
        np.testing.assert_array_equal(rec[n][1], cdB.tstep.bdfB)
    np.testing.assert_array_equal(se.to_host(cdA.fld.u), se.to_host(cdB.fld.u))
    for ha, hb in zip(cdA.fld.uh, cdB.fld.uh):
        np.testing.assert_array_equal(se.to_host(ha), se.to_host(hb))
    assert cdA.tstep.time == cdB.tstep.time and cdA.tstep.istep == cdB.tstep.istep == 7
    assert float(JA) == float(JB) and cdA.iters_total == cdB.iters_total
    Mo = o.generateMask("DDDD", mVo)
    ref = o.cd_forward(mVo, mDo, u0, -mVo.y, mVo.x, 1e-2, 1e-2, 7, Mo, tol_rel=1e-14)
    assert _rel(se.to_host(cdA.fld.u), ref) < 1e-11


def test_cd_gradient_matches_oracle_and_checkpointing_is_exact(se):
    _log("cd adjoint 6x6 pts, 3x3 elements, 9 steps")
    cdn, mVo, mDo = _cd_pair(se)
    u0 = np.exp(-((mVo.x - 0.5) ** 2 + mVo.y ** 2) / (2 * 0.3 ** 2))
    Mo = o.generateMask("DDDD", mVo)
    Jo, go = o.cd_adjoint(mVo, mDo, u0, -mVo.y, mVo.x, 1e-2, 1e-2, 9, Mo, tol_rel=1e-14)
    r = se.cd_gradient(cdn, se.to_device(u0), 9)          # sqrt checkpoints: c = 3
    assert r.snapshots == 3 and r.peak_states == 3 and r.steps_forward == 18
    assert abs(r.J - Jo) <= 1e-11 * abs(Jo)
    assert _rel(se.to_host(r.grad), go) < 1e-9
    full = se.cd_gradient(cdn, se.to_device(u0), 9, checkpoints=1)
    assert full.J == r.J
    np.testing.assert_array_equal(se.to_host(full.grad), se.to_host(r.grad))


def test_cd_gradient_fd_cd2d_mesh(se):
    """The cd2d mesh (64x64 elements, N=9, dealiased p=15, v = (-y, x),
    nu = 1e-3, dt = 1e-3, blob sigma 0.1) over 40 steps: J is quadratic in u0,
    so central differences are exact up to the solver tolerance."""
    import torch
    from semhip import workloads as w
    _log("cd FD 64x64 N=9, 40 steps")
    mV = se.Mesh(10, 10, 64, 64, (False, False), se.deform_sinsin(0.1))
    mD = se.Mesh(15, 15, 64, 64, (False, False), se.deform_sinsin(0.1))
    fld = se.Field("DDDD", mV, 2)
    cdn = se.ConvectionDiffusion("cd2d", fld, -mV.y, mV.x, se.TimeStepper(0.0, 1.0, 1e-3, 2), mD,
                                 nu=1e-3, tol=1e-13)
    u0 = w.gaussian_blob(mV)
    N = 40
    r = se.cd_gradient(cdn, u0, N)
    for s in range(2):
        d = w.global_normal_field(mV, 5 + s)
        h = 1e-3
        fd = (se.cd_objective(cdn, u0 + h * d, N) - se.cd_objective(cdn, u0 - h * d, N)) / (2 * h)
        ad = float(torch.sum(r.grad * d))
        _log(f"  dir {s}: fd {fd:.12e} adjoint {ad:.12e} rel {abs(fd - ad) / abs(ad):.2e}")
        assert abs(fd - ad) <= 1e-8 * abs(ad)
