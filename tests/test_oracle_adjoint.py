"""Pins the oracle's adjoint pieces: transposes by inner-product identities,
the variable-coefficient gradient by finite differences and by exact
linearity of the operator in nu, and the convection-diffusion stepper by
closed-form trajectories (linear-in-time advection, an eigenmode's BDF decay)
and its discrete adjoint by central differences (the objective is quadratic in
u0, so the central difference is exact up to rounding). A planted sign error
in the oracle's adjoint must fail the FD check."""
import numpy as np
import pytest

from oracle import sem_oracle as o


def _meshes(nr=6, nd=9, E=3, amp=0.1):
    d = o.deform_sinsin(amp) if amp else o.fixU
    mV = o.Mesh(nr, nr, E, E, deform=d, dense_gs=False)
    mD = o.Mesh(nd, nd, E, E, deform=d, dense_gs=False)
    return mV, mD


def test_advect_transpose_inner_product():
    mV, mD = _meshes()
    rng = np.random.default_rng(1)
    T, w = rng.standard_normal(mV.x.shape), rng.standard_normal(mV.x.shape)
    vx, vy = -mV.y, mV.x
    a = np.sum(o.advect(T, vx, vy, mV, mD) * w)
    b = np.sum(T * o.advect_transpose(w, vx, vy, mV, mD))
    assert abs(a - b) <= 1e-13 * abs(a)


def _nu_field(m, seed):
    rng = np.random.default_rng(seed)
    phi = np.zeros_like(m.x)
    for _ in range(4):
        kx, ky = rng.integers(1, 3, size=2)
        phi += 0.25 * np.cos(np.pi * (kx * m.x + ky * m.y) + rng.uniform(0, 2 * np.pi))
    return m.mult * m.gs(1.0 + 0.5 * np.tanh(phi))


def test_dnu_is_exact_derivative_of_linear_operator():
    """lam' A(nu) u is linear in nu, so a central difference is exact; it must
    equal sum(delta .* B .* (grad lam . grad u))."""
    m = o.Mesh(5, 5, 2, 2, deform=o.deform_sinsin(0.1), dense_gs=False)
    rng = np.random.default_rng(3)
    nu = _nu_field(m, 7)
    u, lam, dlt = (rng.standard_normal(m.x.shape) for _ in range(3))
    f = lambda n: np.sum(lam * o.laplace(u, m.Dr, m.Ds, n * m.G11, n * m.G12, n * m.G22))
    fd = (f(nu + 0.5 * dlt) - f(nu - 0.5 * dlt))
    lx, ly = o.grad(lam, m)
    ux, uy = o.grad(u, m)
    ad = np.sum(dlt * m.B * (lx * ux + ly * uy))
    assert abs(fd - ad) <= 1e-12 * abs(ad)


def test_vardiff_gradient_matches_central_differences():
    m = o.Mesh(6, 6, 3, 3, deform=o.deform_sinsin(0.1), dense_gs=False)
    M = o.generateMask("DDDD", m)
    nu, nus = _nu_field(m, 41), _nu_field(m, 42)
    f = np.ones_like(m.x)
    A = o.vardiff_operator(m, nus, M)
    b = m.gs(M * (m.B * f))
    ustar = o.pcg(b, A, mult=m.mult, tol=1e-14 * np.abs(b).max())
    J, u, lam, g = o.vardiff_gradient(m, nu, f, ustar, M, tol_rel=1e-14)
    rng = np.random.default_rng(43)
    for _ in range(3):
        d = m.mult * m.gs(rng.standard_normal(m.x.shape))
        h = 1e-5   # J is not quadratic in nu: the O(h^2) term is ~1e-6 at h = 1e-4
        Jp = o.vardiff_gradient(m, nu + h * d, f, ustar, M, tol_rel=1e-14)[0]
        Jm = o.vardiff_gradient(m, nu - h * d, f, ustar, M, tol_rel=1e-14)[0]
        fd = (Jp - Jm) / (2 * h)
        ad = np.sum(g * d)
        assert abs(fd - ad) <= 1e-7 * abs(ad), (fd, ad)


def test_advection_of_linear_field_is_exact_in_time():
    """u0 = x, v = (1, 0), nu = 0, no Dirichlet nodes: the exact solution
    x - t is linear in time and space, so BDF1/EXT1 then BDF2/EXT2 reproduce
    it to rounding (fixes the sign of advect and the EXT coefficients)."""
    mV, mD = _meshes(amp=0.0)
    M = o.generateMask("NNNN", mV)
    dt = 0.05
    _, st = o.cd_forward(mV, mD, mV.x.copy(), np.ones_like(mV.x), np.zeros_like(mV.x), 0.0, dt,
                         4, M, tol_rel=1e-15, keep=True)
    for n, u in enumerate(st):
        np.testing.assert_allclose(u, mV.x - n * dt, rtol=0, atol=1e-12)


def test_diffusion_eigenmode_follows_scalar_bdf_recurrence():
    """Pure diffusion of sin(pi(x+1)/2) sin(pi(y+1)/2): each step multiplies
    the mode by the scalar BDF1 then BDF2 amplification for lambda = nu pi^2/2
    (spatial error of the order-9 basis is far below the checked level)."""
    m = o.Mesh(10, 10, 4, 4, dense_gs=False)
    mD = o.Mesh(15, 15, 4, 4, dense_gs=False)
    M = o.generateMask("DDDD", m)
    mode = np.sin(np.pi * (m.x + 1) / 2) * np.sin(np.pi * (m.y + 1) / 2)
    nu, dt, N = 0.1, 0.02, 6
    z = np.zeros_like(m.x)
    _, st = o.cd_forward(m, mD, mode, z, z, nu, dt, N, M, tol_rel=1e-14, keep=True)
    lam = nu * np.pi ** 2 / 2
    g = [1.0, 1.0 / (1.0 + lam * dt)]
    for _ in range(2, N + 1):
        g.append((2 * g[-1] - 0.5 * g[-2]) / (1.5 + lam * dt))
    for n in range(N + 1):
        np.testing.assert_allclose(st[n], g[n] * mode, rtol=0, atol=2e-8)


def _cd_case():
    mV, mD = _meshes()
    M = o.generateMask("DDDD", mV)
    u0 = np.exp(-((mV.x - 0.5) ** 2 + mV.y ** 2) / (2 * 0.3 ** 2))
    return mV, mD, M, u0, -mV.y, mV.x


def test_cd_adjoint_matches_central_differences():
    mV, mD, M, u0, vx, vy = _cd_case()
    dt, N, nu = 1e-2, 10, 1e-2
    J, g = o.cd_adjoint(mV, mD, u0, vx, vy, nu, dt, N, M, tol_rel=1e-14)
    rng = np.random.default_rng(5)
    for _ in range(3):
        d = rng.standard_normal(u0.shape)
        h = 1e-3
        Jp = o.cd_objective(o.cd_forward(mV, mD, u0 + h * d, vx, vy, nu, dt, N, M, tol_rel=1e-14,
                                         keep=True)[1], mV, dt)
        Jm = o.cd_objective(o.cd_forward(mV, mD, u0 - h * d, vx, vy, nu, dt, N, M, tol_rel=1e-14,
                                         keep=True)[1], mV, dt)
        fd = (Jp - Jm) / (2 * h)
        ad = np.sum(g * d)
        assert abs(fd - ad) <= 1e-9 * abs(ad), (fd, ad)


def test_planted_adjoint_sign_error_fails_fd(monkeypatch):
    mV, mD, M, u0, vx, vy = _cd_case()
    dt, N, nu = 1e-2, 4, 1e-2
    good = o.advect_transpose
    monkeypatch.setattr(o, "advect_transpose", lambda *a, **k: -good(*a, **k))
    J, g = o.cd_adjoint(mV, mD, u0, vx, vy, nu, dt, N, M, tol_rel=1e-14)
    monkeypatch.undo()
    d = np.random.default_rng(6).standard_normal(u0.shape)
    h = 1e-3
    Jp = o.cd_objective(o.cd_forward(mV, mD, u0 + h * d, vx, vy, nu, dt, N, M, tol_rel=1e-14,
                                     keep=True)[1], mV, dt)
    Jm = o.cd_objective(o.cd_forward(mV, mD, u0 - h * d, vx, vy, nu, dt, N, M, tol_rel=1e-14,
                                     keep=True)[1], mV, dt)
    fd = (Jp - Jm) / (2 * h)
    assert abs(fd - np.sum(g * d)) > 1e-6 * abs(fd)


def test_reference_preconditioner_leaves_continuous_space():
    """The reference's u ./ B ./ b0 preconditioner yields discontinuous PCG
    iterates at shared nodes (documented defect); Jacobi keeps them continuous."""
    mV, mD, M, u0, vx, vy = _cd_case()
    dt, nu = 1e-2, 1e-2
    ref = o.cd_forward(mV, mD, u0, vx, vy, nu, dt, 1, M, tol_rel=1e-14, precond="reference")
    jac = o.cd_forward(mV, mD, u0, vx, vy, nu, dt, 1, M, tol_rel=1e-14)
    jump = lambda u: np.abs(mV.mult * mV.gs(u) - u).max()
    assert jump(jac) < 1e-13
    assert jump(ref) > 1e-8
