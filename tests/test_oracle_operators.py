"""Pins the oracle's operators and solver by brute force: dense kron assembly
(the p2d_explicit.jl:145-180 pattern), manufactured solutions, a direct solve
of the full-rank global system, and adjoint identities. A planted mistake in
the oracle (a transposed derivative matrix) must fail these checks."""
import numpy as np
import pytest
import scipy.linalg

from oracle import sem_oracle as o


def dense_ops(m):
    """Explicit matrices of p2d_explicit.jl:145-162 for mesh m."""
    nr, ns, Ex, Ey = m.nr, m.ns, m.Ex, m.Ey
    Dx1 = np.kron(np.eye(Ex), m.Dr)
    Dy1 = np.kron(np.eye(Ey), m.Ds)
    Dr = np.kron(np.eye(ns * Ey), Dx1)
    Ds = np.kron(Dy1, np.eye(nr * Ex))
    Drs = np.vstack([Dr, Ds])
    v = lambda a: a.reshape(-1, order="F")
    g11, g12, g22 = np.diag(v(m.G11)), np.diag(v(m.G12)), np.diag(v(m.G22))
    G = np.block([[g11, g12], [g12, g22]])
    A = Drs.T @ G @ Drs
    B = np.diag(v(m.B))
    Q = np.kron(o.semq(Ey, ns, m.ifperiodic[1]), o.semq(Ex, nr, m.ifperiodic[0]))
    return A, B, Q


def test_abu_matches_kron():
    rng = np.random.default_rng(1)
    u = rng.standard_normal((12, 15))
    Br = rng.standard_normal((5, 4))   # acts on groups of 4 rows -> 5 rows
    As = rng.standard_normal((2, 3))   # acts on blocks of 3 columns -> 2 columns
    v = o.ABu(As, Br, u)
    K = np.kron(np.kron(np.eye(5), As), np.kron(np.eye(3), Br))
    ref = (K @ u.reshape(-1, order="F")).reshape(15, 10, order="F")
    np.testing.assert_allclose(v, ref, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(o.ABu(None, Br, u),
                               (np.kron(np.eye(15), np.kron(np.eye(3), Br)) @ u.reshape(-1, order="F")).reshape(15, 15, order="F"))
    np.testing.assert_array_equal(o.ABu(None, None, u), u)


def test_element_contraction_is_a_u_bt():
    """test/Operators.jl:22 identity on one element: ABu(B, A, U) == A U B'."""
    rng = np.random.default_rng(2)
    A = rng.standard_normal((6, 6))
    Bm = rng.standard_normal((6, 6))
    U = rng.standard_normal((6, 6))
    np.testing.assert_allclose(o.ABu(Bm, A, U), A @ U @ Bm.T, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("per", [(False, False), (True, True)])
def test_laplace_matches_dense_assembly(per):
    m = o.Mesh(5, 5, 2, 3, per, deform=o.deform_sinsin(0.1))
    A, B, _ = dense_ops(m)
    rng = np.random.default_rng(3)
    u = rng.standard_normal(m.x.shape)
    ref = (A @ u.reshape(-1, order="F")).reshape(u.shape, order="F")
    np.testing.assert_allclose(o.laplace(u, m.Dr, m.Ds, m.G11, m.G12, m.G22), ref, rtol=1e-12, atol=1e-12)
    H = o.hlmz(u, 0.7, 1.3, m)
    np.testing.assert_allclose(H, 0.7 * ref + 1.3 * m.B * u, rtol=1e-12, atol=1e-12)


def test_laplace_null_space_and_symmetry():
    m = o.Mesh(8, 8, 3, 3, deform=o.deform_sinsin(0.1), dense_gs=False)
    one = np.ones(m.x.shape)
    assert np.abs(o.lapl(one, m)).max() < 1e-12
    # a linear function is also in the null space of the weak Laplacian's
    # element interior rows only; the constant is the exact check above.
    rng = np.random.default_rng(4)
    u = rng.standard_normal(m.x.shape)
    v = rng.standard_normal(m.x.shape)
    s1 = np.sum(v * o.lapl(u, m))
    s2 = np.sum(u * o.lapl(v, m))
    assert abs(s1 - s2) < 1e-11 * abs(s1)
    assert np.sum(u * o.lapl(u, m)) > 0


def test_planted_transposed_derivative_fails_dense_check():
    m = o.Mesh(5, 5, 2, 2, deform=o.deform_sinsin(0.1))
    A, _, _ = dense_ops(m)
    rng = np.random.default_rng(5)
    u = rng.standard_normal(m.x.shape)
    ref = (A @ u.reshape(-1, order="F")).reshape(u.shape, order="F")
    wrong = o.laplace(u, m.Dr.T, m.Ds, m.G11, m.G12, m.G22)
    assert np.abs(wrong - ref).max() > 1e-3


def test_jacobi_diag_matches_dense_assembled_diagonal():
    m = o.Mesh(5, 5, 3, 2, deform=o.deform_sinsin(0.1))
    M = o.generateMask("DDDD", m)
    nu, k = 0.8, 1.7
    A, B, Q = dense_ops(m)
    H = nu * A + k * B
    dg = np.diag(Q.T @ H @ Q)
    dl = (Q @ dg).reshape(m.x.shape, order="F")
    dinv = o.jacobi_diag(m, nu, k, M)
    np.testing.assert_allclose(dinv[M != 0], 1.0 / dl[M != 0], rtol=1e-12)
    assert np.all(dinv[M == 0] == 0)


def _poisson_solve(n, Ex, tol_rel=1e-12):
    m = o.Mesh(n, n, Ex, Ex, dense_gs=False)
    M = o.generateMask("DDDD", m)
    f = 2 * np.pi ** 2 * np.sin(np.pi * m.x) * np.sin(np.pi * m.y)
    b = m.gs(M * o.mass(f, m))
    info = o.PCGInfo()
    u = o.pcg(b, lambda v: o.opLHS(v, 1.0, 0.0, m, M), mult=m.mult,
              tol=tol_rel * np.abs(b).max(), info=info)
    return m, u, info


def test_manufactured_solution_converges_spectrally():
    errs = []
    for n in (5, 7, 9, 11):  # points per direction; the 4x4-element case is n = 9 (N = 8)
        m, u, info = _poisson_solve(n, 4)
        assert info.converged
        errs.append(np.abs(u - np.sin(np.pi * m.x) * np.sin(np.pi * m.y)).max())
    assert errs[2] < 2e-10    # N = 8: spectral accuracy
    for i in range(len(errs) - 1):
        assert errs[i + 1] < errs[i] / 20


def test_pcg_matches_direct_global_solve():
    """p2d_explicit.jl:166-180: the local rank-deficient PCG equals the
    full-rank restricted global system solved directly."""
    m = o.Mesh(5, 5, 3, 3, deform=o.deform_sinsin(0.1))
    M = o.generateMask("DDDD", m)
    nu, k = 1.0, 1.0
    A, B, Q = dense_ops(m)
    rng = np.random.default_rng(6)
    f = rng.standard_normal(m.x.shape)
    vf = f.reshape(-1, order="F")
    ng = Q.shape[1]
    nxg = m.Ex * (m.nr - 1) + 1
    gx, gy = np.arange(ng) % nxg, np.arange(ng) // nxg
    nyg = ng // nxg
    interior = (gx > 0) & (gx < nxg - 1) & (gy > 0) & (gy < nyg - 1)
    Rg = np.eye(ng)[interior]
    AA = Rg @ Q.T @ (nu * A + k * B) @ Q @ Rg.T
    bb = Rg @ Q.T @ B @ vf
    uu = (Q @ Rg.T @ scipy.linalg.solve(AA, bb, assume_a="pos")).reshape(f.shape, order="F")
    b = m.gs(M * o.mass(f, m))
    for opM in (None, "jacobi"):
        pre = None
        if opM == "jacobi":
            dinv = o.jacobi_diag(m, nu, k, M)
            pre = lambda r: dinv * r
        info = o.PCGInfo()
        u = o.pcg(b, lambda v: o.opLHS(v, nu, k, m, M), opM=pre, mult=m.mult,
                  tol=1e-13 * np.abs(b).max(), info=info)
        assert info.converged
        np.testing.assert_allclose(u, uu, rtol=0, atol=1e-11 * np.abs(uu).max())


def test_pcg_maxiter_warns_and_returns():
    m, _, _ = _poisson_solve(5, 2)
    M = o.generateMask("DDDD", m)
    b = m.gs(M * o.mass(np.ones(m.x.shape), m))
    info = o.PCGInfo()
    o.pcg(b, lambda v: o.opLHS(v, 1.0, 0.0, m, M), mult=m.mult, tol=1e-30, maxiter=3, info=info)
    assert info.iters == 3 and not info.converged


def test_grad_exact_and_gradT_adjoint():
    m = o.Mesh(6, 6, 2, 2, deform=o.deform_sinsin(0.1), dense_gs=False)
    ux, uy = o.grad(2 * m.x - 3 * m.y, m)
    np.testing.assert_allclose(ux, 2, atol=1e-11)
    np.testing.assert_allclose(uy, -3, atol=1e-11)
    rng = np.random.default_rng(8)
    u = rng.standard_normal(m.x.shape)
    vx = rng.standard_normal(m.x.shape)
    vy = rng.standard_normal(m.x.shape)
    gx, gy = o.grad(u, m)
    lhs = np.sum(gx * vx + gy * vy)
    rhs = np.sum(u * o.gradT(vx, vy, m))
    assert abs(lhs - rhs) < 1e-11 * abs(lhs)


def test_advect_of_linear_field_is_mass():
    """v = (1, 0), T = x on the undeformed mesh: the dealiased convective term
    of a constant derivative is J' B_D J 1 = B (GLL quadrature is exact)."""
    mV = o.Mesh(6, 6, 2, 2, dense_gs=False)
    mD = o.Mesh(9, 9, 2, 2, dense_gs=False)
    C = o.advect(mV.x, np.ones(mV.x.shape), np.zeros(mV.x.shape), mV, mD)
    np.testing.assert_allclose(C, mV.B, rtol=1e-12, atol=1e-14)


def test_dealiased_laplace_and_mass8_match_kron_brute_force():
    """laplace with interpolation (lapl.jl:83-103) and the 8-argument mass
    (mass.jl:32-50) against explicit Kronecker matrices on 2x2 elements; with
    the interpolation matrices set to the identity they reduce to laplace and
    B .* u."""
    m = o.Mesh(4, 4, 2, 2, deform=o.deform_sinsin(0.1), dense_gs=True)
    md = o.Mesh(6, 6, 2, 2, deform=o.deform_sinsin(0.1), dense_gs=False)
    Jr = o.interpMat(md.zr, m.zr)
    Js = o.interpMat(md.zs, m.zs)
    rng = np.random.default_rng(9)
    u = rng.standard_normal(m.x.shape)
    G = [rng.standard_normal(md.x.shape) for _ in range(3)]
    Bd = rng.standard_normal(md.x.shape)
    v = lambda a: a.reshape(-1, order="F")
    Dx = np.kron(np.eye(m.ns * m.Ey), np.kron(np.eye(m.Ex), m.Dr))
    Dy = np.kron(np.kron(np.eye(m.Ey), m.Ds), np.eye(m.nr * m.Ex))
    # (Js kron Jr) per element, in the local layout
    Jfull = np.kron(np.kron(np.eye(m.Ey), Js), np.kron(np.eye(m.Ex), Jr))
    g11, g12, g22 = (np.diag(v(g)) for g in G)
    Jur, Jus = Jfull @ Dx, Jfull @ Dy
    A = Jur.T @ (g11 @ Jur + g12 @ Jus) + Jus.T @ (g12 @ Jur + g22 @ Jus)
    ref = (A @ v(u)).reshape(u.shape, order="F")
    got = o.laplace_dealiased(u, Jr, Js, m.Dr, m.Ds, *G)
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    I = np.eye(m.nr)
    np.testing.assert_allclose(o.laplace_dealiased(u, I, I, m.Dr, m.Ds, m.G11, m.G12, m.G22),
                               o.laplace(u, m.Dr, m.Ds, m.G11, m.G12, m.G22), rtol=0, atol=1e-12)
    M = o.generateMask("DDDD", m)
    got = o.mass8(u, M, Bd, Jr, Js, m.QQtx, m.QQty, m.mult)
    ref = M * m.gs((Jfull.T @ np.diag(v(Bd)) @ Jfull @ v(u)).reshape(u.shape, order="F"))
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    np.testing.assert_allclose(o.mass8(u, M, m.B, I, I, m.QQtx, m.QQty, m.mult),
                               M * m.gs(m.B * u), rtol=0, atol=1e-14)
