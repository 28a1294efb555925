"""Pins the oracle's multigrid preconditioner (oracle/sem_mg.py) by what the
mathematics fixes: transfers that reproduce polynomials, restriction as the
exact transpose of prolongation, the Chebyshev error polynomial on
eigenvectors of the dense assembled pencil, the power estimate against dense
eigenvalues, a symmetric positive V-cycle, and PCG with it against a direct
solve. Planted mistakes (a restriction without the mult weights, a
prolongation with the sub-element order swapped) must fail."""
import numpy as np
import pytest
import scipy.linalg

from oracle import sem_oracle as o
from oracle import sem_mg as mg


def _rand_global(m, rng):
    return rng.standard_normal(m.nglobal)[m.l2g]


def _dense_pencil(m, M, nu, k):
    """Assembled global A (free nodes) and its diagonal, by applying opLHS to
    unit vectors of the global numbering."""
    free = np.unique(m.l2g[M != 0])
    A = np.zeros((free.size, free.size))
    for c, g in enumerate(free):
        e = (m.l2g == g).astype(float)
        y = o.opLHS(e, nu, k, m, M)
        for r_, g2 in enumerate(free):
            A[r_, c] = y[m.l2g == g2][0]
    return free, A


def test_prolongation_reproduces_polynomials():
    m = o.Mesh(8, 8, 4, 4, dense_gs=False)
    P = mg.MG(m, 1.0, 1.0, "DDDD")
    assert [lv.msh.nr for lv in P.levels] == [8, 4, 2, 2]
    assert [lv.msh.Ex for lv in P.levels] == [4, 4, 4, 2]
    for l in range(len(P.levels) - 1):
        f, c = P.levels[l], P.levels[l + 1]
        deg = c.msh.nr - 1
        poly = lambda x, y: (1 + x + 2 * y) ** deg + x ** deg * y ** deg - 3 * x * y
        got = mg.prolong(poly(c.msh.x, c.msh.y), c, f)
        np.testing.assert_allclose(got, poly(f.msh.x, f.msh.y), rtol=0, atol=1e-12 * 3 ** deg)


def test_planted_swapped_subelements_fail_reproduction():
    m = o.Mesh(2, 2, 4, 4, dense_gs=False)
    P = mg.MG(m, 1.0, 1.0, "DDDD")
    f, c = P.levels[0], P.levels[1]
    assert f.Jx.shape[0] == 2
    poly = lambda x, y: 1 + 2 * x + 3 * y + 4 * x * y
    good = mg.prolong(poly(c.msh.x, c.msh.y), c, f)
    np.testing.assert_allclose(good, poly(f.msh.x, f.msh.y), atol=1e-13)
    f.Jx = f.Jx[::-1].copy()
    bad = mg.prolong(poly(c.msh.x, c.msh.y), c, f)
    assert np.abs(bad - poly(f.msh.x, f.msh.y)).max() > 0.1


@pytest.mark.parametrize("per", [(False, False), (True, False)])
def test_restriction_is_transpose_of_prolongation(per):
    rng = np.random.default_rng(3)
    m = o.Mesh(6, 6, 4, 4, per, deform=o.deform_sinsin(0.1), dense_gs=False)
    bc = "DDDD"
    P = mg.MG(m, 1.0, 0.5, bc)
    for l in range(len(P.levels) - 1):
        f, c = P.levels[l], P.levels[l + 1]
        r = f.M * f.msh.gs(rng.standard_normal(f.msh.x.shape))        # assembled, masked
        v = c.M * _rand_global(c.msh, rng)                           # continuous, masked
        lhs = mg.wdot(mg.restrict(r, f, c), v, c.msh)
        rhs = mg.wdot(r, mg.prolong(v, c, f), f.msh)
        assert abs(lhs - rhs) <= 1e-12 * (abs(lhs) + abs(rhs))


def test_planted_unweighted_restriction_fails_transpose():
    rng = np.random.default_rng(4)
    m = o.Mesh(5, 5, 4, 4, dense_gs=False)
    P = mg.MG(m, 1.0, 1.0, "DDDD")
    f, c = P.levels[0], P.levels[1]
    r = f.M * f.msh.gs(rng.standard_normal(f.msh.x.shape))
    v = c.M * _rand_global(c.msh, rng)
    saved = f.msh.mult
    f.msh.mult = np.ones_like(saved)            # forget the partition of unity
    bad = mg.restrict(r, f, c)
    f.msh.mult = saved
    lhs = mg.wdot(bad, v, c.msh)
    rhs = mg.wdot(r, mg.prolong(v, c, f), f.msh)
    assert abs(lhs - rhs) > 1e-3 * abs(rhs)


def test_chebyshev_error_polynomial_on_eigenvectors():
    m = o.Mesh(4, 4, 2, 2, deform=o.deform_sinsin(0.1), dense_gs=False)
    M = o.generateMask("DDDD", m)
    nu, k = 1.0, 0.3
    free, A = _dense_pencil(m, M, nu, k)
    D = np.diag(np.diag(A))
    lam, V = scipy.linalg.eigh(A, D)
    dinv = o.jacobi_diag(m, nu, k, M)
    Aop = lambda v: o.opLHS(v, nu, k, m, M)
    lmax, lmin, deg = 1.05 * lam[-1], lam[-1] / 7.0, 4
    theta, delta = 0.5 * (lmax + lmin), 0.5 * (lmax - lmin)
    T = np.polynomial.chebyshev.Chebyshev.basis(deg)
    for q in (0, len(lam) // 2, len(lam) - 1):
        vg = np.zeros(m.nglobal)
        vg[free] = V[:, q]
        v = vg[m.l2g]
        b = Aop(v)
        x = mg.chebyshev(Aop, dinv, b, None, lmin, lmax, deg)
        tau = T((theta - lam[q]) / delta) / T(theta / delta)
        np.testing.assert_allclose(v - x, tau * v, rtol=0, atol=1e-11 * np.abs(v).max())


def test_power_estimate_against_dense_eigenvalues():
    m = o.Mesh(5, 5, 3, 3, deform=o.deform_sinsin(0.1), dense_gs=False)
    M = o.generateMask("DDDD", m)
    free, A = _dense_pencil(m, M, 1.0, 1.0)
    lam = scipy.linalg.eigh(A, np.diag(np.diag(A)), eigvals_only=True)
    P = mg.MG(m, 1.0, 1.0, "DDDD", mg.MGParams(npow=40))
    est = P.levels[0].lmax / P.params.safety
    assert 0.9 * lam[-1] <= est <= lam[-1] * (1 + 1e-12)


def test_vcycle_is_symmetric_positive():
    rng = np.random.default_rng(5)
    m = o.Mesh(8, 8, 4, 4, deform=o.deform_sinsin(0.1), dense_gs=False)
    P = mg.MG(m, 1.0, 1.0, "DDDD")
    M = P.levels[0].M
    a = M * m.gs(rng.standard_normal(m.x.shape))
    b = M * m.gs(rng.standard_normal(m.x.shape))
    Va, Vb = P(a), P(b)
    # the output is continuous and masked
    np.testing.assert_array_equal(Va, M * Va)
    np.testing.assert_allclose(Va, m.gs(m.mult * Va), rtol=0, atol=1e-13 * np.abs(Va).max())
    lhs, rhs = mg.wdot(Va, b, m), mg.wdot(a, Vb, m)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    assert mg.wdot(Va, a, m) > 0 and mg.wdot(Vb, b, m) > 0


def test_pcg_with_vcycle_matches_direct_solve():
    m = o.Mesh(6, 6, 4, 4, deform=o.deform_sinsin(0.1), dense_gs=False)
    M = o.generateMask("DDDD", m)
    nu, k = 1.0, 1.0
    free, A = _dense_pencil(m, M, nu, k)
    rng = np.random.default_rng(7)
    b = M * m.gs(m.B * _rand_global(m, rng))
    bg = np.array([b[m.l2g == g][0] for g in free])
    xg = np.zeros(m.nglobal)
    xg[free] = scipy.linalg.solve(A, bg, assume_a="pos")
    P = mg.MG(m, nu, k, "DDDD")
    info = o.PCGInfo()
    x = o.pcg(b, lambda v: o.opLHS(v, nu, k, m, M), opM=P, mult=m.mult,
              tol=1e-13 * np.abs(b).max(), info=info)
    assert info.converged and info.iters <= 12, info.iters
    np.testing.assert_allclose(x, xg[m.l2g], rtol=0, atol=1e-11 * np.abs(xg).max())
    infoj = o.PCGInfo()
    dinv = o.jacobi_diag(m, nu, k, M)
    o.pcg(b, lambda v: o.opLHS(v, nu, k, m, M), opM=lambda r: dinv * r, mult=m.mult,
          tol=1e-13 * np.abs(b).max(), info=infoj)
    assert infoj.iters > 3 * info.iters
