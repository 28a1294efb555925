"""Pins the oracle's 1-D spectral primitives (GLL, derivMat, interpMat, semq,
semmesh, bdfExtK) with closed forms, library routines and exactness
properties, so that a dropped term, a wrong index or a transposed operand fails.
"""
import math

import numpy as np
import pytest

from oracle import sem_oracle as o


@pytest.mark.parametrize("n", list(range(2, 18)))
def test_gll_against_library_roots_and_quadrature(n):
    z, w = o.gausslobatto(n)
    N = n - 1
    assert z[0] == -1.0 and z[-1] == 1.0
    assert np.all(np.diff(z) > 0)
    np.testing.assert_array_equal(z, -z[::-1])
    # interior nodes are the roots of P_N' (numpy's Legendre class, an independent routine)
    if n > 2:
        roots = np.sort(np.polynomial.legendre.Legendre.basis(N).deriv().roots().real)
        np.testing.assert_allclose(z[1:-1], roots, rtol=0, atol=5e-15)
    # GLL quadrature is exact for degree <= 2N-1
    for k in range(0, 2 * N):
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        assert abs(np.dot(w, z ** k) - exact) < 4e-15 * max(1, n)
    # it is NOT exact for degree 2N (checks that the rule is Lobatto, not Gauss)
    k = 2 * N
    assert abs(np.dot(w, z ** k) - 2.0 / (k + 1)) > 1e-12


def test_gll_closed_forms():
    z, w = o.gausslobatto(3)
    np.testing.assert_allclose(z, [-1, 0, 1], atol=0)
    np.testing.assert_allclose(w, [1 / 3, 4 / 3, 1 / 3], rtol=1e-15)
    z, w = o.gausslobatto(4)
    s = 1 / math.sqrt(5)
    np.testing.assert_allclose(z, [-1, -s, s, 1], rtol=1e-15)
    np.testing.assert_allclose(w, [1 / 6, 5 / 6, 5 / 6, 1 / 6], rtol=1e-15)
    z, w = o.gausslobatto(5)
    s = math.sqrt(3 / 7)
    np.testing.assert_allclose(z, [-1, -s, 0, s, 1], rtol=1e-15, atol=1e-16)
    np.testing.assert_allclose(w, [1 / 10, 49 / 90, 32 / 45, 49 / 90, 1 / 10], rtol=1e-14)


def test_gll_weights_by_golub_welsch_lobatto():
    """Independent construction: Lobatto rule from the Jacobi matrix of the
    Legendre recurrence with the last entry modified (Golub 1973)."""
    for n in (5, 8, 9, 10, 16):
        N = n - 1
        beta = np.array([k / math.sqrt(4 * k * k - 1) for k in range(1, N)])
        # modified last off-diagonal for Lobatto endpoints +-1
        J = np.diag(beta, 1) + np.diag(beta, -1)
        Jf = np.zeros((n, n))
        Jf[:N, :N] = J
        bN2 = N / (2 * N - 1)  # beta_N^2 for Lobatto with endpoints +-1
        Jf[N - 1, N] = Jf[N, N - 1] = math.sqrt(bN2)
        lam, V = np.linalg.eigh(Jf)
        wg = 2 * V[0, :] ** 2
        z, w = o.gausslobatto(n)
        np.testing.assert_allclose(lam, z, atol=2e-14)
        np.testing.assert_allclose(wg, w, rtol=2e-12)


@pytest.mark.parametrize("n", [2, 3, 5, 8, 9, 10, 16])
def test_derivmat_exact_on_polynomials(n):
    z, _ = o.gausslobatto(n)
    D = o.derivMat(z)
    N = n - 1
    np.testing.assert_allclose(D @ np.ones(n), 0, atol=1e-13 * n * n)
    for k in range(1, n):
        np.testing.assert_allclose(D @ z ** k, k * z ** (k - 1), atol=2e-13 * n ** 2)
    # GLL corner closed forms
    assert abs(D[0, 0] + N * (N + 1) / 4) < 1e-13 * n * n
    assert abs(D[-1, -1] - N * (N + 1) / 4) < 1e-13 * n * n
    # transpose is wrong: D' x != 1 in general
    if n > 2:
        assert np.abs(D.T @ z - 1).max() > 1e-3


def test_derivmat_on_equispaced_points_matches_polyfit():
    x = np.linspace(-1, 1, 6)
    D = o.derivMat(x)
    for j in range(6):
        e = np.zeros(6)
        e[j] = 1
        c = np.polyfit(x, e, 5)
        np.testing.assert_allclose(D[:, j], np.polyval(np.polyder(c), x), atol=1e-10)


def test_interpmat_reproduces_polynomials():
    zi, _ = o.gausslobatto(8)
    zo, _ = o.gausslobatto(12)
    J = o.interpMat(zo, zi)
    assert J.shape == (12, 8)
    np.testing.assert_allclose(J.sum(axis=1), 1, atol=1e-14)
    for k in range(8):
        np.testing.assert_allclose(J @ zi ** k, zo ** k, atol=1e-13)
    np.testing.assert_allclose(o.interpMat(zi, zi), np.eye(8), atol=1e-15)
    # scalar target (the bdfExtK use)
    a = o.interpMat(2.0, np.array([1.0, 0.0]))
    np.testing.assert_allclose(a, [[2.0, -1.0]], atol=1e-15)


@pytest.mark.parametrize("E,n,per", [(1, 4, False), (3, 5, False), (3, 5, True), (4, 9, False), (1, 3, True)])
def test_semq_structure(E, n, per):
    Q = o.semq(E, n, per)
    ng = E * (n - 1) + (0 if per else 1)
    assert Q.shape == (E * n, ng)
    assert np.all(Q.sum(axis=1) == 1)  # every local node has exactly one global id
    # column sums = multiplicity: 2 on interfaces, 1 elsewhere
    colsum = Q.sum(axis=0)
    for g in range(ng):
        shared = (g % (n - 1) == 0) and (per or 0 < g < ng - 1)
        if E == 1 and per and g == 0:
            shared = True
        assert colsum[g] == (2 if shared else 1)


def test_semmesh_coordinates():
    for E, n in [(1, 2), (4, 9), (5, 8), (256, 8)]:
        z, w = o.semmesh(E, n)
        assert z[0] == -1.0 and z[-1] == 1.0
        assert abs(w.sum() - 2.0) < 1e-13
        # duplicated element end-points
        for e in range(E - 1):
            assert abs(z[e * n + n - 1] - z[(e + 1) * n]) < 1e-15
        assert np.all(np.diff(z) >= 0)


def test_bdfextk_closed_forms():
    dt = 1e-3
    a, b = o.bdfExtK(np.zeros(4))  # steady
    np.testing.assert_allclose(a, [1, 0, 0])
    np.testing.assert_allclose(b, [0, 0, 0, 0])
    a, b = o.bdfExtK(np.array([dt, 0, 0, 0]))
    np.testing.assert_allclose(a, [1, 0, 0], atol=1e-12)
    np.testing.assert_allclose(b * dt, [1, -1, 0, 0], atol=1e-12)
    a, b = o.bdfExtK(np.array([2 * dt, dt, 0, 0]))
    np.testing.assert_allclose(a, [2, -1, 0], atol=1e-12)
    np.testing.assert_allclose(b * dt, [1.5, -2, 0.5, 0], atol=1e-12)
    a, b = o.bdfExtK(np.array([3 * dt, 2 * dt, dt, 0]))
    np.testing.assert_allclose(a, [3, -3, 1], atol=1e-10)
    np.testing.assert_allclose(b * dt, [11 / 6, -3, 1.5, -1 / 3], atol=1e-10)
    a, b = o.bdfExtK(np.array([2 * dt, dt, 0]), k=2)
    np.testing.assert_allclose(a, [2, -1], atol=1e-12)
    np.testing.assert_allclose(b * dt, [1.5, -2, 0.5], atol=1e-12)


def test_timestepper_ramps_order():
    ts = o.TimeStepper(0.0, 1.0, 0.1, k=2)
    np.testing.assert_allclose(ts.bdfB, [0, 0, 0])
    ts.update()
    np.testing.assert_allclose(ts.bdfB * 0.1, [1, -1, 0], atol=1e-12)
    ts.update()
    np.testing.assert_allclose(ts.bdfB * 0.1, [1.5, -2, 0.5], atol=1e-12)
    np.testing.assert_allclose(ts.bdfA, [2, -1], atol=1e-12)
    assert ts.fstep == 10
