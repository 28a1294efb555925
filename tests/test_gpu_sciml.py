"""The Spectral submodule's operator interface over the device kernels
(semhip.sciml): the identities of the reference's own test/Operators.jl
(mul!(v, D, u) = d .* u; mul!(v, T, u) = A U B') against NumPy, the inverse,
adjoint, fusion and SciML call forms, and the fused SEM operator and the
multigrid V-cycle through mul!."""
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


def test_reference_operator_tests(se):
    """test/Operators.jl:1-24 with seeded inputs."""
    _log("sciml reference identities")
    S = se.sciml
    nr, ns = 8, 12
    rng = np.random.default_rng(0)
    u, d = rng.random((nr, ns)), rng.random((nr, ns))
    U, V = S.Field(u), S.Field(rng.random((nr, ns)))
    D = S.DiagonalOp(S.Field(d))
    np.testing.assert_allclose(se.to_host(S.mul_(V, D, U)), d * u, rtol=1e-15)
    A, B = rng.random((nr, nr)), rng.random((ns, ns))
    T = S.TensorProductOp2D(A, B)
    np.testing.assert_allclose(se.to_host(S.mul_(V, T, U)), A @ u @ B.T, rtol=1e-13)
    np.testing.assert_allclose(se.to_host(S.mul_(V, T, U)), A @ u @ B.T, rtol=1e-13)


def test_operator_algebra(se):
    _log("sciml algebra")
    S = se.sciml
    nr, ns = 6, 9
    rng = np.random.default_rng(1)
    u = rng.standard_normal((nr, ns))
    d1, d2 = 1 + rng.random((nr, ns)), 1 + rng.random((nr, ns))
    A, B = rng.standard_normal((nr, nr)), rng.standard_normal((ns, ns))
    A2, B2 = rng.standard_normal((nr, nr)), rng.standard_normal((ns, ns))
    U = S.Field(u)
    D1, D2 = S.DiagonalOp(S.Field(d1)), S.DiagonalOp(S.Field(d2))
    T1, T2 = S.TensorProductOp2D(A, B), S.TensorProductOp2D(A2, B2)
    h = lambda t: se.to_host(t)
    np.testing.assert_allclose(h(S.inv(D1) * U), u / d1, rtol=1e-15)
    np.testing.assert_allclose(h(S.ldiv_(S.Field(0 * u), D1, U)), u / d1, rtol=1e-15)
    np.testing.assert_allclose(h((D1 + D2) * U), (d1 + d2) * u, rtol=1e-15)
    np.testing.assert_allclose(h((D1 * D2) * U), d1 * d2 * u, rtol=1e-15)
    np.testing.assert_allclose(h((2.0 * D1 - D2) * U), (2 * d1 - d2) * u, rtol=1e-14)
    np.testing.assert_allclose(h((T1 * T2) * U), A @ A2 @ u @ (B @ B2).T, rtol=1e-12)
    np.testing.assert_allclose(h((T1 + T2) * U), A @ u @ B.T + A2 @ u @ B2.T, rtol=1e-12)
    np.testing.assert_allclose(h(T1.adjoint() * U), A.T @ u @ B, rtol=1e-12)
    np.testing.assert_allclose(h((T1 * D1) * U), A @ (d1 * u) @ B.T, rtol=1e-12)
    # the SciML call forms A(u, p, t) and A(du, u, p, t)
    np.testing.assert_allclose(h(T1(U, None, 0.0)), A @ u @ B.T, rtol=1e-12)
    du = S.Field(0 * u)
    T1(du, U, None, 0.0)
    np.testing.assert_allclose(h(du), A @ u @ B.T, rtol=1e-12)
    # rectangular tensor product (an interpolation operator)
    J = rng.standard_normal((nr + 3, nr))
    K = rng.standard_normal((ns + 2, ns))
    np.testing.assert_allclose(h(S.TensorProductOp2D(J, K) * U), J @ u @ K.T, rtol=1e-12)
    M = S.MatrixOp(A)
    v = rng.standard_normal(nr)
    np.testing.assert_allclose(h(M * v).reshape(-1), A @ v, rtol=1e-12)


def test_sem_operator_and_vcycle_through_mul(se):
    _log("sciml SEM operator")
    S = se.sciml
    n, E = 7, 4
    m = se.Mesh(n, n, E, E, (False, False), se.deform_sinsin(0.1))
    mo = o.Mesh(n, n, E, E, deform=o.deform_sinsin(0.1), dense_gs=False)
    M = se.generateMask("DDDD", m)
    Mo = o.generateMask("DDDD", mo)
    op = se.SEMOperator(1.0, 1.0, m, M)
    rng = np.random.default_rng(2)
    u = rng.standard_normal(mo.x.shape)
    v = S.Field(0 * u)
    S.mul_(v, op, S.Field(u))
    ref = o.opLHS(u, 1.0, 1.0, mo, Mo)
    assert np.abs(se.to_host(v) - ref).max() <= 1e-12 * np.abs(ref).max()
    # composed with a diagonal: (op * D) u
    d = 1 + rng.random(mo.x.shape)
    w = S.mul_(S.Field(0 * u), S.ComposeOp(op, S.DiagonalOp(S.Field(d))), S.Field(u))
    ref2 = o.opLHS(d * u, 1.0, 1.0, mo, Mo)
    assert np.abs(se.to_host(w) - ref2).max() <= 1e-12 * np.abs(ref2).max()
    # the V-cycle as an operator
    P = se.MG(op)
    r = S.Field(Mo * mo.gs(rng.standard_normal(mo.x.shape)))
    z1 = S.mul_(S.Field(0 * u), P, r)
    np.testing.assert_array_equal(se.to_host(z1), se.to_host(P(r)))
