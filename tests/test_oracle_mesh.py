"""Pins the oracle's mesh data: numbering/connectivity (bit-exact against the
dense semq matrices), mask (against the restriction-matrix construction of
mesh.jl:153-171), multiplicity, gather-scatter (structural vs the dense GEMM
form of gatherScatter.jl:13) and the geometric factors (closed forms on the
undeformed mesh, the analytic Jacobian of the sin-sin map)."""
import numpy as np
import pytest

from oracle import sem_oracle as o

CASES = [
    (3, 3, 2, 3, (False, False)),
    (5, 5, 3, 2, (True, False)),
    (4, 4, 2, 3, (False, True)),
    (4, 4, 3, 3, (True, True)),
    (3, 3, 1, 1, (True, True)),
    (9, 9, 4, 4, (False, False)),
]


@pytest.mark.parametrize("nr,ns,Ex,Ey,per", CASES)
def test_l2g_matches_dense_q(nr, ns, Ex, Ey, per):
    Qx = o.semq(Ex, nr, per[0])
    Qy = o.semq(Ey, ns, per[1])
    Q = np.kron(Qy, Qx)  # vec(u) column-major: x fastest
    l2g, ng = o.local_to_global(Ex, Ey, nr, ns, per)
    assert ng == Q.shape[1]
    g_dense = np.argmax(Q, axis=1)
    np.testing.assert_array_equal(l2g.reshape(-1, order="F"), g_dense)


@pytest.mark.parametrize("nr,ns,Ex,Ey,per", CASES)
def test_gs_structural_equals_dense(nr, ns, Ex, Ey, per):
    m = o.Mesh(nr, ns, Ex, Ey, per)
    rng = np.random.default_rng(7)
    u = rng.standard_normal((nr * Ex, ns * Ey))
    d = o.gatherScatter(u, m.QQtx, m.QQty)
    s = o.gatherScatter_structural(u, m)
    np.testing.assert_array_equal(d, s)
    # QQ' is Q Q': equals scattering the gathered global sums
    Q = np.kron(o.semq(Ey, ns, per[1]), o.semq(Ex, nr, per[0]))
    ref = (Q @ (Q.T @ u.reshape(-1, order="F"))).reshape(u.shape, order="F")
    np.testing.assert_allclose(s, ref, rtol=1e-15, atol=1e-15)


@pytest.mark.parametrize("nr,ns,Ex,Ey,per", CASES)
def test_mult_is_inverse_multiplicity(nr, ns, Ex, Ey, per):
    m = o.Mesh(nr, ns, Ex, Ey, per)
    counts = np.bincount(m.l2g.reshape(-1), minlength=m.nglobal)
    np.testing.assert_array_equal(m.mult, 1.0 / counts[m.l2g])
    assert set(np.unique(m.mult)).issubset({1.0, 0.5, 0.25})


@pytest.mark.parametrize("bc", ["DDDD", "DNDD", "NNDD", "NDNN", "NNNN"])
@pytest.mark.parametrize("per", [(False, False), (True, False), (False, True)])
def test_mask_matches_restriction_construction(bc, per):
    m = o.Mesh(4, 4, 3, 2, per, dense_gs=False)
    nx, ny = 12, 8
    ix = list(range(nx))
    iy = list(range(ny))
    if bc[0] == "D":
        ix = ix[1:]
    if bc[1] == "D":
        ix = ix[:-1]
    if bc[2] == "D":
        iy = iy[1:]
    if bc[3] == "D":
        iy = iy[:-1]
    if per[0]:
        ix = list(range(nx))
    if per[1]:
        iy = list(range(ny))
    Rx = np.eye(nx)[ix, :]
    Ry = np.eye(ny)[iy, :]
    M = np.outer(np.diag(Rx.T @ Rx), np.diag(Ry.T @ Ry))
    np.testing.assert_array_equal(o.generateMask(bc, m), (M == 1).astype(float))


def test_undeformed_geometry_closed_forms():
    Ex = Ey = 4
    n = 9
    m = o.Mesh(n, n, Ex, Ey, dense_gs=False)
    h = 2.0 / Ex
    np.testing.assert_allclose(m.Jac, (h / 2) ** 2, rtol=1e-13)
    np.testing.assert_allclose(m.rx, 2 / h, rtol=1e-13)
    np.testing.assert_allclose(m.sy, 2 / h, rtol=1e-13)
    np.testing.assert_allclose(m.ry, 0, atol=1e-12)
    np.testing.assert_allclose(m.sx, 0, atol=1e-12)
    ww = np.outer(np.tile(m.wr, Ex), np.tile(m.ws, Ey))
    np.testing.assert_allclose(m.G11, ww, rtol=1e-12)
    np.testing.assert_allclose(m.G22, ww, rtol=1e-12)
    np.testing.assert_allclose(m.G12, 0, atol=1e-12)
    assert abs(m.B.sum() - 4.0) < 1e-12


def test_deformed_jacobian_converges_to_analytic():
    amp = 0.1
    errs = []
    for n in (5, 7, 9, 11):
        Ex = 4
        m = o.Mesh(n, n, Ex, Ex, deform=o.deform_sinsin(amp), dense_gs=False)
        xe, _ = o.semmesh(Ex, n)
        X, Y = o.ndgrid(xe, xe)
        h = 2.0 / Ex
        Jex = (h / 2) ** 2 * (1 + amp * np.pi * np.sin(np.pi * (X + Y)))
        errs.append(np.abs(m.Jac - Jex).max() / (h / 2) ** 2)
        assert abs(m.B.sum() - 4.0) < 10 * errs[-1] + 1e-12
    assert errs[-1] < 1e-7
    assert all(errs[i + 1] < errs[i] / 5 for i in range(len(errs) - 1))
