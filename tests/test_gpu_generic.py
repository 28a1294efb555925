"""Meshes the fused strip kernels do not take (nr != ns, or more than 16
points per direction; Mesh(nr, ns, ...) of mesh.jl:66-133 takes any order):
geometry, the element operator, the assembled operator, the Jacobi diagonal
and PCG on the generic element path against the oracle (FP64 within 1e-12;
solves within 1e-9 of the oracle's, iteration counts within 2)."""
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
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


CASES = [
    (5, 9, 4, 3, (False, False), "DDDD"),
    (18, 18, 2, 3, (False, False), "DDDD"),
    (20, 7, 3, 2, (True, False), "DDDD"),
    (6, 4, 5, 4, (False, True), "DNDN"),
    (40, 40, 1, 2, (False, False), "DDDD"),
]


@pytest.mark.parametrize("nr,ns,Ex,Ey,per,bc", CASES)
def test_generic_operator_matches_oracle(se, nr, ns, Ex, Ey, per, bc):
    _log(f"generic apply nr={nr} ns={ns} E={Ex}x{Ey} per={per} {bc}")
    m = se.Mesh(nr, ns, Ex, Ey, per, se.deform_sinsin(0.1))
    mo = o.Mesh(nr, ns, Ex, Ey, per, o.deform_sinsin(0.1), dense_gs=False)
    for name in ("Jac", "G11", "G12", "G22", "B"):
        ref = getattr(mo, name)
        assert np.abs(se.to_host(m.field(name)) - ref).max() <= 1e-12 * np.abs(ref).max()
    np.testing.assert_array_equal(se.to_host(m.mult), mo.mult)
    M = se.generateMask(bc, m)
    Mo = o.generateMask(bc, mo)
    rng = np.random.default_rng(nr * 31 + ns)
    u = rng.standard_normal(mo.x.shape)
    nu, k = 0.8, 1.3
    he = se.to_host(se.hlmz(se.to_device(u), nu, k, m))
    assert _rel(he, o.hlmz(u, nu, k, mo)) < 1e-12
    op = se.SEMOperator(nu, k, m, M)
    y = se.to_host(op(se.to_device(u)))
    assert _rel(y, o.opLHS(u, nu, k, mo, Mo)) < 1e-12
    assert np.all(y[Mo == 0] == 0)
    d = se.to_host(op.jacobi())
    do = o.jacobi_diag(mo, nu, k, Mo)
    assert _rel(d, do) < 1e-12
    v = rng.standard_normal(mo.x.shape)
    w = se.wdot(se.to_device(u), se.to_device(v), m)
    assert abs(w - np.sum(u * v * mo.mult)) <= 1e-12 * np.sum(np.abs(u * v))


@pytest.mark.parametrize("nr,ns,Ex,Ey,per,bc", CASES[:3])
def test_generic_pcg_matches_oracle(se, nr, ns, Ex, Ey, per, bc):
    _log(f"generic pcg nr={nr} ns={ns}")
    m = se.Mesh(nr, ns, Ex, Ey, per, se.deform_sinsin(0.1))
    mo = o.Mesh(nr, ns, Ex, Ey, per, o.deform_sinsin(0.1), dense_gs=False)
    M = se.generateMask(bc, m)
    Mo = o.generateMask(bc, mo)
    rng = np.random.default_rng(ns)
    b = mo.gs(Mo * o.mass(rng.standard_normal(mo.x.shape), mo))
    tol = 1e-11 * np.abs(b).max()
    op = se.SEMOperator(1.0, 1.0, m, M)
    d = op.jacobi()
    info = se.PCGInfo()
    x = se.to_host(se.pcg(se.to_device(b), op, opM=se.Diagonal(d), mult=m.mult, tol=tol,
                          info=info))
    do = o.jacobi_diag(mo, 1.0, 1.0, Mo)
    info_o = o.PCGInfo()
    xo = o.pcg(b, lambda v: o.opLHS(v, 1.0, 1.0, mo, Mo), opM=lambda r: do * r, mult=mo.mult,
               tol=tol, info=info_o)
    assert info.converged and abs(info.iters - info_o.iters) <= 2, (info.iters, info_o.iters)
    assert _rel(x, xo) < 1e-9
    with pytest.raises(se.SemError, match="UNSUPPORTED"):
        se.SEMOperator(1.0, 1.0, m, M, precision="f32")


def test_generic_block_beyond_lds_is_rejected(se):
    """41 x 41 blocks need more than 64 KiB of LDS in the element kernel: the
    operator and hlmz return SEM_ERR_UNSUPPORTED before any launch."""
    _log("generic 41x41 rejected")
    m = se.Mesh(41, 41, 1, 1, (False, False), se.deform_sinsin(0.1))
    M = se.generateMask("DDDD", m)
    op = se.SEMOperator(1.0, 1.0, m, M)
    with pytest.raises(se.SemError, match="UNSUPPORTED"):
        op(m.x)
    with pytest.raises(se.SemError, match="UNSUPPORTED"):
        se.hlmz(m.x, 1.0, 1.0, m)
