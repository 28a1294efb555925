"""The library on the GPU against the committed golden fixtures: numbering,
multiplicity and mask bit-exact; deformed geometry, one Helmholtz apply within
1e-12; one periodic gather-scatter value-exact; the 4x4-element N=8 Poisson Jacobi-PCG
solution within the solve tolerance and its iteration count within 1%."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden.npz")


@pytest.fixture(scope="module")
def gold():
    with np.load(GOLD, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="module")
def se():
    import semhip
    return semhip


def _rel(a, b):
    return np.abs(np.asarray(a) - b).max() / max(np.abs(b).max(), 1e-300)


def test_poisson_4x4_N8_structure_and_solve(se, gold):
    m = se.Mesh(9, 9, 4, 4, (False, False), se.fixU)
    M = se.generateMask("DDDD", m)
    np.testing.assert_array_equal(se.to_host(m.l2g.double()).astype(np.int32), gold["c1_l2g"])
    np.testing.assert_array_equal(se.to_host(m.mult), gold["c1_mult"])
    np.testing.assert_array_equal(se.to_host(M), gold["c1_mask"])
    x, y = se.to_host(m.x), se.to_host(m.y)
    f = 2 * np.pi ** 2 * np.sin(np.pi * x) * np.sin(np.pi * y)
    b = se.gatherScatter(se.mask(se.mass(se.to_device(f), m), M), m)
    op = se.SEMOperator(1.0, 0.0, m, M)
    info = se.PCGInfo()
    xs = se.pcg(b, op, opM=se.Diagonal(op.jacobi()), mult=m.mult,
                tol=1e-10 * float(b.abs().max()), info=info)
    assert info.converged
    assert abs(info.iters - int(gold["c1_pcg_iters"])) <= max(1, 0.01 * int(gold["c1_pcg_iters"]))
    assert _rel(se.to_host(xs), gold["c1_pcg_x"]) < 1e-9


def test_deformed_geometry_and_hlmz(se, gold):
    m = se.Mesh(8, 8, 2, 2, (False, False), se.deform_sinsin(0.1))
    for k in ("x", "y", "B", "G11", "G12", "G22", "Jac"):
        assert _rel(se.to_host(m.field(k)), gold[f"d_{k}"]) < 1e-12, k
    y = se.to_host(se.hlmz(se.to_device(gold["d_u"]), 1.3, 0.7, m))
    assert _rel(y, gold["d_hlmz"]) < 1e-12


def test_periodic_gather_scatter_value_exact(se, gold):
    m = se.Mesh(5, 5, 3, 2, (True, False), se.deform_affine01)
    y = se.to_host(se.gatherScatter(se.to_device(gold["p_u"]), m))
    np.testing.assert_array_equal(y, gold["p_gs"])
