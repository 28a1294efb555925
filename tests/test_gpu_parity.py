"""GPU parity of the HIP path against the oracle (called through the C ABI
via the semhip mirror). Integer/structural data (numbering, mask, mult) and
the assembly order are checked bit-exactly; FP64 operators within 1e-12
relative; converged solves within the tolerance stated per test."""
import os
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


@pytest.fixture
def rows_env():
    old = os.environ.get("SEM_AX_ROWS")
    yield
    if old is None:
        os.environ.pop("SEM_AX_ROWS", None)
    else:
        os.environ["SEM_AX_ROWS"] = old


def _pair(se, n, E, per=(False, False), deform="sinsin", Ey=None):
    Ey = Ey or E
    if deform == "sinsin":
        dg, do = se.deform_sinsin(0.1), o.deform_sinsin(0.1)
    elif deform == "affine":
        dg, do = se.deform_affine01, o.deform_affine01
    else:
        dg, do = se.fixU, o.fixU
    m = se.Mesh(n, n, E, Ey, per, dg)
    mo = o.Mesh(n, n, E, Ey, per, do, dense_gs=False)
    return m, mo


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("n,E,Ey,per,deform", [
    (9, 4, 4, (False, False), "none"),
    (8, 3, 2, (True, False), "sinsin"),
    (5, 2, 3, (True, True), "affine"),
    (16, 2, 2, (False, True), "sinsin"),
])
def test_mesh_setup_matches_oracle(se, n, E, Ey, per, deform):
    _log(f"mesh n={n} E={E}x{Ey} per={per} {deform}")
    m, mo = _pair(se, n, E, per, deform, Ey)
    # fields that vanish identically (ry, sx, G12 on affine maps) are rounding
    # noise on both sides, so each is scaled by the largest field of its group
    groups = [["x", "y"], ["Jac"], ["rx", "ry", "sx", "sy"], ["B"], ["G11", "G12", "G22"]]
    for grp in groups:
        scale = max(np.abs(getattr(mo, nm)).max() for nm in grp)
        for name in grp:
            got = se.to_host(m.field(name))
            ref = getattr(mo, name)
            # 1e-12: sums of D entries up to N(N+1)/4 = 60 at N = 15 (the
            # reference's own GEMM summation order is not pinned)
            assert np.abs(got - ref).max() <= 1e-12 * scale, name
    # structural data are bit-exact
    np.testing.assert_array_equal(se.to_host(m.mult), mo.mult)
    np.testing.assert_array_equal(se.to_host(m.l2g).astype(np.int64), mo.l2g)
    assert m.nglobal == mo.nglobal
    for bc in ("DDDD", "DNDN", "NNDD", "NNNN"):
        np.testing.assert_array_equal(se.to_host(se.generateMask(bc, m)), o.generateMask(bc, mo))


@pytest.mark.parametrize("nr,ns,E,Ey,per", [
    (4, 6, 3, 2, (False, True)), (3, 5, 1, 4, (True, True)), (5, 3, 2, 7, (True, True)),
    (2, 2, 1, 1, (True, True)), (6, 9, 4, 3, (True, False)),
])
def test_gather_scatter_map_rectangular_orders(se, nr, ns, E, Ey, per):
    """The mesh's QQ' map (built on the device from the structured numbering)
    for nr != ns and single-element periodic directions: value-exact against
    the dense QQtx * u * QQty' and equal to the map built from l2g by the
    host's counting sort (sem_gsmap_create)."""
    _log(f"gs nr={nr} ns={ns} E={E}x{Ey} per={per}")
    m = se.Mesh(nr, ns, E, Ey, per)
    mo = o.Mesh(nr, ns, E, Ey, per)
    np.testing.assert_array_equal(se.to_host(m.l2g).astype(np.int64), mo.l2g)
    rng = np.random.default_rng(12)
    u = rng.standard_normal(mo.x.shape)
    ref = o.gatherScatter(u, mo.QQtx, mo.QQty)
    got = se.to_host(se.gatherScatter(se.to_device(u), m))
    np.testing.assert_array_equal(got, ref)
    np.testing.assert_array_equal(got, np.asarray(se.gatherScatter(u, mo.l2g)))


@pytest.mark.parametrize("n,E,Ey,per", [
    (8, 3, 2, (False, False)), (5, 4, 3, (True, False)), (4, 3, 3, (True, True)),
    (3, 1, 1, (True, True)), (9, 4, 4, (False, True)),
])
def test_gather_scatter_map_is_value_exact(se, n, E, Ey, per):
    _log(f"gs n={n} E={E}x{Ey} per={per}")
    m = se.Mesh(n, n, E, Ey, per)
    mo = o.Mesh(n, n, E, Ey, per)   # dense QQtx/QQty as the reference
    rng = np.random.default_rng(11)
    u = rng.standard_normal(mo.x.shape)
    ref = o.gatherScatter(u, mo.QQtx, mo.QQty)
    got = se.to_host(se.gatherScatter(se.to_device(u), m))
    np.testing.assert_array_equal(got, ref)
    # reference's (u, QQtx, QQty) form and the (u, l2g) form
    np.testing.assert_array_equal(se.gatherScatter(u, mo.QQtx, mo.QQty), ref)
    np.testing.assert_array_equal(se.gatherScatter(u, mo.l2g), ref)
    # in place
    ud = se.to_device(u)
    se.gatherScatter(ud, m, out=ud)
    np.testing.assert_array_equal(se.to_host(ud), ref)


def test_element_operators_match_oracle(se):
    _log("element ops")
    m, mo = _pair(se, 8, 3, (False, False), "sinsin", 2)
    rng = np.random.default_rng(12)
    u = rng.standard_normal(mo.x.shape)
    nu = 1.0 + rng.random(mo.x.shape)
    ud = se.to_device(u)
    assert _rel(se.to_host(se.lapl(ud, m)), o.lapl(u, mo)) < 1e-12
    # B .* u is exact given B; B itself agrees with the oracle to rounding
    np.testing.assert_array_equal(se.to_host(se.mass(ud, m)), se.to_host(m.B) * u)
    assert _rel(se.to_host(se.mass(ud, m)), o.mass(u, mo)) < 1e-13
    assert _rel(se.to_host(se.hlmz(ud, 0.7, 1.3, m)), o.hlmz(u, 0.7, 1.3, mo)) < 1e-12
    assert _rel(se.to_host(se.lapl(ud, se.to_device(nu), m)), o.lapl(u, mo, nu)) < 1e-12
    assert _rel(se.to_host(se.hlmz(ud, se.to_device(nu), se.to_device(nu), m)),
                o.hlmz(u, nu, nu, mo)) < 1e-12
    # numpy in, numpy out (host arrays round-trip through the device)
    assert _rel(se.lapl(u, m), o.lapl(u, mo)) < 1e-12


def test_abu_matches_oracle(se):
    _log("ABu")
    rng = np.random.default_rng(13)
    u = rng.standard_normal((12, 15))
    Br = rng.standard_normal((5, 4))
    As = rng.standard_normal((2, 3))
    assert _rel(se.ABu(As, Br, u), o.ABu(As, Br, u)) < 1e-14
    assert _rel(se.ABu(None, Br, u), o.ABu(None, Br, u)) < 1e-14
    assert _rel(se.ABu(As, None, u), o.ABu(As, None, u)) < 1e-14
    v = rng.standard_normal((15, 10))
    assert _rel(se.ABu(As, Br, v, trans=True), o.ABu(As.T, Br.T, v)) < 1e-14


APPLY_CASES = [
    # n, Ex, Ey, per, bc, rows per segment (> 0: swept one after another)
    (8, 70, 5, (False, False), "DDDD", 1),
    (8, 70, 5, (False, False), "DDDD", 2),
    (8, 70, 5, (False, False), "DNDN", 3),
    (8, 70, 7, (True, False), "NNDD", 5),
    (8, 33, 6, (False, True), "DDNN", 4),
    (8, 32, 6, (True, True), "NNNN", 6),
    (9, 4, 4, (False, False), "DDDD", 4),
    (16, 20, 3, (False, False), "DDDD", 2),
    (16, 17, 4, (True, True), "NNNN", 3),
    (2, 300, 3, (False, False), "DDDD", 2),
    (3, 90, 4, (True, False), "NNDD", 2),
    (10, 30, 5, (False, False), "DDDD", 2),
    (5, 1, 3, (True, True), "NNNN", 2),
    (7, 2, 1, (True, True), "NNNN", 1),
    # 0: the default plan (element rows side by side where the strip allows);
    # -2 / -4: two / four rows side by side, forced
    (8, 70, 5, (False, False), "DDDD", 0),
    (8, 70, 5, (True, True), "NNNN", -2),
    (8, 33, 6, (False, True), "DDNN", -4),
    (8, 17, 3, (True, False), "NNDD", -4),
    (8, 16, 2, (True, True), "NNNN", -4),
    (4, 70, 3, (True, True), "NNNN", -4),
    (2, 300, 3, (False, False), "DDDD", 0),
    (7, 40, 9, (False, False), "DNDN", 0),
    (6, 22, 7, (True, False), "DDDD", 0),
    (9, 29, 5, (False, True), "NNDD", -4),
    (14, 20, 3, (True, True), "NNNN", 0),
]


@pytest.mark.parametrize("n,Ex,Ey,per,bc,S", APPLY_CASES)
def test_fused_apply_matches_oracle_and_assembly_order(se, rows_env, n, Ex, Ey, per, bc, S):
    _log(f"apply n={n} E={Ex}x{Ey} per={per} bc={bc} S={S}")
    os.environ["SEM_AX_ROWS"] = str(S)
    m, mo = _pair(se, n, Ex, per, "sinsin", Ey)
    M = se.generateMask(bc, m)
    Mo = o.generateMask(bc, mo)
    rng = np.random.default_rng(100 + n)
    u = rng.standard_normal(mo.x.shape)
    nu, k = 0.9, 1.1
    op = se.SEMOperator(nu, k, m, M)
    y = se.to_host(op(se.to_device(u)))
    ref = o.opLHS(u, nu, k, mo, Mo)
    assert _rel(y, ref) < 1e-12
    # assembly order: identical bits to (GPU element values) -> oracle QQ' -> mask
    hl = se.to_host(se.hlmz(se.to_device(u), nu, k, m))
    np.testing.assert_array_equal(y, Mo * o.gatherScatter_structural(hl, mo))
    # masked nodes are exact zeros
    assert np.all(y[Mo == 0] == 0)


def test_fused_apply_is_deterministic(se):
    _log("determinism")
    m = se.Mesh(8, 8, 64, 64, deform=se.deform_sinsin(0.1))
    M = se.generateMask("DDDD", m)
    op = se.SEMOperator(1.0, 1.0, m, M)
    import torch
    g = torch.Generator(device="cuda").manual_seed(5)
    u = se.empty_field(*m.shape)
    u.copy_(torch.randn(m.shape, generator=g, device="cuda", dtype=torch.float64))
    a = op(u)
    for _ in range(3):
        assert torch.equal(op(u), a)


def test_fused_apply_256x256_N7_full_size(se):
    """256x256 elements, N=7 (4.19 M local nodes), deformed, Helmholtz: the
    headline operator against the oracle at full size."""
    _log("apply 256x256 N=7")
    m, mo = _pair(se, 8, 256, (False, False), "sinsin")
    M = se.generateMask("DDDD", m)
    Mo = o.generateMask("DDDD", mo)
    rng = np.random.default_rng(2)
    u = rng.standard_normal(mo.x.shape)
    op = se.SEMOperator(1.0, 1.0, m, M)
    y = se.to_host(op(se.to_device(u)))
    assert _rel(y, o.opLHS(u, 1.0, 1.0, mo, Mo)) < 1e-12


def test_jacobi_matches_oracle(se):
    _log("jacobi")
    m, mo = _pair(se, 8, 5, (False, False), "sinsin", 4)
    M = se.generateMask("DDDD", m)
    op = se.SEMOperator(0.8, 1.7, m, M)
    d = se.to_host(op.jacobi())
    ref = o.jacobi_diag(mo, 0.8, 1.7, o.generateMask("DDDD", mo))
    assert _rel(d, ref) < 1e-13


def _poisson_4x4_n9(se):
    m, mo = _pair(se, 9, 4, (False, False), "none")
    M = se.generateMask("DDDD", m)
    Mo = o.generateMask("DDDD", mo)
    f = 2 * np.pi ** 2 * np.sin(np.pi * mo.x) * np.sin(np.pi * mo.y)
    b = mo.gs(Mo * o.mass(f, mo))
    return m, mo, M, Mo, b


def test_pcg_poisson_4x4_N8_matches_oracle_and_exact_solution(se):
    _log("pcg Poisson 4x4 N=8")
    m, mo, M, Mo, b = _poisson_4x4_n9(se)
    tol = 1e-10 * np.abs(b).max()
    info_o = o.PCGInfo()
    uo = o.pcg(b, lambda v: o.opLHS(v, 1.0, 0.0, mo, Mo), mult=mo.mult, tol=tol, info=info_o)
    op = se.SEMOperator(1.0, 0.0, m, M)
    info = se.PCGInfo()
    u = se.to_host(se.pcg(se.to_device(b), op, mult=m.mult, tol=tol, info=info))
    assert info.converged and abs(info.iters - info_o.iters) <= 2
    assert _rel(u, uo) < 1e-9
    assert np.abs(u - np.sin(np.pi * mo.x) * np.sin(np.pi * mo.y)).max() < 2e-10
    # generic closure path of the mirror: same recursion, same answer
    info_g = se.PCGInfo()
    ug = se.to_host(se.pcg(se.to_device(b), lambda v: se.mask(se.gatherScatter(
        se.hlmz(v, 1.0, 0.0, m), m), M), mult=m.mult, tol=tol, info=info_g))
    assert abs(info_g.iters - info_o.iters) <= 2 and _rel(ug, uo) < 1e-9


@pytest.mark.parametrize("per,bc", [((False, False), "DDDD"), ((True, False), "NNDD")])
def test_pcg_jacobi_helmholtz_matches_oracle(se, per, bc):
    _log(f"pcg jacobi {per} {bc}")
    m, mo = _pair(se, 8, 12, per, "sinsin", 10)
    M = se.generateMask(bc, m)
    Mo = o.generateMask(bc, mo)
    rng = np.random.default_rng(3)
    f = rng.standard_normal(mo.x.shape)
    b = mo.gs(Mo * o.mass(f, mo))
    dinv = o.jacobi_diag(mo, 1.0, 1.0, Mo)
    op = se.SEMOperator(1.0, 1.0, m, M)
    d = op.jacobi()
    for rel, band in ((1e-10, 0.01), (1e-12, None)):
        tol = rel * np.abs(b).max()
        info_o = o.PCGInfo()
        uo = o.pcg(b, lambda v: o.opLHS(v, 1.0, 1.0, mo, Mo), opM=lambda r: dinv * r,
                   mult=mo.mult, tol=tol, info=info_o)
        info = se.PCGInfo()
        u = se.to_host(se.pcg(se.to_device(b), op, opM=se.Diagonal(d), mult=m.mult, tol=tol,
                              info=info))
        assert info.converged
        if band is not None:
            # iteration counts compared away from the rounding floor: the fused
            # solve sums its inner products in a different order
            assert abs(info.iters - info_o.iters) <= max(2, int(band * info_o.iters))
        assert _rel(u, uo) < 10 * rel


def test_pcg_relative_tolerance_on_device_and_dot_acc(se):
    """rtol: the fused solve's stop test norm(r, Inf) <= rtol norm(b, Inf)
    (norm(b, Inf) formed on the device) runs the same iterations as the
    absolute tol = rtol norm(b, Inf); the generic path and the host-driven
    path agree; sem_dot_acc accumulates sum(a .* b .* w) on the device."""
    import torch
    _log("pcg rtol / dot_acc")
    m, mo = _pair(se, 8, 12, (False, False), "sinsin", 10)
    M = se.generateMask("DDDD", m)
    Mo = o.generateMask("DDDD", mo)
    rng = np.random.default_rng(8)
    b = mo.gs(Mo * o.mass(rng.standard_normal(mo.x.shape), mo))
    op = se.SEMOperator(1.0, 1.0, m, M)
    d = se.Diagonal(op.jacobi())
    tol = 1e-9 * np.abs(b).max()
    ia, ir, ig = se.PCGInfo(), se.PCGInfo(), se.PCGInfo()
    xa = se.to_host(se.pcg(se.to_device(b), op, opM=d, mult=m.mult, tol=tol, info=ia))
    xr = se.to_host(se.pcg(se.to_device(b), op, opM=d, mult=m.mult, rtol=1e-9, info=ir))
    xg = se.to_host(se.pcg(se.to_device(b), lambda v: op(v), opM=d, mult=m.mult, rtol=1e-9,
                           info=ig))
    assert ia.iters == ir.iters and ir.converged and ir.res <= tol
    np.testing.assert_array_equal(xa, xr)
    assert abs(ig.iters - ir.iters) <= 2 and _rel(xg, xr) < 1e-8
    zero = se.PCGInfo()
    se.pcg(se.to_device(0 * b), op, opM=d, mult=m.mult, rtol=1e-9, info=zero)
    assert zero.iters == 0 and zero.converged
    acc = torch.zeros(1, dtype=torch.float64, device="cuda")
    u, v = rng.standard_normal(mo.x.shape), rng.standard_normal(mo.x.shape)
    se.dot_acc(acc, se.to_device(u), se.to_device(v), m.B, 0.5)
    se.dot_acc(acc, se.to_device(u), se.to_device(u), None, 2.0)
    ref = 0.5 * np.sum(u * v * mo.B) + 2.0 * np.sum(u * u)
    assert abs(float(acc) - ref) <= 1e-12 * np.sum(np.abs(u * v * mo.B) + 2 * u * u)


def test_pcg_maxiter_and_zero_rhs(se):
    _log("pcg maxiter / zero rhs")
    m, mo, M, Mo, b = _poisson_4x4_n9(se)
    op = se.SEMOperator(1.0, 0.0, m, M)
    info = se.PCGInfo()
    with pytest.warns(RuntimeWarning, match="warning: res"):
        se.pcg(se.to_device(b), op, mult=m.mult, tol=1e-30, maxiter=3, info=info)
    assert info.iters == 3 and not info.converged
    info2 = se.PCGInfo()
    x = se.pcg(se.to_device(0 * b), op, mult=m.mult, tol=1e-12, info=info2)
    assert info2.iters == 0 and float(x.abs().max()) == 0.0


def test_f32_operator_and_mixed_solve(se):
    _log("f32 operator + mixed")
    m, mo = _pair(se, 8, 16, (False, False), "sinsin")
    M = se.generateMask("DDDD", m)
    Mo = o.generateMask("DDDD", mo)
    rng = np.random.default_rng(4)
    u = rng.standard_normal(mo.x.shape)
    op32 = se.SEMOperator(1.0, 1.0, m, M, precision="f32")
    y32 = se.to_host(op32(se.to_device(u)).double())
    assert _rel(y32, o.opLHS(u, 1.0, 1.0, mo, Mo)) < 1e-5
    f = rng.standard_normal(mo.x.shape)
    b = mo.gs(Mo * o.mass(f, mo))
    op64 = se.SEMOperator(1.0, 1.0, m, M)
    d = op64.jacobi()
    tol = 1e-10 * np.abs(b).max()
    info = se.PCGInfo()
    x = se.to_host(se.pcg_mixed(se.to_device(b), op64, op32, d=d, tol=tol, inner_tol=1e-5,
                                info=info))
    assert info.converged
    info64 = se.PCGInfo()
    x64 = se.to_host(se.pcg(se.to_device(b), op64, opM=se.Diagonal(d), mult=m.mult, tol=tol,
                            info=info64))
    r = b - o.opLHS(x, 1.0, 1.0, mo, Mo)
    assert np.abs(r).max() <= 1.01 * tol
    assert _rel(x, x64) < 1e-8


def test_grad_gradT_adjoint_dnu_match_oracle(se):
    _log("grad/gradT/dnu")
    m, mo = _pair(se, 7, 3, (False, False), "sinsin", 2)
    rng = np.random.default_rng(9)
    u = rng.standard_normal(mo.x.shape)
    lam = rng.standard_normal(mo.x.shape)
    gx, gy = se.grad(u, m)
    ox, oy = o.grad(u, mo)
    assert _rel(gx, ox) < 1e-12 and _rel(gy, oy) < 1e-12
    assert _rel(se.gradT(u, lam, m), o.gradT(u, lam, mo)) < 1e-12
    g = se.to_host(se.adjoint_dnu(se.to_device(u), se.to_device(lam), m))
    lx, ly = o.grad(lam, mo)
    assert _rel(g, mo.B * (lx * ox + ly * oy)) < 1e-12


def test_advect_and_transpose_match_oracle(se):
    _log("advect")
    mV, moV = _pair(se, 10, 3, (False, False), "sinsin", 2)
    mD, moD = _pair(se, 15, 3, (False, False), "sinsin", 2)
    rng = np.random.default_rng(10)
    T = rng.standard_normal(moV.x.shape)
    vx, vy = -moV.y, moV.x
    ref = o.advect(T, vx, vy, moV, moD)
    assert _rel(se.advect(T, vx, vy, mV, mD), ref) < 1e-12
    w = rng.standard_normal(moV.x.shape)
    At_w = se.advect(w, vx, vy, mV, mD, transpose=True)
    # <w, A T> = <A' w, T>
    lhs = np.sum(w * ref)
    rhs = np.sum(At_w * T)
    assert abs(lhs - rhs) < 1e-11 * abs(lhs)


PCG_SEAM_CASES = [
    # n, Ex, Ey, per, bc, rows per segment, preconditioner
    (8, 70, 6, (False, False), "DDDD", 3, "jacobi"),
    (8, 70, 6, (True, False), "NNDD", 2, "jacobi"),
    (8, 40, 7, (False, True), "DDNN", 4, "identity"),
    (6, 50, 4, (True, True), "NNNN", 1, "jacobi"),
    (16, 20, 5, (False, False), "DDDD", 2, "jacobi"),
    # -2: two element rows side by side per workgroup
    (8, 70, 6, (True, False), "NNDD", -2, "jacobi"),
    (8, 40, 7, (False, True), "DDNN", -2, "identity"),
    (6, 50, 5, (True, True), "NNNN", -2, "jacobi"),
]


@pytest.mark.parametrize("n,Ex,Ey,per,bc,S,pre", PCG_SEAM_CASES)
def test_pcg_seam_configs_match_oracle(se, rows_env, n, Ex, Ey, per, bc, S, pre):
    """The fused PCG leaves workgroup-seam copies of A p unsummed and the
    update kernel finishes them; every seam type is exercised here."""
    _log(f"pcg seams n={n} E={Ex}x{Ey} per={per} bc={bc} S={S} {pre}")
    os.environ["SEM_AX_ROWS"] = str(S)
    m, mo = _pair(se, n, Ex, per, "sinsin", Ey)
    M = se.generateMask(bc, m)
    Mo = o.generateMask(bc, mo)
    rng = np.random.default_rng(20 + n)
    f = rng.standard_normal(mo.x.shape)
    b = mo.gs(Mo * o.mass(f, mo))
    tol = 1e-10 * np.abs(b).max()
    k = 1.0   # Helmholtz keeps the all-Neumann case SPD
    pre_o = None
    op = se.SEMOperator(1.0, k, m, M)
    opM = None
    if pre == "jacobi":
        dinv = o.jacobi_diag(mo, 1.0, k, Mo)
        pre_o = lambda r: dinv * r
        opM = se.Diagonal(op.jacobi())
    info_o = o.PCGInfo()
    uo = o.pcg(b, lambda v: o.opLHS(v, 1.0, k, mo, Mo), opM=pre_o, mult=mo.mult, tol=tol,
               info=info_o)
    info = se.PCGInfo()
    u = se.to_host(se.pcg(se.to_device(b), op, opM=opM, mult=m.mult, tol=tol, info=info))
    assert info.converged
    assert abs(info.iters - info_o.iters) <= max(2, info_o.iters // 100)
    assert _rel(u, uo) < 1e-9
    # true residual of the returned solution
    r = b - o.opLHS(u, 1.0, k, mo, Mo)
    assert np.abs(r).max() < 10 * tol


# relative error of one apply against the FP64 oracle, per contraction form:
# FP16 / BF16 inputs round u, D and the fluxes to 11 / 8 significant bits;
# the split forms carry about 22 / 16 bits.
MFMA_TOL = {"f32": 2e-6, "f16": 5e-3, "bf16": 4e-2, "f16x3": 2e-5, "bf16x3": 2e-4}


@pytest.mark.parametrize("prec", ["f16", "bf16", "f16x3", "bf16x3"])
def test_mfma_operator_matches_oracle(se, rows_env, prec):
    """N = 15 (16x16 blocks): MFMA contractions, FP32 fields, fused GS/mask."""
    _log(f"mfma apply {prec}")
    os.environ["SEM_AX_ROWS"] = "2"
    m, mo = _pair(se, 16, 20, (False, False), "sinsin", 5)
    M = se.generateMask("DDDD", m)
    Mo = o.generateMask("DDDD", mo)
    rng = np.random.default_rng(30)
    u = rng.standard_normal(mo.x.shape)
    ref = o.opLHS(u, 1.0, 1.0, mo, Mo)
    op = se.SEMOperator(1.0, 1.0, m, M, precision=prec)
    y = se.to_host(op(se.to_device(u)).double())
    assert _rel(y, ref) < MFMA_TOL[prec]
    assert np.all(y[Mo == 0] == 0)
    # assembled output is continuous: every copy of a node holds the same value
    g = mo.l2g.reshape(-1)
    v = y.reshape(-1)
    first = np.full(mo.nglobal, np.nan)
    first[g[::-1]] = v[::-1]
    np.testing.assert_array_equal(first[g], v)
    # element-local form (no assembly)
    op_e = se.SEMOperator(1.0, 1.0, m, None, precision=prec, assemble=False)
    ye = se.to_host(op_e(se.to_device(u)).double())
    assert _rel(ye, o.hlmz(u, 1.0, 1.0, mo)) < MFMA_TOL[prec]


# relative error of one apply on the smoothest mode sin(pi (x+1)/2) sin(pi
# (y+1)/2) at 32x32 elements, N = 15: the single-pass FP16 / BF16 forms deflate
# each contracted vector by its end-node linear interpolant, which keeps this
# error near the FP32 operator's instead of 3-4 orders above it (without the
# deflation the BF16 form is off by a factor of about 100 on this mode).
SMOOTH_TOL = {"f32": 1e-2, "f16": 6e-2, "bf16": 0.4}


@pytest.mark.parametrize("prec", ["f32", "f16", "bf16"])
def test_low_precision_operator_on_smooth_mode(se, prec):
    _log(f"smooth-mode apply {prec}")
    m, mo = _pair(se, 16, 32, (False, False), "sinsin")
    M = se.generateMask("DDDD", m)
    Mo = o.generateMask("DDDD", mo)
    u = Mo * np.sin(np.pi * (mo.x + 1) / 2) * np.sin(np.pi * (mo.y + 1) / 2)
    ref = o.opLHS(u, 1.0, 1.0, mo, Mo)
    op = se.SEMOperator(1.0, 1.0, m, M, precision=prec)
    y = se.to_host(op(se.to_device(u, dtype=op.dtype)).double())
    assert _rel(y, ref) < SMOOTH_TOL[prec]


@pytest.mark.parametrize("prec", ["f32", "f16", "bf16", "f16x3", "bf16x3"])
def test_mixed_solve_full_size_reaches_fp64_tolerance(se, prec):
    """128x128 elements, N = 15, deformed, Helmholtz, f ~ N(0,1) per global
    node (seed 3), inner tolerance 1e-5: every variant reaches
    norm(b - A64 x, Inf) <= 1e-10 norm(b, Inf) and agrees with the FP64 solve."""
    import torch
    from semhip import workloads as W
    _log(f"mixed full size {prec}")
    m = se.Mesh(16, 16, 128, 128, (False, False), se.deform_sinsin(0.1))
    M = se.generateMask("DDDD", m)
    op64 = se.SEMOperator(1.0, 1.0, m, M)
    b = W.helmholtz_rhs(se, m, M, seed=3)
    tol = 1e-10 * float(b.abs().max())
    d = op64.jacobi()
    op_lo = se.SEMOperator(1.0, 1.0, m, M, precision=prec)
    info = se.PCGInfo()
    x = se.pcg_mixed(b, op64, op_lo, d=d, tol=tol, inner_tol=1e-5, max_outer=200, info=info)
    assert info.converged, (prec, info.outer, info.iters, info.res)
    assert float((b - op64(x)).abs().max()) <= 1.01 * tol
    x64 = se.pcg(b, op64, opM=se.MG(op64), mult=m.mult, tol=tol)
    assert float((x - x64).abs().max()) <= 1e-8 * float(x64.abs().max())
    del x, x64
    torch.cuda.synchronize()


@pytest.mark.parametrize("prec", ["f32", "f16", "bf16", "f16x3", "bf16x3"])
def test_mixed_solve_reaches_fp64_tolerance(se, prec):
    """FP64 defect correction around each low-precision inner solve reaches
    the FP64 tolerance; the error against the FP64 solve is reported."""
    _log(f"mixed {prec}")
    m, mo = _pair(se, 16, 12, (False, False), "sinsin")
    M = se.generateMask("DDDD", m)
    Mo = o.generateMask("DDDD", mo)
    rng = np.random.default_rng(31)
    f = rng.standard_normal(mo.x.shape)
    b = mo.gs(Mo * o.mass(f, mo))
    tol = 1e-10 * np.abs(b).max()
    op64 = se.SEMOperator(1.0, 1.0, m, M)
    d = op64.jacobi()
    op_lo = se.SEMOperator(1.0, 1.0, m, M, precision=prec)
    info = se.PCGInfo()
    x = se.to_host(se.pcg_mixed(se.to_device(b), op64, op_lo, d=d, tol=tol, inner_tol=1e-5,
                                max_outer=60, info=info))
    assert info.converged, (prec, info.outer, info.res)
    r = b - o.opLHS(x, 1.0, 1.0, mo, Mo)
    assert np.abs(r).max() <= 1.01 * tol
    x64 = se.to_host(se.pcg(se.to_device(b), op64, opM=se.Diagonal(d), mult=m.mult, tol=tol))
    assert _rel(x, x64) < 1e-8


def test_mixed_solve_reports_stall_distinctly(se):
    """A tolerance below the FP64 rounding floor cannot be met: the solver
    stops after four replacements without progress with SEM_WARN_STALLED
    (a StallWarning, info.stalled), not the maxiter warning, and returns its
    best iterate."""
    from semhip import _lib as SL
    _log("mixed stall")
    m, mo = _pair(se, 16, 6, (False, False), "sinsin")
    M = se.generateMask("DDDD", m)
    Mo = o.generateMask("DDDD", mo)
    rng = np.random.default_rng(5)
    b = mo.gs(Mo * o.mass(rng.standard_normal(mo.x.shape), mo))
    op64 = se.SEMOperator(1.0, 1.0, m, M)
    op32 = se.SEMOperator(1.0, 1.0, m, M, precision="f32")
    info = se.PCGInfo()
    with pytest.warns(SL.StallWarning):
        x = se.to_host(se.pcg_mixed(se.to_device(b), op64, op32, d=op64.jacobi(), tol=0.0,
                                    inner_tol=1e-5, max_outer=200, info=info))
    assert info.stalled and not info.converged and info.outer < 200
    assert np.abs(b - o.opLHS(x, 1.0, 1.0, mo, Mo)).max() <= 1e-12 * np.abs(b).max()


@pytest.mark.parametrize("Ex", [896, 1040])
def test_pcg_long_rows_match_oracle(se, Ex):
    """Rows of 16 * 896 = 14336 nodes (the longest whose fact table the
    update kernel copies into LDS) and 16 * 1040 = 16640 (table read from
    global memory): 40 Jacobi-PCG iterations agree with the oracle's."""
    _log(f"pcg long rows Ex={Ex}")
    m, mo = _pair(se, 16, Ex, (False, False), "sinsin", Ey=2)
    M = se.generateMask("DDDD", m)
    Mo = o.generateMask("DDDD", mo)
    rng = np.random.default_rng(Ex)
    b = mo.gs(Mo * o.mass(rng.standard_normal(mo.x.shape), mo))
    op = se.SEMOperator(1.0, 1.0, m, M)
    dinv = o.jacobi_diag(mo, 1.0, 1.0, Mo)
    info_o = o.PCGInfo()
    xo = o.pcg(b, lambda v: o.opLHS(v, 1.0, 1.0, mo, Mo), opM=lambda r: dinv * r,
               mult=mo.mult, tol=1e-300, maxiter=40, info=info_o)
    info = se.PCGInfo()
    with pytest.warns(RuntimeWarning):
        x = se.to_host(se.pcg(se.to_device(b), op, opM=se.Diagonal(op.jacobi()), mult=m.mult,
                              tol=1e-300, maxiter=40, info=info))
    assert info.iters == info_o.iters == 40
    assert _rel(x, xo) < 1e-10


@pytest.mark.parametrize("fma16", ["0", "1"])
def test_p16_fp64_and_fp32_contraction_paths(se, rows_env, fma16):
    """N = 15: the FP64/FP32 operators default to the exact FP64/FP32 MFMA
    contractions; SEM_AX_FMA16=1 selects the per-lane FMA path. Both match the
    oracle to FP64 / FP32 rounding."""
    _log(f"p16 contraction path fma16={fma16}")
    old = os.environ.get("SEM_AX_FMA16")
    os.environ["SEM_AX_FMA16"] = fma16
    try:
        m, mo = _pair(se, 16, 17, (True, False), "sinsin", 4)
        M = se.generateMask("NNDD", m)
        Mo = o.generateMask("NNDD", mo)
        rng = np.random.default_rng(33)
        u = rng.standard_normal(mo.x.shape)
        ref = o.opLHS(u, 0.9, 1.1, mo, Mo)
        y64 = se.to_host(se.SEMOperator(0.9, 1.1, m, M)(se.to_device(u)))
        assert _rel(y64, ref) < 1e-12
        hl = se.to_host(se.hlmz(se.to_device(u), 0.9, 1.1, m))
        np.testing.assert_array_equal(y64, Mo * o.gatherScatter_structural(hl, mo))
        y32 = se.to_host(se.SEMOperator(0.9, 1.1, m, M, precision="f32")(se.to_device(u)).double())
        assert _rel(y32, ref) < MFMA_TOL["f32"]
    finally:
        if old is None:
            os.environ.pop("SEM_AX_FMA16", None)
        else:
            os.environ["SEM_AX_FMA16"] = old


@pytest.mark.parametrize("dealiased", [False, True])
def test_twelve_argument_lapl_and_eight_argument_mass(se, dealiased):
    """lapl(u, M, Jr, Js, QQtx, QQty, Dr, Ds, G11, G12, G22, mult) (lapl.jl:54-68,
    83-103) and mass(u, M, B, Jr, Js, QQtx, QQty, mult) (mass.jl:32-50) through
    the library against the oracle, with the dense 0/1 QQ factors of the
    reference; the two-mesh forms forward to the first mesh (lapl.jl:47-52)."""
    _log(f"12-arg lapl / 8-arg mass dealiased={dealiased}")
    mo = o.Mesh(5, 5, 3, 2, deform=o.deform_sinsin(0.1), dense_gs=True)
    Mo = o.generateMask("DNDD", mo)
    rng = np.random.default_rng(12)
    u = rng.standard_normal(mo.x.shape)
    if dealiased:
        md = o.Mesh(8, 8, 3, 2, deform=o.deform_sinsin(0.1), dense_gs=False)
        Jr, Js = o.interpMat(md.zr, mo.zr), o.interpMat(md.zs, mo.zs)
        G = [rng.standard_normal(md.x.shape) for _ in range(3)]
        Bd = rng.standard_normal(md.x.shape)
    else:
        Jr = Js = np.zeros((0, 0))
        G = [mo.G11, mo.G12, mo.G22]
        Bd = mo.B
    ref = o.lapl12(u, Mo, Jr, Js, mo.QQtx, mo.QQty, mo.Dr, mo.Ds, *G, mo.mult)
    got = se.lapl(u, Mo, Jr, Js, mo.QQtx, mo.QQty, mo.Dr, mo.Ds, *G, mo.mult)
    assert _rel(got, ref) < 1e-12
    ref = o.mass8(u, Mo, Bd, Jr, Js, mo.QQtx, mo.QQty, mo.mult)
    got = se.mass(u, Mo, Bd, Jr, Js, mo.QQtx, mo.QQty, mo.mult)
    assert _rel(got, ref) < 1e-12
    m = se.Mesh(5, 5, 3, 2, (False, False), se.deform_sinsin(0.1))
    np.testing.assert_array_equal(se.lapl(u, m, m), se.lapl(u, m))
    np.testing.assert_array_equal(se.mass(u, m, m), se.mass(u, m))
    np.testing.assert_array_equal(se.hlmz(u, 1.2, 0.3, m, m), se.hlmz(u, 1.2, 0.3, m))


def test_fused_pcg_and_mixed_solves_are_deterministic(se):
    """Fixed-order reductions everywhere: repeated solves (FP64 fused PCG and
    the FP32 mixed solver) give bit-identical solutions and iteration counts."""
    import torch
    _log("solve determinism")
    m = se.Mesh(8, 8, 24, 24, (False, False), se.deform_sinsin(0.1))
    M = se.generateMask("DDDD", m)
    op = se.SEMOperator(1.0, 1.0, m, M)
    b = se.gatherScatter(se.mask(se.mass(se.workloads.global_normal_field(m, 4), m), M), m)
    d = op.jacobi()
    tol = 1e-10 * float(b.abs().max())
    runs = []
    for _ in range(3):
        info = se.PCGInfo()
        x = se.pcg(b, op, opM=se.Diagonal(d), mult=m.mult, tol=tol, info=info)
        runs.append((x.clone(), info.iters))
    assert all(torch.equal(r[0], runs[0][0]) and r[1] == runs[0][1] for r in runs)
    op32 = se.SEMOperator(1.0, 1.0, m, M, precision="f32")
    xs = [se.pcg_mixed(b, op, op32, d=d, tol=tol, inner_tol=1e-3) for _ in range(2)]
    assert torch.equal(xs[0], xs[1])


@pytest.mark.parametrize("n,Ex", [(16, 128), (8, 256)])
def test_fp32_solves_with_paired_row_update(se, n, Ex):
    """Meshes whose local row count is a multiple of 2048 take the FP32 update
    kernel that handles rows in pairs: an FP32 fused PCG and the mixed solver
    on such a mesh against the FP64 solve."""
    _log(f"paired-row FP32 update n={n} Ex={Ex}")
    m = se.Mesh(n, n, Ex, 4, (False, False), se.deform_sinsin(0.1))
    assert m.shape[0] % 2048 == 0
    M = se.generateMask("DDDD", m)
    op64 = se.SEMOperator(1.0, 1.0, m, M)
    op32 = se.SEMOperator(1.0, 1.0, m, M, precision="f32")
    b = se.gatherScatter(se.mask(se.mass(se.workloads.global_normal_field(m, 6), m), M), m)
    d = op64.jacobi()
    tol = 1e-10 * float(b.abs().max())
    x64 = se.to_host(se.pcg(b, op64, opM=se.Diagonal(d), mult=m.mult, tol=tol))
    info = se.PCGInfo()
    x32 = se.pcg(b.float(), op32, opM=se.Diagonal(d.float()), mult=m.mult,
                 tol=1e-5 * float(b.abs().max()), info=info)
    assert info.converged
    assert _rel(se.to_host(x32.double()), x64) < 1e-4
    info = se.PCGInfo()
    xm = se.to_host(se.pcg_mixed(b, op64, op32, d=d, tol=tol, inner_tol=1e-3, info=info))
    assert info.converged
    assert _rel(xm, x64) < 1e-8
