"""GPU parity of the multigrid preconditioner and of PCG with it (sem_mg_*,
sem_pcg_mg) against the oracle's restatement (oracle/sem_mg.py), called
through the C ABI via the semhip mirror. Tolerances: the hierarchy and the
structural data are exact, the eigenvalue estimates and one V-cycle within
1e-10 relative (same arithmetic, different summation order), converged solves
within 1e-10 of the oracle's direct solve; at the full 256x256, N=7 size the
true residual of the solve is checked."""
import sys

import numpy as np
import pytest

from oracle import sem_oracle as o
from oracle import sem_mg as omg

pytestmark = pytest.mark.gpu


def _log(msg):
    print(f"[case] {msg}", file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def se():
    import semhip
    return semhip


def _rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


def _setup(se, n, E, Ey, per, bc, nu, k, deform=True):
    dg, do = (se.deform_sinsin(0.1), o.deform_sinsin(0.1)) if deform else (se.fixU, o.fixU)
    m = se.Mesh(n, n, E, Ey, per, dg)
    mo = o.Mesh(n, n, E, Ey, per, do, dense_gs=False)
    M = se.generateMask(bc, m)
    op = se.SEMOperator(nu, k, m, M)
    return m, mo, M, op


CASES = [
    (8, 8, 8, (False, False), "DDDD", 1.0, 1.0),
    (9, 4, 4, (False, False), "DDDD", 1.0, 0.0),
    (6, 8, 4, (True, False), "DDDD", 1.0, 0.5),
    (5, 4, 8, (False, False), "DNDN", 0.7, 1.0),
]


@pytest.mark.parametrize("n,E,Ey,per,bc,nu,k", CASES)
def test_hierarchy_and_vcycle_match_oracle(se, n, E, Ey, per, bc, nu, k):
    _log(f"mg hierarchy/vcycle n={n} E={E}x{Ey} per={per} bc={bc}")
    m, mo, M, op = _setup(se, n, E, Ey, per, bc, nu, k)
    P = se.MG(op)
    Po = omg.MG(mo, nu, k, bc)
    lv = P.levels()
    assert [(p, ex, ey) for p, ex, ey, _ in lv] == [
        (L.msh.nr, L.msh.Ex, L.msh.Ey) for L in Po.levels]
    for (_, _, _, lm), L in zip(lv, Po.levels):
        assert abs(lm - L.lmax) <= 1e-10 * L.lmax
    # coarse geometry: the level meshes sample the fine geometry
    rng = np.random.default_rng(11)
    r = Po.levels[0].M * mo.gs(rng.standard_normal(mo.x.shape))
    for i, (_, _, _, lm) in enumerate(lv):
        Po.levels[i].lmax = lm            # the same smoothing intervals on both sides
    z = se.to_host(P(se.to_device(r)))
    zo = Po(r)
    assert _rel(z, zo) < 1e-10
    np.testing.assert_array_equal(z, Po.levels[0].M * z)


@pytest.mark.parametrize("n,E,Ey,per,bc,nu,k", CASES)
def test_pcg_mg_matches_oracle(se, n, E, Ey, per, bc, nu, k):
    _log(f"pcg mg n={n} E={E}x{Ey} per={per} bc={bc}")
    m, mo, M, op = _setup(se, n, E, Ey, per, bc, nu, k)
    Mo = o.generateMask(bc, mo)
    rng = np.random.default_rng(12)
    b = Mo * mo.gs(mo.B * rng.standard_normal(mo.nglobal)[mo.l2g])
    tol = 1e-11 * np.abs(b).max()
    P = se.MG(op)
    info = se.PCGInfo()
    x = se.to_host(se.pcg(se.to_device(b), op, opM=P, mult=m.mult, tol=tol, info=info))
    assert info.converged
    Po = omg.MG(mo, nu, k, bc)
    info_o = o.PCGInfo()
    xo = o.pcg(b, lambda v: o.opLHS(v, nu, k, mo, Mo), opM=Po, mult=mo.mult, tol=tol,
               info=info_o)
    assert abs(info.iters - info_o.iters) <= 1, (info.iters, info_o.iters)
    assert _rel(x, xo) < 1e-9
    r = b - o.opLHS(x, nu, k, mo, Mo)
    assert np.abs(r).max() <= 1.05 * tol
    # far fewer iterations than Jacobi
    info_j = se.PCGInfo()
    se.pcg(se.to_device(b), op, opM=se.Diagonal(op.jacobi()), mult=m.mult, tol=tol, info=info_j)
    assert info_j.iters > 2 * info.iters


def test_pcg_mg_full_size_true_residual(se):
    """256 x 256 elements, N = 7, deformed, Helmholtz (nu = k = 1), seeded
    N(0,1) forcing per global node: the solve to 1e-10 relative meets the
    tolerance on the true FP64 residual and agrees with the Jacobi solve."""
    import torch
    _log("pcg mg 256x256 N=7")
    from semhip import workloads as W
    m = se.Mesh(8, 8, 256, 256, (False, False), se.deform_sinsin(0.1))
    M = se.generateMask("DDDD", m)
    op = se.SEMOperator(1.0, 1.0, m, M)
    b = W.helmholtz_rhs(se, m, M, seed=2)
    tol = 1e-10 * float(b.abs().max())
    P = se.MG(op)
    info = se.PCGInfo()
    x = se.pcg(b, op, opM=P, mult=m.mult, tol=tol, info=info)
    assert info.converged and info.iters < 40, info.iters
    r = b - op(x)
    assert float(r.abs().max()) <= 1.05 * tol
    info_j = se.PCGInfo()
    xj = se.pcg(b, op, opM=se.Diagonal(op.jacobi()), mult=m.mult, tol=tol, info=info_j)
    assert float((x - xj).abs().max()) <= 1e-8 * float(xj.abs().max())
    del x, xj, r
    torch.cuda.synchronize()


@pytest.mark.parametrize("n,E", [(8, 256), (16, 128)])
def test_pcg_mg_fp32_cycle_full_size(se, n, E):
    """The FP32 V-cycle (precision='f32') inside the FP64 PCG, at 256x256
    N=7 and 128x128 N=15: the solve to 1e-10 relative meets the tolerance on
    the true FP64 residual, takes at most two iterations more than with the
    FP64 cycle, and agrees with the FP64-cycle solution to 1e-8; one FP32
    V-cycle agrees with the FP64 one to FP32 accuracy (1e-5 relative)."""
    import torch
    from semhip import workloads as W
    _log(f"pcg mg fp32 cycle {E}x{E} n={n}")
    m = se.Mesh(n, n, E, E, (False, False), se.deform_sinsin(0.1))
    M = se.generateMask("DDDD", m)
    op = se.SEMOperator(1.0, 1.0, m, M)
    b = W.helmholtz_rhs(se, m, M, seed=2)
    tol = 1e-10 * float(b.abs().max())
    P64, P32 = se.MG(op), se.MG(op, precision="f32")
    z64, z32 = P64(b), P32(b)
    assert float((z32 - z64).abs().max()) <= 1e-5 * float(z64.abs().max())
    i64, i32 = se.PCGInfo(), se.PCGInfo()
    x64 = se.pcg(b, op, opM=P64, mult=m.mult, tol=tol, info=i64)
    x32 = se.pcg(b, op, opM=P32, mult=m.mult, tol=tol, info=i32)
    _log(f"  iterations fp64 cycle {i64.iters} ({i64.elapsed_ms:.1f} ms), "
         f"fp32 cycle {i32.iters} ({i32.elapsed_ms:.1f} ms)")
    assert i32.converged and i32.iters <= i64.iters + 2
    assert float((b - op(x32)).abs().max()) <= 1.05 * tol
    assert float((x32 - x64).abs().max()) <= 1e-8 * float(x64.abs().max())
    del x64, x32, P64, P32
    torch.cuda.synchronize()


@pytest.mark.parametrize("n,E,Ey,per,bc,prec", [
    (8, 48, 40, (False, False), "DDDD", "f64"),
    (8, 40, 36, (True, False), "NNDD", "f64"),
    (8, 48, 40, (False, True), "DNDN", "f32"),
    (16, 24, 24, (False, False), "DDDD", "f64"),
])
def test_fused_smoother_bits_equal_unfused(se, monkeypatch, n, E, Ey, per, bc, prec):
    """The smoother's vector updates fused into the level applies (the
    default, on the two-row apply levels p = 8, 4, 2) give exactly the bits
    of the separate vector kernels (SEM_MG_FUSE=0 at creation): one V-cycle,
    and a PCG solve with the same iterations and the same solution."""
    import torch
    from semhip import workloads as W
    _log(f"fused smoother n={n} E={E}x{Ey} per={per} bc={bc} {prec}")
    m = se.Mesh(n, n, E, Ey, per, se.deform_sinsin(0.1))
    M = se.generateMask(bc, m)
    op = se.SEMOperator(1.0, 1.0, m, M)
    b = W.helmholtz_rhs(se, m, M, seed=4)
    tol = 1e-10 * float(b.abs().max())
    Pf = se.MG(op, precision=prec)
    monkeypatch.setenv("SEM_MG_FUSE", "0")
    Pu = se.MG(op, precision=prec)
    monkeypatch.delenv("SEM_MG_FUSE")
    zf, zu = Pf(b), Pu(b)
    assert torch.equal(zf, zu)
    zf2, zu2 = Pf(b), Pu(b)   # the level buffers have traded roles in between
    assert torch.equal(zf2, zf) and torch.equal(zu2, zu)
    i_f, i_u = se.PCGInfo(), se.PCGInfo()
    xf = se.pcg(b, op, opM=Pf, mult=m.mult, tol=tol, info=i_f)
    xu = se.pcg(b, op, opM=Pu, mult=m.mult, tol=tol, info=i_u)
    _log(f"  iterations {i_f.iters} / {i_u.iters}")
    assert i_f.converged and i_f.iters == i_u.iters
    assert torch.equal(xf, xu)
    del Pf, Pu
    torch.cuda.synchronize()


@pytest.mark.parametrize("n,E,Ey,per,bc,prec", [
    (8, 64, 64, (False, False), "DDDD", "f64"),
    (8, 48, 40, (True, False), "NNDD", "f64"),
    (4, 32, 64, (False, True), "DNDN", "f64"),
    (8, 64, 64, (False, False), "DDDD", "f32"),
])
def test_block_coarse_cycle_matches_launch_per_pass(se, monkeypatch, n, E, Ey, per, bc, prec):
    """The bilinear coarse levels cycled by one workgroup (the default) against
    one launch per pass (SEM_MG_BLOCK=0 at creation): the same operations in
    the same order except for the compiler's choice of fused multiply-adds in
    the element operator, so one V-cycle agrees to 1e-12 (FP64) / 1e-5 (FP32)
    relative, and the PCG solves take the same iterations to the same
    solution within 1e-10."""
    import torch
    from semhip import workloads as W
    _log(f"block coarse cycle n={n} E={E}x{Ey} per={per} bc={bc} {prec}")
    m = se.Mesh(n, n, E, Ey, per, se.deform_sinsin(0.1))
    M = se.generateMask(bc, m)
    op = se.SEMOperator(1.0, 1.0, m, M)
    b = W.helmholtz_rhs(se, m, M, seed=6)
    tol = 1e-10 * float(b.abs().max())
    Pb = se.MG(op, precision=prec)
    monkeypatch.setenv("SEM_MG_BLOCK", "0")
    Pl = se.MG(op, precision=prec)
    monkeypatch.delenv("SEM_MG_BLOCK")
    zb, zl = Pb(b), Pl(b)
    rel = float((zb - zl).abs().max()) / float(zl.abs().max())
    _log(f"  V-cycle relative difference {rel:.2e}, bitwise {torch.equal(zb, zl)}")
    assert rel <= (1e-12 if prec == "f64" else 1e-5)
    ib, il = se.PCGInfo(), se.PCGInfo()
    xb = se.pcg(b, op, opM=Pb, mult=m.mult, tol=tol, info=ib)
    xl = se.pcg(b, op, opM=Pl, mult=m.mult, tol=tol, info=il)
    _log(f"  iterations {ib.iters} / {il.iters}")
    assert ib.converged and abs(ib.iters - il.iters) <= 1
    assert float((xb - xl).abs().max()) <= 1e-10 * float(xl.abs().max())
    assert float((b - op(xb)).abs().max()) <= 1.05 * tol
    del Pb, Pl
    torch.cuda.synchronize()


def test_mg_rejects_field_coefficients(se):
    _log("mg rejects field nu")
    m = se.Mesh(5, 5, 4, 4)
    M = se.generateMask("DDDD", m)
    nu = m.zeros() + 1.0
    op = se.SEMOperator(nu, 1.0, m, M)
    with pytest.raises(se.SemError, match="UNSUPPORTED"):
        se.MG(op)
