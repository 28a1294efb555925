"""Writes tests/golden/golden.npz: regression fixtures produced by the oracle
(oracle/sem_oracle.py) at the sizes the reference's examples use.

Contents (all float64 unless noted; fields in the local column-major layout):
  gll{p}_z, gll{p}_w, gll{p}_D      GLL nodes, weights, derivMat for p = 8, 9, 10, 16
  c1_l2g (int32), c1_mult, c1_mask  4x4 elements, p = 9, undeformed, bc DDDD
  c1_pcg_x, c1_pcg_iters (int64)    Poisson -Lap u = 2 pi^2 sin sin, Jacobi PCG to
                                    1e-10 * norm(b, Inf) (the 4x4-element N=8 Poisson solve)
  d_x, d_y, d_B, d_G11, d_G12, d_G22, d_Jac
                                    2x2 elements, p = 8, sin-sin map (amplitude 0.1)
  d_u, d_hlmz                       hlmz(u, 1.3, 0.7) on that mesh, u ~ N(0,1) seed 7
  p_u, p_gs                         gatherScatter on 3x2 elements, p = 5, periodic x,
                                    affine map, u ~ N(0,1) seed 8
  bdf_t{k}, bdf_a{k}, bdf_b{k}      bdfExtK of three time histories

Usage: python tests/golden/make_golden.py   (rewrites golden.npz next to it)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import sem_oracle as o  # noqa: E402

BDF_T = [[1e-3, 0.0, 0.0], [2e-3, 1e-3, 0.0], [0.7, 0.5, 0.2, 0.1]]


def golden():
    g = {}
    for p in (8, 9, 10, 16):
        z, w = o.gausslobatto(p)
        g[f"gll{p}_z"], g[f"gll{p}_w"], g[f"gll{p}_D"] = z, w, o.derivMat(z)
    m1 = o.Mesh(9, 9, 4, 4, deform=o.fixU, dense_gs=False)
    M1 = o.generateMask("DDDD", m1)
    g["c1_l2g"] = m1.l2g.astype(np.int32)
    g["c1_mult"], g["c1_mask"] = m1.mult, M1
    f = 2 * np.pi ** 2 * np.sin(np.pi * m1.x) * np.sin(np.pi * m1.y)
    b = m1.gs(M1 * o.mass(f, m1))
    op = lambda v: o.opLHS(v, 1.0, 0.0, m1, M1)
    d = o.jacobi_diag(m1, 1.0, 0.0, M1)
    info = o.PCGInfo()
    g["c1_pcg_x"] = o.pcg(b, op, opM=lambda r: d * r, mult=m1.mult,
                          tol=1e-10 * np.abs(b).max(), info=info)
    g["c1_pcg_iters"] = np.array(info.iters, dtype=np.int64)
    md = o.Mesh(8, 8, 2, 2, deform=o.deform_sinsin(0.1), dense_gs=False)
    for k in ("x", "y", "B", "G11", "G12", "G22", "Jac"):
        g[f"d_{k}"] = getattr(md, k)
    u = np.random.default_rng(7).standard_normal(md.x.shape)
    g["d_u"], g["d_hlmz"] = u, o.hlmz(u, 1.3, 0.7, md)
    mp = o.Mesh(5, 5, 3, 2, ifperiodic=(True, False), deform=o.deform_affine01, dense_gs=True)
    u = np.random.default_rng(8).standard_normal(mp.x.shape)
    g["p_u"], g["p_gs"] = u, mp.gs(u)
    for i, t in enumerate(BDF_T):
        a, bb = o.bdfExtK(t, k=len(t) - 1)
        g[f"bdf_t{i}"], g[f"bdf_a{i}"], g[f"bdf_b{i}"] = np.array(t), a, bb
    return g


if __name__ == "__main__":
    out = os.path.join(HERE, "golden.npz")
    np.savez_compressed(out, **golden())
    print("wrote", out, os.path.getsize(out), "bytes")
