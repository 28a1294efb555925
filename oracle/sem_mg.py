"""CPU oracle of the multigrid preconditioner for the SEM Helmholtz operator.

TEST INFRASTRUCTURE ONLY, like ``sem_oracle``: only the test suite,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it. The product path never does.

The reference has no multigrid. Its author sketched an element-local fast
diagonalisation preconditioner (lapl.jl:105-118, commented out;
p2d_explicit.jl:108-141) as the next step past the identity/Jacobi
preconditioners of diffusion.jl:47-49. This module restates the
preconditioner the HIP library builds instead: a symmetric V-cycle of
p-multigrid (N -> N/2 -> ... -> 1 on the same elements) followed by
h-multigrid on bilinear elements (2x2 elements -> 1) with Chebyshev-Jacobi
smoothing, applied as opM of the reference's pcg (pcg.jl:16-60).

Definitions shared with the GPU implementation:

* Level operator: mask(QQ'(nu A + k B) u) rediscretised on the level's mesh
  (opLHS, diffusion.jl:36-45, on ``sem_oracle.Mesh``).
* Level meshes: polynomial order N_{l+1} = max(1, N_l // 2); once N = 1,
  elements are merged 2x2 while both element counts are even and the merged
  mesh keeps at least 2 elements per direction (and max_levels allows). The
  coarse nodal coordinates are the fine geometry evaluated at the coarse nodes
  (Lagrange interpolation inside the fine element that holds the node).
* Prolongation (coarse -> fine): per fine element, u_f = Jx u_c Jy' with
  Jx = interpMat(r_f, z_c), r_f the fine element's GLL nodes in the coarse
  element's reference coordinate ((z_f + 2 sub - 1) / 2 for the sub-element
  sub of an h-merge).
* Restriction: the transpose on assembled vectors:
  r_c = mask_c(QQ'_c(Jx' (mult_f .* r_f) Jy)).
* Smoother: Chebyshev iteration of degree k for D^-1 A on
  [lmax / ratio, lmax] with lmax = 1.1 x (Rayleigh quotient after npow power
  iterations from a hashed start vector), D the assembled diagonal.
* V-cycle: pre-smooth from 0 (degree k), residual, restrict, recurse,
  prolong-add, residual, post-smooth (degree k, the same polynomial); the
  coarsest level applies the Chebyshev polynomial of degree k_coarse from 0.
  Pre- and post-smoothers are the same polynomial in D^-1 A, so the V-cycle
  is a symmetric linear operator and plain CG applies.

Pinned in tests/test_oracle_mg.py: prolongation reproduces polynomials
exactly, restriction is the mult-weighted transpose of prolongation, the
Chebyshev smoother's error polynomial equals T_k((theta - lam)/delta) /
T_k(theta/delta) on eigenvectors of the dense assembled pencil, the power
estimate against dense eigenvalues, V-cycle symmetry and positivity, and PCG
with the V-cycle against a dense direct solve.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import sem_oracle as o

__all__ = ["MGParams", "MGLevel", "MG", "hash_start", "chebyshev", "prolong", "restrict",
           "interp_rows", "power_lmax", "coarse_coords"]


@dataclass
class MGParams:
    degree: int = 3          # Chebyshev degree of the pre- and post-smoother
    ratio: float = 20.0      # smoothing range: [lmax / ratio, lmax]
    degree_coarse: int = 8   # Chebyshev degree on the coarsest level
    npow: int = 20           # power iterations for lmax
    safety: float = 1.1      # lmax = safety x Rayleigh quotient
    max_levels: int = 32


def hash_start(msh: o.Mesh, M):
    """Deterministic, continuous start vector of the power iteration:
    M .* (h(g) / 2^32 - 1/2) with h(g) = (g * 2654435761 + 12345) mod 2^32 of
    the global node id g (Knuth's multiplicative hash)."""
    g = msh.l2g.astype(np.uint64)
    h = (g * np.uint64(2654435761) + np.uint64(12345)) % np.uint64(2 ** 32)
    return M * (h.astype(np.float64) / 4294967296.0 - 0.5)


def _level_operator(msh, nu, k, M):
    return lambda v: o.opLHS(v, nu, k, msh, M)


def _assembled_diag(msh, nu, k, M):
    """Assembled diagonal d (zero on masked nodes) and its inverse."""
    dinv = o.jacobi_diag(msh, nu, k, M)
    with np.errstate(divide="ignore"):
        d = np.where(dinv != 0, 1.0 / np.where(dinv != 0, dinv, 1.0), 0.0)
    return d, dinv


def wdot(a, b, msh):
    return float(np.sum(a * b * msh.mult))


def power_lmax(A, dinv, d, v0, msh, npow):
    """Rayleigh quotient v'Av / v'Dv after npow steps of v <- D^-1 A v /
    norm(D^-1 A v, Inf) from v0."""
    v = v0
    for _ in range(npow):
        w = dinv * A(v)
        v = w / np.max(np.abs(w))
    return wdot(v, A(v), msh) / wdot(v * v, d, msh)


def chebyshev(A, dinv, b, x, lmin, lmax, degree, r=None):
    """Chebyshev iteration for D^-1 A on [lmin, lmax] (Saad, Iterative
    Methods for Sparse Linear Systems, 2nd ed., Alg. 12.1 with the Jacobi
    preconditioner). x None means x0 = 0 (then r = b); otherwise r is the
    residual b - A x (formed here when not given). Returns the new x; the
    residual of the final x is not formed."""
    theta = 0.5 * (lmax + lmin)
    delta = 0.5 * (lmax - lmin)
    sigma = theta / delta
    rho = 1.0 / sigma
    if x is None:
        x = np.zeros_like(b)
        r = np.array(b, copy=True)
    elif r is None:
        r = b - A(x)
    d = dinv * r / theta
    for i in range(degree):
        x = x + d
        if i == degree - 1:
            break
        r = r - A(d)
        rho_n = 1.0 / (2.0 * sigma - rho)
        d = (rho_n * rho) * d + (2.0 * rho_n / delta) * (dinv * r)
        rho = rho_n
    return x


def interp_rows(xo, xi):
    """interpMat(xo, xi) (interp.jl:10-35) with an exact unit row wherever an
    output point equals an input point, so that the copies of a shared node
    interpolate to identical values."""
    xo = np.atleast_1d(np.asarray(xo, dtype=np.float64))
    J = o.interpMat(xo, xi)
    for i, x in enumerate(xo):
        hit = np.nonzero(np.asarray(xi) == x)[0]
        if hit.size:
            J[i, :] = 0.0
            J[i, hit[0]] = 1.0
    return J


def _jh(pf: int, pc: int, zf, zc, h: int):
    """Interpolation matrices (nsub, pf, pc): fine GLL nodes of each
    sub-element in the coarse element's reference coordinate."""
    if h == 0:
        return interp_rows(zf, zc)[None, :, :]
    return np.stack([interp_rows((zf + 2.0 * s - 1.0) / 2.0, zc) for s in range(2)])


def prolong(xc, lev_c: "MGLevel", lev_f: "MGLevel"):
    """u_f = Jx u_c Jy' per fine element (the coarse field evaluated at the
    fine nodes)."""
    mc, mf = lev_c.msh, lev_f.msh
    Jx, Jy = lev_f.Jx, lev_f.Jy
    x4 = xc.reshape(mc.Ex, mc.nr, mc.Ey, mc.ns)
    y6 = np.einsum("sia,xayb,tjb->xsiytj", Jx, x4, Jy)
    return y6.reshape(mf.nr * mf.Ex, mf.ns * mf.Ey)


def restrict(rf, lev_f: "MGLevel", lev_c: "MGLevel"):
    """r_c = mask_c(QQ'_c(Jx' (mult_f .* r_f) Jy)): the transpose of prolong
    on assembled vectors."""
    mc, mf = lev_c.msh, lev_f.msh
    Jx, Jy = lev_f.Jx, lev_f.Jy
    nsx, nsy = Jx.shape[0], Jy.shape[0]
    w6 = (mf.mult * rf).reshape(mc.Ex, nsx, mf.nr, mc.Ey, nsy, mf.ns)
    rc = np.einsum("sia,xsiytj,tjb->xayb", Jx, w6, Jy).reshape(mc.nr * mc.Ex, mc.ns * mc.Ey)
    return o.mask(mc.gs(rc), lev_c.M)


def coarse_coords(mf: o.Mesh, pc: int, h: int, zc):
    """The fine geometry (x, y) evaluated at the coarse GLL nodes: each coarse
    node is located in the fine element that holds it (for an h-merge, the
    sub-element on the node's side of the coarse element's middle; a node on
    the middle line takes sub-element 0) and interpolated there."""
    zf = mf.zr
    Exc, Eyc = mf.Ex >> h, mf.Ey >> h
    if h == 0:
        sub = np.zeros(pc, dtype=int)
        J = interp_rows(zc, zf)
    else:
        sub = (zc > 0).astype(int)
        J = np.stack([interp_rows(2.0 * zc[a] + 1.0 - 2.0 * sub[a], zf)[0] for a in range(pc)])
    out = []
    for f in (mf.x, mf.y):
        f4 = f.reshape(mf.Ex, mf.nr, mf.Ey, mf.ns)
        c = np.zeros((Exc, pc, Eyc, pc))
        for a in range(pc):
            for b in range(pc):
                blk = f4[(np.arange(Exc) << h) + sub[a]][:, :, (np.arange(Eyc) << h) + sub[b], :]
                c[:, a, :, b] = np.einsum("i,xiyj,j->xy", J[a], blk, J[b])
        out.append(c.reshape(Exc * pc, Eyc * pc))
    return out


@dataclass
class MGLevel:
    msh: o.Mesh
    M: np.ndarray
    nu: float
    k: float
    A: object = None
    d: Optional[np.ndarray] = None
    dinv: Optional[np.ndarray] = None
    lmax: float = 0.0
    h: int = 0                         # 1: this level's elements were merged 2x2 into the next
    Jx: Optional[np.ndarray] = None    # transfer matrices from the next-coarser level
    Jy: Optional[np.ndarray] = None


class MG:
    """The V-cycle preconditioner for opLHS(., nu, k, msh, M) (scalar nu, k)."""

    def __init__(self, msh: o.Mesh, nu: float, k: float, bc: str,
                 params: Optional[MGParams] = None, lmax: Optional[List[float]] = None):
        self.params = params or MGParams()
        self.bc = bc
        levels = [MGLevel(msh, o.generateMask(bc, msh), float(nu), float(k))]
        while len(levels) < self.params.max_levels:
            mf = levels[-1].msh
            N = mf.nr - 1
            if N > 1:
                pc, h = N // 2 + 1, 0
            elif mf.Ex % 2 == 0 and mf.Ey % 2 == 0 and mf.Ex >= 4 and mf.Ey >= 4:
                pc, h = 2, 1
            else:
                break
            zc, _ = o.gausslobatto(pc)
            xc, yc = coarse_coords(mf, pc, h, zc)
            mc = o.Mesh(pc, pc, mf.Ex >> h, mf.Ey >> h, mf.ifperiodic,
                        lambda x, y, xc=xc, yc=yc: (xc, yc), dense_gs=False)
            lev = levels[-1]
            lev.h = h
            levels.append(MGLevel(mc, o.generateMask(bc, mc), float(nu), float(k)))
            # transfer matrices live on the fine level of each pair
            lev.Jx = _jh(mf.nr, pc, mf.zr, zc, h)
            lev.Jy = _jh(mf.ns, pc, mf.zs, zc, h)
        for i, lev in enumerate(levels):
            lev.A = _level_operator(lev.msh, lev.nu, lev.k, lev.M)
            lev.d, lev.dinv = _assembled_diag(lev.msh, lev.nu, lev.k, lev.M)
            if lmax is not None:
                lev.lmax = lmax[i]
            else:
                v0 = hash_start(lev.msh, lev.M)
                lev.lmax = self.params.safety * power_lmax(lev.A, lev.dinv, lev.d, v0, lev.msh,
                                                           self.params.npow)
        self.levels = levels

    def _smooth(self, lev, b, x, r=None, degree=None):
        p = self.params
        return chebyshev(lev.A, lev.dinv, b, x, lev.lmax / p.ratio, lev.lmax,
                         degree or p.degree, r)

    def vcycle(self, b, l: int = 0):
        lev = self.levels[l]
        if l == len(self.levels) - 1:
            return self._smooth(lev, b, None, degree=self.params.degree_coarse)
        x = self._smooth(lev, b, None)
        r = b - lev.A(x)
        xc = self.vcycle(restrict(r, lev, self.levels[l + 1]), l + 1)
        x = x + prolong(xc, self.levels[l + 1], lev)
        r = b - lev.A(x)
        return self._smooth(lev, b, x, r)

    def __call__(self, r):
        return self.vcycle(r, 0)
