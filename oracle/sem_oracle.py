"""CPU oracle for the SEM operator-apply-and-solve path of SpectralElements.jl.

TEST INFRASTRUCTURE ONLY. This module is a NumPy restatement of the reference
algorithm. Only the test suite, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only as the checker or
the timed CPU baseline. The product path (``semhip`` over ``libsemhip.so``)
never imports it and never falls back to it.

Layout: every nodal field is a 2-D float64 array ``u[i, j]`` of shape
``(nr*Ex, ns*Ey)``; element ``(ex, ey)`` owns rows ``ex*nr .. ex*nr+nr-1`` and
columns ``ey*ns .. ey*ns+ns-1`` (0-based). This is the reference's column-major
local array (mesh.jl:94, mesh.jl:100); reshapes use Fortran order so they mean
what Julia's ``reshape`` means.

Each function cites the reference file:line it restates. Pinning of each
function (closed forms, dense brute force, invariants) lives in
``tests/test_oracle_*.py``. Parity status per function:

* gausslobatto: pinned by textbook closed forms, quadrature exactness and an
  independent eigenvalue construction; bit-level agreement with
  FastGaussQuadrature (third party, absent here) is parity unpinned.
* everything else: pinned (see the tests named in each docstring).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Callable, Optional, Sequence

import numpy as np

__all__ = [
    "gausslobatto", "derivMat", "interpMat", "semq", "semmesh", "ndgrid",
    "jac", "ABu", "Mesh", "generateMask", "laplace", "lapl", "mass", "hlmz",
    "gatherScatter", "gatherScatter_structural", "mask", "pcg", "PCGInfo",
    "grad", "gradT", "advect", "bdfExtK", "TimeStepper", "jacobi_diag",
    "deform_sinsin", "deform_affine01", "fixU", "local_to_global",
    "advect_transpose", "vardiff_operator", "vardiff_gradient", "cd_forward",
    "laplace_dealiased", "lapl12", "mass8",
    "cd_adjoint", "cd_objective", "annulus", "diffusion_simulate",
    "gordonHall", "semInterp",
]


# --------------------------------------------------------------------------
# L0: 1-D spectral primitives
# --------------------------------------------------------------------------
def _legendre_and_deriv(N: int, x):
    """P_N(x) and P_N'(x) by the three-term recurrence (textbook)."""
    x = np.asarray(x, dtype=np.float64)
    p0 = np.ones_like(x)
    if N == 0:
        return p0, np.zeros_like(x)
    p1 = x.copy()
    for k in range(2, N + 1):
        p0, p1 = p1, ((2 * k - 1) * x * p1 - (k - 1) * p0) / k
    # P_N' = N (x P_N - P_{N-1}) / (x^2 - 1), valid away from +-1
    with np.errstate(divide="ignore", invalid="ignore"):
        dp = N * (x * p1 - p0) / (x * x - 1.0)
    return p1, dp


def gausslobatto(n: int):
    """GLL nodes and weights on [-1, 1] with ``n`` points.

    Restates FastGaussQuadrature.gausslobatto(n) as called at mesh.jl:70-71 and
    semmesh.jl:11: nodes are -1, +1 and the roots of P'_N (N = n-1), ascending;
    weights w_i = 2 / (N (N+1) P_N(z_i)^2). Interior roots start from the
    Chebyshev-Gauss-Lobatto guess -cos(pi i / N) and are polished by Newton on
    (1-x^2) P_N'(x), whose derivative is -N(N+1) P_N(x).
    """
    if n < 2:
        raise ValueError("gausslobatto needs n >= 2")
    N = n - 1
    z = -np.cos(np.pi * np.arange(n) / N)
    z[0], z[-1] = -1.0, 1.0
    if n > 2:
        xi = z[1:-1].copy()
        for _ in range(100):
            P, dP = _legendre_and_deriv(N, xi)
            f = (1.0 - xi * xi) * dP
            df = -N * (N + 1) * P
            step = f / df
            xi = xi - step
            if np.max(np.abs(step)) < 1e-16:
                break
        # symmetrise: z_i = -z_{N-i} exactly
        xi = 0.5 * (xi - xi[::-1])
        z[1:-1] = xi
    if n % 2 == 1:
        z[N // 2] = 0.0
    P, _ = _legendre_and_deriv(N, z)
    w = 2.0 / (N * (N + 1) * P * P)
    return z, w


def derivMat(x):
    """Lagrange derivative matrix on points x (derivMat.jl:9-35).

    Barycentric weights a_i = 1/prod_{j!=i}(x_i-x_j) (:13-18); diagonal
    D_ii = sum_{j!=i} 1/(x_i-x_j) (:21-27); off-diagonal
    D_ij = a_j / (a_i (x_i - x_j)) (:30-32). Pinned in
    tests/test_oracle_spectral.py (D.1 = 0, exact derivative of x^k, GLL corner
    closed forms).
    """
    x = np.asarray(x, dtype=np.float64).reshape(-1)
    n = x.size
    a = np.ones(n)
    for i in range(n):
        for j in range(i):
            a[i] = a[i] * (x[i] - x[j])
        for j in range(i + 1, n):
            a[i] = a[i] * (x[i] - x[j])
    a = 1.0 / a
    D = x[:, None] - x[None, :]
    for i in range(n):
        D[i, i] = 1.0
    D = 1.0 / D
    for i in range(n):
        D[i, i] = 0.0
        D[i, i] = _julia_sum(D[i, :])
    for j in range(n):
        for i in range(n):
            if i != j:
                D[i, j] = a[j] / (a[i] * (x[i] - x[j]))
    return D


def _julia_sum(v):
    """Julia's ``sum`` of a short vector: pairwise with a blocksize of 1024,
    which is a plain left-to-right loop for fewer than 1024 entries."""
    s = 0.0
    for t in np.asarray(v).reshape(-1):
        s = s + float(t)
    return s


def interpMat(xo, xi):
    """Lagrange interpolation matrix from points xi to xo (interp.jl:10-35)."""
    xo = np.atleast_1d(np.asarray(xo, dtype=np.float64)).reshape(-1)
    xi = np.atleast_1d(np.asarray(xi, dtype=np.float64)).reshape(-1)
    no, ni = xo.size, xi.size
    a = np.ones(ni)
    for i in range(ni):
        for j in range(i):
            a[i] = a[i] * (xi[i] - xi[j])
        for j in range(i + 1, ni):
            a[i] = a[i] * (xi[i] - xi[j])
    a = 1.0 / a
    J = np.zeros((no, ni))
    s = np.ones(ni)
    t = np.ones(ni)
    for i in range(no):
        x = xo[i]
        for j in range(1, ni):
            s[j] = s[j - 1] * (x - xi[j - 1])
            t[ni - 1 - j] = t[ni - j] * (x - xi[ni - j])
        J[i, :] = a * s * t
    return J


def semq(E: int, n: int, bc: bool):
    """Scatter operator Q (global -> local) as a dense 0/1 matrix (semq.jl:6-25).

    Local row e*n + a maps to global column e*(n-1) + a; if periodic the last
    local row maps to column 0 and the last column is dropped (:19-22).
    """
    Q = np.zeros((E * n, E * (n - 1) + 1))
    i = j = 0
    for _ in range(E):
        Q[i:i + n, j:j + n] += np.eye(n)
        i += n
        j += n - 1
    if bc:
        Q[-1, 0] = 1.0
        Q = Q[:, :-1]
    return Q


def _julia_range(E: int):
    """range(start=-1, stop=1, length=E+1) entries, correctly rounded
    (Julia's StepRangeLen evaluates them in twice precision)."""
    return np.array([(2.0 * k - E) / E for k in range(E + 1)])


def semmesh(E: int, n: int):
    """1-D element coordinates with duplicated end-points (semmesh.jl:9-27)."""
    z0, w0 = gausslobatto(n)
    z0 = 0.5 * (z0 + 1.0)
    w0 = 0.5 * w0
    ze = _julia_range(E)
    dz = np.diff(ze)
    z = np.kron(dz, z0) + np.kron(ze[:-1], np.ones(n))
    w = np.kron(dz, w0)
    return z.reshape(E * n), w.reshape(E * n)


def ndgrid(v1, v2):
    """ndgrid.jl:8-13."""
    v1 = np.asarray(v1).reshape(-1, 1)
    v2 = np.asarray(v2).reshape(1, -1)
    return np.repeat(v1, v2.size, axis=1), np.repeat(v2, v1.size, axis=0)


# --------------------------------------------------------------------------
# L1: tensor-product kernel
# --------------------------------------------------------------------------
def _empty(A) -> bool:
    return A is None or (hasattr(A, "size") and A.size == 0) or (isinstance(A, (list, tuple)) and len(A) == 0)


def ABu(As, Br, u):
    """(As kron Br) * u applied block-wise over the local array (ABu.jl:9-37).

    ``[]`` (here None or an empty array) means identity. Br acts on every group
    of size(Br,2) consecutive rows via one GEMM on reshape(u, nb, :) (:14-20);
    As acts on every block of size(As,2) columns (:23-34). Pinned in
    tests/test_oracle_operators.py against explicit kron matrices (the
    test/Operators.jl:22 identity A*U*B').
    """
    u = np.asarray(u, dtype=np.float64)
    m, n = u.shape
    Bu = u
    if not _empty(Br):
        Br = np.asarray(Br, dtype=np.float64)
        mb, nb = Br.shape
        m = (m * mb) // nb
        Bu = Br @ Bu.reshape(nb, -1, order="F")
        Bu = Bu.reshape(m, n, order="F")
    out = Bu
    if not _empty(As):
        As = np.asarray(As, dtype=np.float64)
        ma, na = As.shape
        nblk = n // na
        blocks = out.reshape(m, na, nblk, order="F")
        tmp = np.einsum("rlb,kl->rkb", blocks, As, optimize=True)
        out = tmp.reshape(m, ma * nblk, order="F")
    return np.ascontiguousarray(out)


# --------------------------------------------------------------------------
# L2: geometry and mesh
# --------------------------------------------------------------------------
def jac(x, y, Dr, Ds):
    """Jacobian and inverse metric (jac.jl:24-40)."""
    xr = ABu(None, Dr, x)
    xs = ABu(Ds, None, x)
    yr = ABu(None, Dr, y)
    ys = ABu(Ds, None, y)
    J = xr * ys - xs * yr
    Ji = 1.0 / J
    rx = Ji * ys
    ry = -Ji * xs
    sx = -Ji * yr
    sy = Ji * xr
    return J, Ji, rx, ry, sx, sy


def fixU(*u):
    """Identity deformation (mesh.jl:6-8)."""
    return u


def deform_sinsin(amp: float = 0.1):
    """x' = x + amp sin(pi x) sin(pi y), y' = y + amp sin(pi x) sin(pi y)."""
    def f(x, y):
        s = amp * np.sin(np.pi * x) * np.sin(np.pi * y)
        return x + s, y + s
    return f


def deform_affine01(x, y):
    """usrgeom of p2d_explicit.jl:78 / semPS.jl:306-310: [-1,1]^2 -> [0,1]^2."""
    return 0.5 * (x + 1.0), 0.5 * (y + 1.0)


def annulus(r0: float = 0.5, r1: float = 1.0, span: float = 2.0 * math.pi):
    """annulus(r, s; r0, r1, span) (geom.jl:40-50): R = (r1 - r0)/2 (r + 1) +
    r0, th = span/2 (s + 1), (x, y) = (R cos th, R sin th)."""
    def f(x, y):
        R = (r1 - r0) / 2 * (x + 1) + r0
        th = span / 2 * (y + 1) + 0.0
        return R * np.cos(th), R * np.sin(th)
    return f


def local_to_global(Ex: int, Ey: int, nr: int, ns: int, periodic=(False, False)):
    """0-based global id of every local node (the intended local l2g of
    mesh.jl:87-89, i.e. Qx * reshape(1:nxg*nyg) * Qy'), as an int64 array.

    Global numbering is column-major over the (nxg, nyg) global grid. Pinned
    against the dense ``semq`` matrices in tests/test_oracle_mesh.py.
    """
    def g1(E, n, per):
        e = np.repeat(np.arange(E), n)
        a = np.tile(np.arange(n), E)
        g = e * (n - 1) + a
        ng = E * (n - 1) + 1
        if per:
            ng = E * (n - 1)
            g = g % ng
        return g, ng
    gx, nxg = g1(Ex, nr, periodic[0])
    gy, nyg = g1(Ey, ns, periodic[1])
    return gx[:, None] + nxg * gy[None, :], nxg * nyg


@dataclass
class Mesh:
    """Mesh (mesh.jl:25-64) built as in the constructor mesh.jl:66-133.

    ``dense_gs=True`` keeps the dense QQtx/QQty the reference stores
    (mesh.jl:44-45, 81-82); they are (nr*Ex)^2 and (ns*Ey)^2 doubles.
    """
    nr: int
    ns: int
    Ex: int
    Ey: int
    ifperiodic: Sequence[bool] = (False, False)
    deform: Callable = fixU
    dense_gs: bool = True

    def __post_init__(self):
        nr, ns, Ex, Ey = self.nr, self.ns, self.Ex, self.Ey
        per = tuple(bool(p) for p in self.ifperiodic)
        self.ifperiodic = per
        self.zr, self.wr = gausslobatto(nr)
        self.zs, self.ws = gausslobatto(ns)
        self.Dr = derivMat(self.zr)
        self.Ds = derivMat(self.zs)
        if self.dense_gs:
            Qx = semq(Ex, nr, per[0])
            # mesh.jl:80 passes Ex for the y direction; the intent (and every
            # configuration, which has Ex == Ey) is Ey.
            Qy = semq(Ey, ns, per[1])
            self.QQtx = Qx @ Qx.T
            self.QQty = Qy @ Qy.T
        else:
            self.QQtx = self.QQty = None
        self.l2g, self.nglobal = local_to_global(Ex, Ey, nr, ns, per)
        ones = np.ones((nr * Ex, ns * Ey))
        self.mult = 1.0 / self.gs(ones)
        xe, _ = semmesh(Ex, nr)
        ye, _ = semmesh(Ey, ns)
        x, y = ndgrid(xe, ye)
        x, y = self.deform(x, y)
        self.x, self.y = np.asarray(x, dtype=np.float64), np.asarray(y, dtype=np.float64)
        (self.Jac, self.Jaci, self.rx, self.ry, self.sx, self.sy) = jac(
            self.x, self.y, self.Dr, self.Ds)
        wx = np.kron(np.ones(Ex), self.wr)
        wy = np.kron(np.ones(Ey), self.ws)
        self.B = self.Jac * np.outer(wx, wy)
        self.Bi = 1.0 / self.B
        rx, ry, sx, sy, B = self.rx, self.ry, self.sx, self.sy, self.B
        self.G11 = B * (rx * rx + ry * ry)
        self.G12 = B * (rx * sx + ry * sy)
        self.G22 = B * (sx * sx + sy * sy)

    def gs(self, u):
        if self.dense_gs:
            return gatherScatter(u, self.QQtx, self.QQty)
        return gatherScatter_structural(u, self)


def generateMask(bc: Sequence[str], msh: Mesh):
    """Dirichlet mask M = mx my' as float64 0/1 (mesh.jl:149-175; Field stores
    it as Array{T}, mesh.jl:183)."""
    nx, ny = msh.nr * msh.Ex, msh.ns * msh.Ey
    mx = np.ones(nx)
    my = np.ones(ny)
    if bc[0] == "D":
        mx[0] = 0.0
    if bc[1] == "D":
        mx[-1] = 0.0
    if bc[2] == "D":
        my[0] = 0.0
    if bc[3] == "D":
        my[-1] = 0.0
    if msh.ifperiodic[0]:
        mx[:] = 1.0
    if msh.ifperiodic[1]:
        my[:] = 1.0
    return np.outer(mx, my)


# --------------------------------------------------------------------------
# L3: operators and assembly
# --------------------------------------------------------------------------
def laplace(u, Dr, Ds, G11, G12, G22):
    """Element weak Laplacian Dr'(G11 ur + G12 us) + (G12 ur + G22 us) Ds
    (lapl.jl:70-81)."""
    ur = ABu(None, Dr, u)
    us = ABu(Ds, None, u)
    wr = G11 * ur + G12 * us
    ws = G12 * ur + G22 * us
    return ABu(None, Dr.T, wr) + ABu(Ds.T, None, ws)


def lapl(u, msh: Mesh, nu=None):
    """lapl(u, msh) (lapl.jl:26-36) and lapl(u, nu, msh) = nu .* lapl (:38-45)."""
    Au = laplace(u, msh.Dr, msh.Ds, msh.G11, msh.G12, msh.G22)
    if nu is not None:
        Au = nu * Au
    return Au


def laplace_dealiased(u, Jr, Js, Dr, Ds, G11, G12, G22):
    """8-argument laplace (lapl.jl:83-103): derivatives on the GLL grid,
    interpolated by (Js kron Jr), G applied on the interpolation grid,
    projected back by the transposes, then Dr' / Ds'."""
    ur = ABu(None, Dr, u)
    us = ABu(Ds, None, u)
    Jur = ABu(Js, Jr, ur)
    Jus = ABu(Js, Jr, us)
    vr = G11 * Jur + G12 * Jus
    vs = G12 * Jur + G22 * Jus
    wr = ABu(Js.T, Jr.T, vr)
    ws = ABu(Js.T, Jr.T, vs)
    return ABu(None, Dr.T, wr) + ABu(Ds.T, None, ws)


def lapl12(u, M, Jr, Js, QQtx, QQty, Dr, Ds, G11, G12, G22, mult):
    """12-argument lapl with gather-scatter and mask (lapl.jl:54-68); Jr = Js
    = [] is the plain laplace (ABu of [] is the identity); the Zygote hook is
    the identity in the forward pass."""
    if _empty(Jr) and _empty(Js):
        Au = laplace(u, Dr, Ds, G11, G12, G22)
    else:
        Au = laplace_dealiased(u, Jr, Js, Dr, Ds, G11, G12, G22)
    Au = gatherScatter(Au, QQtx, QQty)
    return mask(Au, M)


def mass8(u, M, B, Jr, Js, QQtx, QQty, mult):
    """8-argument mass (mass.jl:32-50): Js'Jr'(B .* (Js kron Jr) u) (B empty:
    no weight), then gather-scatter and mask."""
    Ju = ABu(Js, Jr, u)
    BJu = Ju if _empty(B) else B * Ju
    Bu = ABu(None if _empty(Js) else Js.T, None if _empty(Jr) else Jr.T, BJu)
    Bu = gatherScatter(Bu, QQtx, QQty)
    return mask(Bu, M)


def mass(u, msh: Mesh):
    """B .* u (mass.jl:12-22)."""
    return msh.B * u


def hlmz(u, nu, k, msh: Mesh):
    """nu .* lapl(u) .+ k .* mass(u) (hlmz.jl:12-19)."""
    Hu = nu * lapl(u, msh)
    Hu = Hu + k * mass(u, msh)
    return Hu


def gatherScatter(u, QQtx, QQty):
    """QQ' u = QQtx * u * QQty' via ABu(QQty, QQtx, u) (gatherScatter.jl:8-16)."""
    return ABu(QQty, QQtx, u)


def gatherScatter_structural(u, msh: Mesh):
    """Same result as the dense form without the dense matrices: x-pairs are
    summed first, then y-pairs, which is exactly what the 0/1 GEMMs produce
    (each output is a sum of at most two nonzero products per pass). Pinned
    against the dense form in tests/test_oracle_mesh.py."""
    def pass1(v, E, n, per, axis):
        v = np.moveaxis(v, axis, 0).copy()
        out = v.copy()
        for e in range(E - 1):
            i = e * n + n - 1
            s = v[i] + v[i + 1]
            out[i] = s
            out[i + 1] = s
        if per:
            s = v[E * n - 1] + v[0]
            out[E * n - 1] = s
            out[0] = s
            if E == 1 and n == 1:
                out[0] = v[0]
        return np.moveaxis(out, 0, axis)
    v = pass1(np.asarray(u, dtype=np.float64), msh.Ex, msh.nr, msh.ifperiodic[0], 0)
    v = pass1(v, msh.Ey, msh.ns, msh.ifperiodic[1], 1)
    return v


def mask(u, M):
    """M .* u, or a copy when M is empty (mask.jl:10-18)."""
    if _empty(M):
        return np.array(u, copy=True)
    return M * u


def opLHS(u, nu, k, msh: Mesh, M):
    """Diffusion LHS: mask(gatherScatter(hlmz(u, nu, k))) (diffusion.jl:36-45)."""
    return mask(msh.gs(hlmz(u, nu, k, msh)), M)


def jacobi_diag(msh: Mesh, nu, k, M):
    """Inverse assembled diagonal of M QQ'(nu A + k B), zero on masked nodes.

    Not in the reference (its Diffusion preconditioner is the identity,
    diffusion.jl:47-49). Element diagonal for node (a, b):
    nu [sum_i Dr[i,a]^2 G11[i,b] + 2 Dr[a,a] Ds[b,b] G12[a,b]
        + sum_j Ds[j,b]^2 G22[a,j]] + k B[a,b],
    assembled with QQ' and masked. Pinned against diag of the dense assembled
    operator in tests/test_oracle_operators.py.
    """
    nr, ns = msh.nr, msh.ns
    Dr2 = msh.Dr ** 2
    Ds2 = msh.Ds ** 2
    t1 = ABu(None, Dr2.T, msh.G11)
    t3 = ABu(Ds2.T, None, msh.G22)
    dr = np.tile(np.diag(msh.Dr), msh.Ex)
    ds = np.tile(np.diag(msh.Ds), msh.Ey)
    t2 = 2.0 * np.outer(dr, ds) * msh.G12
    d = nu * (t1 + t2 + t3) + k * msh.B
    d = msh.gs(d)
    with np.errstate(divide="ignore"):
        dinv = np.where(M != 0, 1.0 / d, 0.0)
    return dinv


# --------------------------------------------------------------------------
# L4: solver
# --------------------------------------------------------------------------
@dataclass
class PCGInfo:
    iters: int = 0
    res: float = 0.0
    converged: bool = False
    history: list = field(default_factory=list)


def pcg(b, opA, opM=None, mult=None, ifv=False, tol=1e-8, maxiter=None, info: Optional[PCGInfo] = None):
    """Left-preconditioned CG from x0 = 0 (pcg.jl:16-60).

    Stops when norm(r, Inf) <= tol (absolute, :36); inner products are
    sum(a .* b .* mult) (:45, :49, :52); at k == maxiter it warns and returns
    the current x (:39). ``info`` receives the iteration count, the final
    infinity norm and whether the tolerance was met.
    """
    b = np.asarray(b, dtype=np.float64)
    if opM is None:
        opM = lambda v: v
    if mult is None:
        mult = np.ones_like(b)
    if maxiter is None:
        maxiter = b.size
    x = np.zeros_like(b)
    ra = b - np.zeros_like(b)
    hp = np.zeros_like(b)
    rp = np.zeros_like(b)
    u = np.zeros_like(b)
    k = 0
    converged = True
    hist = []
    while np.max(np.abs(ra)) > tol:
        hist.append(float(np.max(np.abs(ra))))
        ha = opM(ra)
        if k == maxiter:
            converged = False
            break
        k += 1
        hpp, rpp = hp, rp
        hp, rp = ha, ra
        t = np.sum(rp * hp * mult)
        if k == 1:
            u = np.array(hp, copy=True)
        else:
            u = hp + (t / np.sum(rpp * hpp * mult)) * u
        Au = opA(u)
        a = t / np.sum(u * Au * mult)
        x = x + a * u
        ra = rp - a * Au
    if info is not None:
        info.iters = k
        info.res = float(np.max(np.abs(ra)))
        info.converged = converged
        info.history = hist
    return x


# --------------------------------------------------------------------------
# gradient, advection, time stepping
# --------------------------------------------------------------------------
def grad(u, msh: Mesh):
    """(ux, uy) = (rx ur + sx us, ry ur + sy us) (grad.jl:15-34)."""
    ur = ABu(None, msh.Dr, u)
    us = ABu(msh.Ds, None, u)
    return msh.rx * ur + msh.sx * us, msh.ry * ur + msh.sy * us


def gradT(ux, uy, msh: Mesh):
    """Transpose of grad applied to (ux, uy): Dr'(rx ux + ry uy) + (sx ux + sy uy) Ds.

    grad.jl:44-63 returns the two halves separately for one input; this is the
    adjoint of ``grad`` as a map L -> 2L, which is what the adjoint gradient
    needs. Pinned by <grad u, (vx, vy)> = <u, gradT(vx, vy)> in the tests.
    """
    Dr, Ds = msh.Dr, msh.Ds
    return (ABu(None, Dr.T, msh.rx * ux + msh.ry * uy)
            + ABu(Ds.T, None, msh.sx * ux + msh.sy * uy))


def advect(T, vx, vy, mshV: Mesh, mshD: Mesh, Jr=None, Js=None):
    """Dealiased convective term Js'Jr'[B_D (Jvx JTx + Jvy JTy)] (advect.jl:45-78)."""
    if Jr is None:
        Jr = interpMat(mshD.zr, mshV.zr)
    if Js is None:
        Js = interpMat(mshD.zs, mshV.zs)
    Tx, Ty = grad(T, mshV)
    JTx = ABu(Js, Jr, Tx)
    JTy = ABu(Js, Jr, Ty)
    Jux = ABu(Js, Jr, vx)
    Juy = ABu(Js, Jr, vy)
    JCu = Jux * JTx + Juy * JTy
    JCu = JCu * mshD.B
    return ABu(Js.T, Jr.T, JCu)


def advect_transpose(w, vx, vy, mshV: Mesh, mshD: Mesh, Jr=None, Js=None):
    """Transpose of ``advect`` as a linear map in T: the ABu pullback rule
    (ABu.jl:39-41) applied to each factor of advect.jl:54-61, then gradT.
    Pinned in the tests by <advect(T), w> = <T, advect_transpose(w)>."""
    if Jr is None:
        Jr = interpMat(mshD.zr, mshV.zr)
    if Js is None:
        Js = interpMat(mshD.zs, mshV.zs)
    Jw = ABu(Js, Jr, w) * mshD.B
    Jux = ABu(Js, Jr, vx)
    Juy = ABu(Js, Jr, vy)
    return gradT(ABu(Js.T, Jr.T, Jux * Jw), ABu(Js.T, Jr.T, Juy * Jw), mshV)


def _julia_unique(t):
    out = []
    for v in t:
        if not any(v == w for w in out):
            out.append(v)
    return np.array(out, dtype=np.float64)


def bdfExtK(t, k=3):
    """BDFk / EXTk coefficients from the time history t = [t_{n+1}, t_n, ...]
    (time.jl:31-53). Returns (a: k, b: k+1)."""
    t = _julia_unique(np.asarray(t, dtype=np.float64))
    kk = t.size - 1
    t1 = t[0]
    t0 = t[1:]
    a = interpMat(t1, t0).reshape(-1) if kk > 0 else np.zeros(0)
    b = derivMat(t)[0, :].copy()
    if kk < k:
        a = np.concatenate([a, np.zeros(k - kk)])
        b = np.concatenate([b, np.zeros(k - kk)])
    else:
        a = a[:k]
        b = b[:k + 1]
    if kk == 0:
        a[0] = 1.0
    return a.reshape(k), b.reshape(k + 1)


class TimeStepper:
    """time.jl:70-112.

    The reference constructor always calls bdfExtK(time) with its default k=3
    (time.jl:87), so its k != 3 steppers fail at the first bdfExtK! broadcast;
    here the constructor uses the stepper's own k, which equals the reference
    for k = 3 and makes the BDF2/EXT2 (k = 2) stepper usable.
    """

    def __init__(self, Ti, Tf, dt, k=3):
        self.time = np.full(k + 1, float(Ti))
        self.bdfA, self.bdfB = bdfExtK(self.time, k=k)
        self.istep = 0
        self.fstep = int(math.ceil(Tf / dt)) if dt != 0 else 0
        self.dt, self.Ti, self.Tf = float(dt), float(Ti), float(Tf)

    def update(self):
        t = self.time
        for i in range(t.size - 1, 0, -1):
            t[i] = t[i - 1]
        t[0] = t[1]
        self.istep += 1
        t[0] += self.dt
        self.bdfA, self.bdfB = bdfExtK(t, k=t.size - 1)


# --------------------------------------------------------------------------
# steady variable-coefficient diffusion and its adjoint gradient
# --------------------------------------------------------------------------
def vardiff_operator(msh: Mesh, nu, M):
    """v -> M QQ'[Dr'(nu G11 vr + nu G12 vs) + (nu G12 vr + nu G22 vs) Ds]:
    the 12-argument lapl of lapl.jl:54-68 with nu folded into G
    (semPS.jl:161-165; p2d_explicit.jl:105-107)."""
    G11, G12, G22 = nu * msh.G11, nu * msh.G12, nu * msh.G22
    return lambda v: mask(msh.gs(laplace(v, msh.Dr, msh.Ds, G11, G12, G22)), M)


def vardiff_gradient(msh: Mesh, nu, f, ustar, M, tol_rel=1e-13):
    """J(nu) = 1/2 sum_local B (u - u*)^2 with A(nu) u = M QQ'(B f); returns
    (J, u, lam, dJ/dnu) where dJ/dnu = -B (grad lam . grad u) per local node
    (the pullback of resi = rhs - lhs(u) at lam, adjoint.jl:13-17) and lam
    solves A lam = M QQ'(B (u - u*))."""
    A = vardiff_operator(msh, nu, M)
    b = msh.gs(M * (msh.B * f))
    u = pcg(b, A, mult=msh.mult, tol=tol_rel * np.abs(b).max())
    e = u - ustar
    J = 0.5 * np.sum(msh.B * e * e)
    g = msh.gs(M * (msh.B * e))
    lam = pcg(g, A, mult=msh.mult, tol=tol_rel * max(np.abs(g).max(), 1e-300))
    lx, ly = grad(lam, msh)
    ux, uy = grad(u, msh)
    return J, u, lam, -msh.B * (lx * ux + ly * uy)


# --------------------------------------------------------------------------
# unsteady convection-diffusion (convectionDiffusion.jl) and its adjoint
# --------------------------------------------------------------------------
def cd_step_matrices(mV: Mesh, mD: Mesh):
    Jr = interpMat(mD.zr, mV.zr)
    Js = interpMat(mD.zs, mV.zs)
    return Jr, Js


def _cd_precond(mV: Mesh, nu, b0, M, precond):
    """'jacobi': the inverse assembled diagonal of the step operator (keeps
    every iterate continuous). 'reference': u ./ B ./ b0 per local copy
    (convectionDiffusion.jl:87-91); the per-copy B makes the iterates
    discontinuous at shared nodes, so the solve is only defined up to the
    kernel of QQ' and the adjoint is not its exact transpose."""
    if precond == "reference":
        return 1.0 / (mV.B * b0)
    return jacobi_diag(mV, nu, b0, M)


def cd_forward(mV: Mesh, mD: Mesh, u0, vx, vy, nu, dt, nsteps, M, k=2, tol_rel=1e-12,
               keep=False, precond="jacobi"):
    """BDFk/EXTk convection-diffusion (convectionDiffusion.jl:76-179) with f = 0
    and homogeneous Dirichlet data, stepped by count. Each step: histories
    shift (mesh.jl:199-215), coefficients from bdfExtK (time.jl:31-53), RHS
    = QQ'(M(-sum b_{1+i} B uh_i + sum a_i exH_i)), exH_i = -advect(uh_i), and
    a PCG solve of M QQ' hlmz(u, nu, b_1) preconditioned by ``_cd_precond``.
    Returns the states u^0..u^N if keep."""
    Jr, Js = cd_step_matrices(mV, mD)
    ts = TimeStepper(0.0, nsteps * dt, dt, k=k)
    uh = [np.zeros_like(u0) for _ in range(k)]
    u = np.array(u0, dtype=np.float64)
    states = [u.copy()] if keep else None
    for _ in range(nsteps):
        for i in range(k - 1, 0, -1):
            uh[i] = uh[i - 1]
        uh[0] = u.copy()
        ts.update()
        a, b = ts.bdfA, ts.bdfB
        rhs = np.zeros_like(u)
        for i in range(k):
            if b[1 + i] != 0.0:
                rhs = rhs - b[1 + i] * mass(uh[i], mV)
            if a[i] != 0.0:
                rhs = rhs + a[i] * (-advect(uh[i], vx, vy, mV, mD, Jr, Js))
        rhs = mV.gs(mask(rhs, M))
        opL = lambda v: mask(mV.gs(hlmz(v, nu, b[0], mV)), M)
        pre = _cd_precond(mV, nu, b[0], M, precond)
        u = pcg(rhs, opL, opM=lambda r: pre * r, mult=mV.mult,
                tol=tol_rel * max(np.abs(rhs).max(), 1e-300))
        if keep:
            states.append(u.copy())
    return (u, states) if keep else u


def _cd_coefficients(dt, nsteps, k):
    """Per-step (bdfA, bdfB) of the TimeStepper (time.jl:70-112), step 1..N."""
    ts = TimeStepper(0.0, nsteps * dt, dt, k=k)
    out = []
    for _ in range(nsteps):
        ts.update()
        out.append((np.array(ts.bdfA, dtype=np.float64), np.array(ts.bdfB, dtype=np.float64)))
    return out


def cd_objective(states, mV: Mesh, dt):
    """J = sum_{n>=1} dt/2 sum_local B (u^n)^2: the time-integrated energy of
    the trajectory, the objective the convection-diffusion adjoint differentiates."""
    return sum(0.5 * dt * float(np.sum(mV.B * u * u)) for u in states[1:])


def cd_adjoint(mV: Mesh, mD: Mesh, u0, vx, vy, nu, dt, nsteps, M, k=2, tol_rel=1e-12,
               precond="jacobi"):
    """Discrete adjoint of cd_forward for cd_objective: returns (J, dJ/du0) with
    dJ/du0 per local node of u0 (u0 enters only through B u0 and advect(u0),
    so every local copy has its own derivative).

    Step n solves H_n u^n = M QQ'(sum_i L_{n,i} uh_i), L_{n,i} = -b_{1+i} B
    - a_i C, C = advect. Reverse sweep: lam^n = H_n^{-1} M QQ'(dt B u^n +
    acc_n), then acc_{n-1-i} += L_{n,i}' lam^n with C' the transpose of
    advect; at level 0 the sum is the gradient with respect to u0."""
    Jr, Js = cd_step_matrices(mV, mD)
    _, states = cd_forward(mV, mD, u0, vx, vy, nu, dt, nsteps, M, k=k, tol_rel=tol_rel,
                           keep=True, precond=precond)
    coefs = _cd_coefficients(dt, nsteps, k)
    J = cd_objective(states, mV, dt)
    acc = {n: np.zeros_like(u0) for n in range(nsteps + 1)}
    for n in range(nsteps, 0, -1):
        a, b = coefs[n - 1]
        mu = mV.gs(mask(dt * mV.B * states[n] + acc[n], M))
        opL = lambda v, b0=b[0]: mask(mV.gs(hlmz(v, nu, b0, mV)), M)
        pre = _cd_precond(mV, nu, b[0], M, precond)
        lam = pcg(mu, opL, opM=lambda r: pre * r, mult=mV.mult,
                  tol=tol_rel * max(np.abs(mu).max(), 1e-300))
        Ct = advect_transpose(lam, vx, vy, mV, mD, Jr, Js)
        for i in range(k):
            m = n - 1 - i
            if m < 0:
                continue
            if b[1 + i] != 0.0 or a[i] != 0.0:
                acc[m] = acc[m] - b[1 + i] * (mV.B * lam) + a[i] * (-Ct)
    return J, acc[0]


# --------------------------------------------------------------------------
# the Diffusion equation (diffusion.jl)
# --------------------------------------------------------------------------
def diffusion_simulate(msh: Mesh, bc, Ti=0.0, Tf=0.0, dt=0.0, k=3, setIC=None, setBC=None,
                       setForcing=None, setVisc=None, nsteps=None, tol=1e-8, maxiter=None,
                       callback=None):
    """simulate!(Diffusion(bc, msh; Ti, Tf, dt, k), callback!, setIC!, setBC!,
    setForcing!, setVisc!) (diffusion.jl:20-34, :110-137).

    Setters have the reference signature setter(array, x, y, t) and write in
    place; None leaves the array as it is (fixU!). Each step (evolve!,
    diffusion.jl:81-106): histories shift (mesh.jl:199-215), the stepper
    advances (time.jl:101-112), ub, f and nu are set at the new time, rhs =
    QQ' mask(B f - nu lapl(ub) - sum_i b_{1+i} B uh_i) (:51-65), u = pcg(rhs,
    mask(QQ' hlmz(., nu, b_1))) with the identity preconditioner, mult
    weights and the absolute tolerance tol (:36-49, :67-77, pcg.jl:16-60),
    then u += ub. The loop is the reference's ``while time <= Tf`` with its
    break after a steady step (time < 1e-12), or ``nsteps`` steps by count.
    The field keeps k histories (the reference's Field always keeps 3, which
    equals k for its default k = 3). Returns (u, per-step PCG iteration
    counts); callback(u, t, istep) runs after the initial condition and after
    every step."""
    M = generateMask(bc, msh)
    ts = TimeStepper(Ti, Tf, dt, k=k)
    x, y = msh.x, msh.y
    u = np.zeros_like(x)
    uh = [np.zeros_like(x) for _ in range(k)]
    ub = np.zeros_like(x)
    f = np.zeros_like(x)
    nu = np.zeros_like(x)
    iters = []
    if setIC is not None:
        setIC(u, x, y, ts.time[0])
    if callback is not None:
        callback(u, ts.time[0], ts.istep)

    def evolve():
        nonlocal u
        for i in range(k - 1, 0, -1):
            uh[i] = uh[i - 1]
        uh[0] = u.copy()
        ts.update()
        t = ts.time[0]
        for setter, arr in ((setBC, ub), (setForcing, f), (setVisc, nu)):
            if setter is not None:
                setter(arr, x, y, t)
        a, b = ts.bdfA, ts.bdfB
        rhs = mass(f, msh)
        rhs = rhs - nu * lapl(ub, msh)
        for i in range(k):
            rhs = rhs - b[1 + i] * mass(uh[i], msh)
        rhs = msh.gs(mask(rhs, M))
        info = PCGInfo()
        u = pcg(rhs, lambda v: mask(msh.gs(hlmz(v, nu, b[0], msh)), M), mult=msh.mult,
                tol=tol, maxiter=maxiter, info=info)
        u = u + ub
        iters.append(info.iters)
        if callback is not None:
            callback(u, ts.time[0], ts.istep)

    if nsteps is not None:
        for _ in range(int(nsteps)):
            evolve()
    else:
        while ts.time[0] <= Tf:
            evolve()
            if ts.time[0] < 1e-12:
                break
    return u, iters


# --------------------------------------------------------------------------
# mesh generation and point evaluation (geom.jl, interp.jl)
# --------------------------------------------------------------------------
def gordonHall(xrm, xrp, xsm, xsp, yrm, yrp, ysm, ysp, zr, zs):
    """Transfinite interpolation of four boundary curves (geom.jl:8-31):
    x = Jer [xrm'; xrp'] + [xsm xsp] Jes' - Jer C Jes' with Jer =
    interpMat(zr, [-1, 1]), Jes = interpMat(zs, [-1, 1]) and C the corner
    values C[a, b] = x(r = ze[a], s = ze[b]), taken from the r-edges. (The
    reference lays C out as C[b, a] (geom.jl:16-23), the transpose; this is
    the interpolant the construction intends, and the two agree whenever
    x(1, -1) = x(-1, 1).) Pinned in tests/test_oracle_geom.py by exact
    reproduction of bilinear maps, interpolation of the edge data and a
    closed-form Coons patch."""
    ze = np.array([-1.0, 1.0])
    Jer = interpMat(zr, ze)
    Jes = interpMat(zs, ze)
    out = []
    for rm, rp, sm, sp in ((xrm, xrp, xsm, xsp), (yrm, yrp, ysm, ysp)):
        rm, rp, sm, sp = (np.asarray(v, dtype=np.float64).reshape(-1) for v in (rm, rp, sm, sp))
        C = np.array([[rm[0], rm[-1]], [rp[0], rp[-1]]])
        er = Jer @ np.vstack([rm, rp])
        es = np.column_stack([sm, sp]) @ Jes.T
        out.append(er + es - Jer @ C @ Jes.T)
    return out[0], out[1]


def semInterp(u, xp, yp, rxe, rye, sxe, sye, zr, zs, xe, ye, tol=1e-8, maxiter=50):
    """Value of the element field u at the physical point (xp, yp)
    (interp.jl:47-102): Newton on the element map from the closest GLL node,
    r += rx dx + ry dy, s += sx dx + sy dy with the inverse Jacobian
    interpolated at (r, s), until |(dx, dy)| < tol. (The reference's update
    reads rx dx + ry dx and sx dy + sy dy, interp.jl:80-81, which is not
    Newton's step for a general map; this is the step it intends.) Returns
    NaN when a step leaves [-1, 1]^2; after maxiter steps it returns the value
    at the last iterate, as the reference does. Pinned in
    tests/test_oracle_geom.py by recovering known (r, s) points of deformed
    elements and exact values on affine elements."""
    d2 = (xe - xp) ** 2 + (ye - yp) ** 2
    ia, ib = np.unravel_index(np.argmin(d2), d2.shape)
    r, s = zr[ia], zs[ib]
    i = 0
    while True:
        Jr = interpMat(r, zr)
        Js = interpMat(s, zs)
        if i == maxiter:
            break
        x = float(ABu(Js, Jr, xe)[0, 0])
        y = float(ABu(Js, Jr, ye)[0, 0])
        dx, dy = xp - x, yp - y
        if math.sqrt(dx * dx + dy * dy) < tol:
            break
        rx = float(ABu(Js, Jr, rxe)[0, 0])
        ry = float(ABu(Js, Jr, rye)[0, 0])
        sx = float(ABu(Js, Jr, sxe)[0, 0])
        sy = float(ABu(Js, Jr, sye)[0, 0])
        dr = rx * dx + ry * dy
        ds = sx * dx + sy * dy
        if abs(r + dr) > 1.0 or abs(s + ds) > 1.0:
            return float("nan")
        r += dr
        s += ds
        i += 1
    return float(ABu(Js, Jr, u)[0, 0])
