"""Python mirror of SpectralElements.jl's operator-apply-and-solve API over
libsemhip.so. Names, argument order and meaning follow the reference:

    Mesh(nr, ns, Ex, Ey, ifperiodic, deform)          mesh.jl:66-133
    generateMask(bc, msh)                             mesh.jl:149-175
    ABu(As, Br, u)                                    ABu.jl:9-37
    lapl(u, msh) / lapl(u, nu, msh)                   lapl.jl:26-45
    mass(u, msh)                                      mass.jl:12-22
    hlmz(u, nu, k, msh)                               hlmz.jl:12-19
    gatherScatter(u, msh) / (u, QQtx, QQty) / (u, l2g) gatherScatter.jl:8-34
    mask(u, M)                                        mask.jl:10-18
    pcg(b, opA; opM, mult, ifv, tol, maxiter)         pcg.jl:16-60
    pcg_(x, b, opA; ...)  (pcg!)                      pcg.jl:64-79
    grad(u, msh), gradT(...)                          grad.jl:15-63
    advect(T, vx, vy, mshV, mshD)                     advect.jl:45-78

Fields are float64 CUDA (ROCm) tensors of shape (nr*Ex, ns*Ey) stored
column-major (strides (1, nr*Ex)), i.e. the reference's memory layout; NumPy
arrays are accepted too and round-trip through the device. Every operator
returns a new array, as the reference does; the ``out=`` argument writes in
place. Operators recognised by ``pcg`` (an ``SEMOperator`` with an identity or
``Diagonal`` preconditioner) run the fused GPU PCG; any other closure runs
the same recursion with device vector algebra.
"""
from __future__ import annotations

import ctypes as C
import weakref
from typing import Optional, Sequence

import numpy as np
import torch

from . import _lib as L

_F64 = torch.float64


# ------------------------------------------------------------------ context
class _Ctx:
    def __init__(self, device: int):
        self.device = device
        h = C.c_void_p()
        L.call("sem_ctx_create", device, None, C.byref(h))
        self.h = h

    def __del__(self):
        try:
            L.load().sem_ctx_destroy(self.h)
        except Exception:
            pass


_ctxs: dict = {}


def ctx() -> C.c_void_p:
    """The library context of the current device, bound to torch's current stream."""
    if not torch.cuda.is_available():
        raise RuntimeError("semhip needs a ROCm GPU (torch.cuda.is_available() is False)")
    dev = torch.cuda.current_device()
    c = _ctxs.get(dev)
    if c is None:
        L.load()
        c = _ctxs[dev] = _Ctx(dev)
    s = torch.cuda.current_stream(dev).cuda_stream
    L.call("sem_ctx_set_stream", c.h, C.c_void_p(s))
    return c.h


def synchronize():
    L.call("sem_ctx_synchronize", ctx())


def _p(t) -> C.c_void_p:
    return C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p(None)


# ------------------------------------------------------------ field layout
def empty_field(L_: int, NC: int, dtype=_F64, device=None) -> torch.Tensor:
    """Uninitialised (L_, NC) column-major device field."""
    device = device or torch.device("cuda", torch.cuda.current_device())
    return torch.empty((NC, L_), dtype=dtype, device=device).t()


def _is_colmajor(t: torch.Tensor) -> bool:
    return t.dim() == 2 and t.stride(0) == 1 and (t.stride(1) == t.shape[0] or t.shape[1] == 1)


def to_device(u, dtype=_F64) -> torch.Tensor:
    """A column-major CUDA copy/view of u (numpy or torch, 2-D)."""
    if isinstance(u, np.ndarray):
        a = np.asfortranarray(u, dtype=np.float64 if dtype == _F64 else np.float32)
        t = empty_field(a.shape[0], a.shape[1], dtype=dtype)
        t.copy_(torch.from_numpy(a))
        return t
    if not isinstance(u, torch.Tensor):
        raise TypeError(f"expected a numpy array or torch tensor, got {type(u)}")
    if u.dim() != 2:
        raise ValueError(f"fields are 2-D (nr*Ex, ns*Ey) arrays, got shape {tuple(u.shape)}")
    if u.device.type != "cuda" or u.dtype != dtype or not _is_colmajor(u):
        t = empty_field(u.shape[0], u.shape[1], dtype=dtype)
        t.copy_(u)
        return t
    return u


def to_host(t: torch.Tensor) -> np.ndarray:
    return np.asfortranarray(t.detach().cpu().numpy())


def _like(u: torch.Tensor, dtype=None) -> torch.Tensor:
    return empty_field(u.shape[0], u.shape[1], dtype=dtype or u.dtype, device=u.device)


def _wrap_out(res: torch.Tensor, was_numpy: bool):
    return to_host(res) if was_numpy else res


def _hip_copy(dst: int, src: int, nbytes: int):
    """Device-to-device copy of library-owned memory into a torch buffer."""
    global _hiprt
    try:
        _hiprt
    except NameError:
        _hiprt = C.CDLL("libamdhip64.so")
        _hiprt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hiprt.hipMemcpy.restype = C.c_int
    torch.cuda.current_stream().synchronize()
    r = _hiprt.hipMemcpy(C.c_void_p(dst), C.c_void_p(src), C.c_size_t(nbytes), 3)
    if r != 0:
        raise RuntimeError(f"hipMemcpy failed with {r}")


# ------------------------------------------------------------ L0 primitives
def gausslobatto(n: int):
    """GLL nodes and weights (FastGaussQuadrature.gausslobatto, mesh.jl:70-71)."""
    z = np.zeros(n)
    w = np.zeros(n)
    L.call("sem_gausslobatto", n, z.ctypes.data_as(L._dp), w.ctypes.data_as(L._dp))
    return z, w


def derivMat(x) -> np.ndarray:
    """derivMat(x), derivMat.jl:9-35."""
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
    D = np.zeros((x.size, x.size), order="F")
    L.call("sem_derivmat", x.size, x.ctypes.data_as(L._dp), D.ctypes.data_as(L._dp))
    return D


def interpMat(xo, xi) -> np.ndarray:
    """interpMat(xo, xi), interp.jl:10-35."""
    xo = np.ascontiguousarray(np.atleast_1d(xo), dtype=np.float64).reshape(-1)
    xi = np.ascontiguousarray(np.atleast_1d(xi), dtype=np.float64).reshape(-1)
    J = np.zeros((xo.size, xi.size), order="F")
    L.call("sem_interpmat", xo.size, xo.ctypes.data_as(L._dp), xi.size,
           xi.ctypes.data_as(L._dp), J.ctypes.data_as(L._dp))
    return J


def semmesh(E: int, n: int):
    """semmesh(E, n), semmesh.jl:9-27."""
    z = np.zeros(E * n)
    w = np.zeros(E * n)
    L.call("sem_semmesh", E, n, z.ctypes.data_as(L._dp), w.ctypes.data_as(L._dp))
    return z, w


def ndgrid(v1, v2):
    """ndgrid.jl:8-13 on the device."""
    v1 = torch.as_tensor(np.asarray(v1, dtype=np.float64)).cuda()
    v2 = torch.as_tensor(np.asarray(v2, dtype=np.float64)).cuda()
    m, n = v1.numel(), v2.numel()
    x = empty_field(m, n)
    y = empty_field(m, n)
    x.copy_(v1.reshape(m, 1).expand(m, n))
    y.copy_(v2.reshape(1, n).expand(m, n))
    return x, y


# ------------------------------------------------------------------ deform
def fixU(*u):
    """Identity deformation (mesh.jl:6-8)."""
    return u


class deform_sinsin:
    """x' = x + amp sin(pi x) sin(pi y), y' = y + amp sin(pi x) sin(pi y),
    evaluated on the GPU during mesh setup."""

    kind = L.SEM_DEFORM_SINSIN

    def __init__(self, amp: float = 0.1):
        self.amp = float(amp)

    def __call__(self, x, y):
        s = self.amp * torch.sin(np.pi * x) * torch.sin(np.pi * y)
        return x + s, y + s


class _Affine01:
    """usrgeom of p2d_explicit.jl:78 / semPS.jl:306-310: [-1,1]^2 -> [0,1]^2."""

    kind = L.SEM_DEFORM_AFFINE01
    amp = 0.0

    def __call__(self, x, y):
        return 0.5 * (x + 1.0), 0.5 * (y + 1.0)


deform_affine01 = _Affine01()


class annulus:
    """annulus(r, s; r0, r1, span) (geom.jl:40-50): R = (r1 - r0)/2 (r + 1) +
    r0, th = span/2 (s + 1), (x, y) = (R cos th, R sin th). A user
    deformation: evaluated in FP64 on the device, then handed to the mesh."""

    def __init__(self, r0: float = 0.5, r1: float = 1.0, span: float = 2.0 * np.pi):
        self.r0, self.r1, self.span = float(r0), float(r1), float(span)

    def __call__(self, x, y):
        R = (self.r1 - self.r0) / 2 * (x + 1) + self.r0
        th = self.span / 2 * (y + 1) + 0.0
        return R * torch.cos(th), R * torch.sin(th)


# -------------------------------------------------------------------- mesh
class Mesh:
    """Mesh(nr, ns, Ex, Ey, ifperiodic=[false,false], deform=fixU), mesh.jl:66-133.

    Geometric factors live on the GPU (computed there in FP64). Attributes
    mirror struct Mesh (mesh.jl:25-64): zr, zs, wr, ws, Dr, Ds are host
    arrays; x, y, Jac, Jaci, rx, ry, sx, sy, B, Bi, G11, G12, G22, mult are
    device fields (copied out of the library on first access); l2g is the
    local-to-global numbering (int32, 0-based). The dense QQtx/QQty of the
    reference are not formed: gather-scatter uses the sorted map.
    """

    def __init__(self, nr: int, ns: int, Ex: int, Ey: int,
                 ifperiodic: Sequence[bool] = (False, False), deform=fixU):
        self.nr, self.ns, self.Ex, self.Ey = int(nr), int(ns), int(Ex), int(Ey)
        self.ifperiodic = [bool(ifperiodic[0]), bool(ifperiodic[1])]
        self.deform = deform
        h = C.c_void_p()
        kind = getattr(deform, "kind", None)
        user = kind is None and deform is not fixU and deform is not None
        kind = L.SEM_DEFORM_NONE if kind is None else kind
        amp = float(getattr(deform, "amp", 0.0))
        L.call("sem_mesh_create", ctx(), self.nr, self.ns, self.Ex, self.Ey,
               int(self.ifperiodic[0]), int(self.ifperiodic[1]), kind, amp, C.byref(h))
        self.h = h
        self._fin = weakref.finalize(self, L.load().sem_mesh_destroy, h)
        nl, ng = C.c_int64(), C.c_int64()
        L.call("sem_mesh_sizes", h, None, None, None, None, C.byref(nl), C.byref(ng))
        self.nlocal, self.nglobal = nl.value, ng.value
        self.shape = (self.nr * self.Ex, self.ns * self.Ey)
        self.zr, self.wr = gausslobatto(self.nr)
        self.zs, self.ws = gausslobatto(self.ns)
        self.Dr, self.Ds = derivMat(self.zr), derivMat(self.zs)
        self._cache = {}
        if user:
            x, y = deform(self.field("x").clone(), self.field("y").clone())
            x, y = to_device(x), to_device(y)
            L.call("sem_mesh_set_coords", ctx(), h, _p(x), _p(y))
            self._cache.clear()

    # device pointer of a library-owned mesh array
    def _ptr(self, name: str) -> int:
        p = C.c_void_p()
        L.call("sem_mesh_field", self.h, L.FIELD_IDS[name], C.byref(p))
        return p.value

    def field(self, name: str) -> torch.Tensor:
        t = self._cache.get(name)
        if t is None:
            t = empty_field(*self.shape)
            _hip_copy(t.data_ptr(), self._ptr(name), 8 * self.nlocal)
            self._cache[name] = t
        return t

    def __getattr__(self, name):
        if name in L.FIELD_IDS:
            return self.field(name)
        if name == "l2g":
            t = self._cache.get("l2g")
            if t is None:
                p = C.c_void_p()
                L.call("sem_mesh_l2g", self.h, C.byref(p))
                t = torch.empty((self.shape[1], self.shape[0]), dtype=torch.int32,
                                device="cuda").t()
                _hip_copy(t.data_ptr(), p.value, 4 * self.nlocal)
                self._cache["l2g"] = t
            return t
        raise AttributeError(name)

    def zeros(self, dtype=_F64) -> torch.Tensor:
        t = empty_field(*self.shape, dtype=dtype)
        t.zero_()
        return t


# -------------------------------------------------------------------- mask
def generateMask(bc, msh: Mesh) -> torch.Tensor:
    """Dirichlet mask M (float64 0/1 field), mesh.jl:149-175. The returned
    tensor remembers its structure so that mask() and SEMOperator use the
    structural form."""
    bc4 = "".join(bc).encode()
    bits = C.c_int()
    L.call("sem_mask_bits", msh.h, bc4, C.byref(bits))
    M = empty_field(*msh.shape)
    L.call("sem_mask_fill", ctx(), msh.h, bits.value, _p(M))
    M._sem_bits = bits.value
    M._sem_mesh = msh
    return M


def mask(u, M, out=None):
    """M .* u, or a copy when M is empty (mask.jl:10-18)."""
    was_np = isinstance(u, np.ndarray)
    u = to_device(u)
    y = out if out is not None else _like(u)
    if M is None or (hasattr(M, "numel") and M.numel() == 0) or (isinstance(M, (list, tuple)) and len(M) == 0):
        L.call("sem_mask", ctx(), u.numel(), None, _p(u), _p(y))
    elif getattr(M, "_sem_bits", None) is not None and M._sem_mesh.shape == tuple(u.shape):
        L.call("sem_mask_bc", ctx(), M._sem_mesh.h, M._sem_bits, _p(u), _p(y))
    else:
        Md = to_device(M)
        L.call("sem_mask", ctx(), u.numel(), _p(Md), _p(u), _p(y))
    return _wrap_out(y, was_np)


# --------------------------------------------------------------- operators
def _dev_mat(A):
    if A is None or (hasattr(A, "__len__") and len(A) == 0):
        return None, 0, 0
    a = np.asfortranarray(np.asarray(A.cpu() if isinstance(A, torch.Tensor) else A,
                                     dtype=np.float64))
    if a.ndim == 1:
        a = a.reshape(1, -1)
    t = torch.from_numpy(np.ascontiguousarray(a.T)).cuda()  # column-major storage
    return t, a.shape[0], a.shape[1]


def ABu(As, Br, u, trans: bool = False):
    """(As kron Br) * u block-wise (ABu.jl:9-37); [] / None is the identity.
    trans=True applies (As' kron Br'), the pullback of ABu.jl:39-41."""
    was_np = isinstance(u, np.ndarray)
    u = to_device(u)
    m, n = u.shape
    At, ma, na = _dev_mat(As)
    Bt, mb, nb = _dev_mat(Br)
    rmb, rnb = (nb, mb) if trans else (mb, nb)
    rma, rna = (na, ma) if trans else (ma, na)
    m1 = (m // rnb) * rmb if Bt is not None else m
    n1 = (n // rna) * rma if At is not None else n
    y = empty_field(m1, n1)
    L.call("sem_abu", ctx(), m, n, _p(u), _p(At), ma, na, _p(Bt), mb, nb, int(trans), _p(y))
    return _wrap_out(y, was_np)


def lapl(u, *args, out=None):
    """lapl(u, msh) (lapl.jl:26-36); lapl(u, nu, msh) = nu .* lapl(u, msh)
    (:38-45); lapl(u, msh1, msh2), whose dealiasing is a placeholder in the
    reference and forwards to lapl(u, msh1) (:47-52); and the 12-argument
    lapl(u, M, Jr, Js, QQtx, QQty, Dr, Ds, G11, G12, G22, mult) with
    gather-scatter and mask (:54-68; Jr = Js = [] is the plain laplace, else
    the dealiased laplace of :83-103)."""
    if len(args) == 11:
        return _lapl12(u, *args, out=out)
    if len(args) == 1:
        nu, msh = None, args[0]
    elif len(args) == 2 and isinstance(args[0], Mesh) and isinstance(args[1], Mesh):
        nu, msh = None, args[0]
    elif len(args) == 2:
        nu, msh = args
    else:
        raise TypeError("lapl(u, msh), lapl(u, nu, msh), lapl(u, msh1, msh2) or the "
                        "12-argument form")
    if nu is None:
        return hlmz(u, 1.0, 0.0, msh, out=out)
    return hlmz(u, nu, 0.0, msh, out=out)


def _is_empty(A) -> bool:
    return A is None or (hasattr(A, "__len__") and len(A) == 0) or (
        hasattr(A, "numel") and A.numel() == 0) or (hasattr(A, "size") and not callable(A.size)
                                                    and A.size == 0)


def gprod(G11, G12, G22, a, b):
    """(G11 .* a + G12 .* b, G12 .* a + G22 .* b) (lapl.jl:75-76)."""
    a, b = to_device(a), to_device(b)
    G = [to_device(g) for g in (G11, G12, G22)]
    wr, ws = _like(a), _like(a)
    L.call("sem_gprod", ctx(), a.numel(), _p(G[0]), _p(G[1]), _p(G[2]), _p(a), _p(b), _p(wr),
           _p(ws))
    return wr, ws


def axpby(alpha, x, beta, y, out=None):
    """alpha .* x .+ beta .* y."""
    x, y = to_device(x), to_device(y)
    o = out if out is not None else _like(x)
    L.call("sem_axpby", ctx(), x.numel(), float(alpha), _p(x), float(beta), _p(y), _p(o))
    return o


def _lapl12(u, M, Jr, Js, QQtx, QQty, Dr, Ds, G11, G12, G22, mult, out=None):
    """lapl.jl:54-68 from library calls: ABu derivatives (and interpolation),
    the pointwise G product, the transposed contractions, QQ' through the map
    built from the 0/1 matrices, and the mask. (The Zygote hook multiplying the
    cotangent by mult is the identity in the forward pass.)"""
    was_np = isinstance(u, np.ndarray)
    u = to_device(u)
    ur = ABu(None, Dr, u)
    us = ABu(Ds, None, u)
    deal = not (_is_empty(Jr) and _is_empty(Js))
    if deal:
        ur, us = ABu(Js, Jr, ur), ABu(Js, Jr, us)
    vr, vs = gprod(G11, G12, G22, ur, us)
    if deal:
        vr, vs = ABu(Js, Jr, vr, trans=True), ABu(Js, Jr, vs, trans=True)
    Au = axpby(1.0, ABu(None, Dr, vr, trans=True), 1.0, ABu(Ds, None, vs, trans=True))
    Au = gatherScatter(Au, QQtx, QQty)
    y = mask(Au, M, out=out)
    return _wrap_out(y, was_np)


def mass(u, *args, out=None):
    """B .* u: mass(u, msh) (mass.jl:12-22); mass(u, msh1, msh2), which
    forwards to mass(u, msh1) in the reference (:25-30); and the 8-argument
    mass(u, M, B, Jr, Js, QQtx, QQty, mult) = mask(QQ'(Js'Jr'(B .* Js Jr u)))
    (:32-50; B = [] means no weight)."""
    if len(args) == 7:
        return _mass8(u, *args, out=out)
    if len(args) not in (1, 2) or not all(isinstance(a, Mesh) for a in args):
        raise TypeError("mass(u, msh), mass(u, msh1, msh2) or the 8-argument form")
    msh = args[0]
    was_np = isinstance(u, np.ndarray)
    u = to_device(u)
    y = out if out is not None else _like(u)
    L.call("sem_mass", ctx(), msh.h, _p(u), _p(y))
    return _wrap_out(y, was_np)


def _mass8(u, M, B, Jr, Js, QQtx, QQty, mult, out=None):
    was_np = isinstance(u, np.ndarray)
    Ju = ABu(Js, Jr, to_device(u))
    BJu = Ju if _is_empty(B) else mask(Ju, to_device(B))
    Bu = ABu(Js, Jr, BJu, trans=True)
    Bu = gatherScatter(Bu, QQtx, QQty)
    y = mask(Bu, M, out=out)
    return _wrap_out(y, was_np)


def _scalar_or_field(v):
    if isinstance(v, (int, float, np.floating, np.integer)):
        return float(v), None
    if isinstance(v, torch.Tensor) and v.numel() == 1:
        return float(v.item()), None
    return 0.0, to_device(v)


def hlmz(u, nu, k, msh: Mesh, msh2: Optional[Mesh] = None, out=None):
    """nu .* lapl(u) + k .* mass(u) (hlmz.jl:12-19); nu, k scalars or fields.
    hlmz(u, nu, k, msh1, msh2) forwards to the first mesh, as the reference's
    placeholder for dealiasing does (hlmz.jl:22-30)."""
    was_np = isinstance(u, np.ndarray)
    u = to_device(u)
    y = out if out is not None else _like(u)
    nus, nuf = _scalar_or_field(nu)
    ks, kf = _scalar_or_field(k)
    L.call("sem_hlmz", ctx(), msh.h, _p(u), nus, _p(nuf), ks, _p(kf), _p(y))
    return _wrap_out(y, was_np)


class GSMap:
    """A gather-scatter map built from a local->global table (0-based)."""

    def __init__(self, l2g, nglobal: Optional[int] = None):
        a = np.asfortranarray(np.asarray(l2g.cpu() if isinstance(l2g, torch.Tensor) else l2g))
        flat = a.reshape(-1, order="F").astype(np.int32)
        ng = int(nglobal if nglobal is not None else flat.max() + 1)
        h = C.c_void_p()
        L.call("sem_gsmap_create", ctx(), flat.size, flat.ctypes.data_as(C.c_void_p), ng,
               C.byref(h))
        self.h = h
        self.nlocal, self.nglobal = flat.size, ng
        self._fin = weakref.finalize(self, L.load().sem_gsmap_destroy, h)

    def __call__(self, u, out=None):
        u = to_device(u)
        y = out if out is not None else _like(u)
        L.call("sem_gsmap_apply", ctx(), self.h, _p(u), _p(y))
        return y


_qq_maps: dict = {}


def _map_from_dense(QQtx, QQty) -> GSMap:
    """Connectivity of dense 0/1 QQ' factors (as stored by mesh.jl:81-82)."""
    key = (id(QQtx), id(QQty))
    mp = _qq_maps.get(key)
    if mp is not None:
        return mp[0]

    def comp(Q):
        Q = np.asarray(Q.cpu() if isinstance(Q, torch.Tensor) else Q)
        n = Q.shape[0]
        lab = -np.ones(n, dtype=np.int64)
        g = 0
        for i in range(n):
            if lab[i] < 0:
                idx = np.nonzero(Q[i])[0]
                lab[idx] = g
                g += 1
        return lab, g
    gx, nx = comp(QQtx)
    gy, ny = comp(QQty)
    l2g = gx[:, None] + nx * gy[None, :]
    mp = GSMap(l2g, nx * ny)
    _qq_maps[key] = (mp, QQtx, QQty)
    return mp


def gatherScatter(u, *args, out=None):
    """QQ' u: gatherScatter(u, msh) (gatherScatter.jl:18-21),
    gatherScatter(u, QQtx, QQty) (:8-16) or gatherScatter(u, l2g) (:27-34)."""
    was_np = isinstance(u, np.ndarray)
    u = to_device(u)
    y = out if out is not None else _like(u)
    if len(args) == 1 and isinstance(args[0], Mesh):
        L.call("sem_gather_scatter", ctx(), args[0].h, _p(u), _p(y))
    elif len(args) == 1 and isinstance(args[0], GSMap):
        args[0](u, out=y)
    elif len(args) == 1:
        GSMap(args[0])(u, out=y)
    elif len(args) == 2:
        _map_from_dense(*args)(u, out=y)
    else:
        raise TypeError("gatherScatter(u, msh), (u, QQtx, QQty) or (u, l2g)")
    return _wrap_out(y, was_np)


def grad(u, msh: Mesh):
    """(ux, uy), grad.jl:15-34."""
    was_np = isinstance(u, np.ndarray)
    u = to_device(u)
    ux, uy = _like(u), _like(u)
    L.call("sem_grad", ctx(), msh.h, _p(u), _p(ux), _p(uy))
    return (to_host(ux), to_host(uy)) if was_np else (ux, uy)


def gradT(ux, uy, msh: Mesh):
    """Adjoint of grad as a map L -> 2L: Dr'(rx ux + ry uy) + (sx ux + sy uy) Ds."""
    was_np = isinstance(ux, np.ndarray)
    ux, uy = to_device(ux), to_device(uy)
    y = _like(ux)
    L.call("sem_grad_t", ctx(), msh.h, _p(ux), _p(uy), _p(y))
    return _wrap_out(y, was_np)


def advect(T, vx, vy, mshV: Mesh, mshD: Mesh, transpose: bool = False):
    """Dealiased (vx, vy) . grad T weighted by B_D, advect.jl:45-78; with
    transpose=True the transpose of that linear map in T."""
    was_np = isinstance(T, np.ndarray)
    T, vx, vy = to_device(T), to_device(vx), to_device(vy)
    y = _like(T)
    L.call("sem_advect_t" if transpose else "sem_advect", ctx(), mshV.h, mshD.h, _p(T), _p(vx),
           _p(vy), _p(y))
    return _wrap_out(y, was_np)


def adjoint_dnu(u, lam, msh: Mesh):
    """B .* (grad lam . grad u): d(lam' A(nu) u)/d nu_local for A(nu) with
    G_ij <- nu G_ij (semPS.jl:161-165)."""
    u, lam = to_device(u), to_device(lam)
    g = _like(u)
    L.call("sem_adjoint_dnu", ctx(), msh.h, _p(u), _p(lam), _p(g))
    return g


def wdot(a, b, msh: Mesh) -> float:
    """sum(a .* b .* mult) (pcg.jl:45)."""
    a, b = to_device(a), to_device(b)
    r = C.c_double()
    L.call("sem_wdot", ctx(), msh.h, _p(a), _p(b), C.byref(r))
    return r.value


def dot_acc(acc: torch.Tensor, a, b, w=None, scale: float = 1.0) -> torch.Tensor:
    """acc += scale * sum(a .* b .* w) on the device (acc: a one-element FP64
    device tensor), without a host read: sums accumulated over a time loop."""
    a, b = to_device(a), to_device(b)
    wd = to_device(w) if w is not None else None
    L.call("sem_dot_acc", ctx(), a.numel(), _p(a), _p(b), _p(wd), float(scale), _p(acc))
    return acc


def dot(a, b, w=None) -> float:
    """sum(a .* b .* w) (w None: sum(a .* b)) over all local nodes."""
    a, b = to_device(a), to_device(b)
    wd = to_device(w) if w is not None else None
    r = C.c_double()
    L.call("sem_dot", ctx(), a.numel(), _p(a), _p(b), _p(wd), C.byref(r))
    return r.value


def combine(msh: Mesh, terms, M=None, assemble: bool = False, out=None):
    """[gatherScatter](mask(sum_j c_j (B .* if weighted_j) x_j, M)) for
    terms = [(c_j, x_j, weighted_j), ...] (at most 8): the composition of
    makeRHS! (convectionDiffusion.jl:93-110, diffusion.jl:51-65). M must come
    from generateMask (or be None: no mask). ``out`` may be one of the x_j."""
    if len(terms) > 8:
        raise ValueError("combine: at most 8 terms")
    xs = [to_device(x) for _, x, _ in terms]
    y = out if out is not None else empty_field(*msh.shape)
    n = len(terms)
    coef = (C.c_double * max(n, 1))(*[float(c) for c, _, _ in terms])
    ptrs = (C.c_void_p * max(n, 1))(*[x.data_ptr() for x in xs])
    wts = (C.c_int * max(n, 1))(*[int(bool(w)) for _, _, w in terms])
    bits = -1
    if M is not None:
        bits = getattr(M, "_sem_bits", None)
        if bits is None:
            raise ValueError("combine needs a mask from generateMask")
    L.call("sem_combine", ctx(), msh.h, n, coef, ptrs, wts, int(bits), int(bool(assemble)),
           _p(y))
    return y


def vardiff_factors(nu, msh: Mesh):
    """(nu .* G11, nu .* G12, nu .* G22): the geometric factors of the
    variable-coefficient operator -div(nu grad u) (semPS.jl:161-163)."""
    nu = to_device(nu)
    g = [empty_field(*msh.shape) for _ in range(3)]
    L.call("sem_vardiff_factors", ctx(), msh.h, _p(nu), _p(g[0]), _p(g[1]), _p(g[2]))
    return tuple(g)


def norm_inf(a) -> float:
    a = to_device(a)
    r = C.c_double()
    L.call("sem_norm_inf", ctx(), a.numel(), _p(a), C.byref(r))
    return r.value


# --------------------------------------------------------- fused operator
_PREC = {"f64": L.SEM_F64, "f32": L.SEM_F32, "f16": L.SEM_F16_MFMA, "bf16": L.SEM_BF16_MFMA,
         "f16x3": 4, "bf16x3": 5}


class SEMOperator:
    """v -> mask(gatherScatter(hlmz(v, nu, k, msh)), M): the Diffusion LHS
    (diffusion.jl:36-45) as one fused device operator. G = (G11, G12, G22)
    overrides the mesh factors (the 12-argument lapl of lapl.jl:54-68, e.g.
    with nu folded in as semPS.jl:161-165). precision: 'f64', 'f32', 'f16' or
    'bf16' (the last two: FP16/BF16-input MFMA contractions, nr == 16).
    assemble=False gives the element-local hlmz."""

    def __init__(self, nu, k, msh: Mesh, M=None, precision: str = "f64", G=None, B=None,
                 assemble: bool = True):
        self.msh, self.M = msh, M
        self.precision = precision
        nus, nuf = _scalar_or_field(nu)
        ks, kf = _scalar_or_field(k)
        self._keep = [t for t in (nuf, kf) if t is not None]
        d = L.sem_op_desc()
        d.nu, d.k = nus, ks
        d.nu_f = nuf.data_ptr() if nuf is not None else None
        d.k_f = kf.data_ptr() if kf is not None else None
        if G is not None:
            g = [to_device(t) for t in G]
            self._keep += g
            d.G11, d.G12, d.G22 = (t.data_ptr() for t in g)
        if B is not None:
            bt = to_device(B)
            self._keep.append(bt)
            d.B = bt.data_ptr()
        bits = 0
        if M is not None:
            bits = getattr(M, "_sem_bits", None)
            if bits is None:
                raise ValueError("SEMOperator needs a mask from generateMask (structured M)")
        d.bc_bits = bits
        d.assemble = 1 if assemble else 0
        self.desc = d
        h = C.c_void_p()
        L.call("sem_operator_create", ctx(), msh.h, C.byref(d), _PREC[precision], C.byref(h))
        self.h = h
        self._fin = weakref.finalize(self, L.load().sem_operator_destroy, h)
        self.dtype = _F64 if precision == "f64" else torch.float32

    def __call__(self, u, out=None):
        was_np = isinstance(u, np.ndarray)
        u = to_device(u, dtype=self.dtype)
        y = out if out is not None else _like(u)
        L.call("sem_operator_apply", ctx(), self.h, _p(u), _p(y))
        return _wrap_out(y, was_np)

    __mul__ = __call__   # `opA * u` (SpectralElements.jl:21)

    def jacobi(self) -> torch.Tensor:
        """Inverse assembled diagonal, zero on masked nodes (FP64)."""
        d = self.msh.zeros()
        L.call("sem_operator_jacobi", ctx(), self.h, _p(d))
        return d


class Diagonal:
    """Diagonal preconditioner z = d .* r (callable like the reference's opM).
    local: True when d may differ between the copies of a shared node (as
    1 ./ (B * b0) with the element-local B, convectionDiffusion.jl:87-91), so
    that z is not continuous; None: decided once per mesh by checking
    QQ'(mult .* d) == d."""

    def __init__(self, d, local: Optional[bool] = None):
        self.d = to_device(d)
        self.local = local
        self._local_on = {}

    def is_local(self, msh) -> bool:
        if self.local is not None:
            return self.local
        key = id(msh)
        if key not in self._local_on:
            d = self.d.to(_F64)
            a = gatherScatter(d * msh.mult, msh)
            scale = float(d.abs().max()) or 1.0
            self._local_on[key] = bool(float((a - d).abs().max()) > 1e-12 * scale)
        return self._local_on[key]

    def __call__(self, r):
        return self.d * r

    __mul__ = __call__


class MG:
    """Multigrid V-cycle preconditioner for an SEMOperator (a stronger opM
    than the identity / Jacobi of diffusion.jl:47-49): p-multigrid on the
    element grid, then h-multigrid on bilinear elements, Chebyshev-Jacobi
    smoothing. ``mg(r)`` applies one V-cycle to an assembled residual;
    ``pcg(b, op, opM=mg)`` runs PCG with it on the GPU (sem_pcg_mg). The
    operator must be FP64, assembled, with scalar nu and k. precision='f32'
    runs the V-cycle in FP32 (level operators, smoothing, transfers); the PCG
    stays FP64 and takes the flexible (Polak-Ribiere) beta."""

    def __init__(self, op: SEMOperator, degree: int = 3, ratio: float = 20.0,
                 degree_coarse: int = 8, npow: int = 20, safety: float = 1.1,
                 max_levels: int = 32, precision: str = "f64"):
        self.op = op
        if precision not in ("f64", "f32"):
            raise ValueError("MG precision: 'f64' or 'f32'")
        prm = L.sem_mg_params(int(degree), float(ratio), int(degree_coarse), int(npow),
                              float(safety), int(max_levels), _PREC[precision])
        h = C.c_void_p()
        L.call("sem_mg_create", ctx(), op.h, C.byref(prm), C.byref(h))
        self.h = h
        self._fin = weakref.finalize(self, L.load().sem_mg_destroy, h)

    def levels(self):
        """[(p, Ex, Ey, lmax)] from the finest level to the coarsest."""
        n = C.c_int()
        L.call("sem_mg_num_levels", self.h, C.byref(n))
        out = []
        for lv in range(n.value):
            p, ex, ey, lm = C.c_int(), C.c_int(), C.c_int(), C.c_double()
            L.call("sem_mg_level_info", self.h, lv, C.byref(p), C.byref(ex), C.byref(ey),
                   C.byref(lm))
            out.append((p.value, ex.value, ey.value, lm.value))
        return out

    def set_lmax(self, level: int, lmax: float):
        L.call("sem_mg_set_lmax", self.h, int(level), float(lmax))

    def __call__(self, r, out=None):
        was_np = isinstance(r, np.ndarray)
        r = to_device(r)
        z = out if out is not None else _like(r)
        L.call("sem_mg_vcycle", ctx(), self.h, _p(r), _p(z))
        return _wrap_out(z, was_np)

    __mul__ = __call__


class PCGInfo:
    def __init__(self):
        self.iters = 0
        self.res = 0.0
        self.converged = False
        self.elapsed_ms = 0.0
        self.outer = 0
        self.stalled = False


def _is_identity(opM) -> bool:
    return opM is None or getattr(opM, "_sem_identity", False)


def identity(v):
    return v


identity._sem_identity = True


def pcg(b, opA, opM=None, mult=None, ifv: bool = False, tol: float = 1e-8,
        maxiter: Optional[int] = None, info: Optional[PCGInfo] = None, x=None,
        rtol: Optional[float] = None):
    """Left-preconditioned CG from x0 = 0 (pcg.jl:16-60): stops when
    norm(r, Inf) <= tol, inner products weighted by mult (None: unweighted,
    the reference's mult = ones), warns and returns at k == maxiter. An
    SEMOperator with the operator mesh's mult and an identity, Diagonal or MG
    opM runs fused on the GPU; anything else runs the same recursion with
    device vector algebra. rtol (instead of tol): stop when norm(r, Inf) <=
    rtol * norm(b, Inf); the fused Jacobi/identity solve forms norm(b, Inf) on
    the device, without a host read."""
    was_np = isinstance(b, np.ndarray)
    info = info if info is not None else PCGInfo()
    fused_diag = isinstance(opA, SEMOperator) and _same_mult(mult, opA.msh) and (
        _is_identity(opM) or isinstance(opM, Diagonal))
    if rtol is not None:
        tol = -float(rtol) if fused_diag else float(rtol) * norm_inf(b)
    if isinstance(opA, SEMOperator) and isinstance(opM, MG) and opM.op is opA and (
            _same_mult(mult, opA.msh)):
        bd = to_device(b)
        if maxiter is None:
            maxiter = bd.numel()
        xd = x if x is not None else _like(bd)
        pi = L.sem_pcg_info()
        st = L.call("sem_pcg_mg", ctx(), opA.h, opM.h, _p(bd), _p(xd), float(tol),
                    int(min(maxiter, 2**31 - 1)), C.byref(pi))
        info.iters, info.res, info.converged = pi.iters, pi.res, bool(pi.converged)
        info.elapsed_ms = pi.elapsed_ms
        if st == L.SEM_WARN_MAXITER:
            L.maxiter_warning(pi.res)
        if ifv:
            print(f"PCG iter: {pi.iters}, res: {pi.res}")
        return _wrap_out(xd, was_np)
    if fused_diag:
        bd = to_device(b, dtype=opA.dtype)
        if maxiter is None:
            maxiter = bd.numel()
        xd = x if x is not None else _like(bd)
        pi = L.sem_pcg_info()
        if isinstance(opM, Diagonal):
            d = opM.d if opA.dtype == _F64 else opM.d.to(torch.float32)
            kind = L.SEM_PRECOND_DIAG_LOCAL if opM.is_local(opA.msh) else L.SEM_PRECOND_DIAG
            st = L.call("sem_pcg", ctx(), opA.h, kind, _p(d), _p(bd), _p(xd),
                        float(tol), int(min(maxiter, 2**31 - 1)), C.byref(pi))
        else:
            st = L.call("sem_pcg", ctx(), opA.h, L.SEM_PRECOND_IDENTITY, None, _p(bd), _p(xd),
                        float(tol), int(min(maxiter, 2**31 - 1)), C.byref(pi))
        info.iters, info.res, info.converged = pi.iters, pi.res, bool(pi.converged)
        info.elapsed_ms = pi.elapsed_ms
        if st == L.SEM_WARN_MAXITER:
            L.maxiter_warning(pi.res)
        if ifv:
            print(f"PCG iter: {pi.iters}, res: {pi.res}")
        return _wrap_out(xd, was_np)
    return _pcg_generic(b, opA, opM, mult, ifv, tol, maxiter, info, x, was_np)


def _same_mult(mult, msh) -> bool:
    """True when mult holds the mesh's own 1/multiplicity (the fused solvers
    weight their inner products with it structurally). None is the
    reference's default mult = ones (pcg.jl:18), which is not."""
    if mult is None:
        return False
    if mult is msh._cache.get("mult"):
        return True
    if tuple(getattr(mult, "shape", ())) != tuple(msh.shape):
        return False
    m = mult if isinstance(mult, torch.Tensor) else torch.as_tensor(np.asarray(mult))
    return bool(torch.equal(m.to(device=msh.mult.device, dtype=_F64), msh.mult))


def _pcg_generic(b, opA, opM, mult, ifv, tol, maxiter, info, x, was_np):
    """pcg.jl:16-60 with arbitrary closures; vector algebra on the device."""
    bd = to_device(b)
    opM = opM or identity
    mult = to_device(mult) if mult is not None else None
    if maxiter is None:
        maxiter = bd.numel()

    def dot(a, c):
        return float(torch.sum(a * c * mult)) if mult is not None else float(torch.sum(a * c))
    xk = torch.zeros_like(bd)
    ra = bd.clone()
    hp = rp = u = None
    k = 0
    conv = True
    while float(ra.abs().max()) > tol:
        ha = opM(ra)
        if k == maxiter:
            L.maxiter_warning(float(ra.abs().max()))
            conv = False
            break
        k += 1
        hpp, rpp = hp, rp
        hp, rp = ha, ra
        t = dot(rp, hp)
        u = hp.clone() if k == 1 else hp + (t / dot(rpp, hpp)) * u
        Au = to_device(opA(u))
        a = t / dot(u, Au)
        xk = xk + a * u
        ra = rp - a * Au
    res = float(ra.abs().max())
    info.iters, info.res, info.converged = k, res, conv and res <= tol
    if ifv:
        print(f"PCG iter: {k}, res: {float(torch.linalg.vector_norm(ra))}")
    if x is not None:
        x.copy_(xk)
        xk = x
    return _wrap_out(xk, was_np)


def pcg_(x, b, opA, **kw):
    """pcg!(x, b, opA; ...) (pcg.jl:64-79): x .= pcg(b, opA; ...)."""
    r = pcg(b, opA, **kw)
    if isinstance(x, np.ndarray):
        x[...] = r if isinstance(r, np.ndarray) else to_host(r)
    else:
        x.copy_(r)
    return None


def pcg_mixed(b, op64: SEMOperator, op_lo: SEMOperator, d=None, tol: float = 1e-10,
              inner_tol: float = 1e-2, max_outer: int = 200, max_inner: int = 1000000,
              info: Optional[PCGInfo] = None):
    """Mixed-precision solve: a low-precision fused PCG (FP32 vectors, op_lo's
    contractions) with FP64 residual replacement every time its residual
    drops by inner_tol, x accumulated in FP64, until the true FP64 residual
    norm(b - op64(x), Inf) <= tol. d is the FP64 Jacobi diagonal or None."""
    bd = to_device(b)
    x = _like(bd)
    pi = L.sem_pcg_info()
    dd = to_device(d) if d is not None else None
    st = L.call("sem_pcg_mixed", ctx(), op64.h, op_lo.h, _p(dd), _p(bd), _p(x), float(tol),
                float(inner_tol), int(max_outer), int(max_inner), C.byref(pi))
    if info is not None:
        info.iters, info.res, info.converged = pi.iters, pi.res, bool(pi.converged)
        info.elapsed_ms, info.outer = pi.elapsed_ms, pi.outer
        info.stalled = st == L.SEM_WARN_STALLED
    if st == L.SEM_WARN_MAXITER:
        L.maxiter_warning(pi.res)
    elif st == L.SEM_WARN_STALLED:
        L.stall_warning(pi.res)
    return x


# ------------------------------------------ mesh generation and evaluation
def _dev_vec(v) -> torch.Tensor:
    if isinstance(v, torch.Tensor):
        return v.detach().to(device="cuda", dtype=_F64).contiguous().reshape(-1)
    return torch.as_tensor(np.asarray(v, dtype=np.float64).reshape(-1)).cuda()


def gordonHall(xrm, xrp, xsm, xsp, yrm, yrp, ysm, ysp, zr, zs):
    """Transfinite interpolation of four boundary curves (geom.jl:8-31) on
    the GPU: xrm, xrp, yrm, yrp sample the r = -1 and r = +1 edges at the
    points zs; xsm, xsp, ysm, ysp the s = -1 and s = +1 edges at zr. Returns
    the (len(zr), len(zs)) fields x, y (column-major device arrays)."""
    zr, zs = _dev_vec(zr), _dev_vec(zs)
    e = [_dev_vec(v) for v in (xrm, xrp, xsm, xsp, yrm, yrp, ysm, ysp)]
    nr, ns = zr.numel(), zs.numel()
    for v, n in zip(e, (ns, ns, nr, nr, ns, ns, nr, nr)):
        if v.numel() != n:
            raise ValueError("gordonHall: r-edges need len(zs) values, s-edges len(zr)")
    x, y = empty_field(nr, ns), empty_field(nr, ns)
    L.call("sem_gordon_hall", ctx(), nr, ns, _p(zr), _p(zs), *[_p(v) for v in e], _p(x), _p(y))
    return x, y


class gordonHallMap:
    """A mesh deformation (the deform argument of Mesh) by transfinite
    interpolation of four boundary curves of the domain. Each curve is a
    callable t -> (x(t), y(t)) on t in [-1, 1]: rm / rp the r = -1 / +1
    sides (t = s), sm / sp the s = -1 / +1 sides (t = r). The undeformed
    node coordinates (semmesh, in [-1, 1]) are the (r, s) of the map."""

    def __init__(self, rm, rp, sm, sp):
        self.rm, self.rp, self.sm, self.sp = rm, rp, sm, sp

    def __call__(self, x, y):
        r = x[:, 0].contiguous()
        s = y[0, :].contiguous()
        xrm, yrm = self.rm(s)
        xrp, yrp = self.rp(s)
        xsm, ysm = self.sm(r)
        xsp, ysp = self.sp(r)
        return gordonHall(xrm, xrp, xsm, xsp, yrm, yrp, ysm, ysp, r, s)


def semInterp(u, *args, tol: float = 1e-8, maxiter: int = 50):
    """Field value at physical points (interp.jl:47-102).

    semInterp(u, xp, yp, msh): u a nodal field of msh, xp / yp arrays of
    points; each point is located in the element whose map inverts onto it
    (Newton from the closest GLL node); returns a device array of values,
    NaN for points in no element.
    semInterp(u, xp, yp, rxe, rye, sxe, sye, zr, zs, xe, ye[, tol, maxiter]):
    the reference's one-element, one-point form; returns a float."""
    if len(args) == 3 and isinstance(args[2], Mesh):
        xp, yp, msh = args
        xpd, ypd = _dev_vec(xp), _dev_vec(yp)
        if xpd.numel() != ypd.numel():
            raise ValueError("semInterp: xp and yp differ in length")
        ud = to_device(u)
        out = torch.empty(xpd.numel(), dtype=_F64, device="cuda")
        L.call("sem_mesh_interp", ctx(), msh.h, _p(ud), xpd.numel(), _p(xpd), _p(ypd),
               float(tol), int(maxiter), _p(out), None)
        return out
    if len(args) in (10, 12):
        xp, yp, rxe, rye, sxe, sye, zr, zs, xe, ye = args[:10]
        if len(args) == 12:
            tol, maxiter = args[10], args[11]
        f = [to_device(np.asarray(v.cpu() if isinstance(v, torch.Tensor) else v, dtype=np.float64))
             for v in (u, xe, ye, rxe, rye, sxe, sye)]
        nr, ns = f[0].shape
        pts = torch.tensor([[float(xp)], [float(yp)]], dtype=_F64, device="cuda")
        out = torch.empty(1, dtype=_F64, device="cuda")
        L.call("sem_interp_points", ctx(), nr, ns, 1, 1, _p(f[0]), _p(f[1]), _p(f[2]), _p(f[3]),
               _p(f[4]), _p(f[5]), _p(f[6]), 1, _p(pts[0]), _p(pts[1]), float(tol),
               int(maxiter), _p(out), None)
        return float(out.item())
    raise TypeError("semInterp(u, xp, yp, msh) or the reference's 12-argument form")
