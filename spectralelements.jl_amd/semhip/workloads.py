"""Seeded synthetic inputs of the benchmark problems, generated on the device.

Random draws are per global node in the reference's global (column-major)
order (mesh.jl:87-88) and scattered to every local copy, so each field is
continuous. numpy's Generator(PCG64(seed)) is the generator.
"""
from __future__ import annotations

import numpy as np
import torch

from .core import Mesh, empty_field, gatherScatter


def global_normal_field(m: Mesh, seed: int) -> torch.Tensor:
    """f ~ N(0, 1) i.i.d. per global node, scattered to the local copies."""
    rng = np.random.default_rng(seed)
    g = torch.from_numpy(rng.standard_normal(m.nglobal)).cuda()
    f = empty_field(*m.shape)
    f.copy_(g[m.l2g.long()])
    return f


def fourier_modes(seed: int, nmodes: int = 16, kmax: int = 4):
    """(kx, ky, theta) of nmodes random-phase modes, k in [1, kmax]^2."""
    rng = np.random.default_rng(seed)
    k = rng.integers(1, kmax + 1, size=(nmodes, 2))
    th = rng.uniform(0.0, 2.0 * np.pi, size=nmodes)
    return k, th


def smooth_coefficient(x, y, seed: int, nmodes: int = 16, amp: float = 0.25, kmax: int = 4):
    """nu = 1 + 0.5 tanh(phi), phi = sum_m amp cos(pi (kx_m x + ky_m y) + theta_m);
    x, y torch or numpy arrays of node coordinates."""
    k, th = fourier_modes(seed, nmodes, kmax)
    lib = torch if isinstance(x, torch.Tensor) else np
    phi = 0.0 * x
    for (kx, ky), t in zip(k, th):
        phi = phi + amp * lib.cos(np.pi * (float(kx) * x + float(ky) * y) + float(t))
    return 1.0 + 0.5 * lib.tanh(phi)


def continuous(u, m: Mesh) -> torch.Tensor:
    """mult .* QQ' u: the average of each global node's copies, written to
    every copy (bit-identical on all copies)."""
    return gatherScatter(u, m) * m.mult


def coefficient_field(m: Mesh, seed: int) -> torch.Tensor:
    """The variable coefficient of the steady adjoint problem, continuous."""
    nu = smooth_coefficient(m.x, m.y, seed)
    out = empty_field(*m.shape)
    out.copy_(continuous(nu, m))
    return out


def gaussian_blob(m: Mesh, x0: float = 0.5, y0: float = 0.0, sigma: float = 0.1) -> torch.Tensor:
    """u0 = exp(-((x - x0)^2 + (y - y0)^2) / (2 sigma^2)) (cd2d initial condition)."""
    u = torch.exp(-((m.x - x0) ** 2 + (m.y - y0) ** 2) / (2.0 * sigma * sigma))
    out = empty_field(*m.shape)
    out.copy_(u)
    return out


def helmholtz_rhs(se, m: Mesh, M, seed: int) -> torch.Tensor:
    """b = QQ'(M .* (B .* f)) for f ~ N(0, 1) per global node: the assembled
    right-hand side of the benchmark solves (makeRHS!, diffusion.jl:51-65,
    with a random forcing)."""
    f = global_normal_field(m, seed)
    return se.gatherScatter(se.mask(se.mass(f, m), M), m)
