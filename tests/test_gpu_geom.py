"""GPU parity of gordonHall (transfinite mesh generation, geom.jl:8-31) and
semInterp (point evaluation, interp.jl:47-102) against the oracle, through
the C ABI: transfinite coordinates within 1e-14, a mesh built from a
transfinite map against the oracle mesh on the same map (geometry within
1e-12), and field values at physical points within 1e-10 (Newton to 1e-12)."""
import math
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


def _sector_edges(zr, zs, lib=np):
    th = lambda s: np.pi / 4 * (s + 1)
    rad = lambda r: 0.75 + 0.25 * r
    return (0.5 * lib.cos(th(zs)), 1.0 * lib.cos(th(zs)), rad(zr), 0 * zr,
            0.5 * lib.sin(th(zs)), 1.0 * lib.sin(th(zs)), 0 * zr, rad(zr))


def test_gordon_hall_matches_oracle(se):
    _log("gordonHall")
    zr, _ = o.gausslobatto(9)
    zs, _ = o.gausslobatto(7)
    e = _sector_edges(zr, zs)
    x, y = se.gordonHall(*e, zr, zs)
    xo, yo = o.gordonHall(*e, zr, zs)
    np.testing.assert_allclose(se.to_host(x), xo, rtol=0, atol=1e-14)
    np.testing.assert_allclose(se.to_host(y), yo, rtol=0, atol=1e-14)


def test_mesh_from_transfinite_map_matches_oracle(se):
    """A 12 x 10 element mesh, N = 7, of the quarter annulus sector by a
    transfinite map; geometric factors against the oracle mesh built on the
    oracle's transfinite coordinates."""
    import torch
    _log("gordonHall mesh")
    th = lambda s: np.pi / 4 * (s + 1)
    rad = lambda r: 0.75 + 0.25 * r
    gmap = se.gordonHallMap(
        lambda s: (0.5 * torch.cos(th(s)), 0.5 * torch.sin(th(s))),
        lambda s: (1.0 * torch.cos(th(s)), 1.0 * torch.sin(th(s))),
        lambda r: (rad(r), 0 * r),
        lambda r: (0 * r, rad(r)))
    m = se.Mesh(8, 8, 12, 10, (False, False), gmap)

    def odef(x, y):
        r, s = x[:, 0], y[0, :]
        return o.gordonHall(0.5 * np.cos(th(s)), np.cos(th(s)), rad(r), 0 * r,
                            0.5 * np.sin(th(s)), np.sin(th(s)), 0 * r, rad(r), r, s)
    mo = o.Mesh(8, 8, 12, 10, (False, False), odef, dense_gs=False)
    for name in ("x", "y"):
        np.testing.assert_allclose(se.to_host(m.field(name)), getattr(mo, name), atol=1e-14)
    for grp in (["Jac"], ["B"], ["G11", "G12", "G22"]):
        scale = max(np.abs(getattr(mo, nm)).max() for nm in grp)
        for name in grp:
            assert np.abs(se.to_host(m.field(name)) - getattr(mo, name)).max() <= 1e-12 * scale
    # area of the sector: pi/4 (1 - 0.25), up to the spectral geometry error
    assert abs(float(m.B.sum()) - np.pi / 4 * 0.75) < 1e-6


def test_sem_interp_points_match_oracle(se):
    _log("semInterp mesh")
    n, E = 8, 6
    m = se.Mesh(n, n, E, E, (False, False), se.deform_sinsin(0.1))
    mo = o.Mesh(n, n, E, E, deform=o.deform_sinsin(0.1), dense_gs=False)
    u = np.cos(2 * mo.x) * np.exp(mo.y) + mo.x * mo.y
    rng = np.random.default_rng(8)
    # points given by reference coordinates inside random elements
    pts, want = [], []
    for _ in range(40):
        ex, ey = rng.integers(0, E, 2)
        r0, s0 = rng.uniform(-1, 1, 2)
        blk = np.s_[ex * n:(ex + 1) * n, ey * n:(ey + 1) * n]
        Jr, Js = o.interpMat(r0, mo.zr), o.interpMat(s0, mo.zs)
        xp = float(o.ABu(Js, Jr, mo.x[blk])[0, 0])
        yp = float(o.ABu(Js, Jr, mo.y[blk])[0, 0])
        pts.append((xp, yp))
        want.append(float(o.ABu(Js, Jr, u[blk])[0, 0]))
    pts = np.array(pts)
    got = se.to_host(se.semInterp(se.to_device(u), pts[:, 0], pts[:, 1], m, tol=1e-12)
                     .reshape(-1, 1)).reshape(-1)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-10 * np.abs(want).max())
    # outside the domain: NaN
    out = se.to_host(se.semInterp(se.to_device(u), [3.0, -2.0], [0.0, 5.0], m)
                     .reshape(-1, 1)).reshape(-1)
    assert np.all(np.isnan(out))
    # the reference's one-element form against the oracle's
    blk = np.s_[2 * n:3 * n, 3 * n:4 * n]
    xp, yp = pts[0]
    e = [mo.rx[blk], mo.ry[blk], mo.sx[blk], mo.sy[blk]]
    r1 = se.semInterp(u[blk], 0.5 * (mo.x[blk].max() + mo.x[blk].min()),
                      0.5 * (mo.y[blk].max() + mo.y[blk].min()), *e, mo.zr, mo.zs, mo.x[blk],
                      mo.y[blk], 1e-12, 50)
    r2 = o.semInterp(u[blk], 0.5 * (mo.x[blk].max() + mo.x[blk].min()),
                     0.5 * (mo.y[blk].max() + mo.y[blk].min()), *e, mo.zr, mo.zs, mo.x[blk],
                     mo.y[blk], tol=1e-12)
    assert math.isfinite(r1) and abs(r1 - r2) <= 1e-10 * max(1.0, abs(r2))
