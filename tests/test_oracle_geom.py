"""Pins the oracle's gordonHall (geom.jl:8-31) and semInterp
(interp.jl:47-102): a transfinite map reproduces any bilinear map exactly and
interpolates its edge data; Newton inversion recovers known reference points
of deformed elements and exact values on affine elements. Planted mistakes
(the reference's transposed corner matrix, its dx/dy mix-up in the Newton
step) must fail."""
import math

import numpy as np

from oracle import sem_oracle as o


def _bilinear(c, r, s):
    """c[a][b] = value at (r = ze[a], s = ze[b])."""
    R, S = np.meshgrid(r, s, indexing="ij")
    rm, rp, sm, sp = (1 - R) / 2, (1 + R) / 2, (1 - S) / 2, (1 + S) / 2
    return rm * sm * c[0][0] + rp * sm * c[1][0] + rm * sp * c[0][1] + rp * sp * c[1][1]


def _edges_of(fx, fy, zr, zs):
    return (fx(-1.0, zs), fx(1.0, zs), fx(zr, -1.0), fx(zr, 1.0),
            fy(-1.0, zs), fy(1.0, zs), fy(zr, -1.0), fy(zr, 1.0))


def test_gordon_hall_reproduces_bilinear_maps():
    zr, _ = o.gausslobatto(7)
    zs, _ = o.gausslobatto(5)
    cx = [[0.0, 0.3], [2.0, 2.5]]     # x(1,-1) = 2 != x(-1,1) = 0.3
    cy = [[0.0, 1.7], [-0.4, 1.0]]
    fx = lambda r, s: _bilinear(cx, np.atleast_1d(r), np.atleast_1d(s)).reshape(-1)
    fy = lambda r, s: _bilinear(cy, np.atleast_1d(r), np.atleast_1d(s)).reshape(-1)
    x, y = o.gordonHall(*_edges_of(fx, fy, zr, zs), zr, zs)
    np.testing.assert_allclose(x, _bilinear(cx, zr, zs), atol=1e-14)
    np.testing.assert_allclose(y, _bilinear(cy, zr, zs), atol=1e-14)


def test_planted_transposed_corners_fail_bilinear_reproduction():
    zr, _ = o.gausslobatto(6)
    zs = zr
    cx = [[0.0, 0.3], [2.0, 2.5]]
    fx = lambda r, s: _bilinear(cx, np.atleast_1d(r), np.atleast_1d(s)).reshape(-1)
    e = _edges_of(fx, fx, zr, zs)
    ze = np.array([-1.0, 1.0])
    Jer, Jes = o.interpMat(zr, ze), o.interpMat(zs, ze)
    xrm, xrp, xsm, xsp = e[:4]
    C_ref_layout = np.array([[xrm[0], xrp[0]], [xrm[-1], xrp[-1]]])   # geom.jl:16-17
    bad = Jer @ np.vstack([xrm, xrp]) + np.column_stack([xsm, xsp]) @ Jes.T \
        - Jer @ C_ref_layout @ Jes.T
    assert np.abs(bad - _bilinear(cx, zr, zs)).max() > 0.1


def test_gordon_hall_interpolates_curved_edges():
    zr, _ = o.gausslobatto(9)
    zs, _ = o.gausslobatto(8)
    # quarter annulus r0 = 0.5, r1 = 1: r-edges are rays, s-edges arcs
    th = lambda s: np.pi / 4 * (np.asarray(s) + 1)
    rad = lambda r: 0.75 + 0.25 * np.asarray(r)
    xrm, yrm = 0.5 * np.cos(th(zs)), 0.5 * np.sin(th(zs))
    xrp, yrp = 1.0 * np.cos(th(zs)), 1.0 * np.sin(th(zs))
    xsm, ysm = rad(zr), 0 * zr
    xsp, ysp = 0 * zr, rad(zr)
    x, y = o.gordonHall(xrm, xrp, xsm, xsp, yrm, yrp, ysm, ysp, zr, zs)
    np.testing.assert_allclose(x[0, :], xrm, atol=1e-15)
    np.testing.assert_allclose(x[-1, :], xrp, atol=1e-15)
    np.testing.assert_allclose(x[:, 0], xsm, atol=1e-15)
    np.testing.assert_allclose(y[:, -1], ysp, atol=1e-15)
    # interior nodes lie inside the annulus sector
    R = np.hypot(x, y)
    assert R.min() >= 0.5 - 1e-12 and R.max() <= 1.0 + 1e-12


def _element(n, deform):
    m = o.Mesh(n, n, 1, 1, deform=deform, dense_gs=False)
    return m


def test_sem_interp_recovers_reference_points_on_deformed_element():
    rot = lambda x, y: (0.8 * x - 0.6 * y + 0.1 * np.sin(np.pi * x) * y,
                        0.6 * x + 0.8 * y + 0.05 * x * x)
    m = _element(8, rot)
    rng = np.random.default_rng(3)
    u = np.cos(m.x) * np.exp(m.y)
    for _ in range(5):
        r0, s0 = rng.uniform(-0.9, 0.9, 2)
        Jr, Js = o.interpMat(r0, m.zr), o.interpMat(s0, m.zs)
        xp = float(o.ABu(Js, Jr, m.x)[0, 0])
        yp = float(o.ABu(Js, Jr, m.y)[0, 0])
        want = float(o.ABu(Js, Jr, u)[0, 0])
        got = o.semInterp(u, xp, yp, m.rx, m.ry, m.sx, m.sy, m.zr, m.zs, m.x, m.y, tol=1e-13)
        assert abs(got - want) <= 1e-11 * max(1.0, abs(want))


def test_sem_interp_exact_on_affine_element_and_outside_is_nan():
    aff = lambda x, y: (2.0 + 0.5 * x + 0.2 * y, -1.0 + 0.1 * x + 0.7 * y)
    m = _element(6, aff)
    f = lambda x, y: x ** 3 * y ** 2 - 2 * x * y ** 4 + 1.0   # degree 5 = N
    u = f(m.x, m.y)
    for (r0, s0) in ((0.3, -0.2), (-0.95, 0.9), (0.0, 0.0)):
        xp, yp = aff(r0, s0)
        got = o.semInterp(u, xp, yp, m.rx, m.ry, m.sx, m.sy, m.zr, m.zs, m.x, m.y, tol=1e-13)
        assert abs(got - f(xp, yp)) <= 1e-12 * max(1.0, abs(f(xp, yp)))
    xp, yp = aff(1.5, 0.0)
    assert math.isnan(o.semInterp(u, xp, yp, m.rx, m.ry, m.sx, m.sy, m.zr, m.zs, m.x, m.y))


def test_planted_reference_newton_step_fails_on_rotated_element():
    rot = lambda x, y: (0.6 * x - 0.8 * y, 0.8 * x + 0.6 * y)
    m = _element(5, rot)
    u = m.x + 2 * m.y
    r0, s0 = 0.4, -0.3
    xp, yp = rot(r0, s0)
    good = o.semInterp(u, xp, yp, m.rx, m.ry, m.sx, m.sy, m.zr, m.zs, m.x, m.y, tol=1e-12)
    assert abs(good - (xp + 2 * yp)) < 1e-12
    # the update of interp.jl:80-81 (rx dx + ry dx, sx dy + sy dy)
    d2 = (m.x - xp) ** 2 + (m.y - yp) ** 2
    ia, ib = np.unravel_index(np.argmin(d2), d2.shape)
    r, s = m.zr[ia], m.zs[ib]
    ok = False
    for _ in range(50):
        Jr, Js = o.interpMat(r, m.zr), o.interpMat(s, m.zs)
        dx = xp - float(o.ABu(Js, Jr, m.x)[0, 0])
        dy = yp - float(o.ABu(Js, Jr, m.y)[0, 0])
        if math.hypot(dx, dy) < 1e-12:
            ok = True
            break
        rx, ry = float(o.ABu(Js, Jr, m.rx)[0, 0]), float(o.ABu(Js, Jr, m.ry)[0, 0])
        sx, sy = float(o.ABu(Js, Jr, m.sx)[0, 0]), float(o.ABu(Js, Jr, m.sy)[0, 0])
        r, s = r + rx * dx + ry * dx, s + sx * dy + sy * dy
        if abs(r) > 1 or abs(s) > 1:
            break
    assert not ok
