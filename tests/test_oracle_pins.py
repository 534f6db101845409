"""Closed-form pins for oracle functions that the reference's own tests leave open -- CPU.

* The capillary quadratic fit (oracle/fs_tension.c fit_add / fit_evaluate, restating
  FluidSim/UniformCPU/FluidSurfaceTensionCPU.h:204-254): points that lie exactly on a parabola
  y' = a + b x' + c x'^2 in the frame of the fit normal must return the parabola's curvature at its
  vertex-frame origin, -2c / (1 + b^2)^(3/2), clamped to [-2, 2] and divided by the cell size.  Collinear
  or coincident points must be refused.
* The interface segment (segment, :287-324) and the PLIC line offset (interface_geometry::offset,
  FluidInterfaceGeometry.h:51-61): the line n.p = offset must cut the unit cell into a liquid part of the
  cell's fraction (polygon clipping), and the segment must be that chord: its centroid is the chord's
  midpoint and its length the chord's length.
* Curvature completion (completeCurvature, :326-376): a cell whose own estimate is invalid takes the
  mean of its valid mirrored 3x3 neighbours; with none valid, the segment fit.  Under-resolved drops
  (R = 1.5 ... 2.5 cells) go through both branches; their curvature must have the drop's sign and stay
  within the stated band around 1/R.
* addVortex (oracle/fs_oracle.c orc_add_vortex, FluidGridCPU.cpp:998-1028): face positions are
  origin + (x h, (y + 1/2) h) for u and origin + ((x + 1/2) h, y h) for v, the velocity is
  (-dy, dx) w exp(-r^2 / 2R^2): zero on the row/column through the centre, odd about it, and one face
  value equal to the closed form.
Each check is chosen so that a sign, an index or an offset slip in the C restatement fails it.
"""
import math

import numpy as np
import pytest

from test_oracle_tension import disk


def _rotate_back(xt, yn, n):
    """Inverse of the fit's frame change x' = -px ny + py nx, y' = px nx + py ny (an involution)."""
    nx, ny = n
    return -xt * ny + yn * nx, xt * nx + yn * ny


@pytest.mark.parametrize("seed", range(6))
def test_fit_returns_the_curvature_of_an_exact_parabola(oracle, seed):
    rng = np.random.default_rng(seed)
    angle = rng.uniform(0, 2 * math.pi)
    n = (math.cos(angle), math.sin(angle))
    a, b, c = rng.uniform(-0.4, 0.4), rng.uniform(-0.8, 0.8), rng.uniform(-0.6, 0.6)
    xt = np.array([-1.3, -0.7, -0.1, 0.45, 1.1, 0.8])[: 3 + seed % 4]
    yn = a + b * xt + c * xt * xt
    px, py = _rotate_back(xt, yn, n)
    h = 0.25
    expect = max(-2.0, min(2.0, -2.0 * c / (1.0 + b * b) ** 1.5)) / h
    k, valid = oracle.fit_curvature(px, py, n, h)
    assert valid
    assert k == pytest.approx(expect, rel=1e-10, abs=1e-12)
    # any positive weights fit the same exact parabola
    kw, valid_w = oracle.fit_curvature(px, py, n, h, weight=rng.uniform(0.1, 2.0, xt.size))
    assert valid_w and kw == pytest.approx(expect, rel=1e-10, abs=1e-12)


def test_fit_clamps_and_refuses_degenerate_points(oracle):
    n = (0.0, 1.0)
    xt = np.array([-1.0, 0.0, 1.0])
    # curvature -2c/(1+b^2)^1.5 = -6 is clamped to -2 (then divided by the cell size)
    px, py = _rotate_back(xt, 3.0 * xt * xt, n)
    k, valid = oracle.fit_curvature(px, py, n, 0.5)
    assert valid and k == -2.0 / 0.5
    px, py = _rotate_back(xt, -3.0 * xt * xt, n)
    assert oracle.fit_curvature(px, py, n, 0.5) == (2.0 / 0.5, True)
    # a straight line has zero curvature whatever its slope
    px, py = _rotate_back(xt, 0.3 + 0.4 * xt, n)
    k, valid = oracle.fit_curvature(px, py, n, 1.0)
    assert valid and abs(k) < 1e-14
    # two distinct abscissae cannot fix a parabola; no points at all give nothing
    px, py = _rotate_back(np.array([0.2, 0.2, 0.9]), np.array([0.1, 0.3, 0.0]), n)
    assert oracle.fit_curvature(px, py, n, 1.0)[1] is False
    assert oracle.fit_curvature(np.zeros(0), np.zeros(0), n, 1.0)[1] is False


def _clip_below(n, offset):
    """The part of the unit square with n.p <= offset (Sutherland-Hodgman), as a vertex list."""
    square = [(0.0, 0.0), (1.0, 0.0), (1.0, 1.0), (0.0, 1.0)]
    out = []
    for i in range(4):
        p, q = square[i], square[(i + 1) % 4]
        fp, fq = n[0] * p[0] + n[1] * p[1] - offset, n[0] * q[0] + n[1] * q[1] - offset
        if fp <= 0:
            out.append(p)
        if (fp < 0 < fq) or (fq < 0 < fp):
            t = fp / (fp - fq)
            out.append((p[0] + t * (q[0] - p[0]), p[1] + t * (q[1] - p[1])))
    return out


def _area(poly):
    s = 0.0
    for i in range(len(poly)):
        (x0, y0), (x1, y1) = poly[i], poly[(i + 1) % len(poly)]
        s += x0 * y1 - x1 * y0
    return 0.5 * abs(s)


def _chord(n, offset):
    """End points of the line n.p = offset inside the unit square (edge intersections)."""
    pts = []
    for x in (0.0, 1.0):
        if abs(n[1]) > 1e-12:
            y = (offset - n[0] * x) / n[1]
            if -1e-12 <= y <= 1 + 1e-12:
                pts.append((x, min(1.0, max(0.0, y))))
    for y in (0.0, 1.0):
        if abs(n[0]) > 1e-12:
            x = (offset - n[1] * y) / n[0]
            if -1e-12 <= x <= 1 + 1e-12:
                pts.append((min(1.0, max(0.0, x)), y))
    pts = np.array(pts)
    # the two points farthest apart (corner hits appear twice)
    best, pair = -1.0, None
    for i in range(len(pts)):
        for j in range(i + 1, len(pts)):
            d = np.hypot(*(pts[i] - pts[j]))
            if d > best:
                best, pair = d, (pts[i], pts[j])
    return pair


@pytest.mark.parametrize("seed", range(4))
def test_segment_is_the_plic_chord_of_the_cell(oracle, seed):
    rng = np.random.default_rng(100 + seed)
    for _ in range(60):
        angle = rng.uniform(0, 2 * math.pi)
        n = (math.cos(angle), math.sin(angle))
        if min(abs(n[0]), abs(n[1])) < 1e-3:
            continue
        f = rng.uniform(0.02, 0.98)
        offset = oracle.line_offset(n, f)
        assert _area(_clip_below(n, offset)) == pytest.approx(f, abs=1e-12)
        cx, cy, length = oracle.interface_segment(n, f)
        p, q = _chord(n, offset)
        mid = 0.5 * (p + q)
        assert cx == pytest.approx(mid[0] - 0.5, abs=1e-12)
        assert cy == pytest.approx(mid[1] - 0.5, abs=1e-12)
        assert length == pytest.approx(float(np.hypot(*(p - q))), abs=1e-12)
    # axis-aligned normals: a horizontal chord across the whole cell at the fraction's height
    cx, cy, length = oracle.interface_segment((0.0, 1.0), 0.3)
    assert (cx, cy, length) == pytest.approx((0.0, 0.3 - 0.5, 1.0), abs=1e-15)
    cx, cy, length = oracle.interface_segment((-1.0, 0.0), 0.25)
    assert (cx, cy, length) == pytest.approx((1.0 - 0.25 - 0.5, 0.0, 1.0), abs=1e-15)


def _mirror(i, extent):
    period = 2 * extent
    q = i % period
    return period - 1 - q if q >= extent else q


CIRCLES = [(1.5, -1, -1), (1.5, 8.37, 8.61), (1.5, 8.5, 8.5), (2.0, -1, -1), (2.0, 8.37, 8.61),
           (2.5, 8.5, 8.5), (2.5, 8.37, 8.61), (5.0, 7.9, 8.3)]


@pytest.mark.parametrize("radius,cx,cy", CIRCLES)
@pytest.mark.parametrize("bubble", [False, True])
def test_completion_takes_the_neighbour_mean_then_the_segment_fit(oracle, radius, cx, cy, bubble):
    W = H = 17
    f = disk(W, H, radius, bubble=bubble, cx=cx, cy=cy)
    _, raw, curv = oracle.capillary_geometry(W, H, 1.0, f)
    rawv = raw[:, 1].reshape(H, W) != 0
    rawk = raw[:, 0].reshape(H, W)
    fin = curv.reshape(H, W, 2)
    by_mean = by_fit = 0
    for y in range(H):
        for x in range(W):
            if rawv[y, x]:
                assert fin[y, x, 1] and fin[y, x, 0] == rawk[y, x]
                continue
            if not fin[y, x, 1]:
                continue
            vals = [rawk[_mirror(y + dy, H), _mirror(x + dx, W)] for dy in (-1, 0, 1) for dx in (-1, 0, 1)
                    if rawv[_mirror(y + dy, H), _mirror(x + dx, W)]]
            if vals:
                by_mean += 1
                assert fin[y, x, 0] == pytest.approx(sum(vals) / len(vals), rel=1e-15)
            else:
                by_fit += 1
    # the drop's curvature is positive, the bubble's negative, and near 1/R on every completed cell
    completed = (fin[:, :, 1] != 0) & ~rawv
    k = fin[:, :, 0][completed] * radius * (-1.0 if bubble else 1.0)
    if radius <= 2.5:
        assert completed.sum() > 0
        assert np.all(k > 0.55) and np.all(k < 1.75)
        assert abs(k.mean() - 1.0) < 0.4
    if (radius, cx, cy) == (1.5, 8.5, 8.5):
        # no height-function estimate survives: every value comes from the segment fit
        assert by_fit == 8 and by_mean == 0 and not rawv.any()


def test_add_vortex_face_positions_and_closed_form(oracle):
    W, H, h = 16, 12, 0.5
    ox, oy = 0.25, -0.75
    x0, y0 = 6, 5
    R, w = 1.7, 0.9
    cx = ox + x0 * h             # on the u-face column x0 and the v-face row y0 ...
    cy = oy + (y0 + 0.5) * h     # ... and through the centres of u row y0 and v column x0
    g = oracle.OracleGrid(W, H, h)
    g.configure([100.0], [0.1])
    g.set_origin(ox, oy)
    g.upload(np.ones(W * H), None, None, None, 100.0 * h * h * np.ones(W * H), np.zeros(W * H))
    g.add_vortex(cx, cy, R, w)
    s = g.download()
    u = s["u"].reshape(H, W + 1)
    v = s["v"].reshape(H + 1, W)
    # u vanishes on the row through the centre and is odd about it
    assert np.all(u[y0] == 0.0)
    for d in range(1, min(y0, H - 1 - y0) + 1):
        assert np.array_equal(u[y0 + d], -u[y0 - d])
    # v = dx w e^(-r^2/2R^2) with v's x positions at cell centres: zero on no column, odd about cx,
    # which lies on a u face; on v row y0 + k the closed form below holds
    for k in (0, 2):
        for x in (x0, x0 + 3):
            dx = ox + (x + 0.5) * h - cx
            dy = oy + (y0 + k) * h - cy
            expect = dx * w * math.exp(-(dx * dx + dy * dy) / (2.0 * R * R))
            assert v[y0 + k, x] == pytest.approx(expect, rel=1e-14)
    for x in range(W):
        mirror_x = 2 * x0 - 1 - x
        if 0 <= mirror_x < W:
            assert np.array_equal(v[:, x], -v[:, mirror_x])
    # one u face in closed form: u = -dy w e^(-dy^2 / 2R^2) on column x0, d rows below the centre
    d = 3
    dy = d * h
    assert u[y0 + d, x0] == pytest.approx(-dy * w * math.exp(-dy * dy / (2.0 * R * R)), rel=1e-15)
    # a second vortex adds (addVortex accumulates)
    g.add_vortex(cx, cy, R, w)
    assert np.allclose(g.download()["u"].reshape(H, W + 1), 2.0 * u, rtol=1e-15, atol=0)
