"""Pins for the oracle's surface tension (oracle/fs_tension.c + applySurfaceTension in fs_oracle.c) -- CPU.

The reference's own capillarity tests (tests/FluidSurfaceTensionTests.cpp) run on the oracle with their
thresholds: pure phases and grid-aligned flat interfaces never move; a circular drop, a bubble and a
drop against a wall settle to the 2-D Laplace jump sigma/R within 8% with small parasitic currents;
the capillary subdivision of a long step equals the same number of explicit substeps; a solver failure
stops the subdivision.  The geometry alone is pinned by the curvature of an exactly integrated circle.
"""
import math

import numpy as np
import pytest


def disk(w, h, radius, bubble=False, cx=-1.0, cy=-1.0, aspect=1.0):
    """tests/FluidSurfaceTensionTests.cpp:28-58: circle-cell intersection areas (128 rows per cell)."""
    if cx < 0:
        cx = w * 0.5 + 0.13
    if cy < 0:
        cy = h * 0.5 - 0.29
    f = np.zeros(w * h)
    rows = (np.arange(128) + 0.5) / 128.0
    for y in range(h):
        dy = (y + rows - cy) * aspect
        inside = np.abs(dy) < radius
        ext = np.sqrt(np.maximum(radius * radius - dy * dy, 0.0)) * aspect
        for x in range(w):
            seg = np.maximum(0.0, np.minimum(x + 1.0, cx + ext) - np.maximum(float(x), cx - ext)) / 128.0
            area = float(np.sum(seg[inside]))
            f[y * w + x] = np.float32(1.0 - area if bubble else area)
    return f


def capillary_grid(O, W, H, fraction, sigma, h=1.0, gravity=0.0):
    g = O.OracleGrid(W, H, h)
    g.configure([100.0], [0.1])
    g.set_gravity(gravity)
    g.set_surface_tension(sigma)
    vol = h * h
    g.upload(fraction, None, None, None, 100.0 * fraction * vol, 0.1 * (1.0 - fraction) * vol)
    return g


def health(s, mass0):
    lm, gm, f = s["liquid_mass"], s["gas_mass"], s["fraction"]
    assert abs(lm.sum() - mass0) < 1e-4 * max(1.0, mass0)
    assert np.isfinite(f).all() and f.min() >= -1e-6 and f.max() <= 1.000001
    assert np.abs(lm / 100.0 - f).max() < 0.002 and np.abs(gm / 0.1 - (1.0 - f)).max() < 0.002


# ---- tests/FluidSurfaceTensionTests.cpp:99-127 ----
@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_flat_interfaces_do_not_move(oracle, kind):
    W, H = 17, 13
    f = np.array([float(kind) if kind < 2 else float(x < 8 if kind == 2 else y < 6) for y in range(H) for x in range(W)])
    g = capillary_grid(oracle, W, H, f, 10000.0)
    before = g.download()
    for _ in range(8):
        assert g.step(0.005, 2000, 1e-4)["converged"]
    after = g.download()
    assert np.array_equal(after["fraction"], before["fraction"])
    assert not after["u"].any() and not after["v"].any()


def circle_run(O, radius, steps, bubble=False, wall=False, sigma=10000.0, tol=0.001):
    size = 64
    cx = 0.0 if wall else size * 0.5 + 0.13
    cy = size * 0.5 - 0.29
    f = disk(size, size, radius, bubble, cx, cy)
    g = capillary_grid(O, size, size, f, sigma)
    mass0 = g.download()["liquid_mass"].sum()
    peak = 0.0
    for i in range(steps):
        r = g.step(0.005, 2000, tol)
        assert r["converged"], f"step {i}"
        if i % 20 == 0 or i + 1 == steps:
            s = g.download()
            health(s, mass0)
            peak = max(peak, np.abs(s["u"]).max(), np.abs(s["v"]).max())
    s = g.download()
    P = s["pressure"].reshape(size, size)
    ys, xs = np.mgrid[0:size, 0:size]
    r = np.hypot(xs + 0.5 - cx, ys + 0.5 - cy)
    jump = P[r < radius - 2].mean() - P[r > radius + 3].mean()
    return jump, peak, g


# ---- tests/FluidSurfaceTensionTests.cpp:129-193 (CPU cases, :110-112) ----
@pytest.mark.parametrize("bubble,wall", [(False, False), (True, False), (False, True)])
def test_circle_settles_to_laplace_jump(oracle, bubble, wall):
    sigma, radius = 10000.0, 8.0
    jump, peak, _ = circle_run(oracle, radius, 200, bubble, wall, sigma)
    expected = sigma / radius
    assert abs(jump - expected) < 0.08 * expected, f"jump {jump} expected {expected}"
    assert peak < 0.15 * math.sqrt(sigma / (100.0 * radius))


def test_circle_curvature_from_heights(oracle):
    """Interfacial cells of an exactly integrated R = 8 circle carry curvature 1/R within 5%."""
    W = H = 32
    f = disk(W, H, 8.0)
    heights, raw, curv = oracle.capillary_geometry(W, H, 1.0, f)
    valid = curv[:, 1] != 0
    assert valid.sum() > 40
    k = curv[valid, 0]
    assert np.abs(k - 1.0 / 8.0).max() < 0.05 / 8.0
    # pure cells far from the interface carry no curvature
    F = f.reshape(H, W)
    assert not curv.reshape(H, W, 2)[0, 0, 1] and F[0, 0] == 0.0


def test_capillary_step_limit_formula(oracle):
    g = capillary_grid(oracle, 8, 8, np.zeros(64), 40000.0, h=1.0)
    expected = 0.5 * math.sqrt((100.0 + 0.1) * 1.0 / (2.0 * math.pi * 40000.0))
    assert abs(g.capillary_step_limit() - expected) <= 1e-15 * expected
    g.set_surface_tension(0.0)
    assert math.isinf(g.capillary_step_limit())
    with pytest.raises(ValueError):
        g.set_surface_tension(-1.0)


# ---- tests/FluidSurfaceTensionTests.cpp:334-412 (without the acceleration field) ----
def test_subdivision_equals_explicit_substeps_and_stops_on_failure(oracle):
    W = H = 32
    f = disk(W, H, 6.0, False, -1, -1, 1.1)
    auto = capillary_grid(oracle, W, H, f, 40000.0)
    manual = capillary_grid(oracle, W, H, f, 40000.0)
    dt = 0.025
    count = math.ceil(dt / auto.capillary_step_limit())
    assert count > 1
    r = auto.step(dt, 2000, 0.001)
    iterations = 0
    for _ in range(count):
        p = manual.step(dt / count, 2000, 0.001)
        assert p["converged"]
        iterations += p["iterations"]
    assert r["converged"] and r["physical_steps"] == count and r["iterations"] == iterations
    a, b = auto.download(), manual.download()
    for k in a:
        assert np.array_equal(a[k], b[k])
    before = auto.download()
    z = auto.step(0.0)
    assert z["physical_steps"] == 0
    for k, v in auto.download().items():
        assert np.array_equal(v, before[k])
    failed = auto.step(0.025, 0, 1e-12)
    assert not failed["converged"] and failed["physical_steps"] == 1
