"""Seeded synthetic inputs for the benchmark and parity workloads.

Random numbers come from SplitMix64 (Steele, Lea & Flood 2014) mapped to doubles through the top
53 bits, with Box-Muller for normals, so every input is reproducible bit for bit on any machine.
Layouts follow the reference (cells W*H, u (W+1)*H, v W*(H+1), y downward).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List

import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


class SplitMix64:
    def __init__(self, seed: int):
        self.state = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)

    def next_u64(self, n: int) -> np.ndarray:
        with np.errstate(over="ignore"):
            steps = np.arange(1, n + 1, dtype=np.uint64)
            z = self.state + steps * _GAMMA
            self.state = self.state + np.uint64(n) * _GAMMA
            z = (z ^ (z >> np.uint64(30))) * _M1
            z = (z ^ (z >> np.uint64(27))) * _M2
            return z ^ (z >> np.uint64(31))

    def uniform(self, n: int, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
        u = (self.next_u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u

    def normal(self, n: int) -> np.ndarray:
        u1 = self.uniform(n)
        u2 = self.uniform(n)
        return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * math.pi * u2)


@dataclass
class Workload:
    """A complete upload: grid shape, materials and state (host FP64 arrays)."""
    name: str
    width: int
    height: int
    cell_size: float
    liquid_density: List[float]
    gas_density: List[float]
    gravity_y: float
    dt: float
    fraction: np.ndarray
    u: np.ndarray
    v: np.ndarray
    liquid_mass: np.ndarray  # component-major, k * N
    gas_mass: np.ndarray
    meta: dict = field(default_factory=dict)


def two_phase_masses(fraction: np.ndarray, cell_size: float, rho_l: float = 100.0, rho_g: float = 0.1):
    """Water 100 / air 0.1 masses = density * phase fraction * h^2 (tests/FluidTestSetup.h:20-29)."""
    volume = cell_size * cell_size
    return rho_l * fraction * volume, rho_g * (1.0 - fraction) * volume


def zero_walls(u: np.ndarray, v: np.ndarray, W: int, H: int):
    u2 = u.reshape(H, W + 1)
    u2[:, 0] = 0.0
    u2[:, W] = 0.0
    v2 = v.reshape(H + 1, W)
    v2[0, :] = 0.0
    v2[H, :] = 0.0


def add_gaussian(u, v, W, H, h, cx, cy, radius, strength, kind="vortex", window=10.0):
    """FluidGridCPU::addVortex (FluidGridCPU.cpp:998-1028) at world positions (origin 0).

    kind="vortex" adds (-dy, dx) * strength * falloff; kind="source" adds (dx, dy) * ... (a divergent
    Gaussian).  Faces farther than `window` radii from the centre are skipped: their falloff is below
    exp(-50) ~ 2e-22, which cannot change an FP64 sum of O(1e-2) terms.
    """
    reach = window * radius
    u2 = u.reshape(H, W + 1)
    v2 = v.reshape(H + 1, W)
    x0 = max(0, int(math.floor((cx - reach) / h)) - 1)
    x1 = min(W, int(math.ceil((cx + reach) / h)) + 1)
    y0 = max(0, int(math.floor((cy - reach) / h)) - 1)
    y1 = min(H, int(math.ceil((cy + reach) / h)) + 1)
    if x1 <= x0 or y1 <= y0:
        return
    # u faces: (x h, (y + 0.5) h), x in [x0, x1], y in [y0, y1)
    xs = np.arange(x0, min(x1, W) + 1)
    ys = np.arange(y0, min(y1, H))
    px = xs[None, :] * h
    py = (ys[:, None] + 0.5) * h
    dx, dy = px - cx, py - cy
    falloff = np.exp(-(dx * dx + dy * dy) / (2.0 * radius * radius))
    if kind == "vortex":
        u2[np.ix_(ys, xs)] += -dy * strength * falloff
    else:
        u2[np.ix_(ys, xs)] += dx * strength * falloff
    # v faces: ((x + 0.5) h, y h), x in [x0, x1), y in [y0, y1]
    xs = np.arange(x0, min(x1, W))
    ys = np.arange(y0, min(y1, H) + 1)
    px = (xs[None, :] + 0.5) * h
    py = ys[:, None] * h
    dx, dy = px - cx, py - cy
    falloff = np.exp(-(dx * dx + dy * dy) / (2.0 * radius * radius))
    if kind == "vortex":
        v2[np.ix_(ys, xs)] += dx * strength * falloff
    else:
        v2[np.ix_(ys, xs)] += dy * strength * falloff


def dam_break(W: int = 128, H: int = 128, seed: int = 1) -> Workload:
    """Dam break: liquid where x_idx < 0.4 W and y_idx >= 0.4 H (the bottom 60%, y downward), water
    100 / air 0.1, interior face velocities U(-1e-3, 1e-3), gravity +9.81 (downward), dt = 1/120."""
    h = 1.0 / W
    xs = np.arange(W)[None, :]
    ys = np.arange(H)[:, None]
    fraction = ((xs < 0.4 * W) & (ys >= 0.4 * H)).astype(np.float64).ravel()
    rng = SplitMix64(seed)
    u = rng.uniform((W + 1) * H, -1e-3, 1e-3)
    v = rng.uniform(W * (H + 1), -1e-3, 1e-3)
    zero_walls(u, v, W, H)
    lm, gm = two_phase_masses(fraction, h)
    return Workload("dam_break", W, H, h, [100.0], [0.1], 9.81, 1.0 / 120.0, fraction, u, v, lm, gm,
                    {"seed": seed})


def vortex_box(W: int = 1024, H: int = 1024, seed: int = 2, vortices: int = 64, sources: int = 16) -> Workload:
    """Fully liquid box with `vortices` Gaussian vortices (centre U[0,1]^2, radius U[0.01,0.05],
    strength N(0,1)) plus `sources` divergent Gaussians of the same law; walls closed."""
    h = 1.0 / W
    rng = SplitMix64(seed)
    n = vortices + sources
    cx = rng.uniform(n)
    cy = rng.uniform(n) * (H * h)
    cx = cx * (W * h)
    radius = rng.uniform(n, 0.01, 0.05)
    strength = rng.normal(n)
    u = np.zeros((W + 1) * H)
    v = np.zeros(W * (H + 1))
    for k in range(n):
        add_gaussian(u, v, W, H, h, cx[k], cy[k], radius[k], strength[k],
                     "vortex" if k < vortices else "source")
    zero_walls(u, v, W, H)
    fraction = np.ones(W * H)
    lm, gm = two_phase_masses(fraction, h)
    return Workload("vortex_box", W, H, h, [100.0], [0.1], 9.81, 1.0 / 120.0, fraction, u, v, lm, gm,
                    {"seed": seed, "vortices": vortices, "sources": sources})


def obstacle_mask(W: int, H: int, rng: SplitMix64, shapes: int = 40, coverage: float = 0.15) -> np.ndarray:
    """Seeded solid mask of `shapes` circles and axis-aligned boxes (alternating), centres uniform in
    the domain, base radii / half-sides U[0.5, 1.5] of a common scale that is chosen (bisection) so the
    union covers `coverage` of the cells; fluid pockets cut off from the largest fluid region are filled,
    so the fluid stays one 4-connected region.  Returns W*H uint8 (1 = solid)."""
    from scipy import ndimage

    cx = rng.uniform(shapes) * W
    cy = rng.uniform(shapes) * H
    a = rng.uniform(shapes, 0.5, 1.5)
    b = rng.uniform(shapes, 0.5, 1.5)
    xs = np.arange(W)[None, :] + 0.5
    ys = np.arange(H)[:, None] + 0.5

    def union(scale):
        m = np.zeros((H, W), dtype=bool)
        for k in range(shapes):
            ra, rb = a[k] * scale, b[k] * scale
            x0, x1 = max(0, int(cx[k] - ra) - 1), min(W, int(cx[k] + ra) + 2)
            y0, y1 = max(0, int(cy[k] - rb) - 1), min(H, int(cy[k] + rb) + 2)
            if x1 <= x0 or y1 <= y0:
                continue
            dx = xs[:, x0:x1] - cx[k]
            dy = ys[y0:y1, :] - cy[k]
            if k % 2 == 0:
                inside = dx * dx + dy * dy < ra * ra
            else:
                inside = (np.abs(dx) < ra) & (np.abs(dy) < rb)
            m[y0:y1, x0:x1] |= inside
        return m

    lo, hi = 0.0, 0.5 * max(W, H)
    for _ in range(30):
        mid = 0.5 * (lo + hi)
        if union(mid).mean() < coverage:
            lo = mid
        else:
            hi = mid
    solid = union(hi)
    labels, count = ndimage.label(~solid)
    if count > 1:
        sizes = np.bincount(labels.ravel())
        sizes[0] = 0
        solid |= labels != int(np.argmax(sizes))
    return solid.astype(np.uint8).ravel()


def obstacle_tank(W: int = 2048, H: int = 2048, seed: int = 3, shapes: int = 40, coverage: float = 0.15,
                  fill: float = 0.7, vortices: int = 64, sources: int = 16) -> Workload:
    """Tank with a seeded obstacle mask (`obstacle_mask`, about `coverage` of the cells solid), liquid
    f = 1 in the bottom `fill` of the rows (water 100 / air 0.1: a free surface with a 1000:1 density
    jump), and the vortex-box velocity law (Gaussian vortices and sources, walls closed); the mask is
    in meta["solid"].  dt = 1/120."""
    h = 1.0 / max(W, H)
    rng = SplitMix64(seed)
    solid = obstacle_mask(W, H, rng, shapes, coverage)
    n = vortices + sources
    cx = rng.uniform(n) * (W * h)
    cy = rng.uniform(n) * (H * h)
    radius = rng.uniform(n, 0.01, 0.05)
    strength = rng.normal(n)
    u = np.zeros((W + 1) * H)
    v = np.zeros(W * (H + 1))
    for k in range(n):
        add_gaussian(u, v, W, H, h, cx[k], cy[k], radius[k], strength[k],
                     "vortex" if k < vortices else "source")
    zero_walls(u, v, W, H)
    ys = np.arange(H)[:, None] + 0.5
    fraction = np.broadcast_to((ys >= (1.0 - fill) * H).astype(np.float64), (H, W)).ravel().copy()
    fraction[solid != 0] = 0.0
    lm, gm = two_phase_masses(fraction, h)
    return Workload("obstacle_tank", W, H, h, [100.0], [0.1], 9.81, 1.0 / 120.0, fraction, u, v, lm, gm,
                    {"seed": seed, "solid": solid, "solid_fraction": float(solid.mean())})


def sloped_pool(W: int, H: int, cell_size: float = 4.0, seed: int = 0) -> Workload:
    """The reference tests' sloped pool fraction clamp(y + 1 - (0.35 x + 0.2 H), 0, 1)
    (tests/FluidGridTests.cpp:629-641) with u = 3 sin(0.07 i), v = 4 cos(0.11 i)."""
    xs = np.arange(W)[None, :].astype(np.float64)
    ys = np.arange(H)[:, None].astype(np.float64)
    fraction = np.clip(ys + 1.0 - (0.35 * xs + 0.2 * H), 0.0, 1.0).ravel()
    u = 3.0 * np.sin(np.arange((W + 1) * H) * 0.07)
    v = 4.0 * np.cos(np.arange(W * (H + 1)) * 0.11)
    lm, gm = two_phase_masses(fraction, cell_size)
    return Workload("sloped_pool", W, H, cell_size, [100.0], [0.1], 100.0, 1.0 / 120.0, fraction, u, v, lm, gm,
                    {"seed": seed})


def load_into(grid, wl: Workload):
    """Configure materials and upload a workload into any object with the FluidGridBase method set."""
    from .grid import FluidGridState, FluidMixture, MixtureComponent

    liquid = FluidMixture([MixtureComponent(f"liquid{k}", d) for k, d in enumerate(wl.liquid_density)])
    gas = FluidMixture([MixtureComponent(f"gas{k}", d) for k, d in enumerate(wl.gas_density)])
    grid.configureMaterials([], liquid, gas)
    grid.setGravityY(wl.gravity_y)
    N = wl.width * wl.height
    nl, ng = len(wl.liquid_density), len(wl.gas_density)
    grid.upload(FluidGridState(wl.fraction, wl.u, wl.v, np.zeros(0),
                               [wl.liquid_mass[k * N:(k + 1) * N] for k in range(nl)],
                               [wl.gas_mass[k * N:(k + 1) * N] for k in range(ng)]))


def pool_with_drop(W: int, H: int, h: float, fill: float, drop_radius: float, rng: SplitMix64):
    """Liquid pool over the bottom `fill` of the box plus one circular drop above it (cell-centre
    sampled fractions); returns the fraction field."""
    ys = (np.arange(H)[:, None] + 0.5) / H
    xs = (np.arange(W)[None, :] + 0.5) * h
    pool = (ys >= 1.0 - fill).astype(np.float64)
    cx = rng.uniform(1, 0.3, 0.7)[0] * W * h
    cy = (1.0 - fill) * 0.5 * H * h
    r = drop_radius * W * h
    drop = (np.hypot(xs - cx, ys * H * h - cy) < r).astype(np.float64)
    return np.maximum(pool, drop).ravel()


def serving_stream(seed: int = 5, count: int = 24, min_size: int = 128, max_size: int = 2048,
                   rate: float = 4.0, steps: int = 50):
    """Poisson request stream: `count` simulations arriving at `rate` per second, each side
    log-uniform in [min_size, max_size] (independent aspect), cell size 1/max(W, H), water 100 / air
    0.1, liquid filling 30-90% plus a drop, random vortex velocity, gravity 9.81, kinematic viscosity
    log-uniform in [1e-3, 1] (mu = nu rho per phase), surface tension U[0.01, 0.07], `steps` steps of
    dt = half the capillary step limit (one physical substep per step), each projection to a relative
    residual of 1e-6 (max|b - A p| <= 1e-6 max|b|, fs_set_relative_tolerance).  Returns
    [(arrival_s, kwargs)] where kwargs build a `fluidsim_amd.serving.Request`."""
    rng = SplitMix64(seed)
    out = []
    t = 0.0
    lo, hi = math.log(min_size), math.log(max_size)
    for k in range(count):
        t += -math.log1p(-rng.uniform(1)[0]) / rate
        W = int(round(math.exp(rng.uniform(1, lo, hi)[0])))
        H = int(round(math.exp(rng.uniform(1, lo, hi)[0])))
        h = 1.0 / max(W, H)
        fill = rng.uniform(1, 0.3, 0.9)[0]
        fraction = pool_with_drop(W, H, h, fill, 0.08, rng)
        nu = math.exp(rng.uniform(1, math.log(1e-3), math.log(1.0))[0])
        sigma = rng.uniform(1, 0.01, 0.07)[0]
        u = np.zeros((W + 1) * H)
        v = np.zeros(W * (H + 1))
        nv = 8
        cxs, cys = rng.uniform(nv, 0.0, W * h), rng.uniform(nv, 0.0, H * h)
        radii, strengths = rng.uniform(nv, 0.02, 0.08), rng.normal(nv) * 0.05
        for j in range(nv):
            add_gaussian(u, v, W, H, h, cxs[j], cys[j], radii[j], strengths[j])
        zero_walls(u, v, W, H)
        lm, gm = two_phase_masses(fraction, h)
        limit = 0.5 * math.sqrt((100.0 + 0.1) * h ** 3 / (2.0 * math.pi * sigma))
        out.append((t, dict(width=W, height=H, cell_size=h, steps=steps, dt=0.5 * limit, max_iterations=-1,
                            divergence_tolerance=1e-6, relative_tolerance=True,
                            gravity_y=9.81, surface_tension=sigma, liquid_density=[100.0], gas_density=[0.1],
                            liquid_viscosity=[nu * 100.0], gas_viscosity=[nu * 0.1], liquid_fraction=fraction,
                            u=u, v=v, liquid_mass=lm, gas_mass=gm)))
    return out


def batch_instances(seed: int = 4, count: int = 256, min_size: int = 64, max_size: int = 512):
    """Batched-solve instances: `count` grids with each side log-uniform in [min_size, max_size]
    (independent aspect), cell size 1/max(W, H), water 100 / air 0.1 with the liquid filling 30-90% of
    the height plus a drop, random-vortex velocity (8 vortices), a kinematic viscosity log-uniform in
    [1e-3, 1] (mu = nu rho per phase) and an obstacle mask of 8 circles and boxes covering a uniform
    0-25% of the cells (`obstacle_mask`; meta["solid"] is None below 1%).  Returns a list of Workload,
    one per instance, each from its own SplitMix64 substream."""
    out = []
    lo, hi = math.log(min_size), math.log(max_size)
    for k in range(count):
        rng = SplitMix64(seed * 1000003 + k)
        W = int(round(math.exp(rng.uniform(1, lo, hi)[0])))
        H = int(round(math.exp(rng.uniform(1, lo, hi)[0])))
        h = 1.0 / max(W, H)
        fill = rng.uniform(1, 0.3, 0.9)[0]
        fraction = pool_with_drop(W, H, h, fill, 0.08, rng)
        u = np.zeros((W + 1) * H)
        v = np.zeros(W * (H + 1))
        cxs, cys = rng.uniform(8, 0.0, W * h), rng.uniform(8, 0.0, H * h)
        radii, strengths = rng.uniform(8, 0.02, 0.08), rng.normal(8)
        for j in range(8):
            add_gaussian(u, v, W, H, h, cxs[j], cys[j], radii[j], strengths[j])
        zero_walls(u, v, W, H)
        nu = math.exp(rng.uniform(1, math.log(1e-3), math.log(1.0))[0])
        coverage = rng.uniform(1, 0.0, 0.25)[0]
        solid = obstacle_mask(W, H, rng, 8, coverage) if coverage >= 0.01 else None
        if solid is not None:
            fraction[solid != 0] = 0.0
        lm, gm = two_phase_masses(fraction, h)
        out.append(Workload("batch_instance", W, H, h, [100.0], [0.1], 9.81, 1.0 / 240.0, fraction, u, v, lm, gm,
                            {"seed": seed, "index": k, "nu": nu, "solid": solid, "coverage": coverage}))
    return out
