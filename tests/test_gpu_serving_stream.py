"""Config C5 on the GPU: the seeded serving stream through the serving loop (fs_server_*).

The stream is `workloads.serving_stream(seed=5)`: 24 requests with Poisson arrivals at 64/s, sides
log-uniform in 128..2048 (three of them above 1024^2 cells), 50 resident steps each of reference
transport, implicit viscosity (nu 1e-3..1), surface tension (0.01..0.07) and MG-PCG to a 1e-6 relative
residual, plus one request at the top of the size range (2048 x 384).  They are served by four lanes with
bfloat16 multigrid conductances, as the benchmark serves them, and:
  - every request finishes all 50 steps, each step accepted (the loop fails a request otherwise);
  - every result equals the same request stepped alone on a plain grid, bit for bit, with the same
    pressure and viscosity iteration counts;
  - the first steps of the four smallest requests are teacher-forced against the FP64 oracle (both sides
    start each step from the oracle's state): fractions bit for bit (transport is exact), velocity and
    pressure within 1e-3 relative L2 (the viscosity solve stops at 1e-4 relative in the mass norm, in FP32
    on the GPU against FP64 in the oracle), and both projections converge.
"""
import math
import time

import numpy as np
import pytest

from fluidsim_amd import FS_MG_BF16, FluidGridHIP, FluidGridState, FluidMixture, MixtureComponent
from fluidsim_amd import workloads as WL
from fluidsim_amd.serving import Request, Server

pytestmark = pytest.mark.gpu

STEPS = 50
RATE = 64.0


def top_of_range_request():
    """A request built like the stream's, with a 2048-cell side."""
    W, H = 2048, 384
    rng = WL.SplitMix64(5_000_011)
    h = 1.0 / max(W, H)
    fraction = WL.pool_with_drop(W, H, h, 0.6, 0.08, rng)
    u = np.zeros((W + 1) * H)
    v = np.zeros(W * (H + 1))
    cxs, cys = rng.uniform(8, 0.0, W * h), rng.uniform(8, 0.0, H * h)
    radii, strengths = rng.uniform(8, 0.02, 0.08), rng.normal(8) * 0.05
    for j in range(8):
        WL.add_gaussian(u, v, W, H, h, cxs[j], cys[j], radii[j], strengths[j])
    WL.zero_walls(u, v, W, H)
    lm, gm = WL.two_phase_masses(fraction, h)
    nu, sigma = 0.05, 0.04
    limit = 0.5 * math.sqrt((100.0 + 0.1) * h ** 3 / (2.0 * math.pi * sigma))
    return dict(width=W, height=H, cell_size=h, steps=STEPS, dt=0.5 * limit, max_iterations=-1,
                divergence_tolerance=1e-6, relative_tolerance=True, gravity_y=9.81, surface_tension=sigma,
                liquid_density=[100.0], gas_density=[0.1], liquid_viscosity=[nu * 100.0], gas_viscosity=[nu * 0.1],
                liquid_fraction=fraction, u=u, v=v, liquid_mass=lm, gas_mass=gm)


def stream():
    reqs = WL.serving_stream(5, 24, 128, 2048, RATE, STEPS)
    last = reqs[-1][0]
    reqs.append((last + 0.01, top_of_range_request()))
    return reqs


def alone(kw, precision):
    """The request stepped on a plain grid with the same settings."""
    W, H = kw["width"], kw["height"]
    g = FluidGridHIP(W, H, kw["cell_size"])
    g.configureMaterials([], FluidMixture([MixtureComponent("water", 100.0, kw["liquid_viscosity"][0])]),
                         FluidMixture([MixtureComponent("air", 0.1, kw["gas_viscosity"][0])]))
    g.setGravityY(kw["gravity_y"])
    g.setSurfaceTension(kw["surface_tension"])
    g.setRelativeTolerance(kw["relative_tolerance"])
    g.upload(FluidGridState(kw["liquid_fraction"], kw["u"], kw["v"], np.zeros(0), [kw["liquid_mass"]],
                            [kw["gas_mass"]]))
    g.setMultigridPrecision(precision)
    pressure = viscosity = 0
    for step in range(kw["steps"]):
        r = g.step(kw["dt"], kw["max_iterations"], kw["divergence_tolerance"])
        assert r.accepted(), f"{W}x{H} step {step} alone: {r}"
        pressure += r.iterations
        viscosity += r.viscosity_iterations
    s = g.download()
    g.destroy()
    return s, pressure, viscosity


@pytest.fixture(scope="module")
def served():
    reqs = stream()
    out = []
    with Server(lanes=4) as srv:
        t0 = time.perf_counter()
        ids = []
        for t, kw in reqs:
            delay = t0 + t - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            ids.append(srv.submit(Request(multigrid_precision=FS_MG_BF16, **kw)))
        srv.drain()
        stats = srv.stats()
        for (t, kw), rid in zip(reqs, ids):
            rep = srv.wait(rid)
            res = srv.result(rid, kw["width"], kw["height"]) if rep.done else None
            out.append((kw, rep, res))
    print(f"served {len(reqs)} requests: {stats['steps']} steps, {stats['steps_per_s']:.1f} steps/s, "
          f"step p50 {stats['step_p50_ms']:.2f} ms p99 {stats['step_p99_ms']:.2f} ms", flush=True)
    return out, stats


def test_every_request_finishes_every_step(served):
    out, stats = served
    assert len(out) == 25
    assert max(kw["width"] for kw, _, _ in out) == 2048
    assert sum(kw["width"] * kw["height"] >= 1024 * 1024 for kw, _, _ in out) >= 3
    for kw, rep, _ in out:
        assert rep.done, f"{kw['width']}x{kw['height']}: {rep.error}"
        assert rep.steps_done == STEPS and rep.physical_steps == STEPS
        assert rep.pressure_iterations > 0 and rep.viscosity_iterations > 0
    assert stats["completed"] == 25 and stats["failed"] == 0 and stats["steps"] == 25 * STEPS
    assert len({rep.lane for _, rep, _ in out}) == 4
    assert 0 < stats["step_p50_ms"] <= stats["step_p99_ms"]


@pytest.mark.parametrize("part", range(5))
def test_each_result_equals_the_request_alone(served, part):
    out, _ = served
    for kw, rep, res in out[part::5]:
        s, pressure, viscosity = alone(kw, FS_MG_BF16)
        name = f"{kw['width']}x{kw['height']}"
        assert rep.pressure_iterations == pressure, name
        assert rep.viscosity_iterations == viscosity, name
        for key, got in (("fraction", s.liquid_fraction), ("u", s.u), ("v", s.v), ("pressure", s.pressure),
                         ("liquid_mass", s.liquid_mass[0]), ("gas_mass", s.gas_mass[0])):
            assert np.array_equal(res[key], got), f"{name}: {key} differs between the mix and alone"


def rel_l2(a, b):
    n = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (n if n > 0 else 1.0)


def test_small_requests_teacher_forced_against_the_oracle(oracle):
    reqs = sorted((kw for _, kw in stream()), key=lambda kw: kw["width"] * kw["height"])[:4]
    for kw in reqs:
        W, H, h, N = kw["width"], kw["height"], kw["cell_size"], kw["width"] * kw["height"]
        o = oracle.OracleGrid(W, H, h)
        o.configure([100.0], [0.1])
        o.configure_viscosity(kw["liquid_viscosity"], kw["gas_viscosity"])
        o.set_surface_tension(kw["surface_tension"])
        o.set_gravity(kw["gravity_y"])
        o.set_relative_tolerance(True)
        o.upload(kw["liquid_fraction"], kw["u"], kw["v"], None, kw["liquid_mass"], kw["gas_mass"])
        g = FluidGridHIP(W, H, h)
        g.configureMaterials([], FluidMixture([MixtureComponent("water", 100.0, kw["liquid_viscosity"][0])]),
                             FluidMixture([MixtureComponent("air", 0.1, kw["gas_viscosity"][0])]))
        g.setGravityY(kw["gravity_y"])
        g.setSurfaceTension(kw["surface_tension"])
        g.setRelativeTolerance(True)
        for step in range(3):
            s = o.download()
            o.upload(s["fraction"], s["u"], s["v"], s["pressure"], s["liquid_mass"], s["gas_mass"])
            g.upload(FluidGridState(s["fraction"], s["u"], s["v"], s["pressure"], [s["liquid_mass"]],
                                    [s["gas_mass"]]))
            ro = o.step(kw["dt"], 100000, kw["divergence_tolerance"])
            rh = g.step(kw["dt"], kw["max_iterations"], kw["divergence_tolerance"])
            name = f"{W}x{H} step {step}"
            assert ro["converged"] and rh.accepted(), f"{name}: oracle {ro} hip {rh}"
            so, sh = o.download(), g.download()
            assert np.array_equal(sh.liquid_fraction, so["fraction"]), f"{name}: fractions"
            for key, got in (("u", sh.u), ("v", sh.v), ("pressure", sh.pressure)):
                e = rel_l2(got, so[key])
                print(f"{name}: {key} rel L2 {e:.2e}", flush=True)
                assert e < 1e-3, f"{name}: {key} rel L2 {e:.3e}"
        g.destroy()
