"""Eviction and restore under HBM pressure in the serving loop (fs_server_* with evict_after_steps).

With a budget that holds two instances and evict_after_steps = 2, four requests time-slice the device: each
admitted pair runs two steps, is checkpointed to host memory (fs_checkpoint_save), freed and queued again,
and is rebuilt from its checkpoint when readmitted.  Checked here:
  - every request finishes all its steps, the resident footprint never exceeds the budget, and the
    server's and the requests' eviction counts agree;
  - each result equals the same request stepped without interruption on a plain grid, bit for bit, with
    the same pressure and viscosity iteration counts -- also for a request with an obstacle mask, bfloat16
    multigrid conductances and the relative stopping rule, whose settings are not part of the checkpoint
    and are applied again after the restore;
  - with evict_after_steps = 0 nothing is evicted and later requests wait for memory instead.
"""
import numpy as np
import pytest

from fluidsim_amd import FS_MG_BF16, FluidGridHIP, FluidGridState, FluidMixture, MixtureComponent
from fluidsim_amd import workloads as WL
from fluidsim_amd.serving import Request, Server

pytestmark = pytest.mark.gpu

W, H = 160, 120


def request(seed, steps=8, nu=0.0, sigma=0.0, **kw):
    rng = WL.SplitMix64(seed)
    h = 1.0 / max(W, H)
    f = WL.pool_with_drop(W, H, h, 0.5, 0.1, rng)
    u = np.zeros((W + 1) * H)
    v = np.zeros(W * (H + 1))
    for _ in range(4):
        WL.add_gaussian(u, v, W, H, h, rng.uniform(1, 0, W * h)[0], rng.uniform(1, 0, H * h)[0], 0.05,
                        0.2 * rng.normal(1)[0])
    WL.zero_walls(u, v, W, H)
    lm, gm = WL.two_phase_masses(f, h)
    args = dict(width=W, height=H, cell_size=h, steps=steps, dt=0.0015, max_iterations=2000,
                divergence_tolerance=1e-5, gravity_y=9.81, surface_tension=sigma, liquid_viscosity=[nu * 100.0],
                gas_viscosity=[nu * 0.1], liquid_fraction=f, u=u, v=v, liquid_mass=lm, gas_mass=gm)
    args.update(kw)
    return Request(**args)


def plain(r: Request):
    """The request stepped without interruption on a plain grid configured as the server configures it."""
    g = FluidGridHIP(r.width, r.height, r.cell_size)
    g.configureMaterials([], FluidMixture([MixtureComponent("water", 100.0, r.liquid_viscosity[0])]),
                         FluidMixture([MixtureComponent("air", 0.1, r.gas_viscosity[0])]))
    g.setGravityY(r.gravity_y)
    g.setSurfaceTension(r.surface_tension)
    g.upload(FluidGridState(r.liquid_fraction, r.u, r.v, np.zeros(0), [r.liquid_mass], [r.gas_mass]))
    if r.solid is not None:
        g.setSolidMask(r.solid)
    g.setMultigridPrecision(r.multigrid_precision)
    g.setRelativeTolerance(bool(r.relative_tolerance))
    pressure = viscosity = 0
    for _ in range(r.steps):
        res = g.step(r.dt, r.max_iterations, r.divergence_tolerance)
        assert res.accepted()
        pressure += res.iterations
        viscosity += res.viscosity_iterations
    s = g.download()
    g.destroy()
    return s, pressure, viscosity


def requests():
    mask = WL.obstacle_tank(W, H, seed=9).meta["solid"]
    return [request(50, nu=0.01), request(51, sigma=0.03),
            request(52, nu=0.05, sigma=0.02, solid=mask, multigrid_precision=FS_MG_BF16, relative_tolerance=True,
                    divergence_tolerance=1e-6, max_iterations=-1),
            request(53, nu=0.002)]


def serve(reqs, evict_after):
    one = max(r.estimate_bytes() for r in reqs)
    budget = int(2.5 * one)   # two instances at a time
    with Server(lanes=2, hbm_budget=budget, evict_after_steps=evict_after) as srv:
        ids = [srv.submit(r) for r in reqs]
        srv.drain()
        reps = [srv.wait(i) for i in ids]
        outs = [srv.result(i, r.width, r.height) for i, r in zip(ids, reqs)]
        st = srv.stats()
    return budget, reps, outs, st


def test_evicted_requests_resume_bit_for_bit():
    reqs = requests()
    budget, reps, outs, st = serve(reqs, 2)
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert st["peak_resident_bytes"] <= budget
    assert st["evictions"] >= 2 and sum(rep.evictions for rep in reps) == st["evictions"]
    assert sum(rep.evictions > 0 for rep in reps) >= 2
    for k, (rep, out, r) in enumerate(zip(reps, outs, reqs)):
        assert rep.done and rep.steps_done == r.steps, (k, rep.error)
        s, pressure, viscosity = plain(r)
        assert (rep.pressure_iterations, rep.viscosity_iterations) == (pressure, viscosity), k
        for key, got in (("fraction", s.liquid_fraction), ("u", s.u), ("v", s.v), ("pressure", s.pressure),
                         ("liquid_mass", s.liquid_mass[0]), ("gas_mass", s.gas_mass[0])):
            assert np.array_equal(out[key], got), f"request {k}: {key} differs after eviction"


def test_no_eviction_without_the_setting():
    reqs = requests()
    budget, reps, outs, st = serve(reqs, 0)
    assert st["completed"] == len(reqs) and st["evictions"] == 0
    assert all(rep.evictions == 0 for rep in reps)
    assert st["peak_resident_bytes"] <= budget
    assert max(rep.queue_ms for rep in reps) > 0.0
