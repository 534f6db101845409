"""The serving loop on the GPU (fs_server_*): many resident simulations stepped concurrently.

  - a request's result is the same alone and in a mix of other requests running on other lanes, and
    equals the same steps taken on a plain FluidGridHIP, bit for bit (every instance owns its device state);
  - admission keeps the estimated resident footprint within the HBM budget, and every admitted request
    still completes (continuous batching: finishing one admits the next);
  - a request larger than the budget is refused at submission; a request whose projection is not
    accepted ends as failed with a message and does not disturb the others;
  - reports and statistics count what ran.
"""
import numpy as np
import pytest

from fluidsim_amd import FluidGridHIP, FluidGridState, FluidMixture, MixtureComponent
from fluidsim_amd import FS_TRANSPORT_REFERENCE, FS_TRANSPORT_SEMI_LAGRANGIAN
from fluidsim_amd import workloads as WL
from fluidsim_amd.serving import Request, Server

pytestmark = pytest.mark.gpu


def request(W, H, seed, steps=6, viscosity=0.0, sigma=0.0, transport=FS_TRANSPORT_REFERENCE, **kw):
    rng = WL.SplitMix64(seed)
    h = 1.0 / max(W, H)
    f = WL.pool_with_drop(W, H, h, 0.5, 0.1, rng)
    u = np.zeros((W + 1) * H)
    v = np.zeros(W * (H + 1))
    for j in range(4):
        WL.add_gaussian(u, v, W, H, h, rng.uniform(1, 0, W * h)[0], rng.uniform(1, 0, H * h)[0], 0.05,
                        0.2 * rng.normal(1)[0])
    WL.zero_walls(u, v, W, H)
    lm, gm = WL.two_phase_masses(f, h)
    args = dict(width=W, height=H, cell_size=h, steps=steps, dt=0.002, max_iterations=2000,
                divergence_tolerance=1e-5, gravity_y=9.81, surface_tension=sigma, transport=transport,
                liquid_viscosity=[viscosity * 100.0], gas_viscosity=[viscosity * 0.1], liquid_fraction=f, u=u, v=v,
                liquid_mass=lm, gas_mass=gm)
    args.update(kw)
    return Request(**args)


def direct(r: Request):
    """The same request stepped on a plain grid."""
    g = FluidGridHIP(r.width, r.height, r.cell_size)
    g.configureMaterials([], FluidMixture([MixtureComponent("water", 100.0, r.liquid_viscosity[0])]),
                         FluidMixture([MixtureComponent("air", 0.1, r.gas_viscosity[0])]))
    g.setGravityY(r.gravity_y)
    g.setSurfaceTension(r.surface_tension)
    g.setTransport(r.transport)
    g.upload(FluidGridState(r.liquid_fraction, r.u, r.v, np.zeros(0), [r.liquid_mass], [r.gas_mass]))
    iterations = 0
    for _ in range(r.steps):
        res = g.step(r.dt, r.max_iterations, r.divergence_tolerance)
        assert res.accepted()
        iterations += res.iterations
    s = g.download()
    g.destroy()
    return s, iterations


def same(a: dict, b: dict):
    return all(np.array_equal(a[k], b[k]) for k in ("fraction", "u", "v", "pressure", "liquid_mass", "gas_mass"))


def test_alone_equals_in_the_mix_equals_a_plain_grid():
    probe = request(72, 56, 1, viscosity=0.01, sigma=0.02)
    with Server(lanes=1) as alone:
        rid = alone.submit(probe)
        rep = alone.wait(rid)
        assert rep.done, rep.error
        ref = alone.result(rid, probe.width, probe.height)
    others = [request(96, 64, 2), request(33, 120, 3, viscosity=0.1), request(150, 90, 4, sigma=0.05),
              request(64, 64, 5, transport=FS_TRANSPORT_SEMI_LAGRANGIAN), request(128, 40, 6, viscosity=0.002),
              request(48, 48, 7), request(200, 150, 8, steps=3)]
    with Server(lanes=4) as mix:
        ids = [mix.submit(r) for r in others[:3]]
        pid = mix.submit(probe)
        ids += [mix.submit(r) for r in others[3:]]
        mix.drain()
        prep = mix.wait(pid)
        assert prep.done, prep.error
        assert prep.pressure_iterations == rep.pressure_iterations
        assert prep.viscosity_iterations == rep.viscosity_iterations
        assert same(mix.result(pid, probe.width, probe.height), ref), "result changed by the other requests"
        lanes = {mix.query(i).lane for i in ids + [pid]}
        assert len(lanes) == 4, "requests were not spread over the lanes"
        for i, r in zip(ids, others):
            q = mix.wait(i)
            assert q.done and q.steps_done == r.steps, q.error
        st = mix.stats()
        assert st["completed"] == len(others) + 1 and st["failed"] == 0
        assert st["steps"] == sum(r.steps for r in others) + probe.steps
        assert 0 < st["step_p50_ms"] <= st["step_p99_ms"]
    s, iterations = direct(probe)
    assert iterations == rep.pressure_iterations
    for k, key in (("liquid_fraction", "fraction"), ("u", "u"), ("v", "v"), ("pressure", "pressure")):
        assert np.array_equal(getattr(s, k), ref[key]), f"{k} differs from the plain grid"


def test_admission_respects_the_hbm_budget():
    reqs = [request(160, 120, 10 + k, steps=4) for k in range(6)]
    one = reqs[0].estimate_bytes()
    budget = int(2.5 * one)
    with Server(lanes=3, hbm_budget=budget) as srv:
        ids = [srv.submit(r) for r in reqs]
        srv.drain()
        st = srv.stats()
        assert st["budget_bytes"] == budget
        assert st["peak_resident_bytes"] <= budget
        assert st["peak_resident_bytes"] >= 2 * one          # two fit at once
        assert st["completed"] == 6 and st["resident"] == 0 and st["resident_bytes"] == 0
        for i in ids:
            rep = srv.wait(i)
            assert rep.done and rep.estimated_bytes == one
            assert 0 < rep.device_bytes <= rep.estimated_bytes   # the estimate bounds the allocation
        # later arrivals waited for memory
        assert max(srv.query(i).queue_ms for i in ids) > 0.0


def test_oversized_failed_and_invalid_requests():
    small = request(40, 32, 20, steps=3)
    with Server(lanes=2, hbm_budget=small.estimate_bytes() * 3) as srv:
        with pytest.raises(ValueError):
            srv.submit(request(1024, 1024, 21))          # larger than the whole budget
        with pytest.raises(ValueError):
            srv.submit(request(40, 32, 22, dt=0.0))
        with pytest.raises(ValueError):
            srv.submit(request(40, 32, 22, surface_tension=-1.0))
        bad = srv.submit(request(40, 32, 23, max_iterations=0))   # projection never accepted
        good = srv.submit(small)
        rb, rg = srv.wait(bad), srv.wait(good)
        assert not rb.done and rb.finished and "not accepted" in rb.error and rb.steps_done == 1
        assert rg.done and rg.steps_done == 3
        with pytest.raises(RuntimeError):
            srv.result(bad, 40, 32)
        srv.release(bad)
        with pytest.raises(ValueError):
            srv.query(bad)
        out = srv.result(good, 40, 32)
        assert np.isfinite(out["u"]).all() and out["fraction"].max() <= 1.0 + 1e-9
        st = srv.stats()
        assert st["failed"] == 1 and st["completed"] == 1


def test_requests_keep_arriving_while_others_run():
    """Continuous batching: a request submitted while the lanes are busy joins at the next step boundary
    and finishes before an earlier, longer request."""
    with Server(lanes=1) as srv:
        long_id = srv.submit(request(256, 256, 30, steps=40))
        q = srv.query(long_id)
        while q.steps_done < 2 and not q.finished:
            q = srv.wait(long_id, 0.001)
        short_id = srv.submit(request(32, 32, 31, steps=2))
        rs = srv.wait(short_id)
        rl = srv.query(long_id)
        assert rs.done and rs.lane == 0
        assert rl.steps_done < 40, "the short request waited for the long one to finish"
        assert srv.wait(long_id).done


def test_requests_with_obstacles_and_16bit_levels():
    """A request may carry an obstacle mask and a multigrid storage mode: its result equals the same steps
    on a plain grid configured alike, bit for bit; an invalid mask or mode is refused at submission."""
    from fluidsim_amd import FS_MG_BF16

    W, H = 96, 80
    wl = WL.obstacle_tank(W, H, seed=7)
    mask = wl.meta["solid"]
    probe = request(W, H, 40, steps=4, viscosity=0.01, sigma=0.02, multigrid_precision=FS_MG_BF16, solid=mask)
    with Server(lanes=2) as srv:
        with pytest.raises(ValueError):
            srv.submit(request(W, H, 41, solid=np.ones(W * H, np.uint8)))
        with pytest.raises(ValueError):
            srv.submit(request(W, H, 41, multigrid_precision=9))
        other = srv.submit(request(128, 64, 42, steps=3))
        rid = srv.submit(probe)
        rep = srv.wait(rid)
        assert rep.done, rep.error
        assert srv.wait(other).done
        out = srv.result(rid, W, H)
    g = FluidGridHIP(W, H, probe.cell_size)
    g.configureMaterials([], FluidMixture([MixtureComponent("water", 100.0, probe.liquid_viscosity[0])]),
                         FluidMixture([MixtureComponent("air", 0.1, probe.gas_viscosity[0])]))
    g.setGravityY(probe.gravity_y)
    g.setSurfaceTension(probe.surface_tension)
    g.upload(FluidGridState(probe.liquid_fraction, probe.u, probe.v, np.zeros(0), [probe.liquid_mass],
                            [probe.gas_mass]))
    g.setSolidMask(mask)
    g.setMultigridPrecision(FS_MG_BF16)
    iterations = 0
    for _ in range(probe.steps):
        r = g.step(probe.dt, probe.max_iterations, probe.divergence_tolerance)
        assert r.accepted()
        iterations += r.iterations
    s = g.download()
    g.destroy()
    assert iterations == rep.pressure_iterations
    for k, key in (("u", "u"), ("v", "v"), ("pressure", "pressure"), ("liquid_fraction", "fraction")):
        assert np.array_equal(getattr(s, k), out[key]), k


def test_requests_with_a_smoother_choice():
    """A request may carry pressure and viscosity smoother choices: its result equals the same steps on a plain
    grid with the same choices, bit for bit; an invalid choice is refused at submission."""
    from fluidsim_amd import FS_MG_BF16, FS_SMOOTHER_CHEBYSHEV

    probe = request(150, 110, 50, steps=4, viscosity=0.01, sigma=0.02, pressure_smoother=FS_SMOOTHER_CHEBYSHEV,
                    smoother_sweeps=3, coarse_sweeps=2, viscosity_sweeps=4)
    with Server(lanes=2) as srv:
        for bad in (dict(smoother_sweeps=9), dict(coarse_sweeps=5), dict(multigrid_precision=FS_MG_BF16),
                    dict(viscosity_sweeps=3)):
            with pytest.raises(ValueError):
                srv.submit(request(64, 64, 51, pressure_smoother=FS_SMOOTHER_CHEBYSHEV,
                                   **{"smoother_sweeps": 3, **bad}))
        rid = srv.submit(probe)
        other = srv.submit(request(96, 80, 52, steps=3))
        rep = srv.wait(rid)
        assert rep.done, rep.error
        assert srv.wait(other).done
        out = srv.result(rid, probe.width, probe.height)
    g = FluidGridHIP(probe.width, probe.height, probe.cell_size)
    g.configureMaterials([], FluidMixture([MixtureComponent("water", 100.0, probe.liquid_viscosity[0])]),
                         FluidMixture([MixtureComponent("air", 0.1, probe.gas_viscosity[0])]))
    g.setGravityY(probe.gravity_y)
    g.setSurfaceTension(probe.surface_tension)
    g.upload(FluidGridState(probe.liquid_fraction, probe.u, probe.v, np.zeros(0), [probe.liquid_mass],
                            [probe.gas_mass]))
    g.setPressureSmoother(FS_SMOOTHER_CHEBYSHEV, 3, 2)
    g.setViscositySmoother(4)
    iterations = viscous = 0
    for _ in range(probe.steps):
        r = g.step(probe.dt, probe.max_iterations, probe.divergence_tolerance)
        assert r.accepted()
        iterations += r.iterations
        viscous += r.viscosity_iterations
    s = g.download()
    g.destroy()
    assert iterations == rep.pressure_iterations
    assert viscous == rep.viscosity_iterations > 0
    for k, key in (("u", "u"), ("v", "v"), ("pressure", "pressure"), ("liquid_fraction", "fraction")):
        assert np.array_equal(getattr(s, k), out[key]), k
