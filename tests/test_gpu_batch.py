"""Batched projection on the GPU (fs_batch_project): one persistent launch runs the PCG iterations of
many grids of different sizes, one workgroup per grid.

Each grid's result must follow fs_project's contract and agree with the single-grid projection and with
the FP64 oracle to the step tolerance (only the FP64 reduction order differs, so not bit for bit); a
grid's result must be the same alone and in any batch, bit for bit (every workgroup reads only its own
grid); diagonal and multigrid grids mix; malformed batches are rejected.
"""
import numpy as np
import pytest

from fluidsim_amd import (FS_PRECOND_DIAGONAL, FS_TRANSPORT_NONE, Context, FluidGridHIP, FluidGridState,
                          FluidMixture, MixtureComponent, batch_project)
from fluidsim_amd import workloads as WL

pytestmark = pytest.mark.gpu

SHAPES = [(17, 13), (1, 129), (131, 1), (64, 64), (65, 37), (128, 96), (200, 150), (300, 200), (2, 67), (96, 257)]


def pool_grid(W, H, seed, precond=None, context=None):
    rng = WL.SplitMix64(seed)
    h = 1.0 / max(W, H)
    f = WL.pool_with_drop(W, H, h, 0.3 + 0.6 * rng.uniform(1)[0], 0.1, rng)
    u = np.zeros((W + 1) * H)
    v = np.zeros(W * (H + 1))
    for _ in range(6):
        WL.add_gaussian(u, v, W, H, h, rng.uniform(1, 0, W * h)[0], rng.uniform(1, 0, H * h)[0],
                        rng.uniform(1, 0.02, 0.08)[0], 0.5 * rng.normal(1)[0])
    WL.zero_walls(u, v, W, H)
    lm, gm = WL.two_phase_masses(f, h)
    g = FluidGridHIP(W, H, h, context=context)
    g.configureMaterials([], FluidMixture([MixtureComponent("water", 100.0)]), FluidMixture([MixtureComponent("air", 0.1)]))
    g.setGravityY(9.81)
    g.setTransport(FS_TRANSPORT_NONE)
    if precond is not None:
        g.setPressurePreconditioner(precond)
    state = FluidGridState(f, u, v, np.zeros(0), [lm], [gm])
    g.upload(state)
    g.applyGravity(0.002)
    return g, state


def tau_of(g):
    return 1e-6 * g.maxDivergence()


def test_batch_matches_single_grid_projections_and_the_oracle(oracle):
    grids = [pool_grid(W, H, k)[0] for k, (W, H) in enumerate(SHAPES)]
    singles = [pool_grid(W, H, k)[0] for k, (W, H) in enumerate(SHAPES)]
    taus = [tau_of(g) for g in grids]
    # velocity scale of each problem: the largest face speed before the projection
    scales = []
    for g in grids:
        st = g.download()
        scales.append(max(np.abs(st.u).max(), np.abs(st.v).max()))
    res = batch_project(grids, 0.002, 2000, taus)
    for k, (g, s, r, tau) in enumerate(zip(grids, singles, res, taus)):
        rs = s.project(0.002, 2000, tau)
        assert r.converged and rs.converged, (SHAPES[k], r, rs)
        assert abs(r.iterations - rs.iterations) <= 3, (SHAPES[k], r.iterations, rs.iterations)
        a, b = g.download(), s.download()
        scale = scales[k]
        assert np.abs(a.u - b.u).max() <= 1e-4 * scale and np.abs(a.v - b.v).max() <= 1e-4 * scale, SHAPES[k]
        # the oracle on the same input
        W, H = SHAPES[k]
        _, st = pool_grid(W, H, k)
        o = oracle.OracleGrid(W, H, 1.0 / max(W, H))
        o.configure([100.0], [0.1])
        o.set_gravity(9.81)
        o.upload(st.liquid_fraction, st.u, st.v, None, st.liquid_mass[0], st.gas_mass[0])
        o.apply_gravity(0.002)
        ro = o.project(0.002, 2000, tau)
        so = o.download()
        assert ro["converged"]
        assert np.abs(a.u - so["u"]).max() <= 1e-4 * scale and np.abs(a.v - so["v"]).max() <= 1e-4 * scale
    for g in grids + singles:
        g.destroy()


def test_alone_equals_in_any_batch():
    probe, _ = pool_grid(96, 64, 42)
    ref = batch_project([probe], 0.002, 2000, tau_of(probe))[0]
    ref_state = probe.download()
    again, _ = pool_grid(96, 64, 42)
    others = [pool_grid(W, H, 100 + k)[0] for k, (W, H) in enumerate(SHAPES)]
    mixed = others[:4] + [again] + others[4:]
    out = batch_project(mixed, 0.002, 2000, [tau_of(g) for g in mixed])
    r = out[4]
    assert (r.iterations, r.max_divergence, r.restarts) == (ref.iterations, ref.max_divergence, ref.restarts)
    s = again.download()
    for k in ("u", "v", "pressure"):
        assert np.array_equal(getattr(s, k), getattr(ref_state, k)), k
    for g in others + [probe, again]:
        g.destroy()


def test_split_batch_runs_large_grids_through_the_single_grid_solve():
    """fs_context_set_batch_split: grids above the split get fs_project's own result bit for bit (the
    multi-kernel solve, run beside the persistent launch), the others the persistent launch's result
    bit for bit -- each grid's route depends only on its size."""
    ctx = Context(0)
    split = 10000
    grids = [pool_grid(W, H, 300 + k, context=ctx)[0] for k, (W, H) in enumerate(SHAPES)]
    taus = [tau_of(g) for g in grids]
    large = [W * H > split for W, H in SHAPES]
    assert 0 < sum(large) < len(SHAPES)
    ctx.set_batch_split(split)
    res = batch_project(grids, 0.002, 2000, taus)
    ctx.set_batch_split(0)
    for k, ((W, H), g, r, tau) in enumerate(zip(SHAPES, grids, res, taus)):
        ref_grid, _ = pool_grid(W, H, 300 + k, context=ctx)
        if large[k]:
            rr = ref_grid.project(0.002, 2000, tau)
        else:
            rr = batch_project([ref_grid], 0.002, 2000, tau)[0]
        assert r.converged and r.max_divergence <= tau, (W, H, r)
        assert (r.iterations, r.max_divergence) == (rr.iterations, rr.max_divergence), (W, H, r, rr)
        s, sr = g.download(), ref_grid.download()
        for name in ("u", "v", "pressure"):
            assert np.array_equal(getattr(s, name), getattr(sr, name)), (W, H, name)
        ref_grid.destroy()
    with pytest.raises(ValueError):
        ctx.set_batch_split(-1)
    for g in grids:
        g.destroy()
    ctx.close()


def test_single_stream_context_batches_every_grid():
    """A context without the helper stream (FS_CONTEXT_SINGLE_STREAM) ignores the batch split: a grid above
    it gets the persistent launch's result, bit for bit the one it gets alone in an unsplit batch."""
    ctx = Context(0, single_stream=True)
    W, H = SHAPES[6]                                  # 200 x 150 = 30000 cells, above the split below
    a, _ = pool_grid(W, H, 500, context=ctx)
    b, _ = pool_grid(W, H, 500, context=ctx)
    small = [pool_grid(w, h, 600 + k, context=ctx)[0] for k, (w, h) in enumerate(SHAPES[:3])]
    ctx.set_batch_split(10000)
    r = batch_project(small + [a], 0.002, 2000, [tau_of(g) for g in small + [a]])[-1]
    ctx.set_batch_split(0)
    rb = batch_project([b], 0.002, 2000, tau_of(b))[0]
    assert r.converged and (r.iterations, r.max_divergence) == (rb.iterations, rb.max_divergence)
    assert np.array_equal(a.download().pressure, b.download().pressure)
    for g in small + [a, b]:
        g.destroy()
    ctx.close()


@pytest.mark.parametrize("single,private", [(False, True), (True, True), (True, False)])
def test_context_flags_change_no_result(single, private):
    """FS_CONTEXT_SINGLE_STREAM (rebuild in line) and FS_CONTEXT_PRIVATE_QUEUE (the work stream on a
    hardware queue of its own, as the serving loop's lanes use) place work differently but compute the
    same: steps on such a context equal steps on the default context bit for bit; unknown flags are
    refused."""
    wl = WL.dam_break(96, 80, seed=1)
    ctx = Context(0, single_stream=single, private_queue=private)
    a = FluidGridHIP(wl.width, wl.height, wl.cell_size, context=ctx)
    b = FluidGridHIP(wl.width, wl.height, wl.cell_size)
    for g in (a, b):
        WL.load_into(g, wl)
    for _ in range(4):
        ra, rb = a.step(wl.dt, 300, 1e-6), b.step(wl.dt, 300, 1e-6)
        assert (ra.iterations, ra.max_divergence) == (rb.iterations, rb.max_divergence)
    sa, sb = a.download(), b.download()
    assert np.array_equal(sa.u, sb.u) and np.array_equal(sa.v, sb.v) and np.array_equal(sa.pressure, sb.pressure)
    a.destroy()
    b.destroy()
    ctx.close()
    import ctypes as C
    from fluidsim_amd._abi import lib
    h = C.c_void_p()
    assert lib().fs_context_create_ex(0, 4, C.byref(h)) != 0


def test_diagonal_and_multigrid_grids_mix():
    grids = [pool_grid(W, H, 7 + k, FS_PRECOND_DIAGONAL if k % 2 else None)[0] for k, (W, H) in
             enumerate([(40, 30), (33, 20), (64, 48), (50, 50)])]
    singles = [pool_grid(W, H, 7 + k, FS_PRECOND_DIAGONAL if k % 2 else None)[0] for k, (W, H) in
               enumerate([(40, 30), (33, 20), (64, 48), (50, 50)])]
    taus = [1e-5 * g.maxDivergence() for g in grids]
    res = batch_project(grids, 0.002, 5000, taus)
    for g, s, r, tau in zip(grids, singles, res, taus):
        rs = s.project(0.002, 5000, tau)
        assert r.converged and rs.converged
        assert abs(r.iterations - rs.iterations) <= max(3, rs.iterations // 20)
    for g in grids + singles:
        g.destroy()


def test_batch_contract_and_validation():
    a, _ = pool_grid(20, 16, 1)
    b, _ = pool_grid(24, 12, 2)
    zero = batch_project([a, b], 0.0, 100, 1e-6)
    assert all(not r.converged for r in zero)
    budget = batch_project([a, b], 0.002, 0, 1e-12)
    assert all(not r.converged for r in budget)
    assert batch_project([], 0.002, 100, 1e-6) == []
    with pytest.raises(ValueError):
        batch_project([a, a], 0.002, 100, 1e-6)
    other = FluidGridHIP(8, 8, 1.0, context=Context(0))
    other.configureMaterials([], FluidMixture([MixtureComponent("w", 100.0)]), FluidMixture([MixtureComponent("a", 0.1)]))
    other.upload(FluidGridState(np.ones(64)))
    with pytest.raises(ValueError):
        batch_project([a, other], 0.002, 100, 1e-6)
    fresh = FluidGridHIP(8, 8, 1.0)
    with pytest.raises(RuntimeError):
        batch_project([a, fresh], 0.002, 100, 1e-6)
    for g in (a, b, other, fresh):
        g.destroy()


def test_context_with_batch_buffers_is_released():
    """A context that ran a batched solve (descriptor buffer, pinned result slots) is destroyed cleanly."""
    ctx = Context(0)
    g = FluidGridHIP(40, 24, 1.0 / 40, context=ctx)
    g.configureMaterials([], FluidMixture([MixtureComponent("water", 100.0)]), FluidMixture([MixtureComponent("air", 0.1)]))
    f = np.clip(np.arange(24)[:, None] - 10.0 + 0 * np.arange(40)[None, :], 0.0, 1.0).ravel()
    lm, gm = WL.two_phase_masses(f, 1.0 / 40)
    g.upload(FluidGridState(f, np.zeros(41 * 24), np.zeros(40 * 25), np.zeros(0), [lm], [gm]))
    g.setGravityY(9.81)
    g.applyGravity(0.002)
    assert batch_project([g], 0.002, 500, 1e-6)[0].converged
    g.destroy()
    ctx.close()


def instance_grid(wl):
    g = FluidGridHIP(wl.width, wl.height, wl.cell_size)
    WL.load_into(g, wl)
    if wl.meta["solid"] is not None:
        g.setSolidMask(wl.meta["solid"])
    g.setTransport(FS_TRANSPORT_NONE)
    return g


def test_instances_with_obstacles_batch_like_single_grids():
    """The benchmark's instances (obstacles over 0-25% of the cells): batched results agree with the
    single-grid projections to the step tolerance, and a masked instance is the same alone and in the
    batch, bit for bit."""
    wls = WL.batch_instances(4, 10)
    assert sum(wl.meta["solid"] is not None for wl in wls) >= 8
    grids = [instance_grid(wl) for wl in wls]
    singles = [instance_grid(wl) for wl in wls]
    taus = [tau_of(g) for g in grids]
    dt = wls[0].dt
    res = batch_project(grids, dt, 2000, taus)
    for wl, g, s, r, tau in zip(wls, grids, singles, res, taus):
        rs = s.project(dt, 2000, tau)
        assert r.converged and rs.converged, (wl.width, wl.height, r, rs)
        assert abs(r.iterations - rs.iterations) <= 3, (wl.width, wl.height, r.iterations, rs.iterations)
        a, b = g.download(), s.download()
        scale = max(np.abs(wl.u).max(), np.abs(wl.v).max())
        assert np.abs(a.u - b.u).max() <= 1e-4 * scale and np.abs(a.v - b.v).max() <= 1e-4 * scale
    k = next(i for i, wl in enumerate(wls) if wl.meta["solid"] is not None)
    alone = instance_grid(wls[k])
    ra = batch_project([alone], dt, 2000, taus[k])[0]
    assert (ra.iterations, ra.max_divergence) == (res[k].iterations, res[k].max_divergence)
    for key in ("u", "v", "pressure"):
        assert np.array_equal(getattr(alone.download(), key), getattr(grids[k].download(), key)), key
    for g in grids + singles + [alone]:
        g.destroy()


def viscous_grid(W, H, seed, nu, mask=None, context=None):
    rng = WL.SplitMix64(seed)
    h = 1.0 / max(W, H)
    f = WL.pool_with_drop(W, H, h, 0.3 + 0.6 * rng.uniform(1)[0], 0.1, rng)
    u = np.zeros((W + 1) * H)
    v = np.zeros(W * (H + 1))
    for _ in range(6):
        WL.add_gaussian(u, v, W, H, h, rng.uniform(1, 0, W * h)[0], rng.uniform(1, 0, H * h)[0],
                        rng.uniform(1, 0.02, 0.08)[0], 0.5 * rng.normal(1)[0])
    WL.zero_walls(u, v, W, H)
    lm, gm = WL.two_phase_masses(f, h)
    g = FluidGridHIP(W, H, h, context=context)
    g.configureMaterials([], FluidMixture([MixtureComponent("water", 100.0, nu * 100.0)]),
                         FluidMixture([MixtureComponent("air", 0.1, nu * 0.1)]))
    g.upload(FluidGridState(f, u, v, np.zeros(0), [lm], [gm]))
    if mask is not None:
        g.setSolidMask(mask)
    return g


VISC_SHAPES = [(17, 13, 0.5), (1, 129, 0.01), (131, 1, 0.1), (64, 64, 1.0), (65, 37, 0.003), (200, 150, 0.05),
               (300, 200, 0.2), (2, 67, 0.3)]


def test_batched_viscosity_matches_single_solves():
    """fs_batch_apply_viscosity runs each grid's CG loop in one workgroup of one launch: each result meets
    the same stopping rule as fs_apply_viscosity on the same grid (iteration counts within 2, velocities
    within the 1e-4 solve tolerance), grids without viscosity are untouched, and a grid's result is the
    same alone and in the batch, bit for bit."""
    from fluidsim_amd import batch_apply_viscosity

    dt = 0.01
    grids = [viscous_grid(W, H, 60 + k, nu) for k, (W, H, nu) in enumerate(VISC_SHAPES)]
    singles = [viscous_grid(W, H, 60 + k, nu) for k, (W, H, nu) in enumerate(VISC_SHAPES)]
    inviscid = viscous_grid(40, 30, 99, 0.0)
    before = inviscid.download()
    its = batch_apply_viscosity(grids + [inviscid], dt)
    assert its[-1] == 0
    after = inviscid.download()
    assert np.array_equal(after.u, before.u) and np.array_equal(after.v, before.v)
    for k, (g, s) in enumerate(zip(grids, singles)):
        n = s.applyViscosity(dt)
        assert abs(its[k] - n) <= 2, (VISC_SHAPES[k], its[k], n)
        a, b = g.download(), s.download()
        scale = max(np.abs(b.u).max(), np.abs(b.v).max(), 1e-12)
        assert np.abs(a.u - b.u).max() <= 1e-3 * scale and np.abs(a.v - b.v).max() <= 1e-3 * scale, VISC_SHAPES[k]
    W, H, nu = VISC_SHAPES[5]
    alone = viscous_grid(W, H, 65, nu)
    assert batch_apply_viscosity([alone], dt) == [its[5]]
    assert np.array_equal(alone.download().u, grids[5].download().u)
    assert np.array_equal(alone.download().v, grids[5].download().v)
    for g in grids + singles + [inviscid, alone]:
        g.destroy()


def test_split_batched_viscosity_runs_large_grids_through_the_single_grid_solve():
    """With fs_context_set_batch_split, grids above the split get fs_apply_viscosity's result bit for bit
    (their own solve, beside the persistent launch) and the others the persistent launch's result bit
    for bit."""
    from fluidsim_amd import batch_apply_viscosity

    dt = 0.01
    ctx = Context(0)
    split = 10000
    grids = [viscous_grid(W, H, 60 + k, nu, context=ctx) for k, (W, H, nu) in enumerate(VISC_SHAPES)]
    large = [W * H > split for W, H, _ in VISC_SHAPES]
    assert 0 < sum(large) < len(VISC_SHAPES)
    ctx.set_batch_split(split)
    its = batch_apply_viscosity(grids, dt)
    ctx.set_batch_split(0)
    for k, ((W, H, nu), g) in enumerate(zip(VISC_SHAPES, grids)):
        ref = viscous_grid(W, H, 60 + k, nu, context=ctx)
        n = ref.applyViscosity(dt) if large[k] else batch_apply_viscosity([ref], dt)[0]
        assert its[k] == n, (W, H, its[k], n)
        a, b = g.download(), ref.download()
        assert np.array_equal(a.u, b.u) and np.array_equal(a.v, b.v), (W, H)
        ref.destroy()
    for g in grids:
        g.destroy()
    ctx.close()


def test_batched_viscosity_of_c4_instances_matches_the_oracle(oracle):
    """Config C4's instances (sides 64..512, nu log-uniform 1e-3..1): every instance's batched viscosity
    step agrees with the FP64 oracle's FluidGridGPU::applyViscosity restatement on the same state within
    3e-4 of the velocity scale (the solve stops at a 1e-4 relative residual in the mass norm; the GPU
    iterates in FP32), the bar of the single-grid viscosity tests."""
    from fluidsim_amd import batch_apply_viscosity

    wls = [wl for wl in WL.batch_instances(4, 40) if wl.width * wl.height <= 120_000][:12]
    assert len(wls) == 12 and max(wl.width * wl.height for wl in wls) > 60_000
    dt = wls[0].dt
    grids = []
    for wl in wls:
        nu = wl.meta["nu"]
        g = FluidGridHIP(wl.width, wl.height, wl.cell_size)
        g.configureMaterials([], FluidMixture([MixtureComponent("water", 100.0, nu * 100.0)]),
                             FluidMixture([MixtureComponent("air", 0.1, nu * 0.1)]))
        g.upload(FluidGridState(wl.fraction, wl.u, wl.v, np.zeros(0), [wl.liquid_mass], [wl.gas_mass]))
        grids.append(g)
    its = batch_apply_viscosity(grids, dt)
    for wl, g, n in zip(wls, grids, its):
        nu = wl.meta["nu"]
        o = oracle.OracleGrid(wl.width, wl.height, wl.cell_size)
        o.configure([100.0], [0.1])
        o.configure_viscosity([nu * 100.0], [nu * 0.1])
        o.upload(wl.fraction, wl.u, wl.v, None, wl.liquid_mass, wl.gas_mass)
        st, oit = o.apply_viscosity(dt)
        assert st == oracle.VISC_OK and n > 0, (wl.width, wl.height, st, n)
        so, sg = o.download(), g.download()
        scale = max(np.abs(wl.u).max(), np.abs(wl.v).max())
        err = max(np.abs(sg.u - so["u"]).max(), np.abs(sg.v - so["v"]).max())
        print(f"{wl.width}x{wl.height} nu {nu:.3g}: {n} GPU / {oit} oracle iterations, max error {err / scale:.2e} of scale",
              flush=True)
        assert err <= 3e-4 * scale, (wl.width, wl.height, err, scale)
        g.destroy()
