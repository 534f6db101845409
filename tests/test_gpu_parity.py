"""GPU parity: the HIP backend (through the C ABI) against the FP64 oracle on the same seeded inputs.

Bars: bit-exact for face weights, multigrid level shapes, centres, transfer maps, FP64 conductances and
the coarse Cholesky factor; pressure and velocity within 1e-5 relative L2 at the same stopping rule;
the reference's own absolute tolerances where its tests state one (2e-6 velocity for a different
preconditioner at tol 1e-8, tests/FluidGridTests.cpp:627-660).
"""
import math

import numpy as np
import pytest

from fluidsim_amd import (FS_PRECOND_DIAGONAL, FS_TRANSPORT_NONE, FS_TRANSPORT_SEMI_LAGRANGIAN, FluidGridHIP,
                          FluidGridState, FluidMixture, MixtureComponent)
from fluidsim_amd import workloads as WL

pytestmark = pytest.mark.gpu

REL_L2 = 1e-5


def rel_l2(a, b):
    n = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (n if n > 0 else 1.0)


def oracle_grid(O, wl, preconditioner="multigrid"):
    g = O.OracleGrid(wl.width, wl.height, wl.cell_size)
    g.configure(wl.liquid_density, wl.gas_density)
    g.set_gravity(wl.gravity_y)
    g.set_preconditioner(preconditioner)
    g.upload(wl.fraction, wl.u, wl.v, None, wl.liquid_mass, wl.gas_mass)
    return g


def hip_grid(wl):
    g = FluidGridHIP(wl.width, wl.height, wl.cell_size)
    WL.load_into(g, wl)
    return g


def test_dam_break_projection_matches_oracle(oracle):
    wl = WL.dam_break(128, 128, seed=1)
    o, h = oracle_grid(oracle, wl), hip_grid(wl)
    ro = o.force_project(wl.dt, 200, 1e-6)
    h.applyGravity(wl.dt)
    rh = h.project(wl.dt, 200, 1e-6)
    assert ro["converged"] and rh.converged
    so, sh = o.download(), h.download()
    assert rel_l2(sh.u, so["u"]) < REL_L2
    assert rel_l2(sh.v, so["v"]) < REL_L2
    assert rel_l2(sh.pressure, so["pressure"]) < REL_L2
    h.destroy()


@pytest.mark.parametrize("w,h", [(128, 128), (65, 37), (1, 129), (131, 1), (13, 9), (300, 200), (2, 67), (1, 1),
                                 (4000, 3), (2, 3000), (1000, 700)])
def test_hierarchy_is_bit_exact(oracle, w, h):
    wl = WL.sloped_pool(w, h, 4.0)
    o = oracle_grid(oracle, wl)
    g = hip_grid(wl)
    iu_o, iv_o = o.face_weights()
    iu_h, iv_h = g.faceWeights()
    assert np.array_equal(iu_o, iu_h) and np.array_equal(iv_o, iv_h)
    mg = oracle.Multigrid(w, h)
    assert mg.rebuild(iu_o, iv_o)
    assert g.multigridLevels() == mg.levels
    for l in range(mg.levels):
        assert g.multigridShape(l) == mg.shape(l)
        cxo, cyo = mg.centers(l)
        cxh, cyh = g.multigridCenters(l)
        assert np.array_equal(cxo, cxh) and np.array_equal(cyo, cyh)
        ro, do, _ = mg.weights(l)
        rh, dh, _ = g.multigridWeights(l)
        assert np.array_equal(ro, rh), f"right conductances differ on level {l}"
        assert np.array_equal(do, dh), f"down conductances differ on level {l}"
        if l + 1 < mg.levels:
            for axis in (0, 1):
                for a, b in zip(mg.prolong(l, axis), g.multigridProlong(l, axis)):
                    assert np.array_equal(a, b)
                for a, b in zip(mg.restriction(l, axis), g.multigridRestrict(l, axis)):
                    assert np.array_equal(a, b)
    assert np.array_equal(mg.coarse_factor(), g.multigridCoarseFactor())
    g.destroy()


@pytest.mark.parametrize("w,h", [(1, 129), (131, 1), (2, 67), (13, 9), (64, 64), (65, 37), (256, 192)])
def test_gpu_multigrid_is_a_valid_preconditioner(w, h):
    """tests/FluidGridWebGPUTests.cpp:428-535 bars for an FP32 V-cycle: linear within 5e-4, symmetric,
    positive, constants annihilated within 5e-6."""
    wl = WL.sloped_pool(w, h, 4.0)
    g = hip_grid(wl)
    i = np.arange(w * h)
    a = np.sin(i * 0.31) + np.cos(i * 0.073)
    b = np.cos(i * 0.17) - np.sin(i * 0.019)
    a -= a.mean()
    b -= b.mean()
    ma, mb = g.multigridApply(a), g.multigridApply(b)
    mc = g.multigridApply(1.3 * a - 0.7 * b)
    assert np.array_equal(ma, g.multigridApply(a)), "V-cycle is not history-free"
    ma, mb, mc = ma - ma.mean(), mb - mb.mean(), mc - mc.mean()
    scale = np.abs(mc).max()
    assert np.abs(mc - (1.3 * ma - 0.7 * mb)).max() <= 5e-4 * max(scale, 1e-30)
    ea, eb = a @ ma, b @ mb
    assert ea > 0 and eb > 0
    assert abs(a @ mb - b @ ma) <= 5e-4 * math.sqrt(ea * eb)
    const = g.multigridApply(np.full(w * h, 7.0), 7.0)
    assert np.abs(const - const.mean()).max() <= 5e-6 * max(1.0, np.abs(ma).max())
    g.destroy()


@pytest.mark.parametrize("w,h", [(17, 13), (1, 129), (131, 1), (2, 67), (13, 9), (65, 47), (128, 128)])
@pytest.mark.parametrize("precond", ["multigrid", "diagonal"])
def test_sloped_pool_projection_matches_oracle(oracle, w, h, precond):
    wl = WL.sloped_pool(w, h, 4.0)
    o = oracle_grid(oracle, wl)
    g = hip_grid(wl)
    # FP32 diagonal PCG cannot reach 1e-8 on the thin, 1000:1-density strips (condition number ~1e7,
    # beyond FP32); like the reference's GPU tests it is held to the tolerance FP32 reaches.
    tol, bound = (1e-8, 2e-6) if precond == "multigrid" else (1e-6, 2e-4)
    if precond == "diagonal":
        g.setPressurePreconditioner(FS_PRECOND_DIAGONAL)
    ro = o.project(wl.dt, 4000, 1e-8)
    rh = g.project(wl.dt, 4000, tol)
    assert ro["converged"] and rh.converged, (ro, rh)
    so, sh = o.download(), g.download()
    assert np.abs(sh.u - so["u"]).max() < bound
    assert np.abs(sh.v - so["v"]).max() < bound
    g.destroy()


def test_vortex_box_1024_relative_residual_and_parity(oracle):
    wl = WL.vortex_box(1024, 1024, seed=2)
    g = hip_grid(wl)
    tau = 1e-6 * g.maxDivergence()
    rh = g.project(wl.dt, 200, tau)
    assert rh.converged
    assert rh.final_residual <= 1e-6 * rh.rhs_max
    o = oracle_grid(oracle, wl)
    ro = o.project(wl.dt, 200, tau)
    assert ro["converged"]
    so, sh = o.download(), g.download()
    assert rel_l2(sh.u, so["u"]) < REL_L2
    assert rel_l2(sh.v, so["v"]) < REL_L2
    assert rel_l2(sh.pressure, so["pressure"]) < REL_L2
    # a size-independent property at full size: the corrected field is divergence-free to tau
    assert g.maxDivergence() <= tau
    g.destroy()


def test_hydrostatic_equilibrium_on_gpu():
    """tests/FluidGridTests.cpp:140-160 closed form: p(H-1) - p(0) = rho g h (H-1)."""
    W, H, h = 12, 10, 2.0
    f = np.ones(W * H)
    lm, gm = WL.two_phase_masses(f, h)
    wl = WL.Workload("filled", W, H, h, [100.0], [0.1], 100.0, 1 / 120, f, np.zeros((W + 1) * H),
                     np.zeros(W * (H + 1)), lm, gm)
    g = hip_grid(wl)
    g.setTransport(FS_TRANSPORT_NONE)
    r = g.step(1 / 120, 2000, 1e-8)
    assert r.converged
    s = g.download()
    assert np.abs(s.v).max() < 1e-7
    rise = s.pressure[(H - 1) * W] - s.pressure[0]
    expected = 100.0 * 100.0 * h * (H - 1)
    assert abs(rise - expected) < expected * 1e-8
    g.destroy()


def test_zero_dt_and_zero_budget():
    """tests/FluidGridTests.cpp:421-442 and tests/FluidGridAPITests.cpp:218-222."""
    wl = WL.dam_break(8, 8, seed=3)
    g = hip_grid(wl)
    g.setTransport(FS_TRANSPORT_NONE)
    before = g.download()
    r = g.step(0.0)
    assert not r.converged and r.physical_steps == 0
    after = g.download()
    assert np.array_equal(before.u, after.u) and np.array_equal(before.v, after.v)
    assert not g.step(0.01, 0, 1e-8).converged
    g.destroy()


def test_divergence_free_field_needs_no_iterations():
    W, H, h = 24, 16, 0.5
    xs = np.arange(W + 1) * h
    ys = np.arange(H + 1) * h
    X, Y = np.meshgrid(xs, ys)
    psi = np.sin(X * 0.7) * np.sin(Y * 0.9) * (X * (W * h - X)) * (Y * (H * h - Y))
    u = ((psi[1:, :] - psi[:-1, :]) / h).ravel()
    v = (-(psi[:, 1:] - psi[:, :-1]) / h).ravel()
    f = np.ones(W * H)
    lm, gm = WL.two_phase_masses(f, h)
    wl = WL.Workload("df", W, H, h, [100.0], [0.1], 0.0, 0.01, f, u, v, lm, gm)
    g = hip_grid(wl)
    g.upload(FluidGridState(f, u, v, np.full(W * H, 3.0), [lm], [gm]))
    r = g.project(0.01, 200, 1e-6)
    s = g.download()
    assert r.iterations == 0 and r.converged
    assert np.all(s.pressure == 0.0)
    assert np.array_equal(s.u, u) and np.array_equal(s.v, v)
    g.destroy()


def test_upload_download_round_trip_and_errors():
    wl = WL.dam_break(16, 12, seed=5)
    g = FluidGridHIP(16, 12, wl.cell_size)
    with pytest.raises(RuntimeError):
        g.project(0.01)  # use before upload is a logic error
    WL.load_into(g, wl)
    p = np.linspace(-1, 1, 16 * 12)
    g.upload(FluidGridState(wl.fraction, wl.u, wl.v, p, [wl.liquid_mass], [wl.gas_mass]))
    s = g.download()
    assert np.array_equal(s.u, wl.u) and np.array_equal(s.v, wl.v) and np.array_equal(s.pressure, p)
    assert np.array_equal(s.liquid_fraction, wl.fraction) and np.array_equal(s.liquid_mass[0], wl.liquid_mass)
    with pytest.raises(ValueError):
        g.upload(FluidGridState(np.zeros(5)))
    with pytest.raises(ValueError):
        g.upload(FluidGridState(wl.fraction, liquid_mass=[wl.liquid_mass, wl.liquid_mass]))
    g.destroy()


def test_repeated_solves_are_bit_identical():
    wl = WL.vortex_box(256, 256, seed=9)
    g = hip_grid(wl)
    g.snapshotSave()
    out = []
    for _ in range(2):
        g.snapshotRestore()
        r = g.project(wl.dt, 200, 1e-6 * 1e3)
        s = g.download()
        out.append((r.iterations, s.u.copy(), s.v.copy(), s.pressure.copy()))
    assert out[0][0] == out[1][0]
    for a, b in zip(out[0][1:], out[1][1:]):
        assert np.array_equal(a, b)
    g.destroy()


def test_semi_lagrangian_transport_matches_oracle_bitwise(oracle):
    wl = WL.vortex_box(96, 80, seed=4, vortices=12, sources=0)
    o, g = oracle_grid(oracle, wl), hip_grid(wl)
    o.advect_semi_lagrangian(wl.dt)
    g.advectSemiLagrangian(wl.dt)
    so, sh = o.download(), g.download()
    assert np.array_equal(so["u"], sh.u)
    assert np.array_equal(so["v"], sh.v)
    assert np.array_equal(so["fraction"], sh.liquid_fraction)
    assert np.array_equal(so["liquid_mass"], sh.liquid_mass[0])


def test_semi_lagrangian_steps_track_oracle(oracle):
    wl = WL.dam_break(64, 64, seed=1)
    o, g = oracle_grid(oracle, wl), hip_grid(wl)
    g.setTransport(FS_TRANSPORT_SEMI_LAGRANGIAN)
    for _ in range(20):
        ro = o.step_semi_lagrangian(wl.dt, 200, 1e-6)
        rh = g.step(wl.dt, 200, 1e-6)
        assert ro["converged"] and rh.converged
    so, sh = o.download(), g.download()
    assert rel_l2(sh.u, so["u"]) < 1e-4
    assert rel_l2(sh.v, so["v"]) < 1e-4
    g.destroy()


def test_gravity_and_vortex_edits_match_oracle(oracle):
    wl = WL.dam_break(40, 30, seed=2)
    o, g = oracle_grid(oracle, wl), hip_grid(wl)
    o.set_origin(0.25, -0.5)
    g.setOrigin((0.25, -0.5))
    o.apply_gravity(0.01)
    g.applyGravity(0.01)
    o.add_vortex(0.4, -0.3, 0.05, 1.7)
    g.addVortex((0.4, -0.3), 0.05, 1.7)
    so, sh = o.download(), g.download()
    assert np.allclose(sh.u, so["u"], rtol=1e-14, atol=1e-15)
    assert np.allclose(sh.v, so["v"], rtol=1e-14, atol=1e-15)
    assert o.max_divergence() == pytest.approx(g.maxDivergence(), rel=1e-12)
    g.destroy()


@pytest.mark.parametrize("index", [3, 11, 2])
def test_high_contrast_pool_with_drop_converges_like_the_oracle(oracle, index):
    """A 1000:1 liquid pool plus a drop (the serving stream's requests 3, 11, 2: 314x132, 271x159,
    343x254): the FP32 recurrence loses accuracy after a 1e3 reduction, so the FP64 refinement check must
    restart CG from the true residual; the projection then converges like the FP64 oracle's."""
    t, kw = WL.serving_stream(5, 16, 128, 2048, 4.0, 10)[index]
    W, H, h, dt = kw["width"], kw["height"], kw["cell_size"], kw["dt"]
    g = FluidGridHIP(W, H, h)
    g.configureMaterials([], FluidMixture([MixtureComponent("water", 100.0)]), FluidMixture([MixtureComponent("air", 0.1)]))
    g.setGravityY(9.81)
    g.upload(FluidGridState(kw["liquid_fraction"], kw["u"], kw["v"], np.zeros(0), [kw["liquid_mass"]],
                            [kw["gas_mass"]]))
    o = oracle.OracleGrid(W, H, h)
    o.configure([100.0], [0.1])
    o.set_gravity(9.81)
    o.upload(kw["liquid_fraction"], kw["u"], kw["v"], None, kw["liquid_mass"], kw["gas_mass"])
    g.applyGravity(dt)
    o.apply_gravity(dt)
    rg = g.project(dt, 200, 1e-4)
    ro = o.project(dt, 200, 1e-4)
    assert ro["converged"] and rg.converged and rg.iterations <= 80, (rg, ro["iterations"])
    sg, so = g.download(), o.download()
    scale = max(np.abs(so["u"]).max(), np.abs(so["v"]).max())
    assert np.abs(sg.u - so["u"]).max() < 1e-4 * scale and np.abs(sg.v - so["v"]).max() < 1e-4 * scale
    g.destroy()


def test_every_tile_variant_computes_the_same_values():
    """The autotuner only picks among tilings of identical arithmetic: each smoother and restriction variant
    gives a bit-identical V-cycle and projection."""
    wl = WL.sloped_pool(300, 200, 4.0)
    g = hip_grid(wl)
    i = np.arange(300 * 200)
    a = np.sin(i * 0.31) + np.cos(i * 0.073)
    a -= a.mean()
    levels = [l for l in range(g.multigridLevels()) if g.multigridTiles(l)[0] >= 0]
    assert levels, "no tiled level"
    g.setMultigridTiles(levels[0], 0, 0, 0)
    ref = g.multigridApply(a)
    for l in levels:
        for v in range(11):
            g.setMultigridTiles(l, v, v, v % 4)
            assert g.multigridTiles(l) == (v, v, v % 4)
            assert np.array_equal(g.multigridApply(a), ref), f"level {l} variant {v}"
        g.setMultigridTiles(l, 0, 0, 0)
    g.snapshotSave()
    r0 = g.project(wl.dt, 2000, 1e-6)
    p0 = g.download().pressure
    g.snapshotRestore()
    for l in levels:
        g.setMultigridTiles(l, 7, 6, 3)
    r1 = g.project(wl.dt, 2000, 1e-6)
    assert r0.iterations == r1.iterations and np.array_equal(g.download().pressure, p0)
    with pytest.raises(ValueError):
        g.setMultigridTiles(levels[0], 11, 0, 0)
    g.destroy()


def test_semi_lagrangian_c2_steps_match_oracle_teacher_forced(oracle):
    """C2's stepped run at 512^2 (vortex box, semi-Lagrangian transport, gravity, projection to a 1e-6
    relative residual in every step): ten steps, each started on both sides from the oracle's state, agree
    within the 1e-5 relative L2 bar (u, v and the mean-free pressure); the transport itself is bitwise."""
    wl = WL.vortex_box(512, 512, seed=2)
    o, g = oracle_grid(oracle, wl), hip_grid(wl)
    g.setTransport(FS_TRANSPORT_SEMI_LAGRANGIAN)
    g.setRelativeTolerance(True)
    o.set_relative_tolerance(True)
    oracle.set_threads(8)
    try:
        for step in range(10):
            s = o.download()
            g.upload(FluidGridState(s["fraction"], s["u"], s["v"], s["pressure"], [s["liquid_mass"]],
                                    [s["gas_mass"]]))
            o.upload(s["fraction"], s["u"], s["v"], s["pressure"], s["liquid_mass"], s["gas_mass"])
            ro = o.step_semi_lagrangian(wl.dt, 200, 1e-6)
            rh = g.step(wl.dt, 200, 1e-6)
            assert ro["converged"] and rh.converged, (step, ro, rh)
            so, sh = o.download(), g.download()
            po, ph = so["pressure"] - so["pressure"].mean(), sh.pressure - sh.pressure.mean()
            errs = (rel_l2(sh.u, so["u"]), rel_l2(sh.v, so["v"]), rel_l2(ph, po))
            print(f"step {step}: iterations GPU {rh.iterations} oracle {ro['iterations']}, rel L2 u/v/p "
                  + " / ".join(f"{e:.1e}" for e in errs))
            assert max(errs) < REL_L2, (step, errs)
    finally:
        oracle.set_threads(1)
    g.destroy()
