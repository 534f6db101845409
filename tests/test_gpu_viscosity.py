"""GPU parity of the implicit viscosity solve (WebGPU-only in the reference: ViscositySolverGPU.cpp,
ViscosityKernels.h, FluidGridGPU::applyViscosity).

The HIP solver runs the reference's algorithm in FP32 with FP64 reductions and stops, like the reference,
when the recomputed residual meets r M^-1 r <= tol^2 v0 M v0 + 1e-12 sum(M) (tol 1e-4).  Checks: the
reference's own viscosity tests with their thresholds (analytic decay, independent strain-energy residual,
energy decay, rigid rotation, zero viscosity and zero flow, variable viscosity, non-finite input), the
volume-weighted material blend (2e-5, tests/FluidViscosityIntegrationTests.cpp:89-140), and agreement with
the FP64 oracle solving the same system to the same tolerance: max |v_gpu - v_oracle| <= 3e-4 max |v0|.
"""
import math

import numpy as np
import pytest

from fluidsim_amd import FluidGridHIP, FluidGridState, FluidMixture, MixtureComponent, viscosity_solve

from test_oracle_viscosity import Case, decay_case, strain_residual, variable_case, _stream_state

pytestmark = pytest.mark.gpu


def gpu_solve(case, dt, tol=1e-4, maximum=-1):
    it, rel, u, v = viscosity_solve(case.w, case.h, case.dx, case.material, case.u, case.v, dt, tol, maximum)
    return it, u, v


def oracle_close(O, case, dt, u, v, scale):
    st, it, rel, ou, ov = O.viscosity_solve(case.w, case.h, case.dx, case.material, case.u, case.v, dt, 1e-4, -1)
    assert st == O.VISC_OK
    err = max(np.abs(u - ou).max(), np.abs(v - ov).max())
    assert err <= 3e-4 * scale, f"GPU vs FP64 oracle {err:.3e} (scale {scale:.3e})"
    return err


@pytest.mark.parametrize("w,h,mu,dt,transverse", [(24, 16, 100, 0.02, False), (31, 23, 1000, 0.02, True),
                                                  (65, 49, 100000, 0.1, True), (47, 1, 1000, 0.02, False),
                                                  (2, 33, 1000, 0.02, False)])
def test_decay_matches_analytic_factor_and_oracle(oracle, w, h, mu, dt, transverse):
    """tests/FluidViscosityTests.cpp:179-215"""
    c, sx, sy = decay_case(w, h, mu, transverse)
    old = c.copy()
    it, c.u, c.v = gpu_solve(c, dt)
    c.walls()
    factor = 1.0 / (1.0 + dt * mu * ((sx * sx + sy * sy) if transverse else 2 * sx * sx) / (2 * c.dx * c.dx))
    err = max(np.abs(c.u - factor * old.u).max(), np.abs(c.v - factor * old.v).max())
    assert err < 3e-4
    assert strain_residual(old, c, dt) < 1.5e-4
    assert c.energy() <= old.energy() * (1 + 1e-6)
    assert it > 0 and it % 4 == 0
    oracle_close(oracle, old, dt, c.u, c.v, max(np.abs(old.u).max(), np.abs(old.v).max()))


def test_zero_viscosity_rotation_and_rest():
    """tests/FluidViscosityTests.cpp:217-254"""
    c = Case(25, 19, 2, 0)
    for y in range(c.h):
        for x in range(1, c.w):
            c.u[c.ui(x, y)] = 0.3 * (y - 8)
    for y in range(1, c.h):
        for x in range(c.w):
            c.v[c.vi(x, y)] = -0.3 * (x - 10)
    old = c.copy()
    it, u, v = gpu_solve(c, 0.01)
    assert it == 0 and np.array_equal(u, old.u) and np.array_equal(v, old.v)
    for y in range(4, c.h - 4):
        for x in range(4, c.w - 4):
            c.material[2 * (y * c.w + x) + 1] = 500
    it, u, v = gpu_solve(c, 0.01)
    assert max(np.abs(u - old.u).max(), np.abs(v - old.v).max()) < 3e-5
    c.u[:] = 0
    c.v[:] = 0
    it, u, v = gpu_solve(c, 0.01)
    assert it == 0 and not u.any() and not v.any()


@pytest.mark.parametrize("w,h", [(33, 25), (64, 48)])
def test_variable_viscosity_steps(oracle, w, h):
    """tests/FluidViscosityTests.cpp:256-289 (10 of 40 steps); each step also against the oracle."""
    c = variable_case(w, h)
    for step in range(10):
        old = c.copy()
        it, c.u, c.v = gpu_solve(c, 0.02)
        assert c.energy() <= old.energy() * (1 + 2e-6), f"step {step}"
        assert strain_residual(old, c, 0.02) < 1.8e-4, f"step {step}"
        oracle_close(oracle, old, 0.02, c.u, c.v, max(np.abs(old.u).max(), np.abs(old.v).max()))


def test_zero_dt_budget_and_nonfinite_input():
    """tests/FluidViscosityTests.cpp:291-344"""
    c = Case(33, 25, 1, 10000)
    for y in range(c.h):
        for x in range(1, c.w):
            c.u[c.ui(x, y)] = math.sin(0.7 * x + 0.3 * y)
    it, u, v = gpu_solve(c, 0.0)
    assert it == 0 and np.array_equal(u, c.u)
    with pytest.raises(RuntimeError):
        gpu_solve(c, 0.02, 1e-4, 1)
    old = c.copy()
    it, c.u, c.v = gpu_solve(c, 0.02)
    assert it > 1 and strain_residual(old, c, 0.02) < 1.8e-4
    bad = old.copy()
    bad.u[bad.ui(5, 5)] = float("nan")
    with pytest.raises(RuntimeError):
        gpu_solve(bad, 0.02)


def _mixture_grid(liquids, gases, W=24, H=18, h=0.7):
    g = FluidGridHIP(W, H, h)
    ld = [10.0 + i for i in range(liquids)]
    lv = [20.0 + 100.0 * i for i in range(liquids)]
    gd = [0.1 + 0.01 * i for i in range(gases)]
    g.configureMaterials([], FluidMixture([MixtureComponent(f"l{i}", d, mu) for i, (d, mu) in enumerate(zip(ld, lv))]),
                         FluidMixture([MixtureComponent(f"g{i}", d, 0.0) for i, d in enumerate(gd)]))
    g.setGravityY(0.0)
    f, u, v = _stream_state(W, H)
    area = h * h
    lm = [area * f * d / liquids for d in ld]
    gm = [area * (1 - f) * d / gases for d in gd]
    g.upload(FluidGridState(f, u, v, np.zeros(0), lm, gm))
    return g, ld, lv, gd, f, u, v, lm, gm


@pytest.mark.parametrize("liquids,gases", [(2, 2), (17, 3)])
def test_material_blend_and_grid_viscosity(oracle, liquids, gases):
    """tests/FluidViscosityIntegrationTests.cpp:89-159: the grid's blended solve equals a direct solve with
    the independently blended material within 2e-5; composition untouched; zero dt and zero viscosity are
    no-ops; gas-only viscosity activates the solver."""
    g, ld, lv, gd, f, u, v, lm, gm = _mixture_grid(liquids, gases)
    g.setComponentViscosity("gas", gases - 1, 5.0)
    gv = [0.0] * (gases - 1) + [5.0]
    W, H, h = g.width, g.height, g.cell_size
    N = W * H
    material = np.zeros(2 * N)
    for i in range(N):
        ms = [m[i] for m in lm] + [m[i] for m in gm]
        ds = ld + gd
        mus = lv + gv
        vol = sum(m / d for m, d in zip(ms, ds))
        material[2 * i] = sum(ms) / (h * h)
        material[2 * i + 1] = sum(m / d * mu for m, d, mu in zip(ms, ds, mus)) / vol
    it_ref, rel, eu, ev = viscosity_solve(W, H, h, material, u, v, 0.01)
    before = g.download()
    it = g.applyViscosity(0.01)
    assert it > 0
    after = g.download()
    assert max(np.abs(after.u - eu).max(), np.abs(after.v - ev).max()) < 2e-5
    assert np.array_equal(after.liquid_fraction, before.liquid_fraction)
    for a, b in zip(after.liquid_mass + after.gas_mass, before.liquid_mass + before.gas_mass):
        assert np.array_equal(a, b)
    # the FP64 oracle's grid-level apply on the same state
    o = oracle.OracleGrid(W, H, h)
    o.configure(ld, gd)
    o.configure_viscosity(lv, gv)
    o.upload(f, u, v, None, np.concatenate(lm), np.concatenate(gm))
    st, oit = o.apply_viscosity(0.01)
    assert st == oracle.VISC_OK
    so = o.download()
    scale = max(np.abs(u).max(), np.abs(v).max())
    assert max(np.abs(after.u - so["u"]).max(), np.abs(after.v - so["v"]).max()) <= 3e-4 * scale
    assert g.applyViscosity(0.0) == 0
    same = g.download()
    assert np.array_equal(same.u, after.u) and np.array_equal(same.v, after.v)
    for k in range(liquids):
        g.setComponentViscosity("liquid", k, 0.0)
    for k in range(gases):
        g.setComponentViscosity("gas", k, 0.0)
    assert g.applyViscosity(0.01) == 0
    assert np.array_equal(g.download().u, after.u)
    g.setComponentViscosity("gas", gases - 1, 5.0)
    assert g.applyViscosity(0.01) > 0
    with pytest.raises(ValueError):
        g.setComponentViscosity("gas", 0, -1.0)
    g.destroy()


def test_nonfinite_grid_state_is_rejected_and_uncommitted():
    g, *_ = _mixture_grid(2, 2)
    s = g.download()
    u = s.u.copy()
    u[5 * (g.width + 1) + 5] = float("nan")
    g.upload(FluidGridState(s.liquid_fraction, u, s.v, np.zeros(0), s.liquid_mass, s.gas_mass))
    with pytest.raises(RuntimeError):
        g.applyViscosity(0.01)
    after = g.download()
    assert np.array_equal(after.v, s.v)
    assert np.isnan(after.u[5 * (g.width + 1) + 5])
    g.destroy()


def test_viscous_steps_track_oracle(oracle):
    """Twelve viscous two-phase steps (transport, viscosity, projection) of the reference's continuation
    scenario (tests/FluidViscosityIntegrationTests.cpp:161-172), teacher-forced: each step starts both
    sides from the oracle's state; velocity within 1e-3 relative L2 (the viscosity tolerance is 1e-4
    relative in the mass norm), fractions and masses bit-identical (transport sees the same input)."""
    W, H, h = 24, 18, 1.0
    g, ld, lv, gd, f, u, v, lm, gm = _mixture_grid(2, 2, W, H, h)
    o = oracle.OracleGrid(W, H, h)
    o.configure(ld, gd)
    o.configure_viscosity(lv, [0.0, 0.0])
    o.set_gravity(0.0)
    o.upload(f, u, v, None, np.concatenate(lm), np.concatenate(gm))
    N = W * H
    for step in range(12):
        s = o.download()
        o.upload(s["fraction"], s["u"], s["v"], s["pressure"], s["liquid_mass"], s["gas_mass"])
        g.upload(FluidGridState(s["fraction"], s["u"], s["v"], s["pressure"],
                                [s["liquid_mass"][k * N:(k + 1) * N] for k in range(2)],
                                [s["gas_mass"][k * N:(k + 1) * N] for k in range(2)]))
        ro = o.step(0.004, 2000, 0.001)
        rh = g.step(0.004, 2000, 0.001)
        assert ro["converged"] and rh.converged, f"step {step}"
        assert rh.viscosity_iterations > 0 and ro["viscosity_iterations"] > 0
        so = o.download()
        sh = g.download()
        assert np.array_equal(sh.liquid_fraction, so["fraction"])
        for key, a in (("u", sh.u), ("v", sh.v)):
            e = np.linalg.norm(a - so[key]) / np.linalg.norm(so[key])
            assert e < 1e-3, f"step {step}: {key} rel L2 {e:.3e}"
    g.destroy()


@pytest.mark.parametrize("n", [2, 4, 6])
def test_chebyshev_viscosity_smoother_matches_oracle(oracle, n):
    """fs_set_viscosity_smoother(n): n Chebyshev-weighted Jacobi sweeps per side on every viscosity level (the
    LDS-tile kernel on the three finest levels of this 150 x 110 grid, ragged tiles at both edges, the
    one-workgroup tail below) against the FP64 oracle's orc_viscosity_solve_smoothed on the same state,
    to the 3e-4 bar of the reference's smoother; from four sweeps on no more iterations than the
    reference's smoother; 0 gives back the reference smoother's result bit for bit."""
    W, H, h, dt = 150, 110, 0.05, 0.05
    ref, ld, lv, gd, f, u, v, lm, gm = _mixture_grid(2, 2, W, H, h)
    it_ref = ref.applyViscosity(dt)
    base = ref.download()
    g, *_ = _mixture_grid(2, 2, W, H, h)
    g.setViscositySmoother(n)
    it = g.applyViscosity(dt)
    got = g.download()
    o = oracle.OracleGrid(W, H, h)
    o.configure(ld, gd)
    o.configure_viscosity(lv, [0.0, 0.0])
    o.set_viscosity_smoother(n)
    o.upload(f, u, v, None, np.concatenate(lm), np.concatenate(gm))
    st, oit = o.apply_viscosity(dt)
    assert st == oracle.VISC_OK
    so = o.download()
    scale = max(np.abs(u).max(), np.abs(v).max())
    err = max(np.abs(got.u - so["u"]).max(), np.abs(got.v - so["v"]).max())
    assert err <= 3e-4 * scale, f"GPU vs FP64 oracle {err:.3e} (scale {scale:.3e})"
    assert it_ref > 4 and abs(it - oit) <= 4, (it, it_ref, oit)
    if n >= 4:
        assert it <= it_ref, (it, it_ref)
    g.upload(FluidGridState(f, u, v, np.zeros(0), lm, gm))
    g.setViscositySmoother(0)
    assert g.applyViscosity(dt) == it_ref
    again = g.download()
    assert np.array_equal(again.u, base.u) and np.array_equal(again.v, base.v)
    with pytest.raises(ValueError):
        g.setViscositySmoother(3)
    g.destroy()
    ref.destroy()


def test_batched_chebyshev_viscosity_matches_oracle(oracle):
    """fs_batch_apply_viscosity honours each grid's smoother: a grid set to four Chebyshev sweeps and one on
    the reference's smoother, in one launch, each within the 3e-4 bar of the oracle with its smoother."""
    from fluidsim_amd import batch_apply_viscosity

    W, H, h, dt = 90, 70, 0.05, 0.01
    grids, states = [], []
    for n in (4, 0):
        g, ld, lv, gd, f, u, v, lm, gm = _mixture_grid(2, 2, W, H, h)
        g.setViscositySmoother(n)
        grids.append(g)
        states.append((n, f, u, v, lm, gm))
    its = batch_apply_viscosity(grids, dt)
    for g, it, (n, f, u, v, lm, gm) in zip(grids, its, states):
        o = oracle.OracleGrid(W, H, h)
        o.configure(ld, gd)
        o.configure_viscosity(lv, [0.0, 0.0])
        o.set_viscosity_smoother(n)
        o.upload(f, u, v, None, np.concatenate(lm), np.concatenate(gm))
        st, oit = o.apply_viscosity(dt)
        assert st == oracle.VISC_OK and abs(it - oit) <= 4, (n, it, oit)
        so, got = o.download(), g.download()
        scale = max(np.abs(u).max(), np.abs(v).max())
        assert max(np.abs(got.u - so["u"]).max(), np.abs(got.v - so["v"]).max()) <= 3e-4 * scale, n
        g.destroy()


@pytest.mark.parametrize("w,h,n", [(47, 1, 6), (2, 333, 4), (1100, 2, 6), (3, 700, 2)])
def test_chebyshev_viscosity_smoother_on_thin_grids(oracle, w, h, n):
    """Thin grids with the Chebyshev smoother: one-row and two-column fields through the LDS-tile kernel (the
    1100 x 2 and 3 x 700 fields' finest levels exceed the one-workgroup tail) and the tail, against the FP64
    oracle with the same sweeps, to the 3e-4 bar."""
    c, sx, sy = decay_case(w, h, 1000, h > 1 and w > 1)
    g = FluidGridHIP(w, h, c.dx)
    g.configureMaterials([], FluidMixture([MixtureComponent("l", 2.0, 1000.0)]),
                         FluidMixture([MixtureComponent("g", 2.0, 1000.0)]))
    g.setGravityY(0.0)
    f = np.full(w * h, 0.5)
    area = c.dx * c.dx
    lm, gm = [area * f * 2.0], [area * (1 - f) * 2.0]
    g.upload(FluidGridState(f, c.u, c.v, np.zeros(0), lm, gm))
    g.setViscositySmoother(n)
    it = g.applyViscosity(0.02)
    got = g.download()
    g.destroy()
    o = oracle.OracleGrid(w, h, c.dx)
    o.configure([2.0], [2.0])
    o.configure_viscosity([1000.0], [1000.0])
    o.set_viscosity_smoother(n)
    o.upload(f, c.u, c.v, None, lm[0], gm[0])
    st, oit = o.apply_viscosity(0.02)
    assert st == oracle.VISC_OK and abs(it - oit) <= 4, (it, oit)
    so = o.download()
    scale = max(np.abs(c.u).max(), np.abs(c.v).max(), 1e-12)
    err = max(np.abs(got.u - so["u"]).max(), np.abs(got.v - so["v"]).max())
    assert err <= 3e-4 * scale, f"GPU vs FP64 oracle {err:.3e} (scale {scale:.3e})"


def test_viscosity_smoother_tuner_keeps_state_and_sets_its_choice():
    """fs_tune_viscosity_smoother: picks one of 0, 2, 4, 6, leaves the velocities bit for bit as they were,
    and leaves the grid set to its choice (the next solve equals a grid set to it by hand, bit for bit); a
    grid without viscosity is refused."""
    W, H, h, dt = 150, 110, 0.05, 0.05
    g, *_ = _mixture_grid(2, 2, W, H, h)
    before = g.download()
    n, ms = g.tuneViscositySmoother(dt)
    assert n in (0, 2, 4, 6) and ms > 0.0
    after = g.download()
    assert np.array_equal(after.u, before.u) and np.array_equal(after.v, before.v)
    it = g.applyViscosity(dt)
    tuned = g.download()
    ref, *_ = _mixture_grid(2, 2, W, H, h)
    ref.setViscositySmoother(n)
    assert ref.applyViscosity(dt) == it
    same = ref.download()
    assert np.array_equal(same.u, tuned.u) and np.array_equal(same.v, tuned.v)
    g.destroy()
    ref.destroy()
    dry, *_ = _mixture_grid(1, 1, 24, 18, 0.7)
    dry.setComponentViscosity("liquid", 0, 0.0)
    with pytest.raises(RuntimeError):
        dry.tuneViscositySmoother(dt)
    dry.destroy()
