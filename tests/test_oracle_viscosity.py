"""Pins for the oracle's implicit viscosity solve (oracle/fs_viscosity.c) -- CPU only.

The reference's viscosity exists only as WebGPU shaders (ViscositySolverGPU.cpp, ViscosityKernels.h),
so the FP64 restatement is pinned by the reference's own viscosity tests (tests/FluidViscosityTests.cpp,
tests/FluidViscosityIntegrationTests.cpp) with their thresholds: the analytic backward-Euler decay factor
of a discrete sine mode, an independent strain-energy residual (forces assembled by differentiating cell
and vertex strain energies, not the solver's gather stencil), energy decay, rigid rotation, zero
viscosity, zero timestep, the iteration cap and non-finite input, and the volume-weighted material blend.
"""
import math

import numpy as np
import pytest


class Case:
    """The reference test's Case (tests/FluidViscosityTests.cpp:22-80)."""

    def __init__(self, w, h, rho, mu, dx=1.0):
        self.w, self.h, self.dx = w, h, dx
        self.material = np.zeros(2 * w * h)
        self.material[0::2] = rho
        self.material[1::2] = mu
        self.u = np.zeros((w + 1) * h)
        self.v = np.zeros(w * (h + 1))

    def copy(self):
        c = Case(self.w, self.h, 0.0, 0.0, self.dx)
        c.material, c.u, c.v = self.material.copy(), self.u.copy(), self.v.copy()
        return c

    def rho(self, x, y):
        return self.material[2 * (y * self.w + x)]

    def mu(self, x, y):
        return self.material[2 * (y * self.w + x) + 1]

    def um(self, x, y):
        return 0.5 * (self.rho(x - 1, y) + self.rho(x, y)) * self.dx * self.dx

    def vm(self, x, y):
        return 0.5 * (self.rho(x, y - 1) + self.rho(x, y)) * self.dx * self.dx

    def ui(self, x, y):
        return y * (self.w + 1) + x

    def vi(self, x, y):
        return y * self.w + x

    def energy(self):
        s = 0.0
        for y in range(self.h):
            for x in range(1, self.w):
                s += self.um(x, y) * self.u[self.ui(x, y)] ** 2
        for y in range(1, self.h):
            for x in range(self.w):
                s += self.vm(x, y) * self.v[self.vi(x, y)] ** 2
        return 0.5 * s

    def walls(self):
        U = self.u.reshape(self.h, self.w + 1)
        V = self.v.reshape(self.h + 1, self.w)
        assert (U[:, 0] == 0).all() and (U[:, self.w] == 0).all()
        assert (V[0] == 0).all() and (V[self.h] == 0).all()
        assert np.isfinite(self.u).all() and np.isfinite(self.v).all()


def strain_residual(old, new, dt):
    """Independent strain-energy residual (tests/FluidViscosityTests.cpp:82-138)."""
    ku, kv = strain_forces(new)
    err = 0.0
    for y in range(new.h):
        for x in range(1, new.w):
            i = new.ui(x, y)
            m = new.um(x, y)
            r = m * (new.u[i] - old.u[i]) + dt * ku[i]
            err += r * r / m
    for y in range(1, new.h):
        for x in range(new.w):
            i = new.vi(x, y)
            m = new.vm(x, y)
            r = m * (new.v[i] - old.v[i]) + dt * kv[i]
            err += r * r / m
    return math.sqrt(err / max(2 * old.energy(), 1e-30))


def strain_forces(new):
    """K v from the cell and vertex strain energies (tests/FluidViscosityTests.cpp:82-122)."""
    ku = np.zeros(new.u.size)
    kv = np.zeros(new.v.size)
    for y in range(new.h):
        for x in range(new.w):
            tx = 2 * new.mu(x, y) * (new.u[new.ui(x + 1, y)] - new.u[new.ui(x, y)])
            ty = 2 * new.mu(x, y) * (new.v[new.vi(x, y + 1)] - new.v[new.vi(x, y)])
            ku[new.ui(x, y)] -= tx
            ku[new.ui(x + 1, y)] += tx
            kv[new.vi(x, y)] -= ty
            kv[new.vi(x, y + 1)] += ty
    for y in range(1, new.h):
        for x in range(1, new.w):
            mu = 0.25 * (new.mu(x - 1, y - 1) + new.mu(x, y - 1) + new.mu(x - 1, y) + new.mu(x, y))
            t = mu * (new.u[new.ui(x, y)] - new.u[new.ui(x, y - 1)] + new.v[new.vi(x, y)] - new.v[new.vi(x - 1, y)])
            ku[new.ui(x, y)] += t
            ku[new.ui(x, y - 1)] -= t
            kv[new.vi(x, y)] += t
            kv[new.vi(x - 1, y)] -= t
    return ku, kv


def solve(O, case, dt, tol=1e-4, maximum=-1):
    st, it, rel, u, v = O.viscosity_solve(case.w, case.h, case.dx, case.material, case.u, case.v, dt, tol, maximum)
    return st, it, u, v


def decay_case(w, h, mu, transverse):
    c = Case(w, h, 2.0, mu, dx=0.7)
    kx = math.pi / w
    ky = math.pi / h if h > 1 else 0.0
    sx = 2 * math.sin(0.5 * kx)
    sy = 2 * math.sin(0.5 * ky)
    for y in range(h):
        for x in range(1, w):
            c.u[c.ui(x, y)] = math.sin(kx * x) * (math.cos(ky * (y + 0.5)) if transverse else 1.0)
    if transverse:
        for y in range(1, h):
            for x in range(w):
                c.v[c.vi(x, y)] = -sx / sy * math.cos(kx * (x + 0.5)) * math.sin(ky * y)
    return c, sx, sy


# ---- tests/FluidViscosityTests.cpp:179-215 ----
@pytest.mark.parametrize("w,h,mu,dt,transverse", [(24, 16, 100, 0.02, False), (31, 23, 1000, 0.02, True),
                                                  (65, 49, 100000, 0.1, True), (47, 1, 1000, 0.02, False),
                                                  (2, 33, 1000, 0.02, False)])
def test_backward_euler_decay_of_a_sine_mode(oracle, w, h, mu, dt, transverse):
    c, sx, sy = decay_case(w, h, mu, transverse)
    old = c.copy()
    st, it, c.u, c.v = solve(oracle, c, dt)
    assert st == oracle.VISC_OK
    c.walls()
    factor = 1.0 / (1.0 + dt * mu * ((sx * sx + sy * sy) if transverse else 2 * sx * sx) / (2 * c.dx * c.dx))
    err = max(np.abs(c.u - factor * old.u).max(), np.abs(c.v - factor * old.v).max())
    assert err < 3e-4
    assert strain_residual(old, c, dt) < 1.5e-4
    assert c.energy() <= old.energy() * (1 + 1e-6)
    assert it % 4 == 0 and it > 0


# ---- tests/FluidViscosityTests.cpp:217-254 ----
def test_zero_viscosity_rigid_rotation_and_rest(oracle):
    c = Case(25, 19, 2, 0)
    for y in range(c.h):
        for x in range(1, c.w):
            c.u[c.ui(x, y)] = 0.3 * (y - 8)
    for y in range(1, c.h):
        for x in range(c.w):
            c.v[c.vi(x, y)] = -0.3 * (x - 10)
    old = c.copy()
    st, it, u, v = solve(oracle, c, 0.01)
    assert st == oracle.VISC_OK and it == 0 and np.array_equal(u, old.u) and np.array_equal(v, old.v)
    for y in range(4, c.h - 4):
        for x in range(4, c.w - 4):
            c.material[2 * (y * c.w + x) + 1] = 500
    st, it, u, v = solve(oracle, c, 0.01)
    assert st == oracle.VISC_OK
    assert max(np.abs(u - old.u).max(), np.abs(v - old.v).max()) < 3e-5
    c.u[:] = 0
    c.v[:] = 0
    st, it, u, v = solve(oracle, c, 0.01)
    assert st == oracle.VISC_OK and it == 0 and not u.any() and not v.any()


def variable_case(w, h):
    c = Case(w, h, 1, 0)
    for y in range(h):
        for x in range(w):
            dense = x < w // 2
            c.material[2 * (y * w + x)] = 10.0 if dense else 0.1
            c.material[2 * (y * w + x) + 1] = 10000.0 if dense else 0.0
    for y in range(h):
        for x in range(1, w):
            c.u[c.ui(x, y)] = math.sin(0.7 * x + 0.3 * y)
    for y in range(1, h):
        for x in range(w):
            c.v[c.vi(x, y)] = math.cos(0.2 * x + 0.8 * y)
    return c


# ---- tests/FluidViscosityTests.cpp:256-289 (first 10 of 40 steps) ----
@pytest.mark.parametrize("w,h", [(33, 25), (64, 48)])
def test_variable_viscosity_steps(oracle, w, h):
    c = variable_case(w, h)
    for _ in range(10):
        old = c.copy()
        st, it, c.u, c.v = solve(oracle, c, 0.02)
        assert st == oracle.VISC_OK
        assert c.energy() <= old.energy() * (1 + 2e-6)
        assert strain_residual(old, c, 0.02) < 1.8e-4


# ---- tests/FluidViscosityTests.cpp:291-344 ----
def test_zero_dt_iteration_cap_and_nonfinite_input(oracle):
    c = Case(33, 25, 1, 10000)
    for y in range(c.h):
        for x in range(1, c.w):
            c.u[c.ui(x, y)] = math.sin(0.7 * x + 0.3 * y)
    st, it, u, v = solve(oracle, c, 0.0)
    assert st == oracle.VISC_OK and it == 0 and np.array_equal(u, c.u)
    st, it, u, v = solve(oracle, c, 0.02, 1e-4, 1)
    assert st == oracle.VISC_NOT_CONVERGED and np.array_equal(u, c.u) and np.array_equal(v, c.v)
    old = c.copy()
    st, it, c.u, c.v = solve(oracle, c, 0.02)
    assert st == oracle.VISC_OK and it > 1
    assert strain_residual(old, c, 0.02) < 1.8e-4
    bad = old.copy()
    bad.u[bad.ui(5, 5)] = float("nan")
    st, it, u, v = solve(oracle, bad, 0.02)
    assert st == oracle.VISC_NONFINITE and np.array_equal(v, old.v)


def test_hierarchy_depth_matches_the_pressure_rule(oracle):
    """ViscositySolverGPU.cpp:79-89: halve (w + 1) / 2 until w * h <= 64."""
    assert oracle.lib().orc_viscosity_levels(8, 8) == 1
    assert oracle.lib().orc_viscosity_levels(9, 8) == 2
    assert oracle.lib().orc_viscosity_levels(512, 420) == 7
    assert oracle.lib().orc_viscosity_levels(1, 129) == 3


# ---- tests/FluidViscosityIntegrationTests.cpp:89-113: the volume-weighted blend ----
def test_material_blend_is_volume_weighted(oracle):
    rng = np.random.default_rng(5)
    N = 37
    ld, lv = np.array([10.0, 11.0]), np.array([20.0, 120.0])
    gd, gv = np.array([0.1, 0.11]), np.array([0.0, 5.0])
    lm = rng.uniform(0, 1, 2 * N)
    gm = rng.uniform(0, 0.01, 2 * N)
    h = 0.7
    got = oracle.viscosity_material(h, ld, lv, lm, gd, gv, gm, N)
    for i in range(N):
        ms = [lm[i], lm[N + i], gm[i], gm[N + i]]
        ds = [ld[0], ld[1], gd[0], gd[1]]
        mus = [lv[0], lv[1], gv[0], gv[1]]
        vol = sum(m / d for m, d in zip(ms, ds))
        assert abs(got[2 * i] - sum(ms) / (h * h)) <= 1e-12 * got[2 * i]
        assert abs(got[2 * i + 1] - sum(m / d * mu for m, d, mu in zip(ms, ds, mus)) / vol) <= 1e-12 * got[2 * i + 1]


def _stream_state(W, H):
    """tests/FluidViscosityIntegrationTests.cpp:25-82: layered fractions, a divergence-free vortex."""
    def psi(x, y):
        if x == 0 or x == W or y == 0 or y == H:
            return 0.0
        return 2.0 * math.sin(math.pi * x / W) * math.sin(math.pi * y / H)

    f = np.array([(y + 1) / (H + 1) for y in range(H) for x in range(W)])
    u = np.array([psi(x, y + 1) - psi(x, y) for y in range(H) for x in range(W + 1)])
    v = np.array([psi(x, y) - psi(x + 1, y) for y in range(H + 1) for x in range(W)])
    return f, u, v


def test_grid_step_with_viscosity_dissipates(oracle):
    """A viscous two-phase step (transport, viscosity, gravity, projection) converges, keeps the liquid
    volume and closed walls, runs the viscosity solve, and leaves less kinetic energy than the same step
    without viscosity."""
    W, H, h = 24, 18, 0.7
    f, u, v = _stream_state(W, H)
    lm = np.concatenate([h * h * f * 10.0 / 2, h * h * f * 11.0 / 2])
    gm = np.concatenate([h * h * (1 - f) * 0.1 / 2, h * h * (1 - f) * 0.11 / 2])
    energies = []
    for viscous in (False, True):
        g = oracle.OracleGrid(W, H, h)
        g.configure([10.0, 11.0], [0.1, 0.11])
        if viscous:
            g.configure_viscosity([20.0, 120.0], [0.0, 5.0])
        g.set_gravity(0.0)
        g.upload(f, u, v, None, lm, gm)
        r = g.step(0.004, 2000, 1e-3)
        assert r["converged"]
        assert (r["viscosity_iterations"] > 0) == viscous
        s = g.download()
        assert abs(s["fraction"].sum() - f.sum()) < 1e-10 * f.sum()
        U = s["u"].reshape(H, W + 1)
        V = s["v"].reshape(H + 1, W)
        assert (U[:, 0] == 0).all() and (U[:, W] == 0).all() and (V[0] == 0).all() and (V[H] == 0).all()
        energies.append(float((s["u"] ** 2).sum() + (s["v"] ** 2).sum()))
    assert energies[1] < energies[0]


# ---- Chebyshev-weighted viscosity smoother (orc_viscosity_solve_smoothed; no reference counterpart) ----
def operator_matrix(case, dt):
    """M + dt K on the interior faces, column by column from strain_forces (independent of the oracle)."""
    faces = [("u", case.ui(x, y), case.um(x, y)) for y in range(case.h) for x in range(1, case.w)]
    faces += [("v", case.vi(x, y), case.vm(x, y)) for y in range(1, case.h) for x in range(case.w)]
    n = len(faces)
    A = np.zeros((n, n))
    probe = case.copy()
    for j, (comp, idx, mass) in enumerate(faces):
        probe.u[:] = 0.0
        probe.v[:] = 0.0
        (probe.u if comp == "u" else probe.v)[idx] = 1.0
        ku, kv = strain_forces(probe)
        for i, (ci, ii, _) in enumerate(faces):
            A[i, j] = dt * (ku[ii] if ci == "u" else kv[ii])
        A[j, j] += mass
    return A


def test_viscosity_spectrum_is_bounded_by_four():
    """The Chebyshev interval's upper end: every eigenvalue of D^-1 (M + dt K) lies in (0, 4] for any
    material (Gershgorin: each row's off-diagonal magnitudes sum to at most three times its viscous
    diagonal part), and exceeds 2 -- so the pressure smoother's interval [0.2, 2] would not do here."""
    rng = np.random.default_rng(11)
    top = 0.0
    for trial in range(6):
        c = Case(6, 5, 1.0, 0.0, dx=0.5)
        c.material[0::2] = rng.uniform(0.1, 10.0, c.w * c.h)
        mu = rng.uniform(0.0, 1e4, c.w * c.h)
        if trial % 2:
            mu *= (np.indices((c.h, c.w)).sum(axis=0).ravel() % 2)   # checkerboard of inviscid cells
        c.material[1::2] = mu
        A = operator_matrix(c, 1.0 if trial < 3 else 1e-3)
        assert np.allclose(A, A.T, rtol=1e-12, atol=1e-9)
        d = np.sqrt(np.diag(A))
        lam = np.linalg.eigvalsh(A / d[:, None] / d[None, :])
        assert lam.min() > 0.0
        assert lam.max() <= 4.0 + 1e-9
        top = max(top, lam.max())
    assert top > 2.2


@pytest.mark.parametrize("n", [2, 4, 6])
def test_chebyshev_viscosity_weights_give_the_chebyshev_residual_polynomial(oracle, n):
    """prod_k (1 - w_k lam) = T_n((4.25 - 2 lam) / 3.75) / T_n(4.25 / 3.75): the weights are the reciprocals of
    the Chebyshev nodes of [0.25, 4], so the error factor is at most 1 / T_n(4.25 / 3.75) there."""
    w = [oracle.lib().orc_viscosity_chebyshev_weight(k, n) for k in range(n)]
    cheb = np.polynomial.chebyshev.Chebyshev.basis(n)
    lam = np.linspace(0.0, 4.0, 81)
    prod = np.prod([1.0 - wk * lam for wk in w], axis=0)
    expect = cheb((4.25 - 2.0 * lam) / 3.75) / cheb(4.25 / 3.75)
    assert np.allclose(prod, expect, rtol=0, atol=1e-12)
    inside = lam >= 0.25
    assert np.abs(prod[inside]).max() <= 1.0 / cheb(4.25 / 3.75) + 1e-12
    assert np.all(prod[~inside] > 0) and np.all(prod[~inside] <= 1.0)


def smoothed(O, case, dt, n, tol=1e-4):
    st, it, rel, u, v = O.viscosity_solve(case.w, case.h, case.dx, case.material, case.u, case.v, dt, tol, -1, n)
    return st, it, u, v


@pytest.mark.parametrize("n", [2, 4, 6])
def test_chebyshev_viscosity_smoother_solves_to_the_same_bar(oracle, n):
    """The decay and variable-viscosity checks above, with n Chebyshev-weighted sweeps per side: the same
    analytic factor and independent strain residual, and no more iterations than the reference's smoother
    from four sweeps on."""
    c, sx, sy = decay_case(31, 23, 1000, True)
    old = c.copy()
    st, it, c.u, c.v = smoothed(oracle, c, 0.02, n)
    assert st == oracle.VISC_OK and it > 0 and it % 4 == 0
    c.walls()
    factor = 1.0 / (1.0 + 0.02 * 1000 * (sx * sx + sy * sy) / (2 * c.dx * c.dx))
    assert max(np.abs(c.u - factor * old.u).max(), np.abs(c.v - factor * old.v).max()) < 3e-4
    assert strain_residual(old, c, 0.02) < 1.5e-4
    c = variable_case(64, 48)
    for _ in range(3):
        old = c.copy()
        st, it, c.u, c.v = smoothed(oracle, c, 0.02, n)
        assert st == oracle.VISC_OK
        assert c.energy() <= old.energy() * (1 + 2e-6)
        assert strain_residual(old, c, 0.02) < 1.8e-4
        st0, it0, _, _ = solve(oracle, old, 0.02)
        if n >= 4:
            assert it <= it0


def test_chebyshev_viscosity_smoother_rejects_other_counts(oracle):
    c = variable_case(9, 7)
    for n in (1, 3, 5, 8, -2):
        st, it, u, v = smoothed(oracle, c, 0.02, n)
        assert st == oracle.VISC_BAD_SMOOTHER and np.array_equal(u, c.u)
    g = oracle.OracleGrid(9, 7, 0.1)
    with pytest.raises(ValueError):
        g.set_viscosity_smoother(3)
    g.set_viscosity_smoother(4)
    g.set_viscosity_smoother(0)
