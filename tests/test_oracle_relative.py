"""The relative stopping rule (extension: fs_set_relative_tolerance / orc_set_relative_tolerance) on the
FP64 oracle -- CPU only.

The reference's tolerance is absolute (FluidGridCPU.cpp:1030-1124: PCG to max|r| <= tol h^2/dt, accepted at
max|div u| <= tol).  The relative rule replaces tol by rel * max|b| dt / h^2, b being the mean-free
pressure rhs.  Pinned two ways:
  * it is the absolute rule at the tolerance that formula gives, computed independently from the rhs
    (orc_pressure_rhs): the two projections must agree bit for bit;
  * it is scale invariant: the projection is linear in the velocity, so scaling the initial velocity by k
    scales the rhs, the stopping threshold and every CG iterate by k -- same iteration count, pressure and
    velocity scaled by k (to rounding).  An absolute rule fails this (it iterates more on the larger flow).
"""
import numpy as np
import pytest

from fluidsim_amd import workloads as WL


def pool(oracle, scale=1.0, relative=True):
    W, H = 40, 28
    h = 1.0 / W
    rng = WL.SplitMix64(17)
    f = WL.pool_with_drop(W, H, h, 0.55, 0.12, rng)
    u = np.zeros((W + 1) * H)
    v = np.zeros(W * (H + 1))
    for _ in range(4):
        WL.add_gaussian(u, v, W, H, h, rng.uniform(1, 0, W * h)[0], rng.uniform(1, 0, H * h)[0], 0.1,
                        rng.normal(1)[0])
    WL.zero_walls(u, v, W, H)
    lm, gm = WL.two_phase_masses(f, h)
    g = oracle.OracleGrid(W, H, h)
    g.configure([100.0], [0.1])
    g.set_gravity(0.0)
    g.set_relative_tolerance(relative)
    g.upload(f, scale * u, scale * v, None, lm, gm)
    return g, h


@pytest.mark.parametrize("rel", [1e-3, 1e-6])
def test_relative_rule_is_the_absolute_rule_at_rel_max_b(oracle, rel):
    dt = 0.01
    probe, h = pool(oracle, relative=False)
    bmax = np.abs(probe.pressure_rhs(dt)).max()
    assert bmax > 0
    absolute, _ = pool(oracle, relative=False)
    ra = absolute.project(dt, 500, rel * bmax * dt / (h * h))
    relative, _ = pool(oracle)
    rr = relative.project(dt, 500, rel)
    assert rr["iterations"] == ra["iterations"] > 0
    assert rr["converged"] and ra["converged"]
    sa, sr = absolute.download(), relative.download()
    for k in ("u", "v", "pressure"):
        assert np.array_equal(sa[k], sr[k]), k
    # the residual bound holds relative to max|b|
    assert rr["final_residual"] <= rel * bmax


def test_relative_rule_is_scale_invariant(oracle):
    dt = 0.01
    base, _ = pool(oracle)
    big, _ = pool(oracle, scale=1000.0)
    r1 = base.project(dt, 500, 1e-6)
    r2 = big.project(dt, 500, 1e-6)
    assert r1["converged"] and r2["converged"]
    assert r1["iterations"] == r2["iterations"]
    s1, s2 = base.download(), big.download()
    for k in ("u", "v", "pressure"):
        ref = np.abs(s2[k]).max()
        assert np.abs(s2[k] - 1000.0 * s1[k]).max() <= 1e-9 * ref, k
    # the absolute rule at the small flow's tolerance needs more iterations on the large one
    base_abs, _ = pool(oracle, relative=False)
    big_abs, _ = pool(oracle, scale=1000.0, relative=False)
    tol = 1e-6 * base_abs.max_divergence()
    assert big_abs.project(dt, 500, tol)["iterations"] > base_abs.project(dt, 500, tol)["iterations"]
