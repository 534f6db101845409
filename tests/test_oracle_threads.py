"""The oracle's threaded mode (orc_set_threads) against its serial mode -- CPU only.

The reference runs its projection passes on up to 8 workers (FluidGridCPU.cpp:25-57) and its multigrid
levels on 4 or 8 (PressureMultigrid.cpp:13-21); the threaded oracle does the same so that the CPU baseline
can be timed on the host's cores.  Element-wise passes give identical values; reductions merge per-batch
partials in batch order, which changes rounding only.  The reference pins exactly that with
parallelExecutionMatchesSerialOnOddGrid (tests/FluidGridTests.cpp:683-735): a 513 x 257 projection at
tol 1e-8 and three steps, serial against 8 threads, velocities within 2e-7.  The same scenario here.
"""
import math

import numpy as np
import pytest


def odd_grid(oracle):
    W, H = 513, 257
    f = np.array([min(max(y + 1.0 - (0.3 * x + 30.0), 0.0), 1.0) for y in range(H) for x in range(W)])
    u = np.sin(np.arange((W + 1) * H) * 0.03)
    v = np.cos(np.arange(W * (H + 1)) * 0.017)
    g = oracle.OracleGrid(W, H, 1.0)
    g.configure([100.0], [0.1])
    g.upload(f, u, v, None, 100.0 * f, 0.1 * (1.0 - f))
    return g


@pytest.fixture
def restore_threads(oracle):
    yield
    oracle.set_threads(1)


def test_threaded_projection_and_steps_match_serial(oracle, restore_threads):
    serial, parallel = odd_grid(oracle), odd_grid(oracle)
    oracle.set_threads(1)
    a = serial.project(0.004, 2000, 1e-8)
    oracle.set_threads(8)
    assert oracle.get_threads() == 8
    b = parallel.project(0.004, 2000, 1e-8)
    assert a["converged"] and b["converged"]
    sa, sb = serial.download(), parallel.download()
    assert np.abs(sa["u"] - sb["u"]).max() < 2e-7
    assert np.abs(sa["v"] - sb["v"]).max() < 2e-7
    for tick in range(3):
        oracle.set_threads(1)
        ra = serial.step(0.004, 2000, 1e-8)
        oracle.set_threads(8)
        rb = parallel.step(0.004, 2000, 1e-8)
        assert ra["converged"] and rb["converged"], tick
        sa, sb = serial.download(), parallel.download()
        assert np.abs(sa["fraction"] - sb["fraction"]).max() < 2e-7
        assert np.abs(sa["liquid_mass"] - sb["liquid_mass"]).max() < 2e-5
        assert np.abs(sa["gas_mass"] - sb["gas_mass"]).max() < 2e-8
        assert np.abs(sa["u"] - sb["u"]).max() < 2e-7
        assert np.abs(sa["v"] - sb["v"]).max() < 2e-7


def test_small_grids_stay_serial_and_bitwise(oracle, restore_threads):
    """Below 65536 items every pass is one batch, whatever the thread count (batchWorkers)."""
    W, H = 200, 150
    rng = np.random.default_rng(3)
    f = np.clip(rng.uniform(-0.5, 1.5, W * H), 0.0, 1.0)

    def run(threads):
        oracle.set_threads(threads)
        g = oracle.OracleGrid(W, H, 1.0 / W)
        g.configure([100.0], [0.1])
        g.upload(f, rng_u, rng_v, None, 100.0 * f / W ** 2, 0.1 * (1.0 - f) / W ** 2)
        r = g.project(0.01, 500, 1e-7)
        return r, g.download()

    rng_u = np.sin(np.arange((W + 1) * H) * 0.11)
    rng_v = np.cos(np.arange(W * (H + 1)) * 0.07)
    r1, s1 = run(1)
    r8, s8 = run(8)
    assert r1["iterations"] == r8["iterations"]
    for k in ("u", "v", "pressure"):
        assert np.array_equal(s1[k], s8[k]), k
