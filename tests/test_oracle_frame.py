"""Pins for the oracle's scene frame schedule (oracle/oracle.py frame_schedule, run_frame) -- CPU.

FluidScene::sceneProcess (FluidScene.inl:343-353) splits a frame into ceil(frame_time * steps_per_second)
equal steps, unless the engine's capillary subdivision alone already reaches that count, in which case
the frame is one engine step.  The cases below are worked by hand from those lines.
"""
import math

import numpy as np
import pytest

from test_oracle_tension import capillary_grid, disk


@pytest.mark.parametrize("frame,rate,limit,count", [
    (0.02, 60, math.inf, 2),       # 1.2 -> 2 steps of 0.01
    (0.02, 50, math.inf, 1),       # exactly one step
    (0.05, 60, math.inf, 3),       # 0.05 * 60 is exactly 3.0 in binary64
    (0.1, 60, math.inf, 6),
    (0.02, 60, 0.004, 1),          # capillary steps (5) already exceed the scene rate (2)
    (0.02, 60, 0.015, 1),          # ceil(1.33) = 2 >= 2: the engine subdivides instead
    (0.02, 60, 0.03, 2),           # ceil(0.67) = 1 < 2: the scene subdivides
    (0.02, 240, 0.004, 1),         # ceil(0.02 / 0.004) = 5 >= ceil(4.8) = 5: one engine step
    (0.02, 300, 0.004, 6),         # 6 scene steps beat 5 capillary steps
    (0.07, 100, math.inf, 8),      # 0.07 * 100 = 7.000000000000001 in binary64: ceil gives 8
])
def test_frame_schedule(oracle, frame, rate, limit, count):
    n, step = oracle.frame_schedule(frame, rate, limit)
    assert n == count
    assert step == frame / count


def test_frame_with_capillary_shortcut_is_one_subdivided_step(oracle):
    """With a capillary limit below the scene step the frame is one engine call of ceil(frame/limit)
    physical substeps, identical to calling step(frame) directly."""
    W = H = 24
    f = disk(W, H, 5.0)
    a = capillary_grid(oracle, W, H, f, 40000.0)
    b = capillary_grid(oracle, W, H, f, 40000.0)
    limit = a.capillary_step_limit()
    assert math.ceil(0.02 / limit) >= 2
    stats = oracle.run_frame(a, 0.02, 60, 0.001)
    r = b.step(0.02, oracle.UNBOUNDED_BUDGET, 0.001)
    assert stats["steps_scheduled"] == 1 and stats["steps_run"] == 1 and stats["accepted"]
    assert stats["substeps"] == math.ceil(0.02 / limit) == r["physical_steps"]
    for k, v in a.download().items():
        assert np.array_equal(v, b.download()[k])


def test_frame_statistics_sum_the_steps(oracle):
    W, H = 20, 16
    f = np.array([float(y >= H // 2) for y in range(H) for x in range(W)])
    g = capillary_grid(oracle, W, H, f, 0.0, gravity=100.0)
    stats = oracle.run_frame(g, 0.05, 60, 0.01)
    assert stats["steps_scheduled"] == 3 and stats["steps_run"] == 3 and stats["substeps"] == 3
    assert stats["accepted"]
