"""Pins for the oracle's writeColorPixels and commitEdits (oracle/fs_oracle.c, oracle.prepare_edits) -- CPU.

writeColorPixels (FluidSim/UniformCPU/FluidGridCPU.cpp:1826-1860): a cell holding exactly one component at
its full density shows that component's colour exactly; an empty cell is black; alpha is the liquid fraction
rounded to float (or 1); colours add with the volume fractions.  commitEdits (:253-284) with
CellEditBatch::prepare (FluidSim/FluidGridEdit.h:37-95): cells outside the grid and substances missing from
the phase are refused before anything changes; edited cells become pure liquid or gas with density * h^2 of
one component; the last edit of a cell wins; nothing else changes.
"""
import numpy as np
import pytest

LIQUID = [(7, 100.0), (3, 80.0)]     # (substance id, density)
GAS = [(5, 0.1)]
RGB = np.array([[0.9, 0.2, 0.1], [0.1, 0.5, 0.95], [0.3, 0.3, 0.3]], dtype=np.float32)


def grid(O, W=6, H=5, h=0.5):
    g = O.OracleGrid(W, H, h)
    g.configure([d for _, d in LIQUID], [d for _, d in GAS])
    N = W * H
    rng = np.random.default_rng(3)
    f = rng.uniform(0, 1, N)
    lm = np.concatenate([0.6 * f * 100.0 * h * h, 0.4 * f * 80.0 * h * h])
    gm = (1 - f) * 0.1 * h * h
    u = rng.uniform(-1, 1, (W + 1) * H)
    v = rng.uniform(-1, 1, W * (H + 1))
    g.upload(f, u, v, None, lm, gm)
    return g


def test_pure_cells_show_their_component_colour_exactly(oracle):
    W, H, h = 3, 1, 0.25
    g = oracle.OracleGrid(W, H, h)
    g.configure([d for _, d in LIQUID], [d for _, d in GAS])
    lm = np.zeros(2 * W)
    gm = np.zeros(W)
    lm[0] = 100.0 * h * h        # cell 0: component 0 at full density
    lm[W + 1] = 80.0 * h * h     # cell 1: component 1
    gm[2] = 0.1 * h * h          # cell 2: the gas
    g.upload(np.array([1.0, 1.0, 0.0]), None, None, None, lm, gm)
    px = g.color_pixels(RGB).reshape(W, 4)
    for c in range(3):
        assert np.array_equal(px[c, :3], RGB[c])
    assert np.array_equal(px[:, 3], np.ones(3, dtype=np.float32))
    a = g.color_pixels(RGB, liquid_fraction_alpha=True).reshape(W, 4)
    assert np.array_equal(a[:, 3], np.array([1.0, 1.0, 0.0], dtype=np.float32))


def test_colours_add_by_volume_fraction_and_empty_cells_are_black(oracle):
    W, H, h = 2, 1, 1.0
    g = oracle.OracleGrid(W, H, h)
    g.configure([d for _, d in LIQUID], [d for _, d in GAS])
    # cell 0: a quarter of component 0 and three quarters of the gas; cell 1 empty
    lm = np.array([25.0, 0.0, 0.0, 0.0])
    gm = np.array([0.075, 0.0])
    g.upload(np.array([0.25, 0.0]), None, None, None, lm, gm)
    px = g.color_pixels(RGB, liquid_fraction_alpha=True).reshape(W, 4)
    expect = 0.25 * RGB[0].astype(np.float64) + (0.075 / 0.1) * RGB[2].astype(np.float64)
    assert np.allclose(px[0, :3], expect, rtol=2e-7, atol=0)
    assert px[0, 3] == np.float32(0.25)
    assert np.array_equal(px[1], np.array([0, 0, 0, 0], dtype=np.float32))


def test_prepare_refuses_bad_edits_and_keeps_the_last_edit_of_a_cell(oracle):
    with pytest.raises(IndexError):
        oracle.prepare_edits([(0, 0, "liquid", 7), (6, 0, "liquid", 7)], 6, 5, 0.25, LIQUID, GAS)
    with pytest.raises(IndexError):
        oracle.prepare_edits([(0, -1, "gas", 5)], 6, 5, 0.25, LIQUID, GAS)
    with pytest.raises(ValueError):
        oracle.prepare_edits([(1, 1, "gas", 7)], 6, 5, 0.25, LIQUID, GAS)   # 7 is a liquid substance
    cells = oracle.prepare_edits([(4, 2, "liquid", 3), (1, 0, "gas", 5), (4, 2, "liquid", 7), (0, 2, "gas", 5)],
                                 6, 5, 0.25, LIQUID, GAS)
    # raster order, the second edit of (4, 2) wins, masses are density * h^2, components liquid-first
    assert cells == [(1, 0, 2, 0.1 * 0.25), (0, 2, 2, 0.1 * 0.25), (4, 2, 0, 100.0 * 0.25)]


def test_commit_replaces_only_the_edited_cells(oracle):
    W, H, h = 6, 5, 0.5
    g = grid(oracle, W, H, h)
    before = g.download()
    cells = oracle.prepare_edits([(2, 1, "liquid", 3), (5, 4, "gas", 5), (2, 1, "gas", 5), (0, 0, "liquid", 7)],
                                 W, H, h * h, LIQUID, GAS)
    g.commit_edits(cells)
    after = g.download()
    N = W * H
    f = before["fraction"].copy()
    lm = before["liquid_mass"].copy()
    gm = before["gas_mass"].copy()
    for x, y, k, m in [(2, 1, 2, 0.1 * h * h), (5, 4, 2, 0.1 * h * h), (0, 0, 0, 100.0 * h * h)]:
        i = y * W + x
        f[i] = 1.0 if k < 2 else 0.0
        lm[i] = lm[N + i] = gm[i] = 0.0
        (lm if k < 2 else gm)[(k if k < 2 else k - 2) * N + i] = m
    assert np.array_equal(after["fraction"], f)
    assert np.array_equal(after["liquid_mass"], lm)
    assert np.array_equal(after["gas_mass"], gm)
    assert np.array_equal(after["u"], before["u"]) and np.array_equal(after["v"], before["v"])
    # the density the next projection sees follows the edit (caches invalidated)
    rho = g.density()
    assert rho[0] == pytest.approx(100.0) and rho[1 * W + 2] == pytest.approx(0.1)
