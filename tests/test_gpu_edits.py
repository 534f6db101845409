"""GPU parity of writeColorPixels and commitEdits (fs_color, fs_commit_edits) against the oracle.

Both are pure FP64/FP32 arithmetic in the reference's order (FluidGridCPU.cpp:1826-1860, :253-284), so the
bar is bit-exact.  A projection after the edits must see the edited densities (caches invalidated) and match
the oracle's projection to the usual 1e-5 relative L2.
"""
import ctypes as C

import numpy as np
import pytest

from fluidsim_amd import (GAS, LIQUID, FluidGridHIP, FluidGridState, FluidMixture, MixtureComponent,
                          Substance, SubstanceCatalogue, _abi)

pytestmark = pytest.mark.gpu

CATALOGUE = SubstanceCatalogue([Substance("air", (0.3, 0.3, 0.3)), Substance("oil", (0.9, 0.6, 0.05)),
                                Substance("water", (0.1, 0.4, 0.95)), Substance("steam", (0.8, 0.8, 0.85))])
WATER, OIL, AIR, STEAM = 2, 1, 0, 3
LIQ = FluidMixture([MixtureComponent(WATER, 100.0), MixtureComponent(OIL, 80.0)])
GASES = FluidMixture([MixtureComponent(AIR, 0.1), MixtureComponent(STEAM, 0.05)])


def rgb():
    return np.array([CATALOGUE[c.substance].color for c in LIQ.components + GASES.components], dtype=np.float32)


def state(W, H, h, seed):
    rng = np.random.default_rng(seed)
    N = W * H
    f = np.clip(rng.uniform(-0.3, 1.3, N), 0.0, 1.0)
    share = rng.uniform(0, 1, N)
    lm = [share * f * 100.0 * h * h, (1 - share) * f * 80.0 * h * h]
    gshare = rng.uniform(0, 1, N)
    gm = [gshare * (1 - f) * 0.1 * h * h, (1 - gshare) * (1 - f) * 0.05 * h * h]
    u = rng.uniform(-0.5, 0.5, (W + 1) * H)
    v = rng.uniform(-0.5, 0.5, W * (H + 1))
    u.reshape(H, W + 1)[:, [0, W]] = 0.0
    v.reshape(H + 1, W)[[0, H], :] = 0.0
    return FluidGridState(f, u, v, np.zeros(0), lm, gm)


def pair(O, W, H, h, seed):
    s = state(W, H, h, seed)
    g = FluidGridHIP(W, H, h)
    g.configureMaterials(CATALOGUE, LIQ, GASES)
    g.setGravityY(9.81)
    g.upload(s)
    o = O.OracleGrid(W, H, h)
    o.configure([100.0, 80.0], [0.1, 0.05])
    o.set_gravity(9.81)
    o.upload(s.liquid_fraction, s.u, s.v, None, np.concatenate(s.liquid_mass), np.concatenate(s.gas_mass))
    return g, o


@pytest.mark.parametrize("W,H", [(97, 61), (256, 200), (1, 1)])
def test_color_pixels_are_bit_exact(oracle, W, H):
    g, o = pair(oracle, W, H, 1.0 / max(W, H), 5)
    for alpha in (False, True):
        got = g.writeColorPixels(liquid_fraction_alpha=alpha)
        assert got.dtype == np.float32 and got.size == 4 * W * H
        assert np.array_equal(got, o.color_pixels(rgb(), alpha))
    # caller storage of the right size is reused
    buf = np.zeros(4 * W * H, dtype=np.float32)
    assert g.writeColorPixels(buf) is buf
    g.destroy()


def test_commit_edits_matches_the_oracle_then_projects_like_it(oracle):
    W, H, h = 120, 90, 1.0 / 120
    g, o = pair(oracle, W, H, h, 6)
    rng = np.random.default_rng(9)
    changes = []
    for _ in range(700):   # a brush stroke with repeated cells, unsorted
        x, y = int(rng.integers(0, W)), int(rng.integers(0, H // 2))
        phase, sub = [(LIQUID, WATER), (LIQUID, OIL), (GAS, AIR), (GAS, STEAM)][int(rng.integers(0, 4))]
        changes.append((x, y, phase, sub))
    session = g.edit()
    for c in changes:
        session.setCell(*c)
    session.commit()
    assert session.empty()
    liquid = [(WATER, 100.0), (OIL, 80.0)]
    gas = [(AIR, 0.1), (STEAM, 0.05)]
    o.commit_edits(oracle.prepare_edits(changes, W, H, h * h, liquid, gas))
    sg, so = g.download(), o.download()
    assert np.array_equal(sg.liquid_fraction, so["fraction"])
    assert np.array_equal(np.concatenate(sg.liquid_mass), so["liquid_mass"])
    assert np.array_equal(np.concatenate(sg.gas_mass), so["gas_mass"])
    assert np.array_equal(sg.u, so["u"]) and np.array_equal(sg.v, so["v"])
    assert np.array_equal(g.writeColorPixels(), o.color_pixels(rgb()))
    # the next projection uses the edited densities
    g.applyGravity(1.0 / 120)
    rg = g.project(1.0 / 120, 200, 1e-6)
    ro = o.force_project(1.0 / 120, 200, 1e-6)   # gravity, then the projection
    assert rg.converged and ro["converged"]
    sg, so = g.download(), o.download()
    for a, b in ((sg.u, so["u"]), (sg.v, so["v"])):
        assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-5
    g.destroy()


def test_edit_errors_follow_the_reference_contract(oracle):
    W, H, h = 16, 12, 0.25
    g, _ = pair(oracle, W, H, h, 7)
    before = g.download()
    with pytest.raises(IndexError):
        g.commitEdits([(0, 0, LIQUID, WATER), (16, 0, LIQUID, WATER)])
    with pytest.raises(ValueError):
        g.commitEdits([(1, 1, GAS, WATER)])
    with pytest.raises(ValueError):
        g.commitEdits([(1, 1, LIQUID, WATER)], velocity=(1.0, 0.0))
    # the ABI's own checks (a caller that skips prepare): nothing changes on a refused batch
    cells = (_abi.FsCellEdit * 2)()
    cells[0].x, cells[0].y, cells[0].component, cells[0].mass = 1, 1, 0, 1.0
    cells[1].x, cells[1].y, cells[1].component, cells[1].mass = 3, 99, 0, 1.0
    assert _abi.lib().fs_commit_edits(g.handle, 2, cells, None) == _abi.FS_E_INVALID_ARGUMENT
    cells[1].y, cells[1].component = 2, 4
    assert _abi.lib().fs_commit_edits(g.handle, 2, cells, None) == _abi.FS_E_INVALID_ARGUMENT
    vel = np.array([0.0, 1.0])
    assert _abi.lib().fs_commit_edits(g.handle, 1, cells, vel.ctypes.data_as(C.POINTER(C.c_double))) == \
        _abi.FS_E_UNSUPPORTED
    after = g.download()
    assert np.array_equal(after.liquid_fraction, before.liquid_fraction)
    assert all(np.array_equal(a, b) for a, b in zip(after.liquid_mass, before.liquid_mass))
    # edits and colours need an uploaded state of the configured mixtures
    fresh = FluidGridHIP(W, H, h)
    fresh.configureMaterials(CATALOGUE, LIQ, GASES)
    with pytest.raises(RuntimeError):
        fresh.commitEdits([(0, 0, LIQUID, WATER)])
    with pytest.raises(RuntimeError):
        fresh.writeColorPixels()
    fresh.destroy()
    g.destroy()
