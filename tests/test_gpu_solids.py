"""Solid obstacles on the GPU (fs_set_solid_mask) against the FP64 oracle's solid-mask extension.

Bars: face weights, every multigrid level's conductances and the coarse factor (with its pinned
unknowns) bit-exact; an all-fluid mask bit-identical to no mask; pressure (up to a constant, fluid
cells) and velocity within 1e-5 relative L2 at the same stopping rule; velocity exactly zero on every
face that touches a solid cell; the masked FP32 V-cycle linear, symmetric and positive to the FP32
bars of tests/FluidGridWebGPUTests.cpp:428-535 and close to the FP64 V-cycle.
"""
import math

import numpy as np
import pytest

from fluidsim_amd import FluidGridHIP
from fluidsim_amd import workloads as WL

pytestmark = pytest.mark.gpu

REL_L2 = 1e-5


def rel_l2(a, b):
    n = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (n if n > 0 else 1.0)


def oracle_grid(O, wl, mask=None, preconditioner="multigrid"):
    g = O.OracleGrid(wl.width, wl.height, wl.cell_size)
    g.configure(wl.liquid_density, wl.gas_density)
    g.set_gravity(wl.gravity_y)
    g.set_preconditioner(preconditioner)
    g.upload(wl.fraction, wl.u, wl.v, None, wl.liquid_mass, wl.gas_mass)
    if mask is not None:
        g.set_solid(mask)
    return g


def hip_grid(wl, mask=None):
    g = FluidGridHIP(wl.width, wl.height, wl.cell_size)
    WL.load_into(g, wl)
    if mask is not None:
        g.setSolidMask(mask)
    return g


def closed_faces(mask, W, H):
    m = mask.reshape(H, W).astype(bool)
    cu = np.zeros((H, W + 1), bool)
    cu[:, 1:W] = m[:, 1:] | m[:, :-1]
    cv = np.zeros((H + 1, W), bool)
    cv[1:H, :] = m[1:, :] | m[:-1, :]
    return cu.ravel(), cv.ravel()


def blocked_16():
    mask = np.zeros((16, 16), np.uint8)
    mask[6:8, 10:12] = 1
    mask[12:14, 2:4] = 1
    return mask.ravel()


@pytest.mark.parametrize("case", ["blocked16", "tank96x80", "tank301x203"])
def test_masked_hierarchy_is_bit_exact(oracle, case):
    if case == "blocked16":
        wl = WL.vortex_box(16, 16, seed=9, vortices=4, sources=2)
        mask = blocked_16()
    else:
        W, H = (96, 80) if case == "tank96x80" else (301, 203)
        wl = WL.obstacle_tank(W, H, seed=3)
        mask = wl.meta["solid"]
    o = oracle_grid(oracle, wl, mask)
    g = hip_grid(wl, mask)
    assert g.solidCells() == int(mask.sum())
    iu_o, iv_o = o.face_weights()
    iu_h, iv_h = g.faceWeights()
    assert np.array_equal(iu_o, iu_h) and np.array_equal(iv_o, iv_h)
    mg = o.multigrid()
    assert mg is not None
    assert g.multigridLevels() == mg.levels
    for lev in range(mg.levels):
        ro, do, _ = mg.weights(lev)
        rh, dh, _ = g.multigridWeights(lev)
        assert np.array_equal(ro, rh) and np.array_equal(do, dh), f"level {lev}"
    f = g.multigridCoarseFactor()
    assert np.array_equal(mg.coarse_factor(), f)
    if case == "blocked16":
        assert mg.pinned().sum() == 2
    g.destroy()


def test_all_fluid_mask_is_bit_identical():
    wl = WL.dam_break(128, 128, seed=1)
    a, b = hip_grid(wl), hip_grid(wl, np.zeros(128 * 128, np.uint8))
    for _ in range(2):
        a.applyGravity(wl.dt)
        b.applyGravity(wl.dt)
        ra, rb = a.project(wl.dt, 200, 1e-6), b.project(wl.dt, 200, 1e-6)
        assert ra.iterations == rb.iterations and ra.max_divergence == rb.max_divergence
    sa, sb = a.download(), b.download()
    for k in ("u", "v", "pressure"):
        assert np.array_equal(getattr(sa, k), getattr(sb, k))
    a.destroy()
    b.destroy()


@pytest.mark.parametrize("W,H,precond", [(96, 80, "multigrid"), (301, 203, "multigrid"), (96, 80, "diagonal"),
                                         (16, 16, "multigrid")])
def test_masked_projection_matches_oracle(oracle, W, H, precond):
    if (W, H) == (16, 16):
        wl = WL.vortex_box(16, 16, seed=9, vortices=4, sources=2)
        mask = blocked_16()
    else:
        wl = WL.obstacle_tank(W, H, seed=3)
        mask = wl.meta["solid"]
    o = oracle_grid(oracle, wl, mask, precond)
    g = hip_grid(wl, mask)
    if precond == "diagonal":
        from fluidsim_amd import FS_PRECOND_DIAGONAL
        g.setPressurePreconditioner(FS_PRECOND_DIAGONAL)
    tau = 1e-6 * o.max_divergence()
    ro = o.project(wl.dt, 20000, tau)
    rh = g.project(wl.dt, 20000, tau)
    assert ro["converged"] and rh.converged, (ro, rh)
    assert rh.max_divergence <= tau
    so, sh = o.download(), g.download()
    cu, cv = closed_faces(mask, W, H)
    assert np.all(sh.u[cu] == 0.0) and np.all(sh.v[cv] == 0.0)
    assert rel_l2(sh.u, so["u"]) < REL_L2 and rel_l2(sh.v, so["v"]) < REL_L2
    fluid = mask == 0
    po, ph = so["pressure"][fluid], sh.pressure[fluid]
    assert rel_l2(ph - ph.mean(), po - po.mean()) < REL_L2
    g.destroy()


def test_masked_gpu_multigrid_is_a_valid_preconditioner(oracle):
    W, H = 301, 203
    wl = WL.obstacle_tank(W, H, seed=3)
    mask = wl.meta["solid"]
    g = hip_grid(wl, mask)
    o = oracle_grid(oracle, wl, mask)
    mg = o.multigrid()
    i = np.arange(W * H)
    a = np.sin(i * 0.31) + np.cos(i * 0.073)
    b = np.cos(i * 0.17) - np.sin(i * 0.019)
    a -= a.mean()
    b -= b.mean()
    ma, mb = g.multigridApply(a), g.multigridApply(b)
    mc = g.multigridApply(1.3 * a - 0.7 * b)
    ma, mb, mc = ma - ma.mean(), mb - mb.mean(), mc - mc.mean()
    scale = np.abs(mc).max()
    assert np.abs(mc - (1.3 * ma - 0.7 * mb)).max() <= 5e-4 * scale
    ea, eb = a @ ma, b @ mb
    assert ea > 0 and eb > 0
    assert abs(a @ mb - b @ ma) <= 5e-4 * math.sqrt(ea * eb)
    # the FP32 device V-cycle is the FP64 one to FP32 accuracy (same hierarchy, same pins)
    ref = mg.apply(a)
    ref -= ref.mean()
    assert rel_l2(ma, ref) < 1e-4


def test_invalid_masks_are_rejected():
    wl = WL.vortex_box(32, 24, seed=4, vortices=4, sources=2)
    g = hip_grid(wl)
    with pytest.raises(ValueError):
        g.setSolidMask(np.ones(32 * 24, np.uint8))            # no fluid cell
    split = np.zeros((24, 32), np.uint8)
    split[:, 15] = 1                                            # two fluid regions
    with pytest.raises(ValueError):
        g.setSolidMask(split.ravel())
    with pytest.raises(ValueError):
        g.setSolidMask(np.zeros(10, np.uint8))                  # wrong size
    assert g.solidCells() == 0
    ring = np.ones((24, 32), np.uint8)
    ring[1:-1, 1:-1] = 0
    g.setSolidMask(ring.ravel())
    assert g.solidCells() == int(ring.sum())
    g.setSolidMask(None)
    assert g.solidCells() == 0
    g.destroy()


def test_masked_steps_keep_solid_faces_closed(oracle):
    """Full steps (transport, gravity, projection) on a masked tank: every solid face stays exactly zero
    after each step and, once the first projection has made the flow divergence-free, the liquid volume
    is conserved; the first step matches the oracle's (its transport runs on the divergent initial
    field, which changes the volume the same way in both)."""
    W, H = 64, 48
    wl = WL.obstacle_tank(W, H, seed=5)
    mask = wl.meta["solid"]
    g = hip_grid(wl, mask)
    o = oracle_grid(oracle, wl, mask)
    cu, cv = closed_faces(mask, W, H)
    ro = o.step(wl.dt, 500, 1e-6)
    rh = g.step(wl.dt, 500, 1e-6)
    assert ro["converged"] and rh.accepted()
    so, sh = o.download(), g.download()
    assert rel_l2(sh.u, so["u"]) < REL_L2 and rel_l2(sh.v, so["v"]) < REL_L2
    assert abs(sh.liquid_fraction.sum() - so["fraction"].sum()) <= 1e-12 * so["fraction"].sum()
    vol0 = sh.liquid_fraction.sum()
    for _ in range(4):
        assert g.step(wl.dt, 500, 1e-6).accepted()
        s = g.download()
        assert np.all(s.u[cu] == 0.0) and np.all(s.v[cv] == 0.0)
        assert abs(s.liquid_fraction.sum() - vol0) <= 1e-9 * vol0
    g.destroy()


_TANK_2048 = {}


def tank_2048_oracle(O):
    """The FP64 oracle's projection of the 2048 x 2048 tank, computed once for every precision."""
    if not _TANK_2048:
        wl = WL.obstacle_tank(2048, 2048, seed=3)
        mask = wl.meta["solid"]
        o = oracle_grid(O, wl, mask)
        tau = 1e-6 * o.max_divergence()
        ro = o.project(wl.dt, 2000, tau)
        assert ro["converged"]
        _TANK_2048.update(wl=wl, mask=mask, tau=tau, result=ro, state=o.download())
    return _TANK_2048


@pytest.mark.parametrize("mode", ["fp32", "fp16", "bf16"])
def test_obstacle_tank_2048_matches_oracle(oracle, mode):
    """2048 x 2048, 40 seeded circles and boxes over 15% of the cells, liquid in the bottom 70%: one
    projection to 1e-6 of the initial divergence with FP32, binary16 or bfloat16 multigrid conductances,
    against the FP64 oracle on the same input.  (With bfloat16 level vectors the solve stalls above
    this tolerance on this problem; the benchmark reports that mode.)"""
    from fluidsim_amd import FS_MG_BF16, FS_MG_FP16, FS_MG_FP32

    ref = tank_2048_oracle(oracle)
    wl, mask, tau = ref["wl"], ref["mask"], ref["tau"]
    g = hip_grid(wl, mask)
    g.setMultigridPrecision({"fp32": FS_MG_FP32, "fp16": FS_MG_FP16, "bf16": FS_MG_BF16}[mode])
    rh = g.project(wl.dt, 2000, tau)
    assert rh.converged and rh.max_divergence <= tau
    sh = g.download()
    g.destroy()
    so = ref["state"]
    print(f"obstacle tank 2048^2 {mode}: iterations gpu {rh.iterations} oracle {ref['result']['iterations']}",
          flush=True)
    assert rel_l2(sh.u, so["u"]) < REL_L2 and rel_l2(sh.v, so["v"]) < REL_L2
    fluid = mask == 0
    po, ph = so["pressure"][fluid], sh.pressure[fluid]
    assert rel_l2(ph - ph.mean(), po - po.mean()) < REL_L2
