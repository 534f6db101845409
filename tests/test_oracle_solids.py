"""Pins for the solid-mask extension of the FP64 oracle -- CPU only.

The reference has no solids (its only walls are the domain border, FluidGridCPU.cpp:1050-1061,
1170-1185), so these checks reduce the extension to cases the reference does define or to
mathematics:
  - an all-fluid mask is bit-identical to no mask (the reference path is untouched);
  - a one-cell solid ring around a (W-2) x (H-2) fluid block is the reference problem on the smaller
    grid: the same velocities come out, to solver tolerance;
  - tiny masked grids match a dense FP64 solve of the fluid cells' system (brute force);
  - solid faces carry exactly zero velocity and fluid cells end divergence-free;
  - coarse unknowns that lose every conductance are pinned (zero factor row) and the V-cycle stays a
    usable preconditioner.
"""
import numpy as np
import pytest

from fluidsim_amd import workloads as WL


def grid_for(O, wl, mask=None, preconditioner="multigrid"):
    g = O.OracleGrid(wl.width, wl.height, wl.cell_size)
    g.configure(wl.liquid_density, wl.gas_density)
    g.set_gravity(wl.gravity_y)
    g.set_preconditioner(preconditioner)
    g.upload(wl.fraction, wl.u, wl.v, None, wl.liquid_mass, wl.gas_mass)
    if mask is not None:
        g.set_solid(mask)
    return g


def closed_faces(mask, W, H):
    """Boolean masks of the u and v faces that touch a solid cell (domain walls excluded)."""
    m = mask.reshape(H, W).astype(bool)
    cu = np.zeros((H, W + 1), bool)
    cu[:, 1:W] = m[:, 1:] | m[:, :-1]
    cv = np.zeros((H + 1, W), bool)
    cv[1:H, :] = m[1:, :] | m[:-1, :]
    return cu.ravel(), cv.ravel()


def test_all_fluid_mask_is_bit_identical(oracle):
    wl = WL.dam_break(40, 28, seed=7)
    a = grid_for(oracle, wl)
    b = grid_for(oracle, wl, np.zeros(40 * 28, np.uint8))
    for _ in range(2):
        ra = a.force_project(wl.dt, 500, 1e-6)
        rb = b.force_project(wl.dt, 500, 1e-6)
        assert ra == rb
    sa, sb = a.download(), b.download()
    for k in ("u", "v", "pressure"):
        assert np.array_equal(sa[k], sb[k])


@pytest.mark.parametrize("W,H,seed", [(18, 14, 1), (33, 21, 2)])
def test_solid_ring_is_the_smaller_reference_grid(oracle, W, H, seed):
    rng = WL.SplitMix64(seed)
    h = 0.25
    f = (rng.uniform(W * H) > 0.4).astype(float)
    lm, gm = WL.two_phase_masses(f, h)
    u = rng.uniform((W + 1) * H, -1, 1)
    v = rng.uniform(W * (H + 1), -1, 1)
    WL.zero_walls(u, v, W, H)
    big = WL.Workload("ring", W, H, h, [100.0], [0.1], 0.0, 0.01, f, u, v, lm, gm)
    mask = np.ones((H, W), np.uint8)
    mask[1:-1, 1:-1] = 0
    w, hh = W - 2, H - 2
    fs = f.reshape(H, W)[1:-1, 1:-1].ravel()
    us = u.reshape(H, W + 1)[1:-1, 1:W].ravel().copy()
    vs = v.reshape(H + 1, W)[1:H, 1:-1].ravel().copy()
    WL.zero_walls(us, vs, w, hh)
    lms, gms = WL.two_phase_masses(fs, h)
    small = WL.Workload("inner", w, hh, h, [100.0], [0.1], 0.0, 0.01, fs, us, vs, lms, gms)
    gb = grid_for(oracle, big, mask.ravel())
    gs = grid_for(oracle, small)
    rb = gb.project(big.dt, 4000, 1e-11)
    rs = gs.project(small.dt, 4000, 1e-11)
    assert rb["converged"] and rs["converged"]
    sb, ss = gb.download(), gs.download()
    ub = sb["u"].reshape(H, W + 1)[1:-1, 1:W].ravel()
    vb = sb["v"].reshape(H + 1, W)[1:H, 1:-1].ravel()
    scale = max(np.abs(ss["u"]).max(), np.abs(ss["v"]).max())
    assert np.abs(ub - ss["u"]).max() <= 1e-8 * scale
    assert np.abs(vb - ss["v"]).max() <= 1e-8 * scale
    # pressure: the same up to a constant on the fluid block
    pb = sb["pressure"].reshape(H, W)[1:-1, 1:-1].ravel()
    ps = ss["pressure"]
    d = (pb - pb.mean()) - (ps - ps.mean())
    assert np.abs(d).max() <= 1e-8 * max(1.0, np.abs(ps).max())
    # every face of the ring is closed
    cu, cv = closed_faces(mask.ravel(), W, H)
    assert np.all(sb["u"][cu] == 0.0) and np.all(sb["v"][cv] == 0.0)


def random_connected_mask(W, H, rng, fraction):
    from scipy import ndimage

    m = (rng.uniform(W * H) < fraction).reshape(H, W)
    labels, count = ndimage.label(~m)
    if count > 1:
        sizes = np.bincount(labels.ravel())
        sizes[0] = 0
        m |= labels != int(np.argmax(sizes))
    return m.astype(np.uint8).ravel()


@pytest.mark.parametrize("W,H,seed", [(6, 5, 1), (8, 8, 2), (9, 7, 3), (12, 10, 4)])
def test_masked_tiny_grids_match_dense_solve(oracle, W, H, seed):
    rng = WL.SplitMix64(100 + seed)
    mask = random_connected_mask(W, H, rng, 0.25)
    f = (rng.uniform(W * H) > 0.5).astype(float)
    lm, gm = WL.two_phase_masses(f, 1.0)
    u = rng.uniform((W + 1) * H, -1, 1)
    v = rng.uniform(W * (H + 1), -1, 1)
    WL.zero_walls(u, v, W, H)
    wl = WL.Workload("tiny", W, H, 1.0, [100.0], [0.1], 0.0, 0.01, f, u, v, lm, gm)
    g = grid_for(oracle, wl, mask)
    iu, iv = g.face_weights()
    cu, cv = closed_faces(mask, W, H)
    assert np.all(iu[cu] == 0.0) and np.all(iv[cv] == 0.0)
    assert np.all(iu.reshape(H, W + 1)[:, 1:W][~cu.reshape(H, W + 1)[:, 1:W]] > 0.0)
    fluid = np.flatnonzero(mask == 0)
    r = g.project(wl.dt, 4000, 1e-13)
    assert r["converged"]
    s = g.download()
    # rhs of the fluid cells from the velocities the projection's wall step left
    u0, v0 = u.copy(), v.copy()
    u0[cu] = 0.0
    v0[cv] = 0.0
    scale = 1.0 / wl.dt
    div = ((u0.reshape(H, W + 1)[:, 1:] - u0.reshape(H, W + 1)[:, :-1]) +
           (v0.reshape(H + 1, W)[1:, :] - v0.reshape(H + 1, W)[:-1, :])).ravel()
    b = -div * scale
    A = np.zeros((W * H, W * H))
    for y in range(H):
        for x in range(W):
            i = y * W + x
            for (nx, ny, wgt) in ((x - 1, y, iu[y * (W + 1) + x]), (x + 1, y, iu[y * (W + 1) + x + 1]),
                                  (x, y - 1, iv[y * W + x]), (x, y + 1, iv[(y + 1) * W + x])):
                if 0 <= nx < W and 0 <= ny < H:
                    A[i, i] += wgt
                    A[i, ny * W + nx] -= wgt
    Af = A[np.ix_(fluid, fluid)]
    assert np.abs(b[mask != 0]).max() == 0.0       # no flux into a solid cell
    exact = np.linalg.lstsq(Af, b[fluid], rcond=None)[0]
    p = s["pressure"][fluid]
    d = (p - p.mean()) - (exact - exact.mean())
    assert np.abs(d).max() <= 1e-9 * max(1.0, np.abs(exact).max())
    assert np.all(s["u"][cu] == 0.0) and np.all(s["v"][cv] == 0.0)
    assert r["max_divergence"] <= 1e-13


def test_blocked_coarse_cells_are_pinned(oracle):
    """A solid 2 x 2 block aligned with the coarse grid leaves its coarse cell without conductance: its
    pivot vanishes, it is pinned, and the projection converges."""
    W = H = 16   # two levels: 16 x 16 -> 8 x 8 (the pinned coarse system)
    mask = np.zeros((H, W), np.uint8)
    mask[6:8, 10:12] = 1            # coarse cell (5, 3): every fine cell solid
    mask[12:14, 2:4] = 1            # and a second one, coarse cell (1, 6)
    from scipy import ndimage
    labels, count = ndimage.label(mask == 0)
    assert count == 1
    wl = WL.vortex_box(W, H, seed=9, vortices=4, sources=2)
    g = grid_for(oracle, wl, mask.ravel())
    mg = g.multigrid()
    assert mg is not None
    pinned = mg.pinned()
    # coarse cell (5, 3) is unknown 3 * 8 + 5 - 1 (the first cell is the reference's pinned one)
    assert pinned[3 * 8 + 5 - 1] and pinned[6 * 8 + 1 - 1] and pinned.sum() == 2
    f = mg.coarse_factor()
    assert np.all(f[pinned, :] == 0.0) and np.all(f[:, pinned] == 0.0)
    assert np.all(np.diag(f)[~pinned] > 0.0)
    r = g.project(wl.dt, 500, 1e-10)
    assert r["converged"] and r["iterations"] < 60
    # without pinning the same hierarchy cannot be factored
    raw = oracle.Multigrid(W, H)
    iu, iv = g.face_weights()
    assert not raw.rebuild(iu, iv)
    raw.set_pin(True)
    assert raw.rebuild(iu, iv)
    assert np.array_equal(raw.coarse_factor(), f)


def test_obstacle_tank_projection(oracle):
    wl = WL.obstacle_tank(96, 80, seed=3)
    mask = wl.meta["solid"]
    assert 0.13 < mask.mean() < 0.17
    g = grid_for(oracle, wl, mask)
    r = g.project(wl.dt, 500, 1e-6)
    assert r["converged"] and r["max_divergence"] <= 1e-6
    s = g.download()
    cu, cv = closed_faces(mask, 96, 80)
    assert np.all(s["u"][cu] == 0.0) and np.all(s["v"][cv] == 0.0)
    # the multigrid preconditioner with pinned coarse unknowns beats the diagonal one
    gd = grid_for(oracle, wl, mask, "diagonal")
    rd = gd.project(wl.dt, 5000, 1e-6)
    assert rd["converged"] and r["iterations"] * 3 < rd["iterations"]
    sd = gd.download()
    scale = np.abs(sd["u"]).max()
    assert np.abs(s["u"] - sd["u"]).max() <= 1e-5 * scale


def test_masked_multigrid_is_symmetric_positive_and_linear(oracle):
    """The preconditioner contract of tests/FluidGridTests.cpp:538-625 on a masked hierarchy: cells
    without conductance are neither read (their residual is not restricted) nor written."""
    W, H = 96, 80
    wl = WL.obstacle_tank(W, H, seed=3)
    mask = wl.meta["solid"]
    g = grid_for(oracle, wl, mask)
    mg = g.multigrid()
    rng = np.random.default_rng(0)
    for _ in range(3):
        x = rng.standard_normal(W * H)
        y = rng.standard_normal(W * H)
        x -= x.mean()
        y -= y.mean()
        mx, my = mg.apply(x), mg.apply(y)
        assert abs(y @ mx - x @ my) <= 1e-12 * np.abs(x) @ np.abs(my)
        assert x @ mx > 0.0
        assert np.allclose(mg.apply(2.0 * x + y), 2.0 * mx + my, rtol=0, atol=1e-11 * np.abs(mx).max())
        # solid cells: no output, and their input is ignored
        assert np.all(mx[mask != 0] == 0.0)
        x2 = x.copy()
        x2[mask != 0] += 5.0
        assert np.array_equal(mg.apply(x2), mx)
