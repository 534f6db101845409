"""Why reduced-precision Krylov vectors cost iterations on the obstacle tank -- CPU only.

PCG in the reference's loop (FluidGridCPU.cpp:1728-1824: max-norm stop checked on the true residual)
with the oracle's FP64 operator and V-cycle, on the masked 1000:1 obstacle tank at 128^2, once in FP64 and
once with the preconditioner's output z rounded after every application: to FP32 it needs clearly more
iterations (loss of A-conjugacy, the GPU's FP32 iterate and preconditioner pay the same), to bfloat16 it
stalls.  These are properties of the algorithm in these precisions, independent of any GPU code.
"""
import numpy as np

from fluidsim_amd import workloads as WL


def bf16(v):
    u = np.ascontiguousarray(v, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def fp32(v):
    return v.astype(np.float32).astype(np.float64)


def tank(oracle, size=128):
    wl = WL.obstacle_tank(size, size, seed=3)
    g = oracle.OracleGrid(size, size, wl.cell_size)
    g.configure(wl.liquid_density, wl.gas_density)
    g.set_gravity(wl.gravity_y)
    g.set_solid(wl.meta["solid"])
    g.upload(wl.fraction, wl.u, wl.v, None, wl.liquid_mass, wl.gas_mass)
    tau = 1e-6 * g.max_divergence()
    g.project(wl.dt, 0, tau)   # the projection's wall step closes the solid faces; the pressure stays 0
    b = g.pressure_rhs(wl.dt)
    return g, b, tau * wl.cell_size ** 2 / wl.dt


def pcg(g, b, tol, rounder=None, maxit=300):
    mg = g.multigrid()
    A = lambda p: g.apply_operator(p)[0]   # noqa: E731

    def M(r):
        z = mg.apply(r, r.mean())
        z = z - z.mean()
        return rounder(z) if rounder else z

    x = np.zeros(b.size)
    r = b.copy()
    z = M(r)
    d = z.copy()
    rz = r @ z
    for it in range(1, maxit + 1):
        q = A(d)
        alpha = rz / (d @ q)
        x += alpha * d
        r -= alpha * q
        if np.abs(r).max() <= tol and np.abs(b - A(x)).max() <= tol:
            return it
        z = M(r)
        rz, old = r @ z, rz
        d = z + (rz / old) * d
    return None


def test_rounded_preconditioner_output_costs_iterations(oracle):
    g, b, tol = tank(oracle)
    n64 = pcg(g, b, tol)
    n32 = pcg(g, b, tol, fp32)
    nbf = pcg(g, b, tol, bf16, maxit=4 * n64)
    print(f"FP64 {n64}, z rounded to FP32 {n32}, z rounded to bfloat16 {nbf}")
    assert n64 is not None and n32 is not None
    assert n32 >= 1.15 * n64
    assert nbf is None
