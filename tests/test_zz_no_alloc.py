"""fs_viscosity_solve_noalloc: the allocation-free viscosity solve, through the C ABI.

It runs fs_viscosity_solve's algorithm on device arrays with its workspace in a static device array, so
on the same input it must return the same iteration count, relative residual and velocities as
fs_viscosity_solve, bit for bit.  Two calls in a row, the second on a smaller grid (the array keeps the
first call's values, which the second must not see), then a grid too large for the array, refused before
any device work.
"""
import ctypes as C

import numpy as np
import pytest

from fluidsim_amd import _abi
from fluidsim_amd.grid import Context, viscosity_solve

from test_oracle_viscosity import variable_case

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def noalloc(ctx, case, dt, tol=1e-4, maximum=-1):
    dev = torch.device("cuda:0")
    m = torch.from_numpy(np.ascontiguousarray(case.material, dtype=np.float64)).to(dev)
    u = torch.from_numpy(np.ascontiguousarray(case.u, dtype=np.float64)).to(dev)
    v = torch.from_numpy(np.ascontiguousarray(case.v, dtype=np.float64)).to(dev)
    torch.cuda.synchronize()
    it = C.c_int()
    rel = C.c_double()
    status = _abi.lib().fs_viscosity_solve_noalloc(ctx.handle, case.w, case.h, case.dx, C.c_void_p(m.data_ptr()),
                                                   C.c_void_p(u.data_ptr()), C.c_void_p(v.data_ptr()), dt, tol,
                                                   maximum, C.byref(it), C.byref(rel))
    torch.cuda.synchronize()
    return status, it.value, rel.value, u.cpu().numpy(), v.cpu().numpy()


def test_noalloc_matches_the_allocating_solve_twice():
    ctx = Context(0)
    for w, h, dt in ((96, 72, 0.02), (41, 29, 0.05)):
        case = variable_case(w, h)
        print(f"case {w}x{h} dt {dt}", flush=True)
        it_ref, rel_ref, u_ref, v_ref = viscosity_solve(w, h, case.dx, case.material, case.u, case.v, dt,
                                                        context=ctx)
        status, it, rel, u, v = noalloc(ctx, case, dt)
        assert status == _abi.FS_OK, _abi.lib().fs_last_error()
        assert it == it_ref > 0
        assert rel == rel_ref
        assert np.array_equal(u, u_ref) and np.array_equal(v, v_ref)
    ctx.close()


def test_noalloc_refuses_a_grid_beyond_its_scratch():
    ctx = Context(0)
    dev = torch.device("cuda:0")
    tiny = torch.zeros(8, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    it = C.c_int()
    rel = C.c_double()
    p = C.c_void_p(tiny.data_ptr())
    status = _abi.lib().fs_viscosity_solve_noalloc(ctx.handle, 1500, 1500, 1.0, p, p, p, 0.01, 1e-4, -1,
                                                   C.byref(it), C.byref(rel))
    assert status == _abi.FS_E_INVALID_ARGUMENT
    assert b"scratch" in _abi.lib().fs_last_error()
    ctx.close()
