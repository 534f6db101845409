"""Python mirror of the serving loop (fs_server_* in include/fluidsim_hip.h).

A `Server` keeps many simulations resident on one GPU and steps them concurrently on `lanes` HIP
streams, admitting requests in arrival order within an HBM budget.  Every call goes to
libfluidsim_hip.so; nothing is computed here.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import _abi
from ._abi import FsRequest, FsRequestReport, FsServerConfig, FsServerStats, call, lib


@dataclass
class Request:
    """One simulation to serve: a grid, its materials and initial state, and the steps to run."""
    width: int
    height: int
    cell_size: float = 1.0
    steps: int = 1
    dt: float = 0.01
    max_iterations: int = -1
    divergence_tolerance: float = 1e-6
    gravity_y: float = 0.0
    surface_tension: float = 0.0
    transport: int = _abi.FS_TRANSPORT_REFERENCE
    preconditioner: int = _abi.FS_PRECOND_MULTIGRID
    liquid_density: List[float] = field(default_factory=lambda: [100.0])
    gas_density: List[float] = field(default_factory=lambda: [0.1])
    liquid_viscosity: Optional[List[float]] = None
    gas_viscosity: Optional[List[float]] = None
    liquid_fraction: Optional[np.ndarray] = None
    u: Optional[np.ndarray] = None
    v: Optional[np.ndarray] = None
    liquid_mass: Optional[np.ndarray] = None     # component-major, n_liquid * W * H
    gas_mass: Optional[np.ndarray] = None
    multigrid_precision: int = _abi.FS_MG_FP32   # storage of the multigrid levels (fs_mg_precision)
    solid: Optional[np.ndarray] = None           # W * H obstacle mask, nonzero = solid
    relative_tolerance: bool = False             # divergence_tolerance relative to each projection's max|b|
    pressure_smoother: int = _abi.FS_SMOOTHER_RED_BLACK   # fs_set_pressure_smoother; smoother_sweeps 0 keeps
    smoother_sweeps: int = 0                              # the reference's V(4,4)
    coarse_sweeps: int = 0
    viscosity_sweeps: int = 0                             # fs_set_viscosity_smoother (0: the reference's)

    def estimate_bytes(self) -> int:
        viscous = any(m > 0 for m in (self.liquid_viscosity or []) + (self.gas_viscosity or []))
        features = ((_abi.FS_FEATURE_TRANSPORT if self.transport == _abi.FS_TRANSPORT_REFERENCE else 0)
                    | (_abi.FS_FEATURE_VISCOSITY if viscous else 0)
                    | (_abi.FS_FEATURE_SURFACE_TENSION if self.surface_tension > 0 else 0)
                    | (_abi.FS_FEATURE_SOLID_MASK if self.solid is not None else 0)
                    | (_abi.FS_FEATURE_MG16 if self.multigrid_precision != _abi.FS_MG_FP32 else 0))
        return estimate_bytes(self.width, self.height, len(self.liquid_density), len(self.gas_density), features)


@dataclass
class Report:
    """fs_request_report."""
    state: int
    lane: int
    steps_done: int
    pressure_iterations: int
    viscosity_iterations: int
    physical_steps: int
    max_divergence: float
    queue_ms: float
    run_ms: float
    step_p50_ms: float
    step_p99_ms: float
    step_max_ms: float
    estimated_bytes: int
    device_bytes: int
    setup_ms: float
    teardown_ms: float
    evictions: int
    error: str

    @property
    def done(self) -> bool:
        return self.state == _abi.FS_REQUEST_DONE

    @property
    def finished(self) -> bool:
        return self.state in (_abi.FS_REQUEST_DONE, _abi.FS_REQUEST_FAILED)


def estimate_bytes(width: int, height: int, n_liquid: int, n_gas: int, features: int) -> int:
    b = C.c_size_t()
    call("fs_grid_estimate_bytes", int(width), int(height), int(n_liquid), int(n_gas), int(features), C.byref(b))
    return b.value


def _arr(a, n=None):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float64).ravel()
    if n is not None and a.size != n:
        raise ValueError(f"request array has {a.size} values, expected {n}")
    return a


class Server:
    """fs_server: lanes worker streams, admission within hbm_budget bytes (0 = 80% of free memory); instances
    of at most batch_cells cells (0: none) take their steps together in one fs_batch_step per lane round;
    with evict_after_steps > 0, instances resident for that many steps are checkpointed to host memory when
    a waiting request needs their room, and resume from the checkpoint when readmitted."""

    def __init__(self, lanes: int = 4, hbm_budget: int = 0, device: int = 0, batch_cells: int = 0,
                 evict_after_steps: int = 0):
        cfg = FsServerConfig(int(device), int(lanes), int(hbm_budget), int(batch_cells), int(evict_after_steps))
        h = C.c_void_p()
        call("fs_server_create", C.byref(cfg), C.byref(h))
        self._h = h

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            call("fs_server_destroy", self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, r: Request) -> int:
        W, H = int(r.width), int(r.height)
        N = W * H
        nl, ng = len(r.liquid_density), len(r.gas_density)
        keep = [_arr(r.liquid_density, nl), _arr(r.liquid_viscosity, nl), _arr(r.gas_density, ng),
                _arr(r.gas_viscosity, ng), _arr(r.liquid_fraction, N), _arr(r.u, (W + 1) * H),
                _arr(r.v, W * (H + 1)), _arr(r.liquid_mass, nl * N), _arr(r.gas_mass, ng * N)]
        ptrs = [None if a is None else a.ctypes.data_as(C.POINTER(C.c_double)) for a in keep]
        solid = None
        if r.solid is not None:
            solid = np.ascontiguousarray(r.solid, dtype=np.uint8).ravel()
            if solid.size != N:
                raise ValueError(f"solid mask has {solid.size} cells, expected {N}")
        req = FsRequest(W, H, float(r.cell_size), int(r.steps), float(r.dt), int(r.max_iterations),
                        float(r.divergence_tolerance), float(r.gravity_y), float(r.surface_tension), int(r.transport),
                        int(r.preconditioner), nl, ng, *ptrs, int(r.multigrid_precision),
                        None if solid is None else solid.ctypes.data, int(bool(r.relative_tolerance)),
                        int(r.pressure_smoother), int(r.smoother_sweeps), int(r.coarse_sweeps),
                        int(r.viscosity_sweeps))
        rid = C.c_longlong()
        call("fs_server_submit", self._h, C.byref(req), C.byref(rid))   # copies the arrays
        return rid.value

    @staticmethod
    def _report(rep: FsRequestReport) -> Report:
        vals = [getattr(rep, n) for n, _ in FsRequestReport._fields_]
        vals[-1] = rep.error.decode(errors="replace")
        return Report(*vals)

    def wait(self, rid: int, timeout: float = -1.0) -> Report:
        rep = FsRequestReport()
        call("fs_server_wait", self._h, int(rid), float(timeout), C.byref(rep))
        return self._report(rep)

    def query(self, rid: int) -> Report:
        rep = FsRequestReport()
        call("fs_server_query", self._h, int(rid), C.byref(rep))
        return self._report(rep)

    def result(self, rid: int, width: int, height: int, n_liquid: int = 1, n_gas: int = 1) -> dict:
        N = width * height
        out = dict(fraction=np.zeros(N), u=np.zeros((width + 1) * height), v=np.zeros(width * (height + 1)),
                   pressure=np.zeros(N), liquid_mass=np.zeros(n_liquid * N), gas_mass=np.zeros(n_gas * N))
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double)) if a.size else None
        call("fs_server_result", self._h, int(rid), dp(out["fraction"]), dp(out["u"]), dp(out["v"]),
             dp(out["pressure"]), dp(out["liquid_mass"]), dp(out["gas_mass"]))
        return out

    def release(self, rid: int):
        call("fs_server_release", self._h, int(rid))

    def drain(self):
        call("fs_server_drain", self._h)

    def stats(self) -> dict:
        s = FsServerStats()
        call("fs_server_get_stats", self._h, C.byref(s))
        return {n: getattr(s, n) for n, _ in FsServerStats._fields_}

    def reset_stats(self):
        call("fs_server_reset_stats", self._h)
