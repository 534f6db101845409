"""Host-side mirror of the reference backend interface over the HIP C ABI.

`FluidGridHIP` implements the method set of fluid::FluidGridBase (FluidSim/FluidGridBase.h:20-147)
with the reference's names, argument meaning and error behaviour, so code written against
FluidGridCPU (FluidSim/UniformCPU/FluidGridCPU.h:15-225) reads the same here.  Every call goes
through libfluidsim_hip.so; nothing is computed on the host.
"""
from __future__ import annotations

import ctypes as C
import json
import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import _abi
from ._abi import FsFrameSettings, FsFrameStatistics, FsProjectionResult, FsStepTimings, call, lib


# ---- value types mirroring FluidGridTypes.h / FluidMixture.h ----

@dataclass
class FluidGridState:
    """FluidGridState<double> (FluidGridTypes.h:10-20); empty arrays mean zero."""
    liquid_fraction: np.ndarray = field(default_factory=lambda: np.zeros(0))
    u: np.ndarray = field(default_factory=lambda: np.zeros(0))
    v: np.ndarray = field(default_factory=lambda: np.zeros(0))
    pressure: np.ndarray = field(default_factory=lambda: np.zeros(0))
    liquid_mass: List[np.ndarray] = field(default_factory=list)
    gas_mass: List[np.ndarray] = field(default_factory=list)


@dataclass
class ProjectionResult:
    """FluidProjectionResult (FluidGridTypes.h:59-79) plus solver diagnostics."""
    iterations: int = 0
    max_divergence: float = math.inf
    converged: bool = False
    physical_steps: int = 1
    transport_steps: int = 0
    relaxed: bool = False
    max_volume_error: float = 0.0
    rms_volume_error: float = 0.0
    viscosity_iterations: int = 0
    precision_limited: bool = False
    restarts: int = 0
    initial_residual: float = 0.0
    final_residual: float = 0.0
    rhs_max: float = 0.0

    def accepted(self) -> bool:
        return bool(self.converged or self.relaxed or self.precision_limited)

    @classmethod
    def from_c(cls, r: FsProjectionResult) -> "ProjectionResult":
        return cls(r.iterations, r.max_divergence, bool(r.converged), r.physical_steps, r.transport_steps,
                   bool(r.relaxed), r.max_volume_error, r.rms_volume_error, r.viscosity_iterations,
                   bool(r.precision_limited), r.restarts, r.initial_residual, r.final_residual, r.rhs_max)


@dataclass
class StepTimings:
    """FluidStepTimings (FluidGridTypes.h:49-57), milliseconds."""
    transport_ms: float = 0.0
    viscosity_ms: float = 0.0
    forces_ms: float = 0.0
    pressure_ms: float = 0.0
    selection_ms: float = 0.0
    regrid_ms: float = 0.0


@dataclass
class FluidCheckpoint:
    """FluidCheckpoint (FluidSim/FluidCheckpoint.h:9-15): backend-owned JSON metadata and native payload."""
    metadata: dict = field(default_factory=dict)
    data: bytes = b""


@dataclass
class FrameStatistics:
    """FluidScene::FrameStatistics (FluidScene.h:162-178) plus the frame's schedule and outcome."""
    fluid_ms: float = 0.0
    stages: StepTimings = field(default_factory=StepTimings)
    substeps: int = 0
    transport_steps: int = 0
    pressure_iterations: int = 0
    viscosity_iterations: int = 0
    projection_tolerance: float = 0.0
    max_volume_error: float = 0.0
    rms_volume_error: float = 0.0
    steps_scheduled: int = 0
    steps_run: int = 0
    accepted: bool = False
    max_divergence: float = math.inf


@dataclass
class MixtureComponent:
    """MixtureComponent (FluidMixture.h:14-23)."""
    substance: str
    density: float
    dynamic_viscosity: float = 0.0


@dataclass
class FluidMixture:
    """FluidMixture (FluidMixture.h:25-39)."""
    components: List[MixtureComponent] = field(default_factory=list)

    def hasViscosity(self) -> bool:
        return any(c.dynamic_viscosity != 0.0 for c in self.components)


@dataclass
class Substance:
    """Substance (SubstanceCatalogue.h:14-21): a name, an RGB colour and an optional composition."""
    name: str
    color: tuple = (1.0, 1.0, 1.0)
    composition: list = field(default_factory=list)


class SubstanceCatalogue:
    """SubstanceCatalogue (SubstanceCatalogue.h:23-59): substances addressed by SubstanceId (their index)."""

    def __init__(self, substances=()):
        self._items: List[Substance] = list(substances)

    def add(self, substance: Substance) -> int:
        self._items.append(substance)
        return len(self._items) - 1

    def __getitem__(self, substance_id: int) -> Substance:
        return self._items[substance_id]

    def __len__(self) -> int:
        return len(self._items)

    def size(self) -> int:
        return len(self._items)

    def find(self, name: str) -> int:
        for k, s in enumerate(self._items):
            if s.name == name:
                return k
        raise KeyError(f"no substance named {name!r} in the catalogue")


LIQUID, GAS = "liquid", "gas"   # Phase (FluidMixture.h:8-12)


class FluidGridEdit:
    """FluidGridEdit<Grid> (FluidGridEdit.h:97-149): staged material replacements; only commit() touches
    the grid.  setCell(x, y, phase, substance) stages one change (the last one of a cell wins); commit()
    resolves them (CellEditBatch::prepare) and applies them through fs_commit_edits, then clears."""

    def __init__(self, grid: "FluidGridHIP"):
        self.grid = grid
        self.changes: list = []

    def size(self) -> int:
        return len(self.changes)

    def empty(self) -> bool:
        return not self.changes

    def clear(self):
        self.changes.clear()

    def setCell(self, x: int, y: int, phase: str, substance):
        self.changes.append((int(x), int(y), phase, substance))

    def commit(self, velocity=None):
        if self.empty():
            return
        self.grid.commitEdits(self.changes, velocity)
        self.clear()


# ---- devices ----

class Context:
    """One HIP device and its streams (fs_context): the work stream plus a helper stream for hierarchy
    rebuilds, or the work stream alone with single_stream=True (fs_context_create_ex,
    FS_CONTEXT_SINGLE_STREAM); private_queue=True puts the work stream on a hardware queue of its own
    (FS_CONTEXT_PRIVATE_QUEUE)."""

    def __init__(self, device: int = 0, single_stream: bool = False, private_queue: bool = False):
        h = C.c_void_p()
        flags = (1 if single_stream else 0) | (2 if private_queue else 0)
        if flags:
            call("fs_context_create_ex", int(device), flags, C.byref(h))
        else:
            call("fs_context_create", int(device), C.byref(h))
        self._h = h
        self.device = device

    @property
    def handle(self):
        return self._h

    def synchronize(self):
        call("fs_context_synchronize", self._h)

    def set_batch_split(self, solo_cells: int):
        """fs_context_set_batch_split: batch_project sends grids of more than `solo_cells` cells through
        the multi-kernel solve, concurrently with the persistent launch of the rest (0: batch all)."""
        call("fs_context_set_batch_split", self._h, int(solo_cells))

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().fs_context_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_contexts = {}


def default_context(device: int = 0) -> Context:
    ctx = _default_contexts.get(device)
    if ctx is None:
        ctx = Context(device)
        _default_contexts[device] = ctx
    return ctx


def _dp(a: Optional[np.ndarray]):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(C.c_double))


def viscosity_solve(width: int, height: int, cell_size: float, material, u, v, dt: float, tolerance: float = 1e-4,
                    maximum: int = -1, context: Optional[Context] = None):
    """ViscositySolverGPU::solve (ViscositySolverGPU.cpp:169-294) on host arrays.  material holds (rho, mu)
    per cell.  Returns (iterations, relative_residual, u, v); raises RuntimeError (inputs untouched) on
    breakdown, non-finite input or an exhausted budget."""
    ctx = context or default_context()
    m = np.ascontiguousarray(material, dtype=np.float64).ravel()
    uu = np.array(u, dtype=np.float64).ravel()
    vv = np.array(v, dtype=np.float64).ravel()
    if m.size != 2 * width * height or uu.size != (width + 1) * height or vv.size != width * (height + 1):
        raise ValueError("viscosity_solve dimensions do not match the grid")
    it = C.c_int()
    rel = C.c_double()
    call("fs_viscosity_solve", ctx.handle, int(width), int(height), float(cell_size), _dp(m), _dp(uu), _dp(vv),
         float(dt), float(tolerance), int(maximum), C.byref(it), C.byref(rel))
    return it.value, rel.value, uu, vv


def batch_project(grids, dt: float, max_iterations: int = 200, divergence_tolerance=1e-6):
    """fs_batch_project: project every grid (all on one context) with one persistent launch running all
    their PCG iterations.  `divergence_tolerance` is one value or one per grid.  Returns one
    ProjectionResult per grid, with fs_project's contract."""
    grids = list(grids)
    n = len(grids)
    tol = np.broadcast_to(np.asarray(divergence_tolerance, dtype=np.float64), (n,)).copy()
    handles = (C.c_void_p * max(n, 1))(*[g.handle.value for g in grids])
    res = (FsProjectionResult * max(n, 1))()
    call("fs_batch_project", n, handles, float(dt), int(max_iterations), _dp(tol), res)
    return [ProjectionResult.from_c(res[i]) for i in range(n)]


def tune_batch_smoother(grids, dt: float, max_iterations: int = 200, divergence_tolerance=1e-6):
    """fs_tune_batch_smoother: time fs_batch_project on the grids with every grid at red-black (sweeps,
    coarse) for sweeps 4, 3, 2 and coarse 4, 2, from their current state (unchanged on return), and set
    every grid to the fastest choice under which all converge.  Returns (sweeps, coarse_sweeps, ms)."""
    grids = list(grids)
    n = len(grids)
    tol = np.broadcast_to(np.asarray(divergence_tolerance, dtype=np.float64), (n,)).copy()
    handles = (C.c_void_p * max(n, 1))(*[g.handle.value for g in grids])
    sweeps, coarse, ms = C.c_int(), C.c_int(), C.c_double()
    call("fs_tune_batch_smoother", n, handles, float(dt), int(max_iterations), _dp(tol), C.byref(sweeps),
         C.byref(coarse), C.byref(ms))
    return sweeps.value, coarse.value, ms.value


def batch_step(grids, dt, max_iterations=200, divergence_tolerance=1e-6, raise_on_error=True):
    """fs_batch_step: fs_step on every grid (all on one context) at once, each with its own dt, budget and
    tolerance (one value or one per grid); viscosity solves and projections run batched.  Returns
    (results, statuses): one ProjectionResult per grid with fs_step's contract and one fs_status per grid.
    With raise_on_error, a failing grid raises after every other grid has been stepped."""
    grids = list(grids)
    n = len(grids)
    dts = np.broadcast_to(np.asarray(dt, dtype=np.float64), (n,)).copy()
    tol = np.broadcast_to(np.asarray(divergence_tolerance, dtype=np.float64), (n,)).copy()
    budgets = (C.c_int * max(n, 1))(*np.broadcast_to(np.asarray(max_iterations, dtype=np.int64), (n,)).tolist())
    handles = (C.c_void_p * max(n, 1))(*[g.handle.value for g in grids])
    res = (FsProjectionResult * max(n, 1))()
    statuses = (C.c_int * max(n, 1))()
    if raise_on_error:
        call("fs_batch_step", n, handles, _dp(dts), budgets, _dp(tol), res, statuses)
    else:
        lib().fs_batch_step(n, handles, _dp(dts), budgets, _dp(tol), res, statuses)
    return [ProjectionResult.from_c(res[i]) for i in range(n)], [statuses[i] for i in range(n)]


def batch_apply_viscosity(grids, dt: float):
    """fs_batch_apply_viscosity: every grid's implicit viscosity step (all on one context) with one
    persistent launch running all their CG loops.  Returns each grid's iteration count."""
    grids = list(grids)
    n = len(grids)
    handles = (C.c_void_p * max(n, 1))(*[g.handle.value for g in grids])
    its = (C.c_int * max(n, 1))()
    call("fs_batch_apply_viscosity", n, handles, float(dt), its)
    return [its[i] for i in range(n)]


def _finite_json(value) -> bool:
    """True if every number in a JSON-like value is finite (JSON has no NaN or infinity)."""
    if isinstance(value, float):
        return math.isfinite(value)
    if isinstance(value, dict):
        return all(_finite_json(v) for v in value.values())
    if isinstance(value, (list, tuple)):
        return all(_finite_json(v) for v in value)
    return True


class FluidGridHIP:
    """Uniform MAC grid backend on one MI355X (the FluidGridBase method set)."""

    Diagonal = _abi.FS_PRECOND_DIAGONAL
    Multigrid = _abi.FS_PRECOND_MULTIGRID

    def __init__(self, width: int, height: int, cell_size: float, context: Optional[Context] = None):
        self.context = context or default_context(0)
        h = C.c_void_p()
        call("fs_grid_create", self.context.handle, int(width), int(height), float(cell_size), C.byref(h))
        self._h = h
        self.width, self.height, self.cell_size = int(width), int(height), float(cell_size)
        self.liquid_mixture = FluidMixture()
        self.gas_mixture = FluidMixture()
        self.substances: list = []
        self._surface_tension = 0.0

    # -- lifetime --
    def destroy(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().fs_grid_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    # -- geometry (FluidGridBase.h:31-39) --
    def getWidth(self) -> int:
        return self.width

    def getHeight(self) -> int:
        return self.height

    def getCellSize(self) -> float:
        return self.cell_size

    def getOrigin(self):
        x, y = C.c_double(), C.c_double()
        call("fs_get_origin", self._h, C.byref(x), C.byref(y))
        return (x.value, y.value)

    def setOrigin(self, origin):
        call("fs_set_origin", self._h, float(origin[0]), float(origin[1]))

    def requiresRenderThread(self) -> bool:
        return False

    # -- materials (FluidGridBase.h:42-54) --
    def configureMaterials(self, substances, liquid: FluidMixture, gas: FluidMixture):
        self.substances = list(substances)
        self.liquid_mixture, self.gas_mixture = liquid, gas
        ld = np.array([c.density for c in liquid.components], dtype=np.float64)
        lv = np.array([c.dynamic_viscosity for c in liquid.components], dtype=np.float64)
        gd = np.array([c.density for c in gas.components], dtype=np.float64)
        gv = np.array([c.dynamic_viscosity for c in gas.components], dtype=np.float64)
        call("fs_configure_materials", self._h, ld.size, _dp(ld), _dp(lv), gd.size, _dp(gd), _dp(gv))

    def getSubstances(self):
        return self.substances

    def _substance_id(self, substance):
        """A component's SubstanceId; a name is looked up in the catalogue (mixtures built by name)."""
        if isinstance(substance, str):
            if isinstance(self.substances, SubstanceCatalogue):
                return self.substances.find(substance)
            for k, s in enumerate(self.substances):
                if getattr(s, "name", s) == substance:
                    return k
            raise KeyError(f"no substance named {substance!r} in the catalogue")
        return int(substance)

    def _component_colors(self) -> np.ndarray:
        rgb = []
        for comp in self.liquid_mixture.components + self.gas_mixture.components:
            sub = self.substances[self._substance_id(comp.substance)]
            rgb.extend(float(c) for c in sub.color)
        return np.asarray(rgb, dtype=np.float32)

    # -- rendering and edits (FluidGridBase.h:139-142) --
    def writeColorPixels(self, pixels: Optional[np.ndarray] = None, liquid_fraction_alpha: bool = False) -> np.ndarray:
        """writeColorPixels (FluidGridCPU.cpp:1826-1860): RGBA float32 per cell, the component colours of the
        substance catalogue weighted by volume fraction.  Reuses `pixels` when it has 4*W*H float32 slots."""
        n = 4 * self.width * self.height
        if pixels is None or pixels.dtype != np.float32 or pixels.size != n or not pixels.flags.c_contiguous:
            pixels = np.empty(n, dtype=np.float32)
        rgb = self._component_colors()
        call("fs_color", self._h, rgb.ctypes.data_as(C.POINTER(C.c_float)) if rgb.size else None,
             int(bool(liquid_fraction_alpha)), pixels.ctypes.data_as(C.POINTER(C.c_float)))
        return pixels

    def edit(self) -> FluidGridEdit:
        return FluidGridEdit(self)

    def commitEdits(self, changes, velocity=None):
        """commitEdits (FluidGridCPU.cpp:253-284) of [(x, y, phase, substance)] changes, resolved as
        CellEditBatch::prepare does (FluidGridEdit.h:37-95): IndexError for a cell outside the grid,
        ValueError for a substance the phase does not hold -- both before anything changes."""
        if velocity is not None:
            raise ValueError("Material edit velocity requires uniform WebGPU")
        nl = len(self.liquid_mixture.components)
        cell_volume = self.cell_size * self.cell_size
        cells = (_abi.FsCellEdit * max(1, len(changes)))()
        last_key, component, mass = None, 0, 0.0
        for i, (x, y, phase, substance) in enumerate(changes):
            if not (0 <= x < self.width and 0 <= y < self.height):
                raise IndexError("Fluid grid edit cell is outside the grid")
            key = (phase, substance)
            if key != last_key:
                mixture = self.liquid_mixture if phase == LIQUID else self.gas_mixture
                want = self._substance_id(substance)
                found = [k for k, c in enumerate(mixture.components) if self._substance_id(c.substance) == want]
                if not found:
                    raise ValueError("Fluid grid edit substance is not configured in the selected phase")
                component = found[0] + (0 if phase == LIQUID else nl)
                mass = mixture.components[found[0]].density * cell_volume
                last_key = key
            cells[i].x, cells[i].y, cells[i].component, cells[i].mass = x, y, component, mass
        call("fs_commit_edits", self._h, len(changes), cells, None)

    def getLiquidMixture(self) -> FluidMixture:
        return self.liquid_mixture

    def getGasMixture(self) -> FluidMixture:
        return self.gas_mixture

    def hasViscosity(self) -> bool:
        return self.liquid_mixture.hasViscosity() or self.gas_mixture.hasViscosity()

    def setComponentViscosity(self, phase, index: int, coefficient: float):
        """setComponentViscosity (FluidGridBase.h:48-52); phase 0/"liquid" or 1/"gas"."""
        ph = {"liquid": 0, "gas": 1}.get(phase, phase)
        call("fs_set_component_viscosity", self._h, int(ph), int(index), float(coefficient))
        mixture = self.liquid_mixture if ph == 0 else self.gas_mixture
        mixture.components[index].dynamic_viscosity = float(coefficient)

    def multigridTiles(self, level: int):
        """(pre, post, restriction) autotuned tile variants of a level; -1 inside the tail kernel."""
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        call("fs_mg_level_tiles", self._h, int(level), C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, c.value

    def setMultigridTiles(self, level: int, pre: int = -1, post: int = -1, restriction: int = -1):
        """Force tile variants of a tiled level (-1 keeps the autotuned one); values do not change."""
        call("fs_mg_set_level_tiles", self._h, int(level), int(pre), int(post), int(restriction))

    def applySurfaceTension(self, dt: float):
        """FluidGridCPU::applySurfaceTension (FluidGridCPU.cpp:862-926)."""
        call("fs_apply_surface_tension", self._h, float(dt))

    def capillaryGeometry(self):
        """(heights N x 4, curvature N x 2) of the current fractions (FluidSurfaceTensionCPU)."""
        n = self.width * self.height
        hts = np.zeros(4 * n)
        cur = np.zeros(2 * n)
        call("fs_capillary_geometry", self._h, _dp(hts), _dp(cur))
        return hts.reshape(n, 4), cur.reshape(n, 2)

    def applyViscosity(self, dt: float) -> int:
        """FluidGridGPU::applyViscosity (FluidGridGPU.cpp:1039-1046); returns the CG iterations."""
        n = C.c_int()
        call("fs_apply_viscosity", self._h, float(dt), C.byref(n))
        return n.value

    # -- forces (FluidGridBase.h:70-84) --
    def getGravityY(self) -> float:
        g = C.c_double()
        call("fs_get_gravity", self._h, C.byref(g))
        return g.value

    def setGravityY(self, acceleration: float):
        call("fs_set_gravity", self._h, float(acceleration))

    def getSurfaceTension(self) -> float:
        return self._surface_tension

    def setSurfaceTension(self, coefficient: float):
        call("fs_set_surface_tension", self._h, float(coefficient))
        self._surface_tension = float(coefficient)

    def capillaryStepLimit(self) -> float:
        v = C.c_double()
        call("fs_capillary_step_limit", self._h, C.byref(v))
        return v.value

    # -- solver options (FluidGridCPU.h:39-44) --
    def setPressurePreconditioner(self, kind: int):
        call("fs_set_preconditioner", self._h, int(kind))

    def setTransport(self, mode: int):
        call("fs_set_transport", self._h, int(mode))

    # -- solid obstacles (extension: the reference's only walls are the domain border) --
    def setSolidMask(self, mask):
        """W*H cells, nonzero = solid; None removes the mask (fs_set_solid_mask)."""
        if mask is None:
            call("fs_set_solid_mask", self._h, None)
            return
        m = np.ascontiguousarray(mask, dtype=np.uint8).ravel()
        if m.size != self.width * self.height:
            raise ValueError("solid mask dimensions do not match the grid")
        call("fs_set_solid_mask", self._h, m.ctypes.data)

    # -- multigrid level storage (extension: FS_MG_FP32 / FS_MG_FP16 / FS_MG_BF16) --
    def setMultigridPrecision(self, mode: int):
        call("fs_set_multigrid_precision", self._h, int(mode))

    def setFlexibleCG(self, enabled: bool):
        """Polak-Ribiere (flexible) CG beta instead of the reference's Fletcher-Reeves one."""
        call("fs_set_flexible_cg", self._h, int(bool(enabled)))

    def setRelativeTolerance(self, enabled: bool):
        """Read step/project tolerances as relative residuals: each projection's PCG stops at
        max|b - A p| <= tol max|b| and is accepted at max|div u| <= tol max|b| dt / h^2
        (fs_set_relative_tolerance)."""
        call("fs_set_relative_tolerance", self._h, int(bool(enabled)))

    def getRelativeTolerance(self) -> bool:
        m = C.c_int()
        call("fs_get_relative_tolerance", self._h, C.byref(m))
        return bool(m.value)

    def setPressureSmoother(self, kind: int, sweeps: int, coarse_sweeps: int = 4):
        """Smoother of the multigrid levels above 4096 cells: FS_SMOOTHER_RED_BLACK (1..4 sweeps per side, default
        4) or FS_SMOOTHER_CHEBYSHEV (Chebyshev-weighted damped Jacobi, 1..8 sweeps; FP32 levels only); the smaller
        levels run `coarse_sweeps` (1..4) red-black sweeps."""
        call("fs_set_pressure_smoother", self._h, int(kind), int(sweeps), int(coarse_sweeps))

    def setViscositySmoother(self, chebyshev_sweeps: int):
        """fs_set_viscosity_smoother: 0 restores the reference's V(2,2) damped-Jacobi viscosity smoother; 2, 4 or 6
        smooths every viscosity multigrid level with that many Chebyshev-weighted Jacobi sweeps per side."""
        call("fs_set_viscosity_smoother", self._h, int(chebyshev_sweeps))

    def tuneViscositySmoother(self, dt: float):
        """fs_tune_viscosity_smoother: keeps the fastest converging viscosity smoother (0, 2, 4, 6) for a
        viscosity solve of the current state, which stays unchanged; returns (sweeps, ms)."""
        n, ms = C.c_int(), C.c_double()
        call("fs_tune_viscosity_smoother", self._h, float(dt), C.byref(n), C.byref(ms))
        return n.value, ms.value

    def getPressureSmoother(self):
        """(kind, sweeps, coarse_sweeps)."""
        k, n, c = C.c_int(), C.c_int(), C.c_int()
        call("fs_get_pressure_smoother", self._h, C.byref(k), C.byref(n), C.byref(c))
        return k.value, n.value, c.value

    def tunePressureSmoother(self, dt: float, max_iterations: int = 200, divergence_tolerance: float = 1e-6):
        """Times a projection of the current state with each smoother candidate and keeps the fastest that
        converges; the state is unchanged.  Returns (kind, sweeps, coarse_sweeps, ms per projection)."""
        k, n, c, ms = C.c_int(), C.c_int(), C.c_int(), C.c_double()
        call("fs_tune_pressure_smoother", self._h, float(dt), int(max_iterations), float(divergence_tolerance),
             C.byref(k), C.byref(n), C.byref(c), C.byref(ms))
        return k.value, n.value, c.value, ms.value

    def getMultigridPrecision(self) -> int:
        m = C.c_int()
        call("fs_get_multigrid_precision", self._h, C.byref(m))
        return m.value

    def solidCells(self) -> int:
        n = C.c_longlong()
        call("fs_solid_cells", self._h, C.byref(n))
        return n.value

    # -- state (FluidGridBase.h:118-119) --
    def _sizes(self):
        W, H = self.width, self.height
        return W * H, (W + 1) * H, W * (H + 1)

    def upload(self, state: FluidGridState):
        N, NU, NV = self._sizes()
        nl, ng = len(self.liquid_mixture.components), len(self.gas_mixture.components)

        def field_(a, n):
            if a is None or len(a) == 0:
                return None
            a = np.ascontiguousarray(a, dtype=np.float64).ravel()
            if a.size != n:
                raise ValueError("FluidGridHIP upload dimensions do not match the grid")
            return a

        def comps(lst, k):
            if lst is None or len(lst) == 0:
                return None
            if len(lst) != k:
                raise ValueError("FluidGridHIP upload component count does not match the mixture")
            out = np.zeros(k * N)
            for i, a in enumerate(lst):
                f = field_(a, N)
                if f is not None:
                    out[i * N:(i + 1) * N] = f
            return out

        arrays = [field_(state.liquid_fraction, N), field_(state.u, NU), field_(state.v, NV),
                  field_(state.pressure, N), comps(state.liquid_mass, nl), comps(state.gas_mass, ng)]
        self._keep = arrays
        call("fs_upload", self._h, *[_dp(a) for a in arrays])

    def download(self) -> FluidGridState:
        N, NU, NV = self._sizes()
        nl, ng = len(self.liquid_mixture.components), len(self.gas_mixture.components)
        f, u, v, p = np.zeros(N), np.zeros(NU), np.zeros(NV), np.zeros(N)
        lm, gm = np.zeros(nl * N), np.zeros(ng * N)
        call("fs_download", self._h, _dp(f), _dp(u), _dp(v), _dp(p), _dp(lm) if nl else None,
             _dp(gm) if ng else None)
        return FluidGridState(f, u, v, p, [lm[i * N:(i + 1) * N] for i in range(nl)],
                              [gm[i * N:(i + 1) * N] for i in range(ng)])

    def invalidateCaches(self):
        call("fs_invalidate_caches", self._h)

    # -- simulation (FluidGridBase.h:133-135, FluidGridCPU.h:117-126) --
    def step(self, dt: float, max_iterations: int = 200, divergence_tolerance: float = 1e-6,
             timings: Optional[StepTimings] = None) -> ProjectionResult:
        res = FsProjectionResult()
        t = FsStepTimings()
        call("fs_step", self._h, float(dt), int(max_iterations), float(divergence_tolerance),
             C.byref(t) if timings is not None else None, C.byref(res))
        if timings is not None:
            for name, _ in FsStepTimings._fields_:
                setattr(timings, name, getattr(t, name))
        return ProjectionResult.from_c(res)

    def project(self, dt: float, max_iterations: int = 200, divergence_tolerance: float = 1e-6) -> ProjectionResult:
        res = FsProjectionResult()
        call("fs_project", self._h, float(dt), int(max_iterations), float(divergence_tolerance), C.byref(res))
        return ProjectionResult.from_c(res)

    def applyGravity(self, dt: float):
        call("fs_apply_gravity", self._h, float(dt))

    def addVortex(self, center, radius: float, strength: float):
        call("fs_add_vortex", self._h, float(center[0]), float(center[1]), float(radius), float(strength))

    def advect(self, dt: float) -> int:
        """FluidGridCPU::advect (FluidGridCPU.cpp:362-463): PLIC/VOF transport alone; returns the CFL
        substep count."""
        n = C.c_int()
        call("fs_advect", self._h, float(dt), C.byref(n))
        return n.value

    def advectSemiLagrangian(self, dt: float):
        call("fs_advect_semi_lagrangian", self._h, float(dt))

    def maxDivergence(self) -> float:
        v = C.c_double()
        call("fs_max_divergence", self._h, C.byref(v))
        return v.value

    # -- checkpoints (FluidGridBase.h:123-125, FluidGridGPU.cpp:566-880) --
    def saveCheckpoint(self, checkpoint: FluidCheckpoint) -> bool:
        """Fills `checkpoint` with the library's metadata and FP64 payload; the adapter adds the substance
        catalogue and each component's substance, which the library does not hold."""
        mb, pb = C.c_size_t(), C.c_size_t()
        call("fs_checkpoint_size", self._h, C.byref(mb), C.byref(pb))
        text = C.create_string_buffer(mb.value)
        payload = C.create_string_buffer(max(pb.value, 1))
        call("fs_checkpoint_save", self._h, text, mb.value, payload, pb.value)
        meta = json.loads(text.value.decode())
        meta["substances"] = [s if isinstance(s, (str, int, float)) else str(s) for s in self.substances]
        for key, mixture in (("liquid_mixture", self.liquid_mixture), ("gas_mixture", self.gas_mixture)):
            for entry, component in zip(meta[key], mixture.components):
                entry["substance"] = component.substance
        checkpoint.metadata = meta
        checkpoint.data = payload.raw[:pb.value]
        return True

    def loadCheckpoint(self, checkpoint: FluidCheckpoint) -> bool:
        """Validates and restores a checkpoint; raises FluidSimError (RuntimeError) without changing the
        grid if it is incompatible or corrupt, as FluidGridGPU::loadCheckpoint throws."""
        meta = checkpoint.metadata
        if not isinstance(meta, dict):
            raise _abi.FluidSimError(_abi.FS_E_RUNTIME, "Checkpoint metadata must be a JSON object")
        substances = meta.get("substances", [])
        if not isinstance(substances, list):
            raise _abi.FluidSimError(_abi.FS_E_RUNTIME, "Checkpoint substances must be a list")
        names = {}
        for key in ("liquid_mixture", "gas_mixture"):
            entries = meta.get(key)
            names[key] = []
            for e in entries if isinstance(entries, list) else []:
                sub = e.get("substance", "") if isinstance(e, dict) else ""
                if isinstance(e, dict) and "substance" in e and substances and sub not in substances:
                    raise _abi.FluidSimError(_abi.FS_E_RUNTIME, "Checkpoint mixture references a missing substance")
                names[key].append(sub)
        text = json.dumps(meta, allow_nan=False).encode() if _finite_json(meta) else b"null"
        data = bytes(checkpoint.data)
        buf = C.create_string_buffer(data, max(len(data), 1))
        call("fs_checkpoint_load", self._h, text, len(text), buf, len(data))
        nl, ng = C.c_int(), C.c_int()
        call("fs_component_counts", self._h, C.byref(nl), C.byref(ng))

        def mixture(phase, count, labels):
            out = FluidMixture()
            for k in range(count):
                d, mu = C.c_double(), C.c_double()
                call("fs_get_component", self._h, phase, k, C.byref(d), C.byref(mu))
                out.components.append(MixtureComponent(labels[k] if k < len(labels) else "", d.value, mu.value))
            return out

        self.liquid_mixture = mixture(0, nl.value, names["liquid_mixture"])
        self.gas_mixture = mixture(1, ng.value, names["gas_mixture"])
        self.substances = list(substances)
        sigma = C.c_double()
        call("fs_get_surface_tension", self._h, C.byref(sigma))
        self._surface_tension = sigma.value
        return True

    # -- scene frame (FluidScene.inl:328-410) --
    def advanceFrame(self, frame_time: float = 0.02, steps_per_second: int = 60,
                     projection_tolerance: float = 0.1, profile_stages: bool = False) -> FrameStatistics:
        """One FluidScene frame: equal steps at the scene rate (one step when the capillary subdivision
        already reaches it), each solved to tolerance; stops at the first step not accepted."""
        settings = FsFrameSettings(float(frame_time), int(steps_per_second), float(projection_tolerance),
                                   int(bool(profile_stages)))
        st = FsFrameStatistics()
        status = lib().fs_advance_frame(self._h, C.byref(settings), C.byref(st))
        out = FrameStatistics(st.fluid_ms, StepTimings(*[getattr(st.stages, n) for n, _ in FsStepTimings._fields_]),
                              st.substeps, st.transport_steps, st.pressure_iterations, st.viscosity_iterations,
                              st.projection_tolerance, st.max_volume_error, st.rms_volume_error,
                              st.steps_scheduled, st.steps_run, bool(st.accepted), st.max_divergence)
        _abi.check(status, "fs_advance_frame")
        return out

    # -- device-resident snapshots --
    def snapshotSave(self):
        call("fs_snapshot_save", self._h)

    def snapshotRestore(self):
        call("fs_snapshot_restore", self._h)

    # -- hierarchy introspection --
    def multigridLevels(self) -> int:
        n = C.c_int()
        call("fs_mg_levels", self._h, C.byref(n))
        return n.value

    def multigridShape(self, level: int):
        w, h = C.c_int(), C.c_int()
        call("fs_mg_level_shape", self._h, int(level), C.byref(w), C.byref(h))
        return w.value, h.value

    def multigridCenters(self, level: int):
        w, h = self.multigridShape(level)
        cx, cy = np.zeros(w), np.zeros(h)
        call("fs_mg_level_centers", self._h, int(level), _dp(cx), _dp(cy))
        return cx, cy

    def multigridWeights(self, level: int):
        w, h = self.multigridShape(level)
        r, d, i = np.zeros(w * h), np.zeros(w * h), np.zeros(w * h, dtype=np.float32)
        call("fs_mg_level_weights", self._h, int(level), _dp(r), _dp(d), i.ctypes.data_as(C.POINTER(C.c_float)))
        return r.reshape(h, w), d.reshape(h, w), i.reshape(h, w)

    def multigridProlong(self, level: int, axis: int):
        w, h = self.multigridShape(level)
        n = w if axis == 0 else h
        f, s, wt = np.zeros(n, np.int32), np.zeros(n, np.int32), np.zeros(n)
        call("fs_mg_level_prolong", self._h, int(level), int(axis), f.ctypes.data_as(C.POINTER(C.c_int)),
             s.ctypes.data_as(C.POINTER(C.c_int)), _dp(wt))
        return f, s, wt

    def multigridRestrict(self, level: int, axis: int):
        cw, ch = self.multigridShape(level + 1)
        n = cw if axis == 0 else ch
        cnt, idx, wt = np.zeros(n, np.int32), np.zeros(4 * n, np.int32), np.zeros(4 * n)
        call("fs_mg_level_restrict", self._h, int(level), int(axis), cnt.ctypes.data_as(C.POINTER(C.c_int)),
             idx.ctypes.data_as(C.POINTER(C.c_int)), _dp(wt))
        return cnt, idx.reshape(n, 4), wt.reshape(n, 4)

    def multigridCoarseFactor(self):
        n = C.c_int()
        call("fs_mg_coarse_factor", self._h, None, C.byref(n))
        f = np.zeros(max(n.value, 0) ** 2)
        call("fs_mg_coarse_factor", self._h, _dp(f), C.byref(n))
        return f.reshape(n.value, n.value)

    def multigridApply(self, residual, residual_mean: float = 0.0):
        r = np.ascontiguousarray(residual, dtype=np.float64).ravel()
        out = np.zeros(self.width * self.height)
        call("fs_mg_apply", self._h, _dp(r), float(residual_mean), _dp(out))
        return out

    def faceWeights(self):
        N, NU, NV = self._sizes()
        iu, iv = np.zeros(NU), np.zeros(NV)
        call("fs_face_weights", self._h, _dp(iu), _dp(iv))
        return iu, iv

    def timeKernel(self, kernel: int, reps: int = 100) -> float:
        ms = C.c_double()
        call("fs_time_kernel", self._h, int(kernel), int(reps), C.byref(ms))
        return ms.value

    def deviceBytes(self) -> int:
        b = C.c_size_t()
        call("fs_grid_device_bytes", self._h, C.byref(b))
        return b.value
