"""ctypes wrapper of the FP64 CPU oracle (oracle/fs_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and bench.py's CPU-baseline leg,
always as the checker or the timed CPU baseline, never as the product path.

The wrapped C code restates, single-threaded and in FP64, the reference's
FluidSim/UniformCPU/FluidGridCPU.cpp step (transport, forces, projection),
FluidSim/UniformCPU/PressureMultigrid.cpp, and the WebGPU-only implicit viscosity solve
(FluidSim/UniformWebGPU/ViscositySolverGPU.cpp, ViscosityKernels.h).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class OrcResult(C.Structure):
    _fields_ = [
        ("iterations", C.c_int),
        ("max_divergence", C.c_double),
        ("converged", C.c_int),
        ("physical_steps", C.c_int),
        ("transport_steps", C.c_int),
        ("restarts", C.c_int),
        ("initial_residual", C.c_double),
        ("final_residual", C.c_double),
        ("viscosity_iterations", C.c_int),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


def build():
    """Compile liboracle.so with the committed Makefile (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    L = C.CDLL(_LIB_PATH)
    vp = C.c_void_p
    sig = {
        "orc_mg_create": (vp, [C.c_int, C.c_int]),
        "orc_mg_destroy": (None, [vp]),
        "orc_mg_rebuild": (C.c_int, [vp, _dp, _dp]),
        "orc_mg_apply": (None, [vp, _dp, _dp, C.c_double]),
        "orc_mg_levels": (C.c_int, [vp]),
        "orc_mg_level_shape": (None, [vp, C.c_int, _ip, _ip]),
        "orc_mg_level_centers": (None, [vp, C.c_int, _dp, _dp]),
        "orc_mg_level_weights": (None, [vp, C.c_int, _dp, _dp, _dp]),
        "orc_mg_level_prolong": (None, [vp, C.c_int, C.c_int, _ip, _ip, _dp]),
        "orc_mg_level_restrict": (None, [vp, C.c_int, C.c_int, _ip, _ip, _dp]),
        "orc_mg_coarse_factor": (C.c_int, [vp, _dp]),
        "orc_mg_set_pin": (None, [vp, C.c_int]),
        "orc_mg_coarse_pinned": (C.c_int, [vp, C.c_void_p]),
        "orc_grid_multigrid": (vp, [vp]),
        "orc_grid_create": (vp, [C.c_int, C.c_int, C.c_double]),
        "orc_grid_destroy": (None, [vp]),
        "orc_set_origin": (None, [vp, C.c_double, C.c_double]),
        "orc_set_gravity": (None, [vp, C.c_double]),
        "orc_set_relative_tolerance": (None, [vp, C.c_int]),
        "orc_set_threads": (None, [C.c_int]),
        "orc_set_smoother": (C.c_int, [C.c_int, C.c_int, C.c_int]),
        "orc_chebyshev_weight": (C.c_double, [C.c_int, C.c_int]),
        "orc_get_threads": (C.c_int, []),
        "orc_set_preconditioner": (None, [vp, C.c_int]),
        "orc_configure": (C.c_int, [vp, C.c_int, _dp, C.c_int, _dp]),
        "orc_upload": (C.c_int, [vp, _dp, _dp, _dp, _dp, _dp, _dp]),
        "orc_download": (None, [vp, _dp, _dp, _dp, _dp, _dp, _dp]),
        "orc_invalidate_caches": (None, [vp]),
        "orc_set_solid": (None, [vp, C.c_void_p]),
        "orc_apply_gravity": (None, [vp, C.c_double]),
        "orc_add_vortex": (None, [vp, C.c_double, C.c_double, C.c_double, C.c_double]),
        "orc_max_divergence": (C.c_double, [vp]),
        "orc_project": (OrcResult, [vp, C.c_double, C.c_int, C.c_double]),
        "orc_force_project": (OrcResult, [vp, C.c_double, C.c_int, C.c_double]),
        "orc_advect_semi_lagrangian": (None, [vp, C.c_double]),
        "orc_step_semi_lagrangian": (OrcResult, [vp, C.c_double, C.c_int, C.c_double]),
        "orc_advect_reference": (C.c_int, [vp, C.c_double]),
        "orc_transport_max_step": (C.c_double, [vp, C.c_double]),
        "orc_step": (OrcResult, [vp, C.c_double, C.c_int, C.c_double]),
        "orc_sample_velocity": (None, [vp, C.c_double, C.c_double, _dp, _dp]),
        "orc_viscosity_solve": (C.c_int, [C.c_int, C.c_int, C.c_double, _dp, _dp, _dp, C.c_double, C.c_double,
                                          C.c_int, _ip, _dp]),
        "orc_viscosity_solve_smoothed": (C.c_int, [C.c_int, C.c_int, C.c_double, _dp, _dp, _dp, C.c_double,
                                                   C.c_double, C.c_int, C.c_int, _ip, _dp]),
        "orc_viscosity_chebyshev_weight": (C.c_double, [C.c_int, C.c_int]),
        "orc_set_viscosity_smoother": (C.c_int, [vp, C.c_int]),
        "orc_viscosity_levels": (C.c_int, [C.c_int, C.c_int]),
        "orc_viscosity_material": (None, [C.c_int, C.c_double, C.c_int, _dp, _dp, _dp, C.c_int, _dp, _dp, _dp,
                                          _dp]),
        "orc_configure_viscosity": (None, [vp, _dp, _dp]),
        "orc_apply_viscosity": (C.c_int, [vp, C.c_double, _ip]),
        "orc_has_viscosity": (C.c_int, [vp]),
        "orc_set_surface_tension": (C.c_int, [vp, C.c_double]),
        "orc_capillary_step_limit": (C.c_double, [vp]),
        "orc_apply_surface_tension": (None, [vp, C.c_double]),
        "orc_capillary_geometry": (None, [C.c_int, C.c_int, C.c_double, _dp, _dp, _dp, _dp]),
        "orc_fit_curvature": (None, [C.c_int, _dp, _dp, _dp, C.c_double, C.c_double, C.c_double, _dp]),
        "orc_interface_segment": (None, [C.c_double, C.c_double, C.c_double, _dp]),
        "orc_line_offset": (C.c_double, [C.c_double, C.c_double, C.c_double]),
        "orc_color_pixels": (None, [vp, C.POINTER(C.c_float), C.c_int, C.POINTER(C.c_float)]),
        "orc_commit_edits": (None, [vp, C.c_int, _ip, _ip, _ip, _dp]),
        "orc_density": (None, [vp, _dp]),
        "orc_face_weights": (None, [vp, _dp, _dp]),
        "orc_apply_operator": (C.c_double, [vp, _dp, _dp]),
        "orc_pressure_rhs": (None, [vp, C.c_double, _dp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _ptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(_dp)


def _f64(a):
    if a is None:
        return None
    return np.ascontiguousarray(a, dtype=np.float64)


class Multigrid:
    """PressureMultigrid (PressureMultigrid.h:9-69) restated in FP64."""

    def __init__(self, width, height):
        self.w, self.h = int(width), int(height)
        self._h = lib().orc_mg_create(self.w, self.h)

    @classmethod
    def borrowed(cls, handle, width, height, owner):
        """A view of a hierarchy owned by `owner` (an OracleGrid); never destroyed here."""
        m = cls.__new__(cls)
        m.w, m.h, m._h, m._owner = int(width), int(height), handle, owner
        return m

    def __del__(self):
        if getattr(self, "_h", None) and not hasattr(self, "_owner"):
            lib().orc_mg_destroy(self._h)
            self._h = None

    def set_pin(self, pin=True):
        """Solid-mask extension: pin vanishing coarse pivots instead of failing."""
        lib().orc_mg_set_pin(self._h, int(bool(pin)))

    def pinned(self):
        n = lib().orc_mg_coarse_factor(self._h, None)
        flags = np.zeros(max(n, 1), np.uint8)
        lib().orc_mg_coarse_pinned(self._h, flags.ctypes.data)
        return flags[:max(n, 0)].astype(bool)

    def rebuild(self, inv_u, inv_v):
        inv_u, inv_v = _f64(inv_u), _f64(inv_v)
        assert inv_u.size == (self.w + 1) * self.h and inv_v.size == self.w * (self.h + 1)
        return bool(lib().orc_mg_rebuild(self._h, _ptr(inv_u), _ptr(inv_v)))

    def apply(self, residual, residual_mean=0.0):
        r = _f64(residual)
        assert r.size == self.w * self.h
        out = np.zeros(self.w * self.h)
        lib().orc_mg_apply(self._h, _ptr(r), _ptr(out), float(residual_mean))
        return out

    @property
    def levels(self):
        return lib().orc_mg_levels(self._h)

    def shape(self, level):
        w, h = C.c_int(), C.c_int()
        lib().orc_mg_level_shape(self._h, level, C.byref(w), C.byref(h))
        return w.value, h.value

    def centers(self, level):
        w, h = self.shape(level)
        cx, cy = np.zeros(w), np.zeros(h)
        lib().orc_mg_level_centers(self._h, level, _ptr(cx), _ptr(cy))
        return cx, cy

    def weights(self, level):
        w, h = self.shape(level)
        r, d, i = np.zeros(w * h), np.zeros(w * h), np.zeros(w * h)
        lib().orc_mg_level_weights(self._h, level, _ptr(r), _ptr(d), _ptr(i))
        return r.reshape(h, w), d.reshape(h, w), i.reshape(h, w)

    def prolong(self, level, axis):
        w, h = self.shape(level)
        n = w if axis == 0 else h
        f, s, wt = np.zeros(n, np.int32), np.zeros(n, np.int32), np.zeros(n)
        lib().orc_mg_level_prolong(self._h, level, axis, f.ctypes.data_as(_ip), s.ctypes.data_as(_ip), _ptr(wt))
        return f, s, wt

    def restriction(self, level, axis):
        cw, ch = self.shape(level + 1)
        n = cw if axis == 0 else ch
        cnt, idx, wt = np.zeros(n, np.int32), np.zeros(4 * n, np.int32), np.zeros(4 * n)
        lib().orc_mg_level_restrict(self._h, level, axis, cnt.ctypes.data_as(_ip), idx.ctypes.data_as(_ip), _ptr(wt))
        return cnt, idx.reshape(n, 4), wt.reshape(n, 4)

    def coarse_factor(self):
        n = lib().orc_mg_coarse_factor(self._h, None)
        f = np.zeros(max(n, 0) * max(n, 0))
        lib().orc_mg_coarse_factor(self._h, _ptr(f))
        return f.reshape(max(n, 0), max(n, 0))


class OracleGrid:
    """FluidGridCPU subset (FluidGridCPU.h:15-225) restated in FP64, host arrays as numpy."""

    def __init__(self, width, height, cell_size):
        self.width, self.height, self.cell_size = int(width), int(height), float(cell_size)
        self._h = lib().orc_grid_create(self.width, self.height, self.cell_size)
        self.n_liquid = 0
        self.n_gas = 0

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_grid_destroy(self._h)
            self._h = None

    @property
    def N(self):
        return self.width * self.height

    def set_origin(self, x, y):
        lib().orc_set_origin(self._h, x, y)

    def set_gravity(self, gy):
        lib().orc_set_gravity(self._h, gy)

    def set_relative_tolerance(self, enabled=True):
        lib().orc_set_relative_tolerance(self._h, int(bool(enabled)))

    def set_preconditioner(self, kind):
        lib().orc_set_preconditioner(self._h, {"diagonal": 0, "multigrid": 1}[kind] if isinstance(kind, str) else int(kind))

    def configure(self, liquid_density, gas_density):
        ld, gd = _f64(np.atleast_1d(liquid_density)), _f64(np.atleast_1d(gas_density))
        self.n_liquid, self.n_gas = ld.size, gd.size
        lib().orc_configure(self._h, ld.size, _ptr(ld), gd.size, _ptr(gd))

    def upload(self, fraction=None, u=None, v=None, pressure=None, liquid_mass=None, gas_mass=None):
        args = [_f64(a) for a in (fraction, u, v, pressure, liquid_mass, gas_mass)]
        sizes = [self.N, (self.width + 1) * self.height, self.width * (self.height + 1), self.N,
                 self.n_liquid * self.N, self.n_gas * self.N]
        for a, s in zip(args, sizes):
            if a is not None and a.size != s:
                raise ValueError("upload dimensions do not match the grid")
        self._keep = args
        lib().orc_upload(self._h, *[_ptr(a) for a in args])

    def download(self):
        out = dict(fraction=np.zeros(self.N), u=np.zeros((self.width + 1) * self.height),
                   v=np.zeros(self.width * (self.height + 1)), pressure=np.zeros(self.N),
                   liquid_mass=np.zeros(self.n_liquid * self.N), gas_mass=np.zeros(self.n_gas * self.N))
        lib().orc_download(self._h, *[_ptr(out[k]) for k in
                                      ("fraction", "u", "v", "pressure", "liquid_mass", "gas_mass")])
        return out

    def invalidate_caches(self):
        lib().orc_invalidate_caches(self._h)

    def multigrid(self):
        """The grid's own hierarchy, rebuilt from the current coefficients; None if it fails."""
        h = lib().orc_grid_multigrid(self._h)
        return Multigrid.borrowed(h, self.width, self.height, self) if h else None

    def set_solid(self, mask):
        """Solid-mask extension: W*H cells, nonzero = solid; None removes the mask."""
        if mask is None:
            lib().orc_set_solid(self._h, None)
            return
        m = np.ascontiguousarray(mask, dtype=np.uint8).ravel()
        if m.size != self.N:
            raise ValueError("solid mask dimensions do not match the grid")
        lib().orc_set_solid(self._h, m.ctypes.data)

    def apply_gravity(self, dt):
        lib().orc_apply_gravity(self._h, dt)

    def add_vortex(self, cx, cy, radius, strength):
        lib().orc_add_vortex(self._h, cx, cy, radius, strength)

    def max_divergence(self):
        return lib().orc_max_divergence(self._h)

    def project(self, dt, max_iterations=200, divergence_tolerance=1e-6):
        return lib().orc_project(self._h, dt, max_iterations, divergence_tolerance).as_dict()

    def force_project(self, dt, max_iterations=200, divergence_tolerance=1e-6):
        return lib().orc_force_project(self._h, dt, max_iterations, divergence_tolerance).as_dict()

    def advect_semi_lagrangian(self, dt):
        lib().orc_advect_semi_lagrangian(self._h, dt)

    def step_semi_lagrangian(self, dt, max_iterations=200, divergence_tolerance=1e-6):
        return lib().orc_step_semi_lagrangian(self._h, dt, max_iterations, divergence_tolerance).as_dict()

    def advect(self, dt):
        """advect(dt), FluidGridCPU.cpp:362-463; returns the number of CFL substeps."""
        return lib().orc_advect_reference(self._h, dt)

    def transport_max_step(self, dt):
        return lib().orc_transport_max_step(self._h, dt)

    def step(self, dt, max_iterations=200, divergence_tolerance=1e-6):
        """step(), FluidGridCPU.cpp:300-360: capillary substeps, transport, the WebGPU viscosity stage,
        gravity, surface tension, projection."""
        return lib().orc_step(self._h, dt, max_iterations, divergence_tolerance).as_dict()

    def configure_viscosity(self, liquid_viscosity, gas_viscosity):
        self._mu = (_f64(liquid_viscosity), _f64(gas_viscosity))
        lib().orc_configure_viscosity(self._h, _ptr(self._mu[0]), _ptr(self._mu[1]))

    def set_viscosity_smoother(self, chebyshev_sweeps):
        """orc_set_viscosity_smoother: 0 (the reference's V(2,2) damped Jacobi) or 2, 4, 6 Chebyshev sweeps."""
        if lib().orc_set_viscosity_smoother(self._h, chebyshev_sweeps) != 0:
            raise ValueError("viscosity smoother sweeps must be 0, 2, 4 or 6")

    def apply_viscosity(self, dt):
        """FluidGridGPU::applyViscosity(dt); returns (status, iterations)."""
        it = C.c_int()
        st = lib().orc_apply_viscosity(self._h, dt, C.byref(it))
        return st, it.value

    def set_surface_tension(self, coefficient):
        if lib().orc_set_surface_tension(self._h, coefficient) != 0:
            raise ValueError("Surface tension must be finite and nonnegative")

    def capillary_step_limit(self):
        return lib().orc_capillary_step_limit(self._h)

    def color_pixels(self, rgb, liquid_fraction_alpha=False):
        """writeColorPixels (FluidGridCPU.cpp:1826-1860): float32 RGBA per cell; rgb = 3 per component."""
        c = np.ascontiguousarray(rgb, dtype=np.float32).ravel()
        out = np.zeros(4 * self.N, dtype=np.float32)
        fp = C.POINTER(C.c_float)
        lib().orc_color_pixels(self._h, c.ctypes.data_as(fp), int(bool(liquid_fraction_alpha)),
                               out.ctypes.data_as(fp))
        return out

    def commit_edits(self, cells):
        """commitEdits (FluidGridCPU.cpp:253-284) on resolved cells [(x, y, component, mass)] as
        prepare_edits returns them."""
        n = len(cells)
        xs = np.array([c[0] for c in cells], dtype=np.int32)
        ys = np.array([c[1] for c in cells], dtype=np.int32)
        ks = np.array([c[2] for c in cells], dtype=np.int32)
        ms = np.array([c[3] for c in cells], dtype=np.float64)
        ip = C.POINTER(C.c_int)
        lib().orc_commit_edits(self._h, n, xs.ctypes.data_as(ip), ys.ctypes.data_as(ip), ks.ctypes.data_as(ip),
                               _ptr(ms))

    def apply_surface_tension(self, dt):
        lib().orc_apply_surface_tension(self._h, dt)

    def sample_velocity(self, x, y):
        vx, vy = C.c_double(), C.c_double()
        lib().orc_sample_velocity(self._h, x, y, C.byref(vx), C.byref(vy))
        return vx.value, vy.value

    def density(self):
        d = np.zeros(self.N)
        lib().orc_density(self._h, _ptr(d))
        return d

    def face_weights(self):
        iu = np.zeros((self.width + 1) * self.height)
        iv = np.zeros(self.width * (self.height + 1))
        lib().orc_face_weights(self._h, _ptr(iu), _ptr(iv))
        return iu, iv

    def apply_operator(self, p):
        p = _f64(p)
        q = np.zeros(self.N)
        dot = lib().orc_apply_operator(self._h, _ptr(p), _ptr(q))
        return q, dot

    def pressure_rhs(self, dt):
        r = np.zeros(self.N)
        lib().orc_pressure_rhs(self._h, dt, _ptr(r))
        return r


VISC_OK, VISC_NONFINITE, VISC_BREAKDOWN, VISC_NOT_CONVERGED, VISC_BAD_SMOOTHER = 0, 1, 2, 3, 4


def set_threads(threads):
    """Worker threads of the projection's passes (1: the reference's single-thread order)."""
    lib().orc_set_threads(int(threads))


def get_threads():
    return lib().orc_get_threads()


SMOOTH_RBGS, SMOOTH_CHEBYSHEV = 0, 1


def set_smoother(kind, sweeps, coarse_sweeps=4):
    """Smoother of the multigrid levels above 4096 cells: SMOOTH_RBGS or SMOOTH_CHEBYSHEV, `sweeps` per side;
    levels of at most 4096 cells run `coarse_sweeps` red-black sweeps; process-wide."""
    if lib().orc_set_smoother(int(kind), int(sweeps), int(coarse_sweeps)) != 0:
        raise ValueError(f"invalid smoother {kind} with {sweeps} / {coarse_sweeps} sweeps")


def chebyshev_weight(i, sweeps):
    return lib().orc_chebyshev_weight(int(i), int(sweeps))


def viscosity_solve(width, height, cell_size, material, u, v, dt, tolerance=1e-4, maximum=-1,
                    chebyshev_sweeps=0):
    """ViscositySolverGPU::solve on host arrays (FP64 restatement).  material: (rho, mu) per cell.
    chebyshev_sweeps 2, 4 or 6 replaces the V(2,2) damped-Jacobi smoother by that many Chebyshev-weighted
    sweeps per side (no reference counterpart).  Returns (status, iterations, relative_residual, u, v); u
    and v change only on success."""
    material = _f64(material)
    u = _f64(u).copy()
    v = _f64(v).copy()
    it = C.c_int()
    rel = C.c_double()
    st = lib().orc_viscosity_solve_smoothed(width, height, cell_size, _ptr(material), _ptr(u), _ptr(v), dt,
                                            tolerance, maximum, chebyshev_sweeps, C.byref(it), C.byref(rel))
    return st, it.value, rel.value, u, v


def viscosity_material(cell_size, liquid_density, liquid_viscosity, liquid_mass, gas_density, gas_viscosity,
                       gas_mass, cells):
    ld, lv, lm = _f64(liquid_density), _f64(liquid_viscosity), _f64(liquid_mass)
    gd, gv, gm = _f64(gas_density), _f64(gas_viscosity), _f64(gas_mass)
    out = np.zeros(2 * cells)
    lib().orc_viscosity_material(cells, cell_size, ld.size, _ptr(ld), _ptr(lv), _ptr(lm), gd.size, _ptr(gd),
                                 _ptr(gv), _ptr(gm), _ptr(out))
    return out


def capillary_geometry(width, height, cell_size, fraction):
    """Heights (N x 4), raw and completed curvatures (N x 2: value, valid) of a fraction field."""
    f = _f64(fraction)
    n = width * height
    heights, raw, curv = np.zeros(4 * n), np.zeros(2 * n), np.zeros(2 * n)
    lib().orc_capillary_geometry(width, height, cell_size, _ptr(f), _ptr(heights), _ptr(raw), _ptr(curv))
    return heights.reshape(n, 4), raw.reshape(n, 2), curv.reshape(n, 2)


def fit_curvature(px, py, normal, cell_size=1.0, weight=None):
    """The capillary quadratic fit of points (px, py) about `normal` (FluidSurfaceTensionCPU.h:204-254):
    (curvature, valid)."""
    px, py = _f64(px), _f64(py)
    w = None if weight is None else _f64(weight)
    out = np.zeros(2)
    lib().orc_fit_curvature(px.size, _ptr(px), _ptr(py), None if w is None else _ptr(w), float(normal[0]),
                            float(normal[1]), float(cell_size), _ptr(out))
    return float(out[0]), bool(out[1])


def interface_segment(normal, fraction):
    """Centroid (relative to the cell centre) and length of a cell's interface segment
    (FluidSurfaceTensionCPU.h:287-324)."""
    out = np.zeros(3)
    lib().orc_interface_segment(float(normal[0]), float(normal[1]), float(fraction), _ptr(out))
    return out


def line_offset(normal, fraction):
    """interface_geometry::offset (FluidInterfaceGeometry.h:51-61)."""
    return float(lib().orc_line_offset(float(normal[0]), float(normal[1]), float(fraction)))


def prepare_edits(changes, width, height, cell_volume, liquid, gas):
    """CellEditBatch::prepare (FluidGridEdit.h:37-95): changes [(x, y, phase, substance)] with phase
    "liquid" or "gas"; liquid / gas are the mixtures as [(substance, density)].  Returns the resolved cells
    [(x, y, component, mass)] in raster order, the last change of a cell kept.  Raises IndexError for a
    cell outside the grid and ValueError for a substance the phase does not hold."""
    cells = []
    for x, y, phase, substance in changes:
        if not (0 <= x < width and 0 <= y < height):
            raise IndexError("Fluid grid edit cell is outside the grid")
        mixture = liquid if phase == "liquid" else gas
        found = [k for k, (sub, _) in enumerate(mixture) if sub == substance]
        if not found:
            raise ValueError("Fluid grid edit substance is not configured in the selected phase")
        k = found[0]
        cells.append((x, y, k + (0 if phase == "liquid" else len(liquid)), mixture[k][1] * cell_volume))
    cells.sort(key=lambda c: (c[1], c[0]))   # stable: equal cells keep their order
    out = []
    for c in cells:
        if out and (out[-1][0], out[-1][1]) == (c[0], c[1]):
            out[-1] = c
        else:
            out.append(c)
    return out


# ---- scene frame (FluidSim/FluidScene.inl:328-410) ----

UNBOUNDED_BUDGET = 100000   # iterations that stand for the scene's unbounded (-1) pressure budget


def frame_schedule(frame_time, steps_per_second, capillary_step_limit):
    """(step_count, step) of one scene frame on a uniform GPU grid (FluidScene.inl:343-353): equal steps at
    the scene's minimum rate, or one step when the engine's capillary subdivision already reaches it."""
    import math

    count = int(math.ceil(frame_time * float(steps_per_second)))
    if math.ceil(frame_time / capillary_step_limit) >= count:
        count = 1
    return count, frame_time / count


def run_frame(grid, frame_time=0.02, steps_per_second=60, projection_tolerance=0.1):
    """One frame on an OracleGrid: the schedule above, each step solved to tolerance, statistics summed as
    FluidScene.inl:371-393 does, stopping at the first step that is not accepted."""
    count, step = frame_schedule(frame_time, steps_per_second, grid.capillary_step_limit())
    stats = dict(steps_scheduled=count, steps_run=0, substeps=0, transport_steps=0, pressure_iterations=0,
                 viscosity_iterations=0, accepted=True)
    for _ in range(count):
        r = grid.step(step, UNBOUNDED_BUDGET, projection_tolerance)
        stats["steps_run"] += 1
        stats["substeps"] += r["physical_steps"]
        stats["transport_steps"] += r.get("transport_steps", 0)
        stats["pressure_iterations"] += r["iterations"]
        stats["viscosity_iterations"] += r.get("viscosity_iterations", 0)
        if not r["converged"]:
            stats["accepted"] = False
            break
    return stats
