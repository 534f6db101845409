"""ctypes binding of libfluidsim_hip.so (declared in include/fluidsim_hip.h).

The product path: every call goes to the HIP library.  There is no CPU fallback; if the library is
missing or cannot be loaded, `lib()` raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfluidsim_hip.so")

FS_OK = 0
FS_E_INVALID_ARGUMENT = 1
FS_E_LOGIC = 2
FS_E_UNSUPPORTED = 3
FS_E_DEVICE = 4
FS_E_RUNTIME = 5

FS_TRANSPORT_REFERENCE = 0
FS_TRANSPORT_NONE = 1
FS_TRANSPORT_SEMI_LAGRANGIAN = 2

FS_PRECOND_DIAGONAL = 0
FS_PRECOND_MULTIGRID = 1

FS_MG_FP32 = 0
FS_MG_FP16 = 1
FS_MG_BF16 = 2
FS_MG_BF16_VECTORS = 3
FS_SMOOTHER_RED_BLACK = 0
FS_SMOOTHER_CHEBYSHEV = 1


class FsProjectionResult(C.Structure):
    _fields_ = [
        ("iterations", C.c_int),
        ("max_divergence", C.c_double),
        ("converged", C.c_int),
        ("physical_steps", C.c_int),
        ("transport_steps", C.c_int),
        ("relaxed", C.c_int),
        ("max_volume_error", C.c_double),
        ("rms_volume_error", C.c_double),
        ("viscosity_iterations", C.c_int),
        ("precision_limited", C.c_int),
        ("restarts", C.c_int),
        ("initial_residual", C.c_double),
        ("final_residual", C.c_double),
        ("rhs_max", C.c_double),
    ]


class FsCellEdit(C.Structure):
    _fields_ = [("x", C.c_int), ("y", C.c_int), ("component", C.c_int), ("mass", C.c_double)]


class FsStepTimings(C.Structure):
    _fields_ = [
        ("transport_ms", C.c_double),
        ("viscosity_ms", C.c_double),
        ("forces_ms", C.c_double),
        ("pressure_ms", C.c_double),
        ("selection_ms", C.c_double),
        ("regrid_ms", C.c_double),
    ]


class FsFrameSettings(C.Structure):
    _fields_ = [
        ("frame_time", C.c_double),
        ("steps_per_second", C.c_int),
        ("projection_tolerance", C.c_double),
        ("collect_timings", C.c_int),
    ]


class FsFrameStatistics(C.Structure):
    _fields_ = [
        ("fluid_ms", C.c_double),
        ("stages", FsStepTimings),
        ("substeps", C.c_int),
        ("transport_steps", C.c_int),
        ("pressure_iterations", C.c_int),
        ("viscosity_iterations", C.c_int),
        ("projection_tolerance", C.c_double),
        ("max_volume_error", C.c_double),
        ("rms_volume_error", C.c_double),
        ("steps_scheduled", C.c_int),
        ("steps_run", C.c_int),
        ("accepted", C.c_int),
        ("max_divergence", C.c_double),
    ]


class FsServerConfig(C.Structure):
    _fields_ = [("device", C.c_int), ("lanes", C.c_int), ("hbm_budget_bytes", C.c_size_t),
                ("batch_cells", C.c_longlong), ("evict_after_steps", C.c_int)]


class FsRequest(C.Structure):
    _fields_ = [
        ("width", C.c_int), ("height", C.c_int), ("cell_size", C.c_double), ("steps", C.c_int), ("dt", C.c_double),
        ("max_iterations", C.c_int), ("divergence_tolerance", C.c_double), ("gravity_y", C.c_double),
        ("surface_tension", C.c_double), ("transport", C.c_int), ("preconditioner", C.c_int),
        ("n_liquid", C.c_int), ("n_gas", C.c_int),
        ("liquid_density", C.POINTER(C.c_double)), ("liquid_viscosity", C.POINTER(C.c_double)),
        ("gas_density", C.POINTER(C.c_double)), ("gas_viscosity", C.POINTER(C.c_double)),
        ("liquid_fraction", C.POINTER(C.c_double)), ("u", C.POINTER(C.c_double)), ("v", C.POINTER(C.c_double)),
        ("liquid_mass", C.POINTER(C.c_double)), ("gas_mass", C.POINTER(C.c_double)),
        ("multigrid_precision", C.c_int), ("solid", C.c_void_p), ("relative_tolerance", C.c_int),
        ("pressure_smoother", C.c_int), ("smoother_sweeps", C.c_int), ("coarse_sweeps", C.c_int),
        ("viscosity_sweeps", C.c_int),
    ]


FS_REQUEST_QUEUED, FS_REQUEST_RUNNING, FS_REQUEST_DONE, FS_REQUEST_FAILED = 0, 1, 2, 3

FS_FEATURE_TRANSPORT, FS_FEATURE_VISCOSITY, FS_FEATURE_SURFACE_TENSION = 1, 2, 4
FS_FEATURE_SOLID_MASK, FS_FEATURE_MG16 = 8, 16


class FsRequestReport(C.Structure):
    _fields_ = [
        ("state", C.c_int), ("lane", C.c_int), ("steps_done", C.c_int), ("pressure_iterations", C.c_int),
        ("viscosity_iterations", C.c_int), ("physical_steps", C.c_int), ("max_divergence", C.c_double),
        ("queue_ms", C.c_double), ("run_ms", C.c_double), ("step_p50_ms", C.c_double), ("step_p99_ms", C.c_double),
        ("step_max_ms", C.c_double), ("estimated_bytes", C.c_size_t), ("device_bytes", C.c_size_t),
        ("setup_ms", C.c_double), ("teardown_ms", C.c_double), ("evictions", C.c_int), ("error", C.c_char * 256),
    ]


class FsServerStats(C.Structure):
    _fields_ = [
        ("submitted", C.c_longlong), ("admitted", C.c_longlong), ("completed", C.c_longlong),
        ("failed", C.c_longlong), ("steps", C.c_longlong), ("queued", C.c_int), ("resident", C.c_int),
        ("budget_bytes", C.c_size_t), ("resident_bytes", C.c_size_t), ("peak_resident_bytes", C.c_size_t),
        ("elapsed_s", C.c_double), ("steps_per_s", C.c_double), ("step_p50_ms", C.c_double),
        ("step_p99_ms", C.c_double), ("setup_ms_total", C.c_double), ("teardown_ms_total", C.c_double),
        ("evictions", C.c_longlong),
    ]


_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p

# name -> argtypes (every function returns fs_status except the two string getters)
SIGNATURES = {
    "fs_context_create": [C.c_int, C.POINTER(_vp)],
    "fs_context_create_ex": [C.c_int, C.c_int, C.POINTER(_vp)],
    "fs_context_destroy": [_vp],
    "fs_context_synchronize": [_vp],
    "fs_grid_create": [_vp, C.c_int, C.c_int, C.c_double, C.POINTER(_vp)],
    "fs_grid_destroy": [_vp],
    "fs_grid_shape": [_vp, _ip, _ip, _dp],
    "fs_set_origin": [_vp, C.c_double, C.c_double],
    "fs_get_origin": [_vp, _dp, _dp],
    "fs_configure_materials": [_vp, C.c_int, _dp, _dp, C.c_int, _dp, _dp],
    "fs_set_gravity": [_vp, C.c_double],
    "fs_get_gravity": [_vp, _dp],
    "fs_set_surface_tension": [_vp, C.c_double],
    "fs_capillary_step_limit": [_vp, _dp],
    "fs_set_preconditioner": [_vp, C.c_int],
    "fs_set_transport": [_vp, C.c_int],
    "fs_set_solid_mask": [_vp, C.c_void_p],
    "fs_set_multigrid_precision": [_vp, C.c_int],
    "fs_get_multigrid_precision": [_vp, C.POINTER(C.c_int)],
    "fs_set_flexible_cg": [_vp, C.c_int],
    "fs_set_relative_tolerance": [_vp, C.c_int],
    "fs_set_pressure_smoother": [_vp, C.c_int, C.c_int, C.c_int],
    "fs_get_pressure_smoother": [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)],
    "fs_tune_pressure_smoother": [_vp, C.c_double, C.c_int, C.c_double, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                  C.POINTER(C.c_int), C.POINTER(C.c_double)],
    "fs_get_relative_tolerance": [_vp, _ip],
    "fs_batch_apply_viscosity": [C.c_int, C.c_void_p, C.c_double, C.POINTER(C.c_int)],
    "fs_solid_cells": [_vp, C.POINTER(C.c_longlong)],
    "fs_upload": [_vp, _dp, _dp, _dp, _dp, _dp, _dp],
    "fs_download": [_vp, _dp, _dp, _dp, _dp, _dp, _dp],
    "fs_invalidate_caches": [_vp],
    "fs_step": [_vp, C.c_double, C.c_int, C.c_double, C.POINTER(FsStepTimings), C.POINTER(FsProjectionResult)],
    "fs_project": [_vp, C.c_double, C.c_int, C.c_double, C.POINTER(FsProjectionResult)],
    "fs_apply_gravity": [_vp, C.c_double],
    "fs_add_vortex": [_vp, C.c_double, C.c_double, C.c_double, C.c_double],
    "fs_max_divergence": [_vp, _dp],
    "fs_color": [_vp, C.POINTER(C.c_float), C.c_int, C.POINTER(C.c_float)],
    "fs_commit_edits": [_vp, C.c_int, C.POINTER(FsCellEdit), C.POINTER(C.c_double)],
    "fs_advect": [_vp, C.c_double, _ip],
    "fs_apply_viscosity": [_vp, C.c_double, _ip],
    "fs_set_viscosity_smoother": [_vp, C.c_int],
    "fs_tune_viscosity_smoother": [_vp, C.c_double, _ip, _dp],
    "fs_apply_surface_tension": [_vp, C.c_double],
    "fs_capillary_geometry": [_vp, _dp, _dp],
    "fs_set_component_viscosity": [_vp, C.c_int, C.c_int, C.c_double],
    "fs_viscosity_solve": [_vp, C.c_int, C.c_int, C.c_double, _dp, _dp, _dp, C.c_double, C.c_double, C.c_int, _ip,
                           _dp],
    "fs_viscosity_solve_noalloc": [_vp, C.c_int, C.c_int, C.c_double, _vp, _vp, _vp, C.c_double, C.c_double, C.c_int, _ip,
                           _dp],
    "fs_advect_semi_lagrangian": [_vp, C.c_double],
    "fs_snapshot_save": [_vp],
    "fs_snapshot_restore": [_vp],
    "fs_mg_levels": [_vp, _ip],
    "fs_mg_level_tiles": [_vp, C.c_int, _ip, _ip, _ip],
    "fs_mg_set_level_tiles": [_vp, C.c_int, C.c_int, C.c_int, C.c_int],
    "fs_mg_level_shape": [_vp, C.c_int, _ip, _ip],
    "fs_mg_level_centers": [_vp, C.c_int, _dp, _dp],
    "fs_mg_level_weights": [_vp, C.c_int, _dp, _dp, _fp],
    "fs_mg_level_prolong": [_vp, C.c_int, C.c_int, _ip, _ip, _dp],
    "fs_mg_level_restrict": [_vp, C.c_int, C.c_int, _ip, _ip, _dp],
    "fs_mg_coarse_factor": [_vp, _dp, _ip],
    "fs_mg_apply": [_vp, _dp, C.c_double, _dp],
    "fs_face_weights": [_vp, _dp, _dp],
    "fs_time_kernel": [_vp, C.c_int, C.c_int, _dp],
    "fs_grid_device_bytes": [_vp, C.POINTER(C.c_size_t)],
    "fs_get_surface_tension": [_vp, _dp],
    "fs_get_preconditioner": [_vp, _ip],
    "fs_get_transport": [_vp, _ip],
    "fs_component_counts": [_vp, _ip, _ip],
    "fs_get_component": [_vp, C.c_int, C.c_int, _dp, _dp],
    "fs_checkpoint_size": [_vp, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)],
    "fs_checkpoint_save": [_vp, C.c_char_p, C.c_size_t, _vp, C.c_size_t],
    "fs_checkpoint_load": [_vp, C.c_char_p, C.c_size_t, _vp, C.c_size_t],
    "fs_advance_frame": [_vp, C.POINTER(FsFrameSettings), C.POINTER(FsFrameStatistics)],
    "fs_batch_project": [C.c_int, C.POINTER(_vp), C.c_double, C.c_int, _dp, C.POINTER(FsProjectionResult)],
    "fs_tune_batch_smoother": [C.c_int, C.POINTER(_vp), C.c_double, C.c_int, _dp, C.POINTER(C.c_int),
                               C.POINTER(C.c_int), C.POINTER(C.c_double)],
    "fs_batch_step": [C.c_int, C.POINTER(_vp), _dp, C.POINTER(C.c_int), _dp, C.POINTER(FsProjectionResult),
                      C.POINTER(C.c_int)],
    "fs_context_set_batch_split": [_vp, C.c_longlong],
    "fs_grid_estimate_bytes": [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_size_t)],
    "fs_server_create": [C.POINTER(FsServerConfig), C.POINTER(_vp)],
    "fs_server_destroy": [_vp],
    "fs_server_submit": [_vp, C.POINTER(FsRequest), C.POINTER(C.c_longlong)],
    "fs_server_wait": [_vp, C.c_longlong, C.c_double, C.POINTER(FsRequestReport)],
    "fs_server_query": [_vp, C.c_longlong, C.POINTER(FsRequestReport)],
    "fs_server_result": [_vp, C.c_longlong, _dp, _dp, _dp, _dp, _dp, _dp],
    "fs_server_release": [_vp, C.c_longlong],
    "fs_server_drain": [_vp],
    "fs_server_get_stats": [_vp, C.POINTER(FsServerStats)],
    "fs_server_reset_stats": [_vp],
}

_lib = None


class FluidSimError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"[fs_status {code}] {message}")
        self.code = code


def lib():
    """Loads the HIP library; raises if it is not built.  No fallback exists."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m fluidsim_amd.build` "
            "(the HIP backend has no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    for name, args in SIGNATURES.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = C.c_int
    L.fs_last_error.restype = C.c_char_p
    L.fs_last_error.argtypes = []
    L.fs_version.restype = C.c_char_p
    L.fs_version.argtypes = []
    _lib = L
    return L


def check(status, what=""):
    if status != FS_OK:
        msg = lib().fs_last_error().decode(errors="replace")
        if status == FS_E_INVALID_ARGUMENT:
            raise ValueError(f"{what}: {msg}")
        if status == FS_E_LOGIC:
            raise RuntimeError(f"{what}: {msg}")
        if status == FS_E_UNSUPPORTED:
            raise NotImplementedError(f"{what}: {msg}")
        raise FluidSimError(status, f"{what}: {msg}")


def call(name, *args):
    check(getattr(lib(), name)(*args), name)
