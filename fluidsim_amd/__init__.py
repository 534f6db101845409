"""fluidsim_amd -- MI355X-native uniform-MAC projection backend (HIP, gfx950).

The compute path is libfluidsim_hip.so (C ABI in include/fluidsim_hip.h).  `FluidGridHIP` mirrors the
reference's FluidGridBase method set over that ABI.
"""
from ._abi import (FS_MG_BF16, FS_MG_BF16_VECTORS, FS_MG_FP16, FS_MG_FP32, FS_PRECOND_DIAGONAL, FS_PRECOND_MULTIGRID,
                   FS_SMOOTHER_CHEBYSHEV, FS_SMOOTHER_RED_BLACK, FS_TRANSPORT_NONE, FS_TRANSPORT_REFERENCE, FS_TRANSPORT_SEMI_LAGRANGIAN, FluidSimError, lib)
from .grid import (GAS, LIQUID, Context, FluidCheckpoint, FluidGridEdit, FluidGridHIP, FluidGridState, FluidMixture,
                   FrameStatistics, MixtureComponent, ProjectionResult, StepTimings, Substance, SubstanceCatalogue,
                   batch_apply_viscosity, batch_project, batch_step, default_context, tune_batch_smoother,
                   viscosity_solve)

__all__ = [
    "FS_MG_BF16", "FS_MG_BF16_VECTORS", "FS_MG_FP16", "FS_MG_FP32", "FS_PRECOND_DIAGONAL", "FS_PRECOND_MULTIGRID", "FS_SMOOTHER_CHEBYSHEV", "FS_SMOOTHER_RED_BLACK", "FS_TRANSPORT_NONE", "FS_TRANSPORT_REFERENCE",
    "FS_TRANSPORT_SEMI_LAGRANGIAN", "FluidSimError", "lib", "Context", "FluidGridHIP", "FluidGridState",
    "FluidMixture", "FluidCheckpoint", "FluidGridEdit", "Substance", "SubstanceCatalogue", "LIQUID", "GAS", "FrameStatistics", "MixtureComponent", "ProjectionResult", "StepTimings", "batch_apply_viscosity", "batch_project", "batch_step", "default_context", "tune_batch_smoother", "viscosity_solve",
]
