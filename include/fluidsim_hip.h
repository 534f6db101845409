/*
 * fluidsim_hip.h -- C ABI of libfluidsim_hip.so, the MI355X (gfx950) uniform-MAC backend.
 *
 * The ABI is the drop-in boundary for the reference's uniform grid backends: a C++ adapter class
 * implementing fluid::FluidGridBase (FluidSim/FluidGridBase.h:20-147) forwards each virtual method
 * to one call below, the way FluidGridWebGPU forwards to FluidGridGPU
 * (FluidSim/UniformWebGPU/FluidGridWebGPU.cpp:134-184).  Only plain C types cross it: opaque
 * handles, ints, doubles, pointers to caller-owned host arrays and POD structs.
 *
 * Conventions (mirroring the reference's error contract, FluidGridBase.h:48-116,
 * FluidGridCPU.cpp:198-251, tests/FluidGridAPITests.cpp:186-229):
 *   - every call returns an fs_status; FS_OK is 0.  FS_E_INVALID_ARGUMENT maps to
 *     std::invalid_argument, FS_E_LOGIC to std::logic_error, FS_E_UNSUPPORTED to the base-class
 *     "requires uniform WebGPU" throws, FS_E_DEVICE / FS_E_RUNTIME to std::runtime_error;
 *   - numerical non-convergence is NOT an error: it is reported in fs_projection_result;
 *   - fs_last_error() returns the message of the last failing call on this thread;
 *   - a grid is owned by one thread (FluidGridBase.h:37-39); distinct grids may be used from
 *     distinct threads.
 *
 * Layouts are the reference's (FluidGridCPU.h:128-130, FluidGridTypes.h:10-20): cells W*H
 * row-major with y growing downward, u (W+1)*H, v W*(H+1), component masses component-major
 * (k*W*H + cell).  All host arrays are FP64 like FluidGridState<double>.
 */
#ifndef FLUIDSIM_HIP_H
#define FLUIDSIM_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum fs_status {
    FS_OK = 0,
    FS_E_INVALID_ARGUMENT = 1,
    FS_E_LOGIC = 2,
    FS_E_UNSUPPORTED = 3,
    FS_E_DEVICE = 4,
    FS_E_RUNTIME = 5
} fs_status;

typedef struct fs_context fs_context; /* one HIP device + one stream */
typedef struct fs_grid fs_grid;       /* one simulation instance (FluidGridBase object) */

/* FluidProjectionResult (FluidGridTypes.h:59-79) plus solver diagnostics. */
typedef struct fs_projection_result {
    int iterations;
    double max_divergence;
    int converged;
    int physical_steps;
    int transport_steps;
    int relaxed;
    double max_volume_error;
    double rms_volume_error;
    int viscosity_iterations;
    int precision_limited;
    /* diagnostics (no reference counterpart) */
    int restarts;               /* true-residual restarts inside PCG */
    double initial_residual;    /* max |b - A p| of the predicted pressure */
    double final_residual;      /* max |b - A p| at the last true-residual check */
    double rhs_max;             /* max |b| (b = -div u* h^2/dt, mean removed) */
} fs_projection_result;

/* FluidStepTimings (FluidGridTypes.h:49-57), milliseconds. */
typedef struct fs_step_timings {
    double transport_ms;
    double viscosity_ms;
    double forces_ms;
    double pressure_ms;
    double selection_ms;
    double regrid_ms;
} fs_step_timings;

/* Transport applied by fs_step.  FS_TRANSPORT_REFERENCE (the default) is the reference step's
 * CFL-split PLIC/VOF transport of fractions, component masses and momentum (FluidGridCPU.cpp:362-463),
 * bit-identical to UniformCPU.  FS_TRANSPORT_NONE runs forces + projection only.
 * FS_TRANSPORT_SEMI_LAGRANGIAN is the build's own mode. */
typedef enum fs_transport {
    FS_TRANSPORT_REFERENCE = 0,
    FS_TRANSPORT_NONE = 1,
    FS_TRANSPORT_SEMI_LAGRANGIAN = 2
} fs_transport;

/* FluidGridCPU::PressurePreconditioner (FluidGridCPU.h:39-44). */
typedef enum fs_preconditioner { FS_PRECOND_DIAGONAL = 0, FS_PRECOND_MULTIGRID = 1 } fs_preconditioner;

/* Storage of the multigrid levels that the fused smoothers stream (no reference counterpart: the
 * reference's UniformCPU hierarchy is FP64, PressureMultigrid.h:9-69).  Arithmetic is FP32 in every
 * mode, the CG vectors are FP32 and the dot products, residual norms and accumulated pressure FP64.
 *   FS_MG_FP32          conductances and level vectors FP32 (default);
 *   FS_MG_FP16          conductances of the tiled levels as IEEE binary16, level vectors FP32;
 *   FS_MG_BF16          conductances of the tiled levels as bfloat16, level vectors FP32;
 *   FS_MG_BF16_VECTORS  conductances and the rhs/solution vectors of the coarser tiled levels as
 *                       bfloat16 (the finest level's rhs and solution are the FP32 CG residual and z).
 *                       Rounding every stored vector makes the V-cycle linear only to ~2^-9, so the CG
 *                       takes the flexible (Polak-Ribiere) beta; hard problems (obstacles with 1000:1
 *                       density jumps at 2048^2) can stall.  binary16 vectors are not offered: its range
 *                       cannot hold a level's rhs and solution, which differ by ~cells/conductance.
 * With 16-bit conductances the preconditioner is a fixed SPD operator (rounded once per hierarchy
 * rebuild), so the CG keeps its Fletcher-Reeves beta.  The FP64 hierarchy, the coarse factor and the
 * in-LDS tail are unchanged; a grid whose whole hierarchy fits the tail has no tiled level and runs
 * FP32 whatever the mode. */
typedef enum fs_mg_precision {
    FS_MG_FP32 = 0,
    FS_MG_FP16 = 1,
    FS_MG_BF16 = 2,
    FS_MG_BF16_VECTORS = 3
} fs_mg_precision;

/* Smoother of the pressure multigrid's levels above 4096 cells (no reference counterpart as a setting: the
 * reference's CPU multigrid runs red-black Gauss-Seidel, PressureMultigrid.cpp:204-231, and its WebGPU
 * multigrid Chebyshev-weighted damped Jacobi on levels above 65536 cells, PressureMultigridGPU.cpp:24-44,
 * 619-631).  Both are symmetric, so the V-cycle stays an SPD preconditioner; the choice changes the
 * preconditioner (iteration counts, results within the tolerance), not the equation.  Levels of at most
 * 4096 cells run red-black Gauss-Seidel with their own sweep count (coarse sweeps, 1..4).
 *   FS_SMOOTHER_RED_BLACK  red-black Gauss-Seidel, 1..4 sweeps per side (default: 4, the reference's cycle);
 *   FS_SMOOTHER_CHEBYSHEV  damped Jacobi with the weights 1 / (1.1 + 0.9 cos(pi (k + 1/2) / n)), k = 0..n-1,
 *                          in order before the coarse correction and reversed after it, n = 1..8 sweeps per
 *                          side; FP32 multigrid levels only.  A cell without conductance (solid masks) is
 *                          set to zero, as the red-black sweep sets it. */
typedef enum fs_pressure_smoother { FS_SMOOTHER_RED_BLACK = 0, FS_SMOOTHER_CHEBYSHEV = 1 } fs_pressure_smoother;

/* ---- library / context ---- */
const char* fs_version(void);
/* Message of the last failed call on this thread ("" if none).  No reference counterpart: the C
 * ABI cannot carry C++ exceptions, so it keeps their what() text here. */
const char* fs_last_error(void);
/* Opens HIP device `device` with its work stream and a helper stream (the multigrid hierarchy is rebuilt
 * on the helper while the work stream forms the right-hand side). */
fs_status fs_context_create(int device, fs_context** out);
/* As fs_context_create, with flags (no reference counterpart):
 *   FS_CONTEXT_SINGLE_STREAM  the work stream only (the hierarchy rebuild then runs in line);
 *   FS_CONTEXT_PRIVATE_QUEUE  the work stream on a hardware queue of its own: it is created with a CU
 *                             mask that covers every compute unit (hipExtStreamCreateWithCUMask), and
 *                             the HIP runtime does not share such a queue with other streams.  Without
 *                             it, streams share the process's pool of GPU_MAX_HW_QUEUES queues in an
 *                             order set by which streams the process created and used before, so two
 *                             busy streams can end up on one queue and run one after the other.
 * The serving loop's lanes use both. */
enum { FS_CONTEXT_SINGLE_STREAM = 1, FS_CONTEXT_PRIVATE_QUEUE = 2 };
fs_status fs_context_create_ex(int device, int flags, fs_context** out);
fs_status fs_context_destroy(fs_context* ctx);
fs_status fs_context_synchronize(fs_context* ctx);

/* ---- grid lifetime: FluidGridCPU(int, int, double) (FluidGridCPU.h:63), destroy() (FluidGridBase.h:129) ---- */
fs_status fs_grid_create(fs_context* ctx, int width, int height, double cell_size, fs_grid** out);
fs_status fs_grid_destroy(fs_grid* grid);

/* getWidth/getHeight/getCellSize/getOrigin/setOrigin (FluidGridBase.h:31-35) */
fs_status fs_grid_shape(const fs_grid* grid, int* width, int* height, double* cell_size);
fs_status fs_set_origin(fs_grid* grid, double x, double y);
fs_status fs_get_origin(const fs_grid* grid, double* x, double* y);

/* configureMaterials(SubstanceCatalogue, FluidMixture liquid, FluidMixture gas) (FluidGridBase.h:42):
 * component densities and dynamic viscosities in mixture order (at most 64 per phase).  Viscosities
 * must be finite and non-negative.  Any positive viscosity makes fs_step run the implicit viscosity
 * solve after transport, as the reference's WebGPU backend does (FluidGridGPU.cpp:1134-1145); the
 * reference's CPU backend rejects viscosity (FluidGridCPU.cpp:200-201). */
fs_status fs_configure_materials(fs_grid* grid, int n_liquid, const double* liquid_density,
    const double* liquid_viscosity, int n_gas, const double* gas_density, const double* gas_viscosity);

/* setComponentViscosity(Phase, size_t component, double coefficient) (FluidGridBase.h:48-52,
 * FluidGridGPU.cpp:976-985): phase 0 = liquid, 1 = gas. */
fs_status fs_set_component_viscosity(fs_grid* grid, int phase, int component, double coefficient);

/* getGravityY/setGravityY (FluidGridBase.h:70-71) */
fs_status fs_set_gravity(fs_grid* grid, double gravity_y);
fs_status fs_get_gravity(const fs_grid* grid, double* gravity_y);
/* setSurfaceTension (FluidGridBase.h:76-80, FluidGridCPU.cpp:835-841): finite and non-negative, else
 * FS_E_INVALID_ARGUMENT.  A positive coefficient makes fs_step subdivide dt into capillary substeps and
 * add the capillary face force before each projection (FluidGridCPU.cpp:319-357, 862-926). */
fs_status fs_set_surface_tension(fs_grid* grid, double coefficient);
/* capillaryStepLimit (FluidGridBase.h:84, FluidGridCPU.cpp:843-860): 0.5 sqrt((rho_l,min + rho_g,min) h^3 /
 * (2 pi sigma)), +inf without surface tension or without one of the phases. */
fs_status fs_capillary_step_limit(const fs_grid* grid, double* limit);
/* FluidGridCPU::pressure_preconditioner (FluidGridCPU.h:39-44) */
fs_status fs_set_preconditioner(fs_grid* grid, fs_preconditioner kind);
fs_status fs_set_transport(fs_grid* grid, fs_transport mode);

/* Solid obstacles.  No counterpart in the reference, whose only walls are the domain border
 * (FluidGridCPU.cpp:1050-1061, 1170-1185); this generalises that wall rule.  `mask` holds W*H bytes,
 * nonzero = solid cell; NULL removes the mask.  Every face with a solid cell on either side gets zero
 * pressure face weight and its velocity is zeroed by the projection's wall step, so no flow enters a
 * solid.  The fluid (non-solid) cells must be non-empty and form one 4-connected region, else
 * FS_E_INVALID_ARGUMENT (an enclosed pocket would add a null space per pocket).  Coarse multigrid
 * unknowns left without conductance, or cut off by coarsening, are pinned at zero by the coarse solve.
 * Viscosity, transport and surface tension see the zeroed face velocities; the mask is configuration,
 * not state: checkpoints do not carry it. */
fs_status fs_set_solid_mask(fs_grid* grid, const unsigned char* mask);
/* Multigrid level storage (fs_mg_precision above).  FS_E_INVALID_ARGUMENT for an unknown mode. */
fs_status fs_set_multigrid_precision(fs_grid* grid, fs_mg_precision mode);
fs_status fs_get_multigrid_precision(const fs_grid* grid, fs_mg_precision* mode);
/* Flexible CG (no reference counterpart; the reference's solvePressure uses the Fletcher-Reeves beta,
 * FluidGridCPU.cpp:1807-1816): nonzero makes the PCG use the Polak-Ribiere beta
 * z_new.(r_new - r_old) / z_old.r_old (Notay 2000), which tolerates a preconditioner that is not exactly
 * linear.  FS_MG_BF16_VECTORS turns it on by itself. */
fs_status fs_set_flexible_cg(fs_grid* grid, int enabled);
/* Number of solid cells of the current mask (0 without one). */
fs_status fs_solid_cells(const fs_grid* grid, long long* count);
/* Relative stopping rule (extension, no reference counterpart: the reference's tolerance is absolute,
 * FluidGridCPU.cpp:1030-1124).  When enabled, the divergence_tolerance of fs_step / fs_project /
 * fs_batch_project / fs_advance_frame is a relative residual: each projection's PCG stops at
 * max|b - A p| <= tol * max|b| (b the mean-free pressure rhs) and the projection is accepted when
 * max|div u| <= tol * max|b| * dt / h^2, i.e. tol times the divergence the projection started from. */
fs_status fs_set_relative_tolerance(fs_grid* grid, int enabled);
/* Pressure smoother choice (see fs_pressure_smoother): `kind` with `sweeps` per side on the levels above 4096
 * cells, red-black with `coarse_sweeps` (1..4) per side on the smaller ones.  (FS_SMOOTHER_RED_BLACK, 4, 4)
 * is the reference's V(4,4).  FS_E_UNSUPPORTED for Chebyshev on a grid with 16-bit multigrid levels.
 * Batched solves (fs_batch_project) run red-black smoothing with the grid's sweep counts: a grid whose
 * levels above 4096 cells run Chebyshev projects on its own in fs_batch_step and is refused by
 * fs_batch_project. */
fs_status fs_set_pressure_smoother(fs_grid* grid, fs_pressure_smoother kind, int sweeps, int coarse_sweeps);
fs_status fs_get_pressure_smoother(const fs_grid* grid, fs_pressure_smoother* kind, int* sweeps, int* coarse_sweeps);
/* Autotuner of the smoother and its sweep counts (no reference counterpart): projects the grid's current
 * state with each candidate (red-black 2, 3, 4 and Chebyshev 3, 4, 5, 6, 8 sweeps, each with 4 and with 2 coarse
 * sweeps; each from the same state, restored after every trial: the grid's state is unchanged on return)
 * with fs_project's arguments, times each projection after one warm-up, and selects the fastest candidate
 * whose projection converges.  kind, sweeps, coarse_sweeps and ms (each may be NULL) receive the choice and
 * its time per projection. */
fs_status fs_tune_pressure_smoother(fs_grid* grid, double dt, int max_iterations, double divergence_tolerance,
    fs_pressure_smoother* kind, int* sweeps, int* coarse_sweeps, double* ms);
fs_status fs_get_relative_tolerance(const fs_grid* grid, int* enabled);

/* upload(const State&) (FluidGridBase.h:118, FluidGridCPU.cpp:198-251).  NULL means an empty
 * vector (zeros).  Sizes are implied by the grid and the configured mixtures.  Resets the pressure
 * history. */
fs_status fs_upload(fs_grid* grid, const double* liquid_fraction, const double* u, const double* v,
    const double* pressure, const double* liquid_mass, const double* gas_mass);
/* download() (FluidGridBase.h:119).  NULL skips a field. */
fs_status fs_download(fs_grid* grid, double* liquid_fraction, double* u, double* v, double* pressure,
    double* liquid_mass, double* gas_mass);
/* invalidateCaches() (FluidGridCPU.h:111) */
fs_status fs_invalidate_caches(fs_grid* grid);

/* step(dt, max_iterations, divergence_tolerance, timings) (FluidGridBase.h:133-134,
 * FluidGridCPU.cpp:300-360).  `timings` may be NULL.  max_iterations < 0 solves each projection to
 * tolerance (the uniform WebGPU backend's unbounded budget, FluidGridGPU.cpp:1326), stopping after
 * 100000 PCG iterations; the reference's CPU backend returns an unconverged result instead. */
fs_status fs_step(fs_grid* grid, double dt, int max_iterations, double divergence_tolerance,
    fs_step_timings* timings, fs_projection_result* result);
/* FluidGridCPU::project(dt, max_iterations, divergence_tolerance) (FluidGridCPU.cpp:1030-1124) */
fs_status fs_project(fs_grid* grid, double dt, int max_iterations, double divergence_tolerance,
    fs_projection_result* result);
/* FluidGridCPU::applyGravity(dt) (FluidGridCPU.cpp:988-996) */
fs_status fs_apply_gravity(fs_grid* grid, double dt);
/* FluidGridCPU::addVortex(center, radius, strength) (FluidGridCPU.cpp:998-1028) */
fs_status fs_add_vortex(fs_grid* grid, double center_x, double center_y, double radius, double strength);
/* maxDivergence() (FluidGridBase.h:135, FluidGridCPU.cpp:1126-1152) */
fs_status fs_max_divergence(fs_grid* grid, double* out);
/* FluidGridCPU::advect(dt) (FluidGridCPU.h:121, FluidGridCPU.cpp:362-463): reference transport alone.
 * `substeps` (may be NULL) receives the number of CFL substeps taken.  Velocities that would need more
 * than 1e6 substeps return FS_E_RUNTIME instead of looping. */
fs_status fs_advect(fs_grid* grid, double dt, int* substeps);
/* FluidGridGPU::applyViscosity(dt) (FluidGridGPU.cpp:987-1046): blend each cell's (rho, mu) from the
 * component masses (mu = sum(m mu/rho) / sum(m/rho)), then solve M (v - v0) + dt K v = 0 to the reference
 * tolerance 1e-4.  *iterations receives the dispatched CG iterations (a multiple of 4; 0 when dt <= 0 or
 * no component is viscous).  Breakdown, non-finite input or non-convergence return FS_E_RUNTIME with the
 * velocities unchanged, as the reference throws. */
fs_status fs_apply_viscosity(fs_grid* grid, double dt, int* iterations);
/* Viscosity smoother (no reference counterpart; the reference's solve smooths every multigrid level with two
 * damped-Jacobi sweeps of weight 1/2 per side, ViscosityKernels.h:214-236, ViscositySolverGPU.cpp:134-167).
 * chebyshev_sweeps 2, 4 or 6 smooths with that many Jacobi sweeps per side weighted
 * 1 / (2.125 + 1.875 cos(pi (k + 1/2) / n)) -- the reciprocals of the Chebyshev nodes of [0.25, 4], 4
 * bounding the spectrum of D^-1 A -- so that the CG needs fewer iterations; 0 (default) restores the
 * reference's smoother.  Applies to fs_apply_viscosity, fs_step and fs_batch_apply_viscosity on this grid.
 * FS_E_INVALID_ARGUMENT for another value. */
fs_status fs_set_viscosity_smoother(fs_grid* grid, int chebyshev_sweeps);
/* Viscosity smoother autotuning (no reference counterpart): times fs_apply_viscosity(dt) of the current state
 * with each smoother (0, 2, 4, 6; best of two runs after a warm-up), keeps the fastest that converges and
 * leaves the velocities as they were.  *chebyshev_sweeps and *ms (may be NULL) receive the choice and its
 * time.  FS_E_LOGIC when the grid has no viscosity or no uploaded state. */
fs_status fs_tune_viscosity_smoother(fs_grid* grid, double dt, int* chebyshev_sweeps, double* ms);
/* ViscositySolverGPU::solve(u, v, material, dt, cell_size, tolerance, maximum)
 * (ViscositySolverGPU.cpp:169-294, .h:12-28) on host arrays: material is (rho, mu) per cell (2*W*H),
 * u ((W+1)*H) and v (W*(H+1)) are updated in place on success only.  max_iterations < 0 is the
 * reference's unbounded mode (capped at 100000 here).  Failures return FS_E_RUNTIME. */
fs_status fs_viscosity_solve(fs_context* ctx, int width, int height, double cell_size, const double* material,
    double* u, double* v, double dt, double tolerance, int max_iterations, int* iterations,
    double* relative_residual);
/* No counterpart in the reference.  For callers that cannot allocate device memory (code that must not
 * call hipMalloc, tight serving loops, harnesses that forbid allocation): fs_viscosity_solve on DEVICE
 * arrays material (2*W*H), u ((W+1)*H) and v (W*(H+1)) with the same algorithm, stopping rule and results,
 * making no allocation of any kind.  Its workspace is a static device array of 64 MiB per GPU (held until
 * the process ends): about 123 bytes per cell, up to about 737 x 737 cells; larger grids return
 * FS_E_INVALID_ARGUMENT.  The
 * work runs on the context's stream, which the call drains before returning (so the caller orders its own
 * writes of the inputs before the call).  Every call on a GPU shares that GPU's array, which keeps the last
 * call's values (each call clears what it uses first); calls that overlapped on one GPU would race on it,
 * so the library serialises them within a process. */
fs_status fs_viscosity_solve_noalloc(fs_context* ctx, int width, int height, double cell_size,
    const double* material, double* u, double* v, double dt, double tolerance, int max_iterations, int* iterations,
    double* relative_residual);
/* FluidGridCPU::applySurfaceTension(dt) (FluidGridCPU.cpp:862-926; private in the reference, exposed for
 * parity tests): capillary geometry of the current fractions, then u, v += dt w sigma kappa df / h. */
fs_status fs_apply_surface_tension(fs_grid* grid, double dt);
/* Capillary geometry of the current fractions (FluidSurfaceTensionCPU.h:134-376), for parity tests:
 * heights[4N] = (height x, height y, orientation x, orientation y), curvature[2N] = (value, valid). */
fs_status fs_capillary_geometry(fs_grid* grid, double* heights, double* curvature);
/* Semi-Lagrangian transport alone (build-defined mode, no reference counterpart). */
fs_status fs_advect_semi_lagrangian(fs_grid* grid, double dt);

/* writeColorPixels(std::vector<float>& pixels, bool liquid_fraction_alpha) (FluidGridBase.h:139,
 * FluidGridCPU.cpp:1826-1860): pixels[4*W*H] receives RGBA per cell: RGB = sum over the components (liquid
 * first, then gas, in mixture order) of (mass / (density h^2)) * colour, each term rounded to float and
 * summed in float; A = the liquid fraction when liquid_fraction_alpha, else 1.  component_rgb holds 3
 * floats per component in the same order: the colours of the SubstanceCatalogue entries the mixtures
 * name (SubstanceCatalogue.h:14-21), looked up by the adapter. */
fs_status fs_color(fs_grid* grid, const float* component_rgb, int liquid_fraction_alpha, float* pixels);

/* One resolved material edit: detail::ResolvedCellEdit (FluidGridEdit.h:22-28).  component indexes the
 * liquid components, then the gas components; mass is the component's density times h^2 as
 * CellEditBatch::prepare computes it (FluidGridEdit.h:37-95). */
typedef struct fs_cell_edit {
    int x, y;
    int component;
    double mass;
} fs_cell_edit;
/* commitEdits(CellEditBatch&) (FluidGridBase.h:142, FluidGridCPU.cpp:253-284): each edited cell becomes
 * pure liquid (fraction 1) or pure gas (fraction 0) holding only its component's mass; face velocities
 * are kept.  The edits are applied in raster order and the last edit of a cell wins (as prepare() leaves
 * them).  A cell outside the grid, an unconfigured component or a non-finite/negative mass returns
 * FS_E_INVALID_ARGUMENT before anything changes; before an upload of the configured mixtures,
 * FS_E_LOGIC.  velocity (2 doubles, the optional edit velocity) must be NULL: like UniformCPU, this
 * backend returns FS_E_UNSUPPORTED ("requires uniform WebGPU") for it. */
fs_status fs_commit_edits(fs_grid* grid, int count, const fs_cell_edit* cells, const double* velocity);

/* ---- device-resident snapshots (no reference counterpart): keep the inputs in HBM and restore
 * them between timed solves or serving requests without a host round trip. ---- */
fs_status fs_snapshot_save(fs_grid* grid);
fs_status fs_snapshot_restore(fs_grid* grid);

/* ---- multigrid hierarchy introspection (PressureMultigrid.cpp:24-140), for parity checks ---- */
fs_status fs_mg_levels(fs_grid* grid, int* levels);
fs_status fs_mg_level_shape(fs_grid* grid, int level, int* width, int* height);
/* FP64 cell centres of a level (lengths width, height). */
fs_status fs_mg_level_centers(fs_grid* grid, int level, double* cx, double* cy);
/* FP64 conductances (right, down) and FP32 1/diagonal as stored on the device, interior w*h each.
 * Rebuilds coefficients and hierarchy from the current state first. */
fs_status fs_mg_level_weights(fs_grid* grid, int level, double* right, double* down, float* inv_diag);
/* Prolongation maps of a fine level (axis 0: x, 1: y), lengths = that level's width/height. */
fs_status fs_mg_level_prolong(fs_grid* grid, int level, int axis, int* first, int* second, double* weight);
/* Restriction stencils of the next coarser level: count[n], index[4n], weight[4n]. */
fs_status fs_mg_level_restrict(fs_grid* grid, int level, int axis, int* count, int* index, double* weight);
/* Autotuned kernel tiles of a level (build's own): pre- and post-smoothing variants of the fused
 * four-sweep kernel (0 = 64x32/256 threads, 1 = 32x32/256, 2 = 32x16/256, 3 = 16x16/256, 4 = 64x32/512,
 * 5 = 32x32/512, 6 = 32x32/384, 7 = 64x32/640, 8 = 32x64/640, 9 = 64x32/960, 10 = 32x64/960) and the
 * residual-restriction variant (0 = 32x16, 1 = 16x16, 2 = 16x8, 3 = 8x8 coarse
 * cells); -1 for levels solved inside the single-workgroup tail.  On a level that runs the Chebyshev smoother,
 * pre and post are tiles of the Jacobi kernel instead (0 = 64x32/512, 1 = 32x32/256, 2 = 32x16/256,
 * 3 = 16x16/256, 4 = 32x32/512, 5 = 64x64/1024, 6 = 64x32/1024, 7 = 32x64/1024), and pre 8, 9, 10 mean
 * that the pre-smoothing runs fused with the residual and restriction (16x8/512, 16x16/512, 32x16/1024
 * coarse cells per workgroup; same values). */
fs_status fs_mg_level_tiles(fs_grid* grid, int level, int* pre, int* post, int* restriction);
/* Overrides the autotuned variants of a tiled level (-1 keeps one): every variant computes the same
 * values, so this changes time only -- for profiling one fixed configuration and for testing that claim.
 * The autotuner may run a Chebyshev level's pre-smoothing fused with its residual and restriction (one
 * launch, the same values): pre 0..7 selects the separate launches with that Jacobi tile, 8..10 the fused
 * launch with the tiles fs_mg_level_tiles lists. */
fs_status fs_mg_set_level_tiles(fs_grid* grid, int level, int pre, int post, int restriction);
/* Coarsest dense Cholesky factor, n*n row-major (lower triangle); *n receives the unknown count. */
fs_status fs_mg_coarse_factor(fs_grid* grid, double* factor, int* n);
/* One V-cycle: correction = M^-1 (residual - mean), both W*H FP64 host arrays (converted to FP32 on
 * the device). */
fs_status fs_mg_apply(fs_grid* grid, const double* residual, double residual_mean, double* correction);
/* FP64 face weights inv_density_u ((W+1)*H) and inv_density_v (W*(H+1)) (FluidGridCPU.cpp:1154-1205). */
fs_status fs_face_weights(fs_grid* grid, double* inv_density_u, double* inv_density_v);

/* ---- measurement (no reference counterpart) ---- */
/* Times `reps` back-to-back launches of one hot kernel on the grid's stream with HIP events and
 * returns the mean milliseconds per launch.  kernel: 0 = finest-level post-smoothing (prolongation +
 * the grid's smoother sweeps, as in the V-cycle),
 * 1 = operator + dot (A d), 2 = PCG update, 3 = FP32 copy of one padded vector (reference kernel),
 * 4 = second level's smoother, 5 = finest residual + restriction, 6 = coarse tail, 7 = coarse
 * inverse (sweep operator), 8 = direction update, 9 = r.z reduction, 10 = FP32 copy of a 64 MiB buffer
 * into another (HBM copy-bandwidth reference; 128 MiB moved per launch). */
fs_status fs_time_kernel(fs_grid* grid, int kernel, int reps, double* ms_per_launch);
/* Cumulative device bytes allocated by this grid. */
fs_status fs_grid_device_bytes(const fs_grid* grid, size_t* bytes);

/* Batched implicit viscosity (no reference counterpart; each instance is FluidGridGPU::applyViscosity,
 * FluidGridGPU.cpp:987-1046: blend, backward-Euler solve to 1e-4): one persistent launch runs the CG loops
 * of all `count` grids (one context), one workgroup per grid.  Grids without viscosity, or dt <= 0, are
 * left as they are (0 iterations).  An instance that needs more than 500 iterations is solved again on
 * its own.  iterations (may be NULL) receives each grid's count.  A failing instance (non-finite state,
 * breakdown, no convergence) leaves its own velocities unchanged and makes the call return FS_E_RUNTIME
 * with the first failure's message after every other instance has been applied. */
fs_status fs_batch_apply_viscosity(int count, fs_grid* const* grids, double dt, int* iterations);

/* ---- batched projection (no reference counterpart: the reference projects one grid per call,
 * FluidGridCPU.cpp:1030-1124) ----
 * fs_project on `count` grids of one context at once: every grid's setup (coefficients, hierarchy, rhs,
 * predicted pressure, first direction) is launched on the context's stream, then ONE persistent launch
 * runs all grids' PCG iterations -- one 1024-thread workgroup per grid, looping on the device until that
 * grid converges or stops -- and then every grid's velocity correction.  Each grid's result follows
 * fs_project's contract (same stopping rule, same result fields) and does not depend on the other
 * grids; the FP64 reductions run in another order than fs_project's, so values agree with it to
 * rounding, not bit for bit.  tolerances[i] is grid i's divergence tolerance. */
fs_status fs_batch_project(int count, fs_grid* const* grids, double dt, int max_iterations,
    const double* divergence_tolerances, fs_projection_result* results);
/* Autotuner of the batched solves' red-black sweep counts (no reference counterpart): runs fs_batch_project
 * on the `count` grids with every grid set to (FS_SMOOTHER_RED_BLACK, sweeps, coarse) for sweeps 4, 3, 2
 * and coarse 4, 2 -- each from the grids' current state, restored after every trial: their state is
 * unchanged on return -- times each batch after one warm-up, and sets every grid to the fastest choice
 * under which every grid's projection converges.  sweeps, coarse_sweeps and ms (each may be NULL) receive
 * the choice and its time per batch.  FS_E_UNSUPPORTED for grids with 16-bit multigrid levels. */
fs_status fs_tune_batch_smoother(int count, fs_grid* const* grids, double dt, int max_iterations,
    const double* divergence_tolerances, int* sweeps, int* coarse_sweeps, double* ms);
/* Split of fs_batch_project and fs_batch_apply_viscosity (no reference counterpart): grids of more than
 * `solo_cells` cells are solved by fs_project's / fs_apply_viscosity's own multi-kernel solve, one after
 * another on the context's stream, while the persistent launch runs the smaller ones on a second stream
 * -- one workgroup is slow on a large grid, and the batch lasts as long as its slowest instance.  The
 * route depends on a grid's size only, so a grid's result is the same alone and in any batch.  0 (the
 * default) batches every grid; contexts created with FS_CONTEXT_SINGLE_STREAM always do. */
fs_status fs_context_set_batch_split(fs_context* ctx, long long solo_cells);

/* Batched timestep (no reference counterpart: each instance is FluidGridCPU::step, FluidGridCPU.cpp:300-360,
 * as fs_step runs it): fs_step on `count` grids of one context at once, each with its own dts[i],
 * max_iterations[i] and divergence_tolerances[i].  Per capillary substep, every grid's transport, gravity and
 * surface tension are launched one grid after another on the context's stream, while the implicit viscosity
 * solves and the projections run batched (fs_batch_apply_viscosity's and fs_batch_project's persistent
 * launches, with the context's split; grids with 16-bit multigrid levels project one by one).  results[i]
 * follows fs_step's contract: a grid stops after its first substep whose projection does not converge, and
 * dts[i] <= 0 leaves it untouched with physical_steps = 0.  Values agree with fs_step to rounding (the batched
 * solves reduce in another order), and a grid's result does not depend on the other grids of the batch.  A
 * grid whose step fails (transport, viscosity or solver error) stops there; statuses[i] (may be NULL)
 * receives FS_OK or FS_E_RUNTIME per grid, and the call returns FS_E_RUNTIME with the lowest failing index's
 * message after every other grid has been stepped. */
fs_status fs_batch_step(int count, fs_grid* const* grids, const double* dts, const int* max_iterations,
    const double* divergence_tolerances, fs_projection_result* results, fs_status* statuses);

/* ---- device footprint (no reference counterpart) ----
 * Upper bound of the device bytes a grid of this shape allocates (fs_grid_device_bytes reads the actual
 * sum once it exists): the always-present state, pressure solver and hierarchy, plus the stages named
 * in `features` (allocated on their first use). */
enum {
    FS_FEATURE_TRANSPORT = 1,          /* reference PLIC/VOF transport */
    FS_FEATURE_VISCOSITY = 2,          /* implicit viscosity solve */
    FS_FEATURE_SURFACE_TENSION = 4,    /* capillary geometry */
    FS_FEATURE_SOLID_MASK = 8,         /* fs_set_solid_mask */
    FS_FEATURE_MG16 = 16               /* 16-bit multigrid levels (any fs_mg_precision but FP32) */
};
fs_status fs_grid_estimate_bytes(int width, int height, int n_liquid, int n_gas, int features, size_t* bytes);

/* ---- configuration getters (FluidGridBase.h:43-46, 70-80; FluidGridCPU.h:39-44) ----
 * getSurfaceTension, pressure_preconditioner, the transport mode, and the configured mixture
 * (FluidMixture components: density and dynamic viscosity), so that an adapter can refresh its own
 * view after fs_checkpoint_load.  phase 0 = liquid, 1 = gas. */
fs_status fs_get_surface_tension(const fs_grid* grid, double* coefficient);
fs_status fs_get_preconditioner(const fs_grid* grid, fs_preconditioner* kind);
fs_status fs_get_transport(const fs_grid* grid, fs_transport* mode);
fs_status fs_component_counts(const fs_grid* grid, int* n_liquid, int* n_gas);
fs_status fs_get_component(const fs_grid* grid, int phase, int component, double* density, double* viscosity);

/* ---- resident-instance checkpoint: saveCheckpoint / loadCheckpoint (FluidGridBase.h:123-125,
 * FluidSim/FluidCheckpoint.h:9-15, FluidSim/UniformWebGPU/FluidGridGPU.cpp:566-695 and 697-880) ----
 * A checkpoint is JSON metadata (backend "hip-uniform", version 1, scalar "float64", byte order,
 * geometry, origin, gravity, surface tension, preconditioner, transport, both mixtures' densities and
 * viscosities, the previous solve's iteration count, the payload layout) plus a native payload of FP64
 * values in the fixed order liquid_fraction (W*H), u ((W+1)*H), v (W*(H+1)), pressure (W*H), previous
 * pressure (W*H), liquid masses (n_liquid*W*H), gas masses (n_gas*W*H).  Both pressure histories are
 * kept, so a loaded instance continues bit for bit as the saved one would have.
 *
 * fs_checkpoint_size: bytes of the metadata text including its terminating NUL, and of the payload.
 * fs_checkpoint_save: writes both; FS_E_LOGIC before the first upload, FS_E_INVALID_ARGUMENT if a
 *   buffer is smaller than fs_checkpoint_size reports.  Saving does not change the instance.
 * fs_checkpoint_load: `metadata` is `metadata_length` bytes of JSON text (members it does not know are
 *   ignored, so an adapter may add its own).  Every field is validated before any live state changes;
 *   an incompatible format, another geometry, a missing or invalid member or a payload of the wrong
 *   size returns FS_E_RUNTIME with the instance untouched (the reference throws std::runtime_error).
 *   A load replaces the configuration and the state, resets derived caches and drops the
 *   device-resident snapshot. */
fs_status fs_checkpoint_size(const fs_grid* grid, size_t* metadata_bytes, size_t* payload_bytes);
fs_status fs_checkpoint_save(fs_grid* grid, char* metadata, size_t metadata_capacity, void* payload,
    size_t payload_capacity);
fs_status fs_checkpoint_load(fs_grid* grid, const char* metadata, size_t metadata_length, const void* payload,
    size_t payload_bytes);

/* ---- scene frame: FluidScene::sceneProcess (FluidSim/FluidScene.inl:328-410) ----
 * Settings mirror FluidSceneSettings (FluidScene.h:21-35): frame_time (default 0.02 s),
 * steps_per_second (default 60), projection_tolerance (default 0.1); collect_timings asks each step for
 * its stage timings (the scene's profile_stages). */
typedef struct fs_frame_settings {
    double frame_time;
    int steps_per_second;
    double projection_tolerance;
    int collect_timings;
} fs_frame_settings;

/* FluidScene::FrameStatistics (FluidScene.h:162-178) plus the schedule and the outcome. */
typedef struct fs_frame_statistics {
    double fluid_ms;            /* host wall time of the engine steps */
    fs_step_timings stages;     /* summed stage timings (zero unless collect_timings) */
    int substeps;               /* physical steps (capillary substeps included) */
    int transport_steps;
    int pressure_iterations;
    int viscosity_iterations;
    double projection_tolerance;
    double max_volume_error;
    double rms_volume_error;
    /* no reference counterpart */
    int steps_scheduled;        /* engine steps the frame was split into */
    int steps_run;              /* engine steps taken (fewer when one was not accepted) */
    int accepted;               /* 0 once a projection was not accepted: the scene stops simulating */
    double max_divergence;      /* of the last step */
} fs_frame_statistics;

/* One scene frame on a uniform grid of this backend: step_count = ceil(frame_time * steps_per_second)
 * equal steps, or a single step when ceil(frame_time / capillaryStepLimit) already reaches step_count
 * (the engine's capillary subdivision then provides the rate); each step runs fs_step with an unbounded
 * pressure budget (-1: to tolerance, at most 100000 iterations) and the frame stops at the first step
 * whose projection is not accepted.  `statistics` (may be NULL) is filled also when a step fails. */
fs_status fs_advance_frame(fs_grid* grid, const fs_frame_settings* settings, fs_frame_statistics* statistics);

/* ---- serving loop (no reference counterpart: the reference steps one scene per process,
 * FluidScene.inl:328-410; several scenes at most rotate on one thread) ----
 * A server keeps many simulations resident on one GPU and steps them concurrently.  `lanes` worker
 * threads each own a HIP stream; an admitted request is placed on the lane with the fewest resident
 * instances and stepped there in rotation with the lane's other instances (continuous batching: a
 * request joins its lane's rotation at the next step boundary, and leaves it when done).  Requests are
 * admitted in arrival order while the sum of the admitted instances' fs_grid_estimate_bytes stays within
 * the HBM budget; finishing an instance frees its device memory and admits the next waiting request.
 * Each step is fs_step(dt, max_iterations, divergence_tolerance) on the instance's own grid, so a
 * request's results do not depend on what else is resident.  With batch_cells > 0, the instances of at
 * most batch_cells cells take each round's step together instead: one fs_batch_step over the lane's small
 * instances (batched viscosity and projection in persistent launches), then fs_step for each larger one.
 * The route depends only on the request's own size, and a batched step's result does not depend on the
 * other instances, so the results still do not depend on what else is resident; they agree with fs_step's
 * to rounding.  Small instances are placed on the lane with the fewest resident small instances, larger ones
 * on the lane with the fewest larger ones.
 * Eviction under HBM pressure (FluidCheckpoint's metadata + payload, FluidGridGPU.cpp:567-695): with
 * evict_after_steps > 0, when the request at the head of the queue does not fit in the budget and the
 * instances that have run at least evict_after_steps steps since their admission would make room, the
 * longest-resident of them are marked; at its next step boundary each marked instance is saved with
 * fs_checkpoint_save into host memory, destroyed (its device memory freed), and queued again at the tail.
 * When readmitted it is rebuilt from the checkpoint (fs_checkpoint_load in place of the upload, then the
 * request's mask, precision and stopping rule) and continues: the checkpoint holds the whole stepping state
 * (fractions, velocities, both pressure histories, masses, the previous solve's iteration count), so an
 * evicted request's results equal the uninterrupted ones bit for bit.  A step that fails, or whose projection is not accepted, ends
 * that request as failed (the scene's simulation_failed rule) without affecting others.  Step latency is the
 * wall time of one step's execution on its lane: for a batched instance, its batch's fs_batch_step. */
typedef struct fs_server fs_server;

typedef struct fs_server_config {
    int device;
    int lanes;                    /* concurrent worker streams, 1..16 */
    size_t hbm_budget_bytes;      /* admission budget; 0 = 80% of the device memory free at creation */
    long long batch_cells;        /* instances of at most this many cells step together, 0 = none */
    int evict_after_steps;        /* eviction under HBM pressure after this many resident steps, 0 = never */
} fs_server_config;

typedef struct fs_request {
    int width, height;
    double cell_size;
    int steps;                    /* resident timesteps to run (>= 0) */
    double dt;
    int max_iterations;           /* pressure budget per projection, -1 = to tolerance */
    double divergence_tolerance;
    double gravity_y;
    double surface_tension;
    int transport;                /* fs_transport */
    int preconditioner;           /* fs_preconditioner */
    int n_liquid, n_gas;
    const double* liquid_density;   /* n_liquid */
    const double* liquid_viscosity; /* n_liquid, may be NULL */
    const double* gas_density;      /* n_gas */
    const double* gas_viscosity;    /* n_gas, may be NULL */
    /* initial state in fs_upload's layout; NULL means zeros.  Copied by fs_server_submit. */
    const double* liquid_fraction;
    const double* u;
    const double* v;
    const double* liquid_mass;
    const double* gas_mass;
    /* extensions (zero: FP32 levels, no obstacles) */
    int multigrid_precision;        /* fs_mg_precision */
    const unsigned char* solid;     /* W*H solid mask (fs_set_solid_mask) or NULL; copied by fs_server_submit */
    int relative_tolerance;         /* nonzero: divergence_tolerance is relative (fs_set_relative_tolerance) */
    /* pressure smoother (fs_set_pressure_smoother); smoother_sweeps 0 keeps the reference's V(4,4) */
    int pressure_smoother;          /* fs_pressure_smoother */
    int smoother_sweeps;            /* 0 or 1..4 (red-black) / 1..8 (Chebyshev) on the levels above 4096 cells */
    int coarse_sweeps;              /* 0 (= 4) or 1..4 on the smaller levels */
    int viscosity_sweeps;           /* fs_set_viscosity_smoother: 0 (the reference's) or 2, 4, 6 Chebyshev sweeps */
} fs_request;

enum { FS_REQUEST_QUEUED = 0, FS_REQUEST_RUNNING = 1, FS_REQUEST_DONE = 2, FS_REQUEST_FAILED = 3 };

typedef struct fs_request_report {
    int state;                    /* FS_REQUEST_* */
    int lane;                     /* -1 until admitted */
    int steps_done;
    int pressure_iterations;      /* summed over the steps */
    int viscosity_iterations;
    int physical_steps;           /* capillary substeps included */
    double max_divergence;        /* of the last step */
    double queue_ms;              /* submission -> admission */
    double run_ms;                /* admission -> end */
    double step_p50_ms, step_p99_ms, step_max_ms;
    size_t estimated_bytes;       /* admission footprint */
    size_t device_bytes;          /* allocated by the instance at its end */
    double setup_ms;              /* wall time to create, configure and upload the instance on its lane */
    double teardown_ms;           /* wall time to download its final state and free it */
    int evictions;                /* times the instance was checkpointed to host memory and left the device */
    char error[256];
} fs_request_report;

typedef struct fs_server_stats {
    long long submitted, admitted, completed, failed, steps;
    int queued, resident;
    size_t budget_bytes, resident_bytes, peak_resident_bytes;
    double elapsed_s;             /* since creation or the last fs_server_reset_stats */
    double steps_per_s;           /* steps completed in that window / elapsed_s */
    double step_p50_ms, step_p99_ms;   /* over the steps of that window */
    double setup_ms_total, teardown_ms_total;   /* lane time spent creating and retiring instances */
    long long evictions;          /* instances checkpointed out of the device (since creation) */
} fs_server_stats;

fs_status fs_server_create(const fs_server_config* config, fs_server** out);
/* Stops the lanes after their current step; unfinished requests end as failed. */
fs_status fs_server_destroy(fs_server* server);
/* Validates and copies a request; FS_E_INVALID_ARGUMENT if malformed or larger than the whole budget. */
fs_status fs_server_submit(fs_server* server, const fs_request* request, long long* id);
/* Waits up to timeout_s seconds (< 0: without limit) for the request to finish, then reports it. */
fs_status fs_server_wait(fs_server* server, long long id, double timeout_s, fs_request_report* report);
fs_status fs_server_query(fs_server* server, long long id, fs_request_report* report);
/* Final state of a finished request (fs_download's layout; NULL skips a field); FS_E_LOGIC otherwise. */
fs_status fs_server_result(fs_server* server, long long id, double* liquid_fraction, double* u, double* v,
    double* pressure, double* liquid_mass, double* gas_mass);
/* Forgets a finished request and its stored result. */
fs_status fs_server_release(fs_server* server, long long id);
/* Blocks until no request is queued or running. */
fs_status fs_server_drain(fs_server* server);
fs_status fs_server_get_stats(fs_server* server, fs_server_stats* stats);
fs_status fs_server_reset_stats(fs_server* server);

#ifdef __cplusplus
}
#endif
#endif
