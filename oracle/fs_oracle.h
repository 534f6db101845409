/*
 * fs_oracle.h -- FP64 CPU restatement of the reference UniformCPU pressure projection.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the parity oracle: only tests/, the smoke check
 * and the CPU-baseline leg of bench.py may load it, and only as the checker.  The
 * product path (fluidsim_amd, libfluidsim_hip.so) never links or calls it.
 *
 * Follows, single-threaded (the reference's bitwise-reproducible mode):
 *   FluidSim/UniformCPU/FluidGridCPU.cpp      (project, solvePressure, operator, gravity, vortices,
 *                                             PLIC/VOF transport, step)
 *   FluidSim/UniformCPU/PressureMultigrid.cpp (hierarchy, coarsening, RB-GS V(4,4), Cholesky)
 *   FluidSim/UniformWebGPU/ViscositySolverGPU.cpp + ViscosityKernels.h (implicit viscosity, FP64)
 *   FluidSim/UniformCPU/FluidSurfaceTensionCPU.h (capillary geometry)
 *
 * Layout conventions are the reference's: u is (W+1) x H, v is W x (H+1), cells W x H,
 * all row-major, y grows downward (FluidGridCPU.h:128-130).
 */
#ifndef FS_ORACLE_H
#define FS_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_mg orc_mg;
typedef struct orc_grid orc_grid;

/* Mirrors FluidProjectionResult (FluidGridTypes.h:59-79), fields used on this path. */
typedef struct orc_result {
    int iterations;
    double max_divergence;
    int converged;
    int physical_steps;
    int transport_steps;
    int restarts;            /* true-residual restarts inside the PCG loop (diagnostic) */
    double initial_residual; /* max |b - A p| before the loop (diagnostic) */
    double final_residual;   /* max |b - A p| at exit, recomputed in FP64 (diagnostic) */
    int viscosity_iterations;
} orc_result;

/* ---- PressureMultigrid (PressureMultigrid.h:9-69) ---- */
orc_mg* orc_mg_create(int width, int height);
void orc_mg_destroy(orc_mg* mg);
/* inv_density_u: (W+1)*H face weights, inv_density_v: W*(H+1). Returns 1 on success, 0 on pivot failure. */
int orc_mg_rebuild(orc_mg* mg, const double* inv_density_u, const double* inv_density_v);
/* correction[W*H] = V-cycle(residual - residual_mean). */
void orc_mg_apply(orc_mg* mg, const double* residual, double* correction, double residual_mean);
int orc_mg_levels(const orc_mg* mg);
/* Level shape; returns width, height via pointers. */
void orc_mg_level_shape(const orc_mg* mg, int level, int* width, int* height);
/* Copy out level data.  Arrays are interior-only (w*h) for right/down/inv_diag, w for cx, h for cy. */
void orc_mg_level_centers(const orc_mg* mg, int level, double* cx, double* cy);
void orc_mg_level_weights(const orc_mg* mg, int level, double* right, double* down, double* inv_diag);
/* Prolongation per axis: first[n], second[n], weight[n] for the fine level's width (axis 0) or height (axis 1). */
void orc_mg_level_prolong(const orc_mg* mg, int level, int axis, int* first, int* second, double* weight);
/* Restriction per axis for coarse index c: count[nc], idx[nc*4], wt[nc*4]. */
void orc_mg_level_restrict(const orc_mg* mg, int level, int axis, int* count, int* idx, double* wt);
/* Coarsest dense Cholesky factor (n*n, lower triangle meaningful); returns n. */
int orc_mg_coarse_factor(const orc_mg* mg, double* factor);
/* Solid-mask extension: pin vanishing coarse pivots (set by orc_set_solid for a grid's hierarchy). */
void orc_mg_set_pin(orc_mg* mg, int pin);
/* Pinned coarse unknowns of the last factorisation (n flags, may be NULL); returns their count. */
int orc_mg_coarse_pinned(const orc_mg* mg, unsigned char* pinned);

/* ---- FluidGridCPU subset (FluidGridCPU.h:15-225) ---- */
orc_grid* orc_grid_create(int width, int height, double cell_size);
void orc_grid_destroy(orc_grid* g);
void orc_set_origin(orc_grid* g, double x, double y);
void orc_set_gravity(orc_grid* g, double gravity_y);
/* 0 = Diagonal, 1 = Multigrid (FluidGridCPU.h:39-44). */
void orc_set_preconditioner(orc_grid* g, int kind);
/* Component densities, as configureMaterials (FluidMixture.h:14-39). */
int orc_configure(orc_grid* g, int n_liquid, const double* liquid_density, int n_gas, const double* gas_density);
/* Any pointer may be NULL (zero field).  Masses are component-major: k*N + cell. */
int orc_upload(orc_grid* g, const double* fraction, const double* u, const double* v, const double* pressure,
    const double* liquid_mass, const double* gas_mass);
void orc_download(const orc_grid* g, double* fraction, double* u, double* v, double* pressure,
    double* liquid_mass, double* gas_mass);
void orc_invalidate_caches(orc_grid* g);
/* Solid-mask extension (no counterpart in the reference): W*H bytes, nonzero = solid; NULL clears.
 * Faces touching a solid cell get zero pressure weight and zero velocity in project(); vanishing
 * coarse pivots (disconnected or fully blocked coarse regions) are pinned instead of failing. */
void orc_set_solid(orc_grid* g, const unsigned char* mask);
/* The grid's hierarchy, rebuilt from the current coefficients if stale; NULL if the factorisation fails. */
orc_mg* orc_grid_multigrid(orc_grid* g);
void orc_apply_gravity(orc_grid* g, double dt);
void orc_add_vortex(orc_grid* g, double cx, double cy, double radius, double strength);
double orc_max_divergence(const orc_grid* g);
/* writeColorPixels (FluidGridCPU.cpp:1826-1860): pixels[4N] RGBA; rgb = 3 floats per component. */
void orc_color_pixels(const orc_grid* g, const float* rgb, int liquid_fraction_alpha, float* pixels);
/* commitEdits (FluidGridCPU.cpp:253-284) on resolved, reduced cells (CellEditBatch::prepare). */
void orc_commit_edits(orc_grid* g, int count, const int* x, const int* y, const int* component, const double* mass);
/* project(dt, max_iterations, divergence_tolerance) -- FluidGridCPU.cpp:1030-1124. */
orc_result orc_project(orc_grid* g, double dt, int max_iterations, double divergence_tolerance);
/* Gravity followed by projection: step() with transport and surface tension off. */
orc_result orc_force_project(orc_grid* g, double dt, int max_iterations, double divergence_tolerance);
/* Worker threads for the projection's passes (default 1: the reference's single-thread, bitwise
 * reproducible order).  With n > 1 the passes use the reference's worker counts (up to 8 on passes of at
 * least 65536 items, FluidGridCPU.cpp:25-31; 4 or 8 on multigrid levels of at least 65536 cells,
 * PressureMultigrid.cpp:13-21), at most n.  Process-wide. */
void orc_set_threads(int threads);
int orc_get_threads(void);
/* Smoother of the multigrid levels above 4096 cells (an execution choice, matching fs_set_pressure_smoother):
 * ORC_SMOOTH_RBGS (the reference CPU's red-black Gauss-Seidel, PressureMultigrid.cpp:204-231) or
 * ORC_SMOOTH_CHEBYSHEV (the reference WebGPU's Chebyshev-weighted damped Jacobi,
 * PressureMultigridGPU.cpp:24-28, 619-631), with `sweeps` (1..8) per side; levels of at most 4096 cells run
 * `coarse_sweeps` (1..8) red-black sweeps.  Default RBGS 4, 4: the reference's cycle.  Returns -1 for an
 * invalid choice.  Process-wide. */
enum { ORC_SMOOTH_RBGS = 0, ORC_SMOOTH_CHEBYSHEV = 1 };
int orc_set_smoother(int kind, int sweeps, int coarse_sweeps);
/* Weight i of `sweeps` Chebyshev-weighted Jacobi sweeps (float-rounded, as the reference stores them). */
double orc_chebyshev_weight(int i, int sweeps);
/* Extension matching fs_set_relative_tolerance: tolerances become relative, the PCG stopping at
 * max|b - A p| <= tol max|b| and the projection accepted at max|div u| <= tol max|b| dt / h^2. */
void orc_set_relative_tolerance(orc_grid* g, int enabled);
/* Semi-Lagrangian transport mode (build-defined; the reference has none), then gravity and projection. */
void orc_advect_semi_lagrangian(orc_grid* g, double dt);
orc_result orc_step_semi_lagrangian(orc_grid* g, double dt, int max_iterations, double divergence_tolerance);
/* advect(dt), FluidGridCPU.cpp:362-463: CFL-split PLIC transport of fractions, component masses and
 * momentum.  Returns the number of transport substeps taken. */
int orc_advect_reference(orc_grid* g, double dt);
/* Largest transport substep for dt (FluidGridCPU.cpp:366-389). */
double orc_transport_max_step(const orc_grid* g, double dt);
/* step(dt, max_iterations, tol), FluidGridCPU.cpp:300-360, without surface tension: transport, gravity,
 * projection. */
orc_result orc_step(orc_grid* g, double dt, int max_iterations, double divergence_tolerance);
/* Staggered bilinear velocity sample at a world position (FluidGridCPU.cpp:465-471). */
void orc_sample_velocity(const orc_grid* g, double px, double py, double* vx, double* vy);

/* ---- implicit viscosity (WebGPU-only in the reference; fs_viscosity.c) ---- */
enum { ORC_VISC_OK = 0, ORC_VISC_NONFINITE = 1, ORC_VISC_BREAKDOWN = 2, ORC_VISC_NOT_CONVERGED = 3,
    ORC_VISC_BAD_SMOOTHER = 4 };
/* ViscositySolverGPU::solve (ViscositySolverGPU.cpp:169-294): material is (rho, mu) per cell; u, v are
 * updated in place only on success.  maximum < 0 means unbounded.  Returns an ORC_VISC_* status. */
int orc_viscosity_solve(int width, int height, double cell_size, const double* material, double* u, double* v,
    double dt, double tolerance, int maximum, int* iterations, double* relative_residual);
/* The same solve with chebyshev_sweeps (2, 4 or 6) Chebyshev-weighted Jacobi sweeps per side in place of the
 * V(2,2) damped Jacobi (0: the reference's smoother); no reference counterpart. */
int orc_viscosity_solve_smoothed(int width, int height, double cell_size, const double* material, double* u,
    double* v, double dt, double tolerance, int maximum, int chebyshev_sweeps, int* iterations,
    double* relative_residual);
double orc_viscosity_chebyshev_weight(int k, int n);
int orc_viscosity_levels(int width, int height);
/* Material blend of FluidGridGPU::applyViscosity (FluidGridGPU.cpp:173-218). */
void orc_viscosity_material(int cells, double cell_size, int n_liquid, const double* liquid_density,
    const double* liquid_viscosity, const double* liquid_mass, int n_gas, const double* gas_density,
    const double* gas_viscosity, const double* gas_mass, double* material);
/* Component dynamic viscosities (setComponentViscosity, FluidGridBase.h:48-52). */
void orc_configure_viscosity(orc_grid* g, const double* liquid_viscosity, const double* gas_viscosity);
/* FluidGridGPU::applyViscosity(dt) (FluidGridGPU.cpp:987-1046) with the reference tolerance 1e-4, unbounded.
 * Returns an ORC_VISC_* status; *iterations receives the dispatched iteration count. */
int orc_apply_viscosity(orc_grid* g, double dt, int* iterations);
int orc_has_viscosity(const orc_grid* g);
/* Viscosity smoother of orc_apply_viscosity (orc_viscosity_solve_smoothed's chebyshev_sweeps); returns 0, or
 * -1 for a value other than 0, 2, 4, 6. */
int orc_set_viscosity_smoother(orc_grid* g, int chebyshev_sweeps);

/* ---- surface tension (FluidGridCPU.cpp:835-926, FluidSurfaceTensionCPU.h; fs_tension.c) ---- */
int orc_set_surface_tension(orc_grid* g, double coefficient);
double orc_capillary_step_limit(const orc_grid* g);
void orc_apply_surface_tension(orc_grid* g, double dt);
/* Capillary geometry of a fraction field: heights (x, y, orientation_x, orientation_y per cell), raw
 * and completed curvatures ((value, valid) per cell).  Any output may be NULL. */
void orc_capillary_geometry(int width, int height, double cell_size, const double* fraction, double* heights,
    double* raw, double* curvature);
/* faceCurvature (FluidSurfaceTensionCPU.h:378-386) of two cells of a (value, valid) array. */
double orc_face_curvature(const double* curvature, size_t left, size_t right);
/* The quadratic fit (Fit::add + Fit::evaluate, FluidSurfaceTensionCPU.h:204-254) of n weighted points
 * (weight may be NULL: 1 each) about a normal; out = (curvature, valid). */
void orc_fit_curvature(int n, const double* px, const double* py, const double* weight, double nx, double ny,
    double cell_size, double* out);
/* segment (FluidSurfaceTensionCPU.h:287-324) of one cell: out = (centroid x, y relative to the cell
 * centre, length). */
void orc_interface_segment(double nx, double ny, double fraction, double* out);
/* interface_geometry::offset (FluidInterfaceGeometry.h:51-61). */
double orc_line_offset(double nx, double ny, double fraction);

/* Derived fields (read-only views of cached data, rebuilt on demand). */
void orc_density(orc_grid* g, double* density);
void orc_face_weights(orc_grid* g, double* inv_density_u, double* inv_density_v);
/* q = A p (FP64 operator, FluidGridCPU.cpp:1507-1598); returns p.q */
double orc_apply_operator(orc_grid* g, const double* p, double* q);
/* rhs the projection would form from the current velocities (before mean removal is applied: mean removed). */
void orc_pressure_rhs(orc_grid* g, double dt, double* rhs);

#ifdef __cplusplus
}
#endif
#endif
