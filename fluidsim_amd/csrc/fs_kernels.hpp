// fs_kernels.hpp -- launchers for the HIP kernels (fs_kernels.hip).  Host code only calls these.
#pragma once
#include "fs_device.hpp"

namespace fsim {

// Everything one grid's kernels need, by value (device pointers).
struct GridDev {
    int W, H, s;              // cells, padded stride (W + 2)
    double h;                 // cell size
    int nl, ng;               // liquid / gas component counts
    double* fraction;         // N   (reference layout)
    double* u;                // (W+1) * H
    double* v;                // W * (H+1)
    double* lmass;            // nl * N
    double* gmass;            // ng * N
    double* density;          // N
    double* p;                // padded
    double* p_prev;           // padded
    double* b;                // padded rhs (FP64)
    double* right64;          // padded level-0 conductances (FP64, bit-exact face weights)
    double* down64;
    float* right32;           // padded level-0 conductances (FP32)
    float* down32;
    float* invd32;            // padded 1/diagonal (FP32), level 0 and the diagonal preconditioner
    float* x;                 // padded FP32 correction accumulated since the last flush
    float* r;                 // padded FP32 residual
    float* d;                 // padded FP32 direction
    float* q;                 // padded FP32 A d
    float* zdiag;             // padded FP32 diagonal-preconditioned residual
    PcgScalars* S;            // device scalar block
    double* partials;         // reduction partials (>= 4 * max blocks)
    double* partials_ad;      // d.q partials of the operator kernel (one per block)
    // Where the update and direction kernels read the d.q and r.z partials, and how many blocks' worth:
    // the producers' own per-block partials (every consumer workgroup reduces them: cheap while the
    // grid has few blocks), or, on large grids, one pre-reduced entry written by k_fold_partials (the
    // same fixed-order sum, so the result is identical).
    double* fold_ad;
    int fold_ad_n;
    double* fold_rz;
    int fold_rz_n;
    const unsigned char* solid;   // W * H, nonzero = solid cell (solid-mask extension); null: no mask
};

// Status values beyond PcgStatus used around the loop.
enum : int { kZeroRhs = 3, kInitialOk = 4, kFailed = 5 };

// Prolongation source of the fused post-smoothing kernel (csol == nullptr: no prolongation); the coarse
// solution is stored as V.
template <class V> struct CoarseLinkT {
    const V* csol;       // coarse solution (padded)
    int cs;              // coarse padded stride
    const int* pxf; const int* pxs; const float* pxw;
    const int* pyf; const int* pys; const float* pyw;
};
using CoarseLink = CoarseLinkT<float>;

// Tail of the hierarchy solved by one workgroup entirely in LDS.
struct TailLevel {
    int w, h, s;
    int cw, ch;          // next coarser tail level (0 on the last)
    const float* right;  // global padded conductances
    const float* down;
    const int* pxf; const int* pxs; const float* pxw;   // transfers to the next tail level
    const int* pyf; const int* pys; const float* pyw;
    const int* rxc; const int* rxi; const float* rxw;
    const int* ryc; const int* ryi; const float* ryw;
    int lds_arrays;      // word offset of right, down, rhs, sol, res (each s * (h + 2))
    int lds_tables;      // word offset of this level's transfer tables (3w + 3h + 9cw + 9ch words)
};

constexpr int kMaxTailLevels = 8;
struct TailDesc {
    int levels;
    int coarse_n;              // unknowns of the pinned coarse system (w*h - 1 of the last level)
    const float* ginv;         // inverse of the pinned coarse matrix, FP32 n*n
    // global rhs and solution of the first tail level (the FP32 residual and z when the tail starts at
    // level 0); stored as FP32, or as bfloat16 when the tail kernel runs with a bfloat16 top level
    const void* top_rhs;
    void* top_sol;
    int lds_ginv;              // word offset of ginv in LDS
    const unsigned* tables;    // every level's transfer tables, in their LDS layout (host-built)
    int table_words;
    int lds_tables_base;       // word offset of the table block in LDS
    int lds_words;             // total LDS words
    int sweeps;                // red-black sweeps per side on every tail level (4: the reference's V(4,4))
    TailLevel lev[kMaxTailLevels];
};

// Fills the LDS offsets of `t` and returns its LDS bytes.
size_t tail_layout(TailDesc& t);
// Raises the dynamic-LDS limit of the tail kernel; false if the device refuses.
bool tail_allow_lds(size_t bytes);

// rel_tol > 0 replaces tol by rel_tol * max|b| once the rhs b is formed (a relative residual).
void launch_init_scalars(hipStream_t st, PcgScalars* S, double tol, int max_iterations, double rel_tol = 0.0);
// Copies a coarse-factorisation failure flagged in `from` into the solve's scalar block.
void launch_merge_failure(hipStream_t st, const PcgScalars* from, PcgScalars* S);

// ---- fields, coefficients, hierarchy (FP64) ----
void launch_density(hipStream_t st, const GridDev& g);
void launch_face_weights(hipStream_t st, const GridDev& g);
void launch_level_diag(hipStream_t st, int w, int h, const double* right64, const double* down64,
    float* right32, float* down32, float* invd32);
void launch_coarsen(hipStream_t st, int fw, int fh, const double* fcx, const double* fcy,
    const double* fright, const double* fdown, int cw, int ch, const double* ccx, const double* ccy,
    double* cright, double* cdown);
// Bit-exact Cholesky of the pinned coarse matrix (factorCoarsest), plus its FP32 inverse when ginv is set.
// pin != 0 (grids with a solid mask): a vanishing pivot pins its unknown at zero instead of failing.
void launch_coarse_factor(hipStream_t st, int w, int h, const double* right64, const double* down64,
    double* factor, float* ginv, int n, PcgScalars* S, int pin);
// FP32 inverse of the pinned coarse matrix by the in-register sweep operator (the projection's path).
void launch_coarse_inverse(hipStream_t st, int w, int h, const double* right64, const double* down64, float* ginv,
    int n, PcgScalars* S, int pin);
// The coarse end of the hierarchy rebuild in one single-workgroup launch: for each listed level, the
// coarsening from its finer level and its diagonal (the values of launch_coarsen + launch_level_diag),
// then the inverse of the last level's pinned matrix (the values of launch_coarse_inverse).
struct ChainLevel {
    int fw, fh, cw, ch;
    const double *fcx, *fcy, *fright, *fdown;   // finer level: centres and FP64 conductances
    const double *ccx, *ccy;                    // this level's centres
    double *cright, *cdown;                     // this level's FP64 conductances (written)
    float *right32, *down32, *invd32;           // and their FP32 copies and 1/diagonal (written)
};
struct ChainDesc {
    int levels;
    ChainLevel lev[kMaxTailLevels];
};
void launch_coarse_chain(hipStream_t st, const ChainDesc& chain, float* ginv, int n, PcgScalars* S, int pin);

// ---- projection set-up and finish (FP64) ----
void launch_rhs(hipStream_t st, const GridDev& g, double scale);
void launch_rhs_center(hipStream_t st, const GridDev& g);
void launch_zero_pressure(hipStream_t st, const GridDev& g);
void launch_predict(hipStream_t st, const GridDev& g);
void launch_sub_mean_p(hipStream_t st, const GridDev& g, int only_if_loop);
void launch_true_residual(hipStream_t st, const GridDev& g, int mode);  // 0 initial, 1 verify
void launch_flush(hipStream_t st, const GridDev& g);                     // only when S->verify
void launch_final_flush_sum(hipStream_t st, const GridDev& g);
void launch_correct_velocity(hipStream_t st, const GridDev& g, double dt);
// when_stopped: measure only once the grid's solve has stopped (otherwise a no-op).
void launch_max_divergence(hipStream_t st, const GridDev& g, bool when_stopped = false);
void launch_gravity(hipStream_t st, const GridDev& g, double gy, double dt);
void launch_add_vortex(hipStream_t st, const GridDev& g, double ox, double oy, double cx, double cy, double radius,
    double strength);
void launch_semi_lagrangian(hipStream_t st, const GridDev& g, double ox, double oy, double dt, double* new_u,
    double* new_v, double* scratch_cells);
// Up to eight FP64 device-to-device copies in one launch.
struct CopyList {
    struct Item {
        const double* src;
        double* dst;
        size_t n;
    };
    int count;
    Item item[8];
};
void launch_copy_many(hipStream_t st, const CopyList& list);

void launch_pad_copy(hipStream_t st, const double* src, double* dst, int W, int H, int to_padded);

// ---- rendering and material edits (FluidGridBase.h:139-142) ----
constexpr int kMaxColorComponents = 128;   // liquid + gas components of one grid
// Colour and density of each component, liquid components first.
struct ColorTable {
    float rgb[kMaxColorComponents][3];
    double density[kMaxColorComponents];
};
// writeColorPixels into a device RGBA buffer (4 floats per cell).
void launch_color(hipStream_t st, const GridDev& g, const ColorTable& t, int alpha_from_fraction, float* rgba);
// One resolved material edit (detail::ResolvedCellEdit, FluidGridEdit.h:22-28).
struct EditCell {
    int x, y;
    int component;     // liquid components first, then gas
    int pad;
    double mass;
};
// commitEdits on duplicate-free cells (each cell at most once).
void launch_commit_edits(hipStream_t st, const GridDev& g, const EditCell* cells, int n);

// ---- PCG loop (FP32 vectors, FP64 reductions) ----
void launch_pcg_ad(hipStream_t st, const GridDev& g);
void launch_pcg_update(hipStream_t st, const GridDev& g);
// flexible: Polak-Ribiere beta from z.q (for a V-cycle with bfloat16 vectors, not exactly linear); the
// two kernels of one iteration must agree on it (it changes the partials' layout).
void launch_pcg_rz(hipStream_t st, const GridDev& g, const float* z, int flexible = 0);
void launch_pcg_direction(hipStream_t st, const GridDev& g, const float* z, int flexible = 0);
void launch_precond_diag(hipStream_t st, const GridDev& g);

// ---- multigrid V-cycle pieces (FP32) ----
// nu (1..4, default 4) red-black sweeps in one launch (reads sol[cur], writes sol[cur ^ 1]); with link.csol
// set the coarse correction is prolonged and added first.  fine selects the finest-level symbol.
// variant: interior tile and threads 0 = 64 x 32 / 256, 1 = 32 x 32 / 256, 2 = 32 x 16 / 256,
// 3 = 16 x 16 / 256, 4 = 64 x 32 / 512, 5 = 32 x 32 / 512, 6 = 32 x 32 / 384, 7 = 64 x 32 / 640,
// 8 = 32 x 64 / 640, 9 = 64 x 32 / 960, 10 = 32 x 64 / 960 (same results, different speed).
constexpr int kSmoothVariants = 11;
void launch_rb_multi(hipStream_t st, const LevelView& L, int cur, const float* rhs, const double* shift,
    int first_color, int from_zero, const CoarseLink& link, const int* status, int fine, int variant, int nu = 4);
// Chebyshev-weighted damped Jacobi (the reference WebGPU multigrid's polynomial smoother,
// PressureMultigridGPU.cpp:24-28, 619-631; PressureSmoothingKernels.h:91-176): nu (1..kJacobiMaxSweeps)
// sweeps x <- x + w_k (b - A x) / diag in one launch on an LDS tile with an nu-cell halo, reading sol[cur]
// (from_zero: x = 0) and writing sol[cur ^ 1]; with link.csol set the coarse correction is prolonged and
// added first.  A cell without conductance is set to zero.  weights.w[k] is sweep k's weight (the caller
// reverses them for post-smoothing).  variant: 0 = 64 x 32 / 512, 1 = 32 x 32 / 256, 2 = 32 x 16 / 256,
// 3 = 16 x 16 / 256, 4 = 32 x 32 / 512, 5 = 64 x 64 / 1024, 6 = 64 x 32 / 1024, 7 = 32 x 64 / 1024 (same
// results, different speed).
constexpr int kJacobiMaxSweeps = 8;
constexpr int kJacobiShortHalo = 5;   // tile halo of the kernels used for up to this many sweeps
constexpr int kJacobiVariants = 8;
struct SmoothWeights {
    float w[kJacobiMaxSweeps];
};
void launch_jacobi_multi(hipStream_t st, const LevelView& L, int cur, const float* rhs, const double* shift,
    int from_zero, const SmoothWeights& weights, int nu, const CoarseLink& link, const int* status, int variant);
// Pre-smoothing from zero (launch_jacobi_multi with from_zero) and the residual + restriction into C
// (launch_res_restrict) in one launch, same values; tile variants 0 = 16x8 coarse cells on 512 threads,
// 1 = 16x16/512, 2 = 32x16/1024.  C must cover the fine level (2 C.w >= L.w, 2 C.h >= L.h).
constexpr int kJacobiRestrictVariants = 3;
void launch_jacobi_res_restrict(hipStream_t st, const LevelView& L, int cur, const float* rhs, const double* shift,
    const SmoothWeights& weights, int nu, const LevelView& C, AxisRestrict rx, AxisRestrict ry, const int* status,
    int variant);
// Residual of sol[cur] and its restriction into C.rhs in one launch; variant: coarse tile 0 = 32 x 16,
// 1 = 16 x 16, 2 = 16 x 8, 3 = 8 x 8.
constexpr int kRestrictVariants = 4;
void launch_res_restrict(hipStream_t st, const LevelView& F, int cur, const float* rhs, const double* shift,
    const LevelView& C, AxisRestrict rx, AxisRestrict ry, const int* status, int variant);
// bf16_top: the first tail level's rhs and solution are stored as bfloat16 (else FP32).
void launch_mg_tail(hipStream_t st, const TailDesc& desc, size_t lds_bytes, const double* shift, const int* status,
    int bf16_top = 0);

// ---- 16-bit multigrid levels: conductances as C (half16 or bf16), vectors as V ----
// Conductances of one level rounded (FP64 -> FP32 -> C, each to nearest even) into 16-bit copies.
template <class C> void launch_level_pack(hipStream_t st, int w, int h, const double* right64, const double* down64,
    C* right, C* down);
// The fused smoother and residual-restriction kernels of launch_rb_multi / launch_res_restrict with
// C conductances, V vectors on this level (its rhs argument included) and VC vectors on the next
// coarser level; MODE 0 and 1 only.  Instantiated for (half16, float, float), (bf16, float, float),
// (bf16, float, bf16) and (bf16, bf16, bf16).
template <class C, class V, class VC>
void launch_rb_multi_t(hipStream_t st, const LevelViewT<C, V>& L, int cur, const V* rhs, const double* shift,
    int first_color, int from_zero, const CoarseLinkT<VC>& link, const int* status, int variant, int nu = 4);
template <class C, class V, class VC>
void launch_res_restrict_t(hipStream_t st, const LevelViewT<C, V>& F, int cur, const V* rhs, const double* shift,
    const LevelViewT<C, VC>& Cl, AxisRestrict rx, AxisRestrict ry, const int* status, int variant);

// ---- batched PCG: one 1024-thread workgroup runs the whole iteration loop of one instance ----
constexpr int kMaxBatchLevels = 16;   // tiled levels above an instance's tail
struct BatchLevel {
    LevelView L;        // tiled level l (rhs: r on level 0)
    LevelView C;        // level l + 1 (its rhs receives the restriction, its sol[0] the coarse correction)
    AxisRestrict rx, ry;  // restriction l -> l + 1
    CoarseLink link;    // prolongation l + 1 -> l
    int sweeps;         // red-black sweeps per side on this level (1..4)
};
struct BatchInstance {
    GridDev g;
    int multigrid;      // 0: diagonal preconditioner
    int levels;         // tiled levels (the tail starts at this level)
    const TailDesc* tail;   // device copy of the instance's tail descriptor
    BatchLevel lev[kMaxBatchLevels];
};
// Runs every instance's PCG iterations (the captured iteration body, looped on the device until the
// instance's scalar block leaves kRunning) in one launch of `count` workgroups.
void launch_batch_pcg(hipStream_t st, const BatchInstance* batch, int count, size_t lds_bytes);
bool batch_allow_lds(size_t bytes);

// ---- reference PLIC/VOF transport (FP64, reference layouts, bit-exact) ----
// One directional pass reads the frozen flow (u, v), the fractions and masses of the pass start and the
// transported velocities nu_in/nv_in; it writes new fractions/masses/density into the *_out buffers, the
// recovered velocities into nu_out/nv_out and updates the momenta in place.  The host swaps the pairs.
struct TransportDev {
    int W, H, nl, ng;
    double icv;                            // 1 / h^2
    const double* u; const double* v;      // flow, frozen for the whole transport call
    const double* f; double* f_out;        // liquid fraction N
    const double* lm; double* lm_out;      // liquid masses nl * N
    const double* gm; double* gm_out;      // gas masses ng * N
    double* density;                       // N, rebuilt from the new masses
    const double* nu_in; double* nu_out;   // transported velocities (W+1)*H
    const double* nv_in; double* nv_out;   // W*(H+1)
    double* mom_u; double* mom_v;          // momenta
    double* lt; double* mt;                // liquid volume and total mass transfer of the pass's faces
    double* lcw; double* gcw; double* mcw; // frozen correction weights (nl*N, ng*N, N)
    unsigned char* split;                  // frozen phase selection (fraction > 0.5)
    double* partials;                      // >= one double per 64 x 16 cell tile
    double* out;                           // one device double: max outflow
};
// Density, momenta and transported velocities from u/v (rebuildMomentum), then the largest cell
// outflow into *out (FluidGridCPU.cpp:1425-1476, 366-386).
void launch_transport_begin(hipStream_t st, const TransportDev& t);
// freezeComponentCorrections (FluidGridCPU.cpp:1307-1349).
void launch_transport_freeze(hipStream_t st, const TransportDev& t);
// One X (horizontal != 0) or Y pass of one substep: interfaces + liquid and component transfers,
// component/fraction update with density, momentum transfer and velocity recovery
// (FluidGridCPU.cpp:401-456).
void launch_transport_pass(hipStream_t st, const TransportDev& t, int horizontal, double scale);

// ---- implicit viscosity (WebGPU-only in the reference) ----
constexpr int kMaxComponents = 64;   // per phase
// Component densities and dynamic viscosities: liquids at [0, nl), gases at [kMaxComponents, +ng).
struct VisMixture {
    double density[2 * kMaxComponents];
    double viscosity[2 * kMaxComponents];
};
// Material blend of FluidGridGPU::applyViscosity into (rho, mu) per cell, FP64.
void launch_viscosity_blend(hipStream_t st, int N, double cell_size, int nl, int ng, const double* lm,
    const double* gm, const VisMixture& mix, double* material);
struct ViscositySolver;   // device workspace of one grid size
ViscositySolver* viscosity_create(int width, int height);
// Allocation-free solve (fs_viscosity_solve_noalloc): the workspace is a file-scope device array of
// kViscosityScratchBytes per device.  viscosity_scratch_setup (with `device` current) resolves that
// device's copy once; viscosity_scratch_bytes is the workspace a W x H solve needs;
// viscosity_solve_scratch runs viscosity_solve's algorithm in it on device u, v and material, one call
// at a time per device, and returns after the stream has drained.
constexpr size_t kViscosityScratchBytes = (size_t)64 << 20;
void* viscosity_scratch_setup(int device);
size_t viscosity_scratch_bytes(int width, int height);
int viscosity_solve_scratch(int device, hipStream_t st, int width, int height, double* u, double* v,
    const double* material, double dt, double spacing, double tol, int maximum, int* iterations,
    double* relative_residual);
void viscosity_destroy(ViscositySolver* s);
size_t viscosity_bytes(const ViscositySolver* s);
int viscosity_levels(const ViscositySolver* s);
// 0: the reference's V(2,2) damped Jacobi; 2, 4, 6: Chebyshev-weighted sweeps per side (throws otherwise)
void viscosity_set_smoother(ViscositySolver* s, int chebyshev_sweeps);
double* viscosity_material_buffer(ViscositySolver* s);   // 2 * W * H doubles
// Backward-Euler viscosity on device FP64 u ((W+1) x H) and v (W x (H+1)) with a device FP64 (rho, mu)
// material; maximum < 0 is unbounded.  Returns 0 converged (or nothing to do), 1 non-finite input,
// 2 breakdown, 3 not converged within `maximum`; u, v change only on 0.
int viscosity_solve(ViscositySolver* s, hipStream_t st, double* u, double* v, const double* material, double dt,
    double spacing, double tol, int maximum, int* iterations, double* relative_residual);
// Batched viscosity solves: one 1024-thread workgroup per instance runs its whole CG loop.
// viscosity_batch_prepare launches an instance's set-up on `st` and writes its descriptor into `desc`
// (viscosity_batch_desc_bytes() of host memory; false: too many levels, solve it on its own); after
// launch_batch_viscosity over the descriptors (copied to device memory), viscosity_batch_fetch enqueues
// the read-back and, after a synchronisation, viscosity_batch_result returns viscosity_solve's codes or
// 4 when the instance needs more iterations than the batch runs (u, v untouched: solve it on its own).
size_t viscosity_batch_desc_bytes();
bool viscosity_batch_prepare(ViscositySolver* s, hipStream_t st, double* u, double* v, const double* material,
    double dt, double spacing, double tol, int maximum, void* desc);
// reserve_lds: dynamic LDS to reserve per workgroup (0: none), so that instances spread one per compute unit.
void launch_batch_viscosity(hipStream_t st, const void* descs, int count, size_t reserve_lds);
void viscosity_batch_fetch(ViscositySolver* s, hipStream_t st);
int viscosity_batch_result(const ViscositySolver* s, int maximum, int* iterations, double* relative_residual);

// ---- surface tension (FluidGridCPU::applySurfaceTension, FluidSurfaceTensionCPU) ----
// Heights (x, y) and orientations per cell, raw and completed curvature (value, valid) per cell.
void launch_capillary_geometry(hipStream_t st, int W, int H, double cell_size, const double* fraction,
    double2* heights, double2* orientations, double2* raw, double2* curvature);
// u, v += dt w sigma kappa_face (f_R - f_L) / h on interior faces; w = level-0 right64 / down64.
void launch_capillary_force(hipStream_t st, int W, int H, double cell_size, const double* fraction,
    const double2* curvature, const double* right64, const double* down64, double sigma, double dt, double* u,
    double* v);

// ---- bench helper: one operator application without reductions (for event timing) ----
void launch_copy_f32(hipStream_t st, const float* src, float* dst, size_t n);

}  // namespace fsim
