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
    unsigned* counter;        // last-block arrival counter (kept at zero between kernels)
};

// Status values beyond PcgStatus used around the loop.
enum : int { kZeroRhs = 3, kInitialOk = 4, kFailed = 5 };

// Tail of the hierarchy solved by one workgroup in LDS.
struct TailLevel {
    int w, h, s;
    const float* right;
    const float* down;
    const float* invd;
    // transfers to the next coarser tail level (unused on the last level)
    const int* px_first; const int* px_second; const float* px_w;
    const int* py_first; const int* py_second; const float* py_w;
    const int* rx_count; const int* rx_index; const float* rx_w;
    const int* ry_count; const int* ry_index; const float* ry_w;
};

constexpr int kMaxTailLevels = 8;
struct TailDesc {
    int levels;
    int coarse_n;                 // unknowns of the dense coarse solve (w*h - 1 of the last level)
    const double* factor;         // coarse Cholesky factor (n*n)
    const float* top_rhs;         // global rhs of the first tail level (the FP32 residual on level 0)
    float* top_sol[2];            // global ping-pong solution of the first tail level
    int lds_offset[kMaxTailLevels];  // float offset of each level's six LDS arrays
    TailLevel lev[kMaxTailLevels];
};

size_t tail_lds_bytes(const TailDesc& t);

void launch_init_scalars(hipStream_t st, PcgScalars* S, double tol, int max_iterations);

// ---- fields, coefficients, hierarchy (FP64) ----
void launch_density(hipStream_t st, const GridDev& g);
void launch_face_weights(hipStream_t st, const GridDev& g);
void launch_level_diag(hipStream_t st, int w, int h, const double* right64, const double* down64,
    float* right32, float* down32, float* invd32);
void launch_coarsen(hipStream_t st, int fw, int fh, const double* fcx, const double* fcy,
    const double* fright, const double* fdown, int cw, int ch, const double* ccx, const double* ccy,
    double* cright, double* cdown);
void launch_cholesky(hipStream_t st, int w, int h, const double* right64, const double* down64, double* factor,
    int n, PcgScalars* S);

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
void launch_max_divergence(hipStream_t st, const GridDev& g);
void launch_gravity(hipStream_t st, const GridDev& g, double gy, double dt);
void launch_add_vortex(hipStream_t st, const GridDev& g, double ox, double oy, double cx, double cy, double radius,
    double strength);
void launch_semi_lagrangian(hipStream_t st, const GridDev& g, double ox, double oy, double dt, double* new_u,
    double* new_v, double* scratch_cells);
void launch_pad_copy(hipStream_t st, const double* src, double* dst, int W, int H, int to_padded);

// ---- PCG loop (FP32 vectors, FP64 reductions) ----
void launch_pcg_ad(hipStream_t st, const GridDev& g);
void launch_pcg_update(hipStream_t st, const GridDev& g);
void launch_pcg_rz(hipStream_t st, const GridDev& g, const float* z);
void launch_pcg_direction(hipStream_t st, const GridDev& g, const float* z);
void launch_precond_diag(hipStream_t st, const GridDev& g);

// ---- multigrid V-cycle pieces (FP32) ----
void launch_rb_sweep(hipStream_t st, const LevelView& L, int cur, const float* rhs, const double* shift,
    int first_color, int from_zero, const int* status);
void launch_mg_residual(hipStream_t st, const LevelView& L, int cur, const float* rhs, const double* shift,
    const int* status);
void launch_mg_restrict(hipStream_t st, const LevelView& F, const LevelView& Cl, AxisRestrict rx, AxisRestrict ry,
    const int* status);
void launch_mg_prolong(hipStream_t st, const LevelView& F, int fcur, const LevelView& Cl, int ccur, AxisProlong px,
    AxisProlong py, const int* status);
void launch_mg_tail(hipStream_t st, const TailDesc* dev_desc, const TailDesc& host_desc, int cur,
    const double* shift, const int* status);

// ---- bench helper: one operator application without reductions (for event timing) ----
void launch_copy_f32(hipStream_t st, const float* src, float* dst, size_t n);

}  // namespace fsim
