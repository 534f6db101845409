// fs_device.hpp -- device-side structures and helpers shared by the HIP kernels and the host driver.
//
// Memory layout (all resident in HBM, one set per grid):
//   * Cell fields use a padded row-major layout: cell (x, y) lives at (y + 1) * stride + (x + 1) with
//     stride = W + 2 and H + 2 rows.  The one-cell halo is never written and stays zero, so every
//     5-point stencil runs without branches; wall faces carry zero weight.
//   * Face fields (u, v) keep the reference layout: u is (W+1) x H, v is W x (H+1), row-major.
//   * FP64: physical state (fractions, masses, u, v, accumulated pressure), face weights and the
//     multigrid conductances (built bit-exactly like the reference), the pressure rhs.
//   * FP32: PCG working vectors (x correction, r, z, d, q) and multigrid smoothing data.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fsim {

// Solver status codes kept on the device.
enum PcgStatus : int { kRunning = 0, kConverged = 1, kStopped = 2 };

// The FP32 recurrence residual drifts from the true one.  The FP64 true residual b - A p is checked
// (after folding the FP32 correction into the FP64 pressure) when the recurrence reaches the tolerance,
// when it has dropped kRefineDrop below the residual that started the current window (mixed-precision
// iterative refinement: FP32 CG cannot resolve much more than that on 1000:1 density contrasts), and
// every kReplaceEvery iterations.  A check that finds the true residual above the tolerance replaces r
// with it; CG restarts (beta = 0) if the check was the tolerance one or the true residual exceeds the
// recurrence by kDriftRatio, and otherwise continues on the refreshed residual (reliable updating,
// van der Vorst & Ye 2000).
constexpr int kReplaceEvery = 32;
// Solid-mask extension: a coarse pivot at or below this fraction of its row's assembled diagonal marks
// the last unknown of a disconnected or fully blocked coarse region, which is then pinned at zero (the
// reference pins its first cell the same way).  Grids without a mask never pin: every pivot of their
// weighted Laplacians lies far above the bound, and a non-positive one fails as in the reference.
constexpr double kPinRatio = 1e-9;
constexpr double kRefineDrop = 1e-3;
constexpr double kDriftRatio = 2.0;

// Device-resident PCG scalar block.  Kernels of one iteration communicate only through this block,
// so the host never reads a scalar inside the loop (it polls `status` once per batch).
struct PcgScalars {
    double rz;          // r . z of the current direction (reference: residual_dot_correction)
    double alpha;
    double beta;
    double r_mean;      // mean of the residual handed to the preconditioner
    double z_mean;      // mean of the preconditioned residual (removed before use)
    double tol;         // residual tolerance, max-norm, in rhs units
    double rmax;        // last recurrence residual max-norm
    double true_rmax;   // last FP64 true residual max-norm
    double bmax;        // max |rhs| after mean removal
    double b_mean;      // rhs mean removed from the divergence
    double p_mean;      // pressure mean (removed)
    double dq;          // last d . A d
    double div_max;     // last max |div u| (1/s)
    double r_sum;       // last residual sum
    double initial_rmax;  // max |b - A p| of the predicted pressure
    int iterations;
    int max_iterations;
    int status;
    int verify;         // recurrence says converged: check the true residual this iteration
    int restart;        // next beta is zero (true residual replaced the recurrence)
    int restarts;
    int mg_failed;      // coarse factorisation failed (the hierarchy must be rebuilt next time)
    int replace;        // check not triggered by the tolerance: refresh r from FP64, restart only on drift
    int pad;
    double window_rmax; // true residual max-norm that started the current refinement window
    // r.z of the direction used by the iteration that starts at iteration count c, kept in slot c & 1:
    // the direction kernel reads the previous slot and writes the next one while other workgroups of
    // the same launch still read the previous value.
    double rz_slot[2];
    // > 0: the tolerance is relative, tol = rel_tol * bmax once the rhs is formed (max-norm relative residual)
    double rel_tol;
};

// 16-bit storage formats of multigrid levels (arithmetic stays FP32): IEEE binary16 and bfloat16
// (the upper half of an FP32 word).  Plain structs, so host code can name them without device headers.
struct half16 { unsigned short bits; };
struct bf16 { unsigned short bits; };

// One multigrid level as the smoothing kernels see it: conductances stored as C, the level's rhs and
// solution vectors as V (FP32 everywhere by default).
template <class C, class V> struct LevelViewT {
    int w, h, s;             // interior width/height, padded stride (w + 2)
    const C* right;          // conductance of the face right of each cell (padded)
    const C* down;           // conductance of the face below each cell (padded)
    const float* invd;       // 1 / diagonal (padded, FP32; unused by the 16-bit kernels)
    V* rhs;                  // padded
    V* sol[2];               // ping-pong solution buffers (padded)
    V* res;                  // residual scratch (padded)
};
using LevelView = LevelViewT<float, float>;

// Per-axis prolongation map of a fine level (first/second coarse index, weight of second).
struct AxisProlong {
    const int* first;
    const int* second;
    const float* weight;
};

// Per-axis restriction stencil of a coarse index (<= 4 fine taps, exact transpose of prolongation).
struct AxisRestrict {
    const int* count;   // [n]
    const int* index;   // [n * 4]
    const float* weight;  // [n * 4]
};

__device__ __forceinline__ int pidx(int x, int y, int s) { return (y + 1) * s + (x + 1); }

// PressureMultigrid::coarsen (PressureMultigrid.cpp:142-202) for coarse cell (x, y): parallel paths add,
// series resistances combine across the two or three fine faces between coarse centres.  Writes the
// coarse right face (x + 1 < cw) and down face (y + 1 < ch) of the cell.
__device__ __forceinline__ void coarsen_cell(int x, int y, int fw, int fh, const double* fcx, const double* fcy,
    const double* fright, const double* fdown, int cw, int ch, const double* ccx, const double* ccy,
    double* cright, double* cdown)
{
    const int fs = fw + 2, cs = cw + 2;
    if (x + 1 < cw) {
        const int edge = 2 * x + 2;
        const double left = (fcx[edge - 1] - ccx[x]) / (fcx[edge - 1] - fcx[edge - 2]);
        const double right = edge + 1 < fw ? (ccx[x + 1] - fcx[edge]) / (fcx[edge + 1] - fcx[edge]) : 0.0;
        double conductance = 0.0;
        const int fy_end = min(2 * y + 2, fh);
        for (int fy = 2 * y; fy < fy_end; ++fy) {
            const int i = (fy + 1) * fs + edge;
            double resistance = left / fright[i - 1] + 1.0 / fright[i];
            if (right != 0.0)
                resistance += right / fright[i + 1];
            conductance += 1.0 / resistance;
        }
        cright[(y + 1) * cs + x + 1] = conductance;
    }
    if (y + 1 < ch) {
        const int edge = 2 * y + 2;
        const double up = (fcy[edge - 1] - ccy[y]) / (fcy[edge - 1] - fcy[edge - 2]);
        const double down = edge + 1 < fh ? (ccy[y + 1] - fcy[edge]) / (fcy[edge + 1] - fcy[edge]) : 0.0;
        double conductance = 0.0;
        const int fx_end = min(2 * x + 2, fw);
        for (int fx = 2 * x; fx < fx_end; ++fx) {
            const int i = edge * fs + fx + 1;
            double resistance = up / fdown[i - fs] + 1.0 / fdown[i];
            if (down != 0.0)
                resistance += down / fdown[i + fs];
            conductance += 1.0 / resistance;
        }
        cdown[(y + 1) * cs + x + 1] = conductance;
    }
}

// Level::rebuildDiagonal (PressureMultigrid.cpp:104-117) of one cell from its four FP64 face
// conductances, and the FP32 copies the smoothers read.
__device__ __forceinline__ void level_diag_values(double rl, double rr, double du, double dd, float& invd32,
    float& right32, float& down32)
{
    const double diagonal = rl + rr + du + dd;
    invd32 = (float)(diagonal > 0.0 ? 1.0 / diagonal : 0.0);
    right32 = (float)rr;
    down32 = (float)dd;
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

// Block-wide sum in a fixed order; result valid in thread 0.  `lds` holds blockDim/64 doubles.
template <int NT> __device__ __forceinline__ double block_sum(double v, double* lds)
{
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0)
        lds[wid] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NT / 64; ++i)
            r += lds[i];
    }
    __syncthreads();
    return r;
}

template <int NT> __device__ __forceinline__ double block_max(double v, double* lds)
{
    v = wave_max(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0)
        lds[wid] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NT / 64; ++i)
            r = fmax(r, lds[i]);
    }
    __syncthreads();
    return r;
}

// Reads a partial written by another block of the previous launch.
__device__ __forceinline__ double load_partial(const double* p, int i) { return p[i]; }

// Sum of `n` partials with stride `k` starting at `offset`, in a fixed order; every thread returns it.
// Loads run eight (then four) strides ahead of the additions (independent loads in flight instead of one memory
// latency per partial); the additions keep the one-stride-at-a-time order, so the sum is unchanged.
template <int NT>
__device__ __forceinline__ double finish_sum(const double* partials, int n, int k, int offset, double* lds)
{
    double v = 0.0;
    int i = threadIdx.x;
    for (; i + 7 * NT < n; i += 8 * NT) {
        double t[8];
#pragma unroll
        for (int u = 0; u < 8; ++u)
            t[u] = load_partial(partials, (i + u * NT) * k + offset);
#pragma unroll
        for (int u = 0; u < 8; ++u)
            v += t[u];
    }
    for (; i + 3 * NT < n; i += 4 * NT) {
        double t[4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            t[u] = load_partial(partials, (i + u * NT) * k + offset);
#pragma unroll
        for (int u = 0; u < 4; ++u)
            v += t[u];
    }
    for (; i < n; i += NT)
        v += load_partial(partials, i * k + offset);
    double r = block_sum<NT>(v, lds);
    __shared__ double out;
    if (threadIdx.x == 0)
        out = r;
    __syncthreads();
    r = out;
    __syncthreads();
    return r;
}

template <int NT>
__device__ __forceinline__ double finish_max(const double* partials, int n, int k, int offset, double* lds)
{
    double v = 0.0;
    int i = threadIdx.x;
    for (; i + 7 * NT < n; i += 8 * NT) {
        double t[8];
#pragma unroll
        for (int u = 0; u < 8; ++u)
            t[u] = load_partial(partials, (i + u * NT) * k + offset);
#pragma unroll
        for (int u = 0; u < 8; ++u)
            v = fmax(v, t[u]);
    }
    for (; i + 3 * NT < n; i += 4 * NT) {
        double t[4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            t[u] = load_partial(partials, (i + u * NT) * k + offset);
#pragma unroll
        for (int u = 0; u < 4; ++u)
            v = fmax(v, t[u]);
    }
    for (; i < n; i += NT)
        v = fmax(v, load_partial(partials, i * k + offset));
    double r = block_max<NT>(v, lds);
    __shared__ double out;
    if (threadIdx.x == 0)
        out = r;
    __syncthreads();
    r = out;
    __syncthreads();
    return r;
}

}  // namespace fsim
