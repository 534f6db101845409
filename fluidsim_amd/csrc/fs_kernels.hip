// fs_kernels.hip -- hand-written HIP kernels for gfx950 (MI355X) on the uniform-MAC projection path.
//
// Compiled with -ffp-contract=off: the FP64 kernels that rebuild face weights, multigrid
// conductances and the coarse Cholesky factor evaluate every expression in the same order as the
// reference UniformCPU (FluidGridCPU.cpp, PressureMultigrid.cpp), so their results are bit-identical
// to it.  The PCG working vectors and the multigrid smoothing run in FP32 with FP64 reductions.
//
// Cell kernels use 64 x 16 tiles and 256 threads (four 64-wide wavefronts, one row segment each),
// so every wavefront touches 64 consecutive cells of a row: loads and stores coalesce.
// Reductions write one partial per block; the block that arrives last finishes the reduction in a
// fixed order and updates the device scalar block, so no host round trip sits inside the loop.
#include "fs_kernels.hpp"

#include <math.h>

namespace fsim {

namespace {

constexpr int kTX = 64;
constexpr int kTY = 16;
constexpr int kNT = 256;

inline dim3 cell_grid(int W, int H) { return dim3((W + kTX - 1) / kTX, (H + kTY - 1) / kTY); }
inline unsigned cell_blocks(int W, int H) { return ((W + kTX - 1) / kTX) * ((H + kTY - 1) / kTY); }

__device__ __forceinline__ unsigned block_id() { return blockIdx.y * gridDim.x + blockIdx.x; }
__device__ __forceinline__ unsigned block_count() { return gridDim.x * gridDim.y; }

__device__ __forceinline__ double abs_or_inf(double v) { return isfinite(v) ? fabs(v) : INFINITY; }

// Iterate the (up to) four cells this thread owns in the block's 64 x 16 tile.
#define FS_TILE_BEGIN(WIDTH, HEIGHT)                                                   \
    {                                                                                  \
        const int x = blockIdx.x * kTX + (threadIdx.x & 63);                           \
        for (int k_ = 0; k_ < 4; ++k_) {                                               \
            const int y = blockIdx.y * kTY + (threadIdx.x >> 6) + 4 * k_;              \
            if (x >= (WIDTH) || y >= (HEIGHT))                                         \
                continue;
#define FS_TILE_END \
        }           \
    }

// ------------------------------------------------------------------------------------------
// Scalar block initialisation (a kernel, so it is safe inside captured graphs).
// ------------------------------------------------------------------------------------------
__global__ void k_init_scalars(PcgScalars* S, double tol, int max_iterations)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        PcgScalars z{};
        z.tol = tol;
        z.max_iterations = max_iterations;
        z.status = kRunning;
        z.restart = 1;
        *S = z;
    }
}

// ------------------------------------------------------------------------------------------
// Density and face weights (FluidGridCPU.cpp:1399-1423, :1154-1205), FP64, bit-exact.
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(kNT) void k_density(GridDev g)
{
    const size_t N = (size_t)g.W * g.H;
    const double inverse_cell_volume = 1.0 / (g.h * g.h);
    for (size_t i = (size_t)blockIdx.x * kNT + threadIdx.x; i < N; i += (size_t)gridDim.x * kNT) {
        double total = 0.0;
        for (int k = 0; k < g.nl; ++k)
            total += g.lmass[(size_t)k * N + i];
        for (int k = 0; k < g.ng; ++k)
            total += g.gmass[(size_t)k * N + i];
        g.density[i] = total * inverse_cell_volume;
    }
}

// Conductance of the face right of / below each cell: 2 / (rho_a + rho_b), zero at walls.
__global__ __launch_bounds__(kNT) void k_face_weights(GridDev g)
{
    FS_TILE_BEGIN(g.W, g.H)
    const int i = y * g.W + x;
    const int P = pidx(x, y, g.s);
    g.right64[P] = x + 1 < g.W ? 2.0 / (g.density[i] + g.density[i + 1]) : 0.0;
    g.down64[P] = y + 1 < g.H ? 2.0 / (g.density[i] + g.density[i + g.W]) : 0.0;
    FS_TILE_END
}

// PressureMultigrid Level::rebuildDiagonal (PressureMultigrid.cpp:104-117) in FP64, then FP32 copies.
__global__ __launch_bounds__(kNT) void k_level_diag(int w, int h, const double* right64, const double* down64,
    float* right32, float* down32, float* invd32)
{
    const int s = w + 2;
    FS_TILE_BEGIN(w, h)
    const int i = pidx(x, y, s);
    const double diagonal = right64[i - 1] + right64[i] + down64[i - s] + down64[i];
    invd32[i] = (float)(diagonal > 0.0 ? 1.0 / diagonal : 0.0);
    right32[i] = (float)right64[i];
    down32[i] = (float)down64[i];
    FS_TILE_END
}

// PressureMultigrid::coarsen (PressureMultigrid.cpp:142-202): parallel paths add, series resistances
// combine across the two or three fine faces between coarse centres.  One thread per coarse cell.
__global__ __launch_bounds__(kNT) void k_coarsen(int fw, int fh, const double* fcx, const double* fcy,
    const double* fright, const double* fdown, int cw, int ch, const double* ccx, const double* ccy,
    double* cright, double* cdown)
{
    const int fs = fw + 2, cs = cw + 2;
    FS_TILE_BEGIN(cw, ch)
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
    FS_TILE_END
}

// factorCoarsest (PressureMultigrid.cpp:300-345).  One workgroup; columns in order, the rows of a
// column in parallel.  Every entry sees the same operation sequence as the reference's row loop,
// so the factor is bit-identical.
__global__ __launch_bounds__(kNT) void k_cholesky(int w, int h, const double* right, const double* down,
    double* f, int n, PcgScalars* S)
{
    const int s = w + 2;
    for (int e = threadIdx.x; e < n * n; e += kNT)
        f[e] = 0.0;
    __syncthreads();
    for (int cell = 1 + threadIdx.x; cell <= n; cell += kNT) {
        const int x = cell % w, y = cell / w;
        const int i = (y + 1) * s + x + 1;
        const int row = cell - 1;
        f[row * n + row] = right[i - 1] + right[i] + down[i - s] + down[i];
        if (x > 0 && cell - 1 > 0)
            f[row * n + row - 1] = -right[i - 1];
        if (y > 0 && cell - w > 0)
            f[row * n + row - w] = -down[i - s];
    }
    __syncthreads();
    __shared__ int failed;
    if (threadIdx.x == 0)
        failed = 0;
    __syncthreads();
    for (int col = 0; col < n; ++col) {
        if (threadIdx.x == 0) {
            double value = f[col * n + col];
            for (int k = 0; k < col; ++k)
                value -= f[col * n + k] * f[col * n + k];
            if (value <= 0.0 || !isfinite(value))
                failed = 1;
            else
                f[col * n + col] = sqrt(value);
        }
        __syncthreads();
        if (failed)
            break;
        const double pivot = f[col * n + col];
        for (int row = col + 1 + threadIdx.x; row < n; row += kNT) {
            double value = f[row * n + col];
            for (int k = 0; k < col; ++k)
                value -= f[row * n + k] * f[col * n + k];
            f[row * n + col] = value / pivot;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0 && failed && S) {
        S->mg_failed = 1;
        if (S->status == kRunning)
            S->status = kFailed;
    }
}

// ------------------------------------------------------------------------------------------
// Projection set-up (FluidGridCPU.cpp:1050-1090, :1728-1756), FP64.
// ------------------------------------------------------------------------------------------

// Zero wall velocities, form rhs = -div * h^2 / dt, sum it.
__global__ __launch_bounds__(kNT) void k_rhs(GridDev g, double scale)
{
    __shared__ double lds[kNT / 64];
    if (g.S->status != kRunning)
        return;
    double sum = 0.0;
    const int W = g.W, H = g.H;
    FS_TILE_BEGIN(W, H)
    double* hz = g.u + (size_t)y * (W + 1);
    double* vt = g.v + (size_t)y * W;
    const double uL = x == 0 ? 0.0 : hz[x];
    const double uR = x + 1 == W ? 0.0 : hz[x + 1];
    const double vT = y == 0 ? 0.0 : vt[x];
    const double vB = y + 1 == H ? 0.0 : vt[x + W];
    if (x == 0) hz[0] = 0.0;
    if (x + 1 == W) hz[W] = 0.0;
    if (y == 0) vt[x] = 0.0;
    if (y + 1 == H) vt[x + W] = 0.0;
    const double du = uR - uL;
    const double dv = vB - vT;
    const double rhs = -(du + dv) / g.h * scale;
    g.b[pidx(x, y, g.s)] = rhs;
    sum += rhs;
    FS_TILE_END
    const double bs = block_sum<kNT>(sum, lds);
    if (threadIdx.x == 0)
        g.partials[block_id()] = bs;
    if (arrive_last(g.counter, block_count())) {
        const double total = finish_sum<kNT>(g.partials, block_count(), 1, 0, lds);
        if (threadIdx.x == 0) {
            g.S->b_mean = total / ((double)W * H);
            *g.counter = 0;
        }
    }
}

// Remove the rhs mean and find max |rhs| (the early-out test of solvePressure).
__global__ __launch_bounds__(kNT) void k_rhs_center(GridDev g)
{
    __shared__ double lds[kNT / 64];
    if (g.S->status != kRunning)
        return;
    const double mean = g.S->b_mean;
    double m = 0.0;
    FS_TILE_BEGIN(g.W, g.H)
    const int P = pidx(x, y, g.s);
    const double v = g.b[P] - mean;
    g.b[P] = v;
    m = fmax(m, abs_or_inf(v));
    FS_TILE_END
    const double bm = block_max<kNT>(m, lds);
    if (threadIdx.x == 0)
        g.partials[block_id()] = bm;
    if (arrive_last(g.counter, block_count())) {
        const double total = finish_max<kNT>(g.partials, block_count(), 1, 0, lds);
        if (threadIdx.x == 0) {
            g.S->bmax = total;
            if (total <= g.S->tol)
                g.S->status = kZeroRhs;
            *g.counter = 0;
        }
    }
}

// A divergence-free field needs no correction: pressure and its history restart at zero.
__global__ __launch_bounds__(kNT) void k_zero_pressure(GridDev g)
{
    if (g.S->status != kZeroRhs)
        return;
    FS_TILE_BEGIN(g.W, g.H)
    const int P = pidx(x, y, g.s);
    g.p[P] = 0.0;
    g.p_prev[P] = 0.0;
    FS_TILE_END
}

// Predictor p = 2 p - p_prev, history shift, pressure sum.
__global__ __launch_bounds__(kNT) void k_predict(GridDev g)
{
    __shared__ double lds[kNT / 64];
    if (g.S->status != kRunning)
        return;
    double sum = 0.0;
    FS_TILE_BEGIN(g.W, g.H)
    const int P = pidx(x, y, g.s);
    const double latest = g.p[P];
    const double next = 2.0 * latest - g.p_prev[P];
    g.p[P] = next;
    g.p_prev[P] = latest;
    sum += next;
    FS_TILE_END
    const double bs = block_sum<kNT>(sum, lds);
    if (threadIdx.x == 0)
        g.partials[block_id()] = bs;
    if (arrive_last(g.counter, block_count())) {
        const double total = finish_sum<kNT>(g.partials, block_count(), 1, 0, lds);
        if (threadIdx.x == 0) {
            g.S->p_mean = total / ((double)g.W * g.H);
            *g.counter = 0;
        }
    }
}

__global__ __launch_bounds__(kNT) void k_sub_mean_p(GridDev g, int final_pass)
{
    const int st = g.S->status;
    if (final_pass ? !(st == kConverged || st == kStopped) : st != kRunning)
        return;
    const double mean = g.S->p_mean;
    FS_TILE_BEGIN(g.W, g.H)
    const int P = pidx(x, y, g.s);
    g.p[P] -= mean;
    FS_TILE_END
}

// True residual r = b - A p in FP64 (FP64 weights), stored as the FP32 working residual.
// mode 0: initial residual of the predicted pressure.  mode 1: verification of a recurrence that
// reached the tolerance; accept (converged) or restart from the true residual.
__global__ __launch_bounds__(kNT) void k_true_residual(GridDev g, int mode)
{
    __shared__ double lds[kNT / 64];
    const int st = g.S->status;
    if (st != kRunning || (mode == 1 && !g.S->verify))
        return;
    const int s = g.s;
    double sum = 0.0, m = 0.0;
    FS_TILE_BEGIN(g.W, g.H)
    const int i = pidx(x, y, s);
    const double c = g.p[i];
    const double product = (g.right64[i - 1] * (c - g.p[i - 1]) + g.right64[i] * (c - g.p[i + 1])) +
                           (g.down64[i - s] * (c - g.p[i - s]) + g.down64[i] * (c - g.p[i + s]));
    const double r = g.b[i] - product;
    g.r[i] = (float)r;
    if (mode == 0)
        g.x[i] = 0.0f;
    sum += r;
    m = fmax(m, fabs(r));
    FS_TILE_END
    const double bs = block_sum<kNT>(sum, lds);
    const double bm = block_max<kNT>(m, lds);
    if (threadIdx.x == 0) {
        g.partials[2 * block_id()] = bs;
        g.partials[2 * block_id() + 1] = bm;
    }
    if (arrive_last(g.counter, block_count())) {
        const double total = finish_sum<kNT>(g.partials, block_count(), 2, 0, lds);
        double maximum = finish_max<kNT>(g.partials, block_count(), 2, 1, lds);
        if (threadIdx.x == 0) {
            PcgScalars* S = g.S;
            if (!isfinite(total))
                maximum = INFINITY;
            S->true_rmax = maximum;
            S->r_sum = total;
            if (mode == 0) {
                S->rmax = maximum;
                S->initial_rmax = maximum;
                S->r_mean = total / ((double)g.W * g.H);
                if (maximum <= S->tol || S->max_iterations == 0)
                    S->status = kInitialOk;
            } else {
                if (maximum <= S->tol) {
                    S->status = kConverged;
                } else {
                    S->r_mean = total / ((double)g.W * g.H);
                    S->restart = 1;
                    S->restarts += 1;
                }
                S->verify = 0;
            }
            *g.counter = 0;
        }
    }
}

// Fold the FP32 correction into the FP64 pressure before a true-residual check.
__global__ __launch_bounds__(kNT) void k_flush(GridDev g)
{
    if (g.S->status != kRunning || !g.S->verify)
        return;
    FS_TILE_BEGIN(g.W, g.H)
    const int P = pidx(x, y, g.s);
    g.p[P] += (double)g.x[P];
    g.x[P] = 0.0f;
    FS_TILE_END
}

// After the loop: fold the correction in and sum the pressure for the final mean removal.
__global__ __launch_bounds__(kNT) void k_final_flush_sum(GridDev g)
{
    __shared__ double lds[kNT / 64];
    const int st = g.S->status;
    if (!(st == kConverged || st == kStopped))
        return;
    double sum = 0.0;
    FS_TILE_BEGIN(g.W, g.H)
    const int P = pidx(x, y, g.s);
    const double v = g.p[P] + (double)g.x[P];
    g.p[P] = v;
    g.x[P] = 0.0f;
    sum += v;
    FS_TILE_END
    const double bs = block_sum<kNT>(sum, lds);
    if (threadIdx.x == 0)
        g.partials[block_id()] = bs;
    if (arrive_last(g.counter, block_count())) {
        const double total = finish_sum<kNT>(g.partials, block_count(), 1, 0, lds);
        if (threadIdx.x == 0) {
            g.S->p_mean = total / ((double)g.W * g.H);
            *g.counter = 0;
        }
    }
}

// u -= (dp/h) dt w_u ; v -= (dp/h) dt w_v (FluidGridCPU.cpp:1098-1118).  Each cell owns its left
// u face and its top v face.
__global__ __launch_bounds__(kNT) void k_correct_velocity(GridDev g, double dt)
{
    if (g.S->status == kFailed)
        return;
    const int W = g.W, s = g.s;
    FS_TILE_BEGIN(W, g.H)
    const int P = pidx(x, y, s);
    if (x >= 1)
        g.u[(size_t)y * (W + 1) + x] -= ((g.p[P] - g.p[P - 1]) / g.h) * dt * g.right64[P - 1];
    if (y >= 1)
        g.v[(size_t)y * W + x] -= ((g.p[P] - g.p[P - s]) / g.h) * dt * g.down64[P - s];
    FS_TILE_END
}

// maxDivergence (FluidGridCPU.cpp:1126-1152): non-finite divergence reports infinity.
__global__ __launch_bounds__(kNT) void k_max_divergence(GridDev g)
{
    __shared__ double lds[kNT / 64];
    const int W = g.W;
    double m = 0.0;
    FS_TILE_BEGIN(W, g.H)
    const double* hz = g.u + (size_t)y * (W + 1);
    const double* vt = g.v + (size_t)y * W;
    const double divergence = ((hz[x + 1] - hz[x]) + (vt[x + W] - vt[x])) / g.h;
    m = fmax(m, abs_or_inf(divergence));
    FS_TILE_END
    const double bm = block_max<kNT>(m, lds);
    if (threadIdx.x == 0)
        g.partials[block_id()] = bm;
    if (arrive_last(g.counter, block_count())) {
        const double total = finish_max<kNT>(g.partials, block_count(), 1, 0, lds);
        if (threadIdx.x == 0) {
            g.S->div_max = total;
            *g.counter = 0;
        }
    }
}

// applyGravity (FluidGridCPU.cpp:988-996): interior v rows only.
__global__ __launch_bounds__(kNT) void k_gravity(double* v, int first, int last, double increment)
{
    for (int i = first + blockIdx.x * kNT + threadIdx.x; i < last; i += gridDim.x * kNT)
        v[i] += increment;
}

// addVortex (FluidGridCPU.cpp:998-1028).
__global__ __launch_bounds__(kNT) void k_add_vortex(GridDev g, double ox, double oy, double cx, double cy,
    double radius, double strength)
{
    const int W = g.W, H = g.H;
    const int nu = (W + 1) * H, nv = W * (H + 1);
    for (int t = blockIdx.x * kNT + threadIdx.x; t < nu + nv; t += gridDim.x * kNT) {
        if (t < nu) {
            const int x = t % (W + 1), y = t / (W + 1);
            const double px = ox + x * g.h, py = oy + (y + 0.5) * g.h;
            const double dx = px - cx, dy = py - cy;
            const double r2 = dx * dx + dy * dy;
            const double falloff = exp(-r2 / (2.0 * radius * radius));
            g.u[t] += -dy * strength * falloff;
        } else {
            const int e = t - nu;
            const int x = e % W, y = e / W;
            const double px = ox + (x + 0.5) * g.h, py = oy + y * g.h;
            const double dx = px - cx, dy = py - cy;
            const double r2 = dx * dx + dy * dy;
            const double falloff = exp(-r2 / (2.0 * radius * radius));
            g.v[e] += dx * strength * falloff;
        }
    }
}

// ---- semi-Lagrangian transport mode (build-defined; samples as FluidGridCPU::sampleVelocity) ----
__device__ __forceinline__ double lerp_std(double a, double b, double t)
{
    if ((a <= 0 && b >= 0) || (a >= 0 && b <= 0))
        return t * b + (1 - t) * a;
    if (t == 1)
        return b;
    const double x = a + t * (b - a);
    return (t > 1) == (b > a) ? (b < x ? x : b) : (b > x ? x : b);
}

__device__ __forceinline__ double dclamp(double v, double lo, double hi) { return v < lo ? lo : (hi < v ? hi : v); }

__device__ double sample_scalar(const double* data, int width, int height, double x, double y)
{
    x = dclamp(x, 0.0, (double)(width - 1));
    y = dclamp(y, 0.0, (double)(height - 1));
    // a non-finite sample position (non-finite velocity) reads cell 0 instead of an undefined index
    if (!(x >= 0.0))
        x = 0.0;
    if (!(y >= 0.0))
        y = 0.0;
    const int x0 = (int)x, y0 = (int)y;
    const int x1 = min(x0 + 1, width - 1), y1 = min(y0 + 1, height - 1);
    const double tx = x - x0, ty = y - y0;
    const double a = lerp_std(data[y0 * width + x0], data[y0 * width + x1], tx);
    const double b = lerp_std(data[y1 * width + x0], data[y1 * width + x1], tx);
    return lerp_std(a, b, ty);
}

__device__ __forceinline__ void sample_velocity(const GridDev& g, double ox, double oy, double px, double py,
    double* vx, double* vy)
{
    const double x = (px - ox) / g.h;
    const double y = (py - oy) / g.h;
    *vx = sample_scalar(g.u, g.W + 1, g.H, x, y - 0.5);
    *vy = sample_scalar(g.v, g.W, g.H + 1, x - 0.5, y);
}

__device__ __forceinline__ void backtrace(const GridDev& g, double ox, double oy, double px, double py, double dt,
    double* sx, double* sy)
{
    double vx, vy;
    sample_velocity(g, ox, oy, px, py, &vx, &vy);
    const double mx = px - 0.5 * dt * vx, my = py - 0.5 * dt * vy;
    sample_velocity(g, ox, oy, mx, my, &vx, &vy);
    *sx = px - dt * vx;
    *sy = py - dt * vy;
}

__global__ __launch_bounds__(kNT) void k_sl_faces(GridDev g, double ox, double oy, double dt, double* nu,
    double* nv)
{
    const int W = g.W, H = g.H;
    const int cu = (W + 1) * H, cv = W * (H + 1);
    for (int t = blockIdx.x * kNT + threadIdx.x; t < cu + cv; t += gridDim.x * kNT) {
        if (t < cu) {
            const int x = t % (W + 1), y = t / (W + 1);
            double value = 0.0;
            if (x >= 1 && x < W) {
                double sx, sy;
                backtrace(g, ox, oy, ox + x * g.h, oy + (y + 0.5) * g.h, dt, &sx, &sy);
                value = sample_scalar(g.u, W + 1, H, (sx - ox) / g.h, (sy - oy) / g.h - 0.5);
            }
            nu[t] = value;
        } else {
            const int e = t - cu;
            const int x = e % W, y = e / W;
            double value = 0.0;
            if (y >= 1 && y < H) {
                double sx, sy;
                backtrace(g, ox, oy, ox + (x + 0.5) * g.h, oy + y * g.h, dt, &sx, &sy);
                value = sample_scalar(g.v, W, H + 1, (sx - ox) / g.h - 0.5, (sy - oy) / g.h);
            }
            nv[e] = value;
        }
    }
}

// Cell-centred scalars (fraction, every component mass) share one backtrace per cell.
__global__ __launch_bounds__(kNT) void k_sl_cells(GridDev g, double ox, double oy, double dt, double* out)
{
    const int W = g.W, H = g.H;
    const size_t N = (size_t)W * H;
    const int fields = 1 + g.nl + g.ng;
    for (size_t i = (size_t)blockIdx.x * kNT + threadIdx.x; i < N; i += (size_t)gridDim.x * kNT) {
        const int x = (int)(i % W), y = (int)(i / W);
        double sx, sy;
        backtrace(g, ox, oy, ox + (x + 0.5) * g.h, oy + (y + 0.5) * g.h, dt, &sx, &sy);
        const double fx = (sx - ox) / g.h - 0.5, fy = (sy - oy) / g.h - 0.5;
        for (int f = 0; f < fields; ++f) {
            const double* a = f == 0 ? g.fraction
                : (f <= g.nl ? g.lmass + (size_t)(f - 1) * N : g.gmass + (size_t)(f - 1 - g.nl) * N);
            out[(size_t)f * N + i] = sample_scalar(a, W, H, fx, fy);
        }
    }
}

__global__ __launch_bounds__(kNT) void k_pad_copy(const double* src, double* dst, int W, int H, int to_padded)
{
    const int s = W + 2;
    FS_TILE_BEGIN(W, H)
    if (to_padded)
        dst[pidx(x, y, s)] = src[(size_t)y * W + x];
    else
        dst[(size_t)y * W + x] = src[pidx(x, y, s)];
    FS_TILE_END
}

// ------------------------------------------------------------------------------------------
// PCG loop (FluidGridCPU.cpp:1779-1819) with FP32 vectors and FP64 scalars.
// ------------------------------------------------------------------------------------------

// q = A d (5-point, variable coefficient; walls have zero weight), d.q, then alpha.
__global__ __launch_bounds__(kNT) void k_pcg_ad(GridDev g)
{
    __shared__ double lds[kNT / 64];
    if (g.S->status != kRunning)
        return;
    const int s = g.s;
    const float* d = g.d;
    double dot = 0.0;
    FS_TILE_BEGIN(g.W, g.H)
    const int i = pidx(x, y, s);
    const float c = d[i];
    const float q = (g.right32[i - 1] * (c - d[i - 1]) + g.right32[i] * (c - d[i + 1])) +
                    (g.down32[i - s] * (c - d[i - s]) + g.down32[i] * (c - d[i + s]));
    g.q[i] = q;
    dot += (double)c * (double)q;
    FS_TILE_END
    const double bs = block_sum<kNT>(dot, lds);
    if (threadIdx.x == 0)
        g.partials[block_id()] = bs;
    if (arrive_last(g.counter, block_count())) {
        const double total = finish_sum<kNT>(g.partials, block_count(), 1, 0, lds);
        if (threadIdx.x == 0) {
            PcgScalars* S = g.S;
            S->dq = total;
            if (total <= 0.0 || !isfinite(total)) {
                S->status = kStopped;
            } else {
                const double alpha = S->rz / total;
                if (!isfinite(alpha))
                    S->status = kStopped;
                S->alpha = alpha;
            }
            *g.counter = 0;
        }
    }
}

// x += alpha d ; r -= alpha q ; residual sum and max-norm; convergence test of the recurrence.
__global__ __launch_bounds__(kNT) void k_pcg_update(GridDev g)
{
    __shared__ double lds[kNT / 64];
    if (g.S->status != kRunning)
        return;
    const float alpha = (float)g.S->alpha;
    double sum = 0.0, m = 0.0;
    FS_TILE_BEGIN(g.W, g.H)
    const int i = pidx(x, y, g.s);
    g.x[i] += alpha * g.d[i];
    const float r = g.r[i] - alpha * g.q[i];
    g.r[i] = r;
    sum += (double)r;
    m = fmax(m, (double)fabsf(r));
    FS_TILE_END
    const double bs = block_sum<kNT>(sum, lds);
    const double bm = block_max<kNT>(m, lds);
    if (threadIdx.x == 0) {
        g.partials[2 * block_id()] = bs;
        g.partials[2 * block_id() + 1] = bm;
    }
    if (arrive_last(g.counter, block_count())) {
        const double total = finish_sum<kNT>(g.partials, block_count(), 2, 0, lds);
        double maximum = finish_max<kNT>(g.partials, block_count(), 2, 1, lds);
        if (threadIdx.x == 0) {
            PcgScalars* S = g.S;
            if (!isfinite(total))
                maximum = INFINITY;
            S->iterations += 1;
            S->rmax = maximum;
            S->r_sum = total;
            S->r_mean = total / ((double)g.W * g.H);
            if (maximum <= S->tol)
                S->verify = 1;
            *g.counter = 0;
        }
    }
}

// Diagonal preconditioner z = (r - mean) / diag.
__global__ __launch_bounds__(kNT) void k_precond_diag(GridDev g)
{
    if (g.S->status != kRunning)
        return;
    const double mean = g.S->r_mean;
    FS_TILE_BEGIN(g.W, g.H)
    const int i = pidx(x, y, g.s);
    g.zdiag[i] = (float)(((double)g.r[i] - mean) * (double)g.invd32[i]);
    FS_TILE_END
}

// sum z, r.z and sum r in one pass; r.(z - mean z) = r.z - mean z * sum r; then beta.
__global__ __launch_bounds__(kNT) void k_pcg_rz(GridDev g, const float* z)
{
    __shared__ double lds[kNT / 64];
    if (g.S->status != kRunning)
        return;
    double sz = 0.0, srz = 0.0, sr = 0.0;
    FS_TILE_BEGIN(g.W, g.H)
    const int i = pidx(x, y, g.s);
    const double zv = (double)z[i];
    const double rv = (double)g.r[i];
    sz += zv;
    srz += rv * zv;
    sr += rv;
    FS_TILE_END
    const double b0 = block_sum<kNT>(sz, lds);
    const double b1 = block_sum<kNT>(srz, lds);
    const double b2 = block_sum<kNT>(sr, lds);
    if (threadIdx.x == 0) {
        g.partials[3 * block_id()] = b0;
        g.partials[3 * block_id() + 1] = b1;
        g.partials[3 * block_id() + 2] = b2;
    }
    if (arrive_last(g.counter, block_count())) {
        const double tz = finish_sum<kNT>(g.partials, block_count(), 3, 0, lds);
        const double trz = finish_sum<kNT>(g.partials, block_count(), 3, 1, lds);
        const double tr = finish_sum<kNT>(g.partials, block_count(), 3, 2, lds);
        if (threadIdx.x == 0) {
            PcgScalars* S = g.S;
            const double n = (double)g.W * g.H;
            const double mean = tz / n;
            const double next = trz - mean * tr;
            S->z_mean = mean;
            const double beta = S->restart ? 0.0 : next / S->rz;
            S->beta = beta;
            S->restart = 0;
            if (!isfinite(beta)) {
                S->status = kStopped;
            } else {
                S->rz = next;
                if (!(S->iterations < S->max_iterations && next > 0.0 && isfinite(next)))
                    S->status = kStopped;
            }
            *g.counter = 0;
        }
    }
}

// d = (z - mean z) + beta d
__global__ __launch_bounds__(kNT) void k_pcg_direction(GridDev g, const float* z)
{
    if (g.S->status != kRunning)
        return;
    const float mean = (float)g.S->z_mean;
    const double beta64 = g.S->beta;
    const float beta = (float)beta64;
    FS_TILE_BEGIN(g.W, g.H)
    const int i = pidx(x, y, g.s);
    const float zz = z[i] - mean;
    g.d[i] = beta64 == 0.0 ? zz : zz + beta * g.d[i];
    FS_TILE_END
}

// ------------------------------------------------------------------------------------------
// Multigrid V(4,4) pieces (PressureMultigrid.cpp:204-298), FP32.
// ------------------------------------------------------------------------------------------

// One full red-black Gauss-Seidel sweep per launch.  The block stages a (64+4) x (16+4) tile of the
// solution and coefficients in LDS; the first colour is updated on the tile plus a one-cell ring,
// the second colour on the tile, and the tile is written to the other ping-pong buffer (no block
// reads what another block writes in the same launch).
template <int TX, int TY>
__global__ __launch_bounds__(kNT) void k_rb_sweep(LevelView L, int cur, const float* rhs, const double* shift,
    int c0, int from_zero, const int* status)
{
    if (status && *status != kRunning)
        return;
    constexpr int LX = TX + 4, LY = TY + 4;
    __shared__ float s_sol[LY][LX];
    __shared__ float s_right[LY][LX];
    __shared__ float s_down[LY][LX];
    __shared__ float s_rhs[LY][LX];
    __shared__ float s_invd[LY][LX];
    const int w = L.w, h = L.h, s = L.s;
    const int x0 = blockIdx.x * TX - 2, y0 = blockIdx.y * TY - 2;
    const double sh = shift ? *shift : 0.0;
    const float* sol_in = L.sol[cur];
    for (int e = threadIdx.x; e < LX * LY; e += kNT) {
        const int ly = e / LX, lx = e - ly * LX;
        const int x = x0 + lx, y = y0 + ly;
        float vs = 0.f, vr = 0.f, vd = 0.f, vb = 0.f, vi = 0.f;
        if (x >= -1 && x <= w && y >= -1 && y <= h) {
            const int P = (y + 1) * s + (x + 1);
            vs = sol_in[P];
            vr = L.right[P];
            vd = L.down[P];
            vi = L.invd[P];
            if (x >= 0 && x < w && y >= 0 && y < h)
                vb = shift ? (float)((double)rhs[P] - sh) : rhs[P];
        }
        s_sol[ly][lx] = vs;
        s_right[ly][lx] = vr;
        s_down[ly][lx] = vd;
        s_rhs[ly][lx] = vb;
        s_invd[ly][lx] = vi;
    }
    __syncthreads();
    // first colour on the tile plus ring (rows/cols 1 .. L-2)
    {
        constexpr int PR = (TX + 2) / 2;
        for (int t = threadIdx.x; t < PR * (TY + 2); t += kNT) {
            const int j = t / PR, q = t - j * PR;
            const int ly = 1 + j;
            int lx = 1 + 2 * q;
            const int y = y0 + ly;
            int x = x0 + lx;
            if (((x + y) & 1) != c0) {
                ++lx;
                ++x;
            }
            if (x < 0 || x >= w || y < 0 || y >= h)
                continue;
            float v = s_rhs[ly][lx];
            if (!from_zero)
                v += (s_right[ly][lx - 1] * s_sol[ly][lx - 1] + s_right[ly][lx] * s_sol[ly][lx + 1]) +
                     (s_down[ly - 1][lx] * s_sol[ly - 1][lx] + s_down[ly][lx] * s_sol[ly + 1][lx]);
            s_sol[ly][lx] = v * s_invd[ly][lx];
        }
    }
    __syncthreads();
    // second colour on the tile (rows/cols 2 .. L-3)
    {
        constexpr int PR = TX / 2;
        const int c1 = c0 ^ 1;
        for (int t = threadIdx.x; t < PR * TY; t += kNT) {
            const int j = t / PR, q = t - j * PR;
            const int ly = 2 + j;
            int lx = 2 + 2 * q;
            const int y = y0 + ly;
            int x = x0 + lx;
            if (((x + y) & 1) != c1) {
                ++lx;
                ++x;
            }
            if (x >= w || y >= h)
                continue;
            const float v = s_rhs[ly][lx] +
                ((s_right[ly][lx - 1] * s_sol[ly][lx - 1] + s_right[ly][lx] * s_sol[ly][lx + 1]) +
                    (s_down[ly - 1][lx] * s_sol[ly - 1][lx] + s_down[ly][lx] * s_sol[ly + 1][lx]));
            s_sol[ly][lx] = v * s_invd[ly][lx];
        }
    }
    __syncthreads();
    float* sol_out = L.sol[cur ^ 1];
    for (int e = threadIdx.x; e < TX * TY; e += kNT) {
        const int ly = 2 + e / TX, lx = 2 + e % TX;
        const int x = x0 + lx, y = y0 + ly;
        if (x < w && y < h)
            sol_out[(y + 1) * s + (x + 1)] = s_sol[ly][lx];
    }
}

// res = rhs - A sol
__global__ __launch_bounds__(kNT) void k_mg_residual(LevelView L, int cur, const float* rhs, const double* shift,
    const int* status)
{
    if (status && *status != kRunning)
        return;
    const int s = L.s;
    const float* sol = L.sol[cur];
    const double sh = shift ? *shift : 0.0;
    FS_TILE_BEGIN(L.w, L.h)
    const int i = pidx(x, y, s);
    const float c = sol[i];
    const float product = (L.right[i - 1] * (c - sol[i - 1]) + L.right[i] * (c - sol[i + 1])) +
                          (L.down[i - s] * (c - sol[i - s]) + L.down[i] * (c - sol[i + s]));
    const float b = shift ? (float)((double)rhs[i] - sh) : rhs[i];
    L.res[i] = b - product;
    FS_TILE_END
}

// coarse rhs = R res (separable transpose-of-bilinear gather, <= 4 taps per axis)
__global__ __launch_bounds__(kNT) void k_mg_restrict(LevelView F, LevelView Cl, AxisRestrict rx, AxisRestrict ry,
    const int* status)
{
    if (status && *status != kRunning)
        return;
    const int fs = F.s;
    FS_TILE_BEGIN(Cl.w, Cl.h)
    const int nx = rx.count[x], ny = ry.count[y];
    float value = 0.f;
    for (int iy = 0; iy < ny; ++iy) {
        const float* row = F.res + (ry.index[y * 4 + iy] + 1) * fs + 1;
        float horizontal = 0.f;
        for (int ix = 0; ix < nx; ++ix)
            horizontal += rx.weight[x * 4 + ix] * row[rx.index[x * 4 + ix]];
        value += ry.weight[y * 4 + iy] * horizontal;
    }
    Cl.rhs[pidx(x, y, Cl.s)] = value;
    FS_TILE_END
}

// fine sol += P coarse sol (bilinear)
__global__ __launch_bounds__(kNT) void k_mg_prolong(LevelView F, int fcur, LevelView Cl, int ccur, AxisProlong px,
    AxisProlong py, const int* status)
{
    if (status && *status != kRunning)
        return;
    const float* csol = Cl.sol[ccur];
    float* fsol = F.sol[fcur];
    FS_TILE_BEGIN(F.w, F.h)
    const float wy = py.weight[y];
    const float* top = csol + (py.first[y] + 1) * Cl.s + 1;
    const float* bottom = csol + (py.second[y] + 1) * Cl.s + 1;
    const int a0 = px.first[x], a1 = px.second[x];
    const float wx = px.weight[x];
    const float a = (1.f - wx) * top[a0] + wx * top[a1];
    const float b = (1.f - wx) * bottom[a0] + wx * bottom[a1];
    fsol[pidx(x, y, F.s)] += (1.f - wy) * a + wy * b;
    FS_TILE_END
}

// ---- coarse tail: the remaining levels and the dense coarse solve in one workgroup, in LDS ----
struct TailLds {
    float* right;
    float* down;
    float* invd;
    float* rhs;
    float* sol;
    float* res;
};

__device__ void tail_smooth(const TailLevel& T, TailLds& A, int reverse, int from_zero)
{
    const int w = T.w, h = T.h, s = T.s;
    for (int pass = 0; pass < 2; ++pass) {
        const int color = reverse ? 1 - pass : pass;
        const int pairs = (w + 1) / 2;
        for (int t = threadIdx.x; t < pairs * h; t += kNT) {
            const int y = t / pairs;
            const int x = 2 * (t - y * pairs) + ((color ^ y) & 1);
            if (x >= w)
                continue;
            const int i = (y + 1) * s + x + 1;
            float v = A.rhs[i];
            if (!(from_zero && pass == 0))
                v += (A.right[i - 1] * A.sol[i - 1] + A.right[i] * A.sol[i + 1]) +
                     (A.down[i - s] * A.sol[i - s] + A.down[i] * A.sol[i + s]);
            A.sol[i] = v * A.invd[i];
        }
        __syncthreads();
    }
}

__device__ void tail_residual_restrict(const TailLevel& T, TailLds& F, const TailLevel& Cn, TailLds& Cd)
{
    const int w = T.w, h = T.h, s = T.s;
    for (int e = threadIdx.x; e < w * h; e += kNT) {
        const int y = e / w, x = e - y * w;
        const int i = (y + 1) * s + x + 1;
        const float c = F.sol[i];
        const float product = (F.right[i - 1] * (c - F.sol[i - 1]) + F.right[i] * (c - F.sol[i + 1])) +
                              (F.down[i - s] * (c - F.sol[i - s]) + F.down[i] * (c - F.sol[i + s]));
        F.res[i] = F.rhs[i] - product;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < Cn.w * Cn.h; e += kNT) {
        const int y = e / Cn.w, x = e - y * Cn.w;
        const int nx = T.rx_count[x], ny = T.ry_count[y];
        float value = 0.f;
        for (int iy = 0; iy < ny; ++iy) {
            const float* row = F.res + (T.ry_index[y * 4 + iy] + 1) * s + 1;
            float horizontal = 0.f;
            for (int ix = 0; ix < nx; ++ix)
                horizontal += T.rx_w[x * 4 + ix] * row[T.rx_index[x * 4 + ix]];
            value += T.ry_w[y * 4 + iy] * horizontal;
        }
        Cd.rhs[(y + 1) * Cn.s + x + 1] = value;
    }
    __syncthreads();
}

__device__ void tail_prolong(const TailLevel& T, TailLds& F, const TailLevel& Cn, TailLds& Cd)
{
    const int w = T.w, h = T.h;
    for (int e = threadIdx.x; e < w * h; e += kNT) {
        const int y = e / w, x = e - y * w;
        const float wy = T.py_w[y];
        const float* top = Cd.sol + (T.py_first[y] + 1) * Cn.s + 1;
        const float* bottom = Cd.sol + (T.py_second[y] + 1) * Cn.s + 1;
        const int a0 = T.px_first[x], a1 = T.px_second[x];
        const float wx = T.px_w[x];
        const float a = (1.f - wx) * top[a0] + wx * top[a1];
        const float b = (1.f - wx) * bottom[a0] + wx * bottom[a1];
        F.sol[(y + 1) * T.s + x + 1] += (1.f - wy) * a + wy * b;
    }
    __syncthreads();
}

// solveCoarsest (PressureMultigrid.cpp:347-384): one pinned cell, dense Cholesky substitution in FP64,
// mean-zero gauge.  Wavefront 0 runs the substitutions with cross-lane broadcasts.
__device__ void tail_coarse_solve(const TailLevel& T, TailLds& A, const double* f, int n, double* cbuf)
{
    const int w = T.w, s = T.s;
    __shared__ double s_mean;
    if (threadIdx.x == 0) {
        double mean = 0.0;
        for (int y = 0; y < T.h; ++y)
            for (int x = 0; x < w; ++x)
                mean += (double)A.rhs[(y + 1) * s + x + 1];
        s_mean = mean / (n + 1);
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        const int cell = lane + 1;
        double val = 0.0;
        if (lane < n)
            val = (double)A.rhs[(cell / w + 1) * s + cell % w + 1] - s_mean;
        // forward: L c = rhs (column order; each row accumulates in increasing column order)
        for (int col = 0; col < n; ++col) {
            const double ccol = __shfl(val, col, 64) / f[col * n + col];
            if (lane == col)
                val = ccol;
            else if (lane > col && lane < n)
                val -= f[lane * n + col] * ccol;
        }
        // backward: L^T c = y
        for (int col = n - 1; col >= 0; --col) {
            const double ccol = __shfl(val, col, 64) / f[col * n + col];
            if (lane == col)
                val = ccol;
            else if (lane < col)
                val -= f[col * n + lane] * ccol;
        }
        // mean-zero gauge over all n + 1 cells (the pinned cell holds 0 before the shift)
        double part = lane < n ? val : 0.0;
        part = wave_sum(part);
        const double mean = part / (n + 1);
        if (lane == 0)
            A.sol[s + 1] = (float)(-mean);
        if (lane < n)
            A.sol[(cell / w + 1) * s + cell % w + 1] = (float)(val - mean);
    }
    (void)cbuf;
    __syncthreads();
}

__device__ __forceinline__ TailLds tail_arrays(float* smem, const TailDesc* D, int l)
{
    const TailLevel& T = D->lev[l];
    const int ps = T.s * (T.h + 2);
    float* p = smem + D->lds_offset[l];
    TailLds A;
    A.right = p;
    A.down = p + ps;
    A.invd = p + 2 * ps;
    A.rhs = p + 3 * ps;
    A.sol = p + 4 * ps;
    A.res = p + 5 * ps;
    return A;
}

__global__ __launch_bounds__(kNT) void k_mg_tail(const TailDesc* D, int cur, const double* shift, const int* status)
{
    if (status && *status != kRunning)
        return;
    extern __shared__ float smem[];
    const int nl = D->levels;
    const double sh = shift ? *shift : 0.0;
    for (int l = 0; l < nl; ++l) {
        const TailLevel& T = D->lev[l];
        TailLds A = tail_arrays(smem, D, l);
        const int ps = T.s * (T.h + 2);
        for (int e = threadIdx.x; e < ps; e += kNT) {
            A.right[e] = T.right[e];
            A.down[e] = T.down[e];
            A.invd[e] = T.invd[e];
            A.sol[e] = 0.f;
            A.res[e] = 0.f;
            float b = 0.f;
            if (l == 0) {
                const int y = e / T.s - 1, x = e % T.s - 1;
                if (x >= 0 && x < T.w && y >= 0 && y < T.h)
                    b = shift ? (float)((double)D->top_rhs[e] - sh) : D->top_rhs[e];
            }
            A.rhs[e] = b;
        }
    }
    __syncthreads();
    for (int l = 0; l + 1 < nl; ++l) {
        const TailLevel& T = D->lev[l];
        TailLds A = tail_arrays(smem, D, l);
        TailLds B = tail_arrays(smem, D, l + 1);
        tail_smooth(T, A, 0, 1);
        for (int k = 1; k < 4; ++k)
            tail_smooth(T, A, 0, 0);
        tail_residual_restrict(T, A, D->lev[l + 1], B);
    }
    {
        TailLds A = tail_arrays(smem, D, nl - 1);
        tail_coarse_solve(D->lev[nl - 1], A, D->factor, D->coarse_n, nullptr);
    }
    for (int l = nl - 2; l >= 0; --l) {
        const TailLevel& T = D->lev[l];
        TailLds A = tail_arrays(smem, D, l);
        TailLds B = tail_arrays(smem, D, l + 1);
        tail_prolong(T, A, D->lev[l + 1], B);
        for (int k = 0; k < 4; ++k)
            tail_smooth(T, A, 1, 0);
    }
    const TailLevel& T0 = D->lev[0];
    TailLds A0 = tail_arrays(smem, D, 0);
    float* out = D->top_sol[cur];
    for (int e = threadIdx.x; e < T0.w * T0.h; e += kNT) {
        const int y = e / T0.w, x = e - y * T0.w;
        const int i = (y + 1) * T0.s + x + 1;
        out[i] = A0.sol[i];
    }
}

__global__ void k_copy_f32(const float* src, float* dst, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

inline unsigned linear_blocks(size_t n)
{
    size_t b = (n + kNT - 1) / kNT;
    if (b > 8192)
        b = 8192;
    return (unsigned)(b ? b : 1);
}

}  // namespace

// ------------------------------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------------------------------

size_t tail_lds_bytes(const TailDesc& t)
{
    size_t floats = 0;
    for (int l = 0; l < t.levels; ++l)
        floats += (size_t)6 * t.lev[l].s * (t.lev[l].h + 2);
    return floats * sizeof(float);
}

void launch_init_scalars(hipStream_t st, PcgScalars* S, double tol, int max_iterations)
{
    hipLaunchKernelGGL(k_init_scalars, dim3(1), dim3(64), 0, st, S, tol, max_iterations);
}

void launch_density(hipStream_t st, const GridDev& g)
{
    hipLaunchKernelGGL(k_density, dim3(linear_blocks((size_t)g.W * g.H)), dim3(kNT), 0, st, g);
}

void launch_face_weights(hipStream_t st, const GridDev& g)
{
    hipLaunchKernelGGL(k_face_weights, cell_grid(g.W, g.H), dim3(kNT), 0, st, g);
}

void launch_level_diag(hipStream_t st, int w, int h, const double* right64, const double* down64, float* right32,
    float* down32, float* invd32)
{
    hipLaunchKernelGGL(k_level_diag, cell_grid(w, h), dim3(kNT), 0, st, w, h, right64, down64, right32, down32, invd32);
}

void launch_coarsen(hipStream_t st, int fw, int fh, const double* fcx, const double* fcy, const double* fright,
    const double* fdown, int cw, int ch, const double* ccx, const double* ccy, double* cright, double* cdown)
{
    hipLaunchKernelGGL(k_coarsen, cell_grid(cw, ch), dim3(kNT), 0, st, fw, fh, fcx, fcy, fright, fdown, cw, ch, ccx,
        ccy, cright, cdown);
}

void launch_cholesky(hipStream_t st, int w, int h, const double* right64, const double* down64, double* factor,
    int n, PcgScalars* S)
{
    if (n <= 0)
        return;
    hipLaunchKernelGGL(k_cholesky, dim3(1), dim3(kNT), 0, st, w, h, right64, down64, factor, n, S);
}

void launch_rhs(hipStream_t st, const GridDev& g, double scale)
{
    hipLaunchKernelGGL(k_rhs, cell_grid(g.W, g.H), dim3(kNT), 0, st, g, scale);
}

void launch_rhs_center(hipStream_t st, const GridDev& g)
{
    hipLaunchKernelGGL(k_rhs_center, cell_grid(g.W, g.H), dim3(kNT), 0, st, g);
}

void launch_zero_pressure(hipStream_t st, const GridDev& g)
{
    hipLaunchKernelGGL(k_zero_pressure, cell_grid(g.W, g.H), dim3(kNT), 0, st, g);
}

void launch_predict(hipStream_t st, const GridDev& g)
{
    hipLaunchKernelGGL(k_predict, cell_grid(g.W, g.H), dim3(kNT), 0, st, g);
}

void launch_sub_mean_p(hipStream_t st, const GridDev& g, int final_pass)
{
    hipLaunchKernelGGL(k_sub_mean_p, cell_grid(g.W, g.H), dim3(kNT), 0, st, g, final_pass);
}

void launch_true_residual(hipStream_t st, const GridDev& g, int mode)
{
    hipLaunchKernelGGL(k_true_residual, cell_grid(g.W, g.H), dim3(kNT), 0, st, g, mode);
}

void launch_flush(hipStream_t st, const GridDev& g)
{
    hipLaunchKernelGGL(k_flush, cell_grid(g.W, g.H), dim3(kNT), 0, st, g);
}

void launch_final_flush_sum(hipStream_t st, const GridDev& g)
{
    hipLaunchKernelGGL(k_final_flush_sum, cell_grid(g.W, g.H), dim3(kNT), 0, st, g);
}

void launch_correct_velocity(hipStream_t st, const GridDev& g, double dt)
{
    hipLaunchKernelGGL(k_correct_velocity, cell_grid(g.W, g.H), dim3(kNT), 0, st, g, dt);
}

void launch_max_divergence(hipStream_t st, const GridDev& g)
{
    hipLaunchKernelGGL(k_max_divergence, cell_grid(g.W, g.H), dim3(kNT), 0, st, g);
}

void launch_gravity(hipStream_t st, const GridDev& g, double gy, double dt)
{
    const int first = g.W, last = g.W * g.H;
    if (last <= first)
        return;
    hipLaunchKernelGGL(k_gravity, dim3(linear_blocks((size_t)(last - first))), dim3(kNT), 0, st, g.v, first, last,
        gy * dt);
}

void launch_add_vortex(hipStream_t st, const GridDev& g, double ox, double oy, double cx, double cy, double radius,
    double strength)
{
    const size_t n = (size_t)(g.W + 1) * g.H + (size_t)g.W * (g.H + 1);
    hipLaunchKernelGGL(k_add_vortex, dim3(linear_blocks(n)), dim3(kNT), 0, st, g, ox, oy, cx, cy, radius, strength);
}

void launch_semi_lagrangian(hipStream_t st, const GridDev& g, double ox, double oy, double dt, double* new_u,
    double* new_v, double* scratch_cells)
{
    const size_t nf = (size_t)(g.W + 1) * g.H + (size_t)g.W * (g.H + 1);
    hipLaunchKernelGGL(k_sl_faces, dim3(linear_blocks(nf)), dim3(kNT), 0, st, g, ox, oy, dt, new_u, new_v);
    hipLaunchKernelGGL(k_sl_cells, dim3(linear_blocks((size_t)g.W * g.H)), dim3(kNT), 0, st, g, ox, oy, dt,
        scratch_cells);
}

void launch_pad_copy(hipStream_t st, const double* src, double* dst, int W, int H, int to_padded)
{
    hipLaunchKernelGGL(k_pad_copy, cell_grid(W, H), dim3(kNT), 0, st, src, dst, W, H, to_padded);
}

void launch_pcg_ad(hipStream_t st, const GridDev& g)
{
    hipLaunchKernelGGL(k_pcg_ad, cell_grid(g.W, g.H), dim3(kNT), 0, st, g);
}

void launch_pcg_update(hipStream_t st, const GridDev& g)
{
    hipLaunchKernelGGL(k_pcg_update, cell_grid(g.W, g.H), dim3(kNT), 0, st, g);
}

void launch_pcg_rz(hipStream_t st, const GridDev& g, const float* z)
{
    hipLaunchKernelGGL(k_pcg_rz, cell_grid(g.W, g.H), dim3(kNT), 0, st, g, z);
}

void launch_pcg_direction(hipStream_t st, const GridDev& g, const float* z)
{
    hipLaunchKernelGGL(k_pcg_direction, cell_grid(g.W, g.H), dim3(kNT), 0, st, g, z);
}

void launch_precond_diag(hipStream_t st, const GridDev& g)
{
    hipLaunchKernelGGL(k_precond_diag, cell_grid(g.W, g.H), dim3(kNT), 0, st, g);
}

void launch_rb_sweep(hipStream_t st, const LevelView& L, int cur, const float* rhs, const double* shift,
    int first_color, int from_zero, const int* status)
{
    constexpr int TX = 64, TY = 16;
    const dim3 grid((L.w + TX - 1) / TX, (L.h + TY - 1) / TY);
    hipLaunchKernelGGL((k_rb_sweep<TX, TY>), grid, dim3(kNT), 0, st, L, cur, rhs, shift, first_color, from_zero,
        status);
}

void launch_mg_residual(hipStream_t st, const LevelView& L, int cur, const float* rhs, const double* shift,
    const int* status)
{
    hipLaunchKernelGGL(k_mg_residual, cell_grid(L.w, L.h), dim3(kNT), 0, st, L, cur, rhs, shift, status);
}

void launch_mg_restrict(hipStream_t st, const LevelView& F, const LevelView& Cl, AxisRestrict rx, AxisRestrict ry,
    const int* status)
{
    hipLaunchKernelGGL(k_mg_restrict, cell_grid(Cl.w, Cl.h), dim3(kNT), 0, st, F, Cl, rx, ry, status);
}

void launch_mg_prolong(hipStream_t st, const LevelView& F, int fcur, const LevelView& Cl, int ccur, AxisProlong px,
    AxisProlong py, const int* status)
{
    hipLaunchKernelGGL(k_mg_prolong, cell_grid(F.w, F.h), dim3(kNT), 0, st, F, fcur, Cl, ccur, px, py, status);
}

void launch_mg_tail(hipStream_t st, const TailDesc* dev_desc, const TailDesc& host_desc, int cur,
    const double* shift, const int* status)
{
    const size_t lds = tail_lds_bytes(host_desc);
    hipLaunchKernelGGL(k_mg_tail, dim3(1), dim3(kNT), lds, st, dev_desc, cur, shift, status);
}

void launch_copy_f32(hipStream_t st, const float* src, float* dst, size_t n)
{
    hipLaunchKernelGGL(k_copy_f32, dim3(linear_blocks(n)), dim3(kNT), 0, st, src, dst, n);
}

}  // namespace fsim
