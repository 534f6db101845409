// fs_kernels.hip -- hand-written HIP kernels for gfx950 (MI355X) on the uniform-MAC projection path.
//
// Compiled with -ffp-contract=off: the FP64 kernels that rebuild face weights, multigrid
// conductances and the coarse Cholesky factor evaluate every expression in the same order as the
// reference UniformCPU (FluidGridCPU.cpp, PressureMultigrid.cpp), so their results are bit-identical
// to it.  The PCG working vectors and the multigrid smoothing run in FP32 with FP64 reductions.
//
// Cell kernels use 64 x 16 tiles and 256 threads (four 64-wide wavefronts, one row segment each),
// so every wavefront touches 64 consecutive cells of a row: loads and stores coalesce.
// Reductions write one partial per block; a one-workgroup finish kernel reduces the partials in a
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

// The four cells a thread owns in its 64 x 16 tile, as clamped (always valid) padded indices plus an
// in-range mask.  Kernels load through these indices unconditionally and predicate only the stores:
// a guarded load makes the compiler wait for it at the branch join, which serialises a thread's
// rows into one memory latency each.
struct Tile4 {
    int P[4];
    bool in[4];
};

__device__ __forceinline__ Tile4 tile4(int W, int H, int s)
{
    Tile4 t;
    const int x = blockIdx.x * kTX + (threadIdx.x & 63);
    const int xc = min(x, W - 1);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int y = blockIdx.y * kTY + (threadIdx.x >> 6) + 4 * k;
        t.in[k] = x < W && y < H;
        t.P[k] = pidx(xc, min(y, H - 1), s);
    }
    return t;
}

// ------------------------------------------------------------------------------------------
// Scalar block initialisation (a kernel, so it is safe inside captured graphs).
// ------------------------------------------------------------------------------------------
__global__ void k_init_scalars(PcgScalars* S, double tol, int max_iterations, double rel_tol)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        PcgScalars z{};
        z.tol = tol;
        z.rel_tol = rel_tol;
        z.max_iterations = max_iterations;
        z.status = kRunning;
        z.restart = 1;
        *S = z;
    }
}

__global__ void k_merge_failure(const PcgScalars* from, PcgScalars* S)
{
    if (threadIdx.x == 0 && blockIdx.x == 0 && from->mg_failed) {
        S->mg_failed = 1;
        if (S->status == kRunning)
            S->status = kFailed;
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

// Conductance of the face right of / below each cell: 2 / (rho_a + rho_b), zero at walls and, with a
// solid mask, zero on every face that touches a solid cell.
__global__ __launch_bounds__(kNT) void k_face_weights(GridDev g)
{
    const int W = g.W, H = g.H;
    const int x = blockIdx.x * kTX + (threadIdx.x & 63);
    const int xc = min(x, W - 1), xr = min(x + 1, W - 1);
    double c[4], r[4], d[4];
    bool open_r[4], open_d[4];
    int yy[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int y = blockIdx.y * kTY + (threadIdx.x >> 6) + 4 * k;
        yy[k] = y;
        const int yc = min(y, H - 1), yd = min(y + 1, H - 1);
        c[k] = g.density[yc * W + xc];
        r[k] = g.density[yc * W + xr];
        d[k] = g.density[yd * W + xc];
        open_r[k] = open_d[k] = true;
        if (g.solid) {
            const bool sc = g.solid[yc * W + xc] != 0;
            open_r[k] = !(sc || g.solid[yc * W + xr] != 0);
            open_d[k] = !(sc || g.solid[yd * W + xc] != 0);
        }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int y = yy[k];
        if (x < W && y < H) {
            const int P = pidx(x, y, g.s);
            g.right64[P] = x + 1 < W && open_r[k] ? 2.0 / (c[k] + r[k]) : 0.0;
            g.down64[P] = y + 1 < H && open_d[k] ? 2.0 / (c[k] + d[k]) : 0.0;
        }
    }
}

// PressureMultigrid Level::rebuildDiagonal (PressureMultigrid.cpp:104-117) in FP64, then FP32 copies.
__global__ __launch_bounds__(kNT) void k_level_diag(int w, int h, const double* right64, const double* down64,
    float* right32, float* down32, float* invd32)
{
    const int s = w + 2;
    const Tile4 T = tile4(w, h, s);
    double rl[4], rr[4], du[4], dd[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int i = T.P[k];
        rl[k] = right64[i - 1];
        rr[k] = right64[i];
        du[k] = down64[i - s];
        dd[k] = down64[i];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (T.in[k]) {
            const int i = T.P[k];
            level_diag_values(rl[k], rr[k], du[k], dd[k], invd32[i], right32[i], down32[i]);
        }
    }
}

// PressureMultigrid::coarsen (PressureMultigrid.cpp:142-202), one thread per coarse cell.
__global__ __launch_bounds__(kNT) void k_coarsen(int fw, int fh, const double* fcx, const double* fcy,
    const double* fright, const double* fdown, int cw, int ch, const double* ccx, const double* ccy,
    double* cright, double* cdown)
{
    const int t = blockIdx.x * kNT + threadIdx.x;
    if (t >= cw * ch)
        return;
    const int y = t / cw, x = t - y * cw;
    coarsen_cell(x, y, fw, fh, fcx, fcy, fright, fdown, cw, ch, ccx, ccy, cright, cdown);
}

// ------------------------------------------------------------------------------------------
// Projection set-up (FluidGridCPU.cpp:1050-1090, :1728-1756), FP64.
// ------------------------------------------------------------------------------------------

// Zero wall velocities, form rhs = -div * h^2 / dt, sum it.
__global__ __launch_bounds__(kNT) void k_rhs(GridDev g, double scale)
{
    __shared__ double lds[kNT / 64];
    const int W = g.W, H = g.H;
    const int x = blockIdx.x * kTX + (threadIdx.x & 63);
    const int xc = min(x, W - 1);
    double uL[4], uR[4], vT[4], vB[4];
    unsigned closed[4];   // solid mask: bit 0 left, 1 right, 2 top, 3 bottom face touches a solid cell
    int yy[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int y = blockIdx.y * kTY + (threadIdx.x >> 6) + 4 * k;
        yy[k] = y;
        const int yc = min(y, H - 1);
        const double* hz = g.u + (size_t)yc * (W + 1);
        const double* vt = g.v + (size_t)yc * W;
        uL[k] = hz[xc];
        uR[k] = hz[xc + 1];
        vT[k] = vt[xc];
        vB[k] = vt[xc + W];
        closed[k] = 0u;
        if (g.solid) {
            const unsigned char* m = g.solid + (size_t)yc * W;
            const bool sc = m[xc] != 0;
            closed[k] = (sc || m[max(xc - 1, 0)] ? 1u : 0u) | (sc || m[min(xc + 1, W - 1)] ? 2u : 0u) |
                        (sc || m[xc - (yc > 0 ? W : 0)] ? 4u : 0u) | (sc || m[xc + (yc + 1 < H ? W : 0)] ? 8u : 0u);
        }
    }
    if (g.S->status != kRunning)
        return;
    double sum = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int y = yy[k];
        if (x < W && y < H) {
            // walls are zeroed and read as zero (FluidGridCPU.cpp:1050-1061); faces touching a solid
            // cell likewise (both cells of such a face write the same zero and neither reads it)
            double* hz = g.u + (size_t)y * (W + 1);
            double* vt = g.v + (size_t)y * W;
            const bool cl = x == 0 || (closed[k] & 1u), cr = x + 1 == W || (closed[k] & 2u);
            const bool ct = y == 0 || (closed[k] & 4u), cb = y + 1 == H || (closed[k] & 8u);
            const double l = cl ? 0.0 : uL[k];
            const double r = cr ? 0.0 : uR[k];
            const double t = ct ? 0.0 : vT[k];
            const double b = cb ? 0.0 : vB[k];
            if (cl) hz[x] = 0.0;
            if (cr) hz[x + 1] = 0.0;
            if (ct) vt[x] = 0.0;
            if (cb) vt[x + W] = 0.0;
            const double du = r - l;
            const double dv = b - t;
            const double rhs = -(du + dv) / g.h * scale;
            g.b[pidx(x, y, g.s)] = rhs;
            sum += rhs;
        }
    }
    const double bs = block_sum<kNT>(sum, lds);
    if (threadIdx.x == 0)
        g.partials[block_id()] = bs;
}

// Remove the rhs mean and find max |rhs| (the early-out test of solvePressure).
__global__ __launch_bounds__(kNT) void k_rhs_center(GridDev g)
{
    __shared__ double lds[kNT / 64];
    const Tile4 T = tile4(g.W, g.H, g.s);
    double bv[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
        bv[k] = g.b[T.P[k]];
    if (g.S->status != kRunning)
        return;
    const double mean = g.S->b_mean;
    double m = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (T.in[k]) {
            const double v = bv[k] - mean;
            g.b[T.P[k]] = v;
            m = fmax(m, abs_or_inf(v));
        }
    }
    const double bm = block_max<kNT>(m, lds);
    if (threadIdx.x == 0)
        g.partials[block_id()] = bm;
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
    const Tile4 T = tile4(g.W, g.H, g.s);
    double pv[4], qv[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        pv[k] = g.p[T.P[k]];
        qv[k] = g.p_prev[T.P[k]];
    }
    if (g.S->status != kRunning)
        return;
    double sum = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (T.in[k]) {
            const double next = 2.0 * pv[k] - qv[k];
            g.p[T.P[k]] = next;
            g.p_prev[T.P[k]] = pv[k];
            sum += next;
        }
    }
    const double bs = block_sum<kNT>(sum, lds);
    if (threadIdx.x == 0)
        g.partials[block_id()] = bs;
}

__global__ __launch_bounds__(kNT) void k_sub_mean_p(GridDev g, int final_pass)
{
    const Tile4 T = tile4(g.W, g.H, g.s);
    double pv[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
        pv[k] = g.p[T.P[k]];
    const int st = g.S->status;
    if (final_pass ? !(st == kConverged || st == kStopped) : st != kRunning)
        return;
    const double mean = g.S->p_mean;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (T.in[k])
            g.p[T.P[k]] = pv[k] - mean;
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
    const Tile4 T = tile4(g.W, g.H, s);
    double r[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int i = T.P[k];
        const double c = g.p[i];
        const double product = (g.right64[i - 1] * (c - g.p[i - 1]) + g.right64[i] * (c - g.p[i + 1])) +
                               (g.down64[i - s] * (c - g.p[i - s]) + g.down64[i] * (c - g.p[i + s]));
        r[k] = g.b[i] - product;
    }
    double sum = 0.0, m = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (T.in[k]) {
            g.r[T.P[k]] = (float)r[k];
            if (mode == 0)
                g.x[T.P[k]] = 0.0f;
            sum += r[k];
            m = fmax(m, fabs(r[k]));
        }
    }
    const double bs = block_sum<kNT>(sum, lds);
    const double bm = block_max<kNT>(m, lds);
    if (threadIdx.x == 0) {
        g.partials[2 * block_id()] = bs;
        g.partials[2 * block_id() + 1] = bm;
    }
}

// Fold the FP32 correction into the FP64 pressure before a true-residual check.
__global__ __launch_bounds__(kNT) void k_flush(GridDev g)
{
    if (g.S->status != kRunning || !g.S->verify)
        return;
    const Tile4 T = tile4(g.W, g.H, g.s);
    double pv[4];
    float xv[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        pv[k] = g.p[T.P[k]];
        xv[k] = g.x[T.P[k]];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (T.in[k]) {
            g.p[T.P[k]] = pv[k] + (double)xv[k];
            g.x[T.P[k]] = 0.0f;
        }
    }
}

// After the loop: fold the correction in and sum the pressure for the final mean removal.
__global__ __launch_bounds__(kNT) void k_final_flush_sum(GridDev g)
{
    __shared__ double lds[kNT / 64];
    const Tile4 T = tile4(g.W, g.H, g.s);
    double pv[4];
    float xv[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        pv[k] = g.p[T.P[k]];
        xv[k] = g.x[T.P[k]];
    }
    const int st = g.S->status;
    if (!(st == kConverged || st == kStopped))
        return;
    double sum = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (T.in[k]) {
            const double v = pv[k] + (double)xv[k];
            g.p[T.P[k]] = v;
            g.x[T.P[k]] = 0.0f;
            sum += v;
        }
    }
    const double bs = block_sum<kNT>(sum, lds);
    if (threadIdx.x == 0)
        g.partials[block_id()] = bs;
}

// u -= (dp/h) dt w_u ; v -= (dp/h) dt w_v (FluidGridCPU.cpp:1098-1118).  Each cell owns its left
// u face and its top v face.
__global__ __launch_bounds__(kNT) void k_correct_velocity(GridDev g, double dt)
{
    const int W = g.W, s = g.s;
    const Tile4 T = tile4(W, g.H, s);
    const int x = blockIdx.x * kTX + (threadIdx.x & 63);
    const int xc = min(x, W - 1);
    double pc[4], pl[4], pu[4], wl[4], wu[4], uv[4], vv[4];
    int yy[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int P = T.P[k];
        const int y = blockIdx.y * kTY + (threadIdx.x >> 6) + 4 * k;
        yy[k] = y;
        const int yc = min(y, g.H - 1);
        pc[k] = g.p[P];
        pl[k] = g.p[P - 1];
        pu[k] = g.p[P - s];
        wl[k] = g.right64[P - 1];
        wu[k] = g.down64[P - s];
        uv[k] = g.u[(size_t)yc * (W + 1) + xc];
        vv[k] = g.v[(size_t)yc * W + xc];
    }
    // after the solve only: a failed hierarchy leaves the velocity alone, and a solve still running (the
    // correction launched behind a batch of iterations that did not finish it) is corrected later
    const int st = g.S->status;
    if (st == kFailed || st == kRunning)
        return;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (!T.in[k])
            continue;
        const int y = yy[k];
        // u -= ((p - p_left)/h) dt w ; v -= ((p - p_up)/h) dt w  (FluidGridCPU.cpp:1098-1118)
        if (x >= 1)
            g.u[(size_t)y * (W + 1) + x] = uv[k] - ((pc[k] - pl[k]) / g.h) * dt * wl[k];
        if (y >= 1)
            g.v[(size_t)y * W + x] = vv[k] - ((pc[k] - pu[k]) / g.h) * dt * wu[k];
    }
}

// maxDivergence (FluidGridCPU.cpp:1126-1152): non-finite divergence reports infinity.
// when_stopped: only once the solve has stopped (the measurement closing a projection).
__global__ __launch_bounds__(kNT) void k_max_divergence(GridDev g, int when_stopped)
{
    __shared__ double lds[kNT / 64];
    if (when_stopped && g.S->status == kRunning)
        return;
    const int W = g.W, H = g.H;
    const int x = blockIdx.x * kTX + (threadIdx.x & 63);
    const int xc = min(x, W - 1);
    double dv[4];
    bool in[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int y = blockIdx.y * kTY + (threadIdx.x >> 6) + 4 * k;
        in[k] = x < W && y < H;
        const int yc = min(y, H - 1);
        const double* hz = g.u + (size_t)yc * (W + 1);
        const double* vt = g.v + (size_t)yc * W;
        dv[k] = ((hz[xc + 1] - hz[xc]) + (vt[xc + W] - vt[xc])) / g.h;
    }
    double m = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (in[k])
            m = fmax(m, abs_or_inf(dv[k]));
    const double bm = block_max<kNT>(m, lds);
    if (threadIdx.x == 0)
        g.partials[block_id()] = bm;
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
    const int s = g.s;
    const float* d = g.d;
    const Tile4 T = tile4(g.W, g.H, s);
    float c[4], q[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int i = T.P[k];
        c[k] = d[i];
        const float dl = d[i - 1], dr = d[i + 1], du = d[i - s], dd = d[i + s];
        const float wl = g.right32[i - 1], wr = g.right32[i], wu = g.down32[i - s], wd = g.down32[i];
        q[k] = (wl * (c[k] - dl) + wr * (c[k] - dr)) + (wu * (c[k] - du) + wd * (c[k] - dd));
    }
    if (g.S->status != kRunning)
        return;
    double dot = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (T.in[k]) {
            g.q[T.P[k]] = q[k];
            dot += (double)c[k] * (double)q[k];
        }
    }
    const double bs = block_sum<kNT>(dot, lds);
    if (threadIdx.x == 0)
        g.partials_ad[block_id()] = bs;
}

// Large grids: one workgroup reduces the per-block partials (stride components) into out[0 .. stride),
// with the same fixed-order finish_sum every consumer workgroup would otherwise repeat over all blocks.
__global__ __launch_bounds__(kNT) void k_fold_partials(const double* partials, int nb, int stride, double* out)
{
    __shared__ double lds[kNT / 64];
    for (int c = 0; c < stride; ++c) {
        const double v = finish_sum<kNT>(partials, nb, stride, c, lds);
        if (threadIdx.x == 0)
            out[c] = v;
    }
}

// x += alpha d ; r -= alpha q ; residual sum and max-norm; convergence test of the recurrence.
// Every workgroup first reduces the operator's d.q partials itself (same order, same value) and forms
// alpha; workgroup 0 records it, clears the restart flag the previous direction consumed, and stops
// the solve on a non-positive d.q.
__global__ __launch_bounds__(kNT) void k_pcg_update(GridDev g)
{
    __shared__ double lds[kNT / 64];
    const Tile4 T = tile4(g.W, g.H, g.s);
    float xv[4], dv[4], rv[4], qv[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int i = T.P[k];
        xv[k] = g.x[i];
        dv[k] = g.d[i];
        rv[k] = g.r[i];
        qv[k] = g.q[i];
    }
    if (g.S->status != kRunning)
        return;
    const double dq = finish_sum<kNT>(g.fold_ad, g.fold_ad_n, 1, 0, lds);
    const double rz = g.S->rz_slot[g.S->iterations & 1];
    bool stop = dq <= 0.0 || !isfinite(dq);
    const double alpha64 = stop ? 0.0 : rz / dq;
    stop = stop || !isfinite(alpha64);
    if (block_id() == 0 && threadIdx.x == 0) {
        g.S->dq = dq;
        g.S->alpha = alpha64;
        g.S->restart = 0;
        if (stop)
            g.S->status = kStopped;
    }
    if (stop)
        return;
    const float alpha = (float)alpha64;
    double sum = 0.0, m = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (T.in[k]) {
            const int i = T.P[k];
            g.x[i] = xv[k] + alpha * dv[k];
            const float r = rv[k] - alpha * qv[k];
            g.r[i] = r;
            sum += (double)r;
            m = fmax(m, (double)fabsf(r));
        }
    }
    const double bs = block_sum<kNT>(sum, lds);
    const double bm = block_max<kNT>(m, lds);
    if (threadIdx.x == 0) {
        g.partials[2 * block_id()] = bs;
        g.partials[2 * block_id() + 1] = bm;
    }
}

// Diagonal preconditioner z = (r - mean) / diag.
__global__ __launch_bounds__(kNT) void k_precond_diag(GridDev g)
{
    const Tile4 T = tile4(g.W, g.H, g.s);
    float rv[4], iv[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        rv[k] = g.r[T.P[k]];
        iv[k] = g.invd32[T.P[k]];
    }
    if (g.S->status != kRunning)
        return;
    const double mean = g.S->r_mean;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (T.in[k])
            g.zdiag[T.P[k]] = (float)(((double)rv[k] - mean) * (double)iv[k]);
}

// sum z, r.z and sum r in one pass; r.(z - mean z) = r.z - mean z * sum r; then beta.
// FLEX (flexible CG, for a preconditioner that is not exactly linear): also sum z.q and q, from which
// the direction kernel forms the Polak-Ribiere beta z_new.(r_new - r_old) / r_old.z_old with
// r_new - r_old = -alpha q (Notay 2000).
template <bool FLEX> __global__ __launch_bounds__(kNT) void k_pcg_rz(GridDev g, const float* z)
{
    __shared__ double lds[kNT / 64];
    constexpr int NP = FLEX ? 5 : 3;
    const Tile4 T = tile4(g.W, g.H, g.s);
    float zv[4], rv[4], qv[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        zv[k] = z[T.P[k]];
        rv[k] = g.r[T.P[k]];
        qv[k] = FLEX ? g.q[T.P[k]] : 0.f;
    }
    if (g.S->status != kRunning)
        return;
    double sz = 0.0, srz = 0.0, sr = 0.0, szq = 0.0, sq = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (T.in[k]) {
            const double zz = (double)zv[k], rr = (double)rv[k];
            sz += zz;
            srz += rr * zz;
            sr += rr;
            if (FLEX) {
                szq += zz * (double)qv[k];
                sq += (double)qv[k];
            }
        }
    }
    const double b0 = block_sum<kNT>(sz, lds);
    const double b1 = block_sum<kNT>(srz, lds);
    const double b2 = block_sum<kNT>(sr, lds);
    const double b3 = FLEX ? block_sum<kNT>(szq, lds) : 0.0;
    const double b4 = FLEX ? block_sum<kNT>(sq, lds) : 0.0;
    if (threadIdx.x == 0) {
        g.partials[NP * block_id()] = b0;
        g.partials[NP * block_id() + 1] = b1;
        g.partials[NP * block_id() + 2] = b2;
        if (FLEX) {
            g.partials[NP * block_id() + 3] = b3;
            g.partials[NP * block_id() + 4] = b4;
        }
    }
}

// d = (z - mean z) + beta d.  Every workgroup reduces the r.z partials itself (same order, same
// value): beta = next / previous r.z (zero after a restart), or with FLEX the Polak-Ribiere
// -alpha (z - mean z).q / previous r.z; workgroup 0 records the new r.z in the next slot and stops the
// solve at the iteration budget or on a non-positive r.z.
template <bool FLEX> __global__ __launch_bounds__(kNT) void k_pcg_direction(GridDev g, const float* z)
{
    __shared__ double lds[kNT / 64];
    const Tile4 T = tile4(g.W, g.H, g.s);
    float zv[4], dv[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        zv[k] = z[T.P[k]];
        dv[k] = g.d[T.P[k]];
    }
    if (g.S->status != kRunning)
        return;
    constexpr int NP = FLEX ? 5 : 3;
    const double tz = finish_sum<kNT>(g.fold_rz, g.fold_rz_n, NP, 0, lds);
    const double trz = finish_sum<kNT>(g.fold_rz, g.fold_rz_n, NP, 1, lds);
    const double tr = finish_sum<kNT>(g.fold_rz, g.fold_rz_n, NP, 2, lds);
    const double n = (double)g.W * g.H;
    const double zmean = tz / n;
    const double next = trz - zmean * tr;
    const int c = g.S->iterations;
    double numerator = next;
    if (FLEX) {
        const double tzq = finish_sum<kNT>(g.fold_rz, g.fold_rz_n, NP, 3, lds);
        const double tq = finish_sum<kNT>(g.fold_rz, g.fold_rz_n, NP, 4, lds);
        numerator = -g.S->alpha * (tzq - zmean * tq);
    }
    const double beta64 = g.S->restart ? 0.0 : numerator / g.S->rz_slot[(c + 1) & 1];
    const bool stop = !isfinite(beta64) || !(c < g.S->max_iterations && next > 0.0 && isfinite(next));
    if (block_id() == 0 && threadIdx.x == 0) {
        g.S->z_mean = zmean;
        g.S->beta = beta64;
        if (isfinite(beta64))
            g.S->rz_slot[c & 1] = next;
        if (stop)
            g.S->status = kStopped;
    }
    if (stop)
        return;
    const float mean = (float)zmean;
    const float beta = (float)beta64;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (T.in[k]) {
            const float zz = zv[k] - mean;
            g.d[T.P[k]] = beta64 == 0.0 ? zz : zz + beta * dv[k];
        }
    }
}

// ------------------------------------------------------------------------------------------
// Reduction finish: one workgroup reads the producer's per-block partials in block order and
// updates the scalar block.  Running it as its own launch keeps cross-XCD visibility at kernel
// boundaries instead of a device-scope fence in every producer block.
// ------------------------------------------------------------------------------------------
enum FinishOp : int {
    kFinRhs = 0, kFinRhsCenter, kFinPredict, kFinTrueRes0, kFinTrueRes1, kFinFinalSum, kFinDiv, kFinAd,
    kFinUpdate, kFinRz, kFinDivDone
};

__global__ __launch_bounds__(kNT) void k_finish(GridDev g, int op, int nb)
{
    __shared__ double lds[kNT / 64];
    PcgScalars* S = g.S;
    const int st = S->status;
    bool run;
    switch (op) {
    case kFinTrueRes1: run = st == kRunning && S->verify; break;
    case kFinFinalSum: run = st == kConverged || st == kStopped; break;
    case kFinDiv: run = true; break;
    case kFinDivDone: run = st != kRunning; break;
    default: run = st == kRunning; break;
    }
    if (!run)
        return;
    const double n = (double)g.W * g.H;
    const double* P = g.partials;
    switch (op) {
    case kFinRhs: {
        const double t = finish_sum<kNT>(P, nb, 1, 0, lds);
        if (threadIdx.x == 0)
            S->b_mean = t / n;
        break;
    }
    case kFinRhsCenter: {
        const double t = finish_max<kNT>(P, nb, 1, 0, lds);
        if (threadIdx.x == 0) {
            S->bmax = t;
            if (S->rel_tol > 0.0)
                S->tol = S->rel_tol * t;
            if (t <= S->tol)
                S->status = kZeroRhs;
        }
        break;
    }
    case kFinPredict:
    case kFinFinalSum: {
        const double t = finish_sum<kNT>(P, nb, 1, 0, lds);
        if (threadIdx.x == 0)
            S->p_mean = t / n;
        break;
    }
    case kFinTrueRes0:
    case kFinTrueRes1: {
        const double total = finish_sum<kNT>(P, nb, 2, 0, lds);
        double maximum = finish_max<kNT>(P, nb, 2, 1, lds);
        if (threadIdx.x == 0) {
            if (!isfinite(total))
                maximum = INFINITY;
            S->true_rmax = maximum;
            S->r_sum = total;
            if (op == kFinTrueRes0) {
                S->rmax = maximum;
                S->initial_rmax = maximum;
                S->window_rmax = maximum;
                S->r_mean = total / n;
                if (maximum <= S->tol || S->max_iterations == 0)
                    S->status = kInitialOk;
            } else {
                if (maximum <= S->tol) {
                    S->status = kConverged;
                } else {
                    S->r_mean = total / n;
                    if (!S->replace || maximum > kDriftRatio * S->rmax) {
                        S->restart = 1;
                        S->restarts += 1;
                    }
                    S->rmax = maximum;
                    S->window_rmax = maximum;
                }
                S->verify = 0;
                S->replace = 0;
            }
        }
        break;
    }
    case kFinDiv:
    case kFinDivDone: {
        const double t = finish_max<kNT>(P, nb, 1, 0, lds);
        if (threadIdx.x == 0)
            S->div_max = t;
        break;
    }
    case kFinAd: {
        const double total = finish_sum<kNT>(P, nb, 1, 0, lds);
        if (threadIdx.x == 0) {
            S->dq = total;
            if (total <= 0.0 || !isfinite(total)) {
                S->status = kStopped;
            } else {
                const double alpha = S->rz / total;
                if (!isfinite(alpha))
                    S->status = kStopped;
                S->alpha = alpha;
            }
        }
        break;
    }
    case kFinUpdate: {
        const double total = finish_sum<kNT>(P, nb, 2, 0, lds);
        double maximum = finish_max<kNT>(P, nb, 2, 1, lds);
        if (threadIdx.x == 0) {
            if (!isfinite(total))
                maximum = INFINITY;
            S->iterations += 1;
            S->rmax = maximum;
            S->r_sum = total;
            S->r_mean = total / n;
            if (maximum <= S->tol) {
                S->verify = 1;
            } else if (maximum <= kRefineDrop * S->window_rmax || S->iterations % kReplaceEvery == 0) {
                S->verify = 1;
                S->replace = 1;
            }
        }
        break;
    }
    case kFinRz: {
        const double tz = finish_sum<kNT>(P, nb, 3, 0, lds);
        const double trz = finish_sum<kNT>(P, nb, 3, 1, lds);
        const double tr = finish_sum<kNT>(P, nb, 3, 2, lds);
        if (threadIdx.x == 0) {
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
        }
        break;
    }
    default:
        break;
    }
}

// Several FP64 device-to-device copies in one launch (blockIdx.y selects the copy).
__global__ __launch_bounds__(kNT) void k_copy_many(CopyList list)
{
    const CopyList::Item it = list.item[blockIdx.y];
    for (size_t i = (size_t)blockIdx.x * kNT + threadIdx.x; i < it.n; i += (size_t)gridDim.x * kNT)
        it.dst[i] = it.src[i];
}

__global__ void k_copy_f32(const float* src, float* dst, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// writeColorPixels (FluidGridCPU.cpp:1826-1860): per cell, RGBA with the component colours weighted by
// their volume fractions mass / (density h^2), in the reference's precision and order (FP64 fraction,
// each channel term rounded to FP32 and summed in FP32, liquid components before gas components).
__global__ __launch_bounds__(kNT) void k_color(GridDev g, ColorTable t, int alpha_from_fraction, float4* out)
{
    const size_t N = (size_t)g.W * g.H;
    const double cell_volume = g.h * g.h;
    for (size_t i = (size_t)blockIdx.x * kNT + threadIdx.x; i < N; i += (size_t)gridDim.x * kNT) {
        float4 px = make_float4(0.0f, 0.0f, 0.0f, alpha_from_fraction ? (float)g.fraction[i] : 1.0f);
        for (int k = 0; k < g.nl + g.ng; ++k) {
            const double mass = k < g.nl ? g.lmass[(size_t)k * N + i] : g.gmass[(size_t)(k - g.nl) * N + i];
            const double volume_fraction = mass / (t.density[k] * cell_volume);
            px.x += (float)(volume_fraction * (double)t.rgb[k][0]);
            px.y += (float)(volume_fraction * (double)t.rgb[k][1]);
            px.z += (float)(volume_fraction * (double)t.rgb[k][2]);
        }
        out[i] = px;
    }
}

// commitEdits (FluidGridCPU.cpp:253-284) for resolved, duplicate-free cells: the cell becomes pure liquid
// or pure gas holding one component's mass, every other component's mass zero.
__global__ __launch_bounds__(kNT) void k_commit_edits(GridDev g, const EditCell* cells, int n)
{
    const size_t N = (size_t)g.W * g.H;
    for (int e = blockIdx.x * kNT + threadIdx.x; e < n; e += gridDim.x * kNT) {
        const EditCell c = cells[e];
        const size_t i = (size_t)c.y * g.W + c.x;
        const bool liquid = c.component < g.nl;
        g.fraction[i] = liquid ? 1.0 : 0.0;
        for (int k = 0; k < g.nl; ++k)
            g.lmass[(size_t)k * N + i] = k == c.component ? c.mass : 0.0;
        for (int k = 0; k < g.ng; ++k)
            g.gmass[(size_t)k * N + i] = g.nl + k == c.component ? c.mass : 0.0;
    }
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

void launch_init_scalars(hipStream_t st, PcgScalars* S, double tol, int max_iterations, double rel_tol)
{
    hipLaunchKernelGGL(k_init_scalars, dim3(1), dim3(64), 0, st, S, tol, max_iterations, rel_tol);
}

void launch_merge_failure(hipStream_t st, const PcgScalars* from, PcgScalars* S)
{
    hipLaunchKernelGGL(k_merge_failure, dim3(1), dim3(64), 0, st, from, S);
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
    hipLaunchKernelGGL(k_coarsen, dim3((cw * ch + kNT - 1) / kNT), dim3(kNT), 0, st, fw, fh, fcx, fcy, fright, fdown,
        cw, ch, ccx, ccy, cright, cdown);
}

void launch_rhs(hipStream_t st, const GridDev& g, double scale)
{
    hipLaunchKernelGGL(k_rhs, cell_grid(g.W, g.H), dim3(kNT), 0, st, g, scale);
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(kNT), 0, st, g, (int)kFinRhs, (int)cell_blocks(g.W, g.H));
}

void launch_rhs_center(hipStream_t st, const GridDev& g)
{
    hipLaunchKernelGGL(k_rhs_center, cell_grid(g.W, g.H), dim3(kNT), 0, st, g);
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(kNT), 0, st, g, (int)kFinRhsCenter, (int)cell_blocks(g.W, g.H));
}

void launch_zero_pressure(hipStream_t st, const GridDev& g)
{
    hipLaunchKernelGGL(k_zero_pressure, cell_grid(g.W, g.H), dim3(kNT), 0, st, g);
}

void launch_predict(hipStream_t st, const GridDev& g)
{
    hipLaunchKernelGGL(k_predict, cell_grid(g.W, g.H), dim3(kNT), 0, st, g);
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(kNT), 0, st, g, (int)kFinPredict, (int)cell_blocks(g.W, g.H));
}

void launch_sub_mean_p(hipStream_t st, const GridDev& g, int final_pass)
{
    hipLaunchKernelGGL(k_sub_mean_p, cell_grid(g.W, g.H), dim3(kNT), 0, st, g, final_pass);
}

void launch_true_residual(hipStream_t st, const GridDev& g, int mode)
{
    hipLaunchKernelGGL(k_true_residual, cell_grid(g.W, g.H), dim3(kNT), 0, st, g, mode);
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(kNT), 0, st, g, mode == 0 ? (int)kFinTrueRes0 : (int)kFinTrueRes1,
        (int)cell_blocks(g.W, g.H));
}

void launch_flush(hipStream_t st, const GridDev& g)
{
    hipLaunchKernelGGL(k_flush, cell_grid(g.W, g.H), dim3(kNT), 0, st, g);
}

void launch_final_flush_sum(hipStream_t st, const GridDev& g)
{
    hipLaunchKernelGGL(k_final_flush_sum, cell_grid(g.W, g.H), dim3(kNT), 0, st, g);
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(kNT), 0, st, g, (int)kFinFinalSum, (int)cell_blocks(g.W, g.H));
}

void launch_correct_velocity(hipStream_t st, const GridDev& g, double dt)
{
    hipLaunchKernelGGL(k_correct_velocity, cell_grid(g.W, g.H), dim3(kNT), 0, st, g, dt);
}

void launch_color(hipStream_t st, const GridDev& g, const ColorTable& t, int alpha_from_fraction, float* rgba)
{
    hipLaunchKernelGGL(k_color, dim3(linear_blocks((size_t)g.W * g.H)), dim3(kNT), 0, st, g, t, alpha_from_fraction,
        (float4*)rgba);
}

void launch_commit_edits(hipStream_t st, const GridDev& g, const EditCell* cells, int n)
{
    if (n > 0)
        hipLaunchKernelGGL(k_commit_edits, dim3(linear_blocks((size_t)n)), dim3(kNT), 0, st, g, cells, n);
}

void launch_max_divergence(hipStream_t st, const GridDev& g, bool when_stopped)
{
    hipLaunchKernelGGL(k_max_divergence, cell_grid(g.W, g.H), dim3(kNT), 0, st, g, (int)when_stopped);
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(kNT), 0, st, g, (int)(when_stopped ? kFinDivDone : kFinDiv),
        (int)cell_blocks(g.W, g.H));
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
    if (g.fold_ad_n == 1)
        hipLaunchKernelGGL(k_fold_partials, dim3(1), dim3(kNT), 0, st, (const double*)g.partials_ad,
            (int)cell_blocks(g.W, g.H), 1, g.fold_ad);
}

void launch_pcg_update(hipStream_t st, const GridDev& g)
{
    hipLaunchKernelGGL(k_pcg_update, cell_grid(g.W, g.H), dim3(kNT), 0, st, g);
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(kNT), 0, st, g, (int)kFinUpdate, (int)cell_blocks(g.W, g.H));
}

void launch_pcg_rz(hipStream_t st, const GridDev& g, const float* z, int flexible)
{
    if (flexible)
        hipLaunchKernelGGL(k_pcg_rz<true>, cell_grid(g.W, g.H), dim3(kNT), 0, st, g, z);
    else
        hipLaunchKernelGGL(k_pcg_rz<false>, cell_grid(g.W, g.H), dim3(kNT), 0, st, g, z);
    if (g.fold_rz_n == 1)
        hipLaunchKernelGGL(k_fold_partials, dim3(1), dim3(kNT), 0, st, (const double*)g.partials,
            (int)cell_blocks(g.W, g.H), flexible ? 5 : 3, g.fold_rz);
}

void launch_pcg_direction(hipStream_t st, const GridDev& g, const float* z, int flexible)
{
    if (flexible)
        hipLaunchKernelGGL(k_pcg_direction<true>, cell_grid(g.W, g.H), dim3(kNT), 0, st, g, z);
    else
        hipLaunchKernelGGL(k_pcg_direction<false>, cell_grid(g.W, g.H), dim3(kNT), 0, st, g, z);
}

void launch_precond_diag(hipStream_t st, const GridDev& g)
{
    hipLaunchKernelGGL(k_precond_diag, cell_grid(g.W, g.H), dim3(kNT), 0, st, g);
}

void launch_copy_many(hipStream_t st, const CopyList& list)
{
    if (list.count <= 0)
        return;
    size_t most = 0;
    for (int k = 0; k < list.count; ++k)
        most = std::max(most, list.item[k].n);
    if (most == 0)
        return;
    const unsigned bx = (unsigned)std::min<size_t>((most + kNT - 1) / kNT, 2048);
    hipLaunchKernelGGL(k_copy_many, dim3(bx, list.count), dim3(kNT), 0, st, list);
}

void launch_copy_f32(hipStream_t st, const float* src, float* dst, size_t n)
{
    hipLaunchKernelGGL(k_copy_f32, dim3(linear_blocks(n)), dim3(kNT), 0, st, src, dst, n);
}

}  // namespace fsim
