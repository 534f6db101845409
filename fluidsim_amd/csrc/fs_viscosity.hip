// fs_viscosity.hip -- implicit viscosity on gfx950: backward Euler M (v - v0) + dt K v = 0 on the MAC faces.
//
// Algorithm of the reference's WebGPU solver (FluidSim/UniformWebGPU/ViscositySolverGPU.cpp:59-294,
// shaders FluidSim/UniformWebGPU/ViscosityKernels.h:10-630): M = face mass 0.5 h^2 (rho_a + rho_b),
// K = symmetric strain operator (normal stress 2 mu on cell centres, shear on interior vertices), CG
// preconditioned by a V(2,2) damped-Jacobi (weight 1/2) multigrid with a 16-sweep coarse solve, stop
// when the recomputed residual meets r M^-1 r <= tol^2 v0 M v0 + 1e-12 sum(M), checked every four
// iterations; a recurrence that claims convergence before the recomputed residual restarts CG from it.
//
// Layout and numerics: every level stores (w+1) x (h+1) slots of float2 (u face left of
// cell p, v face on top of cell p) and float4 coefficients (cell mu, vertex mu, u mass, v mass); reads
// clamp to the slot field like the reference's accessors.  The iterate x starts at v0 in FP32 (the
// reference's formulation: iterating on v - v0 instead loses three digits to cancellation when dt K
// dominates M) and dot products are FP64.  A solve that needs no iteration leaves u, v untouched.  Kernels read a device scalar block and
// exit at once when the solve has stopped, so a batch of four iterations is launched without host
// round trips; the host reads the block once per batch, as the reference does.
#include "fs_internal.hpp"
#include "fs_kernels.hpp"

#include <math.h>

#include <cmath>

#include <algorithm>
#include <mutex>
#include <stdexcept>
#include <vector>

namespace fsim {

namespace {

constexpr int kVX = 64;
constexpr int kVY = 4;
constexpr int kVT = 256;

struct VisScalars {
    double rz;        // r . z of the current direction
    double alpha;
    double beta;
    double target;    // tol^2 v0 M v0 + 1e-12 sum(M)
    double rnorm;     // recurrence r M^-1 r
    double energy;    // v0 M v0
    double mass_sum;
    double truth;     // recomputed r M^-1 r
    int steps;        // completed alpha steps since the last (re)start
    int bad;          // breakdown or non-finite value
    float dt;         // time step of this solve: the iteration kernels read it here, so one captured
                      // batch of iterations serves every dt
    int status;       // batched solve: result code (see viscosity_solve)
    int iterations;   // iterations run (batched solve; single solve: counted by k_vis_decide)
    int done;         // single solve, set on the device: 0 running, 1 converged (4 once unpacked),
                      // 2 breakdown, 3 budget spent, 5 no iteration needed, 6 non-finite input
    int restart;      // single solve: restart from the recomputed residual before the next batch
    int pad;
};

struct VLev {
    int w, h;
    float4* coef;
    float2* mat;    // cell material (rho, mu), w * h
    float2* b;      // rhs (levels > 0)
    float2* x;
    float2* xn;
    float2* r;
};

// Jacobi weights of a level's smoothing, per side: the reference's two sweeps of weight 1/2, or `sweeps`
// (2, 4 or 6) Chebyshev-weighted sweeps (fs_set_viscosity_smoother).  Pre-smoothing takes w[0..sweeps-1],
// post-smoothing the same weights in reverse.
constexpr int kVisMaxSweeps = 6;
struct VisSmooth {
    int sweeps;
    float w[kVisMaxSweeps];
};

__device__ __forceinline__ int sidx(int w, int h, int x, int y)
{
    return min(max(y, 0), h) * (w + 1) + min(max(x, 0), w);
}

__device__ __forceinline__ bool valid_u(int w, int h, int x, int y) { return x > 0 && x < w && y < h; }
__device__ __forceinline__ bool valid_v(int w, int h, int x, int y) { return y > 0 && y < h && x < w; }

__device__ __forceinline__ bool stopped(const VisScalars* S)
{
    return S->bad != 0 || S->rnorm <= S->target || S->done != 0;
}

// Index of slot (x, y) in an array: the global slot field clamped like the reference's accessors, or an
// LDS tile that holds the clamped values (see k_vis_smooth).  The stencils below are written once over
// it, so the fused and the per-sweep kernels evaluate the same expressions.
struct GlobalSlots {
    int w, h;
    __device__ __forceinline__ int operator()(int x, int y) const { return sidx(w, h, x, y); }
};

// shear and stressProduct (ViscosityKernels.h:26-44)
template <class I> __device__ __forceinline__ float shear_at(const float4* c, const float2* s, I at, int x, int y)
{
    const float2 p = s[at(x, y)];
    return c[at(x, y)].y * ((p.x - s[at(x, y - 1)].x) + (p.y - s[at(x - 1, y)].y));
}

// product (ViscosityKernels.h:53-57): M s + dt K s on valid faces.
template <class I>
__device__ __forceinline__ float2 product_at(const float4* c, const float2* s, I at, int w, int h, int x, int y,
    float dt)
{
    const int p = at(x, y);
    const float2 cv = s[p];
    const float4 a = c[p];
    const float t = shear_at(c, s, at, x, y);
    const float nu = 2.0f * (c[at(x - 1, y)].x * (cv.x - s[at(x - 1, y)].x) + a.x * (cv.x - s[at(x + 1, y)].x));
    const float nv = 2.0f * (c[at(x, y - 1)].x * (cv.y - s[at(x, y - 1)].y) + a.x * (cv.y - s[at(x, y + 1)].y));
    const float ku = nu + (t - shear_at(c, s, at, x, y + 1));
    const float kv = nv + (t - shear_at(c, s, at, x + 1, y));
    float2 out;
    out.x = valid_u(w, h, x, y) ? a.z * cv.x + dt * ku : 0.0f;
    out.y = valid_v(w, h, x, y) ? a.w * cv.y + dt * kv : 0.0f;
    return out;
}

// diagonal (ViscosityKernels.h:17-24)
template <class I> __device__ __forceinline__ float2 diagonal_at(const float4* c, I at, int x, int y, float dt)
{
    const float4 a = c[at(x, y)];
    float2 d;
    d.x = a.z + dt * (2.0f * (a.x + c[at(x - 1, y)].x) + a.y + c[at(x, y + 1)].y);
    d.y = a.w + dt * (2.0f * (a.x + c[at(x, y - 1)].x) + a.y + c[at(x + 1, y)].y);
    return d;
}

__device__ float2 product(const float4* c, const float2* s, int w, int h, int x, int y, float dt)
{
    return product_at(c, s, GlobalSlots{w, h}, w, h, x, y, dt);
}

__device__ __forceinline__ float2 diagonal(const float4* c, int w, int h, int x, int y, float dt)
{
    return diagonal_at(c, GlobalSlots{w, h}, x, y, dt);
}

#define VIS_SLOT(L)                                            \
    const int x = blockIdx.x * kVX + (threadIdx.x & 63);       \
    const int y = blockIdx.y * kVY + (threadIdx.x >> 6);       \
    const bool inside = x <= (L).w && y <= (L).h;              \
    const int p = inside ? y * ((L).w + 1) + x : 0;

__device__ __forceinline__ unsigned vblock() { return blockIdx.y * gridDim.x + blockIdx.x; }

// ---- material and coefficients ----

// FluidGridGPU's blend + pack (FluidGridGPU.cpp:173-218): rho = sum(m) / h^2,
// mu = sum(m_k mu_k / rho_k) / sum(m_k / rho_k), FP64.
__global__ __launch_bounds__(kVT) void k_vis_blend(int N, double area, int nl, int ng, const double* lm,
    const double* gm, VisMixture mix, double* material)
{
    for (int i = blockIdx.x * kVT + threadIdx.x; i < N; i += gridDim.x * kVT) {
        double mass = 0.0, volume = 0.0, mu_volume = 0.0;
        for (int k = 0; k < nl; ++k) {
            const double m = lm[(size_t)k * N + i];
            mass += m;
            volume += m / mix.density[k];
            mu_volume += m * (mix.viscosity[k] / mix.density[k]);
        }
        for (int k = 0; k < ng; ++k) {
            const double m = gm[(size_t)k * N + i];
            mass += m;
            volume += m / mix.density[kMaxComponents + k];
            mu_volume += m * (mix.viscosity[kMaxComponents + k] / mix.density[kMaxComponents + k]);
        }
        material[2 * i] = mass / area;
        material[2 * i + 1] = mu_volume / volume;
    }
}

__global__ __launch_bounds__(kVT) void k_vis_material(int N, const double* material, float2* mat)
{
    for (int i = blockIdx.x * kVT + threadIdx.x; i < N; i += gridDim.x * kVT)
        mat[i] = make_float2((float)material[2 * i], (float)material[2 * i + 1]);
}

// "viscosity coarse material" (ViscosityKernels.h:115-135)
__global__ __launch_bounds__(kVT) void k_vis_coarsen(VLev F, VLev C)
{
    const int x = blockIdx.x * kVX + (threadIdx.x & 63);
    const int y = blockIdx.y * kVY + (threadIdx.x >> 6);
    if (x >= C.w || y >= C.h)
        return;
    float2 v = make_float2(0.0f, 0.0f);
    float n = 0.0f;
    for (int dy = 0; dy < 2; ++dy)
        for (int dx = 0; dx < 2; ++dx) {
            const int qx = 2 * x + dx, qy = 2 * y + dy;
            if (qx < F.w && qy < F.h) {
                const float2 m = F.mat[qy * F.w + qx];
                v.x += m.x;
                v.y += m.y;
                n += 1.0f;
            }
        }
    C.mat[y * C.w + x] = make_float2(v.x / n, v.y / n);
}

// "viscosity coefficients" (ViscosityKernels.h:137-152)
__global__ __launch_bounds__(kVT) void k_vis_coef(VLev L, float spacing)
{
    VIS_SLOT(L)
    if (!inside)
        return;
    auto m = [&](int cx, int cy) { return L.mat[min(max(cy, 0), L.h - 1) * L.w + min(max(cx, 0), L.w - 1)]; };
    const float2 c = m(x, y);
    float vertex = 0.0f;
    if (x > 0 && y > 0 && x < L.w && y < L.h)
        vertex = 0.25f * (c.y + m(x - 1, y).y + m(x, y - 1).y + m(x - 1, y - 1).y);
    const float half_area = 0.5f * spacing * spacing;
    L.coef[p] = make_float4(c.y, vertex, half_area * (c.x + m(x - 1, y).x), half_area * (c.x + m(x, y - 1).x));
}

// ---- set-up of the correction system ----

// "viscosity initial velocity and RHS" (ViscosityKernels.h:154-163): v0 on valid faces (FP32).
__global__ __launch_bounds__(kVT) void k_vis_pack(VLev L, const double* u, const double* v, float2* v0)
{
    VIS_SLOT(L)
    if (!inside)
        return;
    const int W = L.w, H = L.h;
    const double su = u[min(y, H - 1) * (W + 1) + x];
    const double sv = v[y * W + min(x, W - 1)];
    v0[p] = make_float2(valid_u(W, H, x, y) ? (float)su : 0.0f, valid_v(W, H, x, y) ? (float)sv : 0.0f);
}

// r0 = M v0 - A v0 = -dt K v0 (the rhs of the correction system) and the mode-0 reduction:
// r0 M^-1 r0, v0 M v0, sum of masses, non-finite flag (ViscosityKernels.h:395-470).
__global__ __launch_bounds__(kVT) void k_vis_initial(VLev L, const float2* v0, float2* r0, float dt, double* partials)
{
    __shared__ double lds[kVT / 64];
    VIS_SLOT(L)
    double a = 0.0, b = 0.0, m = 0.0, bad = 0.0;
    if (inside) {
        const float4 c = L.coef[p];
        const float2 q = product(L.coef, v0, L.w, L.h, x, y, dt);
        const float2 s = v0[p];
        float2 r;
        r.x = valid_u(L.w, L.h, x, y) ? c.z * s.x - q.x : 0.0f;
        r.y = valid_v(L.w, L.h, x, y) ? c.w * s.y - q.y : 0.0f;
        r0[p] = r;
        a = (double)r.x / c.z * r.x + (double)r.y / c.w * r.y;
        b = (double)c.z * s.x * s.x + (double)c.w * s.y * s.y;
        m = (double)c.z + (double)c.w;
        if (!isfinite(a) || !isfinite(b) || !isfinite(m)) {
            a = b = m = 0.0;
            bad = 1.0;
        }
    }
    const unsigned nb = gridDim.x * gridDim.y;
    const double sa = block_sum<kVT>(a, lds);
    const double sb = block_sum<kVT>(b, lds);
    const double sm = block_sum<kVT>(m, lds);
    const double sf = block_max<kVT>(bad, lds);
    if (threadIdx.x == 0) {
        partials[vblock()] = sa;
        partials[nb + vblock()] = sb;
        partials[2 * nb + vblock()] = sm;
        partials[3 * nb + vblock()] = sf;
    }
}

// ---- V-cycle ----

// "viscosity true residual" on a level (ViscosityKernels.h:187-195)
__global__ __launch_bounds__(kVT) void k_vis_residual(VLev L, const float2* b, const VisScalars* S)
{
    if (stopped(S))
        return;
    const float dt = S->dt;
    VIS_SLOT(L)
    if (!inside)
        return;
    const float2 q = product(L.coef, L.x, L.w, L.h, x, y, dt);
    const float2 rhs = b[p];
    L.r[p] = make_float2(valid_u(L.w, L.h, x, y) ? rhs.x - q.x : 0.0f, valid_v(L.w, L.h, x, y) ? rhs.y - q.y : 0.0f);
}

// Face-location bilinear transfer (ViscosityKernels.h:286-316).
__device__ __forceinline__ void transfer_location(int fx, int fy, int comp, int& bx, int& by, float& tx, float& ty)
{
    const float ox = comp == 1 ? 0.5f : 0.0f, oy = comp == 1 ? 0.0f : 0.5f;
    const float lx = ((float)fx + ox) * 0.5f - ox, ly = ((float)fy + oy) * 0.5f - oy;
    const float fx0 = floorf(lx), fy0 = floorf(ly);
    bx = (int)fx0;
    by = (int)fy0;
    tx = lx - fx0;
    ty = ly - fy0;
}

__device__ __forceinline__ void coarse_point(int x, int y, int comp, int cw, int ch, int& ox, int& oy)
{
    ox = min(max(x, 0), cw - (comp == 1 ? 1 : 0));
    oy = min(max(y, 0), ch - (comp == 1 ? 0 : 1));
}

__device__ float transfer_weight(int fx, int fy, int cx, int cy, int comp, int cw, int ch)
{
    int bx, by;
    float tx, ty;
    transfer_location(fx, fy, comp, bx, by, tx, ty);
    float result = 0.0f;
    for (int yy = 0; yy < 2; ++yy)
        for (int xx = 0; xx < 2; ++xx) {
            int px, py;
            coarse_point(bx + xx, by + yy, comp, cw, ch, px, py);
            if (px == cx && py == cy)
                result += (xx == 1 ? tx : 1.0f - tx) * (yy == 1 ? ty : 1.0f - ty);
        }
    return result;
}

// "viscosity transpose restriction" (ViscosityKernels.h:318-357): C.b = P^T F.r
__global__ __launch_bounds__(kVT) void k_vis_restrict(VLev F, VLev C, const VisScalars* S)
{
    if (stopped(S))
        return;
    VIS_SLOT(C)
    if (!inside)
        return;
    float val[2] = {0.0f, 0.0f};
    for (int comp = 0; comp < 2; ++comp) {
        if ((comp == 0 && (x == 0 || x == C.w || y == C.h)) || (comp == 1 && (y == 0 || y == C.h || x == C.w)))
            continue;
        const int ex = comp == 1 ? 2 : 1, ey = comp == 1 ? 1 : 2;
        for (int dy = -1; dy <= ey; ++dy)
            for (int dx = -1; dx <= ex; ++dx) {
                const int qx = 2 * x + dx, qy = 2 * y + dy;
                if (qx < 0 || qy < 0 || qx > F.w || qy > F.h)
                    continue;
                const float2 r = F.r[qy * (F.w + 1) + qx];
                val[comp] += transfer_weight(qx, qy, x, y, comp, C.w, C.h) * (comp == 0 ? r.x : r.y);
            }
    }
    C.b[p] = make_float2(val[0], val[1]);
}

// NS sweeps of the "viscosity symmetric smoother" (ViscosityKernels.h:214-236: damped Jacobi, weight 1/2 in
// the reference's two sweeps per side; the weights of VisSmooth otherwise) fused on an LDS tile:
// pre-smoothing from zero (result in L.x), or (POST) the "viscosity prolongation" F.x + P C.x on valid
// faces (ViscosityKernels.h:359-391) followed by the sweeps (result in L.xn).  The tile holds an NS-slot
// halo of coefficients, rhs and the starting solution, each slot loaded from its clamped position; sweep k
// covers the tile less k + 1 rings, the last one the interior, and every stencil read maps through the same
// clamp to the tile, so each sweep reads exactly the values a per-sweep kernel would read from the slot
// field (results identical to one launch per sweep).
constexpr int kSX = 32, kSY = 16;   // interior slots per workgroup

template <bool POST, int NS>
__global__ __launch_bounds__(kVT) void k_vis_smooth(VLev L, VLev C, const float2* b, const VisScalars* S,
    const VisSmooth M)
{
    constexpr int LX = kSX + 2 * NS, LY = kSY + 2 * NS, LN = LX * LY;   // NS 2: 36 x 20, NS 6: 44 x 28
    __shared__ float4 sc[LN];
    __shared__ float2 sb[LN];
    __shared__ float2 s0[LN];   // starting solution of the first sweep (POST: prolonged); even sweeps' input
    __shared__ float2 s1[LN];   // odd sweeps' input
    if (stopped(S))
        return;
    const float dt = S->dt;
    const int w = L.w, h = L.h;
    const int x0 = blockIdx.x * kSX - NS, y0 = blockIdx.y * kSY - NS;
    const auto at = [=](int x, int y) { return (min(max(y, 0), h) - y0) * LX + (min(max(x, 0), w) - x0); };
    for (int e = threadIdx.x; e < LN; e += kVT) {
        const int gx = x0 + e % LX, gy = y0 + e / LX;
        const int g = sidx(w, h, gx, gy);
        sc[e] = L.coef[g];
        sb[e] = b[g];
        if (POST)
            s0[e] = L.x[g];
    }
    __syncthreads();
    if (POST) {
        // prolongation on every in-field slot of the tile (interior and every halo ring)
        for (int e = threadIdx.x; e < LN; e += kVT) {
            const int x = x0 + e % LX, y = y0 + e / LX;
            if (x < 0 || y < 0 || x > w || y > h)
                continue;
            const float2 sv = s0[e];
            float val[2] = {sv.x, sv.y};
            for (int comp = 0; comp < 2; ++comp) {
                int bx, by;
                float tx, ty;
                transfer_location(x, y, comp, bx, by, tx, ty);
                for (int dy = 0; dy < 2; ++dy)
                    for (int dx = 0; dx < 2; ++dx) {
                        int px, py;
                        coarse_point(bx + dx, by + dy, comp, C.w, C.h, px, py);
                        const float2 c = C.x[py * (C.w + 1) + px];
                        val[comp] +=
                            (dx == 1 ? tx : 1.0f - tx) * (dy == 1 ? ty : 1.0f - ty) * (comp == 0 ? c.x : c.y);
                    }
            }
            s0[e] = make_float2(valid_u(w, h, x, y) ? val[0] : 0.0f, valid_v(w, h, x, y) ? val[1] : 0.0f);
        }
        __syncthreads();
    }
    // sweeps 0..NS-2 on the tile less k + 1 rings
#pragma unroll
    for (int k = 0; k + 1 < NS; ++k) {
        const float wk = M.w[POST ? NS - 1 - k : k];
        const float2* src = (k & 1) ? s1 : s0;
        float2* dst = (k & 1) ? s0 : s1;
        const int RX = LX - 2 * (k + 1), RY = LY - 2 * (k + 1);
        for (int e = threadIdx.x; e < RX * RY; e += kVT) {
            const int x = x0 + k + 1 + e % RX, y = y0 + k + 1 + e / RX;
            if (x < 0 || y < 0 || x > w || y > h)
                continue;
            const int i = at(x, y);
            const float2 d = diagonal_at(sc, at, x, y, dt);
            const float2 rhs = sb[i];
            float2 v;
            if (!POST && k == 0) {
                v.x = wk * rhs.x / d.x;
                v.y = wk * rhs.y / d.y;
            } else {
                const float2 q = product_at(sc, src, at, w, h, x, y, dt);
                const float2 sv = src[i];
                v.x = sv.x + wk * (rhs.x - q.x) / d.x;
                v.y = sv.y + wk * (rhs.y - q.y) / d.y;
            }
            dst[i] = make_float2(valid_u(w, h, x, y) ? v.x : 0.0f, valid_v(w, h, x, y) ? v.y : 0.0f);
        }
        __syncthreads();
    }
    // the last sweep on the interior, written out
    const float wl = M.w[POST ? 0 : NS - 1];
    const float2* src = ((NS - 1) & 1) ? s1 : s0;
    float2* out = POST ? L.xn : L.x;
    for (int e = threadIdx.x; e < kSX * kSY; e += kVT) {
        const int x = x0 + NS + e % kSX, y = y0 + NS + e / kSX;
        if (x > w || y > h)
            continue;
        const int i = at(x, y);
        const float2 d = diagonal_at(sc, at, x, y, dt);
        const float2 rhs = sb[i];
        const float2 q = product_at(sc, src, at, w, h, x, y, dt);
        const float2 sv = src[i];
        float2 v;
        v.x = sv.x + wl * (rhs.x - q.x) / d.x;
        v.y = sv.y + wl * (rhs.y - q.y) / d.y;
        out[y * (w + 1) + x] = make_float2(valid_u(w, h, x, y) ? v.x : 0.0f, valid_v(w, h, x, y) ? v.y : 0.0f);
    }
}

// "viscosity shared coarse solve" (ViscosityKernels.h:242-282): 16 Jacobi sweeps from zero in LDS.
constexpr int kCoarseSlots = 256;
__global__ __launch_bounds__(kCoarseSlots) void k_vis_coarse(VLev L, const float2* b, const VisScalars* S)
{
    __shared__ float2 vel[kCoarseSlots];
    __shared__ float4 cof[kCoarseSlots];
    if (stopped(S))
        return;
    const float dt = S->dt;
    const int lane = threadIdx.x;
    const int stride = L.w + 1;
    const int n = stride * (L.h + 1);
    const int x = lane % stride, y = lane / stride;
    const int src = lane < n ? lane : n - 1;
    vel[lane] = make_float2(0.0f, 0.0f);
    cof[lane] = L.coef[src];
    const float2 rhs = b[src];
    __syncthreads();
    for (int sweep = 0; sweep < 16; ++sweep) {
        float2 v = make_float2(0.0f, 0.0f);
        if (lane < n) {
            const float2 d = diagonal(cof, L.w, L.h, x, y, dt);
            const float2 q = product(cof, vel, L.w, L.h, x, y, dt);
            const float2 s = vel[lane];
            v.x = valid_u(L.w, L.h, x, y) ? s.x + 0.5f * (rhs.x - q.x) / d.x : 0.0f;
            v.y = valid_v(L.w, L.h, x, y) ? s.y + 0.5f * (rhs.y - q.y) / d.y : 0.0f;
        }
        __syncthreads();
        vel[lane] = v;
        __syncthreads();
    }
    if (lane < n)
        L.x[lane] = vel[lane];
}

// ---- CG ----

// mode-1 reduction: r . z, r M^-1 r (ViscosityKernels.h:395-470)
__global__ __launch_bounds__(kVT) void k_vis_rz(VLev L, const float2* r, const float2* z, const VisScalars* S,
    double* partials)
{
    __shared__ double lds[kVT / 64];
    if (stopped(S))
        return;
    VIS_SLOT(L)
    double a = 0.0, b = 0.0, bad = 0.0;
    if (inside) {
        const float4 c = L.coef[p];
        const float2 rr = r[p], zz = z[p];
        a = (double)rr.x * zz.x + (double)rr.y * zz.y;
        b = (double)rr.x / c.z * rr.x + (double)rr.y / c.w * rr.y;
        if (!isfinite(a) || !isfinite(b)) {
            a = b = 0.0;
            bad = 1.0;
        }
    }
    const unsigned nb = gridDim.x * gridDim.y;
    const double sa = block_sum<kVT>(a, lds);
    const double sb = block_sum<kVT>(b, lds);
    const double sf = block_max<kVT>(bad, lds);
    if (threadIdx.x == 0) {
        partials[vblock()] = sa;
        partials[nb + vblock()] = sb;
        partials[3 * nb + vblock()] = sf;
    }
}

__global__ __launch_bounds__(kVT) void k_vis_direction(VLev L, const float2* z, float2* d, const VisScalars* S)
{
    if (stopped(S))
        return;
    VIS_SLOT(L)
    if (!inside)
        return;
    const float2 zz = z[p];
    if (S->steps > 0) {
        const float beta = (float)S->beta;
        const float2 dd = d[p];
        d[p] = make_float2(zz.x + beta * dd.x, zz.y + beta * dd.y);
    } else {
        d[p] = zz;
    }
}

// q = A d and the mode-2 reduction d . q
__global__ __launch_bounds__(kVT) void k_vis_apply(VLev L, const float2* d, float2* q, const VisScalars* S,
    double* partials)
{
    __shared__ double lds[kVT / 64];
    if (stopped(S))
        return;
    const float dt = S->dt;
    VIS_SLOT(L)
    double a = 0.0, bad = 0.0;
    if (inside) {
        const float2 v = product(L.coef, d, L.w, L.h, x, y, dt);
        q[p] = v;
        const float2 dd = d[p];
        a = (double)dd.x * v.x + (double)dd.y * v.y;
        if (!isfinite(a)) {
            a = 0.0;
            bad = 1.0;
        }
    }
    const unsigned nb = gridDim.x * gridDim.y;
    const double sa = block_sum<kVT>(a, lds);
    const double sf = block_max<kVT>(bad, lds);
    if (threadIdx.x == 0) {
        partials[vblock()] = sa;
        partials[3 * nb + vblock()] = sf;
    }
}

__global__ __launch_bounds__(kVT) void k_vis_update(VLev L, float2* xs, float2* r, const float2* d, const float2* q,
    const VisScalars* S)
{
    if (stopped(S))
        return;
    VIS_SLOT(L)
    if (!inside)
        return;
    const float alpha = (float)S->alpha;
    const float2 dd = d[p], qq = q[p], xx = xs[p], rr = r[p];
    xs[p] = make_float2(xx.x + alpha * dd.x, xx.y + alpha * dd.y);
    r[p] = make_float2(rr.x - alpha * qq.x, rr.y - alpha * qq.y);
}

// recomputed residual t = M v0 - A x and t M^-1 t (mode 3)
__global__ __launch_bounds__(kVT) void k_vis_truth(VLev L, const float2* v0, const float2* xs, float2* t, float dt,
    double* partials, const VisScalars* S)
{
    __shared__ double lds[kVT / 64];
    if (S && S->done != 0)
        return;
    VIS_SLOT(L)
    double a = 0.0, bad = 0.0;
    if (inside) {
        const float4 c = L.coef[p];
        const float2 q = product(L.coef, xs, L.w, L.h, x, y, dt);
        const float2 s0 = v0[p];
        float2 v;
        v.x = valid_u(L.w, L.h, x, y) ? c.z * s0.x - q.x : 0.0f;
        v.y = valid_v(L.w, L.h, x, y) ? c.w * s0.y - q.y : 0.0f;
        t[p] = v;
        a = (double)v.x / c.z * v.x + (double)v.y / c.w * v.y;
        if (!isfinite(a)) {
            a = 0.0;
            bad = 1.0;
        }
    }
    const unsigned nb = gridDim.x * gridDim.y;
    const double sa = block_sum<kVT>(a, lds);
    const double sf = block_max<kVT>(bad, lds);
    if (threadIdx.x == 0) {
        partials[vblock()] = sa;
        partials[3 * nb + vblock()] = sf;
    }
}

enum VisFinish : int { kVisInitial = 0, kVisRz, kVisDq, kVisTruth };

// One workgroup reduces the partials in block order and updates the scalar block
// (the reference's "viscosity CG scalars", ViscosityKernels.h:490-535).
__global__ __launch_bounds__(kVT) void k_vis_finish(VisScalars* S, const double* partials, int nb, int op, double tol2,
    float dt)
{
    __shared__ double lds[kVT / 64];
    if ((op == kVisRz || op == kVisDq) && stopped(S))
        return;
    if (op == kVisTruth && S->done != 0)
        return;
    const double a = finish_sum<kVT>(partials, nb, 1, 0, lds);
    const double b = (op == kVisInitial || op == kVisRz) ? finish_sum<kVT>(partials + nb, nb, 1, 0, lds) : 0.0;
    const double m = op == kVisInitial ? finish_sum<kVT>(partials + 2 * nb, nb, 1, 0, lds) : 0.0;
    const double f = finish_max<kVT>(partials + 3 * nb, nb, 1, 0, lds);
    if (threadIdx.x != 0)
        return;
    VisScalars s = *S;
    const bool bad = f != 0.0;
    switch (op) {
    case kVisInitial:
        s = VisScalars{};
        s.dt = dt;
        s.energy = b;
        s.mass_sum = m;
        s.target = tol2 * b + 1e-12 * m;
        s.rnorm = a;
        s.truth = a;
        s.bad = bad ? 1 : 0;
        s.done = s.bad ? 6 : (s.rnorm <= s.target ? 5 : 0);
        break;
    case kVisRz:
        s.bad |= bad ? 1 : 0;
        s.rnorm = b;
        if (s.rnorm > s.target && !s.bad) {
            if (a <= 0.0 || !isfinite(a)) {
                s.bad = 1;
            } else {
                s.beta = s.steps > 0 ? a / s.rz : 0.0;
                s.rz = a;
            }
        }
        break;
    case kVisDq:
        if (bad || a <= 0.0 || !isfinite(a)) {
            s.bad = 1;
        } else {
            s.alpha = s.rz / a;
            s.steps += 1;
        }
        break;
    default:  // kVisTruth
        s.truth = bad ? INFINITY : a;
        if (bad)
            s.bad = 1;
        break;
    }
    *S = s;
}

// Restart from the recomputed residual ("viscosity refresh residual", ViscosityKernels.h:607-620), when
// k_vis_decide asked for it.
__global__ void k_vis_refresh(VisScalars* S)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        VisScalars s = *S;
        if (!s.restart)
            return;
        s.rnorm = s.truth;
        s.rz = s.alpha = s.beta = 0.0;
        s.steps = 0;
        s.restart = 0;
        *S = s;
    }
}

// The decision after a batch of `batch` iterations and the recomputed residual (ViscositySolverGPU.cpp:
// 264-294, taken on the device so that batches queue without a read-back): breakdown, convergence of the
// recomputed residual, a restart when only the recurrence claims convergence, and the iteration budget
// (`maximum` < 0: none).  A batch queued after the solve ended changes nothing (its kernels exit at once);
// the first such batch marks the converged result as unpacked.
__global__ void k_vis_decide(VisScalars* S, int batch, int maximum)
{
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    VisScalars s = *S;
    if (s.done != 0) {
        if (s.done == 1) {
            s.done = 4;
            *S = s;
        }
        return;
    }
    s.iterations += batch;
    if (s.bad) {
        s.done = 2;
    } else if (s.truth <= s.target) {
        s.done = 1;
    } else {
        if (s.rnorm <= s.target)
            s.restart = 1;
        if (maximum >= 0 && s.iterations >= maximum)
            s.done = 3;
    }
    *S = s;
}

// r = the recomputed residual t, when k_vis_decide asked for a restart.
__global__ __launch_bounds__(kVT) void k_vis_restart_copy(const VisScalars* S, float2* r, const float2* t, size_t n)
{
    if (!S->restart)
        return;
    for (size_t i = (size_t)blockIdx.x * kVT + threadIdx.x; i < n; i += (size_t)gridDim.x * kVT)
        r[i] = t[i];
}

// "viscosity staggered result" (ViscosityKernels.h:165-185): every face of u and v takes the solution
// (walls 0).  The output is the FP32 iterate, as in the reference: writing u0 + (x - v0) instead would
// keep the FP32 rounding noise of u0, which the stiff operator dt K amplifies in the residual.
// S (single solve): only once k_vis_decide has found the solve converged, and only the first time.
__global__ __launch_bounds__(kVT) void k_vis_unpack(int W, int H, const float2* xs, double* u, double* v,
    const VisScalars* S)
{
    if (S && S->done != 1)
        return;
    const int x = blockIdx.x * kVX + (threadIdx.x & 63);
    const int y = blockIdx.y * kVY + (threadIdx.x >> 6);
    if (x <= W && y < H) {
        const int i = y * (W + 1) + x;
        u[i] = valid_u(W, H, x, y) ? (double)xs[i].x : 0.0;
    }
    if (x < W && y <= H) {
        v[y * W + x] = valid_v(W, H, x, y) ? (double)xs[y * (W + 1) + x].y : 0.0;
    }
}

inline dim3 slot_grid(int w, int h) { return dim3((w + 1 + kVX - 1) / kVX, (h + 1 + kVY - 1) / kVY); }
inline unsigned slot_blocks(int w, int h) { return ((w + 1 + kVX - 1) / kVX) * ((h + 1 + kVY - 1) / kVY); }

void vhip(hipError_t e, const char* what)
{
    if (e != hipSuccess)
        throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

template <class T> struct VBuf {
    T* p = nullptr;
    size_t n = 0;
    bool owned = true;   // false: placed in memory the buffer does not own (a scratch arena)
    VBuf() = default;
    VBuf(const VBuf&) = delete;
    VBuf& operator=(const VBuf&) = delete;
    ~VBuf() { release(); }
    // Owned memory goes back to the device cache: the solver's owner drains its streams first.
    void release()
    {
        if (owned)
            device_free(p, n * sizeof(T));
        p = nullptr;
        n = 0;
        owned = true;
    }
    void borrow(T* ptr, size_t count)
    {
        release();
        p = ptr;
        n = count;
        owned = false;
    }
    size_t alloc(size_t count)
    {
        release();
        n = count;
        if (!count)
            return 0;
        p = (T*)device_alloc(count * sizeof(T));
        device_zero(p, count * sizeof(T));
        return count * sizeof(T);
    }
};

// ---------------------------------------------------------------------------------------------
// Batched solve: one 1024-thread workgroup runs one instance's whole CG loop (the iteration body of
// ViscositySolver::solve, every kernel a workgroup-strided loop followed by a barrier, the scalars
// held uniformly in registers), so a batch of uneven instances is one launch.
// ---------------------------------------------------------------------------------------------
constexpr int kBVT = 1024;
constexpr int kMaxVisLevels = 16;
constexpr int kBatchVisBudget = 500;   // iterations inside the batch; an instance that needs more is
                                       // solved again on its own by the caller

struct BatchVisDesc {
    int levels;
    int W, H;
    int maximum;          // the caller's budget
    VLev lev[kMaxVisLevels];
    const float2* v0;
    float2* r;
    float2* d;
    float2* q;
    float2* t;            // the iterate, starting at v0
    double* u;
    double* v;
    VisScalars* S;        // after the initial reduction; receives status and iterations
    VisSmooth M;
};

// FP64 sum over the workgroup in a fixed order, returned to every thread.
__device__ double wg_sum_all(double v, double* lds)
{
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0)
        lds[wid] = v;
    __syncthreads();
    double r = 0.0;
#pragma unroll
    for (int i = 0; i < kBVT / 64; ++i)
        r += lds[i];
    __syncthreads();
    return r;
}

__device__ __forceinline__ void slot_xy(int e, int w, int& x, int& y)
{
    y = e / (w + 1);
    x = e - y * (w + 1);
}

// One sweep src -> dst.
__device__ void bv_sweep(const VLev& L, const float2* b, const float2* src, float2* dst, float dt, float wt)
{
    const int n = (L.w + 1) * (L.h + 1);
    for (int e = threadIdx.x; e < n; e += kBVT) {
        int x, y;
        slot_xy(e, L.w, x, y);
        const float2 d = diagonal(L.coef, L.w, L.h, x, y, dt);
        const float2 rhs = b[e];
        const float2 q = product(L.coef, src, L.w, L.h, x, y, dt);
        const float2 sv = src[e];
        dst[e] = make_float2(valid_u(L.w, L.h, x, y) ? sv.x + wt * (rhs.x - q.x) / d.x : 0.0f,
            valid_v(L.w, L.h, x, y) ? sv.y + wt * (rhs.y - q.y) / d.y : 0.0f);
    }
    __syncthreads();
}

// Pre-smoothing from zero on a level (an even number of sweeps): xn = sweep(0), x = sweep(xn), xn = ...,
// ending in x.
__device__ void bv_pre(const VLev& L, const float2* b, float dt, const VisSmooth& M)
{
    const int n = (L.w + 1) * (L.h + 1);
    for (int e = threadIdx.x; e < n; e += kBVT) {
        int x, y;
        slot_xy(e, L.w, x, y);
        const float2 d = diagonal(L.coef, L.w, L.h, x, y, dt);
        const float2 rhs = b[e];
        L.xn[e] = make_float2(valid_u(L.w, L.h, x, y) ? M.w[0] * rhs.x / d.x : 0.0f,
            valid_v(L.w, L.h, x, y) ? M.w[0] * rhs.y / d.y : 0.0f);
    }
    __syncthreads();
    for (int k = 1; k < M.sweeps; ++k)
        bv_sweep(L, b, (k & 1) ? L.xn : L.x, (k & 1) ? L.x : L.xn, dt, M.w[k]);
}

// Post-smoothing after the prolongation into x (an even number of sweeps, the weights in reverse): xn =
// sweep(x), x = sweep(xn), ..., ending in x.
__device__ void bv_post(const VLev& L, const float2* b, float dt, const VisSmooth& M)
{
    for (int k = 0; k < M.sweeps; ++k)
        bv_sweep(L, b, (k & 1) ? L.xn : L.x, (k & 1) ? L.x : L.xn, dt, M.w[M.sweeps - 1 - k]);
}

__device__ void bv_residual_restrict(const VLev& F, const float2* b, const VLev& C, float dt)
{
    const int n = (F.w + 1) * (F.h + 1);
    for (int e = threadIdx.x; e < n; e += kBVT) {
        int x, y;
        slot_xy(e, F.w, x, y);
        const float2 q = product(F.coef, F.x, F.w, F.h, x, y, dt);
        const float2 rhs = b[e];
        F.r[e] = make_float2(valid_u(F.w, F.h, x, y) ? rhs.x - q.x : 0.0f, valid_v(F.w, F.h, x, y) ? rhs.y - q.y : 0.0f);
    }
    __syncthreads();
    const int nc = (C.w + 1) * (C.h + 1);
    for (int e = threadIdx.x; e < nc; e += kBVT) {
        int x, y;
        slot_xy(e, C.w, x, y);
        float val[2] = {0.0f, 0.0f};
        for (int comp = 0; comp < 2; ++comp) {
            if ((comp == 0 && (x == 0 || x == C.w || y == C.h)) || (comp == 1 && (y == 0 || y == C.h || x == C.w)))
                continue;
            const int ex = comp == 1 ? 2 : 1, ey = comp == 1 ? 1 : 2;
            for (int dy = -1; dy <= ey; ++dy)
                for (int dx = -1; dx <= ex; ++dx) {
                    const int qx = 2 * x + dx, qy = 2 * y + dy;
                    if (qx < 0 || qy < 0 || qx > F.w || qy > F.h)
                        continue;
                    const float2 r = F.r[qy * (F.w + 1) + qx];
                    val[comp] += transfer_weight(qx, qy, x, y, comp, C.w, C.h) * (comp == 0 ? r.x : r.y);
                }
        }
        C.b[e] = make_float2(val[0], val[1]);
    }
    __syncthreads();
}

// F.x += P C.x on valid faces (in place: each slot reads only its own fine value).
__device__ void bv_prolong(const VLev& F, const VLev& C)
{
    const int n = (F.w + 1) * (F.h + 1);
    for (int e = threadIdx.x; e < n; e += kBVT) {
        int x, y;
        slot_xy(e, F.w, x, y);
        const float2 sv = F.x[e];
        float val[2] = {sv.x, sv.y};
        for (int comp = 0; comp < 2; ++comp) {
            int bx, by;
            float tx, ty;
            transfer_location(x, y, comp, bx, by, tx, ty);
            for (int dy = 0; dy < 2; ++dy)
                for (int dx = 0; dx < 2; ++dx) {
                    int px, py;
                    coarse_point(bx + dx, by + dy, comp, C.w, C.h, px, py);
                    const float2 c = C.x[py * (C.w + 1) + px];
                    val[comp] += (dx == 1 ? tx : 1.0f - tx) * (dy == 1 ? ty : 1.0f - ty) * (comp == 0 ? c.x : c.y);
                }
        }
        F.x[e] = make_float2(valid_u(F.w, F.h, x, y) ? val[0] : 0.0f, valid_v(F.w, F.h, x, y) ? val[1] : 0.0f);
    }
    __syncthreads();
}

// 16 Jacobi sweeps from zero on the coarsest level (k_vis_coarse), slots in LDS.
__device__ void bv_coarse(const VLev& L, const float2* b, float dt, float2* vel, float4* cof)
{
    const int lane = threadIdx.x;
    const int stride = L.w + 1;
    const int n = stride * (L.h + 1);
    const bool on = lane < kCoarseSlots;
    const int x = lane % stride, y = lane / stride;
    const int src = lane < n ? lane : n - 1;
    float2 rhs = make_float2(0.0f, 0.0f);
    if (on) {
        vel[lane] = make_float2(0.0f, 0.0f);
        cof[lane] = L.coef[src];
        rhs = b[src];
    }
    __syncthreads();
    for (int sweep = 0; sweep < 16; ++sweep) {
        float2 v = make_float2(0.0f, 0.0f);
        if (on && lane < n) {
            const float2 d = diagonal(cof, L.w, L.h, x, y, dt);
            const float2 q = product(cof, vel, L.w, L.h, x, y, dt);
            const float2 sv = vel[lane];
            v.x = valid_u(L.w, L.h, x, y) ? sv.x + 0.5f * (rhs.x - q.x) / d.x : 0.0f;
            v.y = valid_v(L.w, L.h, x, y) ? sv.y + 0.5f * (rhs.y - q.y) / d.y : 0.0f;
        }
        __syncthreads();
        if (on)
            vel[lane] = v;
        __syncthreads();
    }
    if (on && lane < n)
        L.x[lane] = vel[lane];
    __syncthreads();
}

__device__ void bv_cycle(const BatchVisDesc& D, float dt, float2* vel, float4* cof)
{
    const int nl = D.levels;
    for (int l = 0; l + 1 < nl; ++l) {
        const float2* b = l == 0 ? (const float2*)D.r : D.lev[l].b;
        bv_pre(D.lev[l], b, dt, D.M);
        bv_residual_restrict(D.lev[l], b, D.lev[l + 1], dt);
    }
    bv_coarse(D.lev[nl - 1], nl == 1 ? (const float2*)D.r : D.lev[nl - 1].b, dt, vel, cof);
    for (int l = nl - 2; l >= 0; --l) {
        const float2* b = l == 0 ? (const float2*)D.r : D.lev[l].b;
        bv_prolong(D.lev[l], D.lev[l + 1]);
        bv_post(D.lev[l], b, dt, D.M);
    }
}

// The coarse end of a V-cycle in one workgroup: levels whose sweeps are too small to fill the device cost
// more in launches than in work, so cycle() hands every level from the first with at most kVisTailSlots
// slots down to the coarsest to this kernel (1024: one level of 1196 slots more made a 1417 x 787 solve
// slower, one of 4590 slots 20% slower) (the same sweeps, restriction, Jacobi coarse solve and
// prolongation as the per-level kernels, each a workgroup-strided loop).  b0 is the first level's rhs;
// the result lands in lev[0].x.
constexpr int kVisTailSlots = 1024;

struct VisTail {
    int levels;
    VLev lev[kMaxVisLevels];
    VisSmooth M;
};

__global__ __launch_bounds__(kBVT) void k_vis_tail(const VisTail T, const float2* b0, const VisScalars* S)
{
    __shared__ float2 vel[kCoarseSlots];
    __shared__ float4 cof[kCoarseSlots];
    if (stopped(S))
        return;
    const float dt = S->dt;
    const int nl = T.levels;
    for (int l = 0; l + 1 < nl; ++l) {
        const float2* b = l == 0 ? b0 : T.lev[l].b;
        bv_pre(T.lev[l], b, dt, T.M);
        bv_residual_restrict(T.lev[l], b, T.lev[l + 1], dt);
    }
    bv_coarse(T.lev[nl - 1], nl == 1 ? b0 : T.lev[nl - 1].b, dt, vel, cof);
    for (int l = nl - 2; l >= 0; --l) {
        const float2* b = l == 0 ? b0 : T.lev[l].b;
        bv_prolong(T.lev[l], T.lev[l + 1]);
        bv_post(T.lev[l], b, dt, T.M);
    }
}

__global__ __launch_bounds__(kBVT) void k_batch_vis(const BatchVisDesc* descs)
{
    __shared__ double lds[kBVT / 64];
    __shared__ float2 vel[kCoarseSlots];
    __shared__ float4 cof[kCoarseSlots];
    const BatchVisDesc& D = descs[blockIdx.x];
    VisScalars* S = D.S;
    const VLev L0 = D.lev[0];
    const int W = D.W, H = D.H;
    const int n0 = (W + 1) * (H + 1);
    const float dt = S->dt;
    const double target = S->target;
    double rnorm = S->rnorm;
    int bad = S->bad;
    int status;   // 0 converged (or nothing to do), 1 non-finite input, 2 breakdown, 3 over budget
    int iterations = 0;
    double truth = S->truth;
    if (bad) {
        status = 1;
    } else if (rnorm <= target) {
        status = 0;
    } else {
        double rz = 0.0, alpha = 0.0, beta = 0.0;
        int steps = 0;
        const int budget = D.maximum < 0 ? kBatchVisBudget : min(D.maximum, kBatchVisBudget);
        status = 3;
        while (iterations < budget) {
            const int batch = min(4, budget - iterations);
            for (int i = 0; i < batch; ++i) {
                if (bad || rnorm <= target)
                    continue;   // the iteration kernels exit at once once the solve has stopped
                bv_cycle(D, dt, vel, cof);
                // r . z, r M^-1 r
                double a = 0.0, b2 = 0.0, f = 0.0;
                for (int e = threadIdx.x; e < n0; e += kBVT) {
                    const float4 c = L0.coef[e];
                    const float2 rr = D.r[e], zz = L0.x[e];
                    double ea = (double)rr.x * zz.x + (double)rr.y * zz.y;
                    double eb = (double)rr.x / c.z * rr.x + (double)rr.y / c.w * rr.y;
                    if (!isfinite(ea) || !isfinite(eb)) {
                        ea = eb = 0.0;
                        f = 1.0;
                    }
                    a += ea;
                    b2 += eb;
                }
                a = wg_sum_all(a, lds);
                b2 = wg_sum_all(b2, lds);
                f = wg_sum_all(f, lds);
                bad |= f != 0.0 ? 1 : 0;
                rnorm = b2;
                if (rnorm > target && !bad) {
                    if (a <= 0.0 || !isfinite(a)) {
                        bad = 1;
                    } else {
                        beta = steps > 0 ? a / rz : 0.0;
                        rz = a;
                    }
                }
                if (bad || rnorm <= target)
                    continue;
                // direction, q = A d and d . q
                const float fb = (float)beta;
                for (int e = threadIdx.x; e < n0; e += kBVT) {
                    const float2 zz = L0.x[e];
                    if (steps > 0) {
                        const float2 dd = D.d[e];
                        D.d[e] = make_float2(zz.x + fb * dd.x, zz.y + fb * dd.y);
                    } else {
                        D.d[e] = zz;
                    }
                }
                __syncthreads();
                double dq = 0.0;
                f = 0.0;
                for (int e = threadIdx.x; e < n0; e += kBVT) {
                    int x, y;
                    slot_xy(e, W, x, y);
                    const float2 qv = product(L0.coef, D.d, W, H, x, y, dt);
                    D.q[e] = qv;
                    const float2 dd = D.d[e];
                    double ea = (double)dd.x * qv.x + (double)dd.y * qv.y;
                    if (!isfinite(ea)) {
                        ea = 0.0;
                        f = 1.0;
                    }
                    dq += ea;
                }
                dq = wg_sum_all(dq, lds);
                f = wg_sum_all(f, lds);
                if (f != 0.0 || dq <= 0.0 || !isfinite(dq)) {
                    bad = 1;
                    continue;
                }
                alpha = rz / dq;
                steps += 1;
                const float fa = (float)alpha;
                for (int e = threadIdx.x; e < n0; e += kBVT) {
                    const float2 dd = D.d[e], qq = D.q[e], xx = D.t[e], rr = D.r[e];
                    D.t[e] = make_float2(xx.x + fa * dd.x, xx.y + fa * dd.y);
                    D.r[e] = make_float2(rr.x - fa * qq.x, rr.y - fa * qq.y);
                }
                __syncthreads();
            }
            iterations += batch;
            // recomputed residual q = M v0 - A x and q M^-1 q
            double tr = 0.0, f = 0.0;
            for (int e = threadIdx.x; e < n0; e += kBVT) {
                int x, y;
                slot_xy(e, W, x, y);
                const float4 c = L0.coef[e];
                const float2 qv = product(L0.coef, D.t, W, H, x, y, dt);
                const float2 s0 = D.v0[e];
                float2 vv;
                vv.x = valid_u(W, H, x, y) ? c.z * s0.x - qv.x : 0.0f;
                vv.y = valid_v(W, H, x, y) ? c.w * s0.y - qv.y : 0.0f;
                D.q[e] = vv;
                double ea = (double)vv.x / c.z * vv.x + (double)vv.y / c.w * vv.y;
                if (!isfinite(ea)) {
                    ea = 0.0;
                    f = 1.0;
                }
                tr += ea;
            }
            tr = wg_sum_all(tr, lds);
            f = wg_sum_all(f, lds);
            truth = f != 0.0 ? INFINITY : tr;
            if (f != 0.0)
                bad = 1;
            if (bad) {
                status = 2;
                break;
            }
            if (truth <= target) {
                status = 0;
                // the staggered result (k_vis_unpack)
                for (int e = threadIdx.x; e < n0; e += kBVT) {
                    int x, y;
                    slot_xy(e, W, x, y);
                    const float2 xs = D.t[e];
                    if (y < H)
                        D.u[y * (W + 1) + x] = valid_u(W, H, x, y) ? (double)xs.x : 0.0;
                    if (x < W)
                        D.v[y * W + x] = valid_v(W, H, x, y) ? (double)xs.y : 0.0;
                }
                break;
            }
            if (rnorm <= target) {
                // the recurrence claims convergence the recomputed residual denies: restart from it
                for (int e = threadIdx.x; e < n0; e += kBVT)
                    D.r[e] = D.q[e];
                __syncthreads();
                rnorm = truth;
                rz = alpha = beta = 0.0;
                steps = 0;
            }
        }
    }
    if (threadIdx.x == 0) {
        S->bad = bad;
        S->rnorm = rnorm;
        S->truth = truth;
        S->status = status;
        S->iterations = iterations;
    }
}

}  // namespace

// Device memory handed out in 256-byte aligned pieces from one block the arena does not own.  With a
// null base it only measures (every piece is null).
struct ScratchArena {
    unsigned char* base = nullptr;
    size_t capacity = 0;
    size_t used = 0;
    template <class T> T* take(size_t count)
    {
        const size_t offset = (used + 255) & ~(size_t)255;
        used = offset + count * sizeof(T);
        if (base && used > capacity)
            throw std::runtime_error("viscosity scratch arena exhausted");
        return base ? (T*)(base + offset) : nullptr;
    }
};

struct ViscositySolver {
    struct Level {
        int w, h;
        VBuf<float4> coef;
        VBuf<float2> mat, b, x, xn, r;
    };
    // A grid halves until w * h <= 64: at most 31 times for int dimensions.  Fixed storage, so a solver
    // placed in a scratch arena makes no allocation at all.
    struct LevelList {
        static constexpr int kCapacity = 32;
        Level items[kCapacity];
        int count = 0;
        size_t size() const { return (size_t)count; }
        Level* operator[](size_t i) { return &items[i]; }
        const Level* operator[](size_t i) const { return &items[i]; }
        Level* add() { return &items[count++]; }
    };
    int last_rounds = 0;   // batches of four iterations the last solve ran (0: none, or no solve yet)
    int W, H;
    LevelList levels;
    VBuf<float2> v0, r0, r, d, q, t;
    VBuf<double> material, partials;
    VBuf<VisScalars> S;
    VisScalars* pinned = nullptr;
    VisScalars landing{};           // read-back slot of a solver placed in an arena (no pinned memory)
    ScratchArena* arena = nullptr;  // non-null: every device buffer is a piece of the arena
    size_t bytes = 0;

    template <class T> void place(VBuf<T>& b, size_t count)
    {
        if (arena)
            b.borrow(arena->template take<T>(count), count);
        else
            bytes += b.alloc(count);
    }

    // With an arena: the buffers are pieces of it (zeroed by the caller), no per-grid material buffer, no
    // pinned landing slot and no captured graph.
    ViscositySolver(int w, int h, ScratchArena* in = nullptr) : W(w), H(h), arena(in)
    {
        for (;;) {
            Level* L = levels.add();
            L->w = w;
            L->h = h;
            const size_t ns = (size_t)(w + 1) * (h + 1);
            place(L->coef, ns);
            place(L->mat, (size_t)w * h);
            place(L->b, ns);
            place(L->x, ns);
            place(L->xn, ns);
            place(L->r, ns);
            if (w * h <= 64)
                break;
            w = (w + 1) / 2;
            h = (h + 1) / 2;
        }
        const size_t ns = (size_t)(W + 1) * (H + 1);
        place(v0, ns);
        place(r0, ns);
        place(r, ns);
        place(d, ns);
        place(q, ns);
        place(t, ns);
        if (!arena)
            place(material, 2 * (size_t)W * H);
        place(partials, 4 * (size_t)std::max(slot_blocks(W, H), 1u));
        place(S, 1);
        if (!arena)
            pinned = (VisScalars*)pinned_alloc(sizeof(VisScalars));
    }

    ~ViscositySolver()
    {
        drop_batch_graph();
        pinned_free(pinned, sizeof(VisScalars));
    }

    VLev view(int l)
    {
        Level& L = *levels[l];
        return VLev{L.w, L.h, L.coef.p, L.mat.p, L.b.p, L.x.p, L.xn.p, L.r.p};
    }

    void swap_x(int l) { std::swap(levels[l]->x.p, levels[l]->xn.p); }

    VisScalars read(hipStream_t st)
    {
        VisScalars* slot = pinned ? pinned : &landing;
        vhip(hipMemcpyAsync(slot, S.p, sizeof(VisScalars), hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
        vhip(hipStreamSynchronize(st), "hipStreamSynchronize");
        return *slot;
    }

    void finish(hipStream_t st, int nb, int op, double tol2, float dt = 0.0f)
    {
        hipLaunchKernelGGL(k_vis_finish, dim3(1), dim3(kVT), 0, st, S.p, partials.p, nb, op, tol2, dt);
    }

    // cycle (ViscositySolverGPU.cpp:134-167); the result is levels[l]->x.  Every level swaps its x / xn
    // pair once per cycle, so four cycles leave the pointers where they started.
    // First level of the one-workgroup tail (k_vis_tail): the first level below the finest with at most
    // kVisTailSlots slots, when at least one level lies between it and the coarsest; else none (-1).
    int tail_level() const
    {
        const int nl = (int)levels.size();
        for (int l = 1; l + 1 < nl; ++l)
            if ((size_t)(levels[l]->w + 1) * (levels[l]->h + 1) <= (size_t)kVisTailSlots)
                return nl - l <= kMaxVisLevels ? l : -1;
        return -1;
    }

    // Smoothing weights of every level (fs_set_viscosity_smoother): the reference's two sweeps of weight 1/2,
    // or 2, 4 or 6 Chebyshev-weighted sweeps, 1 / (2.125 + 1.875 cos(pi (k + 1/2) / n)) rounded to float --
    // the reciprocals of the Chebyshev nodes of [0.25, 4], 4 bounding the spectrum of D^-1 A (Gershgorin).
    VisSmooth smooth{2, {0.5f, 0.5f, 0.0f, 0.0f, 0.0f, 0.0f}};
    int smooth_chebyshev = 0;

    void set_smoother(int chebyshev_sweeps)
    {
        if (chebyshev_sweeps != 0 && chebyshev_sweeps != 2 && chebyshev_sweeps != 4 && chebyshev_sweeps != 6)
            throw std::invalid_argument("viscosity smoother sweeps must be 0, 2, 4 or 6");
        if (chebyshev_sweeps == smooth_chebyshev)
            return;
        VisSmooth m{2, {0.5f, 0.5f, 0.0f, 0.0f, 0.0f, 0.0f}};
        if (chebyshev_sweeps) {
            const double pi = 3.14159265358979323846;
            m.sweeps = chebyshev_sweeps;
            for (int k = 0; k < chebyshev_sweeps; ++k)
                m.w[k] = (float)(1.0 / (2.125 + 1.875 * std::cos(pi * (k + 0.5) / chebyshev_sweeps)));
        }
        smooth = m;
        smooth_chebyshev = chebyshev_sweeps;
        drop_batch_graph();   // the captured launches carry the weights
    }

    template <bool POST> void launch_smooth(hipStream_t st, dim3 g2, const VLev& L, const VLev& C, const float2* b)
    {
        if (smooth.sweeps == 6)
            hipLaunchKernelGGL((k_vis_smooth<POST, 6>), g2, dim3(kVT), 0, st, L, C, b, S.p, smooth);
        else if (smooth.sweeps == 4)
            hipLaunchKernelGGL((k_vis_smooth<POST, 4>), g2, dim3(kVT), 0, st, L, C, b, S.p, smooth);
        else
            hipLaunchKernelGGL((k_vis_smooth<POST, 2>), g2, dim3(kVT), 0, st, L, C, b, S.p, smooth);
    }

    void cycle(hipStream_t st, int l, const float2* b)
    {
        VLev L = view(l);
        if (l + 1 == (int)levels.size()) {
            hipLaunchKernelGGL(k_vis_coarse, dim3(1), dim3(kCoarseSlots), 0, st, L, b, S.p);
            return;
        }
        if (l == tail_level()) {
            // the tail leaves each of its levels' results in x: no swaps there
            VisTail T{};
            T.levels = (int)levels.size() - l;
            for (int k = 0; k < T.levels; ++k)
                T.lev[k] = view(l + k);
            T.M = smooth;
            hipLaunchKernelGGL(k_vis_tail, dim3(1), dim3(kBVT), 0, st, T, b, S.p);
            return;
        }
        const dim3 g = slot_grid(L.w, L.h);
        const dim3 g2((L.w + 1 + kSX - 1) / kSX, (L.h + 1 + kSY - 1) / kSY);
        // pre-smoothing: every sweep from zero in one launch (the result lands in L.x, as after an even
        // number of swaps)
        launch_smooth<false>(st, g2, L, L, b);
        hipLaunchKernelGGL(k_vis_residual, g, dim3(kVT), 0, st, L, b, S.p);
        VLev C = view(l + 1);
        hipLaunchKernelGGL(k_vis_restrict, slot_grid(C.w, C.h), dim3(kVT), 0, st, L, C, S.p);
        cycle(st, l + 1, C.b);
        C = view(l + 1);
        // prolongation + every sweep in one launch into L.xn: one swap, as after the odd number of the
        // per-sweep kernels
        launch_smooth<true>(st, g2, L, C, b);
        swap_x(l);
    }

    // One CG iteration (ViscositySolverGPU.cpp:238-262): V-cycle, r.z and beta, direction, A d and alpha,
    // update of the iterate `delta` and the residual.
    void iteration(hipStream_t st, float2* delta)
    {
        const VLev L0 = view(0);
        const dim3 g0 = slot_grid(W, H);
        const int nb = (int)slot_blocks(W, H);
        cycle(st, 0, r.p);
        const VLev Lc = view(0);
        hipLaunchKernelGGL(k_vis_rz, g0, dim3(kVT), 0, st, Lc, (const float2*)r.p, (const float2*)Lc.x, S.p,
            partials.p);
        finish(st, nb, kVisRz, 0.0);
        hipLaunchKernelGGL(k_vis_direction, g0, dim3(kVT), 0, st, Lc, (const float2*)Lc.x, d.p, S.p);
        hipLaunchKernelGGL(k_vis_apply, g0, dim3(kVT), 0, st, L0, (const float2*)d.p, q.p, S.p, partials.p);
        finish(st, nb, kVisDq, 0.0);
        hipLaunchKernelGGL(k_vis_update, g0, dim3(kVT), 0, st, L0, delta, r.p, (const float2*)d.p,
            (const float2*)q.p, S.p);
    }

    // A batch of four iterations captured once into a hipGraph: the kernels take every changing value
    // (dt, alpha, beta, the stop state) from the scalar block, and the x / xn swaps of four cycles return
    // to their starting pointers, so one graph replays every full batch of every solve on this solver.
    hipGraph_t batch_graph = nullptr;
    hipGraphExec_t batch_exec = nullptr;
    hipStream_t batch_stream = nullptr;

    void run_batch(hipStream_t st, float2* delta)
    {
        if (arena) {   // instantiating a graph allocates: the same four iterations, launched one by one
            for (int i = 0; i < 4; ++i)
                iteration(st, delta);
            return;
        }
        if (!batch_exec || batch_stream != st) {
            drop_batch_graph();
            std::lock_guard<std::mutex> lock(device_setup_mutex());
            vhip(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture");
            for (int i = 0; i < 4; ++i)
                iteration(st, delta);
            vhip(hipStreamEndCapture(st, &batch_graph), "hipStreamEndCapture");
            vhip(hipGraphInstantiate(&batch_exec, batch_graph, nullptr, nullptr, 0), "hipGraphInstantiate");
            batch_stream = st;
        }
        vhip(hipGraphLaunch(batch_exec, st), "hipGraphLaunch");
    }

    void drop_batch_graph()
    {
        if (batch_exec)
            (void)hipGraphExecDestroy(batch_exec);
        if (batch_graph)
            (void)hipGraphDestroy(batch_graph);
        batch_exec = nullptr;
        batch_graph = nullptr;
    }

    // Set-up of a solve (ViscositySolverGPU.cpp:169-205): material, level coefficients, v0, the rhs r0 and
    // the initial reduction into the scalar block.
    void prepare(hipStream_t st, double* u, double* v, const double* mat64, float fdt, double spacing, double tol)
    {
        const int N = W * H;
        hipLaunchKernelGGL(k_vis_material, dim3(std::min((N + kVT - 1) / kVT, 4096)), dim3(kVT), 0, st, N, mat64,
            levels[0]->mat.p);
        for (size_t l = 0; l < levels.size(); ++l) {
            if (l > 0) {
                VLev F = view((int)l - 1), C = view((int)l);
                hipLaunchKernelGGL(k_vis_coarsen, dim3((C.w + kVX - 1) / kVX, (C.h + kVY - 1) / kVY), dim3(kVT), 0,
                    st, F, C);
            }
            VLev L = view((int)l);
            hipLaunchKernelGGL(k_vis_coef, slot_grid(L.w, L.h), dim3(kVT), 0, st, L, (float)ldexp(spacing, (int)l));
        }
        VLev L0 = view(0);
        const dim3 g0 = slot_grid(W, H);
        hipLaunchKernelGGL(k_vis_pack, g0, dim3(kVT), 0, st, L0, (const double*)u, (const double*)v, v0.p);
        hipLaunchKernelGGL(k_vis_initial, g0, dim3(kVT), 0, st, L0, (const float2*)v0.p, r0.p, fdt, partials.p);
        finish(st, (int)slot_blocks(W, H), kVisInitial, tol * tol, fdt);
    }

    // Batched solve (k_batch_vis): set-up, and the descriptor of this solver's instance.
    bool batch_prepare(hipStream_t st, double* u, double* v, const double* mat64, double dt, double spacing,
        double tol, int maximum, BatchVisDesc* desc)
    {
        if (levels.size() > (size_t)kMaxVisLevels)
            return false;
        prepare(st, u, v, mat64, (float)dt, spacing, tol);
        const size_t ns = (size_t)(W + 1) * (H + 1);
        vhip(hipMemcpyAsync(r.p, r0.p, ns * sizeof(float2), hipMemcpyDeviceToDevice, st), "copy");
        vhip(hipMemcpyAsync(t.p, v0.p, ns * sizeof(float2), hipMemcpyDeviceToDevice, st), "copy");
        *desc = BatchVisDesc{};
        desc->levels = (int)levels.size();
        desc->W = W;
        desc->H = H;
        desc->maximum = maximum;
        for (size_t l = 0; l < levels.size(); ++l)
            desc->lev[l] = view((int)l);
        desc->M = smooth;
        desc->v0 = v0.p;
        desc->r = r.p;
        desc->d = d.p;
        desc->q = q.p;
        desc->t = t.p;
        desc->u = u;
        desc->v = v;
        desc->S = S.p;
        return true;
    }

    void batch_fetch(hipStream_t st)
    {
        vhip(hipMemcpyAsync(pinned, S.p, sizeof(VisScalars), hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
    }

    // After batch_fetch and a synchronisation: viscosity_solve's status code, or 4 when the instance needs
    // more iterations than the batch runs.
    int batch_result(int maximum, int* iterations, double* relative) const
    {
        const VisScalars& s = *pinned;
        const double denom = std::max(s.energy, 1e-30);
        *iterations = s.iterations;
        *relative = std::sqrt(std::max(s.iterations > 0 ? s.truth : s.rnorm, 0.0) / denom);
        if (s.status == 3 && (maximum < 0 || maximum > s.iterations))
            return 4;
        return s.status;
    }

    // ViscositySolverGPU::Impl::solve (ViscositySolverGPU.cpp:169-294) on device FP64 u, v and an FP64
    // (rho, mu) material; u, v change only on success.
    int solve(hipStream_t st, double* u, double* v, const double* mat64, double dt, double spacing, double tol,
        int maximum, int* iterations_out, double* relative_out)
    {
        *iterations_out = 0;
        *relative_out = 0.0;
        if (dt <= 0.0)
            return 0;
        const float fdt = (float)dt;
        prepare(st, u, v, mat64, fdt, spacing, tol);
        VLev L0 = view(0);
        const dim3 g0 = slot_grid(W, H);
        const int nb = (int)slot_blocks(W, H);
        const size_t ns = (size_t)(W + 1) * (H + 1);
        vhip(hipMemcpyAsync(r.p, r0.p, ns * sizeof(float2), hipMemcpyDeviceToDevice, st), "copy");
        vhip(hipMemcpyAsync(t.p, v0.p, ns * sizeof(float2), hipMemcpyDeviceToDevice, st), "copy");
        float2* delta = t.p;   // the iterate x, starting from v0 (the reference's formulation)
        VBuf<float2>& truth = q;   // q is free between iterations
        // Batches of four iterations, each followed by the recomputed residual and k_vis_decide, are queued
        // without a read-back in between: one batch fewer than the previous solve needed (a batch past the
        // solve's end costs its launches, a missing one a read-back), then one at a time.  Every decision
        // is taken on the device, so the iterations and values are those of reading back after each batch.
        int rounds = last_rounds > 0 ? std::max(1, last_rounds - 1) : 0;
        int queued = 0;
        VisScalars s{};
        for (;;) {
            const int queued_before = queued;
            for (int k = 0; k < rounds && (maximum < 0 || queued < maximum); ++k) {
                const int batch = maximum < 0 ? 4 : std::min(4, maximum - queued);
                if (batch == 4) {
                    run_batch(st, delta);
                } else {
                    for (int i = 0; i < batch; ++i)
                        iteration(st, delta);
                }
                queued += batch;
                L0 = view(0);
                hipLaunchKernelGGL(k_vis_truth, g0, dim3(kVT), 0, st, L0, (const float2*)v0.p,
                    (const float2*)delta, truth.p, fdt, partials.p, (const VisScalars*)S.p);
                finish(st, nb, kVisTruth, 0.0);
                hipLaunchKernelGGL(k_vis_decide, dim3(1), dim3(64), 0, st, S.p, batch, maximum);
                // the recurrence claims convergence the recomputed residual denies: restart from it
                hipLaunchKernelGGL(k_vis_restart_copy, dim3((unsigned)std::min<size_t>((ns + kVT - 1) / kVT, 2048)),
                    dim3(kVT), 0, st, (const VisScalars*)S.p, r.p, (const float2*)truth.p, ns);
                hipLaunchKernelGGL(k_vis_refresh, dim3(1), dim3(64), 0, st, S.p);
                hipLaunchKernelGGL(k_vis_unpack, g0, dim3(kVT), 0, st, W, H, (const float2*)delta, u, v,
                    (const VisScalars*)S.p);
            }
            s = read(st);
            // (k_vis_decide ends the solve once the budget is spent: nothing left to queue means done)
            if (s.done != 0 || maximum == 0 || (rounds > 0 && queued == queued_before))
                break;
            rounds = 1;
        }
        last_rounds = (s.iterations + 3) / 4;
        const double denom = std::max(s.energy, 1e-30);
        *iterations_out = s.iterations;
        const bool iterated = s.done >= 1 && s.done <= 4;
        *relative_out = std::sqrt(std::max(iterated ? s.truth : s.rnorm, 0.0) / denom);
        switch (s.done) {
        case 5:
        case 1:
        case 4: return 0;
        case 6: return 1;   // non-finite velocity or material
        case 2: return 2;   // breakdown
        default: return 3;  // not converged within the caller's budget
        }
    }
};

// ---------------------------------------------------------------------------------------------
// Allocation-free solve: the solver's workspace lives in a file-scope device array, one per GPU.
// ---------------------------------------------------------------------------------------------
namespace {

// Each device that loads this library holds its own copy until the process ends.  Host code never
// passes the array's name to a kernel (on the host it names a stand-in, not device memory): it uses the
// per-device address hipGetSymbolAddress returns.
static __device__ __align__(256) unsigned char scratch[kViscosityScratchBytes];

constexpr int kScratchDevices = 64;
std::mutex g_scratch_mutex;                         // guards the address table
void* g_scratch_address[kScratchDevices] = {};
std::mutex g_scratch_use[kScratchDevices];          // one solve at a time per device

__global__ __launch_bounds__(kVT) void k_zero_words(unsigned* p, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * kVT + threadIdx.x; i < n; i += (size_t)gridDim.x * kVT)
        p[i] = 0u;
}

}  // namespace

void* viscosity_scratch_setup(int device)
{
    if (device < 0 || device >= kScratchDevices)
        throw std::runtime_error("device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_address[device]) {
        void* address = nullptr;
        vhip(hipGetSymbolAddress(&address, HIP_SYMBOL(scratch)), "hipGetSymbolAddress");
        if (!address)
            throw std::runtime_error("hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_address[device] = address;
    }
    return g_scratch_address[device];
}

size_t viscosity_scratch_bytes(int width, int height)
{
    ScratchArena measure;
    ViscositySolver probe(width, height, &measure);
    return measure.used;
}

int viscosity_solve_scratch(int device, hipStream_t st, int width, int height, double* u, double* v,
    const double* material, double dt, double spacing, double tol, int maximum, int* iterations,
    double* relative_residual)
{
    void* base = viscosity_scratch_setup(device);
    const size_t need = viscosity_scratch_bytes(width, height);
    if (need > kViscosityScratchBytes)
        throw std::invalid_argument("grid too large for the allocation-free viscosity solve's scratch");
    std::lock_guard<std::mutex> serial(g_scratch_use[device]);
    // every call starts from zeroed memory, as freshly allocated buffers do (the last call's values
    // stay in the array otherwise)
    const size_t words = (need + 3) / 4;
    hipLaunchKernelGGL(k_zero_words, dim3((unsigned)std::min<size_t>((words + kVT - 1) / kVT, 4096)), dim3(kVT), 0,
        st, (unsigned*)base, words);
    ScratchArena arena;
    arena.base = (unsigned char*)base;
    arena.capacity = kViscosityScratchBytes;
    ViscositySolver solver(width, height, &arena);
    const int status = solver.solve(st, u, v, material, dt, spacing, tol, maximum, iterations, relative_residual);
    // the last kernels (velocity unpack) read the scratch: done before another call may reuse it
    vhip(hipStreamSynchronize(st), "hipStreamSynchronize");
    return status;
}

ViscositySolver* viscosity_create(int width, int height) { return new ViscositySolver(width, height); }
void viscosity_destroy(ViscositySolver* s) { delete s; }
size_t viscosity_bytes(const ViscositySolver* s) { return s ? s->bytes : 0; }
int viscosity_levels(const ViscositySolver* s) { return s ? (int)s->levels.size() : 0; }
void viscosity_set_smoother(ViscositySolver* s, int chebyshev_sweeps) { s->set_smoother(chebyshev_sweeps); }
double* viscosity_material_buffer(ViscositySolver* s) { return s->material.p; }

int viscosity_solve(ViscositySolver* s, hipStream_t st, double* u, double* v, const double* material, double dt,
    double spacing, double tol, int maximum, int* iterations, double* relative_residual)
{
    return s->solve(st, u, v, material, dt, spacing, tol, maximum, iterations, relative_residual);
}

size_t viscosity_batch_desc_bytes() { return sizeof(BatchVisDesc); }

bool viscosity_batch_prepare(ViscositySolver* s, hipStream_t st, double* u, double* v, const double* material,
    double dt, double spacing, double tol, int maximum, void* desc)
{
    return s->batch_prepare(st, u, v, material, dt, spacing, tol, maximum, (BatchVisDesc*)desc);
}

void launch_batch_viscosity(hipStream_t st, const void* descs, int count, size_t reserve_lds)
{
    if (count <= 0)
        return;
    // reserved (unused) dynamic LDS keeps a second instance off a busy compute unit
    if (reserve_lds > 65536 &&
        hipFuncSetAttribute((const void*)k_batch_vis, hipFuncAttributeMaxDynamicSharedMemorySize, (int)reserve_lds) !=
            hipSuccess)
        reserve_lds = 0;
    hipLaunchKernelGGL(k_batch_vis, dim3(count), dim3(kBVT), reserve_lds, st, (const BatchVisDesc*)descs);
}

void viscosity_batch_fetch(ViscositySolver* s, hipStream_t st) { s->batch_fetch(st); }

int viscosity_batch_result(const ViscositySolver* s, int maximum, int* iterations, double* relative_residual)
{
    return s->batch_result(maximum, iterations, relative_residual);
}

void launch_viscosity_blend(hipStream_t st, int N, double cell_size, int nl, int ng, const double* lm,
    const double* gm, const VisMixture& mix, double* material)
{
    hipLaunchKernelGGL(k_vis_blend, dim3(std::min((N + kVT - 1) / kVT, 4096)), dim3(kVT), 0, st, N,
        cell_size * cell_size, nl, ng, lm, gm, mix, material);
}

}  // namespace fsim
