// fs_multigrid.hip -- the multigrid preconditioner of PressureMultigrid.cpp on gfx950.
//
// One V(4,4) cycle per level is three launches:
//   k_rb_multi      four red-black Gauss-Seidel sweeps on a (64+16) x (32+16) LDS tile (halo 8: each
//                   half-sweep shrinks the valid region by one cell), from zero on the way down;
//   k_res_restrict  residual of the smoothed solution and its transpose-of-bilinear restriction;
//   k_rb_multi      prolongation of the coarse correction fused into the load, then four reversed sweeps.
// The deepest levels (up to 64 x 64, as LDS allows) and the coarse solve run in k_mg_tail, one
// 1024-thread workgroup with every array in LDS.  The pinned coarse system (PressureMultigrid.cpp:
// 300-384) is factored bit-exactly by k_coarse_factor, which also forms its inverse so the per-cycle
// coarse solve is a 63 x 63 matrix-vector product instead of two serial substitutions.
//
// Smoothers divide by the sum of the four face conductances formed in FP32 (one rcp); pre- and
// post-smoothing use the same diagonal, so the V-cycle stays symmetric.
#include "fs_kernels.hpp"

#include <math.h>

#include <algorithm>
#include <type_traits>

namespace fsim {

namespace {

constexpr int kNT = 256;

__device__ __forceinline__ float inv_or_zero(float d) { return d > 0.f ? __builtin_amdgcn_rcpf(d) : 0.f; }

// Loads and stores of multigrid storage formats; arithmetic is FP32.  Stores round to nearest even.
__device__ __forceinline__ float ldv(float v) { return v; }
__device__ __forceinline__ float ldv(half16 v) { return (float)__builtin_bit_cast(_Float16, v.bits); }
__device__ __forceinline__ float ldv(bf16 v) { return __uint_as_float((unsigned)v.bits << 16); }
template <class T> __device__ __forceinline__ T stv(float f);
template <> __device__ __forceinline__ float stv<float>(float f) { return f; }
template <> __device__ __forceinline__ half16 stv<half16>(float f)
{
    return half16{__builtin_bit_cast(unsigned short, (_Float16)f)};
}
template <> __device__ __forceinline__ bf16 stv<bf16>(float f)
{
    // round to nearest even on the 16 bits dropped (finite inputs: the levels hold no NaN)
    const unsigned u = __float_as_uint(f);
    return bf16{(unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16)};
}

// XCD-aware tile order.  Workgroups of a launch are dealt to the 8 XCDs round-robin (workgroup b runs
// on XCD b % 8), each XCD with its own L2.  Mapping the workgroups an XCD receives onto a contiguous
// run of row-major tiles keeps neighbouring tiles, whose 8-cell halos overlap, on the same L2.  The
// launch is one-dimensional with exactly gx * gy workgroups; the map is a bijection onto the tiles.
constexpr int kXcds = 8;
__device__ __forceinline__ void xcd_tile(int gx, int gy, int& bx, int& by)
{
    const int n = gx * gy;
    const int b = blockIdx.x;
    const int per = n / kXcds, rem = n % kXcds;
    const int xcd = b % kXcds, q = b / kXcds;
    const int t = xcd * per + min(xcd, rem) + q;
    by = t / gx;
    bx = t - by * gx;
}

// ---------------------------------------------------------------------------------------------
// Four-sweep red-black Gauss-Seidel with optional fused prolongation.
// ---------------------------------------------------------------------------------------------
// Layout: the tile's solution lives in two colour arrays (red, black), each LX/2 wide, so a wavefront
// updating 64 same-colour cells of a row touches 64 consecutive LDS words (no bank conflicts).  Each
// thread owns fixed (row, pair) slots; the four conductances, rhs and 1/diagonal of its cells stay in
// registers for all eight half-sweeps, so an update reads four neighbours from LDS and writes one value.
// MODE 0: pre-smoothing from a zero guess (no solution read); 1: post-smoothing, the prolonged coarse
// correction added to the solution first; 2: plain sweeps of the stored solution.
// Post-smoothing tiles are held to the registers that keep their workgroups per CU: 128 VGPRs for 512
// threads (four waves per SIMD, two workgroups) and 960 (one), 170 for 384 (two) and 640 threads (one).
template <int MODE, int NT> constexpr int rb_waves_per_eu()
{
    return MODE != 1 ? 1 : (NT == 512 || NT == 960) ? 4 : (NT == 384 || NT == 640) ? 3 : 1;
}

// C: conductance storage, V: this level's vectors (rhs, solution), VC: the coarse solution (MODE 1).
template <int TX, int TY, int NS, int FINE, int MODE, int NT, class C = float, class V = float, class VC = float>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(rb_waves_per_eu<MODE, NT>())))
void k_rb_multi(LevelViewT<C, V> L, int cur, const V* rhs, const double* shift, int c0, int nu, CoarseLinkT<VC> link,
    const int* status)
{
    constexpr bool from_zero = MODE == 0;
    constexpr int HALO = 2 * NS;
    constexpr int LX = TX + 2 * HALO, LY = TY + 2 * HALO;
    constexpr int HX = LX / 2;                       // colour-array row width
    constexpr int NE = LX * LY;
    constexpr int SLOTS = HX * (LY - 2);             // rows 1 .. LY-2, every pair
    constexpr int ROUNDS = (SLOTS + NT - 1) / NT;
    static_assert(LX % 2 == 0 && (TX % 2) == 0 && (HALO % 2) == 0, "tile must keep colour parity");
    __shared__ float s_col[2][LY * HX];
    const int w = L.w, h = L.h, s = L.s;
    int bx, by;
    xcd_tile((w + TX - 1) / TX, (h + TY - 1) / TY, bx, by);
    const int x0 = bx * TX - HALO, y0 = by * TY - HALO;  // x0 + y0 is even
    const V* sol_in = L.sol[cur];
    // Every global load below goes through a clamped (valid) interior index and is masked afterwards,
    // so a thread issues all its loads before its first wait.
    float cl[2][ROUNDS], cr[2][ROUNDS], cu[2][ROUNDS], cd[2][ROUNDS], cb[2][ROUNDS];
    int cly[ROUNDS], cj[ROUNDS];
    bool cin[2][ROUNDS];
#pragma unroll
    for (int r = 0; r < ROUNDS; ++r) {
        const int t = threadIdx.x + r * NT;
        const int ly = 1 + t / HX, jj = t - (t / HX) * HX;
        cly[r] = t < SLOTS ? ly : -1;
        cj[r] = jj;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const int lx = 2 * jj + ((c + ly) & 1);
            const int x = x0 + lx, y = y0 + ly;
            cin[c][r] = t < SLOTS && x >= 0 && x < w && y >= 0 && y < h;
            const int P = (min(max(y, 0), h - 1) + 1) * s + (min(max(x, 0), w - 1) + 1);
            cl[c][r] = ldv(L.right[P - 1]);
            cr[c][r] = ldv(L.right[P]);
            cu[c][r] = ldv(L.down[P - s]);
            cd[c][r] = ldv(L.down[P]);
            cb[c][r] = ldv(rhs[P]);
        }
    }
    constexpr int LOADS = (NE + NT - 1) / NT;
    float vs[LOADS];
    bool vin[LOADS];
    int ix[LOADS], iy[LOADS];
#pragma unroll
    for (int it = 0; it < LOADS; ++it) {
        const int e = threadIdx.x + it * NT;
        const int ly = e / LX, lx = e - ly * LX;
        const int x = x0 + lx, y = y0 + ly;
        vin[it] = e < NE && x >= 0 && x < w && y >= 0 && y < h;
        ix[it] = min(max(x, 0), w - 1);
        iy[it] = min(max(y, 0), h - 1);
        vs[it] = MODE == 0 ? 0.f : ldv(sol_in[(iy[it] + 1) * s + (ix[it] + 1)]);
    }
    if (MODE == 1) {
        // per chunk of PC cells, two batched phases: transfer tables, then the four coarse values (chunks
        // bound the registers this prologue adds to the kernel)
        constexpr int PC = NT >= 384 ? 1 : 4;   // large tiles keep to 128 VGPRs without spilling
#pragma unroll
        for (int c0 = 0; c0 < LOADS; c0 += PC) {
            int a0[PC], a1[PC], b0[PC], b1[PC];
            float wx[PC], wy[PC];
#pragma unroll
            for (int k = 0; k < PC; ++k) {
                const int it = min(c0 + k, LOADS - 1);
                wx[k] = link.pxw[ix[it]];
                wy[k] = link.pyw[iy[it]];
                a0[k] = link.pxf[ix[it]];
                a1[k] = link.pxs[ix[it]];
                b0[k] = link.pyf[iy[it]];
                b1[k] = link.pys[iy[it]];
            }
#pragma unroll
            for (int k = 0; k < PC; ++k) {
                if (c0 + k >= LOADS)
                    break;
                const VC* top = link.csol + (b0[k] + 1) * link.cs + 1;
                const VC* bot = link.csol + (b1[k] + 1) * link.cs + 1;
                const float t0 = ldv(top[a0[k]]), t1 = ldv(top[a1[k]]), u0 = ldv(bot[a0[k]]), u1 = ldv(bot[a1[k]]);
                const float a = (1.f - wx[k]) * t0 + wx[k] * t1;
                const float b = (1.f - wx[k]) * u0 + wx[k] * u1;
                vs[c0 + k] += (1.f - wy[k]) * a + wy[k] * b;
            }
        }
    }
    if (status && *status != kRunning)
        return;
    const double sh = shift ? *shift : 0.0;
#pragma unroll
    for (int r = 0; r < ROUNDS; ++r) {
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            if (!cin[c][r]) {
                cl[c][r] = cr[c][r] = cu[c][r] = cd[c][r] = cb[c][r] = 0.f;
            } else if (shift) {
                cb[c][r] = (float)((double)cb[c][r] - sh);
            }
        }
    }
#pragma unroll
    for (int it = 0; it < LOADS; ++it) {
        const int e = threadIdx.x + it * NT;
        if (e < NE) {
            const int ly = e / LX, lx = e - ly * LX;
            s_col[(lx + ly) & 1][ly * HX + (lx >> 1)] = vin[it] ? vs[it] : 0.f;
        }
    }
    __syncthreads();
    // Per-slot invariants of the eight half-sweeps, computed once: the slot's LDS index, the index of
    // its left neighbour for each colour, 1/diagonal of each colour's cell, and (one byte per colour)
    // the cell's distance to the tile edge plus one -- zero for a slot outside the grid -- so the
    // half-sweep at depth k (region shrunk by k + 1 cells) updates the cell iff its byte exceeds k + 1.
    int sdst[ROUNDS], sleft[2][ROUNDS], sreach[ROUNDS];
    float sinv[2][ROUNDS];
#pragma unroll
    for (int r = 0; r < ROUNDS; ++r) {
        const int ly = max(cly[r], 1);        // unused slots address a valid row and stay inactive
        const int jj = cj[r];
        sdst[r] = ly * HX + jj;
        int reach = 0;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const int p = (c + ly) & 1;
            sleft[c][r] = sdst[r] + p - 1;    // rows 1 .. LY-2: every neighbour index stays inside the array
            const int lx = 2 * jj + p;
            const int dist = min(min(ly, LY - 1 - ly), min(lx, LX - 1 - lx));
            if (cin[c][r])
                reach |= (dist + 1) << (8 * c);
            sinv[c][r] = inv_or_zero((cl[c][r] + cr[c][r]) + (cu[c][r] + cd[c][r]));
        }
        sreach[r] = reach;
    }
    // ---- 2 nu half-sweeps (nu <= NS), each region one cell smaller than the last; with nu < NS the
    // outer 2 (NS - nu) rings of the halo are loaded but never updated ----
    const int skip = 2 * (NS - nu);
    // A half-sweep reads only the other colour, so all neighbour values are gathered first and the
    // updates written afterwards: one LDS latency per half-sweep instead of one per round.  The colour
    // is a compile-time constant inside a half-sweep, so per-colour registers are addressed directly.
    auto half = [&](auto color_tag, int k) {
        constexpr int color = decltype(color_tag)::value;
        const int lo = k + 1 + skip;
        const bool neighbours = !(from_zero && k == 0);
        float* mine = s_col[color];
        const float* other = s_col[color ^ 1];
        float n0[ROUNDS], n1[ROUNDS], n2[ROUNDS], n3[ROUNDS];
#pragma unroll
        for (int r = 0; r < ROUNDS; ++r) {
            n0[r] = other[sleft[color][r]];
            n1[r] = other[sleft[color][r] + 1];
            n2[r] = other[sdst[r] - HX];
            n3[r] = other[sdst[r] + HX];
        }
#pragma unroll
        for (int r = 0; r < ROUNDS; ++r) {
            const float a = cl[color][r], b = cr[color][r], u = cu[color][r], d = cd[color][r];
            float v = cb[color][r];
            if (neighbours)
                v += (a * n0[r] + b * n1[r]) + (u * n2[r] + d * n3[r]);
            if (((sreach[r] >> (8 * color)) & 0xff) > lo)
                mine[sdst[r]] = v * sinv[color][r];
        }
        __syncthreads();
    };
#pragma unroll 1
    for (int k = 0; k < 2 * nu; k += 2) {
        if (c0 == 0) {
            half(std::integral_constant<int, 0>{}, k);
            half(std::integral_constant<int, 1>{}, k + 1);
        } else {
            half(std::integral_constant<int, 1>{}, k);
            half(std::integral_constant<int, 0>{}, k + 1);
        }
    }
    V* sol_out = L.sol[cur ^ 1];
    for (int e = threadIdx.x; e < TX * TY; e += NT) {
        const int ly = HALO + e / TX, lx = HALO + e % TX;
        const int x = x0 + lx, y = y0 + ly;
        if (x < w && y < h)
            sol_out[(y + 1) * s + (x + 1)] = stv<V>(s_col[(lx + ly) & 1][ly * HX + (lx >> 1)]);
    }
}

// ---------------------------------------------------------------------------------------------
// Chebyshev-weighted damped Jacobi with optional fused prolongation (the reference WebGPU multigrid's
// polynomial smoother, PressureMultigridGPU.cpp:24-28, 619-631; PressureSmoothingKernels.h:91-176).
// ---------------------------------------------------------------------------------------------
// The tile and a HALO-cell ring live in two LDS buffers (sweep k reads buffer k & 1 and writes the other).
// Each thread owns fixed cells; their four conductances, rhs and 1/diagonal stay in registers for all
// sweeps.  Sweep k updates the cells at least k + 1 + (HALO - nu) cells from the tile edge, so after nu
// sweeps the interior holds exactly what nu global sweeps compute.  Cells outside the grid hold zero in both
// buffers and are never updated.  MODE 0: from zero (the first sweep is w b / diag); 1: post-smoothing,
// the prolonged coarse correction added to the solution first.
template <int TX, int TY, int MODE, int NT, int HALO>
__global__ __launch_bounds__(NT) void k_jacobi_multi(LevelView L, int cur, const float* rhs, const double* shift,
    SmoothWeights wt, int nu, CoarseLink link, const int* status)
{
    constexpr int LX = TX + 2 * HALO, LY = TY + 2 * HALO;
    constexpr int NE = LX * LY;
    constexpr int ROUNDS = (NE + NT - 1) / NT;
    __shared__ float s_x[2][NE];
    const int w = L.w, h = L.h, s = L.s;
    int bx, by;
    xcd_tile((w + TX - 1) / TX, (h + TY - 1) / TY, bx, by);
    const int x0 = bx * TX - HALO, y0 = by * TY - HALO;
    const float* sol_in = L.sol[cur];
    // every global load goes through a clamped (valid) interior index and is masked afterwards
    float cl[ROUNDS], cr[ROUNDS], cu[ROUNDS], cd[ROUNDS], cb[ROUNDS], xv[ROUNDS];
    int ix[ROUNDS], iy[ROUNDS];
    bool cin[ROUNDS];
#pragma unroll
    for (int r = 0; r < ROUNDS; ++r) {
        const int e = threadIdx.x + r * NT;
        const int ly = e / LX, lx = e - ly * LX;
        const int x = x0 + lx, y = y0 + ly;
        cin[r] = e < NE && x >= 0 && x < w && y >= 0 && y < h;
        ix[r] = min(max(x, 0), w - 1);
        iy[r] = min(max(y, 0), h - 1);
        const int P = (iy[r] + 1) * s + (ix[r] + 1);
        cl[r] = L.right[P - 1];
        cr[r] = L.right[P];
        cu[r] = L.down[P - s];
        cd[r] = L.down[P];
        cb[r] = rhs[P];
        xv[r] = MODE == 0 ? 0.f : sol_in[P];
    }
    if (MODE == 1) {
#pragma unroll
        for (int r = 0; r < ROUNDS; ++r) {
            const float wx = link.pxw[ix[r]], wy = link.pyw[iy[r]];
            const float* top = link.csol + (link.pyf[iy[r]] + 1) * link.cs + 1;
            const float* bot = link.csol + (link.pys[iy[r]] + 1) * link.cs + 1;
            const int a0 = link.pxf[ix[r]], a1 = link.pxs[ix[r]];
            const float a = (1.f - wx) * top[a0] + wx * top[a1];
            const float b = (1.f - wx) * bot[a0] + wx * bot[a1];
            xv[r] += (1.f - wy) * a + wy * b;
        }
    }
    if (status && *status != kRunning)
        return;
    const double sh = shift ? *shift : 0.0;
    float inv[ROUNDS];
    int reach[ROUNDS];   // distance to the tile edge + 1; 0 for a cell outside the grid
#pragma unroll
    for (int r = 0; r < ROUNDS; ++r) {
        const int e = threadIdx.x + r * NT;
        const int ly = e / LX, lx = e - ly * LX;
        if (!cin[r]) {
            cl[r] = cr[r] = cu[r] = cd[r] = cb[r] = xv[r] = 0.f;
        } else if (shift) {
            cb[r] = (float)((double)cb[r] - sh);
        }
        inv[r] = inv_or_zero((cl[r] + cr[r]) + (cu[r] + cd[r]));
        reach[r] = cin[r] ? min(min(ly, LY - 1 - ly), min(lx, LX - 1 - lx)) + 1 : 0;
        if (e < NE) {
            s_x[0][e] = xv[r];
            s_x[1][e] = xv[r];
        }
    }
    __syncthreads();
    const int skip = HALO - nu;
#pragma unroll 1
    for (int k = 0; k < nu; ++k) {
        float wk = wt.w[0];   // uniform select: no dynamic indexing of the argument block
#pragma unroll
        for (int j = 1; j < kJacobiMaxSweeps; ++j)
            wk = k == j ? wt.w[j] : wk;
        const float* src = s_x[k & 1];
        float* dst = s_x[(k & 1) ^ 1];
        const bool neighbours = !(MODE == 0 && k == 0);
#pragma unroll
        for (int r = 0; r < ROUNDS; ++r) {
            if (reach[r] > k + 1 + skip) {
                const int e = threadIdx.x + r * NT;
                float v;
                if (neighbours) {
                    const float c = src[e];
                    const float product = (cl[r] * (c - src[e - 1]) + cr[r] * (c - src[e + 1])) +
                                          (cu[r] * (c - src[e - LX]) + cd[r] * (c - src[e + LX]));
                    v = c + wk * (cb[r] - product) * inv[r];
                } else {
                    v = wk * cb[r] * inv[r];
                }
                dst[e] = inv[r] > 0.f ? v : 0.f;
            }
        }
        __syncthreads();
    }
    const float* fin = s_x[nu & 1];
    float* sol_out = L.sol[cur ^ 1];
    for (int e = threadIdx.x; e < TX * TY; e += NT) {
        const int ly = HALO + e / TX, lx = HALO + e % TX;
        const int x = x0 + lx, y = y0 + ly;
        if (x < w && y < h)
            sol_out[(y + 1) * s + (x + 1)] = fin[ly * LX + lx];
    }
}

// ---------------------------------------------------------------------------------------------
// Residual + restriction (restrictResidual, PressureMultigrid.cpp:233-276).  A coarse tile of
// CTX x CTY cells gathers fine taps in [2c - 2, 2c + 3]; the fine residual is formed in LDS on
// that window (plus a one-cell ring of solution values).
// ---------------------------------------------------------------------------------------------
template <int CTX, int CTY, class C = float, class V = float, class VC = float>
__global__ __launch_bounds__(kNT) void k_res_restrict(LevelViewT<C, V> F, int cur, const V* rhs, const double* shift,
    LevelViewT<C, VC> Cl, AxisRestrict rx, AxisRestrict ry, const int* status)
{
    constexpr int LX = 2 * CTX + 6, LY = 2 * CTY + 6;
    __shared__ float s_sol[LY * LX];
    __shared__ float s_right[LY * LX];
    __shared__ float s_down[LY * LX];
    __shared__ float s_res[LY * LX];
    __shared__ int s_rxc[CTX], s_rxi[CTX * 4], s_ryc[CTY], s_ryi[CTY * 4];
    __shared__ float s_rxw[CTX * 4], s_ryw[CTY * 4];
    const int w = F.w, h = F.h, s = F.s;
    int bx, by;
    xcd_tile((Cl.w + CTX - 1) / CTX, (Cl.h + CTY - 1) / CTY, bx, by);
    const int cx0 = bx * CTX, cy0 = by * CTY;
    const int fx0 = 2 * cx0 - 3, fy0 = 2 * cy0 - 3;
    const V* sol = F.sol[cur];
    constexpr int NE = LX * LY;
    constexpr int LOADS = (NE + kNT - 1) / kNT;
    float vs[LOADS], vr[LOADS], vd[LOADS], vb[LOADS];
    bool vin[LOADS];
#pragma unroll
    for (int it = 0; it < LOADS; ++it) {
        const int e = threadIdx.x + it * kNT;
        const int ly = e / LX, lx = e - ly * LX;
        const int x = fx0 + lx, y = fy0 + ly;
        vin[it] = e < NE && x >= 0 && x < w && y >= 0 && y < h;
        const int P = (min(max(y, 0), h - 1) + 1) * s + (min(max(x, 0), w - 1) + 1);
        vs[it] = ldv(sol[P]);
        vr[it] = ldv(F.right[P]);
        vd[it] = ldv(F.down[P]);
        vb[it] = ldv(rhs[P]);
    }
    // restriction tables of this coarse tile (missing taps: weight 0 at a valid LDS cell)
    int tc[2], ti[2][4];
    float tw[2][4];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        const int e = threadIdx.x - a * CTX;  // threads [0, CTX) take x, [CTX, CTX + CTY) take y
        const bool isx = a == 0;
        const int n_axis = isx ? CTX : CTY;
        const int c = (isx ? cx0 : cy0) + min(max(e, 0), n_axis - 1);
        const int limit = isx ? Cl.w : Cl.h;
        const AxisRestrict& R = isx ? rx : ry;
        const int cc = min(c, limit - 1);
        const int n = R.count[cc];
        tc[a] = c < limit ? n : 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            ti[a][k] = R.index[cc * 4 + k];
            tw[a][k] = R.weight[cc * 4 + k];
        }
    }
    if (status && *status != kRunning)
        return;
    const double sh = shift ? *shift : 0.0;
#pragma unroll
    for (int it = 0; it < LOADS; ++it) {
        const int e = threadIdx.x + it * kNT;
        if (e < NE) {
            s_sol[e] = vin[it] ? vs[it] : 0.f;
            s_right[e] = vin[it] ? vr[it] : 0.f;
            s_down[e] = vin[it] ? vd[it] : 0.f;
            s_res[e] = vin[it] ? (shift ? (float)((double)vb[it] - sh) : vb[it]) : 0.f;
        }
    }
    if (threadIdx.x < CTX) {
        const int e = threadIdx.x;
        s_rxc[e] = tc[0];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const bool on = k < tc[0];
            s_rxi[e * 4 + k] = on ? ti[0][k] - fx0 : 1;
            s_rxw[e * 4 + k] = on ? tw[0][k] : 0.f;
        }
    } else if (threadIdx.x < CTX + CTY) {
        const int e = threadIdx.x - CTX;
        s_ryc[e] = tc[1];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const bool on = k < tc[1];
            s_ryi[e * 4 + k] = on ? ti[1][k] - fy0 : 1;
            s_ryw[e * 4 + k] = on ? tw[1][k] : 0.f;
        }
    }
    __syncthreads();
    // residual in place of the rhs on the inner window
    for (int e = threadIdx.x; e < (LX - 2) * (LY - 2); e += kNT) {
        const int ly = 1 + e / (LX - 2), lx = 1 + e % (LX - 2);
        const int x = fx0 + lx, y = fy0 + ly;
        if (x < 0 || x >= w || y < 0 || y >= h)
            continue;
        const int i = ly * LX + lx;
        const float c = s_sol[i];
        const float product = (s_right[i - 1] * (c - s_sol[i - 1]) + s_right[i] * (c - s_sol[i + 1])) +
                              (s_down[i - LX] * (c - s_sol[i - LX]) + s_down[i] * (c - s_sol[i + LX]));
        // a cell without conductance (solid masks only) is not restricted: the V-cycle stays symmetric
        const bool active = (s_right[i - 1] + s_right[i]) + (s_down[i - LX] + s_down[i]) > 0.f;
        s_res[i] = active ? s_res[i] - product : 0.f;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < CTX * CTY; e += kNT) {
        const int ly = e / CTX, lx = e - ly * CTX;
        const int x = cx0 + lx, y = cy0 + ly;
        if (x >= Cl.w || y >= Cl.h)
            continue;
        // four taps per axis (missing taps carry weight 0 and a valid index): 16 independent reads
        float tap[4][4];
#pragma unroll
        for (int iy = 0; iy < 4; ++iy) {
            const float* row = s_res + s_ryi[ly * 4 + iy] * LX;
#pragma unroll
            for (int ix = 0; ix < 4; ++ix)
                tap[iy][ix] = row[s_rxi[lx * 4 + ix]];
        }
        float value = 0.f;
#pragma unroll
        for (int iy = 0; iy < 4; ++iy) {
            float horizontal = 0.f;
#pragma unroll
            for (int ix = 0; ix < 4; ++ix)
                horizontal += s_rxw[lx * 4 + ix] * tap[iy][ix];
            value += s_ryw[ly * 4 + iy] * horizontal;
        }
        Cl.rhs[(y + 1) * Cl.s + x + 1] = stv<VC>(value);
    }
}

// ---------------------------------------------------------------------------------------------
// Pre-smoothing from zero with the Chebyshev-weighted Jacobi sweeps, fused with the residual and the
// restriction that follow it, for one coarse tile of CTX x CTY cells.  The smoothing tile is the coarse
// tile's fine image widened by the restriction's reach and the residual's ring (RING = 3: taps in
// [2c - 2, 2c + 3], one more cell for the operator) and by HALO (>= nu) cells of smoothing halo; with
// k_jacobi_multi's update rule its cells at least that halo from the edge hold exactly what nu global
// sweeps compute.  The fine image's cells are written as the pre-smoothed solution, the residual of the
// restriction window is formed from LDS with k_res_restrict's expression and restricted with its sums in
// the same order (a missing tap points at the same cell) -- the values of the two separate launches, in
// one.  It pays on the levels where launches, not bytes, set the time; the autotuner decides per level.
// ---------------------------------------------------------------------------------------------
template <int CTX, int CTY, int NT, int HALO>
__global__ __launch_bounds__(NT) void k_jacobi_res_restrict(LevelView L, int cur, const float* rhs,
    const double* shift, SmoothWeights wt, int nu, LevelView Cl, AxisRestrict rx, AxisRestrict ry,
    const int* status)
{
    constexpr int RING = 3;
    constexpr int EDGE = RING + HALO;
    constexpr int LX = 2 * CTX + 2 * EDGE, LY = 2 * CTY + 2 * EDGE;
    constexpr int NE = LX * LY;
    constexpr int ROUNDS = (NE + NT - 1) / NT;
    static_assert(CTX + CTY <= NT, "one thread per restriction table row");
    __shared__ float s_x[2][NE];
    __shared__ int s_rxi[CTX * 4], s_ryi[CTY * 4];
    __shared__ float s_rxw[CTX * 4], s_ryw[CTY * 4];
    const int w = L.w, h = L.h, s = L.s;
    int bx, by;
    xcd_tile((Cl.w + CTX - 1) / CTX, (Cl.h + CTY - 1) / CTY, bx, by);
    const int cx0 = bx * CTX, cy0 = by * CTY;
    const int x0 = 2 * cx0 - EDGE, y0 = 2 * cy0 - EDGE;
    // every global load goes through a clamped (valid) interior index and is masked afterwards
    float cl[ROUNDS], cr[ROUNDS], cu[ROUNDS], cd[ROUNDS], cb[ROUNDS];
    bool cin[ROUNDS];
#pragma unroll
    for (int r = 0; r < ROUNDS; ++r) {
        const int e = threadIdx.x + r * NT;
        const int ly = e / LX, lx = e - ly * LX;
        const int x = x0 + lx, y = y0 + ly;
        cin[r] = e < NE && x >= 0 && x < w && y >= 0 && y < h;
        const int P = (min(max(y, 0), h - 1) + 1) * s + (min(max(x, 0), w - 1) + 1);
        cl[r] = L.right[P - 1];
        cr[r] = L.right[P];
        cu[r] = L.down[P - s];
        cd[r] = L.down[P];
        cb[r] = rhs[P];
    }
    // restriction tables of this coarse tile, as k_res_restrict stages them (indices relative to x0, y0;
    // a missing tap: weight 0 at the cell k_res_restrict's missing taps read)
    int tc[2], ti[2][4];
    float tw[2][4];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        const int e = threadIdx.x - a * CTX;  // threads [0, CTX) take x, [CTX, CTX + CTY) take y
        const bool isx = a == 0;
        const int n_axis = isx ? CTX : CTY;
        const int c = (isx ? cx0 : cy0) + min(max(e, 0), n_axis - 1);
        const int limit = isx ? Cl.w : Cl.h;
        const AxisRestrict& R = isx ? rx : ry;
        const int cc = min(c, limit - 1);
        const int n = R.count[cc];
        tc[a] = c < limit ? n : 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            ti[a][k] = R.index[cc * 4 + k];
            tw[a][k] = R.weight[cc * 4 + k];
        }
    }
    if (status && *status != kRunning)
        return;
    const double sh = shift ? *shift : 0.0;
    float inv[ROUNDS];
    int reach[ROUNDS];   // distance to the tile edge + 1; 0 for a cell outside the grid
#pragma unroll
    for (int r = 0; r < ROUNDS; ++r) {
        const int e = threadIdx.x + r * NT;
        const int ly = e / LX, lx = e - ly * LX;
        if (!cin[r]) {
            cl[r] = cr[r] = cu[r] = cd[r] = cb[r] = 0.f;
        } else if (shift) {
            cb[r] = (float)((double)cb[r] - sh);
        }
        inv[r] = inv_or_zero((cl[r] + cr[r]) + (cu[r] + cd[r]));
        reach[r] = cin[r] ? min(min(ly, LY - 1 - ly), min(lx, LX - 1 - lx)) + 1 : 0;
        if (e < NE) {
            s_x[0][e] = 0.f;
            s_x[1][e] = 0.f;
        }
    }
    constexpr int MISS = EDGE - 2;   // k_res_restrict's missing-tap cell, 2c0 - 2 on each axis
    if (threadIdx.x < CTX) {
        const int e = threadIdx.x;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const bool on = k < tc[0];
            s_rxi[e * 4 + k] = on ? ti[0][k] - x0 : MISS;
            s_rxw[e * 4 + k] = on ? tw[0][k] : 0.f;
        }
    } else if (threadIdx.x < CTX + CTY) {
        const int e = threadIdx.x - CTX;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const bool on = k < tc[1];
            s_ryi[e * 4 + k] = on ? ti[1][k] - y0 : MISS;
            s_ryw[e * 4 + k] = on ? tw[1][k] : 0.f;
        }
    }
    __syncthreads();
    const int skip = HALO - nu;
#pragma unroll 1
    for (int k = 0; k < nu; ++k) {
        float wk = wt.w[0];   // uniform select: no dynamic indexing of the argument block
#pragma unroll
        for (int j = 1; j < kJacobiMaxSweeps; ++j)
            wk = k == j ? wt.w[j] : wk;
        const float* src = s_x[k & 1];
        float* dst = s_x[(k & 1) ^ 1];
        const bool neighbours = k != 0;
#pragma unroll
        for (int r = 0; r < ROUNDS; ++r) {
            if (reach[r] > k + 1 + skip) {
                const int e = threadIdx.x + r * NT;
                float v;
                if (neighbours) {
                    const float c = src[e];
                    const float product = (cl[r] * (c - src[e - 1]) + cr[r] * (c - src[e + 1])) +
                                          (cu[r] * (c - src[e - LX]) + cd[r] * (c - src[e + LX]));
                    v = c + wk * (cb[r] - product) * inv[r];
                } else {
                    v = wk * cb[r] * inv[r];
                }
                dst[e] = inv[r] > 0.f ? v : 0.f;
            }
        }
        __syncthreads();
    }
    const float* fin = s_x[nu & 1];
    float* res = s_x[(nu & 1) ^ 1];
    float* sol_out = L.sol[cur ^ 1];
    // the pre-smoothed solution of the coarse tile's fine image, and the residual of the restriction
    // window ([2 c0 - 2, 2 c1 + 2) on each axis; zero outside the grid and at cells without conductance)
#pragma unroll
    for (int r = 0; r < ROUNDS; ++r) {
        const int e = threadIdx.x + r * NT;
        const int ly = e / LX, lx = e - ly * LX;
        if (e >= NE)
            continue;
        if (cin[r] && lx >= EDGE && lx < EDGE + 2 * CTX && ly >= EDGE && ly < EDGE + 2 * CTY)
            sol_out[(y0 + ly + 1) * s + (x0 + lx + 1)] = fin[e];
        if (lx >= EDGE - 2 && lx < EDGE + 2 * CTX + 2 && ly >= EDGE - 2 && ly < EDGE + 2 * CTY + 2) {
            float v = 0.f;
            if (cin[r]) {
                const float c = fin[e];
                const float product = (cl[r] * (c - fin[e - 1]) + cr[r] * (c - fin[e + 1])) +
                                      (cu[r] * (c - fin[e - LX]) + cd[r] * (c - fin[e + LX]));
                const bool active = (cl[r] + cr[r]) + (cu[r] + cd[r]) > 0.f;
                v = active ? cb[r] - product : 0.f;
            }
            res[e] = v;
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < CTX * CTY; e += NT) {
        const int ly = e / CTX, lx = e - ly * CTX;
        const int x = cx0 + lx, y = cy0 + ly;
        if (x >= Cl.w || y >= Cl.h)
            continue;
        float tap[4][4];
#pragma unroll
        for (int iy = 0; iy < 4; ++iy) {
            const float* row = res + s_ryi[ly * 4 + iy] * LX;
#pragma unroll
            for (int ix = 0; ix < 4; ++ix)
                tap[iy][ix] = row[s_rxi[lx * 4 + ix]];
        }
        float value = 0.f;
#pragma unroll
        for (int iy = 0; iy < 4; ++iy) {
            float horizontal = 0.f;
#pragma unroll
            for (int ix = 0; ix < 4; ++ix)
                horizontal += s_rxw[lx * 4 + ix] * tap[iy][ix];
            value += s_ryw[ly * 4 + iy] * horizontal;
        }
        Cl.rhs[(y + 1) * Cl.s + x + 1] = value;
    }
}

// ---------------------------------------------------------------------------------------------
// Coarse factorisation (factorCoarsest, PressureMultigrid.cpp:300-345) and inverse.
// Right-looking updates applied column by column: every entry receives its subtractions in the
// same k order as the reference's row loop, so the factor is bit-identical to it.
// ---------------------------------------------------------------------------------------------
constexpr int kMaxCoarse = 63;
constexpr int kFactorThreads = 256;

__global__ __launch_bounds__(kFactorThreads) void k_coarse_factor(int w, int h, const double* right,
    const double* down, double* factor, float* ginv, int n, PcgScalars* S, int pin)
{
    __shared__ double a[kMaxCoarse * kMaxCoarse];   // Cholesky working matrix (lower triangle)
    __shared__ double m[kMaxCoarse * kMaxCoarse];   // Gauss-Jordan working matrix
    __shared__ double q[kMaxCoarse];
    __shared__ double assembled[kMaxCoarse];        // diagonal before elimination (pinning bound)
    __shared__ int failed;
    const int s = w + 2;
    const int tid = threadIdx.x;
    const int tc = tid & 63, tr = tid >> 6;
    constexpr int RSTEP = kFactorThreads / 64;
    for (int e = tid; e < n * n; e += kFactorThreads) {
        a[e] = 0.0;
        m[e] = 0.0;
    }
    if (tid == 0)
        failed = 0;
    __syncthreads();
    for (int cell = 1 + tid; cell <= n; cell += kFactorThreads) {
        const int x = cell % w, y = cell / w;
        const int i = (y + 1) * s + x + 1;
        const int row = cell - 1;
        const double diag = right[i - 1] + right[i] + down[i - s] + down[i];
        a[row * n + row] = diag;
        m[row * n + row] = diag;
        assembled[row] = diag;
        if (x > 0 && cell - 1 > 0) {
            a[row * n + row - 1] = -right[i - 1];
            m[row * n + row - 1] = -right[i - 1];
            m[(row - 1) * n + row] = -right[i - 1];
        }
        if (y > 0 && cell - w > 0) {
            a[row * n + row - w] = -down[i - s];
            m[row * n + row - w] = -down[i - s];
            m[(row - w) * n + row] = -down[i - s];
        }
    }
    __syncthreads();
    // Column k: pivot sqrt, quotients L[i][k] = a[i][k] / pivot into q (one barrier), then the trailing
    // update a[i][c] -= L[i][k] L[c][k] (second barrier).  Each entry still receives its subtractions
    // in increasing k, as the reference's row loop applies them.
    unsigned long long pinned = 0ull;   // the same in every thread
    for (int k = 0; k < n; ++k) {
        const double v = a[k * n + k];
        // a pinned unknown gets a zero column: the trailing update below then subtracts exact zeros
        const bool pin_k = pin && isfinite(v) && fabs(v) <= kPinRatio * assembled[k];
        if (!pin_k && (v <= 0.0 || !isfinite(v))) {
            if (tid == 0)
                failed = 1;
            break;
        }
        if (pin_k)
            pinned |= 1ull << k;
        const double pivot = pin_k ? 0.0 : sqrt(v);
        if (tid > k && tid < n) {
            const double l = pin_k ? 0.0 : a[tid * n + k] / pivot;
            a[tid * n + k] = l;
            q[tid] = l;
        }
        __syncthreads();
        if (tid == 0)
            a[k * n + k] = pivot;  // after the barrier: every thread has read the unfactored diagonal
        {
            // rows i = k+1+tr+4m; gather every operand first, then write (independent LDS accesses)
            const int c = k + 1 + tc;
            const int cc = min(c, n - 1);
            const double qc = q[cc];
            constexpr int M = (kMaxCoarse + RSTEP - 1) / RSTEP;
            double li[M], av[M];
#pragma unroll
            for (int mm = 0; mm < M; ++mm) {
                const int ic = min(k + 1 + tr + RSTEP * mm, n - 1);
                li[mm] = q[ic];
                av[mm] = a[ic * n + cc];
            }
#pragma unroll
            for (int mm = 0; mm < M; ++mm) {
                const int i = k + 1 + tr + RSTEP * mm;
                if (i < n && c <= i)
                    a[i * n + c] = av[mm] - li[mm] * qc;
            }
        }
        __syncthreads();
    }
    __syncthreads();
    if (failed) {
        if (tid == 0 && S) {
            S->mg_failed = 1;
            if (S->status == kRunning)
                S->status = kFailed;
        }
        return;
    }
    for (int e = tid; e < n * n; e += kFactorThreads) {
        const int r = e / n, c = e - r * n;
        factor[e] = c <= r && !((pinned >> r) & 1ull) ? a[e] : 0.0;
    }
    if (!ginv || pinned)
        return;   // the Gauss-Jordan inverse below assumes no pinned unknown (callers pass no ginv with pin)
    // In-place Gauss-Jordan inversion of the (SPD, so pivot-free) pinned matrix.
    for (int k = 0; k < n; ++k) {
        const double p = m[k * n + k];
        __syncthreads();  // every thread holds the pivot before row k is rewritten
        if (tid < n) {
            q[tid] = m[tid * n + k];
            m[k * n + tid] = tid == k ? 1.0 / p : m[k * n + tid] / p;
        }
        __syncthreads();
        {
            constexpr int M = (kMaxCoarse + RSTEP - 1) / RSTEP;
            const int tcc = min(tc, n - 1);
            const double pk = m[k * n + tcc];
            double fv[M], mv[M];
#pragma unroll
            for (int mm = 0; mm < M; ++mm) {
                const int ic = min(tr + RSTEP * mm, n - 1);
                fv[mm] = q[ic];
                mv[mm] = m[ic * n + tcc];
            }
#pragma unroll
            for (int mm = 0; mm < M; ++mm) {
                const int i = tr + RSTEP * mm;
                if (i < n && i != k && tc < n)
                    m[i * n + tc] = tc == k ? -fv[mm] / p : mv[mm] - fv[mm] * pk;
            }
        }
        __syncthreads();
    }
    for (int e = tid; e < n * n; e += kFactorThreads)
        ginv[e] = (float)m[e];
}

// ---------------------------------------------------------------------------------------------
// Inverse of the pinned coarse matrix for the per-cycle coarse solve (the hot path).  Goodnight's
// sweep operator, in place, one pivot per step: the matrix lives in registers, column j in lane j
// and rows split over sixteen wavefronts (4 per wavefront, so that a step's row updates are a short
// chain).  A step broadcasts the pivot row through LDS (double-buffered, one barrier) and the pivot
// column inside each wavefront by lane reads, so the 63 steps cost 63 barriers and no LDS matrix
// traffic.  Pivots are the Schur complement diagonals (the squared Cholesky pivots); a non-positive or
// non-finite one marks the hierarchy failed exactly where the Cholesky factorisation would.
// ---------------------------------------------------------------------------------------------
constexpr int kSweepRows = 4;     // rows per wavefront
constexpr int kSweepWaves = 16;   // kSweepWaves x kSweepRows >= kMaxCoarse

// v as held by lane k (k uniform across the wavefront).
__device__ __forceinline__ double lane_value(double v, int k)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, k);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), k);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

__device__ void coarse_inverse(int w, int h, const double* right, const double* down, float* ginv, int n,
    PcgScalars* S, int pin)
{
    __shared__ double rowbuf[2][64];
    __shared__ double assembled[64];   // diagonal before elimination (pinning bound)
    const int s = w + 2;
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    // pinned matrix entry (i, j) of unknowns i, j (cells i + 1, j + 1)
    double a[kSweepRows];
    const int cj = lane + 1;
    const int xj = cj % w, yj = cj / w;
#pragma unroll
    for (int r = 0; r < kSweepRows; ++r) {
        const int ci = wv * kSweepRows + r + 1;
        double v = 0.0;
        if (ci <= n && cj <= n) {
            const int xi = ci % w, yi = ci / w;
            const int pi = (yi + 1) * s + xi + 1;
            if (ci == cj) {
                v = right[pi - 1] + right[pi] + down[pi - s] + down[pi];
                assembled[lane] = v;
            } else if (yi == yj && (xi - xj == 1 || xj - xi == 1)) {
                v = -right[(yi + 1) * s + min(xi, xj) + 1];
            } else if (xi == xj && (yi - yj == 1 || yj - yi == 1)) {
                v = -down[(min(yi, yj) + 1) * s + xi + 1];
            }
        }
        a[r] = v;
    }
    bool failed = false;
    // Solid masks: an unknown whose Schur pivot vanishes is pinned -- never swept, and its row and
    // column of the inverse are zero (the swept block is then the inverse of the other unknowns' matrix)
    unsigned long long pinned = 0ull;   // the same in every thread
    for (int kb = 0; kb < kSweepWaves && !failed; ++kb) {
#pragma unroll
        for (int r0 = 0; r0 < kSweepRows; ++r0) {
            const int k = kb * kSweepRows + r0;
            if (k >= n || failed)
                break;
            double* row = rowbuf[k & 1];
            if (wv == kb)
                row[lane] = a[r0];
            __syncthreads();   // also orders the assembled[] writes before the first read
            const double d = row[k];
            if (pin && isfinite(d) && fabs(d) <= kPinRatio * assembled[k]) {
                pinned |= 1ull << k;
                continue;
            }
            if (!(d > 0.0) || !isfinite(d)) {
                failed = true;   // d is the same in every thread: the whole block leaves together
                break;
            }
            const double inv_d = 1.0 / d;
            const double t = lane == k ? inv_d : row[lane] * inv_d;
#pragma unroll
            for (int r = 0; r < kSweepRows; ++r) {
                const int i = wv * kSweepRows + r;
                const double b = lane_value(a[r], k);
                const double base = lane == k ? 0.0 : a[r];
                a[r] = i == k ? t : base - b * t;
            }
        }
    }
    if (failed) {
        if (threadIdx.x == 0 && S) {
            S->mg_failed = 1;
            if (S->status == kRunning)
                S->status = kFailed;
        }
        return;
    }
#pragma unroll
    for (int r = 0; r < kSweepRows; ++r) {
        const int i = wv * kSweepRows + r;
        if (i < n && lane < n)
            ginv[i * n + lane] = ((pinned >> i) & 1ull) || ((pinned >> lane) & 1ull) ? 0.f : (float)a[r];
    }
}

__global__ __launch_bounds__(64 * kSweepWaves) void k_coarse_inverse(int w, int h, const double* right,
    const double* down, float* ginv, int n, PcgScalars* S, int pin)
{
    coarse_inverse(w, h, right, down, ginv, n, S, pin);
}

// Coarse end of the hierarchy rebuild: each level's cells coarsened, then its diagonal and FP32 copies,
// each phase workgroup-strided and closed by a barrier (which also makes this workgroup's global
// writes visible to its next phase); then the coarse inverse.
__global__ __launch_bounds__(64 * kSweepWaves) void k_coarse_chain(const ChainDesc C, float* ginv, int n,
    PcgScalars* S, int pin)
{
    for (int l = 0; l < C.levels; ++l) {
        const ChainLevel& L = C.lev[l];
        const int cells = L.cw * L.ch;
        for (int t = threadIdx.x; t < cells; t += 64 * kSweepWaves) {
            const int y = t / L.cw, x = t - y * L.cw;
            coarsen_cell(x, y, L.fw, L.fh, L.fcx, L.fcy, L.fright, L.fdown, L.cw, L.ch, L.ccx, L.ccy, L.cright,
                L.cdown);
        }
        __syncthreads();
        const int s = L.cw + 2;
        for (int t = threadIdx.x; t < cells; t += 64 * kSweepWaves) {
            const int y = t / L.cw, x = t - y * L.cw;
            const int i = (y + 1) * s + x + 1;
            level_diag_values(L.cright[i - 1], L.cright[i], L.cdown[i - s], L.cdown[i], L.invd32[i], L.right32[i],
                L.down32[i]);
        }
        __syncthreads();
    }
    if (n > 0) {
        const ChainLevel& L = C.lev[C.levels - 1];
        coarse_inverse(L.cw, L.ch, L.cright, L.cdown, ginv, n, S, pin);
    }
}

// ---------------------------------------------------------------------------------------------
// Coarse tail: the deepest levels and the coarse solve in one workgroup.
// ---------------------------------------------------------------------------------------------
constexpr int kTailThreads = 1024;

struct TailArrays {
    float* right;
    float* down;
    float* rhs;
    float* sol;
    float* res;
    int* pxf; int* pxs; float* pxw;
    int* pyf; int* pys; float* pyw;
    int* rxc; int* rxi; float* rxw;
    int* ryc; int* ryi; float* ryw;
};

__device__ __forceinline__ TailArrays tail_arrays(float* lds, const TailLevel& T)
{
    TailArrays A;
    const int ps = T.s * (T.h + 2);
    float* p = lds + T.lds_arrays;
    A.right = p;
    A.down = p + ps;
    A.rhs = p + 2 * ps;
    A.sol = p + 3 * ps;
    A.res = p + 4 * ps;
    float* t = lds + T.lds_tables;
    A.pxf = (int*)t; t += T.w;
    A.pxs = (int*)t; t += T.w;
    A.pxw = t; t += T.w;
    A.pyf = (int*)t; t += T.h;
    A.pys = (int*)t; t += T.h;
    A.pyw = t; t += T.h;
    A.rxc = (int*)t; t += T.cw;
    A.rxi = (int*)t; t += 4 * T.cw;
    A.rxw = t; t += 4 * T.cw;
    A.ryc = (int*)t; t += T.ch;
    A.ryi = (int*)t; t += 4 * T.ch;
    A.ryw = t;
    return A;
}

// Tail loops: a thread owns at most ITER items of a level (ITER chosen per level from its size, so a
// small level does not execute the unrolled work of a large one), and a wavefront whose first item is
// past the end skips the loop body entirely.  Every LDS read goes through a clamped (valid) index and
// only the stores are predicated, so the reads of all items issue together.
constexpr int kTailIter = 4;   // at most 4 * 1024 items per phase (the tail selection guarantees it)

__device__ __forceinline__ bool wave_has_items(int items) { return (int)(threadIdx.x & ~63u) < items; }

// t / d for 0 <= t < 2^22 without an integer division: (t + 1/2) / d lies at least 1/(2d) from an
// integer, and the float product's error is below (t / d) 2^-22 < 1/(2d).
__device__ __forceinline__ int tail_div(int t, float inv_d) { return (int)(((float)t + 0.5f) * inv_d); }

// `sweeps` (uniform, 1..4) red-black sweeps of one tail level with each thread's cells, conductances, rhs
// and 1/diagonal held in registers across all the half-sweeps: a half-sweep is four LDS reads, one write
// and a barrier per item.
template <int ITER>
__device__ void tail_smooth4_n(const TailLevel& T, const TailArrays& A, int reverse, int from_zero, int sweeps)
{
    const int w = T.w, h = T.h, s = T.s;
    const int pairs = (w + 1) >> 1;
    const int items = pairs * h;
    const bool active = wave_has_items(items);
    const float inv_pairs = 1.0f / (float)pairs;
    int idx[2][ITER];
    bool on[2][ITER];
    float wl[2][ITER], wr[2][ITER], wu[2][ITER], wd[2][ITER], rb[2][ITER], inv[2][ITER];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
#pragma unroll
        for (int it = 0; it < ITER; ++it) {
            const int t = min((int)threadIdx.x + it * kTailThreads, items - 1);
            const int y = tail_div(t, inv_pairs);
            const int x = 2 * (t - y * pairs) + ((c ^ y) & 1);
            on[c][it] = active && (int)threadIdx.x + it * kTailThreads < items && x < w;
            const int i = (y + 1) * s + min(x, w - 1) + 1;
            idx[c][it] = i;
            wl[c][it] = A.right[i - 1];
            wr[c][it] = A.right[i];
            wu[c][it] = A.down[i - s];
            wd[c][it] = A.down[i];
            rb[c][it] = A.rhs[i];
        }
    }
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int it = 0; it < ITER; ++it)
            inv[c][it] = inv_or_zero((wl[c][it] + wr[c][it]) + (wu[c][it] + wd[c][it]));
    for (int k = 0; k < sweeps; ++k) {
        for (int pass = 0; pass < 2; ++pass) {
            const int color = reverse ? 1 - pass : pass;
            const bool neighbours = !(from_zero && k == 0 && pass == 0);
            if (active) {
                float v[ITER];
#pragma unroll
                for (int it = 0; it < ITER; ++it) {
                    const int i = idx[color][it];
                    float acc = rb[color][it];
                    const float n = (wl[color][it] * A.sol[i - 1] + wr[color][it] * A.sol[i + 1]) +
                                    (wu[color][it] * A.sol[i - s] + wd[color][it] * A.sol[i + s]);
                    if (neighbours)
                        acc += n;
                    v[it] = acc * inv[color][it];
                }
#pragma unroll
                for (int it = 0; it < ITER; ++it)
                    if (on[color][it])
                        A.sol[idx[color][it]] = v[it];
            }
            __syncthreads();
        }
    }
}

// `sweeps` sweeps (PressureMultigrid.cpp:204-231 applied four times in the reference's cycle); reverse =
// post-smoothing colour order.
__device__ void tail_smooth4(const TailLevel& T, const TailArrays& A, int reverse, int from_zero, int sweeps)
{
    const int items = ((T.w + 1) >> 1) * T.h;
    if (items <= kTailThreads)
        tail_smooth4_n<1>(T, A, reverse, from_zero, sweeps);
    else if (items <= 2 * kTailThreads)
        tail_smooth4_n<2>(T, A, reverse, from_zero, sweeps);
    else
        tail_smooth4_n<kTailIter>(T, A, reverse, from_zero, sweeps);
}

template <int ITER> __device__ void tail_residual_n(const TailLevel& T, const TailArrays& F)
{
    const int w = T.w, h = T.h, s = T.s;
    const int cells = w * h;
    if (!wave_has_items(cells))
        return;
    const float inv_w = 1.0f / (float)w;
    float rv[ITER];
    int idx[ITER];
    bool on[ITER];
#pragma unroll
    for (int it = 0; it < ITER; ++it) {
        const int e = min((int)threadIdx.x + it * kTailThreads, cells - 1);
        on[it] = (int)threadIdx.x + it * kTailThreads < cells;
        const int y = tail_div(e, inv_w), x = e - y * w;
        const int i = (y + 1) * s + x + 1;
        idx[it] = i;
        const float c = F.sol[i];
        const float product = (F.right[i - 1] * (c - F.sol[i - 1]) + F.right[i] * (c - F.sol[i + 1])) +
                              (F.down[i - s] * (c - F.sol[i - s]) + F.down[i] * (c - F.sol[i + s]));
        const bool active = (F.right[i - 1] + F.right[i]) + (F.down[i - s] + F.down[i]) > 0.f;
        rv[it] = active ? F.rhs[i] - product : 0.f;   // cells without conductance are not restricted
    }
#pragma unroll
    for (int it = 0; it < ITER; ++it)
        if (on[it])
            F.res[idx[it]] = rv[it];
}

template <int ITER> __device__ void tail_restrict_n(const TailLevel& T, const TailArrays& F, const TailLevel& Cn,
    const TailArrays& Cd)
{
    // missing taps carry weight 0 and index -1, which addresses the zero halo
    const int s = T.s;
    const int cells_c = Cn.w * Cn.h;
    if (!wave_has_items(cells_c))
        return;
    const float inv_cw = 1.0f / (float)Cn.w;
    float val[ITER];
    int dst[ITER];
#pragma unroll
    for (int it = 0; it < ITER; ++it) {
        const int e = min((int)threadIdx.x + it * kTailThreads, cells_c - 1);
        const int y = tail_div(e, inv_cw), x = e - y * Cn.w;
        float value = 0.f;
#pragma unroll
        for (int iy = 0; iy < 4; ++iy) {
            const float* row = F.res + (F.ryi[y * 4 + iy] + 1) * s + 1;
            float horizontal = 0.f;
#pragma unroll
            for (int ix = 0; ix < 4; ++ix)
                horizontal += F.rxw[x * 4 + ix] * row[F.rxi[x * 4 + ix]];
            value += F.ryw[y * 4 + iy] * horizontal;
        }
        val[it] = value;
        dst[it] = (y + 1) * Cn.s + x + 1;
    }
#pragma unroll
    for (int it = 0; it < ITER; ++it)
        if ((int)threadIdx.x + it * kTailThreads < cells_c)
            Cd.rhs[dst[it]] = val[it];
}

__device__ void tail_residual_restrict(const TailLevel& T, const TailArrays& F, const TailLevel& Cn,
    const TailArrays& Cd)
{
    const int cells = T.w * T.h;
    if (cells <= kTailThreads)
        tail_residual_n<1>(T, F);
    else if (cells <= 2 * kTailThreads)
        tail_residual_n<2>(T, F);
    else
        tail_residual_n<kTailIter>(T, F);
    __syncthreads();
    if (Cn.w * Cn.h <= kTailThreads)
        tail_restrict_n<1>(T, F, Cn, Cd);
    else
        tail_restrict_n<2>(T, F, Cn, Cd);
    __syncthreads();
}

template <int ITER> __device__ void tail_prolong_n(const TailLevel& T, const TailArrays& F, const TailLevel& Cn,
    const TailArrays& Cd)
{
    const int w = T.w, h = T.h;
    const int cells = w * h;
    if (!wave_has_items(cells))
        return;
    const float inv_w = 1.0f / (float)w;
    float add[ITER], cur[ITER];
    int idx[ITER];
    bool on[ITER];
#pragma unroll
    for (int it = 0; it < ITER; ++it) {
        const int e = min((int)threadIdx.x + it * kTailThreads, cells - 1);
        on[it] = (int)threadIdx.x + it * kTailThreads < cells;
        const int y = tail_div(e, inv_w), x = e - y * w;
        const float wy = F.pyw[y];
        const float* top = Cd.sol + (F.pyf[y] + 1) * Cn.s + 1;
        const float* bottom = Cd.sol + (F.pys[y] + 1) * Cn.s + 1;
        const int a0 = F.pxf[x], a1 = F.pxs[x];
        const float wx = F.pxw[x];
        const float a = (1.f - wx) * top[a0] + wx * top[a1];
        const float b = (1.f - wx) * bottom[a0] + wx * bottom[a1];
        add[it] = (1.f - wy) * a + wy * b;
        idx[it] = (y + 1) * T.s + x + 1;
        cur[it] = F.sol[idx[it]];
    }
#pragma unroll
    for (int it = 0; it < ITER; ++it)
        if (on[it])
            F.sol[idx[it]] = cur[it] + add[it];
}

__device__ void tail_prolong(const TailLevel& T, const TailArrays& F, const TailLevel& Cn, const TailArrays& Cd)
{
    const int cells = T.w * T.h;
    if (cells <= kTailThreads)
        tail_prolong_n<1>(T, F, Cn, Cd);
    else if (cells <= 2 * kTailThreads)
        tail_prolong_n<2>(T, F, Cn, Cd);
    else
        tail_prolong_n<kTailIter>(T, F, Cn, Cd);
    __syncthreads();
}

// solveCoarsest (PressureMultigrid.cpp:347-384): remove the rhs mean, solve the pinned system
// (here with its precomputed inverse), shift to a mean-zero gauge.  Wavefront 0, one lane per row;
// the caller orders the result for the other threads.
__device__ void tail_coarse_solve(const TailLevel& T, const TailArrays& A, const float* G, int n)
{
    const int lane = threadIdx.x;
    const int w = T.w, s = T.s;
    const int cells = n + 1;
    float b = 0.f;
    if (lane < cells)
        b = A.rhs[(lane / w + 1) * s + lane % w + 1];
    const double mean = wave_sum((double)b) / cells;
    // y_j = rhs(cell j + 1) - mean, held by lane j + 1
    const float y = (float)((double)b - mean);
    // c = G y in the order j = 0 .. n-1, eight terms at a time: their G loads and y reads (v_readlane of a
    // uniform lane, no LDS) issue before the additions
    float c = 0.f;
    const int grow = min(lane, max(n - 1, 0)) * n;
    for (int j0 = 0; j0 < n; j0 += 8) {
        float g[8], yv[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int j = min(j0 + q, n - 1);
            g[q] = G[grow + j];
            yv[q] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(y), j + 1));
        }
#pragma unroll
        for (int q = 0; q < 8; ++q)
            if (j0 + q < n && lane < n)
                c += g[q] * yv[q];
    }
    const double gauge = wave_sum(lane < n ? (double)c : 0.0) / cells;
    if (lane == 0)
        A.sol[s + 1] = (float)(-gauge);
    if (lane < n) {
        const int cell = lane + 1;
        A.sol[(cell / w + 1) * s + cell % w + 1] = (float)((double)c - gauge);
    }
}

// The tail V-cycle of one instance on a 1024-thread workgroup: stage every tail level, its transfer
// tables and the coarse inverse in `lds`, cycle, and write the first tail level's solution back.
// TV: storage of the first tail level's global rhs and solution (FP32, or bfloat16).
template <class TV = float> __device__ void tail_run(const TailDesc& D, const double* shift, const int* status,
    float* lds)
{
    const int nl = D.levels;
    // Flat index over every level's padded arrays; all loads issue before the LDS stores.  The level of an
    // element, its source pointers and its LDS slot are picked by comparisons against uniform level
    // offsets (no per-lane reads of the descriptor).
    constexpr int TL = 8;
    int start[kMaxTailLevels + 1];
    start[0] = 0;
#pragma unroll
    for (int k = 0; k < kMaxTailLevels; ++k)
        start[k + 1] = start[k] + (k < nl ? D.lev[k].s * (D.lev[k].h + 2) : 0);
    const int total = start[kMaxTailLevels];
    const TailLevel& T0 = D.lev[0];
    const float inv_s0 = 1.0f / (float)T0.s;
    float vr[TL], vd[TL], vb[TL];
    int slot[TL], size[TL];
    bool bin[TL];
#pragma unroll
    for (int it = 0; it < TL; ++it) {
        const int e0 = threadIdx.x + it * kTailThreads;
        const int e = min(e0, total - 1);
        const float* pr = T0.right;
        const float* pd = T0.down;
        int first = 0, lds_at = T0.lds_arrays, n_at = start[1];
#pragma unroll
        for (int k = 1; k < kMaxTailLevels; ++k) {
            if (k < nl && e >= start[k]) {
                pr = D.lev[k].right;
                pd = D.lev[k].down;
                first = start[k];
                lds_at = D.lev[k].lds_arrays;
                n_at = start[k + 1] - start[k];
            }
        }
        const int le = e - first;
        vr[it] = pr[le];
        vd[it] = pd[le];
        const int y = tail_div(e, inv_s0) - 1, x = e - (y + 1) * T0.s - 1;   // meaningful on level 0 only
        const bool interior0 = first == 0 && x >= 0 && x < T0.w && y >= 0 && y < T0.h;
        const float b = ldv(((const TV*)D.top_rhs)[min(e, start[1] - 1)]);
        vb[it] = interior0 ? b : 0.f;
        bin[it] = interior0;
        slot[it] = e0 < total ? lds_at + le : -1;
        size[it] = n_at;
    }
    // transfer tables (host-built, already in LDS layout) and the coarse inverse: flat copies with
    // clamped, unconditional loads, issued with the level loads above
    constexpr int TT = 2, TG = 4;
    unsigned tv[TT];
    float gv[TG];
    const int nn = D.coarse_n * D.coarse_n;
#pragma unroll
    for (int it = 0; it < TT; ++it)
        tv[it] = D.tables[min(threadIdx.x + it * kTailThreads, (unsigned)max(D.table_words - 1, 0))];
#pragma unroll
    for (int it = 0; it < TG; ++it)
        gv[it] = D.ginv[min(threadIdx.x + it * kTailThreads, (unsigned)max(nn - 1, 0))];
    if (status && *status != kRunning)
        return;
    const double sh = shift ? *shift : 0.0;
#pragma unroll
    for (int it = 0; it < TL; ++it) {
        if (slot[it] < 0)
            continue;
        float* p = lds + slot[it];
        const int ps = size[it];
        p[0] = vr[it];                 // right
        p[ps] = vd[it];                // down
        p[2 * ps] = (shift && bin[it]) ? (float)((double)vb[it] - sh) : vb[it];   // rhs
        p[3 * ps] = 0.f;               // sol
        p[4 * ps] = 0.f;               // res
    }
    {
        unsigned* tables = (unsigned*)(lds + D.lds_tables_base);
#pragma unroll
        for (int it = 0; it < TT; ++it) {
            const int e = threadIdx.x + it * kTailThreads;
            if (e < D.table_words)
                tables[e] = tv[it];
        }
        float* Gl = lds + D.lds_ginv;
#pragma unroll
        for (int it = 0; it < TG; ++it) {
            const int e = threadIdx.x + it * kTailThreads;
            if (e < nn)
                Gl[e] = gv[it];
        }
    }
    float* G = lds + D.lds_ginv;
    const int n = D.coarse_n;
    __syncthreads();
    for (int l = 0; l + 1 < nl; ++l) {
        const TailLevel& T = D.lev[l];
        const TailArrays A = tail_arrays(lds, T);
        tail_smooth4(T, A, 0, 1, D.sweeps);
        tail_residual_restrict(T, A, D.lev[l + 1], tail_arrays(lds, D.lev[l + 1]));
    }
    if (threadIdx.x < 64)
        tail_coarse_solve(D.lev[nl - 1], tail_arrays(lds, D.lev[nl - 1]), G, n);
    __syncthreads();
    for (int l = nl - 2; l >= 0; --l) {
        const TailLevel& T = D.lev[l];
        const TailArrays A = tail_arrays(lds, T);
        tail_prolong(T, A, D.lev[l + 1], tail_arrays(lds, D.lev[l + 1]));
        tail_smooth4(T, A, 1, 0, D.sweeps);
    }
    const TailArrays A0 = tail_arrays(lds, T0);
    const float inv_w0 = 1.0f / (float)T0.w;
    for (int e = threadIdx.x; e < T0.w * T0.h; e += kTailThreads) {
        const int y = tail_div(e, inv_w0), x = e - y * T0.w;
        const int i = (y + 1) * T0.s + x + 1;
        ((TV*)D.top_sol)[i] = stv<TV>(A0.sol[i]);
    }
}

template <class TV>
__global__ __launch_bounds__(kTailThreads) void k_mg_tail(const TailDesc D, const double* shift, const int* status)
{
    extern __shared__ float lds[];
    tail_run<TV>(D, shift, status, lds);
}

constexpr int kMultiNS = 4;

// ---------------------------------------------------------------------------------------------
// Batched PCG: one 1024-thread workgroup per instance runs the instance's whole iteration loop (the
// kernels of one captured PCG iteration become phases separated by workgroup barriers), so a batch of
// small grids is one launch instead of a dozen launches per iteration per grid.  Vectors stay in the
// instance's own HBM buffers; the scalar block is kept in LDS and written back at the end.  Every value
// is computed with the expression of the corresponding single-instance kernel; only the FP64
// reductions run in another order (thread-strided sums, then the 16 wavefronts in order).
// ---------------------------------------------------------------------------------------------
constexpr int kBT = kTailThreads;
constexpr int kBU = 4;   // items per thread whose loads are issued together in the workgroup-strided loops

// t / d in the workgroup-strided loops: the float-reciprocal quotient (tail_div) below 2^22 items, else
// an integer division; `fast` is uniform across the workgroup.
__device__ __forceinline__ int bdiv(int t, int d, float inv_d, bool fast) { return fast ? tail_div(t, inv_d) : t / d; }
constexpr int kBW = kBT / 64;

// Workgroup sum / max, valid in every thread.
__device__ __forceinline__ double wg_sum(double v, double* slot)
{
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0)
        slot[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = 0.0;
#pragma unroll
    for (int i = 0; i < kBW; ++i)
        r += slot[i];
    __syncthreads();
    return r;
}

__device__ __forceinline__ double wg_max(double v, double* slot)
{
    v = wave_max(v);
    if ((threadIdx.x & 63) == 0)
        slot[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = 0.0;
#pragma unroll
    for (int i = 0; i < kBW; ++i)
        r = fmax(r, slot[i]);
    __syncthreads();
    return r;
}

// 2 * sweeps red-black half-sweeps of one level in place on `sol` (k_rb_multi's update): colours from
// `first`, alternating; the first half-sweep of a pre-smoothing ignores the (zero) neighbours.
__device__ void wg_sweeps(const LevelView& L, float* sol, const float* rhs, const double* shift, int first,
    bool from_zero, int sweeps)
{
    const int w = L.w, h = L.h, s = L.s;
    const int pairs = (w + 1) >> 1;
    const int items = pairs * h;
    const double sh = shift ? *shift : 0.0;
    const float inv_pairs = 1.0f / (float)pairs;
    const bool fast = items < (1 << 22);
    for (int k = 0; k < 2 * sweeps; ++k) {
        const int color = first ^ (k & 1);
        const bool neighbours = !(from_zero && k == 0);
        // kBU items per thread at a time: every load of the group issues before the first store (the
        // items of one colour never read each other), through clamped valid indices
        for (int t0 = threadIdx.x; t0 < items; t0 += kBU * kBT) {
            int idx[kBU];
            bool on[kBU];
            float a[kBU], b[kBU], u[kBU], d[kBU], rv[kBU], n0[kBU], n1[kBU], n2[kBU], n3[kBU];
#pragma unroll
            for (int q = 0; q < kBU; ++q) {
                const int t = t0 + q * kBT;
                const int tt = min(t, items - 1);
                const int y = bdiv(tt, pairs, inv_pairs, fast);
                const int x = 2 * (tt - y * pairs) + ((color + y) & 1);
                on[q] = t < items && x < w;
                const int i = (y + 1) * s + min(x, w - 1) + 1;
                idx[q] = i;
                a[q] = L.right[i - 1];
                b[q] = L.right[i];
                u[q] = L.down[i - s];
                d[q] = L.down[i];
                rv[q] = rhs[i];
                n0[q] = sol[i - 1];
                n1[q] = sol[i + 1];
                n2[q] = sol[i - s];
                n3[q] = sol[i + s];
            }
#pragma unroll
            for (int q = 0; q < kBU; ++q) {
                float v = shift ? (float)((double)rv[q] - sh) : rv[q];
                if (neighbours)
                    v += (a[q] * n0[q] + b[q] * n1[q]) + (u[q] * n2[q] + d[q] * n3[q]);
                if (on[q])
                    sol[idx[q]] = v * inv_or_zero((a[q] + b[q]) + (u[q] + d[q]));
            }
        }
        __syncthreads();
    }
}

// Residual of level F's sol[1] into F.res, then its transpose-of-bilinear restriction into C.rhs
// (k_res_restrict's arithmetic; missing taps have weight 0 and index -1, the zero halo).
__device__ void wg_res_restrict(const LevelView& F, const float* rhs, const double* shift, const LevelView& C,
    const AxisRestrict& rx, const AxisRestrict& ry)
{
    const int w = F.w, h = F.h, s = F.s;
    const float* sol = F.sol[1];
    const double sh = shift ? *shift : 0.0;
    const int cells = w * h;
    const float inv_w = 1.0f / (float)w;
    const bool fast = cells < (1 << 22);
    for (int e0 = threadIdx.x; e0 < cells; e0 += kBU * kBT) {
        int idx[kBU];
        float val[kBU];
#pragma unroll
        for (int q = 0; q < kBU; ++q) {
            const int e = min(e0 + q * kBT, cells - 1);
            const int y = bdiv(e, w, inv_w, fast), x = e - y * w;
            const int i = (y + 1) * s + x + 1;
            idx[q] = i;
            const float c = sol[i];
            const float product = (F.right[i - 1] * (c - sol[i - 1]) + F.right[i] * (c - sol[i + 1])) +
                                  (F.down[i - s] * (c - sol[i - s]) + F.down[i] * (c - sol[i + s]));
            const float b = shift ? (float)((double)rhs[i] - sh) : rhs[i];
            const bool active = (F.right[i - 1] + F.right[i]) + (F.down[i - s] + F.down[i]) > 0.f;
            val[q] = active ? b - product : 0.f;   // cells without conductance are not restricted
        }
#pragma unroll
        for (int q = 0; q < kBU; ++q)
            if (e0 + q * kBT < cells)
                F.res[idx[q]] = val[q];
    }
    __syncthreads();
    const int ccells = C.w * C.h;
    const float inv_cw = 1.0f / (float)C.w;
    for (int e0 = threadIdx.x; e0 < ccells; e0 += kBU * kBT) {
        int dst[kBU];
        float val[kBU];
#pragma unroll
        for (int q = 0; q < kBU; ++q) {
            const int e = min(e0 + q * kBT, ccells - 1);
            const int y = bdiv(e, C.w, inv_cw, fast), x = e - y * C.w;
            float value = 0.f;
#pragma unroll
            for (int iy = 0; iy < 4; ++iy) {
                const float* row = F.res + (ry.index[y * 4 + iy] + 1) * s + 1;
                float horizontal = 0.f;
#pragma unroll
                for (int ix = 0; ix < 4; ++ix)
                    horizontal += rx.weight[x * 4 + ix] * row[rx.index[x * 4 + ix]];
                value += ry.weight[y * 4 + iy] * horizontal;
            }
            val[q] = value;
            dst[q] = (y + 1) * C.s + x + 1;
        }
#pragma unroll
        for (int q = 0; q < kBU; ++q)
            if (e0 + q * kBT < ccells)
                C.rhs[dst[q]] = val[q];
    }
    __syncthreads();
}

// sol[0] = sol[1] + bilinear prolongation of the coarse correction (k_rb_multi MODE 1's load).
__device__ void wg_prolong(const LevelView& L, const CoarseLink& link)
{
    const int w = L.w, h = L.h, s = L.s;
    const int cells = w * h;
    const float inv_w = 1.0f / (float)w;
    const bool fast = cells < (1 << 22);
    for (int e0 = threadIdx.x; e0 < cells; e0 += kBU * kBT) {
        int idx[kBU];
        float val[kBU];
#pragma unroll
        for (int q = 0; q < kBU; ++q) {
            const int e = min(e0 + q * kBT, cells - 1);
            const int y = bdiv(e, w, inv_w, fast), x = e - y * w;
            const int i = (y + 1) * s + x + 1;
            idx[q] = i;
            const float wx = link.pxw[x], wy = link.pyw[y];
            const float* top = link.csol + (link.pyf[y] + 1) * link.cs + 1;
            const float* bot = link.csol + (link.pys[y] + 1) * link.cs + 1;
            const int a0 = link.pxf[x], a1 = link.pxs[x];
            const float a = (1.f - wx) * top[a0] + wx * top[a1];
            const float b = (1.f - wx) * bot[a0] + wx * bot[a1];
            float v = L.sol[1][i];
            v += (1.f - wy) * a + wy * b;
            val[q] = v;
        }
#pragma unroll
        for (int q = 0; q < kBU; ++q)
            if (e0 + q * kBT < cells)
                L.sol[0][idx[q]] = val[q];
    }
    __syncthreads();
}

__global__ __launch_bounds__(kBT) void k_batch_pcg(const BatchInstance* batch)
{
    extern __shared__ float lds[];
    __shared__ PcgScalars S;
    __shared__ double slot[3][kBW];
    const BatchInstance& I = batch[blockIdx.x];
    const GridDev& g = I.g;
    const int W = g.W, H = g.H, s = g.s, N = W * H;
    const double n = (double)N;
    const float inv_W = 1.0f / (float)W;
    const bool fast = N < (1 << 22);
    if (threadIdx.x == 0) {
        S = *g.S;
        S.rz = S.rz_slot[S.iterations & 1];   // r.z of the first direction (written by its kernel)
        S.restart = 0;                         // consumed by that direction
    }
    __syncthreads();
    const float* z = I.multigrid ? (I.levels > 0 ? I.lev[0].L.sol[0] : (const float*)I.tail->top_sol) : g.zdiag;
    while (S.status == kRunning) {
        // q = A d and d.q (k_pcg_ad), alpha (kFinAd)
        double dot = 0.0;
        for (int e0 = threadIdx.x; e0 < N; e0 += kBU * kBT) {
            int idx[kBU];
            float cv[kBU], qv[kBU];
#pragma unroll
            for (int u = 0; u < kBU; ++u) {
                const int e = min(e0 + u * kBT, N - 1);
                const int y = bdiv(e, W, inv_W, fast), i = (y + 1) * s + (e - y * W) + 1;
                idx[u] = i;
                const float c = g.d[i];
                const float dl = g.d[i - 1], dr = g.d[i + 1], du = g.d[i - s], dd = g.d[i + s];
                const float wl = g.right32[i - 1], wr = g.right32[i], wu = g.down32[i - s], wd = g.down32[i];
                cv[u] = c;
                qv[u] = (wl * (c - dl) + wr * (c - dr)) + (wu * (c - du) + wd * (c - dd));
            }
#pragma unroll
            for (int u = 0; u < kBU; ++u) {
                if (e0 + u * kBT < N) {
                    g.q[idx[u]] = qv[u];
                    dot += (double)cv[u] * (double)qv[u];
                }
            }
        }
        dot = wg_sum(dot, slot[0]);
        if (threadIdx.x == 0) {
            S.dq = dot;
            if (dot <= 0.0 || !isfinite(dot)) {
                S.status = kStopped;
            } else {
                const double alpha = S.rz / dot;
                if (!isfinite(alpha))
                    S.status = kStopped;
                S.alpha = alpha;
            }
        }
        __syncthreads();
        if (S.status != kRunning)
            break;
        // x += alpha d, r -= alpha q (k_pcg_update); recurrence checks (kFinUpdate)
        const float alpha = (float)S.alpha;
        double sum = 0.0, m = 0.0;
        for (int e0 = threadIdx.x; e0 < N; e0 += kBU * kBT) {
            int idx[kBU];
            float xv[kBU], rv[kBU];
#pragma unroll
            for (int u = 0; u < kBU; ++u) {
                const int e = min(e0 + u * kBT, N - 1);
                const int y = bdiv(e, W, inv_W, fast), i = (y + 1) * s + (e - y * W) + 1;
                idx[u] = i;
                xv[u] = g.x[i] + alpha * g.d[i];
                rv[u] = g.r[i] - alpha * g.q[i];
            }
#pragma unroll
            for (int u = 0; u < kBU; ++u) {
                if (e0 + u * kBT < N) {
                    g.x[idx[u]] = xv[u];
                    g.r[idx[u]] = rv[u];
                    sum += (double)rv[u];
                    m = fmax(m, (double)fabsf(rv[u]));
                }
            }
        }
        sum = wg_sum(sum, slot[0]);
        m = wg_max(m, slot[1]);
        if (threadIdx.x == 0) {
            if (!isfinite(sum))
                m = INFINITY;
            S.iterations += 1;
            S.rmax = m;
            S.r_sum = sum;
            S.r_mean = sum / n;
            if (m <= S.tol) {
                S.verify = 1;
            } else if (m <= kRefineDrop * S.window_rmax || S.iterations % kReplaceEvery == 0) {
                S.verify = 1;
                S.replace = 1;
            }
        }
        __syncthreads();
        // flush and FP64 true residual (k_flush, k_true_residual mode 1, kFinTrueRes1)
        if (S.verify) {
            double tsum = 0.0, tmax = 0.0;
            for (int e = threadIdx.x; e < N; e += kBT) {
                const int y = bdiv(e, W, inv_W, fast), i = (y + 1) * s + (e - y * W) + 1;
                g.p[i] = g.p[i] + (double)g.x[i];
                g.x[i] = 0.0f;
            }
            __syncthreads();
            for (int e = threadIdx.x; e < N; e += kBT) {
                const int y = bdiv(e, W, inv_W, fast), i = (y + 1) * s + (e - y * W) + 1;
                const double c = g.p[i];
                const double product = (g.right64[i - 1] * (c - g.p[i - 1]) + g.right64[i] * (c - g.p[i + 1])) +
                                       (g.down64[i - s] * (c - g.p[i - s]) + g.down64[i] * (c - g.p[i + s]));
                const double r = g.b[i] - product;
                g.r[i] = (float)r;
                tsum += r;
                tmax = fmax(tmax, fabs(r));
            }
            tsum = wg_sum(tsum, slot[0]);
            tmax = wg_max(tmax, slot[1]);
            if (threadIdx.x == 0) {
                if (!isfinite(tsum))
                    tmax = INFINITY;
                S.true_rmax = tmax;
                S.r_sum = tsum;
                if (tmax <= S.tol) {
                    S.status = kConverged;
                } else {
                    S.r_mean = tsum / n;
                    if (!S.replace || tmax > kDriftRatio * S.rmax) {
                        S.restart = 1;
                        S.restarts += 1;
                    }
                    S.rmax = tmax;
                    S.window_rmax = tmax;
                }
                S.verify = 0;
                S.replace = 0;
            }
            __syncthreads();
            if (S.status != kRunning)
                break;
        }
        // preconditioner: V(4,4) cycle of r - mean (vcycle) or the diagonal (k_precond_diag)
        if (I.multigrid) {
            for (int l = 0; l < I.levels; ++l) {
                const BatchLevel& B = I.lev[l];
                const double* shift = l == 0 ? &S.r_mean : nullptr;
                wg_sweeps(B.L, B.L.sol[1], B.L.rhs, shift, 0, true, B.sweeps);
                wg_res_restrict(B.L, B.L.rhs, shift, B.C, B.rx, B.ry);
            }
            tail_run<float>(*I.tail, I.levels == 0 ? &S.r_mean : nullptr, nullptr, lds);
            __syncthreads();
            for (int l = I.levels - 1; l >= 0; --l) {
                const BatchLevel& B = I.lev[l];
                wg_prolong(B.L, B.link);
                wg_sweeps(B.L, B.L.sol[0], B.L.rhs, l == 0 ? &S.r_mean : nullptr, 1, false, B.sweeps);
            }
        } else {
            const double mean = S.r_mean;
            for (int e = threadIdx.x; e < N; e += kBT) {
                const int y = bdiv(e, W, inv_W, fast), i = (y + 1) * s + (e - y * W) + 1;
                g.zdiag[i] = (float)(((double)g.r[i] - mean) * (double)g.invd32[i]);
            }
            __syncthreads();
        }
        // sum z, r.z, sum r (k_pcg_rz), beta (kFinRz)
        double sz = 0.0, srz = 0.0, sr = 0.0;
        for (int e0 = threadIdx.x; e0 < N; e0 += kBU * kBT) {
            float zv[kBU], rv[kBU];
#pragma unroll
            for (int u = 0; u < kBU; ++u) {
                const int e = min(e0 + u * kBT, N - 1);
                const int y = bdiv(e, W, inv_W, fast), i = (y + 1) * s + (e - y * W) + 1;
                zv[u] = z[i];
                rv[u] = g.r[i];
            }
#pragma unroll
            for (int u = 0; u < kBU; ++u) {
                if (e0 + u * kBT < N) {
                    const double zz = (double)zv[u], rr = (double)rv[u];
                    sz += zz;
                    srz += rr * zz;
                    sr += rr;
                }
            }
        }
        sz = wg_sum(sz, slot[0]);
        srz = wg_sum(srz, slot[1]);
        sr = wg_sum(sr, slot[2]);
        if (threadIdx.x == 0) {
            const double mean = sz / n;
            const double next = srz - mean * sr;
            S.z_mean = mean;
            const double beta = S.restart ? 0.0 : next / S.rz;
            S.beta = beta;
            S.restart = 0;
            if (!isfinite(beta)) {
                S.status = kStopped;
            } else {
                S.rz = next;
                if (!(S.iterations < S.max_iterations && next > 0.0 && isfinite(next)))
                    S.status = kStopped;
            }
        }
        __syncthreads();
        if (S.status != kRunning)
            break;
        // d = (z - mean z) + beta d (k_pcg_direction)
        const float zmean = (float)S.z_mean;
        const double beta64 = S.beta;
        const float beta = (float)beta64;
        for (int e0 = threadIdx.x; e0 < N; e0 += kBU * kBT) {
            int idx[kBU];
            float dv[kBU];
#pragma unroll
            for (int u = 0; u < kBU; ++u) {
                const int e = min(e0 + u * kBT, N - 1);
                const int y = bdiv(e, W, inv_W, fast), i = (y + 1) * s + (e - y * W) + 1;
                idx[u] = i;
                const float zz = z[i] - zmean;
                dv[u] = beta64 == 0.0 ? zz : zz + beta * g.d[i];
            }
#pragma unroll
            for (int u = 0; u < kBU; ++u)
                if (e0 + u * kBT < N)
                    g.d[idx[u]] = dv[u];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        *g.S = S;
}

}  // namespace

size_t tail_layout(TailDesc& t)
{
    int words = 0;
    for (int l = 0; l < t.levels; ++l) {
        TailLevel& T = t.lev[l];
        T.lds_arrays = words;
        words += 5 * T.s * (T.h + 2);
    }
    t.lds_ginv = words;
    words += t.coarse_n * t.coarse_n;
    t.lds_tables_base = words;
    int tw = 0;
    for (int l = 0; l < t.levels; ++l) {
        TailLevel& T = t.lev[l];
        T.lds_tables = words + tw;
        tw += 3 * T.w + 3 * T.h + 9 * T.cw + 9 * T.ch;
    }
    t.table_words = tw;
    words += tw;
    t.lds_words = words;
    return (size_t)words * 4;
}

bool tail_allow_lds(size_t bytes)
{
    return hipFuncSetAttribute((const void*)k_mg_tail<float>, hipFuncAttributeMaxDynamicSharedMemorySize,
               (int)bytes) == hipSuccess &&
           hipFuncSetAttribute((const void*)k_mg_tail<bf16>, hipFuncAttributeMaxDynamicSharedMemorySize,
               (int)bytes) == hipSuccess;
}

// Tile variants (interior cells per workgroup).  The fused sweeps compute every interior cell exactly as
// global sweeps would, whatever the tile, so the choice changes time only; small levels use small
// tiles to spread over more compute units.  Variants 6-8 size the workgroup to the tile's red-black
// slots (32 x 32: 1104 slots on 384 threads; 64 x 32 and 32 x 64: 1840 and 1872 on 640), three
// rounds each with at most 4% of the lanes idle, against 28% for 32 x 32 on 512 threads; 9 runs the
// 64 x 32 and 32 x 64 slots in two rounds of 960 threads.
void launch_rb_multi(hipStream_t st, const LevelView& L, int cur, const float* rhs, const double* shift,
    int first_color, int from_zero, const CoarseLink& link, const int* status, int fine, int variant, int nu)
{
    const int mode = from_zero ? 0 : (link.csol ? 1 : 2);
    nu = std::min(std::max(nu, 1), kMultiNS);
#define FS_RB(TX, TY, F, M, NT)                                                                               \
    hipLaunchKernelGGL((k_rb_multi<TX, TY, kMultiNS, F, M, NT>), dim3(((L.w + TX - 1) / TX) * ((L.h + TY - 1) / TY)), \
        dim3(NT), 0, st, L, cur, rhs, shift, first_color, nu, link, status)
#define FS_RB_MODES(TX, TY, F, NT)      \
    do {                                \
        if (mode == 0)                  \
            FS_RB(TX, TY, F, 0, NT);    \
        else if (mode == 1)             \
            FS_RB(TX, TY, F, 1, NT);    \
        else                            \
            FS_RB(TX, TY, F, 2, NT);    \
    } while (0)
    switch (variant) {
    case 1: FS_RB_MODES(32, 32, 0, 256); break;
    case 2: FS_RB_MODES(32, 16, 0, 256); break;
    case 3: FS_RB_MODES(16, 16, 0, 256); break;
    case 4:
        if (fine)
            FS_RB_MODES(64, 32, 1, 512);
        else
            FS_RB_MODES(64, 32, 0, 512);
        break;
    case 5: FS_RB_MODES(32, 32, 0, 512); break;
    case 6: FS_RB_MODES(32, 32, 0, 384); break;
    case 7: FS_RB_MODES(64, 32, 0, 640); break;
    case 8: FS_RB_MODES(32, 64, 0, 640); break;
    case 9: FS_RB_MODES(64, 32, 0, 960); break;
    case 10: FS_RB_MODES(32, 64, 0, 960); break;
    default:
        if (fine)
            FS_RB_MODES(64, 32, 1, 256);
        else
            FS_RB_MODES(64, 32, 0, 256);
        break;
    }
#undef FS_RB_MODES
#undef FS_RB
}

void launch_jacobi_multi(hipStream_t st, const LevelView& L, int cur, const float* rhs, const double* shift,
    int from_zero, const SmoothWeights& weights, int nu, const CoarseLink& link, const int* status, int variant)
{
    nu = std::min(std::max(nu, 1), kJacobiMaxSweeps);
    // the halo the sweeps need: 5 cells up to 5 sweeps (the common case), 8 beyond
#define FS_JM(TX, TY, M, NT, HALO)                                                                            \
    hipLaunchKernelGGL((k_jacobi_multi<TX, TY, M, NT, HALO>),                                                  \
        dim3(((L.w + TX - 1) / TX) * ((L.h + TY - 1) / TY)), dim3(NT), 0, st, L, cur, rhs, shift, weights, nu,  \
        link, status)
#define FS_JM_MODES(TX, TY, NT)                  \
    do {                                         \
        if (nu <= kJacobiShortHalo) {            \
            if (from_zero)                       \
                FS_JM(TX, TY, 0, NT, kJacobiShortHalo); \
            else                                 \
                FS_JM(TX, TY, 1, NT, kJacobiShortHalo); \
        } else if (from_zero) {                  \
            FS_JM(TX, TY, 0, NT, kJacobiMaxSweeps); \
        } else {                                 \
            FS_JM(TX, TY, 1, NT, kJacobiMaxSweeps); \
        }                                        \
    } while (0)
    switch (variant) {
    case 1: FS_JM_MODES(32, 32, 256); break;
    case 2: FS_JM_MODES(32, 16, 256); break;
    case 3: FS_JM_MODES(16, 16, 256); break;
    case 4: FS_JM_MODES(32, 32, 512); break;
    case 5: FS_JM_MODES(64, 64, 1024); break;
    case 6: FS_JM_MODES(64, 32, 1024); break;
    case 7: FS_JM_MODES(32, 64, 1024); break;
    default: FS_JM_MODES(64, 32, 512); break;
    }
#undef FS_JM_MODES
#undef FS_JM
}

void launch_jacobi_res_restrict(hipStream_t st, const LevelView& L, int cur, const float* rhs, const double* shift,
    const SmoothWeights& weights, int nu, const LevelView& C, AxisRestrict rx, AxisRestrict ry, const int* status,
    int variant)
{
    nu = std::min(std::max(nu, 1), kJacobiMaxSweeps);
#define FS_JRR_H(CX, CY, NT, HALO)                                                                            \
    hipLaunchKernelGGL((k_jacobi_res_restrict<CX, CY, NT, HALO>),                                              \
        dim3(((C.w + CX - 1) / CX) * ((C.h + CY - 1) / CY)), dim3(NT), 0, st, L, cur, rhs, shift, weights, nu, C, \
        rx, ry, status)
#define FS_JRR(CX, CY, NT)                      \
    do {                                        \
        if (nu <= kJacobiShortHalo)             \
            FS_JRR_H(CX, CY, NT, kJacobiShortHalo); \
        else                                    \
            FS_JRR_H(CX, CY, NT, kJacobiMaxSweeps); \
    } while (0)
    switch (variant) {
    case 1: FS_JRR(16, 16, 512); break;
    case 2: FS_JRR(32, 16, 1024); break;
    default: FS_JRR(16, 8, 512); break;
    }
#undef FS_JRR
#undef FS_JRR_H
}

void launch_res_restrict(hipStream_t st, const LevelView& F, int cur, const float* rhs, const double* shift,
    const LevelView& C, AxisRestrict rx, AxisRestrict ry, const int* status, int variant)
{
#define FS_RR(CX, CY)                                                                                         \
    hipLaunchKernelGGL((k_res_restrict<CX, CY>), dim3(((C.w + CX - 1) / CX) * ((C.h + CY - 1) / CY)), dim3(kNT), 0, \
        st, F, cur, rhs, shift, C, rx, ry, status)
    switch (variant) {
    case 1: FS_RR(16, 16); break;
    case 2: FS_RR(16, 8); break;
    case 3: FS_RR(8, 8); break;
    default: FS_RR(32, 16); break;
    }
#undef FS_RR
}

void launch_coarse_factor(hipStream_t st, int w, int h, const double* right64, const double* down64,
    double* factor, float* ginv, int n, PcgScalars* S, int pin)
{
    if (n <= 0)
        return;
    hipLaunchKernelGGL(k_coarse_factor, dim3(1), dim3(kFactorThreads), 0, st, w, h, right64, down64, factor, ginv, n,
        S, pin);
}

void launch_coarse_chain(hipStream_t st, const ChainDesc& chain, float* ginv, int n, PcgScalars* S, int pin)
{
    hipLaunchKernelGGL(k_coarse_chain, dim3(1), dim3(64 * kSweepWaves), 0, st, chain, ginv, n, S, pin);
}

void launch_coarse_inverse(hipStream_t st, int w, int h, const double* right64, const double* down64, float* ginv,
    int n, PcgScalars* S, int pin)
{
    if (n <= 0)
        return;
    hipLaunchKernelGGL(k_coarse_inverse, dim3(1), dim3(64 * kSweepWaves), 0, st, w, h, right64, down64, ginv, n, S,
        pin);
}

void launch_mg_tail(hipStream_t st, const TailDesc& desc, size_t lds_bytes, const double* shift, const int* status,
    int bf16_top)
{
    if (bf16_top)
        hipLaunchKernelGGL(k_mg_tail<bf16>, dim3(1), dim3(kTailThreads), lds_bytes, st, desc, shift, status);
    else
        hipLaunchKernelGGL(k_mg_tail<float>, dim3(1), dim3(kTailThreads), lds_bytes, st, desc, shift, status);
}

// ---- 16-bit multigrid levels ----
namespace {

template <class C>
__global__ __launch_bounds__(kNT) void k_level_pack(int w, int h, const double* right64, const double* down64,
    C* right, C* down)
{
    const int s = w + 2;
    const int n = s * (h + 2);
    for (int i = blockIdx.x * kNT + threadIdx.x; i < n; i += gridDim.x * kNT) {
        right[i] = stv<C>((float)right64[i]);
        down[i] = stv<C>((float)down64[i]);
    }
}

}  // namespace

template <class C> void launch_level_pack(hipStream_t st, int w, int h, const double* right64, const double* down64,
    C* right, C* down)
{
    const int n = (w + 2) * (h + 2);
    const int blocks = std::min((n + kNT - 1) / kNT, 4096);
    hipLaunchKernelGGL(k_level_pack<C>, dim3(blocks), dim3(kNT), 0, st, w, h, right64, down64, right, down);
}
template void launch_level_pack<half16>(hipStream_t, int, int, const double*, const double*, half16*, half16*);
template void launch_level_pack<bf16>(hipStream_t, int, int, const double*, const double*, bf16*, bf16*);

template <class C, class V, class VC>
void launch_rb_multi_t(hipStream_t st, const LevelViewT<C, V>& L, int cur, const V* rhs, const double* shift,
    int first_color, int from_zero, const CoarseLinkT<VC>& link, const int* status, int variant, int nu)
{
    nu = std::min(std::max(nu, 1), kMultiNS);
#define FS_RBT(TX, TY, M, NT)                                                                                 \
    hipLaunchKernelGGL((k_rb_multi<TX, TY, kMultiNS, 0, M, NT, C, V, VC>),                                   \
        dim3(((L.w + TX - 1) / TX) * ((L.h + TY - 1) / TY)), dim3(NT), 0, st, L, cur, rhs, shift, first_color, nu, \
        link, status)
#define FS_RBT_MODES(TX, TY, NT)      \
    do {                              \
        if (from_zero)                \
            FS_RBT(TX, TY, 0, NT);    \
        else                          \
            FS_RBT(TX, TY, 1, NT);    \
    } while (0)
    switch (variant) {
    case 1: FS_RBT_MODES(32, 32, 256); break;
    case 2: FS_RBT_MODES(32, 16, 256); break;
    case 3: FS_RBT_MODES(16, 16, 256); break;
    case 4: FS_RBT_MODES(64, 32, 512); break;
    case 5: FS_RBT_MODES(32, 32, 512); break;
    case 6: FS_RBT_MODES(32, 32, 384); break;
    case 7: FS_RBT_MODES(64, 32, 640); break;
    case 8: FS_RBT_MODES(32, 64, 640); break;
    case 9: FS_RBT_MODES(64, 32, 960); break;
    case 10: FS_RBT_MODES(32, 64, 960); break;
    default: FS_RBT_MODES(64, 32, 256); break;
    }
#undef FS_RBT_MODES
#undef FS_RBT
}

template <class C, class V, class VC>
void launch_res_restrict_t(hipStream_t st, const LevelViewT<C, V>& F, int cur, const V* rhs, const double* shift,
    const LevelViewT<C, VC>& Cl, AxisRestrict rx, AxisRestrict ry, const int* status, int variant)
{
#define FS_RRT(CX, CY)                                                                                        \
    hipLaunchKernelGGL((k_res_restrict<CX, CY, C, V, VC>), dim3(((Cl.w + CX - 1) / CX) * ((Cl.h + CY - 1) / CY)),  \
        dim3(kNT), 0, st, F, cur, rhs, shift, Cl, rx, ry, status)
    switch (variant) {
    case 1: FS_RRT(16, 16); break;
    case 2: FS_RRT(16, 8); break;
    case 3: FS_RRT(8, 8); break;
    default: FS_RRT(32, 16); break;
    }
#undef FS_RRT
}

#define FS_MG16(C, V, VC)                                                                                      \
    template void launch_rb_multi_t<C, V, VC>(hipStream_t, const LevelViewT<C, V>&, int, const V*, const double*, \
        int, int, const CoarseLinkT<VC>&, const int*, int, int);                                               \
    template void launch_res_restrict_t<C, V, VC>(hipStream_t, const LevelViewT<C, V>&, int, const V*,          \
        const double*, const LevelViewT<C, VC>&, AxisRestrict, AxisRestrict, const int*, int);
FS_MG16(half16, float, float)
FS_MG16(bf16, float, float)
FS_MG16(bf16, float, bf16)
FS_MG16(bf16, bf16, bf16)
#undef FS_MG16

bool batch_allow_lds(size_t bytes)
{
    return hipFuncSetAttribute((const void*)k_batch_pcg, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) ==
           hipSuccess;
}

void launch_batch_pcg(hipStream_t st, const BatchInstance* batch, int count, size_t lds_bytes)
{
    if (count > 0)
        hipLaunchKernelGGL(k_batch_pcg, dim3(count), dim3(kBT), lds_bytes, st, batch);
}

}  // namespace fsim
