// fs_transport.hip -- the reference's PLIC/VOF transport on gfx950, FP64.
//
// FluidGridCPU::advect (FluidGridCPU.cpp:362-463) splits each CFL substep into an X and a Y pass.
// Each pass is three launches here instead of the reference's six sweeps:
//   faces : MYC interface normal and line offset of the upwind donor (recomputed from its 3 x 3
//           fraction neighbourhood, FluidGridCPU.cpp:473-541), the liquid volume through the face
//           (liquidTransferX/Y, :759-833) and the summed component mass transfer (:1235-1305);
//   cells : every component's mass update, the new fraction and the new density in one read of the
//           cell's two faces (advectComponents :1351-1397, :407-455, rebuildDensity :1399-1423);
//   faces : momentum transfer of both velocity components and velocity recovery (advectMomentumX/Y
//           :560-712, recoverVelocity :714-757).
// Per-component face transfers are recomputed where needed instead of being stored, so a pass moves
// each mass array once.  Fractions, masses and transported velocities are double-buffered, which
// replaces the reference's in-place ordering ("finish all neighbour reads before changing any cell
// masses") with a pure gather.
//
// Every expression is evaluated in the reference's order with -ffp-contract=off, so on the same input
// the results are the reference's bit for bit.
#include "fs_kernels.hpp"

#include <math.h>

namespace fsim {

namespace {

constexpr int kNT = 256;   // 64 x 4 threads
constexpr int kBX = 64;
constexpr int kBY = 4;

__device__ __forceinline__ double dmin(double a, double b) { return b < a ? b : a; }   // std::min
__device__ __forceinline__ double dmax(double a, double b) { return a < b ? b : a; }   // std::max

// liquidArea (FluidGridCPU.cpp:93-133): area of {p : n.p <= offset} inside [x0,x1] x [y0,y1].
__device__ double liquid_area(double nx, double ny, double offset, double x0, double y0, double x1, double y1)
{
    const double width = x1 - x0;
    const double height = y1 - y0;
    const double area = width * height;
    const double min_projection = nx * (nx >= 0.0 ? x0 : x1) + ny * (ny >= 0.0 ? y0 : y1);
    const double cut = offset - min_projection;
    const double span_x = fabs(nx) * width;
    const double span_y = fabs(ny) * height;
    const double small_span = dmin(span_x, span_y);
    const double large_span = dmax(span_x, span_y);
    if (cut <= 0.0)
        return 0.0;
    if (cut >= small_span + large_span)
        return area;
    if (small_span == 0.0)
        return area * cut / large_span;
    if (cut < small_span)
        return area * cut * cut / (2.0 * small_span * large_span);
    if (cut < large_span)
        return area * (cut - 0.5 * small_span) / large_span;
    const double dry_cut = small_span + large_span - cut;
    return area * (1.0 - dry_cut * dry_cut / (2.0 * small_span * large_span));
}

// interfaceOffset (FluidGridCPU.cpp:135-167).
__device__ double interface_offset(double nx, double ny, double fraction)
{
    const double min_projection = dmin(nx, 0.0) + dmin(ny, 0.0);
    const double span_x = fabs(nx);
    const double span_y = fabs(ny);
    const double small_span = dmin(span_x, span_y);
    const double large_span = dmax(span_x, span_y);
    const double triangle_fraction = 0.5 * small_span / large_span;
    double cut;
    if (fraction < triangle_fraction)
        cut = sqrt(2.0 * small_span * large_span * fraction);
    else if (fraction <= 1.0 - triangle_fraction)
        cut = large_span * fraction + 0.5 * small_span;
    else
        cut = small_span + large_span - sqrt(2.0 * small_span * large_span * (1.0 - fraction));
    return min_projection + cut;
}

// interfaceNormal (FluidGridCPU.cpp:473-521): mixed Youngs-centred normal, neighbours clamped to the grid.
__device__ void interface_normal(const double* f, int W, int H, int x, int y, double& nx, double& ny)
{
    const int lx = max(x - 1, 0), rx = min(x + 1, W - 1);
    const double* above = f + (size_t)max(y - 1, 0) * W;
    const double* row = f + (size_t)y * W;
    const double* below = f + (size_t)min(y + 1, H - 1) * W;
    const double tl = above[lx], t = above[x], tr = above[rx];
    const double l = row[lx], r = row[rx];
    const double bl = below[lx], b = below[x], br = below[rx];
    const double left_sum = tl + l + bl;
    const double right_sum = tr + r + br;
    const double top_sum = tl + t + tr;
    const double bottom_sum = bl + b + br;
    double cx = 0.5 * (left_sum - right_sum), cy = 0.5 * (top_sum - bottom_sum);
    const double yx = left_sum + l - right_sum - r, yy = top_sum + t - bottom_sum - b;
    if (cx == 0.0 && cy == 0.0) {
        const bool none = yx == 0.0 && yy == 0.0;
        nx = none ? 1.0 : yx;
        ny = none ? 0.0 : yy;
        return;
    }
    bool youngs;
    if (fabs(cx) <= fabs(cy)) {
        cy = copysign(1.0, cy);
        youngs = fabs(yx) > fabs(cx * yy);
    } else {
        cx = copysign(1.0, cx);
        youngs = fabs(yy) > fabs(cy * yx);
    }
    nx = youngs ? yx : cx;
    ny = youngs ? yy : cy;
}

// componentMassTransfer (FluidGridCPU.cpp:543-551).
__device__ __forceinline__ double component_transfer(double volume_transfer, double donor_fraction, double donor_mass)
{
    return donor_fraction <= 0.0 ? 0.0 : donor_mass * (volume_transfer / donor_fraction);
}

// momentumTransfer (FluidGridCPU.cpp:553-558).
__device__ __forceinline__ double momentum_transfer(double mass_transfer, double before, double after)
{
    return mass_transfer * (mass_transfer >= 0.0 ? before : after);
}

// ------------------------------------------------------------------------------------------
// rebuildMomentum (FluidGridCPU.cpp:1425-1476); z = 0: u faces, z = 1: v faces.
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(kNT) void k_tr_begin(TransportDev t)
{
    const int W = t.W, H = t.H;
    const int x = blockIdx.x * kBX + (threadIdx.x & 63);
    const int y = blockIdx.y * kBY + (threadIdx.x >> 6);
    if (blockIdx.z == 0) {
        if (x > W || y >= H)
            return;
        const size_t face = (size_t)y * (W + 1) + x;
        if (x == 0 || x == W) {
            t.mom_u[face] = 0.0;
            t.nu_out[face] = 0.0;
            return;
        }
        const size_t cell = (size_t)y * W + x;
        const double face_density = 0.5 * (t.density[cell - 1] + t.density[cell]);
        t.mom_u[face] = face_density * t.u[face];
        t.nu_out[face] = t.u[face];
    } else {
        if (x >= W || y > H)
            return;
        const size_t face = (size_t)y * W + x;
        if (y == 0 || y == H) {
            t.mom_v[face] = 0.0;
            t.nv_out[face] = 0.0;
            return;
        }
        const double face_density = 0.5 * (t.density[face - W] + t.density[face]);
        t.mom_v[face] = face_density * t.v[face];
        t.nv_out[face] = t.v[face];
    }
}

// Largest cell outflow (FluidGridCPU.cpp:366-386): one partial per 64 x 16 tile.
__global__ __launch_bounds__(kNT) void k_tr_outflow(TransportDev t)
{
    __shared__ double lds[kNT / 64];
    const int W = t.W, H = t.H;
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int xc = min(x, W - 1);
    double o[4];
    bool in[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int y = blockIdx.y * 16 + (threadIdx.x >> 6) + 4 * k;
        in[k] = x < W && y < H;
        const int yc = min(y, H - 1);
        const double* hz = t.u + (size_t)yc * (W + 1);
        const double* vt = t.v + (size_t)yc * W;
        o[k] = dmax(hz[xc + 1], 0.0) + dmax(-hz[xc], 0.0) + dmax(vt[xc + W], 0.0) + dmax(-vt[xc], 0.0);
    }
    double m = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (in[k])
            m = dmax(m, o[k]);
    const double bm = block_max<kNT>(m, lds);
    if (threadIdx.x == 0)
        t.partials[blockIdx.y * gridDim.x + blockIdx.x] = bm;
}

__global__ __launch_bounds__(kNT) void k_tr_outflow_finish(TransportDev t, int nb)
{
    __shared__ double lds[kNT / 64];
    const double m = finish_max<kNT>(t.partials, nb, 1, 0, lds);
    if (threadIdx.x == 0)
        *t.out = m;
}

// freezeComponentCorrections (FluidGridCPU.cpp:1307-1349).
__global__ __launch_bounds__(kNT) void k_tr_freeze(TransportDev t)
{
    const size_t N = (size_t)t.W * t.H;
    for (size_t i = (size_t)blockIdx.x * kNT + threadIdx.x; i < N; i += (size_t)gridDim.x * kNT) {
        const double fraction = t.f[i];
        const bool selected = fraction > 0.5;
        t.split[i] = (unsigned char)selected;
        const double inverse_fraction = 1.0 / (selected ? fraction : 1.0 - fraction);
        double correction = 0.0;
        for (int k = 0; k < t.nl; ++k) {
            const double weight = selected ? t.lm[k * N + i] * inverse_fraction : 0.0;
            t.lcw[k * N + i] = weight;
            correction += weight * t.icv;
        }
        for (int k = 0; k < t.ng; ++k) {
            const double weight = !selected ? t.gm[k * N + i] * inverse_fraction : 0.0;
            t.gcw[k * N + i] = weight;
            correction += weight * t.icv;
        }
        t.mcw[i] = correction;
    }
}

// ------------------------------------------------------------------------------------------
// Pass, launch 1: liquid volume and total mass through each face of the pass's axis.
// ------------------------------------------------------------------------------------------
template <bool HORIZ>
__global__ __launch_bounds__(kNT) void k_tr_faces(TransportDev t, double scale)
{
    const int W = t.W, H = t.H;
    const size_t N = (size_t)W * H;
    const int x = blockIdx.x * kBX + (threadIdx.x & 63);
    const int y = blockIdx.y * kBY + (threadIdx.x >> 6);
    if (HORIZ ? (x > W || y >= H) : (x >= W || y > H))
        return;
    const size_t face = HORIZ ? (size_t)y * (W + 1) + x : (size_t)y * W + x;
    const bool wall = HORIZ ? (x == 0 || x == W) : (y == 0 || y == H);
    double lt = 0.0, mt = 0.0;
    if (!wall) {
        const double speed = HORIZ ? t.u[face] : t.v[face];
        const int dx = HORIZ ? (speed > 0.0 ? x - 1 : x) : x;
        const int dy = HORIZ ? y : (speed > 0.0 ? y - 1 : y);
        const size_t donor = (size_t)dy * W + dx;
        const double fraction = t.f[donor];
        // liquidTransferX / liquidTransferY (FluidGridCPU.cpp:759-833)
        if (speed != 0.0 && fraction > 0.0) {
            const double volume_transfer = speed * scale;
            const double strip = fabs(volume_transfer);
            if (fraction >= 1.0) {
                lt = volume_transfer;
            } else {
                double nx, ny;
                interface_normal(t.f, W, H, dx, dy, nx, ny);
                const double offset = interface_offset(nx, ny, fraction);
                const double lo = speed > 0.0 ? 1.0 - strip : 0.0;
                const double hi = speed > 0.0 ? 1.0 : strip;
                const double area = HORIZ ? liquid_area(nx, ny, offset, lo, 0.0, hi, 1.0)
                                          : liquid_area(nx, ny, offset, 0.0, lo, 1.0, hi);
                const double floor_ = dmax(0.0, strip - (1.0 - fraction));
                const double ceil_ = dmin(strip, fraction);
                const double bounded = area < floor_ ? floor_ : (ceil_ < area ? ceil_ : area);  // std::clamp
                lt = speed > 0.0 ? bounded : -bounded;
            }
        }
        // rebuildComponentTransfersX/Y (FluidGridCPU.cpp:1235-1305): liquid components, then gas.
        const double total = speed * scale;
        for (int k = 0; k < t.nl; ++k)
            mt += component_transfer(lt, fraction, t.lm[k * N + donor]) * t.icv;
        for (int k = 0; k < t.ng; ++k)
            mt += component_transfer(total - lt, 1.0 - fraction, t.gm[k * N + donor]) * t.icv;
    }
    t.lt[face] = lt;
    t.mt[face] = mt;
}

// ------------------------------------------------------------------------------------------
// Pass, launch 2: component masses, fraction and density of each cell.
// ------------------------------------------------------------------------------------------
template <bool HORIZ>
__global__ __launch_bounds__(kNT) void k_tr_cells(TransportDev t, double scale)
{
    const int W = t.W, H = t.H;
    const size_t N = (size_t)W * H;
    const int x = blockIdx.x * kBX + (threadIdx.x & 63);
    const int y = blockIdx.y * kBY + (threadIdx.x >> 6);
    if (x >= W || y >= H)
        return;
    const size_t i = (size_t)y * W + x;
    const size_t before = HORIZ ? (size_t)y * (W + 1) + x : i;
    const size_t after = before + (HORIZ ? 1 : (size_t)W);
    const size_t stride = HORIZ ? 1 : (size_t)W;
    const double* flow = HORIZ ? t.u : t.v;
    const bool wall_before = HORIZ ? x == 0 : y == 0;
    const bool wall_after = HORIZ ? x + 1 == W : y + 1 == H;
    const double fb = flow[before], fa = flow[after];
    const double lt_b = t.lt[before], lt_a = t.lt[after];
    // upwind donors of the two faces (clamped to the cell itself at walls, where nothing moves)
    const size_t donor_b = (!wall_before && fb > 0.0) ? i - stride : i;
    const size_t donor_a = (!wall_after && !(fa > 0.0)) ? i + stride : i;
    const double frac_b = t.f[donor_b], frac_a = t.f[donor_a];
    const double all_b = fb * scale, all_a = fa * scale;
    const double net_fluid = scale * (fa - fb);
    double total = 0.0;
    for (int k = 0; k < t.nl; ++k) {
        const double* m = t.lm + k * N;
        const double ct_b = wall_before ? 0.0 : component_transfer(lt_b, frac_b, m[donor_b]);
        const double ct_a = wall_after ? 0.0 : component_transfer(lt_a, frac_a, m[donor_a]);
        const double mass = m[i] + (ct_b - ct_a + t.lcw[k * N + i] * net_fluid);
        t.lm_out[k * N + i] = mass;
        total += mass;
    }
    for (int k = 0; k < t.ng; ++k) {
        const double* m = t.gm + k * N;
        const double ct_b = wall_before ? 0.0 : component_transfer(all_b - lt_b, 1.0 - frac_b, m[donor_b]);
        const double ct_a = wall_after ? 0.0 : component_transfer(all_a - lt_a, 1.0 - frac_a, m[donor_a]);
        const double mass = m[i] + (ct_b - ct_a + t.gcw[k * N + i] * net_fluid);
        t.gm_out[k * N + i] = mass;
        total += mass;
    }
    t.density[i] = total * t.icv;
    t.f_out[i] = t.f[i] - (lt_a - lt_b) + (double)t.split[i] * net_fluid;
}

// ------------------------------------------------------------------------------------------
// Pass, launch 3: momentum transfer and velocity recovery; z = 0: u faces, z = 1: v faces.
// ------------------------------------------------------------------------------------------
template <bool HORIZ>
__device__ __forceinline__ double mass_correction(const TransportDev& t, int cx, int cy, double scale)
{
    // weight * scale * (flow out - flow in) of one cell along the pass's axis
    const size_t cell = (size_t)cy * t.W + cx;
    const double w = t.mcw[cell];
    if (HORIZ) {
        const size_t lf = (size_t)cy * (t.W + 1) + cx;
        return w * scale * (t.u[lf + 1] - t.u[lf]);
    }
    return w * scale * (t.v[cell + t.W] - t.v[cell]);
}

template <bool HORIZ>
__global__ __launch_bounds__(kNT) void k_tr_momentum(TransportDev t, double scale)
{
    const int W = t.W, H = t.H;
    const int x = blockIdx.x * kBX + (threadIdx.x & 63);
    const int y = blockIdx.y * kBY + (threadIdx.x >> 6);
    if (blockIdx.z == 0) {
        // u faces
        if (x > W || y >= H)
            return;
        const size_t us = (size_t)W + 1;
        const size_t face = (size_t)y * us + x;
        if (x == 0 || x == W) {
            t.mom_u[face] = 0.0;
            t.nu_out[face] = 0.0;
            return;
        }
        double m_before, m_after;   // momentum entering through the near side, leaving through the far side
        if (HORIZ) {
            const double lm = 0.5 * (t.mt[face - 1] + t.mt[face]);
            const double rm = 0.5 * (t.mt[face] + t.mt[face + 1]);
            m_before = momentum_transfer(lm, t.nu_in[face - 1], t.nu_in[face]);
            m_after = momentum_transfer(rm, t.nu_in[face], t.nu_in[face + 1]);
        } else {
            const size_t left_top = (size_t)y * W + x - 1;
            const size_t right_top = (size_t)y * W + x;
            const double tm = 0.5 * (t.mt[left_top] + t.mt[right_top]);
            const double bm = 0.5 * (t.mt[left_top + W] + t.mt[right_top + W]);
            m_before = y > 0 ? momentum_transfer(tm, t.nu_in[face - us], t.nu_in[face]) : 0.0;
            m_after = y + 1 < H ? momentum_transfer(bm, t.nu_in[face], t.nu_in[face + us]) : 0.0;
        }
        const double correction =
            0.5 * (mass_correction<HORIZ>(t, x - 1, y, scale) + mass_correction<HORIZ>(t, x, y, scale));
        const double momentum = t.mom_u[face] + (m_before - m_after + t.u[face] * correction);
        t.mom_u[face] = momentum;
        const size_t cell = (size_t)y * W + x;
        t.nu_out[face] = momentum / (0.5 * (t.density[cell - 1] + t.density[cell]));
    } else {
        // v faces
        if (x >= W || y > H)
            return;
        const size_t face = (size_t)y * W + x;
        if (y == 0 || y == H) {
            t.mom_v[face] = 0.0;
            t.nv_out[face] = 0.0;
            return;
        }
        double m_before, m_after;
        if (HORIZ) {
            const size_t above_left = (size_t)(y - 1) * (W + 1) + x;
            const size_t below_left = (size_t)y * (W + 1) + x;
            const double lm = 0.5 * (t.mt[above_left] + t.mt[below_left]);
            const double rm = 0.5 * (t.mt[above_left + 1] + t.mt[below_left + 1]);
            m_before = x > 0 ? momentum_transfer(lm, t.nv_in[face - 1], t.nv_in[face]) : 0.0;
            m_after = x + 1 < W ? momentum_transfer(rm, t.nv_in[face], t.nv_in[face + 1]) : 0.0;
        } else {
            const double tm = 0.5 * (t.mt[face - W] + t.mt[face]);
            const double bm = 0.5 * (t.mt[face] + t.mt[face + W]);
            m_before = momentum_transfer(tm, t.nv_in[face - W], t.nv_in[face]);
            m_after = momentum_transfer(bm, t.nv_in[face], t.nv_in[face + W]);
        }
        const double correction =
            0.5 * (mass_correction<HORIZ>(t, x, y - 1, scale) + mass_correction<HORIZ>(t, x, y, scale));
        const double momentum = t.mom_v[face] + (m_before - m_after + t.v[face] * correction);
        t.mom_v[face] = momentum;
        t.nv_out[face] = momentum / (0.5 * (t.density[face - W] + t.density[face]));
    }
}

inline dim3 face_grid(int W, int H)
{
    // covers both (W+1) x H and W x (H+1)
    return dim3((W + 1 + kBX - 1) / kBX, (H + 1 + kBY - 1) / kBY, 1);
}

inline unsigned linear_blocks(size_t n)
{
    size_t b = (n + kNT - 1) / kNT;
    if (b > 8192)
        b = 8192;
    return (unsigned)(b ? b : 1);
}

}  // namespace

void launch_transport_begin(hipStream_t st, const TransportDev& t)
{
    dim3 g = face_grid(t.W, t.H);
    g.z = 2;
    hipLaunchKernelGGL(k_tr_begin, g, dim3(kNT), 0, st, t);
    const dim3 tiles((t.W + 63) / 64, (t.H + 15) / 16);
    hipLaunchKernelGGL(k_tr_outflow, tiles, dim3(kNT), 0, st, t);
    hipLaunchKernelGGL(k_tr_outflow_finish, dim3(1), dim3(kNT), 0, st, t, (int)(tiles.x * tiles.y));
}

void launch_transport_freeze(hipStream_t st, const TransportDev& t)
{
    hipLaunchKernelGGL(k_tr_freeze, dim3(linear_blocks((size_t)t.W * t.H)), dim3(kNT), 0, st, t);
}

void launch_transport_pass(hipStream_t st, const TransportDev& t, int horizontal, double scale)
{
    const dim3 fg = face_grid(t.W, t.H);
    const dim3 cg((t.W + kBX - 1) / kBX, (t.H + kBY - 1) / kBY);
    dim3 mg = fg;
    mg.z = 2;
    if (horizontal) {
        hipLaunchKernelGGL(k_tr_faces<true>, fg, dim3(kNT), 0, st, t, scale);
        hipLaunchKernelGGL(k_tr_cells<true>, cg, dim3(kNT), 0, st, t, scale);
        hipLaunchKernelGGL(k_tr_momentum<true>, mg, dim3(kNT), 0, st, t, scale);
    } else {
        hipLaunchKernelGGL(k_tr_faces<false>, fg, dim3(kNT), 0, st, t, scale);
        hipLaunchKernelGGL(k_tr_cells<false>, cg, dim3(kNT), 0, st, t, scale);
        hipLaunchKernelGGL(k_tr_momentum<false>, mg, dim3(kNT), 0, st, t, scale);
    }
}

}  // namespace fsim
