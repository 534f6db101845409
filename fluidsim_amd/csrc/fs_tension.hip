// fs_tension.hip -- capillary force of the reference's uniform CPU engine on gfx950, FP64.
//
// FluidGridCPU::applySurfaceTension (FluidGridCPU.cpp:862-926) with the geometry of
// FluidSurfaceTensionCPU (FluidSurfaceTensionCPU.h:13-390): three ordered cell passes (height
// functions; height-function curvature with a quadratic-fit fallback; completion of missing
// curvatures from valid neighbours or a segment fit), then the face force
// u += dt w sigma kappa_face (f_R - f_L) / h with the pressure face weights w.  Walls reflect the fraction
// field evenly and negate mirrored heights.  One thread per cell or face; every pass is a gather over
// the previous pass's arrays.  Expressions follow the reference's order with -ffp-contract=off; the
// only library call besides sqrt is pow(., 1.5), whose last bit may differ from the host library's.
#include "fs_kernels.hpp"

#include <float.h>
#include <math.h>

namespace fsim {

namespace {

constexpr int kCX = 64;
constexpr int kCY = 4;
constexpr int kCT = 256;

struct Geo {
    int W, H;
    double h;
    const double* f;
};

__device__ __forceinline__ double tmin(double a, double b) { return b < a ? b : a; }
__device__ __forceinline__ double tmax(double a, double b) { return a < b ? b : a; }
__device__ __forceinline__ double tclamp(double v, double lo, double hi) { return v < lo ? lo : (hi < v ? hi : v); }

__device__ __forceinline__ int mirror(int i, int extent)
{
    const int period = 2 * extent;
    const int q = ((i % period) + period) % period;
    return q >= extent ? period - 1 - q : q;
}

// FluidSurfaceTensionCPU::fraction (:49-63)
__device__ __forceinline__ double frac(const Geo& G, int x, int y)
{
    const double f = G.f[mirror(y, G.H) * G.W + mirror(x, G.W)];
    const double roundoff = 32.0 * DBL_EPSILON;
    return f <= roundoff ? 0.0 : (f >= 1.0 - roundoff ? 1.0 : f);
}

// FluidSurfaceTensionCPU::interfacial (:65-78)
__device__ bool interfacial(const Geo& G, int x, int y)
{
    const double f = frac(G, x, y);
    if (f > 0.0 && f < 1.0)
        return true;
    const double nb[4] = {frac(G, x - 1, y), frac(G, x + 1, y), frac(G, x, y - 1), frac(G, x, y + 1)};
    bool any = false;
#pragma unroll
    for (int i = 0; i < 4; ++i)
        any |= (f >= 1.0 && nb[i] <= 0.0) || (f <= 0.0 && nb[i] >= 1.0);
    return any;
}

// interface_geometry::normal (FluidInterfaceGeometry.h:14-49) then FluidSurfaceTensionCPU::outward (:80-92)
__device__ double2 outward(const Geo& G, int x, int y)
{
    const double tl = frac(G, x - 1, y - 1), t = frac(G, x, y - 1), tr = frac(G, x + 1, y - 1);
    const double l = frac(G, x - 1, y), r = frac(G, x + 1, y);
    const double bl = frac(G, x - 1, y + 1), b = frac(G, x, y + 1), br = frac(G, x + 1, y + 1);
    const double left = tl + l + bl;
    const double right = tr + r + br;
    const double top = tl + t + tr;
    const double bottom = bl + b + br;
    double cx = 0.5 * (left - right), cy = 0.5 * (top - bottom);
    const double yx = left + l - right - r, yy = top + t - bottom - b;
    double nx, ny;
    if (cx == 0.0 && cy == 0.0) {
        const bool none = yx == 0.0 && yy == 0.0;
        nx = none ? 1.0 : yx;
        ny = none ? 0.0 : yy;
    } else {
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
    const double magnitude = tmax(fabs(nx), fabs(ny));
    nx /= magnitude;
    ny /= magnitude;
    const double length = sqrt(nx * nx + ny * ny);
    return make_double2(nx / length, ny / length);
}

// interface_geometry::offset (FluidInterfaceGeometry.h:51-61)
__device__ double line_offset(double2 n, double fraction)
{
    const double a = tmin(fabs(n.x), fabs(n.y));
    const double b = tmax(fabs(n.x), fabs(n.y));
    const double triangle = 0.5 * a / b;
    const double cut = fraction < triangle ? sqrt(2.0 * a * b * fraction)
                       : fraction <= 1.0 - triangle ? b * fraction + 0.5 * a
                                                    : a + b - sqrt(2.0 * a * b * (1.0 - fraction));
    return tmin(n.x, 0.0) + tmin(n.y, 0.0) + cut;
}

// FluidSurfaceTensionCPU::column (:94-132)
__device__ void column(const Geo& G, int x, int y, int ax, int ay, double& height, double& orientation)
{
    int start = -4;
    double base = frac(G, x - 4 * ax, y - 4 * ay);
    bool have = base <= 0.0 || base >= 1.0;
    base = base >= 1.0 ? 1.0 : 0.0;
    double sum = 0.0;
    double nearest = 10.0;
    height = 0.0;
    orientation = 0.0;
    for (int k = -3; k <= 4; ++k) {
        const double f = frac(G, x + k * ax, y + k * ay);
        if (f <= 0.0 || f >= 1.0) {
            const double pure = f >= 1.0 ? 1.0 : 0.0;
            if (have && pure != base) {
                const double h = (double)start + 0.5 + (base == 1.0 ? sum : (double)(k - start - 1) - sum);
                if (fabs(h) < nearest) {
                    nearest = fabs(h);
                    height = h;
                    orientation = base == 1.0 ? 1.0 : -1.0;
                }
            }
            start = k;
            base = pure;
            have = true;
            sum = 0.0;
        } else {
            sum += f;
        }
    }
}

// heights: (x, y, orientation_x, orientation_y) per cell as two double2
struct Heights {
    double2* hv;   // (x, y)
    double2* ov;   // (orientation_x, orientation_y)
};

__device__ __forceinline__ void height_at(const Heights& Hs, const Geo& G, int x, int y, double2& h, double2& o)
{
    const int i = mirror(y, G.H) * G.W + mirror(x, G.W);
    h = Hs.hv[i];
    o = Hs.ov[i];
    if (x < 0 || x >= G.W) {
        h.x = -h.x;
        o.x = -o.x;
    }
    if (y < 0 || y >= G.H) {
        h.y = -h.y;
        o.y = -o.y;
    }
}

struct Fit {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0, s4 = 0.0, y0 = 0.0, y1 = 0.0, y2 = 0.0;

    __device__ void add(double px, double py, double2 n, double weight)
    {
        const double x = -px * n.y + py * n.x;
        const double y = px * n.x + py * n.y;
        const double x2 = x * x;
        s0 += weight;
        s1 += weight * x;
        s2 += weight * x2;
        s3 += weight * x2 * x;
        s4 += weight * x2 * x2;
        y0 += weight * y;
        y1 += weight * x * y;
        y2 += weight * x2 * y;
    }

    // (value, valid)
    __device__ double2 evaluate(double cell_size) const
    {
        if (s0 == 0.0)
            return make_double2(0.0, 0.0);
        const double a = s2 - s1 * s1 / s0;
        const double b = s3 - s1 * s2 / s0;
        const double c = s4 - s2 * s2 / s0;
        const double d = y1 - s1 * y0 / s0;
        const double e = y2 - s2 * y0 / s0;
        const double determinant = a * c - b * b;
        if (determinant <= 1e-5 * tmax(a * c, 1e-12))
            return make_double2(0.0, 0.0);
        const double quadratic = (a * e - b * d) / determinant;
        const double slope = (c * d - b * e) / determinant;
        const double k = -2.0 * quadratic / pow(1.0 + slope * slope, 1.5);
        return make_double2(tclamp(k, -2.0, 2.0) / cell_size, 1.0);
    }
};

// FluidSurfaceTensionCPU::heightCurvature (:177-200)
__device__ double2 height_curvature(const Heights& Hs, const Geo& G, int x, int y, bool horizontal)
{
    const int tx = horizontal ? 0 : 1, ty = horizontal ? 1 : 0;
    double2 ha, hb, hc, oa, ob, oc;
    height_at(Hs, G, x - tx, y - ty, ha, oa);
    height_at(Hs, G, x, y, hb, ob);
    height_at(Hs, G, x + tx, y + ty, hc, oc);
    const double orientation = horizontal ? ob.x : ob.y;
    if (orientation == 0.0 || (horizontal ? oa.x : oa.y) != orientation || (horizontal ? oc.x : oc.y) != orientation)
        return make_double2(0.0, 0.0);
    const double left = horizontal ? ha.x : ha.y;
    const double mid = horizontal ? hb.x : hb.y;
    const double right = horizontal ? hc.x : hc.y;
    const double slope = 0.5 * (right - left);
    const double second = (right - mid) - (mid - left);
    const double k = -orientation * second / pow(1.0 + slope * slope, 1.5);
    return make_double2(tclamp(k, -1.0, 1.0) / G.h, 1.0);
}

// FluidSurfaceTensionCPU::segment (:287-324): centroid offset (x, y) and length (z)
__device__ double3 segment(const Geo& G, int x, int y, double2 n)
{
    const double offset = line_offset(n, frac(G, x, y));
    double low = 10.0, high = -10.0;
    for (int i = 0; i <= 1; ++i) {
        if (fabs(n.x) > 1e-7) {
            const double px = (offset - n.y * i) / n.x;
            if (px >= 0.0 && px <= 1.0) {
                const double s = -px * n.y + i * n.x;
                low = tmin(low, s);
                high = tmax(high, s);
            }
        }
        if (fabs(n.y) > 1e-7) {
            const double py = (offset - n.x * i) / n.y;
            if (py >= 0.0 && py <= 1.0) {
                const double s = (double)(-i) * n.y + py * n.x;
                low = tmin(low, s);
                high = tmax(high, s);
            }
        }
    }
    if (high <= low)
        return make_double3(0.0, 0.0, 0.0);
    const double middle = 0.5 * (low + high);
    return make_double3(n.x * offset - n.y * middle - 0.5, n.y * offset + n.x * middle - 0.5, high - low);
}

#define CELL_XY(G)                                          \
    const int x = blockIdx.x * kCX + (threadIdx.x & 63);    \
    const int y = blockIdx.y * kCY + (threadIdx.x >> 6);    \
    if (x >= (G).W || y >= (G).H)                           \
        return;                                             \
    const int i = y * (G).W + x;

// FluidSurfaceTensionCPU::buildHeight (:134-157)
__global__ __launch_bounds__(kCT) void k_st_heights(Geo G, Heights Hs)
{
    CELL_XY(G)
    double low = 1.0, high = 0.0;
    for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
            const double f = frac(G, x + dx, y + dy);
            low = tmin(low, f);
            high = tmax(high, f);
        }
    double hx = 0.0, ox = 0.0, hy = 0.0, oy = 0.0;
    if (low != high) {
        column(G, x, y, 1, 0, hx, ox);
        column(G, x, y, 0, 1, hy, oy);
    }
    Hs.hv[i] = make_double2(hx, hy);
    Hs.ov[i] = make_double2(ox, oy);
}

// FluidSurfaceTensionCPU::buildCurvature (:257-285)
__global__ __launch_bounds__(kCT) void k_st_raw(Geo G, Heights Hs, double2* raw)
{
    CELL_XY(G)
    double2 k = make_double2(0.0, 0.0);
    if (interfacial(G, x, y)) {
        const double2 n = outward(G, x, y);
        const bool horizontal = fabs(n.x) > fabs(n.y);
        k = height_curvature(Hs, G, x, y, horizontal);
        if (k.y == 0.0)
            k = height_curvature(Hs, G, x, y, !horizontal);
        if (k.y == 0.0) {
            Fit fit;
            for (int j = -1; j <= 1; ++j) {
                double2 hx, ox, hy, oy;
                height_at(Hs, G, x, y + j, hx, ox);
                height_at(Hs, G, x + j, y, hy, oy);
                if (ox.x * n.x > 0.0)
                    fit.add(hx.x, (double)j, n, 1.0);
                if (oy.y * n.y > 0.0)
                    fit.add((double)j, hy.y, n, 1.0);
            }
            k = fit.evaluate(G.h);
        }
    }
    raw[i] = k;
}

// FluidSurfaceTensionCPU::completeCurvature (:326-376)
__global__ __launch_bounds__(kCT) void k_st_complete(Geo G, const double2* raw, double2* curv)
{
    CELL_XY(G)
    double2 k = raw[i];
    if (k.y == 0.0 && interfacial(G, x, y)) {
        double sum = 0.0;
        int count = 0;
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const double2 nb = raw[mirror(y + dy, G.H) * G.W + mirror(x + dx, G.W)];
                if (nb.y != 0.0) {
                    sum += nb.x;
                    ++count;
                }
            }
        if (count > 0) {
            k = make_double2(sum / count, 1.0);
        } else {
            const double2 n = outward(G, x, y);
            Fit fit;
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const double f = frac(G, x + dx, y + dy);
                    if (f <= 0.0 || f >= 1.0)
                        continue;
                    const double2 nq = outward(G, x + dx, y + dy);
                    if (n.x * nq.x + n.y * nq.y <= 0.0)
                        continue;
                    const double3 s = segment(G, x + dx, y + dy, nq);
                    if (s.z > 0.0)
                        fit.add(s.x + dx, s.y + dy, n, s.z);
                }
            k = fit.evaluate(G.h);
        }
    }
    curv[i] = k;
}

// FluidSurfaceTensionCPU::faceCurvature (:378-386)
__device__ __forceinline__ double face_curvature(const double2* curv, int left, int right)
{
    const double2 a = curv[left], b = curv[right];
    if (a.y == 0.0)
        return b.y != 0.0 ? b.x : 0.0;
    return b.y != 0.0 ? 0.5 * (a.x + b.x) : a.x;
}

// Face force (FluidGridCPU.cpp:900-925); z = 0: interior u faces, z = 1: interior v faces.  The face
// weights are the level-0 conductances right64 / down64 (padded cell layout).
__global__ __launch_bounds__(kCT) void k_st_force(Geo G, const double2* curv, const double* right64,
    const double* down64, double sigma, double dt, double* u, double* v)
{
    const int x = blockIdx.x * kCX + (threadIdx.x & 63);
    const int y = blockIdx.y * kCY + (threadIdx.x >> 6);
    const int W = G.W, H = G.H, s = W + 2;
    if (blockIdx.z == 0) {
        if (x < 1 || x >= W || y >= H)
            return;
        const int cell = y * W + x;
        const double jump = sigma * face_curvature(curv, cell - 1, cell) * (G.f[cell] - G.f[cell - 1]);
        const double w = right64[(y + 1) * s + x];   // face between cells x - 1 and x
        const int face = y * (W + 1) + x;
        u[face] += dt * w * jump / G.h;
    } else {
        if (x >= W || y < 1 || y >= H)
            return;
        const int cell = y * W + x;
        const double jump = sigma * face_curvature(curv, cell - W, cell) * (G.f[cell] - G.f[cell - W]);
        const double w = down64[y * s + x + 1];      // face between cells y - 1 and y
        v[cell] += dt * w * jump / G.h;
    }
}

}  // namespace

void launch_capillary_geometry(hipStream_t st, int W, int H, double cell_size, const double* fraction,
    double2* heights, double2* orientations, double2* raw, double2* curvature)
{
    const Geo G{W, H, cell_size, fraction};
    const Heights Hs{heights, orientations};
    const dim3 grid((W + kCX - 1) / kCX, (H + kCY - 1) / kCY);
    hipLaunchKernelGGL(k_st_heights, grid, dim3(kCT), 0, st, G, Hs);
    hipLaunchKernelGGL(k_st_raw, grid, dim3(kCT), 0, st, G, Hs, raw);
    hipLaunchKernelGGL(k_st_complete, grid, dim3(kCT), 0, st, G, (const double2*)raw, curvature);
}

void launch_capillary_force(hipStream_t st, int W, int H, double cell_size, const double* fraction,
    const double2* curvature, const double* right64, const double* down64, double sigma, double dt, double* u,
    double* v)
{
    const Geo G{W, H, cell_size, fraction};
    const dim3 grid((W + kCX - 1) / kCX, (H + kCY - 1) / kCY, 2);
    hipLaunchKernelGGL(k_st_force, grid, dim3(kCT), 0, st, G, curvature, right64, down64, sigma, dt, u, v);
}

}  // namespace fsim
