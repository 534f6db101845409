/*
 * fs_tension.c -- FP64 CPU restatement of the reference's capillary geometry.
 *
 * TEST INFRASTRUCTURE ONLY (the parity checker; never the product path).
 *
 * Follows FluidSim/UniformCPU/FluidSurfaceTensionCPU.h:13-390 (height functions, height-function
 * curvature with a quadratic-fit fallback, completion of missing curvatures) and the MYC normal / line
 * offset of FluidSim/Shared/Adaptive/FluidInterfaceGeometry.h:14-61, single-threaded, in the reference's
 * evaluation order.  Walls reflect the fraction field evenly (a neutral 90-degree contact angle) and
 * negate mirrored heights.
 */
#include "fs_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct { double x, y; } tvec;
typedef struct { double x, y, ox, oy; } theight;
typedef struct { double value; int valid; } tcurv;

typedef struct {
    int W, H;
    double h;
    const double* liquid;
    theight* heights;
    tcurv* raw;
    tcurv* curv;
} tgeom;

static double tmin(double a, double b) { return b < a ? b : a; }   /* std::min */
static double tmax(double a, double b) { return a < b ? b : a; }   /* std::max */
static double tclamp(double v, double lo, double hi) { return v < lo ? lo : (hi < v ? hi : v); }

/* FluidSurfaceTensionCPU::mirror (:42-47) */
static int mirror(int i, int extent)
{
    const int period = 2 * extent;
    const int q = ((i % period) + period) % period;
    return q >= extent ? period - 1 - q : q;
}

/* FluidSurfaceTensionCPU::fraction (:49-63) */
static double frac(const tgeom* G, int x, int y)
{
    const double f = G->liquid[(size_t)mirror(y, G->H) * G->W + mirror(x, G->W)];
    const double roundoff = 32.0 * DBL_EPSILON;
    if (f <= roundoff)
        return 0.0;
    if (f >= 1.0 - roundoff)
        return 1.0;
    return f;
}

/* FluidSurfaceTensionCPU::interfacial (:65-78) */
static int interfacial(const tgeom* G, int x, int y)
{
    const double f = frac(G, x, y);
    if (f > 0.0 && f < 1.0)
        return 1;
    const double nb[4] = {frac(G, x - 1, y), frac(G, x + 1, y), frac(G, x, y - 1), frac(G, x, y + 1)};
    for (int i = 0; i < 4; ++i)
        if ((f >= 1.0 && nb[i] <= 0.0) || (f <= 0.0 && nb[i] >= 1.0))
            return 1;
    return 0;
}

/* interface_geometry::normal (FluidInterfaceGeometry.h:14-49) on the mirrored, snapped samples */
static tvec myc_normal(const tgeom* G, int x, int y)
{
    const double tl = frac(G, x - 1, y - 1), t = frac(G, x, y - 1), tr = frac(G, x + 1, y - 1);
    const double l = frac(G, x - 1, y), r = frac(G, x + 1, y);
    const double bl = frac(G, x - 1, y + 1), b = frac(G, x, y + 1), br = frac(G, x + 1, y + 1);
    const double left = tl + l + bl;
    const double right = tr + r + br;
    const double top = tl + t + tr;
    const double bottom = bl + b + br;
    tvec c = {0.5 * (left - right), 0.5 * (top - bottom)};
    const tvec yg = {left + l - right - r, top + t - bottom - b};
    if (c.x == 0.0 && c.y == 0.0) {
        if (yg.x == 0.0 && yg.y == 0.0) {
            const tvec e = {1.0, 0.0};
            return e;
        }
        return yg;
    }
    if (fabs(c.x) <= fabs(c.y)) {
        c.y = copysign(1.0, c.y);
        if (fabs(yg.x) > fabs(c.x * yg.y))
            return yg;
    } else {
        c.x = copysign(1.0, c.x);
        if (fabs(yg.y) > fabs(c.y * yg.x))
            return yg;
    }
    return c;
}

/* FluidSurfaceTensionCPU::outward (:80-92) */
static tvec outward(const tgeom* G, int x, int y)
{
    tvec n = myc_normal(G, x, y);
    const double magnitude = tmax(fabs(n.x), fabs(n.y));
    n.x /= magnitude;
    n.y /= magnitude;
    const double length = sqrt(n.x * n.x + n.y * n.y);
    const tvec out = {n.x / length, n.y / length};
    return out;
}

/* interface_geometry::offset (FluidInterfaceGeometry.h:51-61) */
static double line_offset(tvec n, double fraction)
{
    const double a = tmin(fabs(n.x), fabs(n.y));
    const double b = tmax(fabs(n.x), fabs(n.y));
    const double triangle = 0.5 * a / b;
    const double cut = fraction < triangle ? sqrt(2.0 * a * b * fraction)
                       : fraction <= 1.0 - triangle ? b * fraction + 0.5 * a
                                                    : a + b - sqrt(2.0 * a * b * (1.0 - fraction));
    return tmin(n.x, 0.0) + tmin(n.y, 0.0) + cut;
}

/* FluidSurfaceTensionCPU::column (:94-132): signed distance to the nearest pure-phase change along an
 * axis and its orientation (+1 liquid below the change, -1 above). */
static void column(const tgeom* G, int x, int y, int ax, int ay, double* height, double* orientation)
{
    int start = -4;
    double base = frac(G, x - 4 * ax, y - 4 * ay);
    int have = base <= 0.0 || base >= 1.0;
    base = base >= 1.0 ? 1.0 : 0.0;
    double sum = 0.0;
    double nearest = 10.0;
    *height = 0.0;
    *orientation = 0.0;
    for (int k = -3; k <= 4; ++k) {
        const double f = frac(G, x + k * ax, y + k * ay);
        if (f <= 0.0 || f >= 1.0) {
            const double pure = f >= 1.0 ? 1.0 : 0.0;
            if (have && pure != base) {
                const double h = (double)start + 0.5 + (base == 1.0 ? sum : (double)(k - start - 1) - sum);
                if (fabs(h) < nearest) {
                    nearest = fabs(h);
                    *height = h;
                    *orientation = base == 1.0 ? 1.0 : -1.0;
                }
            }
            start = k;
            base = pure;
            have = 1;
            sum = 0.0;
        } else {
            sum += f;
        }
    }
}

/* FluidSurfaceTensionCPU::buildHeight (:134-157) */
static void build_height(tgeom* G, int x, int y)
{
    theight* hh = &G->heights[(size_t)y * G->W + x];
    memset(hh, 0, sizeof(*hh));
    double low = 1.0, high = 0.0;
    for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
            const double f = frac(G, x + dx, y + dy);
            low = tmin(low, f);
            high = tmax(high, f);
        }
    if (low == high)
        return;
    double hx, ox, hy, oy;
    column(G, x, y, 1, 0, &hx, &ox);
    column(G, x, y, 0, 1, &hy, &oy);
    hh->x = hx;
    hh->y = hy;
    hh->ox = ox;
    hh->oy = oy;
}

/* FluidSurfaceTensionCPU::heightAt (:159-175) */
static theight height_at(const tgeom* G, int x, int y)
{
    theight h = G->heights[(size_t)mirror(y, G->H) * G->W + mirror(x, G->W)];
    if (x < 0 || x >= G->W) {
        h.x = -h.x;
        h.ox = -h.ox;
    }
    if (y < 0 || y >= G->H) {
        h.y = -h.y;
        h.oy = -h.oy;
    }
    return h;
}

/* FluidSurfaceTensionCPU::heightCurvature (:177-200) */
static tcurv height_curvature(const tgeom* G, int x, int y, int horizontal)
{
    tcurv none = {0.0, 0};
    const int tx = horizontal ? 0 : 1, ty = horizontal ? 1 : 0;
    const theight a = height_at(G, x - tx, y - ty);
    const theight b = height_at(G, x, y);
    const theight c = height_at(G, x + tx, y + ty);
    const double orientation = horizontal ? b.ox : b.oy;
    if (orientation == 0.0 || (horizontal ? a.ox : a.oy) != orientation || (horizontal ? c.ox : c.oy) != orientation)
        return none;
    const double left = horizontal ? a.x : a.y;
    const double mid = horizontal ? b.x : b.y;
    const double right = horizontal ? c.x : c.y;
    const double slope = 0.5 * (right - left);
    const double second = (right - mid) - (mid - left);
    const double k = -orientation * second / pow(1.0 + slope * slope, 1.5);
    tcurv out = {tclamp(k, -1.0, 1.0) / G->h, 1};
    return out;
}

typedef struct { double s0, s1, s2, s3, s4, y0, y1, y2; } tfit;

/* Fit::add (:213-228) */
static void fit_add(tfit* F, double px, double py, tvec n, double weight)
{
    const double x = -px * n.y + py * n.x;
    const double y = px * n.x + py * n.y;
    const double x2 = x * x;
    F->s0 += weight;
    F->s1 += weight * x;
    F->s2 += weight * x2;
    F->s3 += weight * x2 * x;
    F->s4 += weight * x2 * x2;
    F->y0 += weight * y;
    F->y1 += weight * x * y;
    F->y2 += weight * x2 * y;
}

/* Fit::evaluate (:230-254) */
static tcurv fit_evaluate(const tfit* F, double cell_size)
{
    tcurv none = {0.0, 0};
    if (F->s0 == 0.0)
        return none;
    const double a = F->s2 - F->s1 * F->s1 / F->s0;
    const double b = F->s3 - F->s1 * F->s2 / F->s0;
    const double c = F->s4 - F->s2 * F->s2 / F->s0;
    const double d = F->y1 - F->s1 * F->y0 / F->s0;
    const double e = F->y2 - F->s2 * F->y0 / F->s0;
    const double determinant = a * c - b * b;
    if (determinant <= 1e-5 * tmax(a * c, 1e-12))
        return none;
    const double quadratic = (a * e - b * d) / determinant;
    const double slope = (c * d - b * e) / determinant;
    const double k = -2.0 * quadratic / pow(1.0 + slope * slope, 1.5);
    tcurv out = {tclamp(k, -2.0, 2.0) / cell_size, 1};
    return out;
}

/* FluidSurfaceTensionCPU::buildCurvature (:257-285) */
static void build_curvature(tgeom* G, int x, int y)
{
    tcurv* k = &G->raw[(size_t)y * G->W + x];
    k->value = 0.0;
    k->valid = 0;
    if (!interfacial(G, x, y))
        return;
    const tvec n = outward(G, x, y);
    const int horizontal = fabs(n.x) > fabs(n.y);
    *k = height_curvature(G, x, y, horizontal);
    if (!k->valid)
        *k = height_curvature(G, x, y, !horizontal);
    if (k->valid)
        return;
    tfit fit;
    memset(&fit, 0, sizeof(fit));
    for (int j = -1; j <= 1; ++j) {
        const theight hx = height_at(G, x, y + j);
        const theight hy = height_at(G, x + j, y);
        if (hx.ox * n.x > 0.0)
            fit_add(&fit, hx.x, (double)j, n, 1.0);
        if (hy.oy * n.y > 0.0)
            fit_add(&fit, (double)j, hy.y, n, 1.0);
    }
    *k = fit_evaluate(&fit, G->h);
}

/* FluidSurfaceTensionCPU::segment (:287-324): centroid (relative to the cell centre) and length of the
 * cell's interface segment, or length 0. */
static void segment(const tgeom* G, int x, int y, tvec n, double out[3])
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
    if (high <= low) {
        out[0] = out[1] = out[2] = 0.0;
        return;
    }
    const double middle = 0.5 * (low + high);
    out[0] = n.x * offset - n.y * middle - 0.5;
    out[1] = n.y * offset + n.x * middle - 0.5;
    out[2] = high - low;
}

/* FluidSurfaceTensionCPU::completeCurvature (:326-376) */
static void complete_curvature(tgeom* G, int x, int y)
{
    tcurv* k = &G->curv[(size_t)y * G->W + x];
    *k = G->raw[(size_t)y * G->W + x];
    if (k->valid || !interfacial(G, x, y))
        return;
    double sum = 0.0;
    int count = 0;
    for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
            const tcurv nb = G->raw[(size_t)mirror(y + dy, G->H) * G->W + mirror(x + dx, G->W)];
            if (nb.valid) {
                sum += nb.value;
                ++count;
            }
        }
    if (count > 0) {
        k->value = sum / count;
        k->valid = 1;
        return;
    }
    const tvec n = outward(G, x, y);
    tfit fit;
    memset(&fit, 0, sizeof(fit));
    for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
            const double f = frac(G, x + dx, y + dy);
            if (f <= 0.0 || f >= 1.0)
                continue;
            const tvec nq = outward(G, x + dx, y + dy);
            if (n.x * nq.x + n.y * nq.y <= 0.0)
                continue;
            double s[3];
            segment(G, x + dx, y + dy, nq, s);
            if (s[2] > 0.0)
                fit_add(&fit, s[0] + dx, s[1] + dy, n, s[2]);
        }
    *k = fit_evaluate(&fit, G->h);
}

void orc_capillary_geometry(int W, int H, double cell_size, const double* fraction, double* heights,
    double* raw, double* curvature)
{
    const size_t N = (size_t)W * H;
    tgeom G;
    G.W = W;
    G.H = H;
    G.h = cell_size;
    G.liquid = fraction;
    G.heights = (theight*)calloc(N ? N : 1, sizeof(theight));
    G.raw = (tcurv*)calloc(N ? N : 1, sizeof(tcurv));
    G.curv = (tcurv*)calloc(N ? N : 1, sizeof(tcurv));
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x)
            build_height(&G, x, y);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x)
            build_curvature(&G, x, y);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x)
            complete_curvature(&G, x, y);
    for (size_t i = 0; i < N; ++i) {
        if (heights) {
            heights[4 * i] = G.heights[i].x;
            heights[4 * i + 1] = G.heights[i].y;
            heights[4 * i + 2] = G.heights[i].ox;
            heights[4 * i + 3] = G.heights[i].oy;
        }
        if (raw) {
            raw[2 * i] = G.raw[i].value;
            raw[2 * i + 1] = G.raw[i].valid ? 1.0 : 0.0;
        }
        if (curvature) {
            curvature[2 * i] = G.curv[i].value;
            curvature[2 * i + 1] = G.curv[i].valid ? 1.0 : 0.0;
        }
    }
    free(G.heights);
    free(G.raw);
    free(G.curv);
}

/* FluidSurfaceTensionCPU::faceCurvature (:378-386) on the (value, valid) pairs of two cells */
double orc_face_curvature(const double* curvature, size_t left, size_t right)
{
    const int av = curvature[2 * left + 1] != 0.0, bv = curvature[2 * right + 1] != 0.0;
    const double a = curvature[2 * left], b = curvature[2 * right];
    if (!av)
        return bv ? b : 0.0;
    return bv ? 0.5 * (a + b) : a;
}

/* Test hooks: the quadratic fit and the interface segment on their own, so that each is checked against
 * closed forms (exact parabolas, straight chords) rather than only through a whole geometry pass. */

/* Fit::add over n points, then Fit::evaluate (FluidSurfaceTensionCPU.h:204-254); out = (value, valid). */
void orc_fit_curvature(int n, const double* px, const double* py, const double* weight, double nx, double ny,
    double cell_size, double* out)
{
    tfit fit;
    memset(&fit, 0, sizeof(fit));
    const tvec normal = {nx, ny};
    for (int i = 0; i < n; ++i)
        fit_add(&fit, px[i], py[i], normal, weight ? weight[i] : 1.0);
    const tcurv k = fit_evaluate(&fit, cell_size);
    out[0] = k.value;
    out[1] = k.valid ? 1.0 : 0.0;
}

/* FluidSurfaceTensionCPU::segment (:287-324) of a cell with liquid fraction `fraction` cut by a line of
 * normal (nx, ny): out = (centroid x, centroid y relative to the cell centre, length). */
void orc_interface_segment(double nx, double ny, double fraction, double* out)
{
    tgeom G;
    memset(&G, 0, sizeof(G));
    G.W = 1;
    G.H = 1;
    G.h = 1.0;
    G.liquid = &fraction;
    const tvec n = {nx, ny};
    segment(&G, 0, 0, n, out);
}

/* interface_geometry::offset (FluidInterfaceGeometry.h:51-61) */
double orc_line_offset(double nx, double ny, double fraction)
{
    const tvec n = {nx, ny};
    return line_offset(n, fraction);
}
