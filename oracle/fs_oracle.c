/*
 * fs_oracle.c -- FP64 CPU restatement of the reference UniformCPU pressure projection.
 *
 * TEST INFRASTRUCTURE ONLY (the parity checker; never the product path).
 *
 * Every function runs the reference's single-thread order by default: the reference splits passes
 * over worker batches only when a pass has >= 65536 items and more than one thread is allowed
 * (FluidGridCPU.cpp:25-57, PressureMultigrid.cpp:13-21); with one thread each pass is one
 * batch, so that is the bitwise-reproducible mode.  orc_set_threads(n > 1) runs the projection's
 * passes on the reference's worker counts (OpenMP; used to time the CPU baseline on the host's
 * cores): element-wise passes give the same values, and reductions merge per-batch partials in batch
 * order as reduceBatches does, over this file's own batch split.  Sums that the reference forms with
 * std::reduce follow libstdc++'s order (groups of four, ((a0+a1)+(a2+a3)) added to the
 * running total); 4-lane partial sums are reproduced where the reference writes them out.
 *
 * Build with -ffp-contract=off so no multiply-add is fused.
 */
#include "fs_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------------------------ */
/* helpers                                                                                    */
/* ------------------------------------------------------------------------------------------ */

static double dmax(double a, double b) { return a < b ? b : a; }   /* std::max(a, b) */
static int imax(int a, int b) { return a < b ? b : a; }
static int imin(int a, int b) { return b < a ? b : a; }
static double dclamp(double v, double lo, double hi) { return v < lo ? lo : (hi < v ? hi : v); }

static double* dalloc(size_t n) { return (double*)calloc(n ? n : 1, sizeof(double)); }

/* ---- worker batches (Thread::forEachBatch) ---- */
static int g_threads = 1;

void orc_set_threads(int threads) { g_threads = threads < 1 ? 1 : threads; }
int orc_get_threads(void) { return g_threads; }

/* batchWorkers, FluidGridCPU.cpp:25-31: one batch below 65536 items of work, else up to 8 workers. */
static int batch_workers(long long work) { return work < 65536 || g_threads <= 1 ? 1 : imin(8, g_threads); }

/* forRows, PressureMultigrid.cpp:13-21: levels of at least 65536 cells on 4 workers (8 from 1M cells). */
static int row_workers(int w, int h)
{
    const long long cells = (long long)w * h;
    const int workers = imin(cells >= 1048576 ? 8 : 4, g_threads);
    return cells >= 65536 && workers > 1 ? workers : 1;
}

/* First item of batch b when `count` items are split over `workers` batches (bitloop's own split is not
 * available; with one worker the single batch is [0, count)). */
static int batch_first(int count, int b, int workers) { return (int)((long long)count * b / workers); }

/* std::reduce(first, last, 0.0) as libstdc++ evaluates it. */
static double reduce_sum(const double* a, size_t n)
{
    double init = 0.0;
    size_t i = 0;
    for (; n - i >= 4; i += 4) {
        const double v1 = a[i] + a[i + 1];
        const double v2 = a[i + 2] + a[i + 3];
        init = init + (v1 + v2);
    }
    for (; i < n; ++i)
        init = init + a[i];
    return init;
}

/* sumValues, FluidGridCPU.cpp:59-65: std::reduce per batch, batches merged in order. */
static double sum_values(const double* a, size_t n)
{
    const int workers = batch_workers((long long)n);
    double part[8];
#pragma omp parallel for num_threads(workers) schedule(static, 1) if (workers > 1)
    for (int b = 0; b < workers; ++b) {
        const int first = batch_first((int)n, b, workers), last = batch_first((int)n, b + 1, workers);
        part[b] = reduce_sum(a + first, (size_t)(last - first));
    }
    double total = part[0];
    for (int b = 1; b < workers; ++b)
        total = total + part[b];
    return total;
}

/* FluidGridCPU.cpp:1488-1505 maxAbs: non-finite maps to +inf. */
static double max_abs(const double* a, size_t n)
{
    const int workers = batch_workers((long long)n);
    double part[8];
#pragma omp parallel for num_threads(workers) schedule(static, 1) if (workers > 1)
    for (int b = 0; b < workers; ++b) {
        double m = 0.0;
        for (int i = batch_first((int)n, b, workers); i < batch_first((int)n, b + 1, workers); ++i) {
            const double v = isfinite(a[i]) ? fabs(a[i]) : INFINITY;
            m = dmax(m, v);
        }
        part[b] = m;
    }
    double m = part[0];
    for (int b = 1; b < workers; ++b)
        m = dmax(m, part[b]);
    return m;
}

/* std::lerp as libstdc++ evaluates it (cmath __lerp). */
static double lerp_std(double a, double b, double t)
{
    if ((a <= 0 && b >= 0) || (a >= 0 && b <= 0))
        return t * b + (1 - t) * a;
    if (t == 1)
        return b;
    const double x = a + t * (b - a);
    return (t > 1) == (b > a) ? (b < x ? x : b) : (b > x ? x : b);
}

/* FluidGridCPU.cpp:67-90 sampleScalar: bilinear blend at fractional array indices, clamped. */
static double sample_scalar(const double* data, int width, int height, double x, double y)
{
    x = dclamp(x, 0.0, (double)(width - 1));
    y = dclamp(y, 0.0, (double)(height - 1));
    const int x0 = (int)x;
    const int y0 = (int)y;
    const int x1 = imin(x0 + 1, width - 1);
    const int y1 = imin(y0 + 1, height - 1);
    const double tx = x - x0;
    const double ty = y - y0;
    const double a = lerp_std(data[y0 * width + x0], data[y0 * width + x1], tx);
    const double b = lerp_std(data[y1 * width + x0], data[y1 * width + x1], tx);
    return lerp_std(a, b, ty);
}

/* ------------------------------------------------------------------------------------------ */
/* PressureMultigrid (PressureMultigrid.cpp)                                                   */
/* ------------------------------------------------------------------------------------------ */

typedef struct { int first, second; double weight; } orc_interp;
typedef struct { int count; int idx[4]; double wt[4]; } orc_restr;

typedef struct {
    int w, h, stride;
    double *cx, *cy;
    double *right, *down, *inv_diag, *rhs, *sol, *res;
    orc_interp *px, *py;   /* fine-level prolongation maps (this level -> from next coarser) */
    orc_restr *rx, *ry;    /* restriction stencils indexed by coarse index */
    int nrx, nry;
} orc_level;

struct orc_mg {
    int nlev;
    orc_level* lev;
    int n;          /* coarse unknowns = w*h - 1 of the last level */
    double* factor; /* n*n */
    double* crhs;   /* n */
    int pin_singular;        /* solid-mask extension: pin singular coarse unknowns instead of failing */
    unsigned char* pinned;   /* n: unknown held at zero by the coarse solve */
};

/* Solid-mask extension (no counterpart in the reference, whose grids have no solids): a coarse
 * Cholesky pivot at or below this fraction of its row's assembled diagonal marks the unknown as the
 * last one of a disconnected or fully blocked region; it is pinned to zero like the reference's
 * first cell.  Without a mask every pivot of these weighted Laplacians is far above the bound. */
#define ORC_PIN_RATIO 1e-9

/* Level::Level, PressureMultigrid.cpp:24-36 */
static void level_init(orc_level* L, int w, int h, int spacing, int dw, int dh)
{
    memset(L, 0, sizeof(*L));
    L->w = w;
    L->h = h;
    L->stride = w + 2;
    const size_t count = (size_t)L->stride * (h + 2);
    L->cx = dalloc(w);
    L->cy = dalloc(h);
    L->right = dalloc(count);
    L->down = dalloc(count);
    L->inv_diag = dalloc(count);
    L->rhs = dalloc(count);
    L->sol = dalloc(count);
    L->res = dalloc(count);
    /* A shorter final cell keeps odd grid sizes inside the original domain. */
    for (int x = 0; x < w; ++x)
        L->cx[x] = 0.5 * (x * spacing + imin((x + 1) * spacing, dw));
    for (int y = 0; y < h; ++y)
        L->cy[y] = 0.5 * (y * spacing + imin((y + 1) * spacing, dh));
}

static void level_free(orc_level* L)
{
    free(L->cx); free(L->cy); free(L->right); free(L->down); free(L->inv_diag);
    free(L->rhs); free(L->sol); free(L->res);
    free(L->px); free(L->py); free(L->rx); free(L->ry);
}

/* buildTransfers axis, PressureMultigrid.cpp:64-98 */
static void build_axis(const double* fc, int nf, const double* cc, int nc, orc_interp** pro, orc_restr** res)
{
    orc_interp* p = (orc_interp*)calloc((size_t)imax(nf, 1), sizeof(orc_interp));
    orc_restr* r = (orc_restr*)calloc((size_t)imax(nc, 1), sizeof(orc_restr));
    for (int f = 0; f < nf; ++f) {
        const int parent = f / 2;
        const int first = imax(0, parent - (int)(fc[f] < cc[parent]));
        const int second = imin(first + 1, nc - 1);
        const double weight = first == second ? 0.0
            : dclamp((fc[f] - cc[first]) / (cc[second] - cc[first]), 0.0, 1.0);
        p[f].first = first;
        p[f].second = second;
        p[f].weight = weight;
        /* restriction is the exact transpose of prolongation */
        const double a = 1.0 - weight;
        if (a != 0.0) {
            orc_restr* s = &r[first];
            s->idx[s->count] = f;
            s->wt[s->count++] = a;
        }
        if (weight != 0.0) {
            orc_restr* s = &r[second];
            s->idx[s->count] = f;
            s->wt[s->count++] = weight;
        }
    }
    *pro = p;
    *res = r;
}

orc_mg* orc_mg_create(int width, int height)
{
    orc_mg* mg = (orc_mg*)calloc(1, sizeof(orc_mg));
    const int dw = width, dh = height;
    int spacing = 1;
    int cap = 4;
    mg->lev = (orc_level*)calloc((size_t)cap, sizeof(orc_level));
    for (;;) {
        if (mg->nlev == cap) {
            cap *= 2;
            mg->lev = (orc_level*)realloc(mg->lev, (size_t)cap * sizeof(orc_level));
        }
        level_init(&mg->lev[mg->nlev++], width, height, spacing, dw, dh);
        if (width * height <= 64)
            break;
        width = (width + 1) / 2;
        height = (height + 1) / 2;
        spacing *= 2;
    }
    for (int i = 0; i + 1 < mg->nlev; ++i) {
        orc_level* F = &mg->lev[i];
        const orc_level* C = &mg->lev[i + 1];
        build_axis(F->cx, F->w, C->cx, C->w, &F->px, &F->rx);
        build_axis(F->cy, F->h, C->cy, C->h, &F->py, &F->ry);
        F->nrx = C->w;
        F->nry = C->h;
    }
    const orc_level* last = &mg->lev[mg->nlev - 1];
    mg->n = last->w * last->h - 1;
    mg->factor = dalloc((size_t)imax(mg->n, 0) * imax(mg->n, 0));
    mg->crhs = dalloc((size_t)imax(mg->n, 0));
    mg->pinned = (unsigned char*)calloc((size_t)imax(mg->n, 1), 1);
    return mg;
}

void orc_mg_destroy(orc_mg* mg)
{
    if (!mg)
        return;
    for (int i = 0; i < mg->nlev; ++i)
        level_free(&mg->lev[i]);
    free(mg->lev);
    free(mg->factor);
    free(mg->crhs);
    free(mg->pinned);
    free(mg);
}

/* Level::rebuildDiagonal, PressureMultigrid.cpp:104-117 */
static void rebuild_diagonal(orc_level* L)
{
    const int s = L->stride;
    const int workers = row_workers(L->w, L->h);
#pragma omp parallel for num_threads(workers) schedule(static, 1) if (workers > 1)
    for (int b = 0; b < workers; ++b)
    for (int y = batch_first(L->h, b, workers); y < batch_first(L->h, b + 1, workers); ++y) {
        for (int x = 0, i = (y + 1) * s + 1; x < L->w; ++x, ++i) {
            const double diagonal = L->right[i - 1] + L->right[i] + L->down[i - s] + L->down[i];
            L->inv_diag[i] = diagonal > 0.0 ? 1.0 / diagonal : 0.0;
        }
    }
}

/* coarsen, PressureMultigrid.cpp:142-202: parallel paths add; series resistances across density jumps. */
static void coarsen(const orc_level* F, orc_level* C)
{
    int workers = row_workers(C->w, C->h);
#pragma omp parallel for num_threads(workers) schedule(static, 1) if (workers > 1)
    for (int b = 0; b < workers; ++b)
    for (int y = batch_first(C->h, b, workers); y < batch_first(C->h, b + 1, workers); ++y) {
        for (int x = 0; x + 1 < C->w; ++x) {
            const int edge = 2 * x + 2;
            const double left = (F->cx[edge - 1] - C->cx[x]) / (F->cx[edge - 1] - F->cx[edge - 2]);
            const double right = edge + 1 < F->w
                ? (C->cx[x + 1] - F->cx[edge]) / (F->cx[edge + 1] - F->cx[edge]) : 0.0;
            double conductance = 0.0;
            for (int fy = 2 * y; fy < imin(2 * y + 2, F->h); ++fy) {
                const int i = (fy + 1) * F->stride + edge;
                double resistance = left / F->right[i - 1] + 1.0 / F->right[i];
                if (right != 0.0)
                    resistance += right / F->right[i + 1];
                conductance += 1.0 / resistance;
            }
            C->right[(y + 1) * C->stride + x + 1] = conductance;
        }
    }
    workers = row_workers(C->w, C->h - 1);
#pragma omp parallel for num_threads(workers) schedule(static, 1) if (workers > 1)
    for (int b = 0; b < workers; ++b)
    for (int y = batch_first(C->h - 1, b, workers); y < batch_first(C->h - 1, b + 1, workers); ++y) {
        for (int x = 0; x < C->w; ++x) {
            const int edge = 2 * y + 2;
            const double up = (F->cy[edge - 1] - C->cy[y]) / (F->cy[edge - 1] - F->cy[edge - 2]);
            const double down = edge + 1 < F->h
                ? (C->cy[y + 1] - F->cy[edge]) / (F->cy[edge + 1] - F->cy[edge]) : 0.0;
            double conductance = 0.0;
            for (int fx = 2 * x; fx < imin(2 * x + 2, F->w); ++fx) {
                const int i = edge * F->stride + fx + 1;
                double resistance = up / F->down[i - F->stride] + 1.0 / F->down[i];
                if (down != 0.0)
                    resistance += down / F->down[i + F->stride];
                conductance += 1.0 / resistance;
            }
            C->down[(y + 1) * C->stride + x + 1] = conductance;
        }
    }
    rebuild_diagonal(C);
}

/* factorCoarsest, PressureMultigrid.cpp:300-345: first cell pinned, dense Cholesky.  With
 * pin_singular (solid masks) a vanishing pivot pins its unknown: its row and column of the factor
 * are zero and the solve holds it at zero. */
static int factor_coarsest(orc_mg* mg)
{
    const orc_level* L = &mg->lev[mg->nlev - 1];
    const int n = mg->n;
    if (n > 0)
        memset(mg->factor, 0, sizeof(double) * (size_t)n * n);
    memset(mg->pinned, 0, (size_t)imax(n, 1));
    double* f = mg->factor;
    for (int cell = 1; cell <= n; ++cell) {
        const int x = cell % L->w;
        const int y = cell / L->w;
        const int i = (y + 1) * L->stride + x + 1;
        const int row = cell - 1;
        f[row * n + row] = L->right[i - 1] + L->right[i] + L->down[i - L->stride] + L->down[i];
        if (x > 0 && cell - 1 > 0)
            f[row * n + row - 1] = -L->right[i - 1];
        if (y > 0 && cell - L->w > 0)
            f[row * n + row - L->w] = -L->down[i - L->stride];
    }
    for (int row = 0; row < n; ++row) {
        for (int col = 0; col <= row; ++col) {
            double value = f[row * n + col];
            const double assembled = value;
            for (int k = 0; k < col; ++k)
                value -= f[row * n + k] * f[col * n + k];
            if (row == col) {
                if (mg->pin_singular && isfinite(value) && fabs(value) <= ORC_PIN_RATIO * assembled) {
                    mg->pinned[row] = 1;
                    for (int k = 0; k <= row; ++k)
                        f[row * n + k] = 0.0;
                    continue;
                }
                if (value <= 0.0 || !isfinite(value))
                    return 0;
                f[row * n + col] = sqrt(value);
            } else
                f[row * n + col] = mg->pinned[col] ? 0.0 : value / f[col * n + col];
        }
    }
    return 1;
}

/* PressureMultigrid::rebuild, PressureMultigrid.cpp:119-140 */
int orc_mg_rebuild(orc_mg* mg, const double* inv_u, const double* inv_v)
{
    orc_level* F = &mg->lev[0];
    const int workers = row_workers(F->w, F->h);
#pragma omp parallel for num_threads(workers) schedule(static, 1) if (workers > 1)
    for (int b = 0; b < workers; ++b)
    for (int y = batch_first(F->h, b, workers); y < batch_first(F->h, b + 1, workers); ++y) {
        for (int x = 0, i = (y + 1) * F->stride + 1; x < F->w; ++x, ++i) {
            F->right[i] = inv_u[y * (F->w + 1) + x + 1];
            F->down[i] = inv_v[(y + 1) * F->w + x];
        }
    }
    rebuild_diagonal(F);
    for (int i = 0; i + 1 < mg->nlev; ++i)
        coarsen(&mg->lev[i], &mg->lev[i + 1]);
    return factor_coarsest(mg);
}

/* smooth, PressureMultigrid.cpp:204-231: red-black Gauss-Seidel, colour order reversed on the way up. */
static void smooth(orc_level* L, int reverse, int from_zero)
{
    const int s = L->stride;
    const int workers = row_workers(L->w, L->h);
    for (int pass = 0; pass < 2; ++pass) {
        const int color = reverse ? 1 - pass : pass;
#pragma omp parallel for num_threads(workers) schedule(static, 1) if (workers > 1)
        for (int b = 0; b < workers; ++b)
        for (int y = batch_first(L->h, b, workers); y < batch_first(L->h, b + 1, workers); ++y) {
            for (int x = (color ^ (y & 1)), i = (y + 1) * s + x + 1; x < L->w; x += 2, i += 2) {
                double value = L->rhs[i];
                if (!(from_zero && pass == 0)) {
                    value += (L->right[i - 1] * L->sol[i - 1] + L->right[i] * L->sol[i + 1]) +
                             (L->down[i - s] * L->sol[i - s] + L->down[i] * L->sol[i + s]);
                }
                L->sol[i] = value * L->inv_diag[i];
            }
        }
    }
}

/* restrictResidual, PressureMultigrid.cpp:233-276 */
static void restrict_residual(orc_level* F, orc_level* C)
{
    const int s = F->stride;
    int workers = row_workers(F->w, F->h);
#pragma omp parallel for num_threads(workers) schedule(static, 1) if (workers > 1)
    for (int b = 0; b < workers; ++b)
    for (int y = batch_first(F->h, b, workers); y < batch_first(F->h, b + 1, workers); ++y) {
        for (int x = 0, i = (y + 1) * s + 1; x < F->w; ++x, ++i) {
            const double center = F->sol[i];
            const double product = (F->right[i - 1] * (center - F->sol[i - 1]) +
                                       F->right[i] * (center - F->sol[i + 1])) +
                                   (F->down[i - s] * (center - F->sol[i - s]) +
                                       F->down[i] * (center - F->sol[i + s]));
            /* a cell without conductance (only solid masks make one) has a zero row: its residual
             * is not restricted, so the V-cycle neither reads nor writes it and stays symmetric */
            F->res[i] = F->inv_diag[i] > 0.0 ? F->rhs[i] - product : 0.0;
        }
    }
    workers = row_workers(C->w, C->h);
#pragma omp parallel for num_threads(workers) schedule(static, 1) if (workers > 1)
    for (int b = 0; b < workers; ++b)
    for (int y = batch_first(C->h, b, workers); y < batch_first(C->h, b + 1, workers); ++y) {
        for (int x = 0; x < C->w; ++x) {
            const orc_restr* rx = &F->rx[x];
            const orc_restr* ry = &F->ry[y];
            double value = 0.0;
            for (int iy = 0; iy < ry->count; ++iy) {
                const double* row = F->res + (ry->idx[iy] + 1) * s + 1;
                double horizontal = 0.0;
                for (int ix = 0; ix < rx->count; ++ix)
                    horizontal += rx->wt[ix] * row[rx->idx[ix]];
                value += ry->wt[iy] * horizontal;
            }
            C->rhs[(y + 1) * C->stride + x + 1] = value;
        }
    }
}

/* prolongCorrection, PressureMultigrid.cpp:278-298 */
static void prolong_correction(const orc_level* C, orc_level* F)
{
    const int workers = row_workers(F->w, F->h);
#pragma omp parallel for num_threads(workers) schedule(static, 1) if (workers > 1)
    for (int b = 0; b < workers; ++b)
    for (int y = batch_first(F->h, b, workers); y < batch_first(F->h, b + 1, workers); ++y) {
        const orc_interp* py = &F->py[y];
        const double* top = C->sol + (py->first + 1) * C->stride + 1;
        const double* bottom = C->sol + (py->second + 1) * C->stride + 1;
        for (int x = 0, i = (y + 1) * F->stride + 1; x < F->w; ++x, ++i) {
            const orc_interp* px = &F->px[x];
            const double a = (1.0 - px->weight) * top[px->first] + px->weight * top[px->second];
            const double b = (1.0 - px->weight) * bottom[px->first] + px->weight * bottom[px->second];
            F->sol[i] += (1.0 - py->weight) * a + py->weight * b;
        }
    }
}

/* solveCoarsest, PressureMultigrid.cpp:347-384 */
static void solve_coarsest(orc_mg* mg)
{
    orc_level* L = &mg->lev[mg->nlev - 1];
    const int n = mg->n;
    const double* f = mg->factor;
    double* c = mg->crhs;
    double mean = 0.0;
    for (int y = 0; y < L->h; ++y)
        for (int x = 0; x < L->w; ++x)
            mean += L->rhs[(y + 1) * L->stride + x + 1];
    mean /= n + 1;
    for (int row = 0; row < n; ++row) {
        const int cell = row + 1;
        double value = L->rhs[(cell / L->w + 1) * L->stride + cell % L->w + 1] - mean;
        for (int col = 0; col < row; ++col)
            value -= f[row * n + col] * c[col];
        c[row] = mg->pinned[row] ? 0.0 : value / f[row * n + row];
    }
    for (int row = n - 1; row >= 0; --row) {
        double value = c[row];
        for (int col = row + 1; col < n; ++col)
            value -= f[col * n + row] * c[col];
        c[row] = mg->pinned[row] ? 0.0 : value / f[row * n + row];
    }
    /* std::accumulate: sequential */
    double acc = 0.0;
    for (int i = 0; i < n; ++i)
        acc = acc + c[i];
    mean = acc / (n + 1);
    L->sol[L->stride + 1] = -mean;
    for (int cell = 1; cell <= n; ++cell)
        L->sol[(cell / L->w + 1) * L->stride + cell % L->w + 1] = c[cell - 1] - mean;
}

/* Smoother choice of the levels above ORC_SMOOTH_MIN_CELLS cells (an execution choice: the reference's CPU
 * multigrid runs red-black Gauss-Seidel, PressureMultigrid.cpp:204-231, 396; its WebGPU multigrid runs
 * Chebyshev-weighted damped Jacobi on levels above 65536 cells, PressureMultigridGPU.cpp:24-28, 619-631).
 * Smaller levels run red-black Gauss-Seidel with their own sweep count (4 in the reference's V(4,4)).
 * Process-wide, like the worker count. */
#define ORC_SMOOTH_MIN_CELLS 4096
static int g_smooth_kind = ORC_SMOOTH_RBGS;
static int g_smooth_sweeps = 4;
static int g_coarse_sweeps = 4;

int orc_set_smoother(int kind, int sweeps, int coarse_sweeps)
{
    if ((kind != ORC_SMOOTH_RBGS && kind != ORC_SMOOTH_CHEBYSHEV) || sweeps < 1 || sweeps > 8 ||
        coarse_sweeps < 1 || coarse_sweeps > 8)
        return -1;
    g_smooth_kind = kind;
    g_smooth_sweeps = sweeps;
    g_coarse_sweeps = coarse_sweeps;
    return 0;
}

/* The polynomial weights of PressureMultigridGPU's constructor (PressureMultigridGPU.cpp:24-28):
 * 1 / (1.1 + 0.9 cos(pi (i + 1/2) / sweeps)), the reciprocals of the Chebyshev nodes of [0.2, 2] (the
 * spectrum of D^-1 A), rounded to float as the reference stores them. */
double orc_chebyshev_weight(int i, int sweeps)
{
    const double pi = 3.14159265358979323846;
    return (double)(float)(1.0 / (1.1 + 0.9 * cos(pi * (i + 0.5) / sweeps)));
}

/* One damped-Jacobi sweep x <- x + w (b - A x) / diag (polynomial_smooth, PressureSmoothingKernels.h:91-176):
 * from zero it is w b / diag.  A cell without conductance (only solid masks make one; the reference has
 * none) is set to zero, as the Gauss-Seidel sweep's value * inv_diag sets it: keeping a prolonged value
 * there would leak into the mean removal and make the cycle unsymmetric.  The new values go to the
 * level's residual array, which then swaps with the solution. */
static void jacobi(orc_level* L, double weight, int from_zero)
{
    const int s = L->stride;
    const int workers = row_workers(L->w, L->h);
    const double* x = L->sol;
    double* out = L->res;
#pragma omp parallel for num_threads(workers) schedule(static, 1) if (workers > 1)
    for (int b = 0; b < workers; ++b)
    for (int y = batch_first(L->h, b, workers); y < batch_first(L->h, b + 1, workers); ++y) {
        for (int x0 = 0, i = (y + 1) * s + 1; x0 < L->w; ++x0, ++i) {
            if (from_zero) {
                out[i] = weight * L->rhs[i] * L->inv_diag[i];
                continue;
            }
            const double c = x[i];
            const double product = (L->right[i - 1] * (c - x[i - 1]) + L->right[i] * (c - x[i + 1])) +
                                   (L->down[i - s] * (c - x[i - s]) + L->down[i] * (c - x[i + s]));
            out[i] = L->inv_diag[i] > 0.0 ? c + weight * (L->rhs[i] - product) * L->inv_diag[i] : 0.0;
        }
    }
    L->res = L->sol;
    L->sol = out;
}

/* cycle, PressureMultigrid.cpp:386-408: V(4,4) (or the chosen smoother on the large levels) */
static void cycle(orc_mg* mg, int index)
{
    if (index + 1 == mg->nlev) {
        solve_coarsest(mg);
        return;
    }
    orc_level* F = &mg->lev[index];
    orc_level* C = &mg->lev[index + 1];
    const int large = (long)F->w * F->h > ORC_SMOOTH_MIN_CELLS;
    const int kind = large ? g_smooth_kind : ORC_SMOOTH_RBGS;
    const int sweeps = large ? g_smooth_sweeps : g_coarse_sweeps;
    if (kind == ORC_SMOOTH_CHEBYSHEV) {
        for (int k = 0; k < sweeps; ++k)
            jacobi(F, orc_chebyshev_weight(k, sweeps), k == 0);
    } else {
        smooth(F, 0, 1);
        for (int k = 1; k < sweeps; ++k)
            smooth(F, 0, 0);
    }
    restrict_residual(F, C);
    cycle(mg, index + 1);
    prolong_correction(C, F);
    if (kind == ORC_SMOOTH_CHEBYSHEV) {
        for (int k = sweeps - 1; k >= 0; --k)
            jacobi(F, orc_chebyshev_weight(k, sweeps), 0);
    } else {
        for (int k = 0; k < sweeps; ++k)
            smooth(F, 1, 0);
    }
}

/* apply, PressureMultigrid.cpp:410-434 */
void orc_mg_apply(orc_mg* mg, const double* residual, double* correction, double residual_mean)
{
    orc_level* F = &mg->lev[0];
    const int workers = row_workers(F->w, F->h);
#pragma omp parallel for num_threads(workers) schedule(static, 1) if (workers > 1)
    for (int b = 0; b < workers; ++b)
        for (int y = batch_first(F->h, b, workers); y < batch_first(F->h, b + 1, workers); ++y)
            for (int x = 0; x < F->w; ++x)
                F->rhs[(y + 1) * F->stride + x + 1] = residual[y * F->w + x] - residual_mean;
    cycle(mg, 0);
#pragma omp parallel for num_threads(workers) schedule(static, 1) if (workers > 1)
    for (int b = 0; b < workers; ++b)
        for (int y = batch_first(F->h, b, workers); y < batch_first(F->h, b + 1, workers); ++y)
            memcpy(correction + (size_t)y * F->w, F->sol + (y + 1) * F->stride + 1, sizeof(double) * (size_t)F->w);
}

int orc_mg_levels(const orc_mg* mg) { return mg->nlev; }

void orc_mg_level_shape(const orc_mg* mg, int level, int* width, int* height)
{
    *width = mg->lev[level].w;
    *height = mg->lev[level].h;
}

void orc_mg_level_centers(const orc_mg* mg, int level, double* cx, double* cy)
{
    const orc_level* L = &mg->lev[level];
    memcpy(cx, L->cx, sizeof(double) * (size_t)L->w);
    memcpy(cy, L->cy, sizeof(double) * (size_t)L->h);
}

void orc_mg_level_weights(const orc_mg* mg, int level, double* right, double* down, double* inv_diag)
{
    const orc_level* L = &mg->lev[level];
    for (int y = 0; y < L->h; ++y) {
        for (int x = 0; x < L->w; ++x) {
            const int i = (y + 1) * L->stride + x + 1;
            if (right) right[y * L->w + x] = L->right[i];
            if (down) down[y * L->w + x] = L->down[i];
            if (inv_diag) inv_diag[y * L->w + x] = L->inv_diag[i];
        }
    }
}

void orc_mg_level_prolong(const orc_mg* mg, int level, int axis, int* first, int* second, double* weight)
{
    const orc_level* L = &mg->lev[level];
    const orc_interp* p = axis == 0 ? L->px : L->py;
    const int n = axis == 0 ? L->w : L->h;
    for (int i = 0; i < n; ++i) {
        first[i] = p[i].first;
        second[i] = p[i].second;
        weight[i] = p[i].weight;
    }
}

void orc_mg_level_restrict(const orc_mg* mg, int level, int axis, int* count, int* idx, double* wt)
{
    const orc_level* L = &mg->lev[level];
    const orc_restr* r = axis == 0 ? L->rx : L->ry;
    const int n = axis == 0 ? L->nrx : L->nry;
    for (int c = 0; c < n; ++c) {
        count[c] = r[c].count;
        for (int k = 0; k < 4; ++k) {
            idx[c * 4 + k] = k < r[c].count ? r[c].idx[k] : -1;
            wt[c * 4 + k] = k < r[c].count ? r[c].wt[k] : 0.0;
        }
    }
}

void orc_mg_set_pin(orc_mg* mg, int pin) { mg->pin_singular = pin != 0; }

int orc_mg_coarse_pinned(const orc_mg* mg, unsigned char* pinned)
{
    int count = 0;
    for (int i = 0; i < mg->n; ++i) {
        count += mg->pinned[i];
        if (pinned)
            pinned[i] = mg->pinned[i];
    }
    return count;
}

int orc_mg_coarse_factor(const orc_mg* mg, double* factor)
{
    if (factor && mg->n > 0)
        memcpy(factor, mg->factor, sizeof(double) * (size_t)mg->n * mg->n);
    return mg->n;
}

/* ------------------------------------------------------------------------------------------ */
/* FluidGridCPU subset                                                                         */
/* ------------------------------------------------------------------------------------------ */

struct orc_grid {
    int W, H;
    double h;
    double ox, oy;
    double gravity_y;
    int preconditioner; /* 0 diagonal, 1 multigrid */
    int nl, ng;
    double *rho_l, *rho_g;
    double *mu_l, *mu_g;   /* component dynamic viscosities */
    int vis_sweeps;        /* orc_set_viscosity_smoother (0: the reference's smoother) */
    double sigma;          /* surface tension coefficient */
    double *fraction, *u, *v, *p, *p_prev;
    double *lmass, *gmass;
    double *density, *inv_u, *inv_v, *inv_diag;
    double *rhs, *res, *cor, *dir, *prod;
    unsigned char* solid;  /* solid-mask extension: W*H, nonzero = solid cell (NULL: none) */
    orc_mg* mg;
    int density_valid, coeff_valid, mg_valid;
    int relative_tol;      /* extension: tolerances are relative to max|rhs| (orc_set_relative_tolerance) */
};

orc_grid* orc_grid_create(int W, int H, double cell_size)
{
    orc_grid* g = (orc_grid*)calloc(1, sizeof(orc_grid));
    const size_t N = (size_t)W * H;
    g->W = W;
    g->H = H;
    g->h = cell_size;
    g->gravity_y = 100.0; /* FluidGridCPU.h:37 default */
    g->preconditioner = 1;
    g->fraction = dalloc(N);
    g->u = dalloc((size_t)(W + 1) * H);
    g->v = dalloc((size_t)W * (H + 1));
    g->p = dalloc(N);
    g->p_prev = dalloc(N);
    g->density = dalloc(N);
    g->inv_u = dalloc((size_t)(W + 1) * H);
    g->inv_v = dalloc((size_t)W * (H + 1));
    g->inv_diag = dalloc(N);
    g->rhs = dalloc(N);
    g->res = dalloc(N);
    g->cor = dalloc(N);
    g->dir = dalloc(N);
    g->prod = dalloc(N);
    g->mg = orc_mg_create(W, H);
    return g;
}

void orc_grid_destroy(orc_grid* g)
{
    if (!g)
        return;
    free(g->rho_l); free(g->rho_g); free(g->mu_l); free(g->mu_g);
    free(g->fraction); free(g->u); free(g->v); free(g->p); free(g->p_prev);
    free(g->lmass); free(g->gmass);
    free(g->density); free(g->inv_u); free(g->inv_v); free(g->inv_diag);
    free(g->rhs); free(g->res); free(g->cor); free(g->dir); free(g->prod);
    free(g->solid);
    orc_mg_destroy(g->mg);
    free(g);
}

void orc_set_origin(orc_grid* g, double x, double y) { g->ox = x; g->oy = y; }
void orc_set_gravity(orc_grid* g, double gy) { g->gravity_y = gy; }
void orc_set_preconditioner(orc_grid* g, int kind) { g->preconditioner = kind; }

/* Solid-mask extension.  A face with a solid cell on either side is closed like a domain wall: zero
 * pressure face weight, zero velocity after the projection's wall step.  NULL removes the mask. */
void orc_set_solid(orc_grid* g, const unsigned char* mask)
{
    free(g->solid);
    g->solid = NULL;
    if (mask) {
        const size_t N = (size_t)g->W * g->H;
        g->solid = (unsigned char*)malloc(N);
        for (size_t i = 0; i < N; ++i)
            g->solid[i] = mask[i] ? 1 : 0;
    }
    g->mg->pin_singular = mask != NULL;
    orc_invalidate_caches(g);
}

/* Closed faces of the solid mask: the face left of / above cell (x, y). */
static int solid_left(const orc_grid* g, int x, int y)
{
    return g->solid && x > 0 && x < g->W && (g->solid[(size_t)y * g->W + x] || g->solid[(size_t)y * g->W + x - 1]);
}
static int solid_up(const orc_grid* g, int x, int y)
{
    return g->solid && y > 0 && y < g->H &&
           (g->solid[(size_t)y * g->W + x] || g->solid[(size_t)(y - 1) * g->W + x]);
}

void orc_invalidate_caches(orc_grid* g)
{
    g->density_valid = 0;
    g->coeff_valid = 0;
    g->mg_valid = 0;
}

int orc_configure(orc_grid* g, int nl, const double* rl, int ng, const double* rg)
{
    const size_t N = (size_t)g->W * g->H;
    free(g->rho_l); free(g->rho_g); free(g->lmass); free(g->gmass); free(g->mu_l); free(g->mu_g);
    g->nl = nl;
    g->ng = ng;
    g->rho_l = dalloc((size_t)nl);
    g->rho_g = dalloc((size_t)ng);
    g->mu_l = dalloc((size_t)nl);
    g->mu_g = dalloc((size_t)ng);
    for (int k = 0; k < nl; ++k) g->rho_l[k] = rl[k];
    for (int k = 0; k < ng; ++k) g->rho_g[k] = rg[k];
    g->lmass = dalloc((size_t)nl * N);
    g->gmass = dalloc((size_t)ng * N);
    orc_invalidate_caches(g);
    return 0;
}

/* upload, FluidGridCPU.cpp:198-251: empty means zero; the pressure history restarts. */
int orc_upload(orc_grid* g, const double* fraction, const double* u, const double* v, const double* p,
    const double* lm, const double* gm)
{
    const size_t N = (size_t)g->W * g->H;
    const size_t NU = (size_t)(g->W + 1) * g->H, NV = (size_t)g->W * (g->H + 1);
#define ORC_ASSIGN(dst, src, n) do { if (src) memcpy(dst, src, sizeof(double) * (n)); else memset(dst, 0, sizeof(double) * (n)); } while (0)
    ORC_ASSIGN(g->fraction, fraction, N);
    ORC_ASSIGN(g->u, u, NU);
    ORC_ASSIGN(g->v, v, NV);
    ORC_ASSIGN(g->p, p, N);
    ORC_ASSIGN(g->lmass, lm, (size_t)g->nl * N);
    ORC_ASSIGN(g->gmass, gm, (size_t)g->ng * N);
#undef ORC_ASSIGN
    memcpy(g->p_prev, g->p, sizeof(double) * N);
    orc_invalidate_caches(g);
    return 0;
}

void orc_download(const orc_grid* g, double* fraction, double* u, double* v, double* p, double* lm, double* gm)
{
    const size_t N = (size_t)g->W * g->H;
    if (fraction) memcpy(fraction, g->fraction, sizeof(double) * N);
    if (u) memcpy(u, g->u, sizeof(double) * (size_t)(g->W + 1) * g->H);
    if (v) memcpy(v, g->v, sizeof(double) * (size_t)g->W * (g->H + 1));
    if (p) memcpy(p, g->p, sizeof(double) * N);
    if (lm) memcpy(lm, g->lmass, sizeof(double) * (size_t)g->nl * N);
    if (gm) memcpy(gm, g->gmass, sizeof(double) * (size_t)g->ng * N);
}

/* rebuildDensity, FluidGridCPU.cpp:1399-1423 */
static void rebuild_density(orc_grid* g)
{
    if (g->density_valid)
        return;
    const size_t N = (size_t)g->W * g->H;
    const double inverse_cell_volume = 1.0 / (g->h * g->h);
    const int workers = batch_workers((long long)N);
#pragma omp parallel for num_threads(workers) schedule(static, 1) if (workers > 1)
    for (int b = 0; b < workers; ++b)
        for (size_t i = (size_t)batch_first((int)N, b, workers); i < (size_t)batch_first((int)N, b + 1, workers); ++i) {
            double total = 0.0;
            for (int k = 0; k < g->nl; ++k) total += g->lmass[(size_t)k * N + i];
            for (int k = 0; k < g->ng; ++k) total += g->gmass[(size_t)k * N + i];
            g->density[i] = total * inverse_cell_volume;
        }
    g->density_valid = 1;
}

/* rebuildPressureCoefficients, FluidGridCPU.cpp:1154-1205 */
static void rebuild_coefficients(orc_grid* g)
{
    if (g->coeff_valid)
        return;
    rebuild_density(g);
    const int W = g->W, H = g->H;
    const int workers = batch_workers((long long)H * W);
#pragma omp parallel for num_threads(workers) schedule(static, 1) if (workers > 1)
    for (int b = 0; b < workers; ++b)
    for (int y = batch_first(H, b, workers); y < batch_first(H, b + 1, workers); ++y) {
        const double* row = g->density + (size_t)y * W;
        double* hz = g->inv_u + (size_t)y * (W + 1);
        double* vt = g->inv_v + (size_t)y * W;
        hz[0] = hz[W] = 0.0;
        for (int x = 1; x < W; ++x)
            hz[x] = 2.0 / (row[x - 1] + row[x]);
        if (y == 0) {
            for (int x = 0; x < W; ++x) vt[x] = 0.0;
        } else {
            for (int x = 0; x < W; ++x)
                vt[x] = 2.0 / (row[x - W] + row[x]);
        }
    }
    for (int x = 0; x < W; ++x)
        g->inv_v[(size_t)H * W + x] = 0.0;
    if (g->solid) {
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                if (solid_left(g, x, y))
                    g->inv_u[(size_t)y * (W + 1) + x] = 0.0;
                if (solid_up(g, x, y))
                    g->inv_v[(size_t)y * W + x] = 0.0;
            }
    }
#pragma omp parallel for num_threads(workers) schedule(static, 1) if (workers > 1)
    for (int b = 0; b < workers; ++b)
    for (int y = batch_first(H, b, workers); y < batch_first(H, b + 1, workers); ++y) {
        const double* hz = g->inv_u + (size_t)y * (W + 1);
        const double* vt = g->inv_v + (size_t)y * W;
        double* d = g->inv_diag + (size_t)y * W;
        for (int x = 0; x < W; ++x) {
            const double sum = hz[x] + hz[x + 1] + vt[x] + vt[x + W];
            d[x] = sum > 0.0 ? 1.0 / sum : 0.0;
        }
    }
    g->coeff_valid = 1;
}

orc_mg* orc_grid_multigrid(orc_grid* g)
{
    rebuild_coefficients(g);
    if (!g->mg_valid)
        g->mg_valid = orc_mg_rebuild(g->mg, g->inv_u, g->inv_v);
    return g->mg_valid ? g->mg : NULL;
}

void orc_density(orc_grid* g, double* density)
{
    rebuild_density(g);
    memcpy(density, g->density, sizeof(double) * (size_t)g->W * g->H);
}

void orc_face_weights(orc_grid* g, double* iu, double* iv)
{
    rebuild_coefficients(g);
    if (iu) memcpy(iu, g->inv_u, sizeof(double) * (size_t)(g->W + 1) * g->H);
    if (iv) memcpy(iv, g->inv_v, sizeof(double) * (size_t)g->W * (g->H + 1));
}

/* applyGravity, FluidGridCPU.cpp:988-996: interior v rows only. */
void orc_apply_gravity(orc_grid* g, double dt)
{
    const int W = g->W, H = g->H;
    for (int i = W; i < W * (H - 1) + W; ++i)
        g->v[i] += g->gravity_y * dt;
}

/* addVortex, FluidGridCPU.cpp:998-1028 */
void orc_add_vortex(orc_grid* g, double cx, double cy, double radius, double strength)
{
    const int W = g->W, H = g->H;
    for (int y = 0; y < H; ++y) {
        for (int x = 0; x <= W; ++x) {
            const double px = g->ox + x * g->h, py = g->oy + (y + 0.5) * g->h;
            const double dx = px - cx, dy = py - cy;
            const double r2 = dx * dx + dy * dy;
            const double falloff = exp(-r2 / (2.0 * radius * radius));
            g->u[y * (W + 1) + x] += -dy * strength * falloff;
        }
    }
    for (int y = 0; y <= H; ++y) {
        for (int x = 0; x < W; ++x) {
            const double px = g->ox + (x + 0.5) * g->h, py = g->oy + y * g->h;
            const double dx = px - cx, dy = py - cy;
            const double r2 = dx * dx + dy * dy;
            const double falloff = exp(-r2 / (2.0 * radius * radius));
            g->v[y * W + x] += dx * strength * falloff;
        }
    }
}

/* maxDivergence, FluidGridCPU.cpp:1126-1152 */
double orc_max_divergence(const orc_grid* g)
{
    const int W = g->W, H = g->H;
    const int workers = batch_workers((long long)H * W);
    double part[8];
#pragma omp parallel for num_threads(workers) schedule(static, 1) if (workers > 1)
    for (int b = 0; b < workers; ++b) {
        double maximum = 0.0;
        for (int y = batch_first(H, b, workers); y < batch_first(H, b + 1, workers) && isfinite(maximum); ++y) {
            const double* hz = g->u + (size_t)y * (W + 1);
            const double* vt = g->v + (size_t)y * W;
            for (int x = 0; x < W; ++x) {
                const double divergence = ((hz[x + 1] - hz[x]) + (vt[x + W] - vt[x])) / g->h;
                if (!isfinite(divergence)) {
                    maximum = INFINITY;
                    break;
                }
                maximum = dmax(maximum, fabs(divergence));
            }
        }
        part[b] = maximum;
    }
    double maximum = part[0];
    for (int b = 1; b < workers; ++b)
        maximum = dmax(maximum, part[b]);
    return maximum;
}

/* removeMean, FluidGridCPU.cpp:1478-1486 */
static void remove_mean(double* a, size_t n)
{
    const double mean = sum_values(a, n) / (double)n;
    const int workers = batch_workers((long long)n);
#pragma omp parallel for num_threads(workers) schedule(static, 1) if (workers > 1)
    for (int b = 0; b < workers; ++b)
        for (int i = batch_first((int)n, b, workers); i < batch_first((int)n, b + 1, workers); ++i)
            a[i] -= mean;
}

/* applyPressureOperator, FluidGridCPU.cpp:1507-1598: q = A p, returns p.q in the reference's order. */
static double apply_operator(const orc_grid* g, const double* in, double* out)
{
    const int W = g->W, H = g->H;
    const double* iu = g->inv_u;
    const double* iv = g->inv_v;
#define ORC_CELL(x, y) ({ \
        const int i_ = (y) * W + (x); \
        const double c_ = in[i_]; \
        double v_ = 0.0; \
        if ((x) > 0) v_ += iu[(y) * (W + 1) + (x)] * (c_ - in[i_ - 1]); \
        if ((x) + 1 < W) v_ += iu[(y) * (W + 1) + (x) + 1] * (c_ - in[i_ + 1]); \
        if ((y) > 0) v_ += iv[i_] * (c_ - in[i_ - W]); \
        if ((y) + 1 < H) v_ += iv[i_ + W] * (c_ - in[i_ + W]); \
        out[i_] = v_; \
        c_ * v_; })
    if (W < 3 || H < 3) {
        double dot = 0.0;
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x)
                dot += ORC_CELL(x, y);
        return dot;
    }
    const int stride = W;
    /* reduceBatches(height - 2, applyRows, plus, width): rows 1 .. H-2 in batches, each with its own
     * four lanes, merged in batch order */
    const int workers = batch_workers((long long)(H - 2) * W);
    double part[8];
#pragma omp parallel for num_threads(workers) schedule(static, 1) if (workers > 1)
    for (int b = 0; b < workers; ++b) {
    double dots[4] = {0.0, 0.0, 0.0, 0.0};
    for (int y = batch_first(H - 2, b, workers) + 1; y < batch_first(H - 2, b + 1, workers) + 1; ++y) {
        const double* row = in + (size_t)y * stride;
        const double* above = row - stride;
        const double* below = row + stride;
        const double* hz = iu + (size_t)y * (stride + 1);
        const double* up = iv + (size_t)y * stride;
        const double* dn = up + stride;
        double* res = out + (size_t)y * stride;
#define ORC_APPLY(xx) ({ \
        const double c_ = row[xx]; \
        res[xx] = (hz[xx] * (c_ - row[(xx) - 1]) + hz[(xx) + 1] * (c_ - row[(xx) + 1])) + \
                  (up[xx] * (c_ - above[xx]) + dn[xx] * (c_ - below[xx])); \
        c_ * res[xx]; })
        int x = 1;
        for (; x + 3 < stride - 1; x += 4) {
            dots[0] += ORC_APPLY(x);
            dots[1] += ORC_APPLY(x + 1);
            dots[2] += ORC_APPLY(x + 2);
            dots[3] += ORC_APPLY(x + 3);
        }
        for (; x < stride - 1; ++x)
            dots[0] += ORC_APPLY(x);
#undef ORC_APPLY
    }
    part[b] = (dots[0] + dots[1]) + (dots[2] + dots[3]);
    }
    double dot = part[0];
    for (int b = 1; b < workers; ++b)
        dot = dot + part[b];
    for (int x = 0; x < W; ++x) {
        dot += ORC_CELL(x, 0);
        dot += ORC_CELL(x, H - 1);
    }
    for (int y = 1; y < H - 1; ++y) {
        dot += ORC_CELL(0, y);
        dot += ORC_CELL(W - 1, y);
    }
#undef ORC_CELL
    return dot;
}

double orc_apply_operator(orc_grid* g, const double* p, double* q)
{
    rebuild_coefficients(g);
    return apply_operator(g, p, q);
}

typedef struct { double sum, maximum; } orc_stats;

/* 4-lane (sum, max) accumulation over [FIRST, LAST) shared by updatePressureAndResidual and
 * rebuildPressureResidual (one batch of their reduceBatches). */
#define ORC_LANE_STATS(FIRST, LAST, BODY)                                                \
    do {                                                                                 \
        double sums_[4] = {0, 0, 0, 0}, maxs_[4] = {0, 0, 0, 0};                         \
        size_t i = (FIRST);                                                              \
        for (; i + 3 < (LAST); i += 4) {                                                 \
            for (int lane_ = 0; lane_ < 4; ++lane_) {                                    \
                const size_t j = i + lane_;                                              \
                const double value = (BODY);                                             \
                sums_[lane_] += value;                                                   \
                maxs_[lane_] = dmax(maxs_[lane_], fabs(value));                          \
            }                                                                            \
        }                                                                                \
        st.sum = (sums_[0] + sums_[1]) + (sums_[2] + sums_[3]);                          \
        st.maximum = dmax(dmax(maxs_[0], maxs_[1]), dmax(maxs_[2], maxs_[3]));           \
        for (; i < (LAST); ++i) {                                                        \
            const size_t j = i;                                                          \
            const double value = (BODY);                                                 \
            st.sum += value;                                                             \
            st.maximum = dmax(st.maximum, fabs(value));                                  \
        }                                                                                \
    } while (0)

/* updatePressureAndResidual, FluidGridCPU.cpp:1600-1641 */
static orc_stats update_pressure_residual(orc_grid* g, double alpha)
{
    const size_t N = (size_t)g->W * g->H;
    orc_stats st;
    double* p = g->p;
    double* r = g->res;
    const double* d = g->dir;
    const double* q = g->prod;
    const int workers = batch_workers((long long)N);
    orc_stats part[8];
#pragma omp parallel for num_threads(workers) schedule(static, 1) if (workers > 1)
    for (int b = 0; b < workers; ++b) {
        orc_stats st;
        ORC_LANE_STATS((size_t)batch_first((int)N, b, workers), (size_t)batch_first((int)N, b + 1, workers),
            (p[j] += alpha * d[j], r[j] -= alpha * q[j]));
        part[b] = st;
    }
    st = part[0];
    for (int b = 1; b < workers; ++b) {
        st.sum = st.sum + part[b].sum;
        st.maximum = dmax(st.maximum, part[b].maximum);
    }
    if (!isfinite(st.sum))
        st.maximum = INFINITY;
    return st;
}

/* rebuildPressureResidual, FluidGridCPU.cpp:1643-1683 */
static orc_stats rebuild_residual(orc_grid* g)
{
    const size_t N = (size_t)g->W * g->H;
    apply_operator(g, g->p, g->prod);
    orc_stats st;
    double* r = g->res;
    const double* b = g->rhs;
    const double* q = g->prod;
    const int workers = batch_workers((long long)N);
    orc_stats part[8];
#pragma omp parallel for num_threads(workers) schedule(static, 1) if (workers > 1)
    for (int k = 0; k < workers; ++k) {
        orc_stats st;
        ORC_LANE_STATS((size_t)batch_first((int)N, k, workers), (size_t)batch_first((int)N, k + 1, workers),
            (r[j] = b[j] - q[j]));
        part[k] = st;
    }
    st = part[0];
    for (int k = 1; k < workers; ++k) {
        st.sum = st.sum + part[k].sum;
        st.maximum = dmax(st.maximum, part[k].maximum);
    }
    if (!isfinite(st.sum))
        st.maximum = INFINITY;
    return st;
}

/* applyPreconditioner, FluidGridCPU.cpp:1685-1726 */
static double apply_preconditioner(orc_grid* g, double residual_mean)
{
    const size_t N = (size_t)g->W * g->H;
    const int workers = batch_workers((long long)N);
    if (g->preconditioner == 1)
        orc_mg_apply(g->mg, g->res, g->cor, residual_mean);
    else {
#pragma omp parallel for num_threads(workers) schedule(static, 1) if (workers > 1)
        for (int b = 0; b < workers; ++b)
            for (int i = batch_first((int)N, b, workers); i < batch_first((int)N, b + 1, workers); ++i)
                g->cor[i] = (g->res[i] - residual_mean) * g->inv_diag[i];
    }
    const double mean = sum_values(g->cor, N) / (double)N;
    double part[8];
#pragma omp parallel for num_threads(workers) schedule(static, 1) if (workers > 1)
    for (int b = 0; b < workers; ++b) {
        const size_t last = (size_t)batch_first((int)N, b + 1, workers);
        double dots[4] = {0, 0, 0, 0};
        size_t i = (size_t)batch_first((int)N, b, workers);
        for (; i + 3 < last; i += 4) {
            for (int lane = 0; lane < 4; ++lane) {
                g->cor[i + lane] -= mean;
                dots[lane] += g->res[i + lane] * g->cor[i + lane];
            }
        }
        double dot = (dots[0] + dots[1]) + (dots[2] + dots[3]);
        for (; i < last; ++i) {
            g->cor[i] -= mean;
            dot += g->res[i] * g->cor[i];
        }
        part[b] = dot;
    }
    double dot = part[0];
    for (int b = 1; b < workers; ++b)
        dot = dot + part[b];
    return dot;
}

/* solvePressure, FluidGridCPU.cpp:1728-1824 */
static int solve_pressure(orc_grid* g, int max_iterations, double tol, orc_result* info)
{
    const size_t N = (size_t)g->W * g->H;
    if (max_abs(g->rhs, N) <= tol) {
        memset(g->p, 0, sizeof(double) * N);
        memset(g->p_prev, 0, sizeof(double) * N);
        return 0;
    }
    const int workers = batch_workers((long long)N);
#pragma omp parallel for num_threads(workers) schedule(static, 1) if (workers > 1)
    for (int b = 0; b < workers; ++b)
        for (int i = batch_first((int)N, b, workers); i < batch_first((int)N, b + 1, workers); ++i) {
            const double latest = g->p[i];
            g->p[i] = 2.0 * latest - g->p_prev[i];
            g->p_prev[i] = latest;
        }
    remove_mean(g->p, N);
    const orc_stats initial = rebuild_residual(g);
    info->initial_residual = initial.maximum;
    if (initial.maximum <= tol || max_iterations == 0)
        return 0;
    if (g->preconditioner == 1 && !g->mg_valid) {
        g->mg_valid = orc_mg_rebuild(g->mg, g->inv_u, g->inv_v);
        if (!g->mg_valid)
            return -1;
    }
    double residual_mean = initial.sum / (double)N;
    double rz = apply_preconditioner(g, residual_mean);
    memcpy(g->dir, g->cor, sizeof(double) * N);
    int iterations = 0;
    while (iterations < max_iterations && rz > 0.0 && isfinite(rz)) {
        const double denominator = apply_operator(g, g->dir, g->prod);
        if (denominator <= 0.0 || !isfinite(denominator))
            break;
        const double alpha = rz / denominator;
        if (!isfinite(alpha))
            break;
        const orc_stats st = update_pressure_residual(g, alpha);
        residual_mean = st.sum / (double)N;
        ++iterations;
        int restart = 0;
        if (st.maximum <= tol) {
            const orc_stats actual = rebuild_residual(g);
            if (actual.maximum <= tol)
                break;
            residual_mean = actual.sum / (double)N;
            restart = 1;
            info->restarts++;
        }
        const double next = apply_preconditioner(g, residual_mean);
        const double beta = restart ? 0.0 : next / rz;
        if (!isfinite(beta))
            break;
#pragma omp parallel for num_threads(workers) schedule(static, 1) if (workers > 1)
        for (int b = 0; b < workers; ++b)
            for (int i = batch_first((int)N, b, workers); i < batch_first((int)N, b + 1, workers); ++i)
                g->dir[i] = g->cor[i] + beta * g->dir[i];
        rz = next;
    }
    remove_mean(g->p, N);
    return iterations;
}

/* Form rhs = -div * h^2/dt and remove its mean (FluidGridCPU.cpp:1063-1090). */
static void form_rhs(orc_grid* g, double dt)
{
    const int W = g->W, H = g->H;
    const size_t N = (size_t)W * H;
    const double scale = g->h * g->h / dt;
    /* reduceBatches(height, ..., width): a running sum per batch of rows, merged in batch order */
    int workers = batch_workers((long long)H * W);
    double part[8];
#pragma omp parallel for num_threads(workers) schedule(static, 1) if (workers > 1)
    for (int b = 0; b < workers; ++b) {
        double sum = 0.0;
        for (int y = batch_first(H, b, workers); y < batch_first(H, b + 1, workers); ++y) {
            const double* hz = g->u + (size_t)y * (W + 1);
            const double* vt = g->v + (size_t)y * W;
            double* rhs = g->rhs + (size_t)y * W;
            for (int x = 0; x < W; ++x) {
                const double du = hz[x + 1] - hz[x];
                const double dv = vt[x + W] - vt[x];
                rhs[x] = -(du + dv) / g->h * scale;
                sum += rhs[x];
            }
        }
        part[b] = sum;
    }
    double sum = part[0];
    for (int b = 1; b < workers; ++b)
        sum = sum + part[b];
    const double mean = sum / (double)N;
    workers = batch_workers((long long)N);
#pragma omp parallel for num_threads(workers) schedule(static, 1) if (workers > 1)
    for (int b = 0; b < workers; ++b)
        for (int i = batch_first((int)N, b, workers); i < batch_first((int)N, b + 1, workers); ++i)
            g->rhs[i] -= mean;
}

void orc_pressure_rhs(orc_grid* g, double dt, double* rhs)
{
    form_rhs(g, dt);
    memcpy(rhs, g->rhs, sizeof(double) * (size_t)g->W * g->H);
}

/* project, FluidGridCPU.cpp:1030-1124 */
orc_result orc_project(orc_grid* g, double dt, int max_iterations, double tol)
{
    orc_result result;
    memset(&result, 0, sizeof(result));
    result.max_divergence = INFINITY;
    result.physical_steps = 1;
    const int W = g->W, H = g->H;
    if (dt <= 0.0 || W <= 0 || H <= 0 || max_iterations < 0)
        return result;
    const double scale = g->h * g->h / dt;
    rebuild_coefficients(g);
    for (int y = 0; y < H; ++y) {
        g->u[y * (W + 1)] = 0.0;
        g->u[y * (W + 1) + W] = 0.0;
    }
    for (int x = 0; x < W; ++x) {
        g->v[x] = 0.0;
        g->v[H * W + x] = 0.0;
    }
    if (g->solid) {
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                if (solid_left(g, x, y))
                    g->u[(size_t)y * (W + 1) + x] = 0.0;
                if (solid_up(g, x, y))
                    g->v[(size_t)y * W + x] = 0.0;
            }
    }
    form_rhs(g, dt);
    if (g->relative_tol) {
        /* extension (no reference counterpart): stop at max|b - A p| <= tol max|b| and accept the
         * divergence that residual bound allows */
        double bmax = 0.0;
        for (size_t i = 0; i < (size_t)W * H; ++i)
            bmax = dmax(bmax, fabs(g->rhs[i]));
        tol = tol * bmax / scale;
    }
    result.iterations = solve_pressure(g, max_iterations, tol * scale, &result);
    if (result.iterations < 0) {
        orc_result failed;
        memset(&failed, 0, sizeof(failed));
        failed.max_divergence = INFINITY;
        failed.physical_steps = 1;
        return failed;
    }
    const int workers = batch_workers((long long)H * W);
#pragma omp parallel for num_threads(workers) schedule(static, 1) if (workers > 1)
    for (int b = 0; b < workers; ++b)
    for (int y = batch_first(H, b, workers); y < batch_first(H, b + 1, workers); ++y) {
        const double* row = g->p + (size_t)y * W;
        double* hz = g->u + (size_t)y * (W + 1);
        const double* hw = g->inv_u + (size_t)y * (W + 1);
        for (int x = 1; x < W; ++x)
            hz[x] -= ((row[x] - row[x - 1]) / g->h) * dt * hw[x];
        if (y == 0)
            continue;
        double* vt = g->v + (size_t)y * W;
        const double* vw = g->inv_v + (size_t)y * W;
        for (int x = 0; x < W; ++x)
            vt[x] -= ((row[x] - row[x - W]) / g->h) * dt * vw[x];
    }
    /* diagnostic: true residual of the returned pressure */
    {
        const size_t N = (size_t)W * H;
        double* q = (double*)malloc(sizeof(double) * N);
        apply_operator(g, g->p, q);
        double m = 0.0;
        for (size_t i = 0; i < N; ++i)
            m = dmax(m, fabs(g->rhs[i] - q[i]));
        result.final_residual = m;
        free(q);
    }
    result.max_divergence = orc_max_divergence(g);
    result.converged = result.max_divergence <= tol;
    return result;
}

void orc_set_relative_tolerance(orc_grid* g, int enabled) { g->relative_tol = enabled != 0; }

orc_result orc_force_project(orc_grid* g, double dt, int max_iterations, double tol)
{
    if (dt <= 0.0) {
        orc_result r;
        memset(&r, 0, sizeof(r));
        r.max_divergence = INFINITY;
        return r;
    }
    orc_apply_gravity(g, dt);
    return orc_project(g, dt, max_iterations, tol);
}

/* sampleVelocity, FluidGridCPU.cpp:465-471 */
void orc_sample_velocity(const orc_grid* g, double px, double py, double* vx, double* vy)
{
    const double x = (px - g->ox) / g->h;
    const double y = (py - g->oy) / g->h;
    *vx = sample_scalar(g->u, g->W + 1, g->H, x, y - 0.5);
    *vy = sample_scalar(g->v, g->W, g->H + 1, x - 0.5, y);
}

/*
 * Semi-Lagrangian transport mode.  The reference has no such step (its transport is
 * PLIC/VOF); this is the build's own definition, built only from the reference's staggered
 * sampler (sampleVelocity/sampleScalar).  Midpoint (RK2) backtrace in index space:
 *   a = P - 0.5*dt*vel(P);  src = P - dt*vel(a);  new = sample(old field, src)
 * for every interior u and v face and every cell-centred scalar (fraction, masses).
 * Wall faces stay zero.
 */
static void backtrace(const orc_grid* g, double px, double py, double dt, double* sx, double* sy)
{
    double vx, vy;
    orc_sample_velocity(g, px, py, &vx, &vy);
    const double mx = px - 0.5 * dt * vx, my = py - 0.5 * dt * vy;
    orc_sample_velocity(g, mx, my, &vx, &vy);
    *sx = px - dt * vx;
    *sy = py - dt * vy;
}

void orc_advect_semi_lagrangian(orc_grid* g, double dt)
{
    const int W = g->W, H = g->H;
    const size_t N = (size_t)W * H;
    const size_t NU = (size_t)(W + 1) * H, NV = (size_t)W * (H + 1);
    double* nu = dalloc(NU);
    double* nv = dalloc(NV);
    double* tmp = dalloc(N);
    const double inv_h = 1.0 / g->h;
    (void)inv_h;
    for (int y = 0; y < H; ++y) {
        for (int x = 1; x < W; ++x) {
            double sx, sy;
            backtrace(g, g->ox + x * g->h, g->oy + (y + 0.5) * g->h, dt, &sx, &sy);
            nu[y * (W + 1) + x] = sample_scalar(g->u, W + 1, H, (sx - g->ox) / g->h, (sy - g->oy) / g->h - 0.5);
        }
    }
    for (int y = 1; y < H; ++y) {
        for (int x = 0; x < W; ++x) {
            double sx, sy;
            backtrace(g, g->ox + (x + 0.5) * g->h, g->oy + y * g->h, dt, &sx, &sy);
            nv[y * W + x] = sample_scalar(g->v, W, H + 1, (sx - g->ox) / g->h - 0.5, (sy - g->oy) / g->h);
        }
    }
    /* cell-centred scalars share one backtrace per cell */
    double* srcx = dalloc(N);
    double* srcy = dalloc(N);
    for (int y = 0; y < H; ++y) {
        for (int x = 0; x < W; ++x) {
            double sx, sy;
            backtrace(g, g->ox + (x + 0.5) * g->h, g->oy + (y + 0.5) * g->h, dt, &sx, &sy);
            srcx[y * W + x] = (sx - g->ox) / g->h - 0.5;
            srcy[y * W + x] = (sy - g->oy) / g->h - 0.5;
        }
    }
    const int fields = 1 + g->nl + g->ng;
    for (int f = 0; f < fields; ++f) {
        double* a = f == 0 ? g->fraction : (f <= g->nl ? g->lmass + (size_t)(f - 1) * N
                                                        : g->gmass + (size_t)(f - 1 - g->nl) * N);
        for (size_t i = 0; i < N; ++i)
            tmp[i] = sample_scalar(a, W, H, srcx[i], srcy[i]);
        memcpy(a, tmp, sizeof(double) * N);
    }
    memcpy(g->u, nu, sizeof(double) * NU);
    memcpy(g->v, nv, sizeof(double) * NV);
    free(nu); free(nv); free(tmp); free(srcx); free(srcy);
    orc_invalidate_caches(g);
}

orc_result orc_step_semi_lagrangian(orc_grid* g, double dt, int max_iterations, double tol)
{
    if (dt <= 0.0) {
        orc_result r;
        memset(&r, 0, sizeof(r));
        r.max_divergence = INFINITY;
        return r;
    }
    orc_advect_semi_lagrangian(g, dt);
    orc_result r = orc_force_project(g, dt, max_iterations, tol);
    r.transport_steps = 1;
    return r;
}

/* ------------------------------------------------------------------------------------------ */
/* Reference transport: CFL-split X/Y PLIC volume transfers with MYC normals, component masses  */
/* and consistent momentum (FluidGridCPU.cpp:93-167, 362-463, 473-833, 1207-1476).             */
/* ------------------------------------------------------------------------------------------ */

static double dmin(double a, double b) { return b < a ? b : a; }   /* std::min(a, b) */

/* liquidArea, FluidGridCPU.cpp:93-133: liquid area of the half plane n.p <= offset inside a
 * rectangle of the unit cell. */
static double liquid_area(double nx, double ny, double offset, double x0, double y0, double x1, double y1)
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

/* interfaceOffset, FluidGridCPU.cpp:135-167: line offset that cuts `fraction` of the unit cell. */
static double interface_offset(double nx, double ny, double fraction)
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

/* interfaceNormal, FluidGridCPU.cpp:473-521: mixed Youngs-centred (MYC) normal. */
static void interface_normal(const double* f, int W, int H, int x, int y, double* nx, double* ny)
{
    const int lx = imax(x - 1, 0), rx = imin(x + 1, W - 1);
    const double* above = f + (size_t)imax(y - 1, 0) * W;
    const double* row = f + (size_t)y * W;
    const double* below = f + (size_t)imin(y + 1, H - 1) * W;
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
        if (yx == 0.0 && yy == 0.0) { *nx = 1.0; *ny = 0.0; }
        else { *nx = yx; *ny = yy; }
        return;
    }
    if (fabs(cx) <= fabs(cy)) {
        cy = copysign(1.0, cy);
        if (fabs(yx) > fabs(cx * yy)) { *nx = yx; *ny = yy; return; }
    } else {
        cx = copysign(1.0, cx);
        if (fabs(yy) > fabs(cy * yx)) { *nx = yx; *ny = yy; return; }
    }
    *nx = cx;
    *ny = cy;
}

typedef struct {
    double *next_fraction, *next_u, *next_v;
    double *nx, *ny, *offset;
    unsigned char* split;
    double *mcw, *lcw, *gcw;
    double *lt_u, *lt_v, *mt_u, *mt_v, *mom_u, *mom_v, *ct;
} orc_transport;

static void transport_alloc(const orc_grid* g, orc_transport* T)
{
    const int W = g->W, H = g->H;
    const size_t N = (size_t)W * H, NU = (size_t)(W + 1) * H, NV = (size_t)W * (H + 1);
    T->next_fraction = dalloc(N);
    T->next_u = dalloc(NU);
    T->next_v = dalloc(NV);
    T->nx = dalloc(N);
    T->ny = dalloc(N);
    T->offset = dalloc(N);
    T->split = (unsigned char*)calloc(N, 1);
    T->mcw = dalloc(N);
    T->lcw = dalloc((size_t)g->nl * N);
    T->gcw = dalloc((size_t)g->ng * N);
    T->lt_u = dalloc(NU);
    T->lt_v = dalloc(NV);
    T->mt_u = dalloc(NU);
    T->mt_v = dalloc(NV);
    T->mom_u = dalloc(NU);
    T->mom_v = dalloc(NV);
    T->ct = dalloc(NU > NV ? NU : NV);
}

static void transport_free(orc_transport* T)
{
    free(T->next_fraction); free(T->next_u); free(T->next_v);
    free(T->nx); free(T->ny); free(T->offset); free(T->split);
    free(T->mcw); free(T->lcw); free(T->gcw);
    free(T->lt_u); free(T->lt_v); free(T->mt_u); free(T->mt_v);
    free(T->mom_u); free(T->mom_v); free(T->ct);
}

/* rebuildInterfaces, FluidGridCPU.cpp:523-541: only interface cells are rewritten. */
static void rebuild_interfaces(const orc_grid* g, orc_transport* T)
{
    const int W = g->W, H = g->H;
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const size_t i = (size_t)y * W + x;
            const double fraction = g->fraction[i];
            if (fraction <= 0.0 || fraction >= 1.0)
                continue;
            interface_normal(g->fraction, W, H, x, y, &T->nx[i], &T->ny[i]);
            T->offset[i] = interface_offset(T->nx[i], T->ny[i], fraction);
        }
}

/* liquidTransferX / liquidTransferY, FluidGridCPU.cpp:759-833.  `horizontal` selects the axis;
 * `speed` is the face velocity, `donor` the upwind cell. */
static double liquid_transfer(const orc_grid* g, const orc_transport* T, int horizontal, double speed,
    size_t donor, double step)
{
    if (speed == 0.0)
        return 0.0;
    const double fraction = g->fraction[donor];
    if (fraction <= 0.0)
        return 0.0;
    const double volume_transfer = speed * (step / g->h);
    const double strip = fabs(volume_transfer);
    if (fraction >= 1.0)
        return volume_transfer;
    double x0, y0, x1, y1;
    if (horizontal) {
        x0 = speed > 0.0 ? 1.0 - strip : 0.0;
        x1 = speed > 0.0 ? 1.0 : strip;
        y0 = 0.0;
        y1 = 1.0;
    } else {
        x0 = 0.0;
        x1 = 1.0;
        y0 = speed > 0.0 ? 1.0 - strip : 0.0;
        y1 = speed > 0.0 ? 1.0 : strip;
    }
    const double area = liquid_area(T->nx[donor], T->ny[donor], T->offset[donor], x0, y0, x1, y1);
    const double bounded = dclamp(area, dmax(0.0, strip - (1.0 - fraction)), dmin(strip, fraction));
    return speed > 0.0 ? bounded : -bounded;
}

/* rebuildTransfersX / rebuildTransfersY, FluidGridCPU.cpp:1207-1233 */
static void rebuild_transfers(const orc_grid* g, orc_transport* T, int horizontal, double step)
{
    const int W = g->W, H = g->H;
    if (horizontal) {
        for (int y = 0; y < H; ++y)
            for (int x = 0; x <= W; ++x) {
                const size_t face = (size_t)y * (W + 1) + x;
                double t = 0.0;
                if (x != 0 && x != W) {
                    const double speed = g->u[face];
                    t = liquid_transfer(g, T, 1, speed, (size_t)y * W + (speed > 0.0 ? x - 1 : x), step);
                }
                T->lt_u[face] = t;
            }
    } else {
        for (int y = 0; y <= H; ++y)
            for (int x = 0; x < W; ++x) {
                const size_t face = (size_t)y * W + x;
                double t = 0.0;
                if (y != 0 && y != H) {
                    const double speed = g->v[face];
                    t = liquid_transfer(g, T, 0, speed, (size_t)(speed > 0.0 ? y - 1 : y) * W + x, step);
                }
                T->lt_v[face] = t;
            }
    }
}

/* freezeComponentCorrections, FluidGridCPU.cpp:1307-1349 */
static void freeze_corrections(const orc_grid* g, orc_transport* T)
{
    const size_t N = (size_t)g->W * g->H;
    const double inverse_cell_volume = 1.0 / (g->h * g->h);
    for (size_t i = 0; i < N; ++i) {
        const double fraction = g->fraction[i];
        const int selected = fraction > 0.5;
        T->split[i] = (unsigned char)selected;
        const double inverse_fraction = 1.0 / (selected ? fraction : 1.0 - fraction);
        T->mcw[i] = 0.0;
        for (int k = 0; k < g->nl; ++k) {
            const double weight = selected ? g->lmass[(size_t)k * N + i] * inverse_fraction : 0.0;
            T->lcw[(size_t)k * N + i] = weight;
            T->mcw[i] += weight * inverse_cell_volume;
        }
        for (int k = 0; k < g->ng; ++k) {
            const double weight = !selected ? g->gmass[(size_t)k * N + i] * inverse_fraction : 0.0;
            T->gcw[(size_t)k * N + i] = weight;
            T->mcw[i] += weight * inverse_cell_volume;
        }
    }
}

/* componentMassTransfer, FluidGridCPU.cpp:543-551 */
static double component_mass_transfer(double volume_transfer, double donor_fraction, double donor_mass)
{
    if (donor_fraction <= 0.0)
        return 0.0;
    return donor_mass * (volume_transfer / donor_fraction);
}

/* advectComponents + rebuildComponentTransfersX/Y, FluidGridCPU.cpp:1235-1305, 1351-1397 */
static void advect_components(orc_grid* g, orc_transport* T, double step, int horizontal)
{
    const int W = g->W, H = g->H;
    const size_t N = (size_t)W * H;
    const double scale = step / g->h;
    const double inverse_cell_volume = 1.0 / (g->h * g->h);
    const double* flow = horizontal ? g->u : g->v;
    const size_t face_offset = horizontal ? 1 : (size_t)W;
    double* total = horizontal ? T->mt_u : T->mt_v;
    memset(total, 0, sizeof(double) * (horizontal ? (size_t)(W + 1) * H : (size_t)W * (H + 1)));
    const double* lt = horizontal ? T->lt_u : T->lt_v;
    for (int phase = 0; phase < 2; ++phase) {
        const int is_liquid = phase == 0;
        const int count = is_liquid ? g->nl : g->ng;
        for (int k = 0; k < count; ++k) {
            double* mass = (is_liquid ? g->lmass : g->gmass) + (size_t)k * N;
            const double* weights = (is_liquid ? T->lcw : T->gcw) + (size_t)k * N;
            if (horizontal) {
                for (int y = 0; y < H; ++y)
                    for (int x = 0; x <= W; ++x) {
                        const size_t face = (size_t)y * (W + 1) + x;
                        if (x == 0 || x == W) { T->ct[face] = 0.0; continue; }
                        const size_t donor = (size_t)y * W + (g->u[face] > 0.0 ? x - 1 : x);
                        const double all = g->u[face] * scale;
                        const double liquid = lt[face];
                        const double fraction = g->fraction[donor];
                        T->ct[face] = component_mass_transfer(is_liquid ? liquid : all - liquid,
                            is_liquid ? fraction : 1.0 - fraction, mass[donor]);
                        total[face] += T->ct[face] * inverse_cell_volume;
                    }
            } else {
                for (int y = 0; y <= H; ++y)
                    for (int x = 0; x < W; ++x) {
                        const size_t face = (size_t)y * W + x;
                        if (y == 0 || y == H) { T->ct[face] = 0.0; continue; }
                        const size_t donor = g->v[face] > 0.0 ? face - W : face;
                        const double all = g->v[face] * scale;
                        const double liquid = lt[face];
                        const double fraction = g->fraction[donor];
                        T->ct[face] = component_mass_transfer(is_liquid ? liquid : all - liquid,
                            is_liquid ? fraction : 1.0 - fraction, mass[donor]);
                        total[face] += T->ct[face] * inverse_cell_volume;
                    }
            }
            for (int y = 0; y < H; ++y)
                for (int x = 0; x < W; ++x) {
                    const size_t i = (size_t)y * W + x;
                    const size_t before = horizontal ? (size_t)y * (W + 1) + x : i;
                    const size_t after = before + face_offset;
                    const double net = scale * (flow[after] - flow[before]);
                    mass[i] += T->ct[before] - T->ct[after] + weights[i] * net;
                }
        }
    }
    orc_invalidate_caches(g);
}

/* momentumTransfer, FluidGridCPU.cpp:553-558 */
static double momentum_transfer(double mass_transfer, double before, double after)
{
    return mass_transfer * (mass_transfer >= 0.0 ? before : after);
}

/* advectMomentumX, FluidGridCPU.cpp:560-634 */
static void advect_momentum_x(orc_grid* g, orc_transport* T, double step)
{
    const int W = g->W, H = g->H;
    const double scale = step / g->h;
    const double* flow = g->u;
#define ORC_MC(cx, cy) (T->mcw[(size_t)(cy) * W + (cx)] * scale * \
        (flow[(size_t)(cy) * (W + 1) + (cx) + 1] - flow[(size_t)(cy) * (W + 1) + (cx)]))
    for (int y = 0; y < H; ++y)
        for (int x = 0; x <= W; ++x) {
            const size_t face = (size_t)y * (W + 1) + x;
            if (x == 0 || x == W) { T->mom_u[face] = 0.0; continue; }
            const double lm = 0.5 * (T->mt_u[face - 1] + T->mt_u[face]);
            const double rm = 0.5 * (T->mt_u[face] + T->mt_u[face + 1]);
            const double lmom = momentum_transfer(lm, T->next_u[face - 1], T->next_u[face]);
            const double rmom = momentum_transfer(rm, T->next_u[face], T->next_u[face + 1]);
            const double correction = 0.5 * (ORC_MC(x - 1, y) + ORC_MC(x, y));
            T->mom_u[face] += lmom - rmom + g->u[face] * correction;
        }
    for (int y = 0; y <= H; ++y)
        for (int x = 0; x < W; ++x) {
            const size_t face = (size_t)y * W + x;
            if (y == 0 || y == H) { T->mom_v[face] = 0.0; continue; }
            const size_t above_left = (size_t)(y - 1) * (W + 1) + x;
            const size_t below_left = (size_t)y * (W + 1) + x;
            const double lm = 0.5 * (T->mt_u[above_left] + T->mt_u[below_left]);
            const double rm = 0.5 * (T->mt_u[above_left + 1] + T->mt_u[below_left + 1]);
            const double lmom = x > 0 ? momentum_transfer(lm, T->next_v[face - 1], T->next_v[face]) : 0.0;
            const double rmom = x + 1 < W ? momentum_transfer(rm, T->next_v[face], T->next_v[face + 1]) : 0.0;
            const double correction = 0.5 * (ORC_MC(x, y - 1) + ORC_MC(x, y));
            T->mom_v[face] += lmom - rmom + g->v[face] * correction;
        }
#undef ORC_MC
}

/* advectMomentumY, FluidGridCPU.cpp:636-712 */
static void advect_momentum_y(orc_grid* g, orc_transport* T, double step)
{
    const int W = g->W, H = g->H;
    const double scale = step / g->h;
    const double* flow = g->v;
    const size_t us = (size_t)W + 1;
#define ORC_MC(cx, cy) (T->mcw[(size_t)(cy) * W + (cx)] * scale * \
        (flow[(size_t)(cy) * W + (cx) + W] - flow[(size_t)(cy) * W + (cx)]))
    for (int y = 0; y < H; ++y)
        for (int x = 0; x <= W; ++x) {
            const size_t face = (size_t)y * us + x;
            if (x == 0 || x == W) { T->mom_u[face] = 0.0; continue; }
            const size_t left_top = (size_t)y * W + x - 1;
            const size_t right_top = (size_t)y * W + x;
            const double tm = 0.5 * (T->mt_v[left_top] + T->mt_v[right_top]);
            const double bm = 0.5 * (T->mt_v[left_top + W] + T->mt_v[right_top + W]);
            const double tmom = y > 0 ? momentum_transfer(tm, T->next_u[face - us], T->next_u[face]) : 0.0;
            const double bmom = y + 1 < H ? momentum_transfer(bm, T->next_u[face], T->next_u[face + us]) : 0.0;
            const double correction = 0.5 * (ORC_MC(x - 1, y) + ORC_MC(x, y));
            T->mom_u[face] += tmom - bmom + g->u[face] * correction;
        }
    for (int y = 0; y <= H; ++y)
        for (int x = 0; x < W; ++x) {
            const size_t face = (size_t)y * W + x;
            if (y == 0 || y == H) { T->mom_v[face] = 0.0; continue; }
            const double tm = 0.5 * (T->mt_v[face - W] + T->mt_v[face]);
            const double bm = 0.5 * (T->mt_v[face] + T->mt_v[face + W]);
            const double tmom = momentum_transfer(tm, T->next_v[face - W], T->next_v[face]);
            const double bmom = momentum_transfer(bm, T->next_v[face], T->next_v[face + W]);
            const double correction = 0.5 * (ORC_MC(x, y - 1) + ORC_MC(x, y));
            T->mom_v[face] += tmom - bmom + g->v[face] * correction;
        }
#undef ORC_MC
}

/* rebuildMomentum, FluidGridCPU.cpp:1425-1476 */
static void rebuild_momentum(orc_grid* g, orc_transport* T)
{
    const int W = g->W, H = g->H;
    rebuild_density(g);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x <= W; ++x) {
            const size_t face = (size_t)y * (W + 1) + x;
            if (x == 0 || x == W) { T->mom_u[face] = 0.0; T->next_u[face] = 0.0; continue; }
            const size_t cell = (size_t)y * W + x;
            const double face_density = 0.5 * (g->density[cell - 1] + g->density[cell]);
            T->mom_u[face] = face_density * g->u[face];
            T->next_u[face] = g->u[face];
        }
    for (int y = 0; y <= H; ++y)
        for (int x = 0; x < W; ++x) {
            const size_t face = (size_t)y * W + x;
            if (y == 0 || y == H) { T->mom_v[face] = 0.0; T->next_v[face] = 0.0; continue; }
            const double face_density = 0.5 * (g->density[face - W] + g->density[face]);
            T->mom_v[face] = face_density * g->v[face];
            T->next_v[face] = g->v[face];
        }
}

/* recoverVelocity, FluidGridCPU.cpp:714-757 */
static void recover_velocity(orc_grid* g, orc_transport* T)
{
    const int W = g->W, H = g->H;
    rebuild_density(g);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x <= W; ++x) {
            const size_t face = (size_t)y * (W + 1) + x;
            if (x == 0 || x == W) { T->next_u[face] = 0.0; continue; }
            const size_t cell = (size_t)y * W + x;
            const double face_density = 0.5 * (g->density[cell - 1] + g->density[cell]);
            T->next_u[face] = T->mom_u[face] / face_density;
        }
    for (int y = 0; y <= H; ++y)
        for (int x = 0; x < W; ++x) {
            const size_t face = (size_t)y * W + x;
            if (y == 0 || y == H) { T->next_v[face] = 0.0; continue; }
            const double face_density = 0.5 * (g->density[face - W] + g->density[face]);
            T->next_v[face] = T->mom_v[face] / face_density;
        }
}

/* New fractions after one directional pass (FluidGridCPU.cpp:407-428, 437-455). */
static void update_fraction(orc_grid* g, orc_transport* T, double scale, int horizontal)
{
    const int W = g->W, H = g->H;
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const size_t i = (size_t)y * W + x;
            double net_liquid, net_fluid;
            if (horizontal) {
                const size_t lf = (size_t)y * (W + 1) + x;
                net_liquid = T->lt_u[lf + 1] - T->lt_u[lf];
                net_fluid = scale * (g->u[lf + 1] - g->u[lf]);
            } else {
                net_liquid = T->lt_v[i + W] - T->lt_v[i];
                net_fluid = scale * (g->v[i + W] - g->v[i]);
            }
            T->next_fraction[i] = g->fraction[i] - net_liquid + (double)T->split[i] * net_fluid;
        }
    double* swap = g->fraction;
    g->fraction = T->next_fraction;
    T->next_fraction = swap;
}

/* Transport CFL limit, FluidGridCPU.cpp:366-389: half a cell of outflow per substep. */
double orc_transport_max_step(const orc_grid* g, double dt)
{
    const int W = g->W, H = g->H;
    double maximum = 0.0;
    for (int y = 0; y < H; ++y) {
        const double* hz = g->u + (size_t)y * (W + 1);
        const double* vt = g->v + (size_t)y * W;
        for (int x = 0; x < W; ++x) {
            const double outflow = dmax(hz[x + 1], 0.0) + dmax(-hz[x], 0.0) + dmax(vt[x + W], 0.0) +
                                   dmax(-vt[x], 0.0);
            maximum = dmax(maximum, outflow);
        }
    }
    return maximum > 0.0 ? 0.5 * g->h / maximum : dt;
}

/* advect, FluidGridCPU.cpp:362-463.  Returns the number of transport substeps. */
int orc_advect_reference(orc_grid* g, double dt)
{
    orc_transport T;
    transport_alloc(g, &T);
    rebuild_momentum(g, &T);
    const double max_step = orc_transport_max_step(g, dt);
    double remaining = dt;
    int substeps = 0;
    while (remaining > 0.0) {
        const double step = dmin(remaining, max_step);
        const double scale = step / g->h;
        freeze_corrections(g, &T);

        rebuild_interfaces(g, &T);
        rebuild_transfers(g, &T, 1, step);
        advect_components(g, &T, step, 1);
        advect_momentum_x(g, &T, step);
        update_fraction(g, &T, scale, 1);
        recover_velocity(g, &T);

        rebuild_interfaces(g, &T);
        rebuild_transfers(g, &T, 0, step);
        advect_components(g, &T, step, 0);
        advect_momentum_y(g, &T, step);
        update_fraction(g, &T, scale, 0);
        recover_velocity(g, &T);

        remaining -= step;
        ++substeps;
    }
    double* swap = g->u;
    g->u = T.next_u;
    T.next_u = swap;
    swap = g->v;
    g->v = T.next_v;
    T.next_v = swap;
    transport_free(&T);
    return substeps;
}

/* setSurfaceTension (FluidGridCPU.cpp:835-841); returns 0, or -1 for a negative or non-finite value. */
int orc_set_surface_tension(orc_grid* g, double coefficient)
{
    if (!isfinite(coefficient) || coefficient < 0.0)
        return -1;
    g->sigma = coefficient;
    return 0;
}

/* capillaryStepLimit (FluidGridCPU.cpp:843-860) */
double orc_capillary_step_limit(const orc_grid* g)
{
    if (g->sigma == 0.0 || g->nl == 0 || g->ng == 0)
        return INFINITY;
    double ml = g->rho_l[0], mg = g->rho_g[0];
    for (int k = 0; k < g->nl; ++k)
        ml = dmin(ml, g->rho_l[k]);
    for (int k = 0; k < g->ng; ++k)
        mg = dmin(mg, g->rho_g[k]);
    const double pi = 3.14159265358979323846;
    return 0.5 * sqrt((ml + mg) * g->h * g->h * g->h / (2.0 * pi * g->sigma));
}

/* applySurfaceTension (FluidGridCPU.cpp:862-926): capillary face force sigma kappa df dt w / h. */
void orc_apply_surface_tension(orc_grid* g, double dt)
{
    const int W = g->W, H = g->H;
    const size_t N = (size_t)W * H;
    double* curvature = dalloc(2 * N);
    orc_capillary_geometry(W, H, g->h, g->fraction, NULL, NULL, curvature);
    rebuild_coefficients(g);
    for (int y = 0; y < H; ++y) {
        for (int x = 1; x < W; ++x) {
            const size_t cell = (size_t)y * W + x;
            const size_t face = (size_t)y * (W + 1) + x;
            const double jump = g->sigma * orc_face_curvature(curvature, cell - 1, cell) *
                                (g->fraction[cell] - g->fraction[cell - 1]);
            g->u[face] += dt * g->inv_u[face] * jump / g->h;
        }
        if (y == 0)
            continue;
        for (int x = 0; x < W; ++x) {
            const size_t cell = (size_t)y * W + x;
            const double jump = g->sigma * orc_face_curvature(curvature, cell - W, cell) *
                                (g->fraction[cell] - g->fraction[cell - W]);
            g->v[cell] += dt * g->inv_v[cell] * jump / g->h;
        }
    }
    free(curvature);
}

/* step, FluidGridCPU.cpp:300-360 (no acceleration field): capillary subdivision, then per substep
 * transport, viscosity (the WebGPU engine's, when a component is viscous), gravity, surface tension
 * and projection; stops after the first substep whose projection does not converge. */
orc_result orc_step(orc_grid* g, double dt, int max_iterations, double tol)
{
    orc_result result;
    memset(&result, 0, sizeof(result));
    result.max_divergence = INFINITY;
    if (dt <= 0.0)
        return result;
    const int tension = g->sigma > 0.0;
    const double subdivisions = tension ? dmax(1.0, ceil(dt / orc_capillary_step_limit(g))) : 1.0;
    if (subdivisions > 2147483647.0)
        return result;
    const int count = (int)subdivisions;
    const double step = dt / count;
    for (int i = 0; i < count; ++i) {
        orc_advect_reference(g, step);
        int viscosity_iterations = 0;
        if (orc_has_viscosity(g) && orc_apply_viscosity(g, step, &viscosity_iterations) != ORC_VISC_OK) {
            result.converged = 0;
            result.viscosity_iterations += viscosity_iterations;
            return result;
        }
        result.viscosity_iterations += viscosity_iterations;
        orc_apply_gravity(g, step);
        if (tension)
            orc_apply_surface_tension(g, step);
        const orc_result p = orc_project(g, step, max_iterations, tol);
        result.iterations += p.iterations;
        result.max_divergence = p.max_divergence;
        result.converged = p.converged;
        result.restarts = p.restarts;
        result.initial_residual = p.initial_residual;
        result.final_residual = p.final_residual;
        ++result.physical_steps;
        if (!result.converged)
            break;
    }
    return result;
}

/* ---- grid-level implicit viscosity (the WebGPU step order: transport, viscosity, forces, projection,
 * FluidGridGPU.cpp:1134-1158) ---- */
int orc_has_viscosity(const orc_grid* g)
{
    for (int k = 0; k < g->nl; ++k)
        if (g->mu_l[k] > 0.0)
            return 1;
    for (int k = 0; k < g->ng; ++k)
        if (g->mu_g[k] > 0.0)
            return 1;
    return 0;
}

void orc_configure_viscosity(orc_grid* g, const double* mu_l, const double* mu_g)
{
    for (int k = 0; k < g->nl; ++k)
        g->mu_l[k] = mu_l ? mu_l[k] : 0.0;
    for (int k = 0; k < g->ng; ++k)
        g->mu_g[k] = mu_g ? mu_g[k] : 0.0;
}

int orc_set_viscosity_smoother(orc_grid* g, int chebyshev_sweeps)
{
    if (chebyshev_sweeps != 0 && chebyshev_sweeps != 2 && chebyshev_sweeps != 4 && chebyshev_sweeps != 6)
        return -1;
    g->vis_sweeps = chebyshev_sweeps;
    return 0;
}

int orc_apply_viscosity(orc_grid* g, double dt, int* iterations)
{
    if (iterations)
        *iterations = 0;
    if (dt <= 0.0 || !orc_has_viscosity(g))
        return ORC_VISC_OK;
    const size_t N = (size_t)g->W * g->H;
    double* material = dalloc(2 * N);
    orc_viscosity_material((int)N, g->h, g->nl, g->rho_l, g->mu_l, g->lmass, g->ng, g->rho_g, g->mu_g, g->gmass,
        material);
    const int status = orc_viscosity_solve_smoothed(g->W, g->H, g->h, material, g->u, g->v, dt, 1e-4, -1,
        g->vis_sweeps, iterations, NULL);
    free(material);
    return status;
}

/* writeColorPixels, FluidGridCPU.cpp:1826-1860: rgb holds 3 floats per component (liquid, then gas). */
void orc_color_pixels(const orc_grid* g, const float* rgb, int liquid_fraction_alpha, float* pixels)
{
    const size_t N = (size_t)g->W * g->H;
    const double cell_volume = g->h * g->h;
    for (size_t i = 0; i < N; ++i) {
        float* pixel = pixels + i * 4;
        pixel[0] = pixel[1] = pixel[2] = 0.0f;
        pixel[3] = liquid_fraction_alpha ? (float)g->fraction[i] : 1.0f;
        for (int k = 0; k < g->nl + g->ng; ++k) {
            const int liquid = k < g->nl;
            const double density = liquid ? g->rho_l[k] : g->rho_g[k - g->nl];
            const double mass = liquid ? g->lmass[(size_t)k * N + i] : g->gmass[(size_t)(k - g->nl) * N + i];
            const double volume_fraction = mass / (density * cell_volume);
            for (int channel = 0; channel < 3; ++channel)
                pixel[channel] += (float)(volume_fraction * (double)rgb[3 * k + channel]);
        }
    }
}

/* commitEdits, FluidGridCPU.cpp:253-284, on cells already resolved and reduced by CellEditBatch::prepare
 * (FluidGridEdit.h:37-95): component indexes liquids, then gases. */
void orc_commit_edits(orc_grid* g, int count, const int* x, const int* y, const int* component, const double* mass)
{
    const size_t N = (size_t)g->W * g->H;
    for (int n = 0; n < count; ++n) {
        const size_t i = (size_t)y[n] * g->W + x[n];
        const int liquid = component[n] < g->nl;
        g->fraction[i] = liquid ? 1.0 : 0.0;
        for (int k = 0; k < g->nl; ++k)
            g->lmass[(size_t)k * N + i] = 0.0;
        for (int k = 0; k < g->ng; ++k)
            g->gmass[(size_t)k * N + i] = 0.0;
        if (liquid)
            g->lmass[(size_t)component[n] * N + i] = mass[n];
        else
            g->gmass[(size_t)(component[n] - g->nl) * N + i] = mass[n];
    }
    orc_invalidate_caches(g);
}
