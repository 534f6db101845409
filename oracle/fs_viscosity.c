/*
 * fs_viscosity.c -- FP64 CPU restatement of the reference's implicit viscosity solve.
 *
 * TEST INFRASTRUCTURE ONLY (the parity checker; never the product path).
 *
 * The reference has no CPU viscosity: it exists only as WebGPU shaders driven by
 * ViscositySolverGPU (FluidSim/UniformWebGPU/ViscositySolverGPU.cpp:59-294, shaders
 * FluidSim/UniformWebGPU/ViscosityKernels.h:10-630) and the material blend of
 * FluidGridGPU::applyViscosity (FluidSim/UniformWebGPU/FluidGridGPU.cpp:173-218, 987-1037).
 * This file restates that algorithm in FP64, single-threaded:
 *
 *   backward Euler  M (v - v0) + dt K v = 0  on the MAC faces, M = face mass 0.5 h^2 (rho_a + rho_b),
 *   K the symmetric strain operator (normal stresses 2 mu du/dx on cell centres, shear
 *   mu_vertex (du/dy + dv/dx) on interior vertices), solved by CG preconditioned with a V(2,2)
 *   damped-Jacobi (weight 1/2) multigrid whose coarsest level takes 16 Jacobi sweeps; the solve
 *   stops when the recomputed residual satisfies r M^-1 r <= tol^2 v0 M v0 + 1e-12 sum(M), checked
 *   every four iterations.
 *
 * orc_viscosity_solve_smoothed replaces the V(2,2) damped-Jacobi smoother by n = 2, 4 or 6 Jacobi sweeps
 * per side with the weights 1 / (2.125 + 1.875 cos(pi (k + 1/2) / n)) -- the reciprocals of the Chebyshev
 * nodes of [0.25, 4] (pre-smoothing k = 0..n-1, post-smoothing in reverse).  This option has no
 * counterpart in the reference; 4 bounds the spectrum of D^-1 A by Gershgorin's theorem (each row's
 * off-diagonal magnitudes sum to at most three times its viscous diagonal part), so the smoother never
 * amplifies an error component and the V-cycle stays a symmetric positive definite preconditioner.
 *
 * Storage follows the shaders: every level keeps (w+1) x (h+1) "slots", slot p holding the u face
 * on the left of cell p and the v face on top of cell p; reads clamp to the field like the WGSL
 * accessors (FluidSim/Shared/WebGPU/FluidWebGPUCompute.h:241-262).
 */
#include "fs_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int w, h;
    double *rho, *mu;            /* cell material (w*h); level 0 points at the caller's */
    double *mc, *mv, *mu_m, *mv_m; /* slot coefficients: cell mu, vertex mu, u mass, v mass */
    double *bu, *bv;             /* rhs (coarse levels) */
    double *xu, *xv, *nu, *nv;   /* solution and ping-pong */
    double *ru, *rv;             /* residual */
} vlevel;

static int vclampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
static double* vz(size_t n) { return (double*)calloc(n ? n : 1, sizeof(double)); }

static size_t slot(const vlevel* L, int x, int y)
{
    return (size_t)vclampi(y, 0, L->h) * (L->w + 1) + vclampi(x, 0, L->w);
}

static size_t cellc(const vlevel* L, int x, int y)
{
    return (size_t)vclampi(y, 0, L->h - 1) * L->w + vclampi(x, 0, L->w - 1);
}

static int valid_u(const vlevel* L, int x, int y) { return x > 0 && x < L->w && y < L->h; }
static int valid_v(const vlevel* L, int x, int y) { return y > 0 && y < L->h && x < L->w; }

/* "viscosity coefficients" (ViscosityKernels.h:137-152) */
static void coefficients(vlevel* L, double spacing)
{
    for (int y = 0; y <= L->h; ++y)
        for (int x = 0; x <= L->w; ++x) {
            const size_t s = (size_t)y * (L->w + 1) + x;
            const size_t c = cellc(L, x, y);
            double vertex = 0.0;
            if (x > 0 && y > 0 && x < L->w && y < L->h)
                vertex = 0.25 * (L->mu[c] + L->mu[cellc(L, x - 1, y)] + L->mu[cellc(L, x, y - 1)] +
                                 L->mu[cellc(L, x - 1, y - 1)]);
            L->mc[s] = L->mu[c];
            L->mv[s] = vertex;
            L->mu_m[s] = 0.5 * spacing * spacing * (L->rho[c] + L->rho[cellc(L, x - 1, y)]);
            L->mv_m[s] = 0.5 * spacing * spacing * (L->rho[c] + L->rho[cellc(L, x, y - 1)]);
        }
}

/* "viscosity coarse material" (ViscosityKernels.h:115-135): mean of the existing 2 x 2 children. */
static void coarsen_material(const vlevel* F, vlevel* C)
{
    for (int y = 0; y < C->h; ++y)
        for (int x = 0; x < C->w; ++x) {
            double r = 0.0, m = 0.0, n = 0.0;
            for (int dy = 0; dy < 2; ++dy)
                for (int dx = 0; dx < 2; ++dx) {
                    const int qx = 2 * x + dx, qy = 2 * y + dy;
                    if (qx < F->w && qy < F->h) {
                        r += F->rho[(size_t)qy * F->w + qx];
                        m += F->mu[(size_t)qy * F->w + qx];
                        n += 1.0;
                    }
                }
            C->rho[(size_t)y * C->w + x] = r / n;
            C->mu[(size_t)y * C->w + x] = m / n;
        }
}

/* shear and stressProduct (ViscosityKernels.h:26-44) on the field (su, sv). */
static double shear(const vlevel* L, const double* su, const double* sv, int x, int y)
{
    return L->mv[slot(L, x, y)] * ((su[slot(L, x, y)] - su[slot(L, x, y - 1)]) +
                                   (sv[slot(L, x, y)] - sv[slot(L, x - 1, y)]));
}

static void stress_product(const vlevel* L, const double* su, const double* sv, int x, int y, double* ou,
    double* ov)
{
    const size_t p = slot(L, x, y);
    const double cu = su[p], cv = sv[p];
    const double t = shear(L, su, sv, x, y);
    const double nu = 2.0 * (L->mc[slot(L, x - 1, y)] * (cu - su[slot(L, x - 1, y)]) +
                             L->mc[p] * (cu - su[slot(L, x + 1, y)]));
    const double nv = 2.0 * (L->mc[slot(L, x, y - 1)] * (cv - sv[slot(L, x, y - 1)]) +
                             L->mc[p] * (cv - sv[slot(L, x, y + 1)]));
    *ou = nu + (t - shear(L, su, sv, x, y + 1));
    *ov = nv + (t - shear(L, su, sv, x + 1, y));
}

/* product (ViscosityKernels.h:53-57): M s + dt K s on valid faces, 0 elsewhere. */
static void product(const vlevel* L, const double* su, const double* sv, double dt, int x, int y, double* ou,
    double* ov)
{
    double ku, kv;
    stress_product(L, su, sv, x, y, &ku, &kv);
    const size_t p = slot(L, x, y);
    *ou = valid_u(L, x, y) ? L->mu_m[p] * su[p] + dt * ku : 0.0;
    *ov = valid_v(L, x, y) ? L->mv_m[p] * sv[p] + dt * kv : 0.0;
}

/* diagonal (ViscosityKernels.h:17-24) */
static void diagonal(const vlevel* L, double dt, int x, int y, double* du, double* dv)
{
    const size_t p = slot(L, x, y);
    *du = L->mu_m[p] + dt * (2.0 * (L->mc[p] + L->mc[slot(L, x - 1, y)]) + L->mv[p] + L->mv[slot(L, x, y + 1)]);
    *dv = L->mv_m[p] + dt * (2.0 * (L->mc[p] + L->mc[slot(L, x, y - 1)]) + L->mv[p] + L->mv[slot(L, x + 1, y)]);
}

/* "viscosity symmetric smoother" (ViscosityKernels.h:214-236): damped Jacobi, weight 1/2 (w = 0.5). */
static void smooth(vlevel* L, const double* bu, const double* bv, double dt, int zero, double w)
{
    for (int y = 0; y <= L->h; ++y)
        for (int x = 0; x <= L->w; ++x) {
            const size_t p = (size_t)y * (L->w + 1) + x;
            double du, dv, au = 0.0, av = 0.0;
            diagonal(L, dt, x, y, &du, &dv);
            double vu, vv;
            if (zero) {
                vu = w * bu[p] / du;
                vv = w * bv[p] / dv;
            } else {
                product(L, L->xu, L->xv, dt, x, y, &au, &av);
                vu = L->xu[p] + w * (bu[p] - au) / du;
                vv = L->xv[p] + w * (bv[p] - av) / dv;
            }
            L->nu[p] = valid_u(L, x, y) ? vu : 0.0;
            L->nv[p] = valid_v(L, x, y) ? vv : 0.0;
        }
    double* t = L->xu; L->xu = L->nu; L->nu = t;
    t = L->xv; L->xv = L->nv; L->nv = t;
}

/* "viscosity true residual" (ViscosityKernels.h:187-195) */
static void residual(const vlevel* L, const double* su, const double* sv, const double* bu, const double* bv,
    double dt, double* ru, double* rv)
{
    for (int y = 0; y <= L->h; ++y)
        for (int x = 0; x <= L->w; ++x) {
            const size_t p = (size_t)y * (L->w + 1) + x;
            double au, av;
            product(L, su, sv, dt, x, y, &au, &av);
            ru[p] = valid_u(L, x, y) ? bu[p] - au : 0.0;
            rv[p] = valid_v(L, x, y) ? bv[p] - av : 0.0;
        }
}

/* Face-location bilinear transfer (ViscosityKernels.h:286-316).  Component 0 = u (offset (0, 1/2)),
 * 1 = v (offset (1/2, 0)).  coarse_point clamps to [0, size - (0,1)] for u, [0, size - (1,0)] for v. */
static void transfer_location(int fx, int fy, int comp, int* bx, int* by, double* tx, double* ty)
{
    const double ox = comp == 1 ? 0.5 : 0.0, oy = comp == 1 ? 0.0 : 0.5;
    const double lx = (fx + ox) * 0.5 - ox, ly = (fy + oy) * 0.5 - oy;
    *bx = (int)floor(lx);
    *by = (int)floor(ly);
    *tx = lx - floor(lx);
    *ty = ly - floor(ly);
}

static void coarse_point(int x, int y, int comp, int cw, int ch, int* ox, int* oy)
{
    *ox = vclampi(x, 0, cw - (comp == 1 ? 1 : 0));
    *oy = vclampi(y, 0, ch - (comp == 1 ? 0 : 1));
}

static double transfer_weight(int fx, int fy, int cx, int cy, int comp, int cw, int ch)
{
    int bx, by;
    double tx, ty;
    transfer_location(fx, fy, comp, &bx, &by, &tx, &ty);
    double result = 0.0;
    for (int y = 0; y < 2; ++y)
        for (int x = 0; x < 2; ++x) {
            int px, py;
            coarse_point(bx + x, by + y, comp, cw, ch, &px, &py);
            if (px == cx && py == cy)
                result += (x == 1 ? tx : 1.0 - tx) * (y == 1 ? ty : 1.0 - ty);
        }
    return result;
}

/* "viscosity transpose restriction" (ViscosityKernels.h:318-357) */
static void restrict_residual(const vlevel* F, vlevel* C)
{
    for (int y = 0; y <= C->h; ++y)
        for (int x = 0; x <= C->w; ++x) {
            const size_t p = (size_t)y * (C->w + 1) + x;
            double val[2] = {0.0, 0.0};
            for (int comp = 0; comp < 2; ++comp) {
                if ((comp == 0 && (x == 0 || x == C->w || y == C->h)) || (comp == 1 && (y == 0 || y == C->h || x == C->w)))
                    continue;
                const int ex = comp == 1 ? 2 : 1, ey = comp == 1 ? 1 : 2;
                for (int dy = -1; dy <= ey; ++dy)
                    for (int dx = -1; dx <= ex; ++dx) {
                        const int qx = 2 * x + dx, qy = 2 * y + dy;
                        if (qx < 0 || qy < 0 || qx > F->w || qy > F->h)
                            continue;
                        const size_t q = (size_t)qy * (F->w + 1) + qx;
                        val[comp] += transfer_weight(qx, qy, x, y, comp, C->w, C->h) * (comp == 0 ? F->ru[q] : F->rv[q]);
                    }
            }
            C->bu[p] = val[0];
            C->bv[p] = val[1];
        }
}

/* "viscosity prolongation" (ViscosityKernels.h:359-391) */
static void prolong(vlevel* F, const vlevel* C)
{
    for (int y = 0; y <= F->h; ++y)
        for (int x = 0; x <= F->w; ++x) {
            const size_t p = (size_t)y * (F->w + 1) + x;
            double val[2] = {F->xu[p], F->xv[p]};
            for (int comp = 0; comp < 2; ++comp) {
                int bx, by;
                double tx, ty;
                transfer_location(x, y, comp, &bx, &by, &tx, &ty);
                for (int dy = 0; dy < 2; ++dy)
                    for (int dx = 0; dx < 2; ++dx) {
                        int px, py;
                        coarse_point(bx + dx, by + dy, comp, C->w, C->h, &px, &py);
                        const size_t q = (size_t)py * (C->w + 1) + px;
                        val[comp] += (dx == 1 ? tx : 1.0 - tx) * (dy == 1 ? ty : 1.0 - ty) *
                                     (comp == 0 ? C->xu[q] : C->xv[q]);
                    }
            }
            F->nu[p] = valid_u(F, x, y) ? val[0] : 0.0;
            F->nv[p] = valid_v(F, x, y) ? val[1] : 0.0;
        }
    double* t = F->xu; F->xu = F->nu; F->nu = t;
    t = F->xv; F->xv = F->nv; F->nv = t;
}

/* "viscosity shared coarse solve" (ViscosityKernels.h:242-282): 16 Jacobi sweeps from zero. */
static void coarse_solve(vlevel* L, const double* bu, const double* bv, double dt)
{
    const size_t n = (size_t)(L->w + 1) * (L->h + 1);
    memset(L->xu, 0, n * sizeof(double));
    memset(L->xv, 0, n * sizeof(double));
    for (int sweep = 0; sweep < 16; ++sweep)
        smooth(L, bu, bv, dt, 0, 0.5);
}

typedef struct {
    int nlev;
    vlevel* lev;
    double dt;
    int sweeps;         /* per side */
    double weight[6];   /* pre-smoothing weights; post-smoothing takes them in reverse */
} vsolver;

/* Weight k of n Chebyshev-weighted sweeps: 1 / (2.125 + 1.875 cos(pi (k + 1/2) / n)). */
double orc_viscosity_chebyshev_weight(int k, int n)
{
    return 1.0 / (2.125 + 1.875 * cos(3.14159265358979323846 * (k + 0.5) / n));
}

/* cycle (ViscositySolverGPU.cpp:134-167) */
static void cycle(vsolver* S, int level, const double* bu, const double* bv)
{
    vlevel* L = &S->lev[level];
    if (level + 1 == S->nlev) {
        coarse_solve(L, bu, bv, S->dt);
        return;
    }
    for (int k = 0; k < S->sweeps; ++k)
        smooth(L, bu, bv, S->dt, k == 0, S->weight[k]);
    residual(L, L->xu, L->xv, bu, bv, S->dt, L->ru, L->rv);
    vlevel* C = &S->lev[level + 1];
    restrict_residual(L, C);
    cycle(S, level + 1, C->bu, C->bv);
    prolong(L, C);
    for (int k = S->sweeps - 1; k >= 0; --k)
        smooth(L, bu, bv, S->dt, 0, S->weight[k]);
}

static void level_alloc(vlevel* L, int w, int h, int fine)
{
    const size_t ns = (size_t)(w + 1) * (h + 1), nc = (size_t)w * h;
    memset(L, 0, sizeof(*L));
    L->w = w;
    L->h = h;
    if (!fine) {
        L->rho = vz(nc);
        L->mu = vz(nc);
    }
    L->mc = vz(ns); L->mv = vz(ns); L->mu_m = vz(ns); L->mv_m = vz(ns);
    L->bu = vz(ns); L->bv = vz(ns);
    L->xu = vz(ns); L->xv = vz(ns); L->nu = vz(ns); L->nv = vz(ns);
    L->ru = vz(ns); L->rv = vz(ns);
}

static void level_release(vlevel* L, int fine)
{
    if (!fine) {
        free(L->rho);
        free(L->mu);
    }
    free(L->mc); free(L->mv); free(L->mu_m); free(L->mv_m);
    free(L->bu); free(L->bv); free(L->xu); free(L->xv); free(L->nu); free(L->nv);
    free(L->ru); free(L->rv);
}

int orc_viscosity_levels(int width, int height)
{
    int n = 0;
    for (;;) {
        ++n;
        if (width * height <= 64)
            break;
        width = (width + 1) / 2;
        height = (height + 1) / 2;
    }
    return n;
}

static int finite3(double a, double b, double c) { return isfinite(a) && isfinite(b) && isfinite(c); }

/* ViscositySolverGPU::Impl::solve (ViscositySolverGPU.cpp:169-294), not prescribed boundary velocity. */
int orc_viscosity_solve(int W, int H, double spacing, const double* material, double* u, double* v, double dt,
    double tolerance, int maximum, int* iterations_out, double* relative_residual)
{
    return orc_viscosity_solve_smoothed(W, H, spacing, material, u, v, dt, tolerance, maximum, 0, iterations_out,
        relative_residual);
}

/* chebyshev_sweeps 0: the reference's V(2,2) damped Jacobi; 2, 4 or 6: the Chebyshev-weighted sweeps. */
int orc_viscosity_solve_smoothed(int W, int H, double spacing, const double* material, double* u, double* v,
    double dt, double tolerance, int maximum, int chebyshev_sweeps, int* iterations_out, double* relative_residual)
{
    if (chebyshev_sweeps != 0 && chebyshev_sweeps != 2 && chebyshev_sweeps != 4 && chebyshev_sweeps != 6)
        return ORC_VISC_BAD_SMOOTHER;
    if (iterations_out)
        *iterations_out = 0;
    if (relative_residual)
        *relative_residual = 0.0;
    if (dt <= 0.0)
        return ORC_VISC_OK;
    const size_t NC = (size_t)W * H;
    vsolver S;
    S.dt = dt;
    S.sweeps = chebyshev_sweeps ? chebyshev_sweeps : 2;
    for (int k = 0; k < S.sweeps; ++k)
        S.weight[k] = chebyshev_sweeps ? orc_viscosity_chebyshev_weight(k, chebyshev_sweeps) : 0.5;
    S.nlev = orc_viscosity_levels(W, H);
    S.lev = (vlevel*)calloc((size_t)S.nlev, sizeof(vlevel));
    double* rho0 = vz(NC);
    double* mu0 = vz(NC);
    for (size_t i = 0; i < NC; ++i) {
        rho0[i] = material[2 * i];
        mu0[i] = material[2 * i + 1];
    }
    {
        int w = W, h = H;
        for (int l = 0; l < S.nlev; ++l) {
            level_alloc(&S.lev[l], w, h, l == 0);
            w = (w + 1) / 2;
            h = (h + 1) / 2;
        }
    }
    S.lev[0].rho = rho0;
    S.lev[0].mu = mu0;
    for (int l = 0; l < S.nlev; ++l) {
        if (l > 0)
            coarsen_material(&S.lev[l - 1], &S.lev[l]);
        coefficients(&S.lev[l], ldexp(spacing, l));
    }
    vlevel* F = &S.lev[0];
    const size_t ns = (size_t)(W + 1) * (H + 1);
    double *xu = vz(ns), *xv = vz(ns), *bu = vz(ns), *bv = vz(ns), *ru = vz(ns), *rv = vz(ns);
    double *du = vz(ns), *dv = vz(ns), *qu = vz(ns), *qv = vz(ns), *tu = vz(ns), *tv = vz(ns);
    /* "viscosity initial velocity and RHS" (ViscosityKernels.h:154-163) */
    for (int y = 0; y <= H; ++y)
        for (int x = 0; x <= W; ++x) {
            const size_t p = (size_t)y * (W + 1) + x;
            const double su = u[(size_t)vclampi(y, 0, H - 1) * (W + 1) + x];
            const double sv = v[(size_t)y * W + vclampi(x, 0, W - 1)];
            xu[p] = valid_u(F, x, y) ? su : 0.0;
            xv[p] = valid_v(F, x, y) ? sv : 0.0;
            bu[p] = F->mu_m[p] * xu[p];
            bv[p] = F->mv_m[p] * xv[p];
        }
    int status = ORC_VISC_OK;
    int iterations = 0;
    /* reduce(mode 0): r M^-1 r, x M x, sum of masses, non-finite flag */
    residual(F, xu, xv, bu, bv, dt, ru, rv);
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    int bad = 0;
    for (size_t p = 0; p < ns; ++p) {
        const double a = ru[p] / F->mu_m[p] * ru[p] + rv[p] / F->mv_m[p] * rv[p];
        const double b = F->mu_m[p] * xu[p] * xu[p] + F->mv_m[p] * xv[p] * xv[p];
        const double c = F->mu_m[p] + F->mv_m[p];
        if (!finite3(a, b, c))
            bad = 1;
        else {
            s0 += a;
            s1 += b;
            s2 += c;
        }
    }
    const double target = tolerance * tolerance * s1 + 1e-12 * s2;
    double rel = sqrt(s0 / (s1 > 1e-30 ? s1 : 1e-30));
    if (bad) {
        status = ORC_VISC_NONFINITE;
        goto done;
    }
    if (s0 <= target)
        goto done;
    {
        double rz = 0.0, alpha = 0.0, beta = 0.0;
        int breakdown = 0, steps = 0;
        double rnorm = s0;
        while (maximum < 0 || iterations < maximum) {
            const int batch = maximum < 0 ? 4 : (maximum - iterations < 4 ? maximum - iterations : 4);
            for (int it = 0; it < batch; ++it) {
                if (breakdown || rnorm <= target)
                    continue;   /* stopped(): the dispatched slot does nothing */
                /* z = V-cycle(r) */
                cycle(&S, 0, ru, rv);
                double zr = 0.0, rr = 0.0;
                for (size_t p = 0; p < ns; ++p) {
                    zr += ru[p] * F->xu[p] + rv[p] * F->xv[p];
                    rr += ru[p] / F->mu_m[p] * ru[p] + rv[p] / F->mv_m[p] * rv[p];
                }
                rnorm = rr;
                if (!(rnorm > target))
                    continue;
                if (zr <= 0.0 || !isfinite(zr)) {
                    breakdown = 1;
                    continue;
                }
                beta = steps > 0 ? zr / rz : 0.0;
                rz = zr;
                for (size_t p = 0; p < ns; ++p) {
                    du[p] = F->xu[p] + (steps > 0 ? beta * du[p] : 0.0);
                    dv[p] = F->xv[p] + (steps > 0 ? beta * dv[p] : 0.0);
                }
                for (int y = 0; y <= H; ++y)
                    for (int x = 0; x <= W; ++x) {
                        const size_t p = (size_t)y * (W + 1) + x;
                        product(F, du, dv, dt, x, y, &qu[p], &qv[p]);
                    }
                double dq = 0.0;
                for (size_t p = 0; p < ns; ++p)
                    dq += du[p] * qu[p] + dv[p] * qv[p];
                if (dq <= 0.0 || !isfinite(dq)) {
                    breakdown = 1;
                    continue;
                }
                alpha = rz / dq;
                ++steps;
                for (size_t p = 0; p < ns; ++p) {
                    xu[p] += alpha * du[p];
                    xv[p] += alpha * dv[p];
                    ru[p] -= alpha * qu[p];
                    rv[p] -= alpha * qv[p];
                }
            }
            iterations += batch;
            /* recomputed residual (reduce mode 3) */
            residual(F, xu, xv, bu, bv, dt, tu, tv);
            double t0 = 0.0;
            for (size_t p = 0; p < ns; ++p)
                t0 += tu[p] / F->mu_m[p] * tu[p] + tv[p] / F->mv_m[p] * tv[p];
            rel = sqrt(t0 / (s1 > 1e-30 ? s1 : 1e-30));
            if (breakdown || !isfinite(t0)) {
                status = ORC_VISC_BREAKDOWN;
                goto done;
            }
            if (t0 <= target) {
                /* "viscosity staggered result": every face of u and v, walls included */
                for (int y = 0; y < H; ++y)
                    for (int x = 0; x <= W; ++x)
                        u[(size_t)y * (W + 1) + x] = xu[(size_t)y * (W + 1) + x];
                for (int y = 0; y <= H; ++y)
                    for (int x = 0; x < W; ++x)
                        v[(size_t)y * W + x] = xv[(size_t)y * (W + 1) + x];
                goto done;
            }
            if (rnorm <= target) {
                /* "viscosity refresh residual": restart CG from the recomputed residual */
                memcpy(ru, tu, ns * sizeof(double));
                memcpy(rv, tv, ns * sizeof(double));
                rnorm = t0;
                steps = 0;
                rz = alpha = beta = 0.0;
            }
        }
        status = ORC_VISC_NOT_CONVERGED;
    }
done:
    if (iterations_out)
        *iterations_out = iterations;
    if (relative_residual)
        *relative_residual = rel;
    free(xu); free(xv); free(bu); free(bv); free(ru); free(rv);
    free(du); free(dv); free(qu); free(qv); free(tu); free(tv);
    for (int l = 0; l < S.nlev; ++l)
        level_release(&S.lev[l], l == 0);
    free(S.lev);
    free(rho0);
    free(mu0);
    return status;
}

/* The material blend of FluidGridGPU::applyViscosity (FluidGridGPU.cpp:173-218, 995-1031):
 * rho = sum of masses / h^2, mu = sum(m_k mu_k / rho_k) / sum(m_k / rho_k). */
void orc_viscosity_material(int N, double cell_size, int nl, const double* liquid_density,
    const double* liquid_viscosity, const double* liquid_mass, int ng, const double* gas_density,
    const double* gas_viscosity, const double* gas_mass, double* material)
{
    const double area = cell_size * cell_size;
    for (int i = 0; i < N; ++i) {
        double mass = 0.0, volume = 0.0, mu_volume = 0.0;
        for (int k = 0; k < nl; ++k) {
            const double m = liquid_mass[(size_t)k * N + i];
            mass += m;
            volume += m / liquid_density[k];
            mu_volume += m * (liquid_viscosity[k] / liquid_density[k]);
        }
        for (int k = 0; k < ng; ++k) {
            const double m = gas_mass[(size_t)k * N + i];
            mass += m;
            volume += m / gas_density[k];
            mu_volume += m * (gas_viscosity[k] / gas_density[k]);
        }
        material[2 * i] = mass / area;
        material[2 * i + 1] = mu_volume / volume;
    }
}
