/*
 * ilh_oracle.c -- CPU FP64 restatement of ImmersedLayers.jl's immersed-layer
 * hot path, written from the reference's sources and from the published
 * semantics of its CartesianGrids.jl dependency.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for libimmersedhip.so.  Loaded only by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 *
 * What each part follows (reference paths relative to the reference checkout):
 *   orc_ddf              Yang3/Roma/M4prime/Witchhat DDFs used by Regularize,
 *                        default Yang3 (src/ImmersedLayers.jl:83).
 *   orc_regularize       R = RegularizationMatrix with GridScaling weights ds/dx^2
 *                        (src/cache.jl:305-316; src/surface_operators.jl:38-46).
 *   orc_interpolate      E = InterpolationMatrix, unit weights
 *                        (src/surface_operators.jl:80-88).
 *   orc_regularize_normal / orc_normal_interpolate and the _cross variants
 *                        src/surface_operators.jl:98-103,150-162,228-233,277-290.
 *   stencils             src/grid_operators.jl:25-134 (fill!, unscaled stencil,
 *                        ./= dx; Laplacian factor 1/dx^2 from src/cache.jl:324).
 *   LGF / inverse Lap.   src/grid_operators.jl:144-179: L\w = dx^2 (G * w), a
 *                        linear convolution over the field's own index set with
 *                        the 5-point lattice Green's function, G(0,0) = 0.
 *   surface calculus     src/surface_operators.jl:357-725.
 *   Schur builders       src/matrix_operators.jl:9-61,135-155,225-244 (column loops).
 *   composites           test/literate/problems.jl:259-298 (Dirichlet) and
 *                        test/literate/neumann.jl:101-142 (Neumann).
 *   vector cache         src/surface_operators.jl:113-138,236-267,388-455,
 *                        565-591,632-664; create_CL2invCT
 *                        src/matrix_operators.jl:95-125; Stokes flow
 *                        test/literate/stokes.jl:101-166.
 */
#include "ilh_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

static int g_threads = 0;
void orc_set_threads(int n) { g_threads = n; if (n > 0) omp_set_num_threads(n); }
int  orc_get_threads(void) { return g_threads > 0 ? g_threads : omp_get_max_threads(); }

/* ------------------------------------------------------------------ DDFs */
/* Yang, Zhang, Li & He (2009) JCP 228:7821 smoothed 3-point kernel "Yang3";
   Roma et al. (1999) 3-point; M4' (Monaghan); linear hat. r in grid units. */
double orc_ddf(int ddf, double r)
{
    double a = fabs(r);
    switch (ddf) {
    case ORC_YANG3: {
        const double s3 = sqrt(3.0);
        if (a <= 1.0)
            return 17.0 / 48.0 + s3 * M_PI / 108.0 + a / 4.0 - a * a / 4.0
                 + (1.0 - 2.0 * a) / 16.0 * sqrt(-12.0 * a * a + 12.0 * a + 1.0)
                 - s3 / 12.0 * asin(s3 / 2.0 * (2.0 * a - 1.0));
        if (a <= 2.0)
            return 55.0 / 48.0 - s3 * M_PI / 108.0 - 13.0 * a / 12.0 + a * a / 4.0
                 + (2.0 * a - 3.0) / 48.0 * sqrt(-12.0 * a * a + 36.0 * a - 23.0)
                 + s3 / 36.0 * asin(s3 / 2.0 * (2.0 * a - 3.0));
        return 0.0;
    }
    case ORC_ROMA:
        if (a <= 0.5) return (1.0 + sqrt(1.0 - 3.0 * a * a)) / 3.0;
        if (a <= 1.5) return (5.0 - 3.0 * a - sqrt(1.0 - 3.0 * (1.0 - a) * (1.0 - a))) / 6.0;
        return 0.0;
    case ORC_M4PRIME:
        if (a <= 1.0) return 1.0 - 2.5 * a * a + 1.5 * a * a * a;
        if (a <= 2.0) return 0.5 * (2.0 - a) * (2.0 - a) * (1.0 - a);
        return 0.0;
    case ORC_WITCHHAT:
        return a <= 1.0 ? 1.0 - a : 0.0;
    }
    return 0.0;
}

static double ddf_support(int ddf)
{
    switch (ddf) { case ORC_ROMA: return 1.5; case ORC_WITCHHAT: return 1.0; default: return 2.0; }
}

/* ----------------------------------------------------------------- grids */
void orc_dims(const orc_grid* g, int kind, int* nx, int* ny)
{
    switch (kind) {
    case ORC_PRIMAL: *nx = g->NX - 1; *ny = g->NY - 1; break;
    case ORC_DUAL:   *nx = g->NX;     *ny = g->NY;     break;
    case ORC_EDGE_U: *nx = g->NX;     *ny = g->NY - 1; break;
    default:         *nx = g->NX - 1; *ny = g->NY;     break;
    }
}

/* half-cell shifts: dual positions sit at (i - I0 - 1/2) dx */
static void kind_shift(int kind, double* sx, double* sy)
{
    *sx = (kind == ORC_DUAL || kind == ORC_EDGE_U) ? 0.5 : 0.0;
    *sy = (kind == ORC_DUAL || kind == ORC_EDGE_V) ? 0.5 : 0.0;
}

static size_t edge_u_size(const orc_grid* g) { return (size_t)g->NX * (g->NY - 1); }
static size_t edge_size(const orc_grid* g) { return edge_u_size(g) + (size_t)(g->NX - 1) * g->NY; }

/* ------------------------------------------------- regularize / interpolate */
/* r_i = (x_i - X)/dx with x_i = (i + 1 - I0 - shift) dx (0-based i) */
void orc_regularize(const orc_grid* g, int kind, int ddf, int N, const double* X, const double* Y,
                    const double* ds, const double* f, double* out)
{
    int nx, ny; orc_dims(g, kind, &nx, &ny);
    double sx, sy; kind_shift(kind, &sx, &sy);
    const double dx = g->dx, sup = ddf_support(ddf);
    memset(out, 0, sizeof(double) * (size_t)nx * ny);           /* fill!(s, 0.0) */
    for (int p = 0; p < N; ++p) {
        const double xi = X[p] / dx - (1.0 - g->I0x - sx);
        const double yi = Y[p] / dx - (1.0 - g->I0y - sy);
        const double w = ds[p] / (dx * dx) * f[p];
        int i0 = (int)ceil(xi - sup), i1 = (int)floor(xi + sup);
        int j0 = (int)ceil(yi - sup), j1 = (int)floor(yi + sup);
        if (i0 < 0) i0 = 0;
        if (j0 < 0) j0 = 0;
        if (i1 > nx - 1) i1 = nx - 1;
        if (j1 > ny - 1) j1 = ny - 1;
        for (int j = j0; j <= j1; ++j) {
            const double wy = orc_ddf(ddf, j - yi);
            for (int i = i0; i <= i1; ++i)
                out[i + (size_t)nx * j] += w * orc_ddf(ddf, i - xi) * wy;
        }
    }
}

void orc_interpolate(const orc_grid* g, int kind, int ddf, int N, const double* X, const double* Y,
                     const double* grid, double* out)
{
    int nx, ny; orc_dims(g, kind, &nx, &ny);
    double sx, sy; kind_shift(kind, &sx, &sy);
    const double dx = g->dx, sup = ddf_support(ddf);
    for (int p = 0; p < N; ++p) {
        const double xi = X[p] / dx - (1.0 - g->I0x - sx);
        const double yi = Y[p] / dx - (1.0 - g->I0y - sy);
        int i0 = (int)ceil(xi - sup), i1 = (int)floor(xi + sup);
        int j0 = (int)ceil(yi - sup), j1 = (int)floor(yi + sup);
        if (i0 < 0) i0 = 0;
        if (j0 < 0) j0 = 0;
        if (i1 > nx - 1) i1 = nx - 1;
        if (j1 > ny - 1) j1 = ny - 1;
        double acc = 0.0;
        for (int j = j0; j <= j1; ++j) {
            const double wy = orc_ddf(ddf, j - yi);
            for (int i = i0; i <= i1; ++i)
                acc += orc_ddf(ddf, i - xi) * wy * grid[i + (size_t)nx * j];
        }
        out[p] = acc;
    }
}

/* Rf (VectorData -> Edges{Primal}): one regularization per component */
static void reg_edges(const orc_grid* g, int ddf, int N, const double* X, const double* Y,
                      const double* ds, const double* fu, const double* fv, double* q)
{
    orc_regularize(g, ORC_EDGE_U, ddf, N, X, Y, ds, fu, q);
    orc_regularize(g, ORC_EDGE_V, ddf, N, X, Y, ds, fv, q + edge_u_size(g));
}

static void interp_edges(const orc_grid* g, int ddf, int N, const double* X, const double* Y,
                         const double* q, double* su, double* sv)
{
    orc_interpolate(g, ORC_EDGE_U, ddf, N, X, Y, q, su);
    orc_interpolate(g, ORC_EDGE_V, ddf, N, X, Y, q + edge_u_size(g), sv);
}

/* q = Rf (n o f)   (surface_operators.jl:98-103) */
void orc_regularize_normal(const orc_grid* g, int ddf, int N, const double* X, const double* Y,
                           const double* nx, const double* ny, const double* ds, const double* f,
                           double* q)
{
    double* t = (double*)malloc(sizeof(double) * 2 * (size_t)(N > 0 ? N : 1));
    for (int p = 0; p < N; ++p) { t[p] = nx[p] * f[p]; t[N + p] = ny[p] * f[p]; }
    reg_edges(g, ddf, N, X, Y, ds, t, t + N, q);
    free(t);
}

/* vn = n . (Rf^T q)   (surface_operators.jl:228-233) */
void orc_normal_interpolate(const orc_grid* g, int ddf, int N, const double* X, const double* Y,
                            const double* nx, const double* ny, const double* q, double* vn)
{
    double* t = (double*)malloc(sizeof(double) * 2 * (size_t)(N > 0 ? N : 1));
    interp_edges(g, ddf, N, X, Y, q, t, t + N);
    for (int p = 0; p < N; ++p) vn[p] = nx[p] * t[p] + ny[p] * t[N + p];
    free(t);
}

/* q = Rf (n x f e_z) = Rf (n_y f, -n_x f)   (surface_operators.jl:150-162) */
void orc_regularize_normal_cross(const orc_grid* g, int ddf, int N, const double* X, const double* Y,
                                 const double* nx, const double* ny, const double* ds, const double* f,
                                 double* q)
{
    double* t = (double*)malloc(sizeof(double) * 2 * (size_t)(N > 0 ? N : 1));
    for (int p = 0; p < N; ++p) { t[p] = ny[p] * f[p]; t[N + p] = -nx[p] * f[p]; }
    reg_edges(g, ddf, N, X, Y, ds, t, t + N, q);
    free(t);
}

/* wn = e_z . (n x Rf^T q) = n_x (Ef q)_v - n_y (Ef q)_u   (surface_operators.jl:277-290) */
void orc_normal_cross_interpolate(const orc_grid* g, int ddf, int N, const double* X, const double* Y,
                                  const double* nx, const double* ny, const double* q, double* vn)
{
    double* t = (double*)malloc(sizeof(double) * 2 * (size_t)(N > 0 ? N : 1));
    interp_edges(g, ddf, N, X, Y, q, t, t + N);
    for (int p = 0; p < N; ++p) vn[p] = nx[p] * t[N + p] - ny[p] * t[p];
    free(t);
}

/* -------------------------------------------------------------- stencils */
/* Unscaled MAC differences on the staggered layouts (CartesianGrids). */
static void div_unscaled(const orc_grid* g, const double* q, double* p)
{
    const int NX = g->NX, NY = g->NY;
    const double* u = q; const double* v = q + edge_u_size(g);
    for (int j = 0; j < NY - 1; ++j)
        for (int i = 0; i < NX - 1; ++i)
            p[i + (size_t)(NX - 1) * j] = u[i + 1 + (size_t)NX * j] - u[i + (size_t)NX * j]
                                        + v[i + (size_t)(NX - 1) * (j + 1)] - v[i + (size_t)(NX - 1) * j];
}

static void grad_unscaled(const orc_grid* g, const double* p, double* q)
{
    const int NX = g->NX, NY = g->NY;
    double* u = q; double* v = q + edge_u_size(g);
    memset(q, 0, sizeof(double) * edge_size(g));
    for (int j = 0; j < NY - 1; ++j)
        for (int i = 1; i < NX - 1; ++i)
            u[i + (size_t)NX * j] = p[i + (size_t)(NX - 1) * j] - p[i - 1 + (size_t)(NX - 1) * j];
    for (int j = 1; j < NY - 1; ++j)
        for (int i = 0; i < NX - 1; ++i)
            v[i + (size_t)(NX - 1) * j] = p[i + (size_t)(NX - 1) * j] - p[i + (size_t)(NX - 1) * (j - 1)];
}

static void curl_n2e_unscaled(const orc_grid* g, const double* s, double* q)
{
    const int NX = g->NX, NY = g->NY;
    double* u = q; double* v = q + edge_u_size(g);
    for (int j = 0; j < NY - 1; ++j)
        for (int i = 0; i < NX; ++i)
            u[i + (size_t)NX * j] = s[i + (size_t)NX * (j + 1)] - s[i + (size_t)NX * j];
    for (int j = 0; j < NY; ++j)
        for (int i = 0; i < NX - 1; ++i)
            v[i + (size_t)(NX - 1) * j] = s[i + (size_t)NX * j] - s[i + 1 + (size_t)NX * j];
}

static void curl_e2n_unscaled(const orc_grid* g, const double* q, double* w)
{
    const int NX = g->NX, NY = g->NY;
    const double* u = q; const double* v = q + edge_u_size(g);
    memset(w, 0, sizeof(double) * (size_t)NX * NY);
    for (int j = 1; j < NY - 1; ++j)
        for (int i = 1; i < NX - 1; ++i)
            w[i + (size_t)NX * j] = v[i + (size_t)(NX - 1) * j] - v[i - 1 + (size_t)(NX - 1) * j]
                                  - u[i + (size_t)NX * j] + u[i + (size_t)NX * (j - 1)];
}

/* GridScaling divides the unscaled differences by dx (grid_operators.jl:9) */
static void div_arr(double* a, size_t n, double d) { for (size_t k = 0; k < n; ++k) a[k] /= d; }

/* grid_operators.jl:25-31 */
void orc_divergence(const orc_grid* g, const double* q, double* p)
{
    div_unscaled(g, q, p);
    div_arr(p, (size_t)(g->NX - 1) * (g->NY - 1), g->dx);
}
/* grid_operators.jl:45-51 */
void orc_grad(const orc_grid* g, const double* p, double* q)
{
    grad_unscaled(g, p, q);
    div_arr(q, edge_size(g), g->dx);
}
/* grid_operators.jl:79-87 */
void orc_curl_n2e(const orc_grid* g, const double* s, double* q)
{
    curl_n2e_unscaled(g, s, q);
    div_arr(q, edge_size(g), g->dx);
}
/* grid_operators.jl:97-105 */
void orc_curl_e2n(const orc_grid* g, const double* q, double* w)
{
    curl_e2n_unscaled(g, q, w);
    div_arr(w, (size_t)g->NX * g->NY, g->dx);
}
/* grid_operators.jl:114-134: 5-point on interior nodes, factor 1/dx^2 */
void orc_laplacian(const orc_grid* g, int kind, const double* s, double* w)
{
    int nx, ny; orc_dims(g, kind, &nx, &ny);
    const double f = 1.0 / (g->dx * g->dx);
    memset(w, 0, sizeof(double) * (size_t)nx * ny);
    for (int j = 1; j < ny - 1; ++j)
        for (int i = 1; i < nx - 1; ++i) {
            size_t k = i + (size_t)nx * j;
            w[k] = f * (s[k - nx] + s[k - 1] - 4.0 * s[k] + s[k + 1] + s[k + nx]);
        }
}

/* ----------------------------------------------------------------- LGF */
/* G(m,n) = (1/2pi) int_0^pi [1 - cos(m t) e^{-|n| sigma}] / sinh(sigma) dt,
   cosh(sigma) = 2 - cos t.  Split as A(n) + B(m,n) with n = max(|m|,|n|):
     A(n)   = (1/2pi) int_0^pi (1 - e^{-n sigma}) / sinh sigma dt
     B(m,n) = (1/2pi) int_0^T  e^{-n sigma} 2 sin^2(m t/2) / sinh sigma dt,
   T = min(pi, 50/n) (e^{-n sigma} < 1e-19 beyond).  Gauss-Legendre panels:
   geometric [0,1/n],[1/n,2/n],[2/n,4/n],... for A; width 4/n for B. */
#define GL_K 20
static double gl_x[GL_K], gl_w[GL_K];
static int gl_ready = 0;

static void gl_init(void)
{
    if (gl_ready) return;
    for (int i = 0; i < GL_K; ++i) {
        double x = cos(M_PI * (i + 0.75) / (GL_K + 0.5));
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = x;
            for (int k = 2; k <= GL_K; ++k) { double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k; p0 = p1; p1 = p2; }
            double dp = GL_K * (x * p1 - p0) / (x * x - 1.0);
            double dxn = p1 / dp;
            x -= dxn;
            if (fabs(dxn) < 1e-17) break;
        }
        double p0 = 1.0, p1 = x;
        for (int k = 2; k <= GL_K; ++k) { double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k; p0 = p1; p1 = p2; }
        double dp = GL_K * (x * p1 - p0) / (x * x - 1.0);
        gl_x[i] = x;
        gl_w[i] = 2.0 / ((1.0 - x * x) * dp * dp);
    }
    gl_ready = 1;
}

/* u = 2 sin^2(t/2) = cosh(sigma) - 1; sinh sigma = sqrt(u(2+u)); sigma = log1p(u + sinh sigma) */
static void lgf_sig(double t, double* sig, double* sh)
{
    double h = sin(0.5 * t);
    double u = 2.0 * h * h;
    *sh = sqrt(u * (2.0 + u));
    *sig = log1p(u + *sh);
}

static double lgf_A(int n)
{
    double acc = 0.0, a = 0.0, b = 1.0 / n;
    for (;;) {
        if (b > M_PI) b = M_PI;
        double c = 0.5 * (a + b), h = 0.5 * (b - a);
        for (int k = 0; k < GL_K; ++k) {
            double t = c + h * gl_x[k], sig, sh;
            lgf_sig(t, &sig, &sh);
            acc += h * gl_w[k] * (-expm1(-n * sig)) / sh;
        }
        if (b >= M_PI) break;
        a = b; b = 2.0 * b;
    }
    return acc / (2.0 * M_PI);
}

/* nodes/weights of the B integral for a given n (independent of m) */
static int lgf_B_nodes(int n, double* t, double* wt)
{
    double T = 50.0 / n; if (T > M_PI) T = M_PI;
    double hw = 4.0 / n;
    int np = (int)ceil(T / hw - 1e-12);
    if (np < 1) np = 1;
    double h = T / np;
    int c = 0;
    for (int q = 0; q < np; ++q) {
        double mid = (q + 0.5) * h;
        for (int k = 0; k < GL_K; ++k) {
            double tt = mid + 0.5 * h * gl_x[k], sig, sh;
            lgf_sig(tt, &sig, &sh);
            t[c] = tt;
            wt[c] = 0.5 * h * gl_w[k] * exp(-n * sig) / sh;
            ++c;
        }
    }
    return c;
}

double orc_lgf(int m, int n)
{
    gl_init();
    m = abs(m); n = abs(n);
    if (m > n) { int t = m; m = n; n = t; }
    if (n == 0) return 0.0;
    double t[GL_K * 64], wt[GL_K * 64];
    int K = lgf_B_nodes(n, t, wt);
    double B = 0.0;
    for (int k = 0; k < K; ++k) { double s = sin(0.5 * m * t[k]); B += wt[k] * 2.0 * s * s; }
    return lgf_A(n) + B / (2.0 * M_PI);
}

void orc_lgf_table(int mx, int my, double* tab)
{
    gl_init();
    int nmax = mx > my ? mx : my;
    #pragma omp parallel for schedule(dynamic, 4)
    for (int n = 0; n < nmax; ++n) {
        if (n == 0) { tab[0] = 0.0; continue; }
        double t[GL_K * 64], wt[GL_K * 64];
        int K = lgf_B_nodes(n, t, wt);
        double A = lgf_A(n);
        for (int m = 0; m <= n; ++m) {
            double B = 0.0;
            for (int k = 0; k < K; ++k) { double s = sin(0.5 * m * t[k]); B += wt[k] * 2.0 * s * s; }
            double G = A + B / (2.0 * M_PI);
            if (m < mx && n < my) tab[m + (size_t)mx * n] = G;
            if (n < mx && m < my) tab[n + (size_t)mx * m] = G;
        }
    }
}

/* out(i,j) = sum_{k,l} G(i-k, j-l) in(k,l): brute force, tiny grids only */
void orc_conv_direct(int nx, int ny, const double* tab, int tmx, const double* in, double* out)
{
    #pragma omp parallel for
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            double acc = 0.0;
            for (int l = 0; l < ny; ++l)
                for (int k = 0; k < nx; ++k)
                    acc += tab[abs(i - k) + (size_t)tmx * abs(j - l)] * in[k + (size_t)nx * l];
            out[i + (size_t)nx * j] = acc;
        }
}

/* ------------------------------------------------------------------ FFT */
typedef struct { double re, im; } cpx;

static void fft_inplace(cpx* a, int n, const cpx* tw, int sign)
{
    for (int i = 1, j = 0; i < n; ++i) {               /* bit reversal */
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) { cpx t = a[i]; a[i] = a[j]; a[j] = t; }
    }
    for (int len = 2; len <= n; len <<= 1) {
        int step = n / len;
        for (int i = 0; i < n; i += len)
            for (int k = 0; k < len / 2; ++k) {
                cpx w = tw[k * step];
                if (sign > 0) w.im = -w.im;
                cpx x = a[i + k], y = a[i + k + len / 2];
                cpx t = { y.re * w.re - y.im * w.im, y.re * w.im + y.im * w.re };
                a[i + k].re = x.re + t.re; a[i + k].im = x.im + t.im;
                a[i + k + len / 2].re = x.re - t.re; a[i + k + len / 2].im = x.im - t.im;
            }
    }
}

static cpx* make_tw(int n)   /* tw[k] = exp(-2 pi i k / n) */
{
    cpx* tw = (cpx*)malloc(sizeof(cpx) * (size_t)(n > 1 ? n : 2));
    for (int k = 0; k < n; ++k) { double a = -2.0 * M_PI * k / n; tw[k].re = cos(a); tw[k].im = sin(a); }
    return tw;
}

static int next_pow2(int v) { int p = 1; while (p < v) p <<= 1; return p; }

struct orc_lap {
    int nx, ny, Px, Py, tmx, tmy;
    double scale;          /* dx^2 / (Px Py) */
    double* tab;           /* G(m,n), 0<=m<tmx, 0<=n<tmy */
    double* spec;          /* real spectrum of the circular kernel, Px x Py */
    cpx *twx, *twy;
};

/* 2-D FFT of a Px x Py complex array (x fastest); rows only for j < nrows */
static void fft2(cpx* a, int Px, int Py, const cpx* twx, const cpx* twy, int sign, int rows_first, int nrows)
{
    if (rows_first) {
        #pragma omp parallel for
        for (int j = 0; j < nrows; ++j) fft_inplace(a + (size_t)Px * j, Px, twx, sign);
    }
    #pragma omp parallel
    {
        cpx* col = (cpx*)malloc(sizeof(cpx) * Py);
        #pragma omp for
        for (int i = 0; i < Px; ++i) {
            for (int j = 0; j < Py; ++j) col[j] = a[i + (size_t)Px * j];
            fft_inplace(col, Py, twy, sign);
            for (int j = 0; j < Py; ++j) a[i + (size_t)Px * j] = col[j];
        }
        free(col);
    }
    if (!rows_first) {
        #pragma omp parallel for
        for (int j = 0; j < nrows; ++j) fft_inplace(a + (size_t)Px * j, Px, twx, sign);
    }
}

orc_lap* orc_lap_create(const orc_grid* g, int kind)
{
    orc_lap* L = (orc_lap*)calloc(1, sizeof(orc_lap));
    orc_dims(g, kind, &L->nx, &L->ny);
    /* G is even, so a period of 2n-2 already separates every needed offset */
    L->Px = next_pow2(2 * L->nx - 2 > 2 ? 2 * L->nx - 2 : 2);
    L->Py = next_pow2(2 * L->ny - 2 > 2 ? 2 * L->ny - 2 : 2);
    L->tmx = L->Px / 2 + 1; L->tmy = L->Py / 2 + 1;
    if (L->tmx < L->nx) L->tmx = L->nx;
    if (L->tmy < L->ny) L->tmy = L->ny;
    L->tab = (double*)malloc(sizeof(double) * (size_t)L->tmx * L->tmy);
    orc_lgf_table(L->tmx, L->tmy, L->tab);
    L->twx = make_tw(L->Px); L->twy = make_tw(L->Py);
    L->scale = g->dx * g->dx / ((double)L->Px * L->Py);
    size_t P = (size_t)L->Px * L->Py;
    cpx* a = (cpx*)malloc(sizeof(cpx) * P);
    for (int j = 0; j < L->Py; ++j) {
        int b = j <= L->Py / 2 ? j : L->Py - j;
        for (int i = 0; i < L->Px; ++i) {
            int c = i <= L->Px / 2 ? i : L->Px - i;
            a[i + (size_t)L->Px * j].re = L->tab[c + (size_t)L->tmx * b];
            a[i + (size_t)L->Px * j].im = 0.0;
        }
    }
    fft2(a, L->Px, L->Py, L->twx, L->twy, -1, 1, L->Py);
    L->spec = (double*)malloc(sizeof(double) * P);
    for (size_t k = 0; k < P; ++k) L->spec[k] = a[k].re;
    free(a);
    return L;
}

void orc_lap_destroy(orc_lap* L)
{
    if (!L) return;
    free(L->tab); free(L->spec); free(L->twx); free(L->twy); free(L);
}

const double* orc_lap_table(orc_lap* L, int* tmx) { *tmx = L->tmx; return L->tab; }

/* L \ w = dx^2 (G * w) over the field's own index set (zero-padded FFT) */
void orc_inverse_laplacian(orc_lap* L, const double* in, double* out)
{
    const int nx = L->nx, ny = L->ny, Px = L->Px, Py = L->Py;
    size_t P = (size_t)Px * Py;
    cpx* a = (cpx*)calloc(P, sizeof(cpx));
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) a[i + (size_t)Px * j].re = in[i + (size_t)nx * j];
    fft2(a, Px, Py, L->twx, L->twy, -1, 1, ny);
    #pragma omp parallel for
    for (long long k = 0; k < (long long)P; ++k) {
        double s = L->spec[k] * L->scale;
        a[k].re *= s; a[k].im *= s;
    }
    fft2(a, Px, Py, L->twx, L->twy, +1, 0, ny);
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) out[i + (size_t)nx * j] = a[i + (size_t)Px * j].re;
    free(a);
}

/* ----------------------------------------------------- surface calculus */
static double* dalloc(size_t n) { return (double*)calloc(n > 0 ? n : 1, sizeof(double)); }

/* theta = D Rf (n o f) / dx   (surface_operators.jl:530-547) */
void orc_surface_divergence(const orc_grid* g, int ddf, int N, const double* X, const double* Y,
                            const double* nx, const double* ny, const double* ds, const double* f, double* th)
{
    double* q = dalloc(edge_size(g));
    orc_regularize_normal(g, ddf, N, X, Y, nx, ny, ds, f, q);
    div_unscaled(g, q, th);
    div_arr(th, (size_t)(g->NX - 1) * (g->NY - 1), g->dx);
    free(q);
}

/* vn = n . Rf^T G phi / dx   (surface_operators.jl:604-620) */
void orc_surface_grad(const orc_grid* g, int ddf, int N, const double* X, const double* Y,
                      const double* nx, const double* ny, const double* phi, double* vn)
{
    double* q = dalloc(edge_size(g));
    grad_unscaled(g, phi, q);
    orc_normal_interpolate(g, ddf, N, X, Y, nx, ny, q, vn);
    div_arr(vn, N, g->dx);
    free(q);
}

/* w = C^T Rf (n o f) / dx   (surface_operators.jl:357-375) */
void orc_surface_curl(const orc_grid* g, int ddf, int N, const double* X, const double* Y,
                      const double* nx, const double* ny, const double* ds, const double* f, double* w)
{
    double* q = dalloc(edge_size(g));
    orc_regularize_normal(g, ddf, N, X, Y, nx, ny, ds, f, q);
    curl_e2n_unscaled(g, q, w);
    div_arr(w, (size_t)g->NX * g->NY, g->dx);
    free(q);
}

/* vn = n . Rf^T C s / dx   (surface_operators.jl:412-429) */
void orc_surface_curl_adj(const orc_grid* g, int ddf, int N, const double* X, const double* Y,
                          const double* nx, const double* ny, const double* s, double* vn)
{
    double* q = dalloc(edge_size(g));
    curl_n2e_unscaled(g, s, q);
    orc_normal_interpolate(g, ddf, N, X, Y, nx, ny, q, vn);
    div_arr(vn, N, g->dx);
    free(q);
}

/* w = C^T Rf (n x f e_z) / dx   (surface_operators.jl:467-484) */
void orc_surface_curl_cross(const orc_grid* g, int ddf, int N, const double* X, const double* Y,
                            const double* nx, const double* ny, const double* ds, const double* f, double* w)
{
    double* q = dalloc(edge_size(g));
    orc_regularize_normal_cross(g, ddf, N, X, Y, nx, ny, ds, f, q);
    curl_e2n_unscaled(g, q, w);
    div_arr(w, (size_t)g->NX * g->NY, g->dx);
    free(q);
}

/* gamma = e_z . (n x Rf^T C s) / dx   (surface_operators.jl:498-515) */
void orc_surface_curl_cross_adj(const orc_grid* g, int ddf, int N, const double* X, const double* Y,
                                const double* nx, const double* ny, const double* s, double* vn)
{
    double* q = dalloc(edge_size(g));
    curl_n2e_unscaled(g, s, q);
    orc_normal_cross_interpolate(g, ddf, N, X, Y, nx, ny, q, vn);
    div_arr(vn, N, g->dx);
    free(q);
}

/* gamma = e_z . (n x Rf^T G phi) / dx   (surface_operators.jl:708-725) */
void orc_surface_grad_cross(const orc_grid* g, int ddf, int N, const double* X, const double* Y,
                            const double* nx, const double* ny, const double* phi, double* vn)
{
    double* q = dalloc(edge_size(g));
    grad_unscaled(g, phi, q);
    orc_normal_cross_interpolate(g, ddf, N, X, Y, nx, ny, q, vn);
    div_arr(vn, N, g->dx);
    free(q);
}

/* --------------------------------------------------------------- Schur */
/* Column loops exactly as src/matrix_operators.jl (one unit vector per column). */
void orc_schur_columns(const orc_grid* g, int ddf, int N, const double* X, const double* Y,
                       const double* nx, const double* ny, const double* ds, int kind, double scale, double* A)
{
    int gk = (kind == ORC_S_CLINVCT) ? ORC_DUAL : ORC_PRIMAL;
    int gx, gy; orc_dims(g, gk, &gx, &gy);
    orc_lap* L = (kind == ORC_S_NRTRN) ? NULL : orc_lap_create(g, gk);
    double* e = dalloc(N);
    double* col = dalloc(N);
    double* gd = dalloc((size_t)gx * gy);
    double* q = dalloc(edge_size(g));
    for (int c = 0; c < N; ++c) {
        memset(e, 0, sizeof(double) * N);
        e[c] = 1.0;
        double sg = -scale;
        switch (kind) {
        case ORC_S_RTLINVR:            /* matrix_operators.jl:16-26 */
            orc_regularize(g, ORC_PRIMAL, ddf, N, X, Y, ds, e, gd);
            orc_inverse_laplacian(L, gd, gd);
            orc_interpolate(g, ORC_PRIMAL, ddf, N, X, Y, gd, col);
            break;
        case ORC_S_CLINVCT:            /* matrix_operators.jl:47-56 */
            orc_surface_curl(g, ddf, N, X, Y, nx, ny, ds, e, gd);
            orc_inverse_laplacian(L, gd, gd);
            orc_surface_curl_adj(g, ddf, N, X, Y, nx, ny, gd, col);
            break;
        case ORC_S_GLINVD:             /* matrix_operators.jl:142-150 */
            orc_surface_divergence(g, ddf, N, X, Y, nx, ny, ds, e, gd);
            orc_inverse_laplacian(L, gd, gd);
            orc_surface_grad(g, ddf, N, X, Y, nx, ny, gd, col);
            break;
        default:                       /* create_nRTRn, matrix_operators.jl:232-239 */
            orc_regularize_normal(g, ddf, N, X, Y, nx, ny, ds, e, q);
            orc_normal_interpolate(g, ddf, N, X, Y, nx, ny, q, col);
            sg = scale;
            break;
        }
        for (int r = 0; r < N; ++r) A[r + (size_t)N * c] = sg * col[r];
    }
    free(e); free(col); free(gd); free(q);
    orc_lap_destroy(L);
}

/* Direct assembly: the same column definition evaluated on a 16x16 local grid
   around each point (the composition only touches a few cells), then the
   inverse Laplacian applied as dx^2 sum G(a-b) between the two local patches.
   Valid when every point sits at least 5 cells inside the grid. */
#define LOC 16
typedef struct { int off_x, off_y; double v[LOC * LOC]; } patch;   /* one grid kind */

static void local_grid(const orc_grid* g, double X, double Y, orc_grid* lg, int* offx, int* offy)
{
    double xi = X / g->dx - (1.0 - g->I0x), yi = Y / g->dx - (1.0 - g->I0y);
    *offx = (int)floor(xi) - LOC / 2 + 1;
    *offy = (int)floor(yi) - LOC / 2 + 1;
    lg->NX = LOC + 1; lg->NY = LOC + 1; lg->dx = g->dx;
    lg->I0x = g->I0x - *offx; lg->I0y = g->I0y - *offy;
}

/* column of the surface->grid half (regularize side) for point q, on grid kind gk */
static void col_patch(const orc_grid* g, int ddf, const double* X, const double* Y, const double* nx,
                      const double* ny, const double* ds, int q, int kind, patch* out)
{
    orc_grid lg; int ox, oy;
    local_grid(g, X[q], Y[q], &lg, &ox, &oy);
    double one = 1.0;
    double buf[(LOC + 1) * (LOC + 1)];
    switch (kind) {
    case ORC_S_RTLINVR: orc_regularize(&lg, ORC_PRIMAL, ddf, 1, X + q, Y + q, ds + q, &one, buf); break;
    case ORC_S_CLINVCT: orc_surface_curl(&lg, ddf, 1, X + q, Y + q, nx + q, ny + q, ds + q, &one, buf); break;
    default:            orc_surface_divergence(&lg, ddf, 1, X + q, Y + q, nx + q, ny + q, ds + q, &one, buf); break;
    }
    int gk = kind == ORC_S_CLINVCT ? ORC_DUAL : ORC_PRIMAL;
    int lx, ly; orc_dims(&lg, gk, &lx, &ly);
    out->off_x = ox; out->off_y = oy;
    for (int j = 0; j < LOC; ++j)
        for (int i = 0; i < LOC; ++i) out->v[i + LOC * j] = buf[i + (size_t)lx * j];
}

/* row of the grid->surface half (interpolate side) for point p: the linear
   functional evaluated on unit fields of the local grid */
static void row_patch(const orc_grid* g, int ddf, const double* X, const double* Y, const double* nx,
                      const double* ny, int p, int kind, patch* out)
{
    orc_grid lg; int ox, oy;
    local_grid(g, X[p], Y[p], &lg, &ox, &oy);
    int gk = kind == ORC_S_CLINVCT ? ORC_DUAL : ORC_PRIMAL;
    int lx, ly; orc_dims(&lg, gk, &lx, &ly);
    double buf[(LOC + 1) * (LOC + 1)];
    out->off_x = ox; out->off_y = oy;
    memset(buf, 0, sizeof(buf));
    for (int j = 0; j < LOC; ++j)
        for (int i = 0; i < LOC; ++i) {
            double val = 0.0;
            if (i >= 3 && i < LOC - 3 && j >= 3 && j < LOC - 3) {
                buf[i + (size_t)lx * j] = 1.0;
                switch (kind) {
                case ORC_S_RTLINVR: orc_interpolate(&lg, ORC_PRIMAL, ddf, 1, X + p, Y + p, buf, &val); break;
                case ORC_S_CLINVCT: orc_surface_curl_adj(&lg, ddf, 1, X + p, Y + p, nx + p, ny + p, buf, &val); break;
                default:            orc_surface_grad(&lg, ddf, 1, X + p, Y + p, nx + p, ny + p, buf, &val); break;
                }
                buf[i + (size_t)lx * j] = 0.0;
            }
            out->v[i + LOC * j] = val;
        }
}

void orc_schur_direct(const orc_grid* g, int ddf, int N, const double* X, const double* Y,
                      const double* nx, const double* ny, const double* ds, int kind, double scale, double* A)
{
    if (kind == ORC_S_NRTRN) {     /* no inverse Laplacian: the column loop is cheap */
        orc_schur_columns(g, ddf, N, X, Y, nx, ny, ds, kind, scale, A);
        return;
    }
    int gk = kind == ORC_S_CLINVCT ? ORC_DUAL : ORC_PRIMAL;
    int gx, gy; orc_dims(g, gk, &gx, &gy);
    int tmx = gx + LOC, tmy = gy + LOC;
    double* tab = (double*)malloc(sizeof(double) * (size_t)tmx * tmy);
    orc_lgf_table(tmx, tmy, tab);
    patch* cols = (patch*)malloc(sizeof(patch) * (size_t)(N > 0 ? N : 1));
    patch* rows = (patch*)malloc(sizeof(patch) * (size_t)(N > 0 ? N : 1));
    #pragma omp parallel for
    for (int k = 0; k < N; ++k) {
        col_patch(g, ddf, X, Y, nx, ny, ds, k, kind, &cols[k]);
        row_patch(g, ddf, X, Y, nx, ny, k, kind, &rows[k]);
    }
    const double dx2 = g->dx * g->dx;
    #pragma omp parallel
    {
        int ia[LOC * LOC], ja[LOC * LOC], ib[LOC * LOC], jb[LOC * LOC];
        double va[LOC * LOC], vb[LOC * LOC];
        #pragma omp for schedule(dynamic, 8)
        for (int q = 0; q < N; ++q) {
            int nb = 0;
            for (int j = 0; j < LOC; ++j)
                for (int i = 0; i < LOC; ++i) {
                    double v = cols[q].v[i + LOC * j];
                    if (v != 0.0) { ib[nb] = i + cols[q].off_x; jb[nb] = j + cols[q].off_y; vb[nb] = v; ++nb; }
                }
            for (int p = 0; p < N; ++p) {
                int na = 0;
                for (int j = 0; j < LOC; ++j)
                    for (int i = 0; i < LOC; ++i) {
                        double v = rows[p].v[i + LOC * j];
                        if (v != 0.0) { ia[na] = i + rows[p].off_x; ja[na] = j + rows[p].off_y; va[na] = v; ++na; }
                    }
                double acc = 0.0;
                for (int a = 0; a < na; ++a) {
                    double s = 0.0;
                    for (int b = 0; b < nb; ++b)
                        s += tab[abs(ia[a] - ib[b]) + (size_t)tmx * abs(ja[a] - jb[b])] * vb[b];
                    acc += va[a] * s;
                }
                A[p + (size_t)N * q] = -scale * dx2 * acc;
            }
        }
    }
    free(cols); free(rows); free(tab);
}

/* ------------------------------------------------------------------ LU */
/* LU with partial pivoting (LAPACK getrf, what Julia's `\` on a Matrix runs),
   blocked right-looking: each 64-column panel is factored column by column
   with the row interchanges applied to whole rows, then the block row U12 =
   L11^-1 A12 and the trailing update A22 -= L21 U12 (columns in parallel).
   The pivots are those of the unblocked algorithm; the blocking only changes
   the summation order of the updates. */
int orc_lu_factor(int n, double* A, int* piv)
{
    int info = 0;
    const int NB = 64;
    for (int k0 = 0; k0 < n; k0 += NB) {
        const int nb = n - k0 < NB ? n - k0 : NB, k1 = k0 + nb;
        for (int k = k0; k < k1; ++k) {
            int pr = k; double mx = fabs(A[k + (size_t)n * k]);
            for (int i = k + 1; i < n; ++i) { double v = fabs(A[i + (size_t)n * k]); if (v > mx) { mx = v; pr = i; } }
            piv[k] = pr;
            if (mx == 0.0) { if (!info) info = k + 1; continue; }
            if (pr != k)
                for (int j = 0; j < n; ++j) { double t = A[k + (size_t)n * j]; A[k + (size_t)n * j] = A[pr + (size_t)n * j]; A[pr + (size_t)n * j] = t; }
            const double inv = 1.0 / A[k + (size_t)n * k];
            double* ck = A + (size_t)n * k;
            for (int i = k + 1; i < n; ++i) ck[i] *= inv;
            for (int j = k + 1; j < k1; ++j) {
                double* cj = A + (size_t)n * j;
                const double akj = cj[k];
                if (akj == 0.0) continue;
                for (int i = k + 1; i < n; ++i) cj[i] -= ck[i] * akj;
            }
        }
        #pragma omp parallel for schedule(static) if (n - k1 > 128)
        for (int j = k1; j < n; ++j) {
            double* cj = A + (size_t)n * j;
            for (int k = k0; k < k1; ++k) {           /* U12: unit-lower solve */
                const double akj = cj[k];
                if (akj == 0.0) continue;
                const double* ck = A + (size_t)n * k;
                for (int i = k + 1; i < k1; ++i) cj[i] -= ck[i] * akj;
            }
            for (int k = k0; k < k1; ++k) {           /* A22 -= L21 U12 */
                const double akj = cj[k];
                if (akj == 0.0) continue;
                const double* ck = A + (size_t)n * k;
                for (int i = k1; i < n; ++i) cj[i] -= ck[i] * akj;
            }
        }
    }
    return info;
}

void orc_lu_solve(int n, const double* LU, const int* piv, double* b)
{
    for (int k = 0; k < n; ++k) if (piv[k] != k) { double t = b[k]; b[k] = b[piv[k]]; b[piv[k]] = t; }
    for (int j = 0; j < n; ++j) { double bj = b[j]; for (int i = j + 1; i < n; ++i) b[i] -= LU[i + (size_t)n * j] * bj; }
    for (int j = n - 1; j >= 0; --j) { b[j] /= LU[j + (size_t)n * j]; double bj = b[j]; for (int i = 0; i < j; ++i) b[i] -= LU[i + (size_t)n * j] * bj; }
}

/* x = S \ b with a fresh factorisation, as the reference's S\s on a Matrix */
static int dense_solve(int n, const double* S, double* b)
{
    double* A = dalloc((size_t)n * n);
    int* piv = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    memcpy(A, S, sizeof(double) * (size_t)n * n);
    int info = orc_lu_factor(n, A, piv);
    if (!info) orc_lu_solve(n, A, piv, b);
    free(A); free(piv);
    return info;
}

/* The reference keeps the Laplacian plan (LGF table + FFT plans) in the cache
   (src/cache.jl:170-178); the composites reuse the last plans built here. */
static orc_lap* cached_lap(const orc_grid* g, int kind)
{
    static orc_grid keys[2];
    static orc_lap* plans[2];
    static int used[2];
    for (int k = 0; k < 2; ++k)
        if (used[k] && plans[k] && memcmp(&keys[k], g, sizeof(*g)) == 0 && k == kind) return plans[k];
    if (used[kind]) orc_lap_destroy(plans[kind]);
    plans[kind] = orc_lap_create(g, kind);
    keys[kind] = *g;
    used[kind] = 1;
    return plans[kind];
}

/* ---------------------------------------------------------- composites */
/* test/literate/problems.jl:259-298 (no forcing). */
int orc_dirichlet_solve(const orc_grid* g, int ddf, int N, const double* X, const double* Y,
                        const double* nx, const double* ny, const double* ds, const double* S,
                        const double* fplus, const double* fminus, double* f, double* s)
{
    size_t ng = (size_t)(g->NX - 1) * (g->NY - 1);
    double* Sown = NULL;
    if (!S) {
        Sown = dalloc((size_t)N * N);
        orc_schur_direct(g, ddf, N, X, Y, nx, ny, ds, ORC_S_RTLINVR, 1.0, Sown);
        S = Sown;
    }
    orc_lap* L = cached_lap(g, ORC_PRIMAL);
    double* fb = dalloc(N);
    double* fstar = dalloc(ng);
    for (int p = 0; p < N; ++p) fb[p] = fplus[p] - (fminus ? fminus[p] : 0.0);   /* jump */
    orc_surface_divergence(g, ddf, N, X, Y, nx, ny, ds, fb, fstar);
    orc_inverse_laplacian(L, fstar, fstar);
    for (int p = 0; p < N; ++p) { fb[p] = 0.5 * fplus[p]; fb[p] += 0.5 * (fminus ? fminus[p] : 0.0); }
    orc_interpolate(g, ORC_PRIMAL, ddf, N, X, Y, fstar, s);
    for (int p = 0; p < N; ++p) s[p] = fb[p] - s[p];
    int info = dense_solve(N, S, s);
    for (int p = 0; p < N; ++p) s[p] = -s[p];
    orc_regularize(g, ORC_PRIMAL, ddf, N, X, Y, ds, s, f);
    orc_inverse_laplacian(L, f, f);
    for (size_t k = 0; k < ng; ++k) f[k] += fstar[k];
    free(fb); free(fstar); free(Sown);
    return info;
}

/* test/literate/neumann.jl:101-142. */
int orc_neumann_solve(const orc_grid* g, int ddf, int N, const double* X, const double* Y,
                      const double* nx, const double* ny, const double* ds, const double* S,
                      const double* vplus, const double* vminus,
                      double* f, double* df, double* s, double* dsv)
{
    size_t ng = (size_t)(g->NX - 1) * (g->NY - 1), nd = (size_t)g->NX * g->NY;
    double* Sown = NULL;
    if (!S) {
        Sown = dalloc((size_t)N * N);
        orc_schur_direct(g, ddf, N, X, Y, nx, ny, ds, ORC_S_CLINVCT, 1.0, Sown);
        S = Sown;
    }
    orc_lap* Lp = cached_lap(g, ORC_PRIMAL);
    orc_lap* Ld = cached_lap(g, ORC_DUAL);
    double* dvn = dalloc(N); double* vn = dalloc(N);
    double* fstar = dalloc(ng); double* sstar = dalloc(nd);
    for (int p = 0; p < N; ++p) {
        double vm = vminus ? vminus[p] : 0.0;
        dvn[p] = vplus[p] - vm;
        vn[p] = 0.5 * vplus[p]; vn[p] += 0.5 * vm;
    }
    orc_regularize(g, ORC_PRIMAL, ddf, N, X, Y, ds, dvn, fstar);
    orc_inverse_laplacian(Lp, fstar, fstar);
    orc_surface_grad(g, ddf, N, X, Y, nx, ny, fstar, df);
    for (int p = 0; p < N; ++p) df[p] = vn[p] - df[p];
    int info = dense_solve(N, S, df);
    for (int p = 0; p < N; ++p) df[p] = -df[p];
    orc_surface_divergence(g, ddf, N, X, Y, nx, ny, ds, df, f);
    orc_inverse_laplacian(Lp, f, f);
    for (size_t k = 0; k < ng; ++k) f[k] += fstar[k];
    orc_surface_curl(g, ddf, N, X, Y, nx, ny, ds, df, sstar);
    orc_surface_grad_cross(g, ddf, N, X, Y, nx, ny, fstar, dsv);
    info |= dense_solve(N, S, dsv);
    orc_surface_curl_cross(g, ddf, N, X, Y, nx, ny, ds, dsv, s);
    for (size_t k = 0; k < nd; ++k) { s[k] -= sstar[k]; s[k] *= -1.0; }
    orc_inverse_laplacian(Ld, s, s);
    free(dvn); free(vn); free(fstar); free(sstar); free(Sown);
    return info;
}


/* ------------------------------------------- vector cache (Stokes flow) */
/* EdgeGradient{Primal,Dual} = [dudx; dudy; dvdx; dvdy]: dudx and dvdy on
   Nodes{Primal}, dudy and dvdx on Nodes{Dual} (where the MAC differences of
   the edge components land: the x difference of u and the y difference of v
   at the primal nodes, the cross differences at the dual nodes). */
static size_t prim_size(const orc_grid* g) { return (size_t)(g->NX - 1) * (g->NY - 1); }
static size_t dual_size(const orc_grid* g) { return (size_t)g->NX * g->NY; }
size_t orc_tensor_size(const orc_grid* g) { return 2 * (prim_size(g) + dual_size(g)); }

static void tensor_parts(const orc_grid* g, double* A, double** xx, double** xy, double** yx, double** yy)
{
    *xx = A;
    *xy = A + prim_size(g);
    *yx = *xy + dual_size(g);
    *yy = *yx + dual_size(g);
}

/* grad!(A::EdgeGradient, q::Edges{Primal}), unscaled; the ring of the dual
   components that has no edge pair around it stays 0 */
static void grad_tensor_unscaled(const orc_grid* g, const double* q, double* A)
{
    const int NX = g->NX, NY = g->NY;
    const double* u = q; const double* v = q + edge_u_size(g);
    double *xx, *xy, *yx, *yy;
    tensor_parts(g, A, &xx, &xy, &yx, &yy);
    memset(A, 0, sizeof(double) * orc_tensor_size(g));
    for (int j = 0; j < NY - 1; ++j)
        for (int i = 0; i < NX - 1; ++i) {
            xx[i + (size_t)(NX - 1) * j] = u[i + 1 + (size_t)NX * j] - u[i + (size_t)NX * j];
            yy[i + (size_t)(NX - 1) * j] = v[i + (size_t)(NX - 1) * (j + 1)] - v[i + (size_t)(NX - 1) * j];
        }
    for (int j = 1; j < NY - 1; ++j)
        for (int i = 0; i < NX; ++i)
            xy[i + (size_t)NX * j] = u[i + (size_t)NX * j] - u[i + (size_t)NX * (j - 1)];
    for (int j = 0; j < NY; ++j)
        for (int i = 1; i < NX - 1; ++i)
            yx[i + (size_t)NX * j] = v[i + (size_t)(NX - 1) * j] - v[i - 1 + (size_t)(NX - 1) * j];
}

/* divergence!(q::Edges{Primal}, A::EdgeGradient), unscaled, by rows:
   q_u = d/dx A_xx + d/dy A_xy, q_v = d/dx A_yx + d/dy A_yy.  Applied here
   only to symmetric tensors, for which rows and columns give the same. */
static void div_tensor_unscaled(const orc_grid* g, double* A, double* q)
{
    const int NX = g->NX, NY = g->NY;
    double* u = q; double* v = q + edge_u_size(g);
    double *xx, *xy, *yx, *yy;
    tensor_parts(g, A, &xx, &xy, &yx, &yy);
    memset(q, 0, sizeof(double) * edge_size(g));
    for (int j = 0; j < NY - 1; ++j)
        for (int i = 1; i < NX - 1; ++i)
            u[i + (size_t)NX * j] = (xx[i + (size_t)(NX - 1) * j] - xx[i - 1 + (size_t)(NX - 1) * j])
                                  + (xy[i + (size_t)NX * (j + 1)] - xy[i + (size_t)NX * j]);
    for (int j = 1; j < NY - 1; ++j)
        for (int i = 0; i < NX - 1; ++i)
            v[i + (size_t)(NX - 1) * j] = (yx[i + 1 + (size_t)NX * j] - yx[i + (size_t)NX * j])
                                        + (yy[i + (size_t)(NX - 1) * j] - yy[i + (size_t)(NX - 1) * (j - 1)]);
}

/* A = R_t (n o f + f o n - (n . f) I)   (surface_operators.jl:121-138), in the
   reference's order: tensor product, plus its transpose, minus n . f on the
   diagonal, then one regularization per component */
void orc_regularize_normal_symm(const orc_grid* g, int ddf, int N, const double* X, const double* Y,
                                const double* nx, const double* ny, const double* ds, const double* f, double* A)
{
    double* t = dalloc(4 * (size_t)N);
    double *txx = t, *txy = t + N, *tyx = t + 2 * N, *tyy = t + 3 * N;
    for (int p = 0; p < N; ++p) {
        const double fu = f[p], fv = f[N + p];
        const double nf = nx[p] * fu + ny[p] * fv;
        txx[p] = (nx[p] * fu + nx[p] * fu) - nf;
        txy[p] = nx[p] * fv + ny[p] * fu;
        tyx[p] = ny[p] * fu + nx[p] * fv;
        tyy[p] = (ny[p] * fv + ny[p] * fv) - nf;
    }
    double *xx, *xy, *yx, *yy;
    tensor_parts(g, A, &xx, &xy, &yx, &yy);
    orc_regularize(g, ORC_PRIMAL, ddf, N, X, Y, ds, txx, xx);
    orc_regularize(g, ORC_DUAL, ddf, N, X, Y, ds, txy, xy);
    orc_regularize(g, ORC_DUAL, ddf, N, X, Y, ds, tyx, yx);
    orc_regularize(g, ORC_PRIMAL, ddf, N, X, Y, ds, tyy, yy);
    free(t);
}

/* tau = n . R_t^T (A + A^T - tr(A) I)   (surface_operators.jl:245-267) */
void orc_normal_interpolate_symm(const orc_grid* g, int ddf, int N, const double* X, const double* Y,
                                 const double* nx, const double* ny, const double* A, double* tau)
{
    const size_t np = prim_size(g), nd = dual_size(g);
    double* B = dalloc(orc_tensor_size(g));
    double *bxx, *bxy, *byx, *byy;
    tensor_parts(g, B, &bxx, &bxy, &byx, &byy);
    const double* axx = A; const double* axy = A + np; const double* ayx = axy + nd; const double* ayy = ayx + nd;
    for (size_t k = 0; k < np; ++k) {
        const double tr = axx[k] + ayy[k];
        bxx[k] = (axx[k] + axx[k]) - tr;
        byy[k] = (ayy[k] + ayy[k]) - tr;
    }
    for (size_t k = 0; k < nd; ++k) {
        bxy[k] = ayx[k] + axy[k];
        byx[k] = axy[k] + ayx[k];
    }
    double* t = dalloc(4 * (size_t)N);
    orc_interpolate(g, ORC_PRIMAL, ddf, N, X, Y, bxx, t);
    orc_interpolate(g, ORC_DUAL, ddf, N, X, Y, bxy, t + N);
    orc_interpolate(g, ORC_DUAL, ddf, N, X, Y, byx, t + 2 * N);
    orc_interpolate(g, ORC_PRIMAL, ddf, N, X, Y, byy, t + 3 * N);
    for (int p = 0; p < N; ++p) {
        tau[p] = nx[p] * t[p] + ny[p] * t[2 * N + p];
        tau[N + p] = nx[p] * t[N + p] + ny[p] * t[3 * N + p];
    }
    free(t); free(B);
}

/* q = D R_t (n o f + f o n - (n . f) I) / dx   (surface_operators.jl:574-591) */
void orc_surface_divergence_symm(const orc_grid* g, int ddf, int N, const double* X, const double* Y,
                                 const double* nx, const double* ny, const double* ds, const double* f, double* q)
{
    double* A = dalloc(orc_tensor_size(g));
    orc_regularize_normal_symm(g, ddf, N, X, Y, nx, ny, ds, f, A);
    div_tensor_unscaled(g, A, q);
    div_arr(q, edge_size(g), g->dx);
    free(A);
}

/* tau = n . R_t^T (G v + (G v)^T - tr(G v) I) / dx   (surface_operators.jl:632-664) */
void orc_surface_grad_symm(const orc_grid* g, int ddf, int N, const double* X, const double* Y,
                           const double* nx, const double* ny, const double* q, double* tau)
{
    double* A = dalloc(orc_tensor_size(g));
    grad_tensor_unscaled(g, q, A);
    orc_normal_interpolate_symm(g, ddf, N, X, Y, nx, ny, A, tau);
    div_arr(tau, 2 * (size_t)N, g->dx);
    free(A);
}

/* w = C^T R f / dx for VectorData f   (surface_operators.jl:388-400) */
void orc_surface_curl_vector(const orc_grid* g, int ddf, int N, const double* X, const double* Y,
                             const double* ds, const double* f, double* w)
{
    double* q = dalloc(edge_size(g));
    reg_edges(g, ddf, N, X, Y, ds, f, f + N, q);
    curl_e2n_unscaled(g, q, w);
    div_arr(w, dual_size(g), g->dx);
    free(q);
}

/* v = R^T C s / dx   (surface_operators.jl:443-455) */
void orc_surface_curl_vector_adj(const orc_grid* g, int ddf, int N, const double* X, const double* Y,
                                 const double* s, double* v)
{
    double* q = dalloc(edge_size(g));
    curl_n2e_unscaled(g, s, q);
    interp_edges(g, ddf, N, X, Y, q, v, v + N);
    div_arr(v, 2 * (size_t)N, g->dx);
    free(q);
}

/* create_CL2invCT (matrix_operators.jl:95-125): 2N x 2N column-major, one
   unit vector per column, the inverse Laplacian applied twice */
void orc_cl2invct(const orc_grid* g, int ddf, int N, const double* X, const double* Y, const double* ds,
                  double scale, double* A)
{
    const int n2 = 2 * N;
    orc_lap* L = orc_lap_create(g, ORC_DUAL);
    double* e = dalloc(n2);
    double* col = dalloc(n2);
    double* w = dalloc(dual_size(g));
    for (int c = 0; c < n2; ++c) {
        memset(e, 0, sizeof(double) * n2);
        e[c] = 1.0;
        orc_surface_curl_vector(g, ddf, N, X, Y, ds, e, w);
        orc_inverse_laplacian(L, w, w);
        orc_inverse_laplacian(L, w, w);
        orc_surface_curl_vector_adj(g, ddf, N, X, Y, w, col);
        for (int r = 0; r < n2; ++r) A[r + (size_t)n2 * c] = -scale * col[r];
    }
    free(e); free(col); free(w);
    orc_lap_destroy(L);
}

/* create_GLinvD_symm (src/matrix_operators.jl:165-185), 2N x 2N column-major:
   column c = -scale G_s^symm L^-1 D_s^symm e_c, the inverse Laplacian of
   Edges{Primal} taken per component on its own grid */
void orc_glinvd_symm(const orc_grid* g, int ddf, int N, const double* X, const double* Y, const double* nx,
                     const double* ny, const double* ds, double scale, double* A)
{
    const int n2 = 2 * N;
    const size_t nu = edge_u_size(g);
    orc_lap* Lu = orc_lap_create(g, ORC_EDGE_U);
    orc_lap* Lv = orc_lap_create(g, ORC_EDGE_V);
    double* e = dalloc(n2);
    double* col = dalloc(n2);
    double* q = dalloc(edge_size(g));
    for (int c = 0; c < n2; ++c) {
        memset(e, 0, sizeof(double) * n2);
        e[c] = 1.0;
        orc_surface_divergence_symm(g, ddf, N, X, Y, nx, ny, ds, e, q);
        orc_inverse_laplacian(Lu, q, q);
        orc_inverse_laplacian(Lv, q + nu, q + nu);
        orc_surface_grad_symm(g, ddf, N, X, Y, nx, ny, q, col);
        for (int r = 0; r < n2; ++r) A[r + (size_t)n2 * c] = -scale * col[r];
    }
    free(e); free(col); free(q);
    orc_lap_destroy(Lu);
    orc_lap_destroy(Lv);
}

/* test/literate/stokes.jl:101-166 in the reference's order, traction
   unfiltered (the reference's last step, sigma = C^2 sigma, is left to the
   caller).  S2 (2N x 2N) and Ss (N x N) column-major, or NULL to build them
   here; each solve re-factorises them, as S \ vprime does. */
int orc_stokes_solve(const orc_grid* g, int ddf, int N, const double* X, const double* Y,
                     const double* nx, const double* ny, const double* ds, const double* S2, const double* Ss,
                     const double* vplus, const double* vminus, double* v, double* s, double* sigma)
{
    const int n2 = 2 * N;
    const size_t ne = edge_size(g), nd = dual_size(g), np = prim_size(g);
    double *S2own = NULL, *Ssown = NULL;
    if (!S2) {
        S2own = dalloc((size_t)n2 * n2);
        orc_cl2invct(g, ddf, N, X, Y, ds, 1.0, S2own);
        S2 = S2own;
    }
    if (!Ss) {
        Ssown = dalloc((size_t)N * N);
        orc_schur_columns(g, ddf, N, X, Y, nx, ny, ds, ORC_S_CLINVCT, 1.0, Ssown);
        Ss = Ssown;
    }
    orc_lap* Lp = cached_lap(g, ORC_PRIMAL);
    orc_lap* Ld = cached_lap(g, ORC_DUAL);
    double* dv = dalloc(n2); double* vb = dalloc(n2); double* vprime = dalloc(n2); double* dvn = dalloc(N);
    double* sstar = dalloc(nd); double* phi = dalloc(np); double* vphi = dalloc(ne); double* dsv = dalloc(N);
    for (int k = 0; k < n2; ++k) {                       /* prescribed_surface_jump! / average! */
        const double vm = vminus ? vminus[k] : 0.0;
        dv[k] = vplus[k] - vm;
        vb[k] = 0.5 * vplus[k];
        vb[k] += 0.5 * vm;
    }
    orc_surface_divergence_symm(g, ddf, N, X, Y, nx, ny, ds, dv, v);      /* psi* */
    orc_curl_e2n(g, v, sstar);
    for (size_t k = 0; k < nd; ++k) sstar[k] *= -1.0;
    orc_inverse_laplacian(Ld, sstar, sstar);
    orc_inverse_laplacian(Ld, sstar, sstar);
    for (int p = 0; p < N; ++p) dvn[p] = nx[p] * dv[p] + ny[p] * dv[N + p];   /* normal velocity jump */
    orc_regularize(g, ORC_PRIMAL, ddf, N, X, Y, ds, dvn, phi);
    orc_inverse_laplacian(Lp, phi, phi);
    orc_grad(g, phi, vphi);
    interp_edges(g, ddf, N, X, Y, vphi, vprime, vprime + N);
    for (int k = 0; k < n2; ++k) vprime[k] = vb[k] - vprime[k];
    orc_curl_n2e(g, sstar, v);
    interp_edges(g, ddf, N, X, Y, v, vb, vb + N);
    for (int k = 0; k < n2; ++k) vprime[k] -= vb[k];                       /* spurious slip */
    memcpy(sigma, vprime, sizeof(double) * n2);
    int info = dense_solve(n2, S2, sigma);
    orc_surface_curl_vector(g, ddf, N, X, Y, ds, sigma, s);                /* correction streamfunction */
    for (size_t k = 0; k < nd; ++k) s[k] *= -1.0;
    orc_inverse_laplacian(Ld, s, s);
    orc_inverse_laplacian(Ld, s, s);
    for (size_t k = 0; k < nd; ++k) s[k] += sstar[k];
    orc_curl_n2e(g, s, v);
    for (size_t k = 0; k < ne; ++k) v[k] += vphi[k];
    orc_surface_grad_cross(g, ddf, N, X, Y, nx, ny, phi, dsv);            /* potential's streamfunction */
    info |= dense_solve(N, Ss, dsv);
    orc_surface_curl_cross(g, ddf, N, X, Y, nx, ny, ds, dsv, sstar);
    for (size_t k = 0; k < nd; ++k) sstar[k] *= -1.0;
    orc_inverse_laplacian(Ld, sstar, sstar);
    for (size_t k = 0; k < nd; ++k) s[k] += sstar[k];
    free(dv); free(vb); free(vprime); free(dvn); free(sstar); free(phi); free(vphi); free(dsv);
    free(S2own); free(Ssown);
    return info;
}

/* ------------------------------------------------------------------ heat */
/* K(p) = e^{-x} I_p(x), x = 2a, from the power series
   I_p(x) = sum_k (x/2)^{2k+p} / (k! (k+p)!)  (Abramowitz & Stegun 9.6.10),
   summed in log space for the prefactor; kept while K(p) > 1e-17 K(0). */
int orc_heat_kernel(double a, double* K, int maxlen)
{
    if (a <= 0.0) { K[0] = 1.0; return 1; }
    const double x = 2.0 * a, h = 0.5 * x;
    int M = 0;
    for (int p = 0; p < maxlen; ++p) {
        /* term_k = (x/2)^{2k+p}/(k!(k+p)!) e^{-x}, first term in logs */
        double lt = p * log(h) - lgamma(p + 1.0) - x;
        double t = exp(lt), sum = 0.0;
        for (int k = 0; k < 100000; ++k) {
            sum += t;
            t *= h * h / ((k + 1.0) * (k + 1.0 + p));
            if (t < 1e-20 * sum && k > h) break;
        }
        K[p] = sum;
        M = p;
        if (p > 0 && sum <= 1e-17 * K[0]) break;
    }
    return M + 1;
}

void orc_heat_apply(int nx, int ny, const double* K, int M, const double* in, double* out)
{
    double* tmp = (double*)malloc(sizeof(double) * (size_t)nx * ny);
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            double acc = 0.0;
            for (int p = -M; p <= M; ++p) {
                const int q = i - p;
                if (q >= 0 && q < nx) acc += K[abs(p)] * in[q + (size_t)nx * j];
            }
            tmp[i + (size_t)nx * j] = acc;
        }
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            double acc = 0.0;
            for (int p = -M; p <= M; ++p) {
                const int q = j - p;
                if (q >= 0 && q < ny) acc += K[abs(p)] * tmp[i + (size_t)nx * q];
            }
            out[i + (size_t)nx * j] = acc;
        }
    free(tmp);
}

/* S[:, q] = E H R e_q column by column (H = the filter K, or the identity) */
static void heat_schur(const orc_grid* g, int ddf, int N, const double* X, const double* Y, const double* ds,
                       const double* K, int M, double* S)
{
    int nx, ny;
    orc_dims(g, ORC_PRIMAL, &nx, &ny);
    const size_t ng = (size_t)nx * ny;
    double* e = (double*)calloc(N, sizeof(double));
    double* f = (double*)malloc(sizeof(double) * ng);
    double* h = (double*)malloc(sizeof(double) * ng);
    for (int q = 0; q < N; ++q) {
        memset(e, 0, sizeof(double) * N);
        e[q] = 1.0;
        orc_regularize(g, ORC_PRIMAL, ddf, N, X, Y, ds, e, f);
        if (K) orc_heat_apply(nx, ny, K, M, f, h);
        else memcpy(h, f, sizeof(double) * ng);
        orc_interpolate(g, ORC_PRIMAL, ddf, N, X, Y, h, S + (size_t)N * q);
    }
    free(e);
    free(f);
    free(h);
}

/* The same stage matrix assembled directly: entry (p, q) is the interpolation
   functional of p applied to the heat kernel K(|m|) K(|n|) (|m|, |n| <= M)
   convolved with the regularized unit strength of q, summed over the two
   points' local patches; pairs whose patches lie more than M cells apart are
   zero.  Equal to the column loop above whenever every patch and its kernel
   support lie inside the grid (the 5-cell margin of the solves). */
static void heat_schur_direct(const orc_grid* g, int ddf, int N, const double* X, const double* Y, const double* ds,
                              const double* K, int M, double* S)
{
    static const double one = 1.0;
    if (!K) { K = &one; M = 0; }
    patch* cols = (patch*)malloc(sizeof(patch) * (size_t)(N > 0 ? N : 1));
    patch* rows = (patch*)malloc(sizeof(patch) * (size_t)(N > 0 ? N : 1));
    #pragma omp parallel for
    for (int k = 0; k < N; ++k) {
        col_patch(g, ddf, X, Y, NULL, NULL, ds, k, ORC_S_RTLINVR, &cols[k]);
        row_patch(g, ddf, X, Y, NULL, NULL, k, ORC_S_RTLINVR, &rows[k]);
    }
    #pragma omp parallel
    {
        int ia[LOC * LOC], ja[LOC * LOC], ib[LOC * LOC], jb[LOC * LOC];
        double va[LOC * LOC], vb[LOC * LOC];
        #pragma omp for schedule(dynamic, 8)
        for (int q = 0; q < N; ++q) {
            int nb = 0;
            for (int j = 0; j < LOC; ++j)
                for (int i = 0; i < LOC; ++i) {
                    double v = cols[q].v[i + LOC * j];
                    if (v != 0.0) { ib[nb] = i + cols[q].off_x; jb[nb] = j + cols[q].off_y; vb[nb] = v; ++nb; }
                }
            for (int p = 0; p < N; ++p) {
                if (abs(rows[p].off_x - cols[q].off_x) > M + LOC || abs(rows[p].off_y - cols[q].off_y) > M + LOC) {
                    S[p + (size_t)N * q] = 0.0;
                    continue;
                }
                int na = 0;
                for (int j = 0; j < LOC; ++j)
                    for (int i = 0; i < LOC; ++i) {
                        double v = rows[p].v[i + LOC * j];
                        if (v != 0.0) { ia[na] = i + rows[p].off_x; ja[na] = j + rows[p].off_y; va[na] = v; ++na; }
                    }
                double acc = 0.0;
                for (int a = 0; a < na; ++a) {
                    double sum = 0.0;
                    for (int b = 0; b < nb; ++b) {
                        const int di = abs(ia[a] - ib[b]), dj = abs(ja[a] - jb[b]);
                        if (di <= M && dj <= M) sum += K[di] * K[dj] * vb[b];
                    }
                    acc += va[a] * sum;
                }
                S[p + (size_t)N * q] = acc;
            }
        }
    }
    free(cols);
    free(rows);
}

/* the stage matrix E H R of the kernel K (NULL: the identity) by the column
   loop (method 1) or the direct assembly (method 0), N x N column-major */
void orc_heat_schur(const orc_grid* g, int ddf, int N, const double* X, const double* Y, const double* ds,
                    const double* K, int M, int method, double* S)
{
    if (method == 1) heat_schur(g, ddf, N, X, Y, ds, K, M, S);
    else heat_schur_direct(g, ddf, N, X, Y, ds, K, M, S);
}

int orc_heat_advance(const orc_grid* g, int ddf, int N, const double* X, const double* Y, const double* nx_,
                     const double* ny_, const double* ds, double kappa, double dt, int nsteps, double* T,
                     const double* tb_plus, const double* tb_minus)
{
    /* tableau of the 3-stage HERK scheme: c = (0, 1/3, 1) */
    const double a21 = 1.0 / 3.0, a31 = -1.0, a32 = 2.0, b2 = 0.75, b3 = 0.25;
    int nx, ny;
    orc_dims(g, ORC_PRIMAL, &nx, &ny);
    const size_t ng = (size_t)nx * ny;
    const double a13 = kappa * dt / 3.0 / (g->dx * g->dx);
    double K13[512], K23[512];
    const int M13 = orc_heat_kernel(a13, K13, 512) - 1, M23 = orc_heat_kernel(2.0 * a13, K23, 512) - 1;
    double *S13 = NULL, *S23 = NULL, *S0 = NULL;
    int *p13 = NULL, *p23 = NULL, *p0 = NULL;
    int info = 0;
    if (N > 0) {
        S13 = (double*)malloc(sizeof(double) * (size_t)N * N);
        S23 = (double*)malloc(sizeof(double) * (size_t)N * N);
        S0 = (double*)malloc(sizeof(double) * (size_t)N * N);
        p13 = (int*)malloc(sizeof(int) * N);
        p23 = (int*)malloc(sizeof(int) * N);
        p0 = (int*)malloc(sizeof(int) * N);
        heat_schur_direct(g, ddf, N, X, Y, ds, K13, M13, S13);
        heat_schur_direct(g, ddf, N, X, Y, ds, K23, M23, S23);
        heat_schur_direct(g, ddf, N, X, Y, ds, NULL, 0, S0);
        info = orc_lu_factor(N, S13, p13) | orc_lu_factor(N, S23, p23) | orc_lu_factor(N, S0, p0);
    }
    /* r1 = surface_divergence(-kappa [T]), r2 = surface average */
    double* F = (double*)calloc(ng, sizeof(double));
    double* r2 = (double*)calloc(N > 0 ? N : 1, sizeof(double));
    double* dn = (double*)calloc(N > 0 ? N : 1, sizeof(double));
    for (int k = 0; k < N; ++k) {
        const double a = tb_plus[k], b = tb_minus ? tb_minus[k] : 0.0;
        dn[k] = -kappa * (a - b);
        r2[k] = 0.5 * (a + b);
    }
    if (N > 0) orc_surface_divergence(g, ddf, N, X, Y, nx_, ny_, ds, dn, F);
    double *A = (double*)malloc(sizeof(double) * ng), *Yh = (double*)malloc(sizeof(double) * ng);
    double *W1 = (double*)malloc(sizeof(double) * ng), *W2 = (double*)malloc(sizeof(double) * ng);
    double* Rs = (double*)malloc(sizeof(double) * ng);
    double* sg = (double*)malloc(sizeof(double) * (N > 0 ? N : 1));
    /* sigma = -(S^-1 (E Y - r2)) / (dt c);  W = F + R sigma (W may be NULL: Rs only) */
    #define MULT(Yv, S, P, c, W) do { \
        orc_interpolate(g, ORC_PRIMAL, ddf, N, X, Y, Yv, sg); \
        for (int k = 0; k < N; ++k) sg[k] = sg[k] - r2[k]; \
        orc_lu_solve(N, S, P, sg); \
        for (int k = 0; k < N; ++k) sg[k] *= -1.0 / (dt * (c)); \
        orc_regularize(g, ORC_PRIMAL, ddf, N, X, Y, ds, sg, Rs); \
        if (W) for (size_t k = 0; k < ng; ++k) ((double*)(W))[k] = F[k] + Rs[k]; \
    } while (0)
    for (int n = 0; n < nsteps; ++n) {
        if (N == 0) {
            orc_heat_apply(nx, ny, K13, M13, T, A);
            for (size_t k = 0; k < ng; ++k) A[k] += dt * b2 * F[k];
            orc_heat_apply(nx, ny, K23, M23, A, Yh);
            for (size_t k = 0; k < ng; ++k) T[k] = Yh[k] + dt * b3 * F[k];
            continue;
        }
        for (size_t k = 0; k < ng; ++k) A[k] = T[k] + dt * a21 * F[k];
        orc_heat_apply(nx, ny, K13, M13, A, Yh);
        MULT(Yh, S13, p13, a21, W1);
        for (size_t k = 0; k < ng; ++k) A[k] = T[k] + dt * a31 * W1[k];
        orc_heat_apply(nx, ny, K13, M13, A, Yh);
        for (size_t k = 0; k < ng; ++k) Yh[k] += dt * a32 * F[k];
        orc_heat_apply(nx, ny, K23, M23, Yh, A);
        MULT(A, S23, p23, a32, W2);
        orc_heat_apply(nx, ny, K13, M13, T, A);
        for (size_t k = 0; k < ng; ++k) A[k] += dt * b2 * W2[k];
        orc_heat_apply(nx, ny, K23, M23, A, Yh);
        for (size_t k = 0; k < ng; ++k) Yh[k] += dt * b3 * F[k];
        MULT(Yh, S0, p0, b3, (double*)NULL);
        for (size_t k = 0; k < ng; ++k) T[k] = Yh[k] + dt * b3 * Rs[k];
    }
    #undef MULT
    free(S13); free(S23); free(S0); free(p13); free(p23); free(p0);
    free(F); free(r2); free(dn); free(A); free(Yh); free(W1); free(W2); free(Rs); free(sg);
    return info;
}
