// ilh_stencil.hip -- staggered-grid MAC stencils (GridScaling).
//
// divergence!, grad!, curl! (both directions) and laplacian! of
// src/grid_operators.jl:25-134.  Each wrapper in the reference does fill!,
// the unscaled CartesianGrids stencil, then ./= dx; here that is one pass that
// writes every output cell exactly once (zeros where the reference leaves the
// fill! value).  Memory-bound: one read of each input cell (neighbour reads
// hit L1/L2 inside a 64x4 tile) and one write per output cell.
#include "ilh_internal.hpp"

namespace ilh {

static constexpr int BX = 64, BY = 4;

// p(i,j) = (u(i+1,j) - u(i,j) + v(i,j+1) - v(i,j)) / dx on all primal cells
__global__ void k_div(int NX, int NY, double dx, const double* __restrict__ u, const double* __restrict__ v,
                      double* __restrict__ p)
{
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    if (i >= NX - 1 || j >= NY - 1) return;
    const double d = (u[i + 1 + (size_t)NX * j] - u[i + (size_t)NX * j]) +
                     (v[i + (size_t)(NX - 1) * (j + 1)] - v[i + (size_t)(NX - 1) * j]);
    p[i + (size_t)(NX - 1) * j] = d / dx;
}

// u(i,j) = (p(i,j) - p(i-1,j)) / dx on interior x-faces, v likewise in y; 0
// elsewhere.  One pass over the NX x NY index range writes both components
// (u is NX x (NY-1), v is (NX-1) x NY), so each p value is fetched once.
__global__ void k_grad_uv(int NX, int NY, double dx, const double* __restrict__ p, double* __restrict__ u,
                          double* __restrict__ v)
{
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    if (i >= NX || j >= NY) return;
    const size_t np = NX - 1;
    const double pc = (i <= NX - 2 && j <= NY - 2) ? p[i + np * j] : 0.0;
    if (j <= NY - 2) {
        double r = 0.0;
        if (i >= 1 && i <= NX - 2) r = (pc - p[i - 1 + np * j]) / dx;
        u[i + (size_t)NX * j] = r;
    }
    if (i <= NX - 2) {
        double r = 0.0;
        if (j >= 1 && j <= NY - 2) r = (pc - p[i + np * (j - 1)]) / dx;
        v[i + np * j] = r;
    }
}

// u(i,j) = (s(i,j+1) - s(i,j)) / dx,  v(i,j) = (s(i,j) - s(i+1,j)) / dx, in one pass
__global__ void k_curl_n2e_uv(int NX, int NY, double dx, const double* __restrict__ s, double* __restrict__ u,
                              double* __restrict__ v)
{
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    if (i >= NX || j >= NY) return;
    const double sc = s[i + (size_t)NX * j];
    if (j <= NY - 2) u[i + (size_t)NX * j] = (s[i + (size_t)NX * (j + 1)] - sc) / dx;
    if (i <= NX - 2) v[i + (size_t)(NX - 1) * j] = (sc - s[i + 1 + (size_t)NX * j]) / dx;
}

// w(i,j) = ((v(i,j) - v(i-1,j)) - (u(i,j) - u(i,j-1))) / dx on interior dual nodes, ring 0
__global__ void k_curl_e2n(int NX, int NY, double dx, const double* __restrict__ u, const double* __restrict__ v,
                           double* __restrict__ w)
{
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    if (i >= NX || j >= NY) return;
    double r = 0.0;
    if (i >= 1 && i <= NX - 2 && j >= 1 && j <= NY - 2)
        r = ((v[i + (size_t)(NX - 1) * j] - v[i - 1 + (size_t)(NX - 1) * j]) -
             (u[i + (size_t)NX * j] - u[i + (size_t)NX * (j - 1)])) / dx;
    w[i + (size_t)NX * j] = r;
}

// 5-point Laplacian / dx^2 on interior nodes of an nx x ny node field, ring 0
__global__ void k_lap(int nx, int ny, double f, const double* __restrict__ s, double* __restrict__ w)
{
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    if (i >= nx || j >= ny) return;
    const size_t k = i + (size_t)nx * j;
    double r = 0.0;
    if (i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 2)
        r = f * (s[k - nx] + s[k - 1] - 4.0 * s[k] + s[k + 1] + s[k + nx]);
    w[k] = r;
}

static dim3 grid2(int nx, int ny) { return dim3((nx + BX - 1) / BX, (ny + BY - 1) / BY); }

void k_divergence(ilh_plan* pl, double* p, const double* q)
{
    const Grid& g = pl->g;
    k_div<<<grid2(g.NX - 1, g.NY - 1), dim3(BX, BY), 0, pl->ctx->stream>>>(g.NX, g.NY, g.dx, q, q + g.n(GK_EDGE_U), p);
    ILH_HIP(hipGetLastError());
}

void k_grad(ilh_plan* pl, double* q, const double* p)
{
    const Grid& g = pl->g;
    k_grad_uv<<<grid2(g.NX, g.NY), dim3(BX, BY), 0, pl->ctx->stream>>>(g.NX, g.NY, g.dx, p, q, q + g.n(GK_EDGE_U));
    ILH_HIP(hipGetLastError());
}

void k_curl_n2e(ilh_plan* pl, double* q, const double* s)
{
    const Grid& g = pl->g;
    k_curl_n2e_uv<<<grid2(g.NX, g.NY), dim3(BX, BY), 0, pl->ctx->stream>>>(g.NX, g.NY, g.dx, s, q,
                                                                           q + g.n(GK_EDGE_U));
    ILH_HIP(hipGetLastError());
}

void k_curl_e2n(ilh_plan* pl, double* w, const double* q)
{
    const Grid& g = pl->g;
    k_curl_e2n<<<grid2(g.NX, g.NY), dim3(BX, BY), 0, pl->ctx->stream>>>(g.NX, g.NY, g.dx, q, q + g.n(GK_EDGE_U), w);
    ILH_HIP(hipGetLastError());
}

void k_laplacian(ilh_plan* pl, int gk, double* w, const double* s)
{
    const Grid& g = pl->g;
    const int nx = g.nx(gk), ny = g.ny(gk);
    k_lap<<<grid2(nx, ny), dim3(BX, BY), 0, pl->ctx->stream>>>(nx, ny, 1.0 / (g.dx * g.dx), s, w);
    ILH_HIP(hipGetLastError());
}

__global__ void k_fill0(double* a, size_t n)
{
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
        a[k] = 0.0;
}

__global__ void k_axpby(double* y, const double* x, double a, double b, size_t n)
{
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
        y[k] = a * x[k] + b * y[k];
}

void fill_zero(hipStream_t st, double* a, size_t n)
{
    if (!n) return;
    size_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    k_fill0<<<(unsigned)blocks, 256, 0, st>>>(a, n);
    ILH_HIP(hipGetLastError());
}

void axpby(hipStream_t st, double* y, const double* x, double a, double b, size_t n)
{
    if (!n) return;
    size_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    k_axpby<<<(unsigned)blocks, 256, 0, st>>>(y, x, a, b, n);
    ILH_HIP(hipGetLastError());
}

}  // namespace ilh
