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

// Each kernel is instantiated for FP64 and FP32 grid data.  FP32 follows a
// Float32 cache of the reference: the unscaled difference in the data type,
// then the scaling by dx (a Float64 grid constant) in FP64, rounded once.
template <class T> __device__ __forceinline__ T scaled(T d, double s) { return (T)((double)d * s); }
template <class T> __device__ __forceinline__ T divided(T d, double dx) { return (T)((double)d / dx); }
// x positions per thread: two for FP32 (each of the thread's accesses is still
// a contiguous 256-byte run per wave), one for FP64
template <class T> constexpr int vx() { return sizeof(T) == 4 ? 2 : 1; }
template <class T> __device__ __forceinline__ int x_of(int r) { return (blockIdx.x * vx<T>() + r) * BX + threadIdx.x; }

// p(i,j) = (u(i+1,j) - u(i,j) + v(i,j+1) - v(i,j)) / dx on all primal cells
template <class T>
__global__ void k_div(int NX, int NY, double dx, const T* __restrict__ u, const T* __restrict__ v, T* __restrict__ p)
{
    const int j = blockIdx.y * BY + threadIdx.y;
    if (j >= NY - 1) return;
#pragma unroll
    for (int r = 0; r < vx<T>(); ++r) {
        const int i = x_of<T>(r);
        if (i >= NX - 1) break;
        const T d = (u[i + 1 + (size_t)NX * j] - u[i + (size_t)NX * j]) +
                    (v[i + (size_t)(NX - 1) * (j + 1)] - v[i + (size_t)(NX - 1) * j]);
        p[i + (size_t)(NX - 1) * j] = divided(d, dx);
    }
}

// u(i,j) = (p(i,j) - p(i-1,j)) / dx on interior x-faces, v likewise in y; 0
// elsewhere.  One pass over the NX x NY index range writes both components
// (u is NX x (NY-1), v is (NX-1) x NY), so each p value is fetched once.
template <class T>
__global__ void k_grad_uv(int NX, int NY, double dx, const T* __restrict__ p, T* __restrict__ u, T* __restrict__ v)
{
    const int j = blockIdx.y * BY + threadIdx.y;
    if (j >= NY) return;
    const size_t np = NX - 1;
#pragma unroll
    for (int rr = 0; rr < vx<T>(); ++rr) {
        const int i = x_of<T>(rr);
        if (i >= NX) break;
        const T pc = (i <= NX - 2 && j <= NY - 2) ? p[i + np * j] : T(0);
        if (j <= NY - 2) {
            T r = T(0);
            if (i >= 1 && i <= NX - 2) r = divided<T>(pc - p[i - 1 + np * j], dx);
            u[i + (size_t)NX * j] = r;
        }
        if (i <= NX - 2) {
            T r = T(0);
            if (j >= 1 && j <= NY - 2) r = divided<T>(pc - p[i + np * (j - 1)], dx);
            v[i + np * j] = r;
        }
    }
}

// u(i,j) = (s(i,j+1) - s(i,j)) / dx,  v(i,j) = (s(i,j) - s(i+1,j)) / dx, in one pass
template <class T>
__global__ void k_curl_n2e_uv(int NX, int NY, double dx, const T* __restrict__ s, T* __restrict__ u,
                              T* __restrict__ v)
{
    const int j = blockIdx.y * BY + threadIdx.y;
    if (j >= NY) return;
#pragma unroll
    for (int r = 0; r < vx<T>(); ++r) {
        const int i = x_of<T>(r);
        if (i >= NX) break;
        const T sc = s[i + (size_t)NX * j];
        if (j <= NY - 2) u[i + (size_t)NX * j] = divided<T>(s[i + (size_t)NX * (j + 1)] - sc, dx);
        if (i <= NX - 2) v[i + (size_t)(NX - 1) * j] = divided<T>(sc - s[i + 1 + (size_t)NX * j], dx);
    }
}

// w(i,j) = ((v(i,j) - v(i-1,j)) - (u(i,j) - u(i,j-1))) / dx on interior dual nodes, ring 0
template <class T>
__global__ void k_curl_e2n(int NX, int NY, double dx, const T* __restrict__ u, const T* __restrict__ v,
                           T* __restrict__ w)
{
    const int j = blockIdx.y * BY + threadIdx.y;
    if (j >= NY) return;
#pragma unroll
    for (int rr = 0; rr < vx<T>(); ++rr) {
        const int i = x_of<T>(rr);
        if (i >= NX) break;
        T r = T(0);
        if (i >= 1 && i <= NX - 2 && j >= 1 && j <= NY - 2)
            r = divided<T>((v[i + (size_t)(NX - 1) * j] - v[i - 1 + (size_t)(NX - 1) * j]) -
                               (u[i + (size_t)NX * j] - u[i + (size_t)NX * (j - 1)]),
                           dx);
        w[i + (size_t)NX * j] = r;
    }
}

// 5-point Laplacian / dx^2 on interior nodes of an nx x ny node field, ring 0
template <class T>
__global__ void k_lap(int nx, int ny, double f, const T* __restrict__ s, T* __restrict__ w)
{
    const int j = blockIdx.y * BY + threadIdx.y;
    if (j >= ny) return;
#pragma unroll
    for (int rr = 0; rr < vx<T>(); ++rr) {
        const int i = x_of<T>(rr);
        if (i >= nx) break;
        const size_t k = i + (size_t)nx * j;
        T r = T(0);
        if (i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 2)
            r = scaled<T>(s[k - nx] + s[k - 1] - T(4) * s[k] + s[k + 1] + s[k + nx], f);
        w[k] = r;
    }
}

template <class T> static dim3 grid2(int nx, int ny)
{
    constexpr int W = BX * vx<T>();
    return dim3((nx + W - 1) / W, (ny + BY - 1) / BY);
}

template <class T> static void divergence_t(ilh_plan* pl, T* p, const T* q)
{
    const Grid& g = pl->g;
    k_div<T><<<grid2<T>(g.NX - 1, g.NY - 1), dim3(BX, BY), 0, pl->ctx->stream>>>(g.NX, g.NY, g.dx, q,
                                                                             q + g.n(GK_EDGE_U), p);
    ILH_HIP(hipGetLastError());
}

template <class T> static void grad_t(ilh_plan* pl, T* q, const T* p)
{
    const Grid& g = pl->g;
    k_grad_uv<T><<<grid2<T>(g.NX, g.NY), dim3(BX, BY), 0, pl->ctx->stream>>>(g.NX, g.NY, g.dx, p, q,
                                                                         q + g.n(GK_EDGE_U));
    ILH_HIP(hipGetLastError());
}

template <class T> static void curl_n2e_t(ilh_plan* pl, T* q, const T* s)
{
    const Grid& g = pl->g;
    k_curl_n2e_uv<T><<<grid2<T>(g.NX, g.NY), dim3(BX, BY), 0, pl->ctx->stream>>>(g.NX, g.NY, g.dx, s, q,
                                                                             q + g.n(GK_EDGE_U));
    ILH_HIP(hipGetLastError());
}

template <class T> static void curl_e2n_t(ilh_plan* pl, T* w, const T* q)
{
    const Grid& g = pl->g;
    k_curl_e2n<T><<<grid2<T>(g.NX, g.NY), dim3(BX, BY), 0, pl->ctx->stream>>>(g.NX, g.NY, g.dx, q,
                                                                          q + g.n(GK_EDGE_U), w);
    ILH_HIP(hipGetLastError());
}

template <class T> static void laplacian_t(ilh_plan* pl, int gk, T* w, const T* s)
{
    const Grid& g = pl->g;
    const int nx = g.nx(gk), ny = g.ny(gk);
    k_lap<T><<<grid2<T>(nx, ny), dim3(BX, BY), 0, pl->ctx->stream>>>(nx, ny, 1.0 / (g.dx * g.dx), s, w);
    ILH_HIP(hipGetLastError());
}

void k_divergence(ilh_plan* pl, double* p, const double* q) { divergence_t(pl, p, q); }
void k_divergence(ilh_plan* pl, float* p, const float* q) { divergence_t(pl, p, q); }
void k_grad(ilh_plan* pl, double* q, const double* p) { grad_t(pl, q, p); }
void k_grad(ilh_plan* pl, float* q, const float* p) { grad_t(pl, q, p); }
void k_curl_n2e(ilh_plan* pl, double* q, const double* s) { curl_n2e_t(pl, q, s); }
void k_curl_n2e(ilh_plan* pl, float* q, const float* s) { curl_n2e_t(pl, q, s); }
void k_curl_e2n(ilh_plan* pl, double* w, const double* q) { curl_e2n_t(pl, w, q); }
void k_curl_e2n(ilh_plan* pl, float* w, const float* q) { curl_e2n_t(pl, w, q); }
void k_laplacian(ilh_plan* pl, int gk, double* w, const double* s) { laplacian_t(pl, gk, w, s); }
void k_laplacian(ilh_plan* pl, int gk, float* w, const float* s) { laplacian_t(pl, gk, w, s); }

template <class T>
__global__ void k_fill0(T* a, size_t n)
{
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
        a[k] = T(0);
}

__global__ void k_axpby(double* y, const double* x, double a, double b, size_t n)
{
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
        y[k] = a * x[k] + b * y[k];
}

template <class T> static void fill_zero_t(hipStream_t st, T* a, size_t n)
{
    if (!n) return;
    size_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    k_fill0<T><<<(unsigned)blocks, 256, 0, st>>>(a, n);
    ILH_HIP(hipGetLastError());
}

void fill_zero(hipStream_t st, double* a, size_t n) { fill_zero_t(st, a, n); }
void fill_zero(hipStream_t st, float* a, size_t n) { fill_zero_t(st, a, n); }

void axpby(hipStream_t st, double* y, const double* x, double a, double b, size_t n)
{
    if (!n) return;
    size_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    k_axpby<<<(unsigned)blocks, 256, 0, st>>>(y, x, a, b, n);
    ILH_HIP(hipGetLastError());
}

}  // namespace ilh
