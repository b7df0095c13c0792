// ilh_learned.hip -- a learned preconditioner for the Schur complement of
// moving bodies, trained and run on the GPU.
//
// When bodies move, update_system! (src/system.jl:26-50, via
// regenerate_problem, src/problem.jl:95-105) rebuilds R, E and S =
// create_RTLinvR (src/matrix_operators.jl:9-30) every step, and the
// reference then solves S x = b by LU.  S^-1 is concentrated near the diagonal
// along each body's curve (for the 1024^2 workload its entries fall from ~5e3
// on the diagonal to ~3e1 eight points away and ~0.4 sixteen away), so a band
// of 2K+1 entries per row, taken along the curve, is a good preconditioner for
// GMRES on S.  The band depends on the local geometry (the neighbours'
// positions in the point's normal frame, the points' positions within their
// grid cells, the normal and the panel length), and a network predicts it:
//
//   features (6(2K+1)+3) -> 256 -> 256 -> 256 -> 2K+1,  ReLU, BF16 weights
//   and activations, FP32 accumulation (v_mfma_f32_16x16x32_bf16),
//
// trained with Adam (FP32 master weights) on the bands of exact inverses that
// the plain method computes (each column of S^-1 is the solve of a unit
// right-hand side; one geometry of N points gives N training rows).  The
// prediction is the band of ds_p dx (S^-1)_{p,q}, standardized per output.
#include "ilh_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

namespace ilh {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float float4_t __attribute__((ext_vector_type(4)));

// ------------------------------------------------------------ BF16
__device__ __forceinline__ uint16_t f2bf(float f)
{
    // round to nearest, ties to even (finite inputs)
    unsigned u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((unsigned)h << 16); }

// ------------------------------------------------------------ GEMM
// C (M x N, row-major, ldc) = opA(A) opB(B) [+ bias[n]] [ReLU], A and B BF16
// row-major: opA(A)[m][k] = A[m lda + k] (TA = 0) or A[k lda + m] (TA = 1);
// opB(B)[k][n] = B[k ldb + n] (TB = 0) or B[n ldb + k] (TB = 1).  Outputs
// FP32 (Cf) and/or BF16 (Cb).  Split K: work-group z covers k in
// [z kchunk, (z+1) kchunk) and writes its partial product at Cf + z split_stride.
// 64 x 64 tile per work-group of four waves (32 x 32 each, 2 x 2 blocks of
// 16 x 16), K in chunks of 32 staged in LDS k-contiguous for both operands.
// v_mfma_f32_16x16x32_bf16 (checked on gfx950): lane l holds row l % 16 of
// the first operand and column l % 16 of the second, their k entries
// 8 (l / 16) .. +7; item r of the result is D[4 (l / 16) + r][l % 16].
static constexpr int BT = 64, BKC = 32, BKP = BKC + 8;

template <int TA, int TB>
__global__ void __launch_bounds__(256) k_bgemm(int M, int N, int K, int kchunk, const uint16_t* __restrict__ A,
                                               int lda, const uint16_t* __restrict__ B, int ldb,
                                               const float* __restrict__ bias, int relu, float* __restrict__ Cf,
                                               uint16_t* __restrict__ Cb, int ldc, size_t split_stride)
{
    __shared__ __align__(16) uint16_t As[BT][BKP];   // [m][k]
    __shared__ __align__(16) uint16_t Bs[BT][BKP];   // [n][k]
    const int m0 = blockIdx.x * BT, n0 = blockIdx.y * BT;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
    const int kbeg = blockIdx.z * kchunk, kend = min(K, kbeg + kchunk);
    float4_t acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = (float4_t){0.f, 0.f, 0.f, 0.f};
    for (int k0 = kbeg; k0 < kend; k0 += BKC) {
        // 64 x 32 of each operand: 8 entries per thread
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int idx = tid + 256 * e;
            int m, k;
            if (TA == 0) { m = idx / BKC; k = idx % BKC; }
            else { k = idx / BT; m = idx % BT; }
            const int gm = m0 + m, gk = k0 + k;
            uint16_t v = 0;
            if (gm < M && gk < kend) v = TA == 0 ? A[(size_t)gm * lda + gk] : A[(size_t)gk * lda + gm];
            As[m][k] = v;
            int n, kb;
            if (TB == 1) { n = idx / BKC; kb = idx % BKC; }
            else { kb = idx / BT; n = idx % BT; }
            const int gn = n0 + n, gkb = k0 + kb;
            uint16_t w = 0;
            if (gn < N && gkb < kend) w = TB == 1 ? B[(size_t)gn * ldb + gkb] : B[(size_t)gkb * ldb + gn];
            Bs[n][kb] = w;
        }
        __syncthreads();
        const int kq = 8 * (lane >> 4);
        bf16x8_t a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) a[i] = *(const bf16x8_t*)&As[wm + 16 * i + (lane & 15)][kq];
#pragma unroll
        for (int j = 0; j < 2; ++j) b[j] = *(const bf16x8_t*)&Bs[wn + 16 * j + (lane & 15)][kq];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }
    float* Cz = Cf ? Cf + (size_t)blockIdx.z * split_stride : nullptr;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gm = m0 + wm + 16 * i + 4 * (lane >> 4) + r;
                const int gn = n0 + wn + 16 * j + (lane & 15);
                if (gm < M && gn < N) {
                    float v = acc[i][j][r];
                    if (bias) v += bias[gn];
                    if (relu) v = v > 0.f ? v : 0.f;
                    if (Cz) Cz[(size_t)gm * ldc + gn] = v;
                    if (Cb) Cb[(size_t)gm * ldc + gn] = f2bf(v);
                }
            }
}

// k_bgemm<TA, TB> for K slices of at most 256 (every forward layer, the
// weight gradients split over 256-row slices, the input gradients): the
// work-group's whole 64 x K slices of both operands go out in one round of
// 16-byte loads (eight BF16 each), an operand whose memory order is not
// [row][k] transposed into LDS on the way, and the K/32 MFMA steps then run
// from LDS without further barriers.  The MFMA sequence and the epilogue are
// k_bgemm's, so the results are its results bit for bit; k_bgemm waits for
// one load round per 32-wide step.  Needs 16-byte aligned operands, leading
// dimensions and the 16-byte-loaded extents (K or M, N) multiples of 8.
static constexpr int FK = 256, FKP = FK + 8;

// one operand's 64 x KL slice into S[row][k] (rows r0.., k from kbeg): KT = 0
// for memory [row][k] (16-byte loads along k), 1 for [k][row] (along rows,
// transposed into LDS)
template <int KT>
__device__ __forceinline__ void wide_load(const uint16_t* __restrict__ X, int ld, int R, int r0, int kbeg, int KL,
                                          int tid, uint4 (&x)[8])
{
    const int kv = KL >> 3;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int v = tid + 256 * e;
        if (KT == 0) {
            const int r = v / kv, k8 = (v - r * kv) * 8;
            x[e] = v < BT * kv && r0 + r < R ? *(const uint4*)&X[(size_t)(r0 + r) * ld + kbeg + k8]
                                             : make_uint4(0u, 0u, 0u, 0u);
        } else {
            const int k = v >> 3, r8 = (v & 7) * 8;
            x[e] = v < KL * 8 && r0 + r8 < R ? *(const uint4*)&X[(size_t)(kbeg + k) * ld + r0 + r8]
                                             : make_uint4(0u, 0u, 0u, 0u);
        }
    }
}

template <int KT>
__device__ __forceinline__ void wide_store(uint16_t (*S)[FKP], int KL, int tid, const uint4 (&x)[8])
{
    const int kv = KL >> 3;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int v = tid + 256 * e;
        if (KT == 0) {
            if (v < BT * kv) {
                const int r = v / kv, k8 = (v - r * kv) * 8;
                *(uint4*)&S[r][k8] = x[e];
            }
        } else if (v < KL * 8) {
            const int k = v >> 3, r8 = (v & 7) * 8;
            const uint16_t* h = (const uint16_t*)&x[e];
#pragma unroll
            for (int i = 0; i < 8; ++i) S[r8 + i][k] = h[i];
        }
    }
}

template <int TA, int TB>
__global__ void __launch_bounds__(256) k_bgemm_w(int M, int N, int K, int kchunk, const uint16_t* __restrict__ A,
                                                 int lda, const uint16_t* __restrict__ B, int ldb,
                                                 const float* __restrict__ bias, int relu, float* __restrict__ Cf,
                                                 uint16_t* __restrict__ Cb, int ldc, size_t split_stride)
{
    __shared__ __align__(16) uint16_t As[BT][FKP];   // [m][k]
    __shared__ __align__(16) uint16_t Bs[BT][FKP];   // [n][k]
    const int m0 = blockIdx.x * BT, n0 = blockIdx.y * BT;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
    // this split's K slice (empty past the end: the partial sums are then zeros)
    const int kbeg = blockIdx.z * kchunk, KL = max(0, min(K, kbeg + kchunk) - kbeg), K32 = (KL + BKC - 1) / BKC * BKC;
    uint4 xa[8], xb[8];
    wide_load<TA == 0 ? 0 : 1>(A, lda, M, m0, kbeg, KL, tid, xa);
    wide_load<TB == 1 ? 0 : 1>(B, ldb, N, n0, kbeg, KL, tid, xb);
    wide_store<TA == 0 ? 0 : 1>(As, KL, tid, xa);
    wide_store<TB == 1 ? 0 : 1>(Bs, KL, tid, xb);
    for (int e = tid; e < BT * (K32 - KL); e += 256) {   // zeros for k in [KL, K32)
        const int r = e / (K32 - KL), k = KL + e % (K32 - KL);
        As[r][k] = 0;
        Bs[r][k] = 0;
    }
    __syncthreads();
    float4_t acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = (float4_t){0.f, 0.f, 0.f, 0.f};
    const int kq = 8 * (lane >> 4);
    for (int k0 = 0; k0 < K32; k0 += BKC) {
        bf16x8_t a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) a[i] = *(const bf16x8_t*)&As[wm + 16 * i + (lane & 15)][k0 + kq];
#pragma unroll
        for (int j = 0; j < 2; ++j) b[j] = *(const bf16x8_t*)&Bs[wn + 16 * j + (lane & 15)][k0 + kq];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    float* Cz = Cf ? Cf + (size_t)blockIdx.z * split_stride : nullptr;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gm = m0 + wm + 16 * i + 4 * (lane >> 4) + r;
                const int gn = n0 + wn + 16 * j + (lane & 15);
                if (gm < M && gn < N) {
                    float v = acc[i][j][r];
                    if (bias) v += bias[gn];
                    if (relu) v = v > 0.f ? v : 0.f;
                    if (Cz) Cz[(size_t)gm * ldc + gn] = v;
                    if (Cb) Cb[(size_t)gm * ldc + gn] = f2bf(v);
                }
            }
}

static void bgemm(hipStream_t st, int ta, int tb, int M, int N, int K, const uint16_t* A, int lda, const uint16_t* B,
                  int ldb, const float* bias, int relu, float* Cf, uint16_t* Cb, int ldc, int splits = 1,
                  size_t split_stride = 0)
{
    if (M <= 0 || N <= 0) return;
    const int kchunk = ((K + splits - 1) / splits + BKC - 1) / BKC * BKC;
    const dim3 grid((M + BT - 1) / BT, (N + BT - 1) / BT, splits);
    // the one-round form: every K slice at most 256, the 16-byte-loaded
    // extents (K along [row][k] operands, M or N along [k][row] ones) and the
    // leading dimensions multiples of 8, both operands 16-byte aligned
    const bool wide = K > 0 && std::min(K, kchunk) <= FK && lda % 8 == 0 && ldb % 8 == 0 &&
                      (((uintptr_t)A | (uintptr_t)B) & 15) == 0 && (ta == 1 ? M % 8 == 0 : K % 8 == 0) &&
                      (tb == 0 ? N % 8 == 0 : K % 8 == 0);
    if (wide && ta == 0 && tb == 0)
        k_bgemm_w<0, 0><<<grid, 256, 0, st>>>(M, N, K, kchunk, A, lda, B, ldb, bias, relu, Cf, Cb, ldc, split_stride);
    else if (wide && ta == 0 && tb == 1)
        k_bgemm_w<0, 1><<<grid, 256, 0, st>>>(M, N, K, kchunk, A, lda, B, ldb, bias, relu, Cf, Cb, ldc, split_stride);
    else if (wide && ta == 1 && tb == 0)
        k_bgemm_w<1, 0><<<grid, 256, 0, st>>>(M, N, K, kchunk, A, lda, B, ldb, bias, relu, Cf, Cb, ldc, split_stride);
    else if (ta == 0 && tb == 0)
        k_bgemm<0, 0><<<grid, 256, 0, st>>>(M, N, K, kchunk, A, lda, B, ldb, bias, relu, Cf, Cb, ldc, split_stride);
    else if (ta == 0 && tb == 1)
        k_bgemm<0, 1><<<grid, 256, 0, st>>>(M, N, K, kchunk, A, lda, B, ldb, bias, relu, Cf, Cb, ldc, split_stride);
    else if (ta == 1 && tb == 0)
        k_bgemm<1, 0><<<grid, 256, 0, st>>>(M, N, K, kchunk, A, lda, B, ldb, bias, relu, Cf, Cb, ldc, split_stride);
    else
        fail(ILH_ERR_ARG, "bgemm: unsupported transposes");
    ILH_HIP(hipGetLastError());
}

// ------------------------------------------------------------ helpers
// out[e] = sum_{z < nz} part[z stride + e] (fixed order), and optionally the
// column sums of dz (rows x cols, row-major) into bsum (one thread per column)
__global__ void k_sum_splits(const float* __restrict__ part, int nz, size_t stride, size_t n, float* __restrict__ out)
{
    const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (e >= n) return;
    float s = 0.f;
    for (int z = 0; z < nz; ++z) s += part[z * stride + e];
    out[e] = s;
}

// column sums of a BF16 matrix (the bias gradients): a work-group of 1024
// threads per 32 columns, thread (g, c) summing rows g, g + 32, ... of column
// c, then the 32 partial sums of each column added in a fixed order
// (deterministic)
__global__ void __launch_bounds__(1024) k_colsum_bf(const uint16_t* __restrict__ X, int rows, int cols, int ld,
                                                    float* __restrict__ out)
{
    __shared__ float part[32][33];
    const int cl = threadIdx.x & 31, g = threadIdx.x >> 5;
    const int c = blockIdx.x * 32 + cl;
    float s = 0.f;
    if (c < cols)
        for (int r = g; r < rows; r += 32) s += bf2f(X[(size_t)r * ld + c]);
    part[g][cl] = s;
    __syncthreads();
    if (g == 0 && c < cols) {
        float t = 0.f;
        for (int k = 0; k < 32; ++k) t += part[k][cl];
        out[c] = t;
    }
}

// k_colsum_bf's sums over more of the GPU, bit for bit: thread (g, c8) of
// row class g = blockIdx.y adds rows g, g + 32, ... of columns c8 .. c8 + 7
// (16-byte loads; each column's additions are k_colsum_bf's thread (g, c)'s,
// in the same order) into part[g][c]; k_colsum_fin then adds the 32 partial
// sums of each column in g order.  Needs cols and ld multiples of 8 and X
// 16-byte aligned.
__global__ void __launch_bounds__(64) k_colsum_part(const uint16_t* __restrict__ X, int rows, int cols, int ld,
                                                    float* __restrict__ part)
{
    const int c8 = (blockIdx.x * 64 + threadIdx.x) * 8, g = blockIdx.y;
    if (c8 >= cols) return;
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int r = g; r < rows; r += 32) {
        const uint4 v = *(const uint4*)&X[(size_t)r * ld + c8];
        const uint16_t* h = (const uint16_t*)&v;
#pragma unroll
        for (int i = 0; i < 8; ++i) s[i] += bf2f(h[i]);
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) part[(size_t)g * cols + c8 + i] = s[i];
}

__global__ void k_colsum_fin(const float* __restrict__ part, int cols, float* __restrict__ out)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    float t = 0.f;
    for (int k = 0; k < 32; ++k) t += part[(size_t)k * cols + c];
    out[c] = t;
}

// dZ = dA * (A > 0), BF16 (A the ReLU output of the layer)
__global__ void k_relu_bwd(const float* __restrict__ dA, const uint16_t* __restrict__ Aact, size_t n,
                           uint16_t* __restrict__ dZ)
{
    const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (e >= n) return;
    dZ[e] = bf2f(Aact[e]) > 0.f ? f2bf(dA[e]) : (uint16_t)0;
}

__global__ void k_to_bf(const float* __restrict__ a, size_t n, uint16_t* __restrict__ b)
{
    const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (e < n) b[e] = f2bf(a[e]);
}

// loss = mean over rows x O of (Y - T)^2 (per-block partial sums), dY = 2 (Y - T) / (rows O) in BF16 (0 in the
// padded columns)
__global__ void __launch_bounds__(256) k_loss(const float* __restrict__ Y, const float* __restrict__ T, int rows,
                                              int O, int ld, uint16_t* __restrict__ dY, double* __restrict__ part)
{
    __shared__ double red[256];
    double s = 0.0;
    const size_t n = (size_t)rows * ld;
    const float sc = 2.f / ((float)rows * (float)O);
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
        const int c = (int)(e % ld);
        float d = 0.f;
        if (c < O) {
            d = Y[e] - T[e];
            s += (double)d * d;
        }
        dY[e] = f2bf(sc * d);
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m1,
                       float* __restrict__ m2, size_t n, float lr, float b1, float b2, float c1, float c2, float eps)
{
    const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (e >= n) return;
    const float gg = g[e];
    const float a = b1 * m1[e] + (1.f - b1) * gg;
    const float b = b2 * m2[e] + (1.f - b2) * gg * gg;
    m1[e] = a;
    m2[e] = b;
    w[e] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

// Per-point features (6W + 3 values, W = 2K + 1) for the neighbours nbr[p W + j]:
// offsets in grid units in the point's (tangent, normal) frame, sin/cos of
// 2 pi times the neighbours' fractional primal-grid coordinates, the normal,
// ds / dx.
__global__ void k_lp_features(int N, int W, Grid g, const double* __restrict__ X, const double* __restrict__ Y,
                              const double* __restrict__ nxv, const double* __restrict__ nyv,
                              const double* __restrict__ ds, const int* __restrict__ nbr, float* __restrict__ F,
                              int ldf)
{
    // one thread per (point, neighbour): the neighbour's six features (four
    // FP64 sines and cosines), and the point's own three for neighbour 0
    const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (e >= (size_t)N * W) return;
    const int p = (int)(e / W), j = (int)(e % W);
    const double dx = g.dx, pi2 = 6.283185307179586476925286766559;
    const double xp = X[p] / dx, yp = Y[p] / dx, nx = nxv[p], ny = nyv[p];
    float* f = F + (size_t)p * ldf;
    {
        const int q = nbr[(size_t)p * W + j];
        const double xq = X[q] / dx, yq = Y[q] / dx;
        const double ddx = xq - xp, ddy = yq - yp;
        f[j] = (float)(-ddx * ny + ddy * nx);
        f[W + j] = (float)(ddx * nx + ddy * ny);
        double fx = xq - (1.0 - g.I0x), fy = yq - (1.0 - g.I0y);
        fx -= floor(fx);
        fy -= floor(fy);
        f[2 * W + j] = (float)sin(pi2 * fx);
        f[3 * W + j] = (float)cos(pi2 * fx);
        f[4 * W + j] = (float)sin(pi2 * fy);
        f[5 * W + j] = (float)cos(pi2 * fy);
    }
    if (j == 0) {
        f[6 * W] = (float)nx;
        f[6 * W + 1] = (float)ny;
        f[6 * W + 2] = (float)(ds[p] / dx);
    }
}

// column mean and standard deviation (FP64 sums, one work-group per column)
__global__ void __launch_bounds__(256) k_colstats(const float* __restrict__ X, int rows, int cols, int ld,
                                                  float* __restrict__ mu, float* __restrict__ sd)
{
    __shared__ double r1[256], r2[256];
    const int c = blockIdx.x;
    double s = 0.0, s2 = 0.0;
    for (int r = threadIdx.x; r < rows; r += 256) {
        const double v = X[(size_t)r * ld + c];
        s += v;
        s2 += v * v;
    }
    r1[threadIdx.x] = s;
    r2[threadIdx.x] = s2;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            r1[threadIdx.x] += r1[threadIdx.x + off];
            r2[threadIdx.x] += r2[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double m = r1[0] / rows;
        const double var = fmax(r2[0] / rows - m * m, 0.0);
        mu[c] = (float)m;
        sd[c] = (float)(sqrt(var) + 1e-6);
    }
}

// (raw - mu) / sd -> BF16 rows of width ldo (zero padding beyond cols); rows
// picked through idx (or 0..rows-1)
__global__ void k_standardize_bf(const float* __restrict__ raw, int ldr, const int* __restrict__ idx, int rows,
                                 int cols, const float* __restrict__ mu, const float* __restrict__ sd, int ldo,
                                 uint16_t* __restrict__ out)
{
    const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (e >= (size_t)rows * ldo) return;
    const int r = (int)(e / ldo), c = (int)(e % ldo);
    const int src = idx ? idx[r] : r;
    out[e] = c < cols ? f2bf((raw[(size_t)src * ldr + c] - mu[c]) / sd[c]) : (uint16_t)0;
}

// standardized targets of the picked rows (FP32, width ldo, zero padding)
__global__ void k_standardize_f(const float* __restrict__ raw, int ldr, const int* __restrict__ idx, int rows,
                                int cols, const float* __restrict__ mu, const float* __restrict__ sd, int ldo,
                                float* __restrict__ out)
{
    const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (e >= (size_t)rows * ldo) return;
    const int r = (int)(e / ldo), c = (int)(e % ldo);
    const int src = idx ? idx[r] : r;
    out[e] = c < cols ? (raw[(size_t)src * ldr + c] - mu[c]) / sd[c] : 0.f;
}

// targets: T[p][j] = ds_p dx (S^-1)_{p, nbr[p][j]} from the row-major inverse
__global__ void k_band_targets(const double* __restrict__ Sinv, int ld, int N, int W, const int* __restrict__ nbr,
                               const double* __restrict__ ds, double dx, float* __restrict__ T)
{
    const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (e >= (size_t)N * W) return;
    const int p = (int)(e / W);
    T[e] = (float)(Sinv[(size_t)p * ld + nbr[e]] * ds[p] * dx);
}

// band coefficients from the standardized prediction Y (ldy wide)
__global__ void k_band_coef(const float* __restrict__ Y, int ldy, int N, int W, const float* __restrict__ mu,
                            const float* __restrict__ sd, const double* __restrict__ ds, double dx,
                            double* __restrict__ coef)
{
    const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (e >= (size_t)N * W) return;
    const int p = (int)(e / W), j = (int)(e % W);
    coef[e] = ((double)Y[(size_t)p * ldy + j] * sd[j] + mu[j]) / (ds[p] * dx);
}

// z (+)= band(coef) v, skipped once the flags say the solve has converged;
// one wave per row
__global__ void __launch_bounds__(256) k_band_apply(const double* __restrict__ coef, const int* __restrict__ nbr, int W,
                             const double* __restrict__ v, double* __restrict__ z, int n, const int* __restrict__ flags,
                             int accumulate)
{
    if (flags && flags[0]) return;
    const int p = blockIdx.x * 4 + (int)(threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (p >= n) return;
    const double s = band_row(coef + (size_t)p * W, nbr + (size_t)p * W, W, v, lane);
    if (lane == 0) z[p] = accumulate ? z[p] + s : s;
}

void band_apply(hipStream_t st, const double* coef, const int* nbr, int W, const double* v, double* z, int n,
                const int* flags, int accumulate)
{
    if (n <= 0) return;
    k_band_apply<<<(n + 3) / 4, 256, 0, st>>>(coef, nbr, W, v, z, n, flags, accumulate);
    ILH_HIP(hipGetLastError());
}

// neighbours along the curve: the body's points cyclically, j - K .. j + K
void plan_neighbours(ilh_plan* pl, int K)
{
    const int N = pl->N, W = 2 * K + 1;
    if (pl->nbr_K == K && pl->nbr_gen == pl->body_gen) return;
    std::vector<int> off = pl->body_off;
    if (off.empty()) off = {0, N};
    std::vector<int> h((size_t)N * W);
    for (size_t b = 0; b + 1 < off.size(); ++b) {
        const int s = off[b], m = off[b + 1] - off[b];
        for (int a = 0; a < m; ++a)
            for (int j = 0; j < W; ++j) h[(size_t)(s + a) * W + j] = s + (((a + j - K) % m) + m) % m;
    }
    pl->nbr.ensure(sizeof(int) * h.size());
    ILH_HIP(hipMemcpy(pl->nbr.p, h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice));
    pl->nbr_K = K;
    pl->nbr_gen = pl->body_gen;
}

}  // namespace ilh

using namespace ilh;

// ------------------------------------------------------------ the model
struct ilh_lpc {
    ilh_ctx* ctx = nullptr;
    int K = 16, W = 33, F = 201, Fp = 224, H = 256, Op = 64;
    // parameters (FP32 masters): W1 Fp x H, b1 H, W2 H x H, b2, W3 H x H, b3, W4 H x Op, b4 Op
    size_t off_w[4] = {}, off_b[4] = {}, nparam = 0;
    int rows_in[4] = {}, cols_out[4] = {};
    DevBuf P, G, M1, M2, Pb;          // params, gradients, Adam moments, BF16 weights (same offsets as P)
    int adam_t = 0;
    // training set: raw features (rows x F) and raw targets (rows x W)
    std::vector<float> hX, hT;
    int rows = 0;
    DevBuf dX, dT;
    bool stats_ok = false;
    DevBuf mux, sdx, muy, sdy;        // standardization (frozen at the first training step)
    // work buffers
    DevBuf act[4], zf, dA, dZ, Yf, Tf, part, idx, splitbuf, feat, Xb, band;
    DevBuf colpart;                   // the bias gradients' partial column sums (32 x columns)
    size_t cap_rows = 0;
};

namespace {

void lpc_layout(ilh_lpc* m)
{
    const int in[4] = {m->Fp, m->H, m->H, m->H}, out[4] = {m->H, m->H, m->H, m->Op};
    size_t o = 0;
    for (int l = 0; l < 4; ++l) {
        m->rows_in[l] = in[l];
        m->cols_out[l] = out[l];
        m->off_w[l] = o;
        o += (size_t)in[l] * out[l];
        m->off_b[l] = o;
        o += out[l];
    }
    m->nparam = o;
}

void lpc_refresh_bf16(ilh_lpc* m)
{
    k_to_bf<<<(unsigned)((m->nparam + 255) / 256), 256, 0, m->ctx->stream>>>(m->P.as<float>(), m->nparam,
                                                                               m->Pb.as<uint16_t>());
    ILH_HIP(hipGetLastError());
}

void lpc_work(ilh_lpc* m, size_t rows)
{
    if (rows <= m->cap_rows) return;
    for (int l = 0; l < 4; ++l) m->act[l].ensure(sizeof(uint16_t) * rows * std::max(m->Fp, m->H));
    m->zf.ensure(sizeof(float) * rows * m->H);
    m->dA.ensure(sizeof(float) * rows * std::max(m->Fp, m->H));
    m->dZ.ensure(sizeof(uint16_t) * rows * m->H);
    m->Yf.ensure(sizeof(float) * rows * m->Op);
    m->Tf.ensure(sizeof(float) * rows * m->Op);
    m->idx.ensure(sizeof(int) * rows);
    m->cap_rows = rows;
}

// forward pass of `rows` standardized BF16 inputs act[0] (rows x Fp): ReLU
// activations act[1..3] (BF16, rows x H), prediction Yf (FP32, rows x Op)
void lpc_forward(ilh_lpc* m, int rows)
{
    hipStream_t st = m->ctx->stream;
    const uint16_t* Wb = m->Pb.as<uint16_t>();
    const float* Pf = m->P.as<float>();
    for (int l = 0; l < 3; ++l)
        bgemm(st, 0, 0, rows, m->H, m->rows_in[l], m->act[l].as<uint16_t>(), m->rows_in[l], Wb + m->off_w[l], m->H,
              Pf + m->off_b[l], 1, nullptr, m->act[l + 1].as<uint16_t>(), m->H);
    bgemm(st, 0, 0, rows, m->Op, m->H, m->act[3].as<uint16_t>(), m->H, Wb + m->off_w[3], m->Op, Pf + m->off_b[3], 0,
          m->Yf.as<float>(), nullptr, m->Op);
}

// loss of the prediction against Tf and all parameter gradients into G
double lpc_backward(ilh_lpc* m, int rows)
{
    hipStream_t st = m->ctx->stream;
    const uint16_t* Wb = m->Pb.as<uint16_t>();
    float* Gf = m->G.as<float>();
    const int nblk = 256;
    m->part.ensure(sizeof(double) * nblk);
    m->dZ.ensure(sizeof(uint16_t) * (size_t)rows * std::max(m->H, m->Op));
    // dY (BF16, rows x Op) into dZ
    k_loss<<<nblk, 256, 0, st>>>(m->Yf.as<float>(), m->Tf.as<float>(), rows, m->W, m->Op, m->dZ.as<uint16_t>(),
                                 m->part.as<double>());
    ILH_HIP(hipGetLastError());
    const int splits = std::max(1, std::min(16, rows / 256));
    for (int l = 3; l >= 0; --l) {
        const int in = m->rows_in[l], out = m->cols_out[l];
        // dW = act[l]^T dZ  (in x out), split over the rows, partial sums reduced in order
        const size_t nW = (size_t)in * out;
        m->splitbuf.ensure(sizeof(float) * nW * splits);
        bgemm(st, 1, 0, in, out, rows, m->act[l].as<uint16_t>(), in, m->dZ.as<uint16_t>(), out, nullptr, 0,
              m->splitbuf.as<float>(), nullptr, out, splits, nW);
        k_sum_splits<<<(unsigned)((nW + 255) / 256), 256, 0, st>>>(m->splitbuf.as<float>(), splits, nW, nW,
                                                                    Gf + m->off_w[l]);
        if (out % 8 == 0 && ((uintptr_t)m->dZ.p & 15) == 0) {
            m->colpart.ensure(sizeof(float) * 32 * (size_t)out);
            k_colsum_part<<<dim3((unsigned)((out / 8 + 63) / 64), 32), 64, 0, st>>>(m->dZ.as<uint16_t>(), rows, out,
                                                                                    out, m->colpart.as<float>());
            k_colsum_fin<<<(unsigned)((out + 255) / 256), 256, 0, st>>>(m->colpart.as<float>(), out,
                                                                       Gf + m->off_b[l]);
        } else {
            k_colsum_bf<<<(out + 31) / 32, 1024, 0, st>>>(m->dZ.as<uint16_t>(), rows, out, out, Gf + m->off_b[l]);
        }
        ILH_HIP(hipGetLastError());
        if (l > 0) {
            // dA = dZ W^T (rows x in), then dZ = dA * (act[l] > 0)
            bgemm(st, 0, 1, rows, in, out, m->dZ.as<uint16_t>(), out, Wb + m->off_w[l], out, nullptr, 0,
                  m->dA.as<float>(), nullptr, in);
            const size_t n = (size_t)rows * in;
            k_relu_bwd<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(m->dA.as<float>(), m->act[l].as<uint16_t>(), n,
                                                                    m->dZ.as<uint16_t>());
            ILH_HIP(hipGetLastError());
        }
    }
    std::vector<double> hp(nblk);
    ILH_HIP(hipMemcpyAsync(hp.data(), m->part.p, sizeof(double) * nblk, hipMemcpyDeviceToHost, st));
    ILH_HIP(hipStreamSynchronize(st));
    double s = 0.0;
    for (double v : hp) s += v;
    return s / ((double)rows * m->W);
}

void lpc_stats(ilh_lpc* m)
{
    if (m->stats_ok) return;
    if (m->rows < 2) fail(ILH_ERR_STATE, "the learned preconditioner has no training rows");
    hipStream_t st = m->ctx->stream;
    m->mux.ensure(sizeof(float) * m->F);
    m->sdx.ensure(sizeof(float) * m->F);
    m->muy.ensure(sizeof(float) * m->W);
    m->sdy.ensure(sizeof(float) * m->W);
    k_colstats<<<m->F, 256, 0, st>>>(m->dX.as<float>(), m->rows, m->F, m->F, m->mux.as<float>(), m->sdx.as<float>());
    k_colstats<<<m->W, 256, 0, st>>>(m->dT.as<float>(), m->rows, m->W, m->W, m->muy.as<float>(), m->sdy.as<float>());
    ILH_HIP(hipGetLastError());
    m->stats_ok = true;
}

// features of the plan's current geometry (N x F raw, FP32) into m->feat
void lpc_features(ilh_lpc* m, ilh_plan* pl)
{
    plan_neighbours(pl, m->K);
    const int N = pl->N;
    m->feat.ensure(sizeof(float) * (size_t)std::max(N, 1) * m->F);
    k_lp_features<<<(unsigned)(((size_t)N * m->W + 255) / 256), 256, 0, pl->ctx->stream>>>(N, m->W, pl->g, pl->X.as<double>(), pl->Y.as<double>(),
                                                                pl->nx.as<double>(), pl->ny.as<double>(),
                                                                pl->ds.as<double>(), pl->nbr.as<int>(),
                                                                m->feat.as<float>(), m->F);
    ILH_HIP(hipGetLastError());
}

void check_model(ilh_lpc* m, ilh_plan* pl)
{
    if (!m) fail(ILH_ERR_ARG, "null model");
    if (pl && pl->ctx != m->ctx) fail(ILH_ERR_ARG, "plan and model belong to different contexts");
}

}  // namespace

namespace ilh {
// band coefficients (N x W) of the plan's current geometry predicted by the model
void lpc_band(ilh_lpc* m, ilh_plan* pl, double* coef)
{
    if (!m->stats_ok) fail(ILH_ERR_STATE, "the learned preconditioner is not trained");
    const int N = pl->N;
    hipStream_t st = pl->ctx->stream;
    lpc_features(m, pl);
    lpc_work(m, (size_t)N);
    const size_t n = (size_t)N * m->Fp;
    k_standardize_bf<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(m->feat.as<float>(), m->F, nullptr, N, m->F,
                                                                   m->mux.as<float>(), m->sdx.as<float>(), m->Fp,
                                                                   m->act[0].as<uint16_t>());
    ILH_HIP(hipGetLastError());
    lpc_forward(m, N);
    const size_t nw = (size_t)N * m->W;
    k_band_coef<<<(unsigned)((nw + 255) / 256), 256, 0, st>>>(m->Yf.as<float>(), m->Op, N, m->W, m->muy.as<float>(),
                                                              m->sdy.as<float>(), pl->ds.as<double>(), pl->g.dx, coef);
    ILH_HIP(hipGetLastError());
}
int lpc_width(const ilh_lpc* m) { return m->W; }
int lpc_half_band(const ilh_lpc* m) { return m->K; }
}  // namespace ilh

extern "C" int ilh_plan_set_bodies(ilh_plan* pl, int nbodies, const int* offsets)
{
    return guarded_call([&] {
        if (!pl) fail(ILH_ERR_ARG, "null plan");
        if (nbodies < 1 || !offsets) fail(ILH_ERR_ARG, "need at least one body and its offsets");
        std::vector<int> off(offsets, offsets + nbodies + 1);
        if (off[0] != 0 || off[nbodies] != pl->N) fail(ILH_ERR_DIM, "body offsets must run from 0 to N");
        for (int b = 0; b < nbodies; ++b)
            if (off[b + 1] - off[b] < 1) fail(ILH_ERR_ARG, "empty body");
        pl->body_off = off;
        ++pl->body_gen;
    });
}

extern "C" int ilh_lpc_create(ilh_ctx* ctx, int half_band, unsigned long long seed, ilh_lpc** out)
{
    return guarded_call([&] {
        if (!ctx || !out) fail(ILH_ERR_ARG, "null argument");
        if (half_band < 1 || half_band > 31) fail(ILH_ERR_ARG, "half band must be in [1, 31]");
        ILH_HIP(hipSetDevice(ctx->device));
        ilh_lpc* m = new ilh_lpc();
        try {
            m->ctx = ctx;
            m->K = half_band;
            m->W = 2 * half_band + 1;
            m->F = 6 * m->W + 3;
            m->Fp = (m->F + 31) / 32 * 32;
            m->Op = (m->W + 63) / 64 * 64;
            lpc_layout(m);
            // He-uniform weights from a seeded generator, zero biases; the
            // padded input rows and output columns stay zero
            std::vector<float> h(m->nparam, 0.f);
            std::mt19937_64 rng(seed);
            for (int l = 0; l < 4; ++l) {
                const int fan_in = l == 0 ? m->F : m->H, nin = l == 0 ? m->F : m->rows_in[l];
                const int nout = l == 3 ? m->W : m->cols_out[l];
                const float a = std::sqrt(6.0f / fan_in);
                std::uniform_real_distribution<float> U(-a, a);
                for (int i = 0; i < nin; ++i)
                    for (int j = 0; j < nout; ++j) h[m->off_w[l] + (size_t)i * m->cols_out[l] + j] = U(rng);
            }
            for (DevBuf* b : {&m->P, &m->G, &m->M1, &m->M2}) b->ensure(sizeof(float) * m->nparam);
            m->Pb.ensure(sizeof(uint16_t) * m->nparam);
            ILH_HIP(hipMemcpy(m->P.p, h.data(), sizeof(float) * m->nparam, hipMemcpyHostToDevice));
            ILH_HIP(hipMemset(m->M1.p, 0, sizeof(float) * m->nparam));
            ILH_HIP(hipMemset(m->M2.p, 0, sizeof(float) * m->nparam));
            ILH_HIP(hipMemset(m->G.p, 0, sizeof(float) * m->nparam));
            lpc_refresh_bf16(m);
            ILH_HIP(hipStreamSynchronize(ctx->stream));
            ctx_retain(ctx);
        } catch (...) {
            delete m;
            throw;
        }
        *out = m;
    });
}

extern "C" int ilh_lpc_destroy(ilh_lpc* m)
{
    return guarded_call([&] {
        if (!m) return;
        ilh_ctx* c = m->ctx;
        (void)hipSetDevice(c->device);
        (void)hipStreamSynchronize(c->stream);
        delete m;
        ctx_release(c);
    });
}

// Training rows of one geometry: its features and the band of its exact
// inverse (the plain method: the deflated Cholesky of the plan's Schur matrix).
extern "C" int ilh_lpc_add_geometry(ilh_lpc* m, ilh_plan* pl)
{
    return guarded_call([&] {
        check_model(m, pl);
        ILH_HIP(hipSetDevice(pl->ctx->device));
        if (m->stats_ok) fail(ILH_ERR_STATE, "training rows must be added before the first training step");
        need_margin(pl, "learned preconditioner training");
        const int N = pl->N;
        if (N < m->W) fail(ILH_ERR_ARG, "too few points for the band");
        hipStream_t st = pl->ctx->stream;
        DevBuf Sinv, T;
        Sinv.ensure(sizeof(double) * (size_t)N * solve_ld(N));
        exact_inverse_rowmajor(pl, Sinv.as<double>());
        lpc_features(m, pl);
        T.ensure(sizeof(float) * (size_t)N * m->W);
        k_band_targets<<<(unsigned)(((size_t)N * m->W + 255) / 256), 256, 0, st>>>(
            Sinv.as<double>(), solve_ld(N), N, m->W, pl->nbr.as<int>(), pl->ds.as<double>(), pl->g.dx, T.as<float>());
        ILH_HIP(hipGetLastError());
        const size_t r0 = (size_t)m->rows;
        m->hX.resize((r0 + N) * m->F);
        m->hT.resize((r0 + N) * m->W);
        ILH_HIP(hipMemcpyAsync(m->hX.data() + r0 * m->F, m->feat.p, sizeof(float) * (size_t)N * m->F,
                               hipMemcpyDeviceToHost, st));
        ILH_HIP(hipMemcpyAsync(m->hT.data() + r0 * m->W, T.p, sizeof(float) * (size_t)N * m->W, hipMemcpyDeviceToHost,
                               st));
        ILH_HIP(hipStreamSynchronize(st));
        m->rows += N;
        m->dX.ensure(sizeof(float) * m->hX.size());
        m->dT.ensure(sizeof(float) * m->hT.size());
        ILH_HIP(hipMemcpy(m->dX.p, m->hX.data(), sizeof(float) * m->hX.size(), hipMemcpyHostToDevice));
        ILH_HIP(hipMemcpy(m->dT.p, m->hT.data(), sizeof(float) * m->hT.size(), hipMemcpyHostToDevice));
    });
}

extern "C" int ilh_lpc_rows(const ilh_lpc* m) { return m ? m->rows : -1; }

// `steps` Adam steps on minibatches of `batch` rows drawn (seeded, without
// replacement within an epoch) from the training set; *loss: the mean of the
// steps' losses (standardized targets)
extern "C" int ilh_lpc_train(ilh_lpc* m, int steps, int batch, double lr, unsigned long long seed, double* loss)
{
    return guarded_call([&] {
        check_model(m, nullptr);
        ILH_HIP(hipSetDevice(m->ctx->device));
        if (steps < 0 || batch < 1 || !(lr > 0.0)) fail(ILH_ERR_ARG, "bad training arguments");
        lpc_stats(m);
        batch = std::min(batch, m->rows);
        lpc_work(m, (size_t)batch);
        hipStream_t st = m->ctx->stream;
        std::mt19937_64 rng(seed);
        std::vector<int> perm(m->rows);
        for (int i = 0; i < m->rows; ++i) perm[i] = i;
        size_t pos = perm.size();
        double acc = 0.0;
        std::vector<int> pick(batch);
        for (int it = 0; it < steps; ++it) {
            for (int b = 0; b < batch; ++b) {
                if (pos >= perm.size()) {
                    std::shuffle(perm.begin(), perm.end(), rng);
                    pos = 0;
                }
                pick[b] = perm[pos++];
            }
            ILH_HIP(hipMemcpyAsync(m->idx.p, pick.data(), sizeof(int) * batch, hipMemcpyHostToDevice, st));
            const size_t nx = (size_t)batch * m->Fp, ny = (size_t)batch * m->Op;
            k_standardize_bf<<<(unsigned)((nx + 255) / 256), 256, 0, st>>>(
                m->dX.as<float>(), m->F, m->idx.as<int>(), batch, m->F, m->mux.as<float>(), m->sdx.as<float>(), m->Fp,
                m->act[0].as<uint16_t>());
            k_standardize_f<<<(unsigned)((ny + 255) / 256), 256, 0, st>>>(
                m->dT.as<float>(), m->W, m->idx.as<int>(), batch, m->W, m->muy.as<float>(), m->sdy.as<float>(), m->Op,
                m->Tf.as<float>());
            ILH_HIP(hipGetLastError());
            lpc_forward(m, batch);
            acc += lpc_backward(m, batch);
            ++m->adam_t;
            const float b1 = 0.9f, b2 = 0.999f;
            const float c1 = 1.f - std::pow(b1, (float)m->adam_t), c2 = 1.f - std::pow(b2, (float)m->adam_t);
            k_adam<<<(unsigned)((m->nparam + 255) / 256), 256, 0, st>>>(m->P.as<float>(), m->G.as<float>(),
                                                                        m->M1.as<float>(), m->M2.as<float>(),
                                                                        m->nparam, (float)lr, b1, b2, c1, c2, 1e-8f);
            ILH_HIP(hipGetLastError());
            lpc_refresh_bf16(m);
        }
        ILH_HIP(hipStreamSynchronize(st));
        if (loss) *loss = steps ? acc / steps : 0.0;
    });
}

// Sizes of the model: [K, W (band), F (features), Fp, H (width), Op, nparam]
extern "C" int ilh_lpc_dims(const ilh_lpc* m, int* dims)
{
    return guarded_call([&] {
        if (!m || !dims) fail(ILH_ERR_ARG, "null argument");
        dims[0] = m->K;
        dims[1] = m->W;
        dims[2] = m->F;
        dims[3] = m->Fp;
        dims[4] = m->H;
        dims[5] = m->Op;
        dims[6] = (int)m->nparam;
    });
}

// all FP32 master parameters (layout of ilh_lpc_dims: W1 Fp x H row-major, b1,
// W2, b2, W3, b3, W4 H x Op, b4) copied out (get != 0) or in
extern "C" int ilh_lpc_params(ilh_lpc* m, float* params, int get)
{
    return guarded_call([&] {
        check_model(m, nullptr);
        if (!params) fail(ILH_ERR_ARG, "null array");
        ILH_HIP(hipSetDevice(m->ctx->device));
        if (get) {
            ILH_HIP(hipMemcpy(params, m->P.p, sizeof(float) * m->nparam, hipMemcpyDeviceToHost));
        } else {
            ILH_HIP(hipMemcpy(m->P.p, params, sizeof(float) * m->nparam, hipMemcpyHostToDevice));
            lpc_refresh_bf16(m);
            ILH_HIP(hipStreamSynchronize(m->ctx->stream));
        }
    });
}

// The loss and every parameter gradient of one batch given in standardized
// form (X rows x Fp, BF16 bit patterns; T rows x Op, FP32): the training
// step's forward and backward kernels without the update.
extern "C" int ilh_lpc_loss_grad(ilh_lpc* m, int rows, const unsigned short* X, const float* T, double* loss, float* grads)
{
    return guarded_call([&] {
        check_model(m, nullptr);
        if (rows < 1 || !X || !T || !grads) fail(ILH_ERR_ARG, "bad arguments");
        ILH_HIP(hipSetDevice(m->ctx->device));
        lpc_work(m, (size_t)rows);
        hipStream_t st = m->ctx->stream;
        ILH_HIP(hipMemcpyAsync(m->act[0].p, X, sizeof(uint16_t) * (size_t)rows * m->Fp, hipMemcpyHostToDevice, st));
        ILH_HIP(hipMemcpyAsync(m->Tf.p, T, sizeof(float) * (size_t)rows * m->Op, hipMemcpyHostToDevice, st));
        lpc_forward(m, rows);
        const double l = lpc_backward(m, rows);
        ILH_HIP(hipMemcpy(grads, m->G.p, sizeof(float) * m->nparam, hipMemcpyDeviceToHost));
        if (loss) *loss = l;
    });
}

// The plan's features (N x F, FP32) and the predicted band (N x W, FP64; null
// to skip) for its current geometry
extern "C" int ilh_lpc_predict(ilh_lpc* m, ilh_plan* pl, float* features, double* band)
{
    return guarded_call([&] {
        check_model(m, pl);
        ILH_HIP(hipSetDevice(pl->ctx->device));
        const int N = pl->N;
        hipStream_t st = pl->ctx->stream;
        if (band) {
            DevBuf c;
            c.ensure(sizeof(double) * (size_t)N * m->W);
            lpc_band(m, pl, c.as<double>());
            ILH_HIP(hipMemcpyAsync(band, c.p, sizeof(double) * (size_t)N * m->W, hipMemcpyDeviceToHost, st));
            ILH_HIP(hipStreamSynchronize(st));
        } else {
            lpc_features(m, pl);
        }
        if (features) {
            ILH_HIP(hipMemcpyAsync(features, m->feat.p, sizeof(float) * (size_t)N * m->F, hipMemcpyDeviceToHost, st));
            ILH_HIP(hipStreamSynchronize(st));
        }
    });
}

// standardization used by the model ([mu_x (F), sd_x (F), mu_y (W), sd_y (W)])
extern "C" int ilh_lpc_standardization(ilh_lpc* m, float* out)
{
    return guarded_call([&] {
        check_model(m, nullptr);
        if (!m->stats_ok) fail(ILH_ERR_STATE, "not trained yet");
        ILH_HIP(hipSetDevice(m->ctx->device));
        ILH_HIP(hipMemcpy(out, m->mux.p, sizeof(float) * m->F, hipMemcpyDeviceToHost));
        ILH_HIP(hipMemcpy(out + m->F, m->sdx.p, sizeof(float) * m->F, hipMemcpyDeviceToHost));
        ILH_HIP(hipMemcpy(out + 2 * m->F, m->muy.p, sizeof(float) * m->W, hipMemcpyDeviceToHost));
        ILH_HIP(hipMemcpy(out + 2 * m->F + m->W, m->sdy.p, sizeof(float) * m->W, hipMemcpyDeviceToHost));
    });
}
