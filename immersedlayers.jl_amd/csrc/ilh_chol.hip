// ilh_chol.hip -- the dense FP64 algebra of the Schur complements on the
// matrix cores: a v_mfma_f64_16x16x4f64 GEMM, and the blocked Cholesky
// factorisation of the symmetrised Schur matrix with its multi-right-hand-side
// solves.
//
// The reference solves S x = b with `S\b` (LAPACK LU on every call,
// test/literate/problems.jl:291, test/literate/dirichlet.jl:99).  For the
// Schur matrices built from an interpolation E and a regularization R of the
// same grid kind (create_RTLinvR, src/matrix_operators.jl:9-30, and the
// integrating-factor stage matrices E H R), GridScaling makes R = E^T diag(d)
// (d = ds / dx^2), so
//     S = M diag(d),   M = E K E^T  symmetric,
// with K the kernel (-L^-1 or the heat kernel).  -L^-1 is only conditionally
// positive definite: M is positive definite on the vectors orthogonal to the
// constant one and has a single negative eigenvalue along it.  A Householder
// reflection H (H u = e_1, u = 1/sqrt(n)) moves that direction to the first
// coordinate:
//     H M H = [alpha  beta^T ]      Mt = (n-1) x (n-1), positive definite,
//             [beta   Mt     ]
// Mt is factored by blocked Cholesky (no pivoting), the first coordinate is
// eliminated through its scalar Schur complement sigma = alpha - beta^T Mt^-1 beta.
// An exactly singular direction (points whose windows coincide) shows as a
// non-positive pivot; the caller then falls back to LU.
#include "ilh_internal.hpp"

#include <algorithm>
#include <cmath>

namespace ilh {

typedef double double4_t __attribute__((ext_vector_type(4)));

// ------------------------------------------------------------------ GEMM
// C (M x Nc) = beta C + alpha opA(A) opB(B), column-major, beta in {0, 1}
// (BETA0: C is not read).  opA(A)[m][k] = A[m + lda k] (TA = 0) or
// A[k + lda m] (TA = 1); opB(B)[k][n] = B[k + ldb n] (TB = 0) or B[n + ldb k]
// (TB = 1).  LOWER: tiles lying wholly above the diagonal of a square C are
// skipped (their entries are never read).  tri restricts the K range of each
// tile to where a triangular operand is nonzero (its other triangle must hold
// zeros): TRI_A_LOWER opA(A)[m][k] = 0 for k > m, TRI_A_UPPER for k < m,
// TRI_B_LOWER opB(B)[k][n] = 0 for k < n, TRI_B_UPPER for k > n.
//
// 64 x 64 tile per work-group of four waves, each wave a 32 x 32 quadrant of
// 2 x 2 blocks of 16 x 16.  K runs in chunks of 16 through two LDS buffers:
// the next chunk's global loads are in flight while the current one feeds the
// matrix cores.  The product is formed transposed (the B tile as the MFMA's
// first operand), so a lane's result entries run down a column of C and the
// stores of 16 lanes cover 128 contiguous bytes.
// v_mfma_f64_16x16x4f64 lane layout (checked on gfx950): first operand
// X[i][k] at lane i + 16 k, second Y[k][j] at lane j + 16 k; lane l, item r
// of the result holds D[l / 16 + 4 r][l % 16].
static constexpr int GT = 64, GKC = 16;

template <int TA, int TB, bool LOWER, bool BETA0>
// (A or B may be C itself when every work-group reads all the entries it
// overwrites: the in-place panel and diagonal-block products below; hence no
// __restrict__.)
__global__ void __launch_bounds__(256) k_dgemm(int M, int Nc, int K, double alpha, const double* A, int lda,
                                               const double* B, int ldb, double* C, int ldc, int tri)
{
    __shared__ double As[2][GKC][GT + 2];   // As[k][m]
    __shared__ double Bs[2][GKC][GT + 2];   // Bs[k][n]
    const int m0 = blockIdx.x * GT, n0 = blockIdx.y * GT;
    if (LOWER && m0 + GT - 1 < n0) return;
    int klo = 0, khi = K;
    if (tri == TRI_A_LOWER) khi = min(K, m0 + GT);
    else if (tri == TRI_A_UPPER) klo = m0;
    else if (tri == TRI_B_LOWER) klo = n0;
    else if (tri == TRI_B_UPPER) khi = min(K, n0 + GT);
    const int kbase = klo / GKC * GKC;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
    double ra[4], rb[4];
    auto load = [&](int k0) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            int m, k;
            if (TA == 0) { m = tid & 63; k = (tid >> 6) + 4 * j; }
            else { k = tid & 15; m = (tid >> 4) + 16 * j; }
            const int gm = m0 + m, gk = k0 + k;
            ra[j] = (gm < M && gk >= klo && gk < khi) ? (TA == 0 ? A[gm + (size_t)lda * gk] : A[gk + (size_t)lda * gm])
                                                      : 0.0;
            int n, kb;
            if (TB == 0) { kb = tid & 15; n = (tid >> 4) + 16 * j; }
            else { n = tid & 63; kb = (tid >> 6) + 4 * j; }
            const int gn = n0 + n, gkb = k0 + kb;
            rb[j] = (gn < Nc && gkb >= klo && gkb < khi) ? (TB == 0 ? B[gkb + (size_t)ldb * gn] : B[gn + (size_t)ldb * gkb])
                                                          : 0.0;
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (TA == 0) As[buf][(tid >> 6) + 4 * j][tid & 63] = ra[j];
            else As[buf][tid & 15][(tid >> 4) + 16 * j] = ra[j];
            if (TB == 0) Bs[buf][tid & 15][(tid >> 4) + 16 * j] = rb[j];
            else Bs[buf][(tid >> 6) + 4 * j][tid & 63] = rb[j];
        }
    };
    double4_t acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = (double4_t){0.0, 0.0, 0.0, 0.0};
    const int nch = khi > kbase ? (khi - kbase + GKC - 1) / GKC : 0;
    load(kbase);
    store(0);
    __syncthreads();
    for (int c = 0; c < nch; ++c) {
        const int buf = c & 1;
        if (c + 1 < nch) load(kbase + (c + 1) * GKC);
#pragma unroll
        for (int kk = 0; kk < GKC; kk += 4) {
            const int kr = kk + (lane >> 4);
            double a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) a[i] = As[buf][kr][wm + 16 * i + (lane & 15)];
#pragma unroll
            for (int j = 0; j < 2; ++j) b[j] = Bs[buf][kr][wn + 16 * j + (lane & 15)];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[j], a[i], acc[i][j], 0, 0, 0);
        }
        if (c + 1 < nch) store(buf ^ 1);
        __syncthreads();
    }
    // acc[i][j] item r of lane l: C row m0 + wm + 16 i + l % 16, column n0 + wn + 16 j + l / 16 + 4 r
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gm = m0 + wm + 16 * i + (lane & 15);
                const int gn = n0 + wn + 16 * j + (lane >> 4) + 4 * r;
                if (gm < M && gn < Nc) {
                    double* p = C + gm + (size_t)ldc * gn;
                    *p = BETA0 ? alpha * acc[i][j][r] : *p + alpha * acc[i][j][r];
                }
            }
}

void dgemm(hipStream_t st, int ta, int tb, int M, int Nc, int K, double alpha, const double* A, int lda,
           const double* B, int ldb, double beta, double* C, int ldc, bool lower, int tri)
{
    if (M <= 0 || Nc <= 0) return;
    if (beta != 0.0 && beta != 1.0) fail(ILH_ERR_ARG, "dgemm: beta must be 0 or 1");
    const dim3 grid((M + GT - 1) / GT, (Nc + GT - 1) / GT);
    const bool b0 = beta == 0.0;
#define ILH_DG(TA_, TB_, LO_, B0_) \
    k_dgemm<TA_, TB_, LO_, B0_><<<grid, 256, 0, st>>>(M, Nc, K, alpha, A, lda, B, ldb, C, ldc, tri)
    if (K <= 0) {
        if (!b0) return;
        alpha = 0.0;   // C = 0
    }
    if (lower) {
        if (ta != 0 || tb != 1 || b0) fail(ILH_ERR_ARG, "dgemm: the lower-triangle form is C += alpha A B^T");
        ILH_DG(0, 1, true, false);
    } else if (ta == 0 && tb == 0) {
        if (b0) ILH_DG(0, 0, false, true); else ILH_DG(0, 0, false, false);
    } else if (ta == 0 && tb == 1) {
        if (b0) ILH_DG(0, 1, false, true); else ILH_DG(0, 1, false, false);
    } else if (ta == 1 && tb == 0) {
        if (b0) ILH_DG(1, 0, false, true); else ILH_DG(1, 0, false, false);
    } else {
        if (b0) ILH_DG(1, 1, false, true); else ILH_DG(1, 1, false, false);
    }
#undef ILH_DG
    ILH_HIP(hipGetLastError());
}

// ------------------------------------------------------------- Cholesky
static constexpr int CB = 64;   // block size of the factorisation and of the solves

__device__ __forceinline__ double readlane_d(double v, int lane)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Cholesky of the nb x nb diagonal block at (k0, k0) of A (lower triangle
// used), written back into A's lower triangle, and the inverse of the factor
// (nb x nb with the zeros of its upper triangle) into Winv (leading dimension ldw).
// Work-group barriers, not arithmetic, set the time of a small factorisation,
// so the block is processed in 16-wide sub-blocks: one wave factors a 16 x 16
// diagonal sub-block on its own (lane i holds row i; the pivot and the column
// of each step are broadcast with readlane, no barrier), then the work-group
// solves the sub-panel below it and updates the trailing lower triangle.  The
// inverse is built the same way: the four 16 x 16 diagonal inverses in
// parallel (one wave each), then the off-diagonal blocks by their distance
// from the diagonal.  A pivot that is not positive records its 1-based column
// in *info (the first one only) and is replaced by 1, so that the kernel still
// runs to its end.
static constexpr int SB = 16;

__global__ void __launch_bounds__(256) k_chol_diag(double* __restrict__ A, int lda, int k0, int nb,
                                                   double* __restrict__ Winv, int ldw, int* __restrict__ info)
{
    __shared__ double T[CB][CB + 1];
    __shared__ double X[CB][CB + 1];
    __shared__ double Y[3][SB][SB + 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int e = tid; e < CB * CB; e += 256) {
        const int r = e % CB, c = e / CB;
        T[r][c] = (r < nb && c < nb) ? (r >= c ? A[(k0 + r) + (size_t)lda * (k0 + c)] : 0.0) : (r == c ? 1.0 : 0.0);
    }
    __syncthreads();
    for (int o = 0; o < CB; o += SB) {
        // (a) the 16 x 16 diagonal sub-block, wave 0
        if (wave == 0) {
            double d[SB];
            const int ri = lane < SB ? lane : 0;
#pragma unroll
            for (int c = 0; c < SB; ++c) d[c] = T[o + ri][o + c];
#pragma unroll
            for (int j = 0; j < SB; ++j) {
                double piv = readlane_d(d[j], j);
                if (!(piv > 0.0)) {
                    if (lane == 0) atomicCAS(info, 0, k0 + o + j + 1);
                    piv = 1.0;
                }
                const double ljj = sqrt(piv), rinv = 1.0 / ljj;
                const double lij = d[j] * rinv;
#pragma unroll
                for (int l = j + 1; l < SB; ++l) d[l] -= lij * readlane_d(lij, l);
                d[j] = lane > j ? lij : (lane == j ? ljj : 0.0);
            }
            if (lane < SB)
#pragma unroll
                for (int c = 0; c < SB; ++c) T[o + lane][o + c] = c <= lane ? d[c] : 0.0;
        }
        __syncthreads();
        const int rest = CB - o - SB;
        if (rest > 0) {
            // (b) sub-panel: rows o+16.. solve p L11^T = a (one thread per row)
            if (tid < rest) {
                const int r = o + SB + tid;
                double p[SB];
#pragma unroll
                for (int c = 0; c < SB; ++c) {
                    double v = T[r][o + c];
#pragma unroll
                    for (int k = 0; k < c; ++k) v -= p[k] * T[o + c][o + k];
                    p[c] = v / T[o + c][o + c];
                }
#pragma unroll
                for (int c = 0; c < SB; ++c) T[r][o + c] = p[c];
            }
            __syncthreads();
            // (c) trailing lower triangle: T[r][s] -= sum_k P[r][k] P[s][k]
            for (int e = tid; e < rest * rest; e += 256) {
                const int ri = e % rest, si = e / rest;
                if (si > ri) continue;
                const int r = o + SB + ri, sc = o + SB + si;
                double acc = 0.0;
#pragma unroll
                for (int k = 0; k < SB; ++k) acc += T[r][o + k] * T[sc][o + k];
                T[r][sc] -= acc;
            }
            __syncthreads();
        }
    }
    for (int e = tid; e < nb * nb; e += 256) {
        const int r = e % nb, c = e / nb;
        if (r >= c) A[(k0 + r) + (size_t)lda * (k0 + c)] = T[r][c];
    }
    // inverse: the diagonal 16 x 16 inverses, wave b forms block b (lane c < 16: column c)
    {
        const int o = wave * SB, c = lane;
        if (c < SB) {
            double x[SB];
#pragma unroll
            for (int r = 0; r < SB; ++r) {
                double v = r == c ? 1.0 : 0.0;
#pragma unroll
                for (int k = 0; k < r; ++k) v -= T[o + r][o + k] * x[k];
                x[r] = v / T[o + r][o + r];
            }
#pragma unroll
            for (int r = 0; r < SB; ++r) X[o + r][o + c] = x[r];
        }
        for (int e = tid; e < CB * CB; e += 256) {   // zero the strict upper blocks
            const int r = e % CB, cc = e / CB;
            if (r / SB < cc / SB) X[r][cc] = 0.0;
        }
    }
    __syncthreads();
    // off-diagonal blocks (bi, bj), bi - bj = dist: X_ij = -X_ii sum_{k=j}^{i-1} L_ik X_kj
    for (int dist = 1; dist < CB / SB; ++dist) {
        const int nblk = CB / SB - dist;
        for (int e = tid; e < nblk * SB * SB; e += 256) {
            const int b = e / (SB * SB), rr = e % SB, cc = (e / SB) % SB;
            const int bj = b, bi = b + dist;
            double acc = 0.0;
            for (int kb = bj; kb < bi; ++kb)
#pragma unroll
                for (int k = 0; k < SB; ++k) acc += T[bi * SB + rr][kb * SB + k] * X[kb * SB + k][bj * SB + cc];
            Y[b][rr][cc] = acc;
        }
        __syncthreads();
        for (int e = tid; e < nblk * SB * SB; e += 256) {
            const int b = e / (SB * SB), rr = e % SB, cc = (e / SB) % SB;
            const int bj = b, bi = b + dist;
            double acc = 0.0;
#pragma unroll
            for (int k = 0; k < SB; ++k) acc += X[bi * SB + rr][bi * SB + k] * Y[b][k][cc];
            X[bi * SB + rr][bj * SB + cc] = -acc;
        }
        __syncthreads();
    }
    for (int e = tid; e < nb * nb; e += 256) {
        const int r = e % nb, c = e / nb;
        Winv[r + (size_t)ldw * c] = X[r][c];
    }
}

// dst[i + ldd j] = src[i + lds j], i < rows, j < cols
__global__ void k_copy2d(const double* __restrict__ src, int lds, double* __restrict__ dst, int ldd, int rows,
                         int cols)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i < rows && j < cols) dst[i + (size_t)ldd * j] = src[i + (size_t)lds * j];
}

// Recursive Cholesky of the n x n block at (koff, koff) of A (lower triangle),
// which also forms W = L^-1 of that block (W's strict upper triangle must hold
// zeros on entry).  With L = [L11 0; L21 L22] and W = L^-1 = [W11 0; W21 W22]:
//   L21 = A21 W11^T,  A22 -= L21 L21^T,  W21 = -W22 (L21 W11),
// halves split at multiples of CB, the leaves the diagonal-block kernel: the
// work between leaves is large GEMMs on the matrix cores (the triangular
// factors' zero halves skipped).  T: scratch of n/2 x n/2 at least.
static void chol_rec(hipStream_t st, double* A, int lda, double* W, int ldw, int koff, int n, double* T, int* info)
{
    if (n <= CB) {
        k_chol_diag<<<1, 256, 0, st>>>(A, lda, koff, n, W + koff + (size_t)ldw * koff, ldw, info);
        ILH_HIP(hipGetLastError());
        return;
    }
    const int n1 = ((n + 1) / 2 + CB - 1) / CB * CB, n2 = n - n1;
    chol_rec(st, A, lda, W, ldw, koff, n1, T, info);
    double* A21 = A + (koff + n1) + (size_t)lda * koff;
    double* A22 = A + (koff + n1) + (size_t)lda * (koff + n1);
    const double* W11 = W + koff + (size_t)ldw * koff;
    double* W21 = W + (koff + n1) + (size_t)ldw * koff;
    const double* W22 = W + (koff + n1) + (size_t)ldw * (koff + n1);
    dgemm(st, 0, 1, n2, n1, n1, 1.0, A21, lda, W11, ldw, 0.0, T, n2, false, TRI_B_UPPER);   // L21
    k_copy2d<<<dim3((n2 + 255) / 256, n1), 256, 0, st>>>(T, n2, A21, lda, n2, n1);
    ILH_HIP(hipGetLastError());
    dgemm(st, 0, 1, n2, n2, n1, -1.0, A21, lda, A21, lda, 1.0, A22, lda, true);
    chol_rec(st, A, lda, W, ldw, koff + n1, n2, T, info);
    dgemm(st, 0, 0, n2, n1, n1, 1.0, A21, lda, W11, ldw, 0.0, T, n2, false, TRI_B_LOWER);   // L21 W11
    dgemm(st, 0, 0, n2, n1, n2, -1.0, W22, ldw, T, n2, 0.0, W21, ldw, false, TRI_A_LOWER);
}

// Cholesky of the n x n matrix A (lower triangle, leading dimension lda):
// W (n x n, leading dimension n) receives L^-1; A's diagonal blocks hold L's,
// the rest of A is scratch afterwards.  Returns 0, or the 1-based column of
// the first non-positive pivot.
int chol_factor(ilh_plan* pl, double* A, int n, int lda, double* W, double* T, int* info_dev)
{
    hipStream_t st = pl->ctx->stream;
    ILH_HIP(hipMemsetAsync(info_dev, 0, sizeof(int), st));
    ILH_HIP(hipMemsetAsync(W, 0, sizeof(double) * (size_t)n * n, st));
    chol_rec(st, A, lda, W, n, 0, n, T, info_dev);
    int h = 0;
    ILH_HIP(hipMemcpyAsync(&h, info_dev, sizeof(int), hipMemcpyDeviceToHost, st));
    ILH_HIP(hipStreamSynchronize(st));
    return h;
}

// X <- (L L^T)^-1 X = W^T (W X) (X n x nrhs, leading dimension ldx; T n x nrhs scratch)
void chol_solve(ilh_plan* pl, const double* W, int n, double* X, int ldx, int nrhs, double* T)
{
    hipStream_t st = pl->ctx->stream;
    if (n <= 0 || nrhs <= 0) return;
    dgemm(st, 0, 0, n, nrhs, n, 1.0, W, n, X, ldx, 0.0, T, n, false, TRI_A_LOWER);
    dgemm(st, 1, 0, n, nrhs, n, 1.0, W, n, T, n, 0.0, X, ldx, false, TRI_A_UPPER);
}

// --------------------------------------------------- deflated symmetric solve
// w: the Householder vector of H = I - 2 w w^T with H (1/sqrt(n)) = e_1
__global__ void k_house_vec(double* __restrict__ w, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (n == 1) {   // H = I
        w[0] = 0.0;
        return;
    }
    const double u = 1.0 / sqrt((double)n);
    const double nrm = sqrt(2.0 - 2.0 * u);
    w[i] = (i == 0 ? u - 1.0 : u) / nrm;
}

// y = A x (A n x n column-major, symmetric: one wave per row reads the row
// through the column, i.e. y_i = sum_j A[j + lda i] x_j)
__global__ void k_symv(const double* __restrict__ A, int n, int lda, const double* __restrict__ x,
                       double* __restrict__ y)
{
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (i >= n) return;
    const double* col = A + (size_t)lda * i;
    double s = 0.0;
    for (int j = lane; j < n; j += 64) s += col[j] * x[j];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) y[i] = s;
}

// p = 2 (y - (w.y) w), y = A w: one work-group
__global__ void __launch_bounds__(1024) k_house_p(const double* __restrict__ w, const double* __restrict__ y, int n,
                                                  double* __restrict__ p)
{
    __shared__ double red[1024];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 1024) s += w[i] * y[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int off = 512; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    const double t = red[0];
    for (int i = threadIdx.x; i < n; i += 1024) p[i] = 2.0 * (y[i] - t * w[i]);
}

// A <- H A H = A - w p^T - p w^T
__global__ void k_rank2(double* __restrict__ A, int n, int lda, const double* __restrict__ w,
                        const double* __restrict__ p)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= n) return;
    A[i + (size_t)lda * j] -= w[i] * p[j] + p[i] * w[j];
}

// r[j] = sum_{i < m} v[i] X[i + ldx j], one wave per column j
__global__ void k_coldot(const double* __restrict__ v, const double* __restrict__ X, int m, int ldx, int ncols,
                         double* __restrict__ r)
{
    const int j = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (j >= ncols) return;
    const double* col = X + (size_t)ldx * j;
    double s = 0.0;
    for (int i = lane; i < m; i += 64) s += v[i] * col[i];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) r[j] = s;
}

// X[i + ldx j] += c a[i] b[j]  (i < m)
__global__ void k_rank1(double* __restrict__ X, int m, int ldx, double c, const double* __restrict__ a,
                        const double* __restrict__ b)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= m) return;
    X[i + (size_t)ldx * j] += c * a[i] * b[j];
}

// rows 1..m of X -= wt y1^T, y1 = row 0 of X (not written here)
__global__ void k_rank1_row(double* __restrict__ X, int ldx, int m, const double* __restrict__ wt)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= m) return;
    X[1 + i + (size_t)ldx * j] -= wt[i] * X[(size_t)ldx * j];
}

// first row of the deflated solve: X[0, j] = (X[0, j] - r[j]) / sigma
__global__ void k_first_row(double* __restrict__ X, int ldx, int ncols, const double* __restrict__ r,
                            const double* __restrict__ sig)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ncols) return;
    X[(size_t)ldx * j] = (X[(size_t)ldx * j] - r[j]) / sig[0];
}

// sigma = alpha - beta . wt  (alpha = A[0,0], beta = A[1:, 0], wt = Mt^-1 beta)
__global__ void __launch_bounds__(1024) k_sigma(const double* __restrict__ A, const double* __restrict__ wt, int m,
                                                double* __restrict__ sig)
{
    __shared__ double red[1024];
    double s = 0.0;
    for (int i = threadIdx.x; i < m; i += 1024) s += A[1 + i] * wt[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int off = 512; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) sig[0] = A[0] - red[0];
}

bool sym_factor(ilh_plan* pl, SymFactor& F, int n)
{
    hipStream_t st = pl->ctx->stream;
    F.n = n;
    const int m = n - 1;
    F.w.ensure(sizeof(double) * (size_t)n);
    F.vec.ensure(sizeof(double) * (size_t)2 * n + 8);
    F.W.ensure(sizeof(double) * (size_t)std::max(m, 1) * std::max(m, 1));
    F.T.ensure(sizeof(double) * ((size_t)std::max(m, 2) / 2 + CB) * ((size_t)std::max(m, 2) / 2 + CB));
    F.info.ensure(sizeof(int));
    double* A = F.A.as<double>();
    double* w = F.w.as<double>();
    double* y = F.vec.as<double>();
    double* p = y + n;
    k_house_vec<<<(n + 255) / 256, 256, 0, st>>>(w, n);
    k_symv<<<(n + 3) / 4, 256, 0, st>>>(A, n, n, w, y);
    k_house_p<<<1, 1024, 0, st>>>(w, y, n, p);
    k_rank2<<<dim3((n + 255) / 256, n), 256, 0, st>>>(A, n, n, w, p);
    ILH_HIP(hipGetLastError());
    if (m > 0) {
        const int bad = chol_factor(pl, A + 1 + (size_t)n, m, n, F.W.as<double>(), F.T.as<double>(), F.info.as<int>());
        if (bad) return false;
    }
    // wt = Mt^-1 beta, sigma = alpha - beta . wt
    F.wt.ensure(sizeof(double) * (size_t)std::max(m, 1));
    if (m > 0) {
        ILH_HIP(hipMemcpyAsync(F.wt.p, A + 1, sizeof(double) * m, hipMemcpyDeviceToDevice, st));
        chol_solve(pl, F.W.as<double>(), m, F.wt.as<double>(), m, 1, F.T.as<double>());
    }
    F.sig.ensure(sizeof(double));
    k_sigma<<<1, 1024, 0, st>>>(A, F.wt.as<double>(), m, F.sig.as<double>());
    ILH_HIP(hipGetLastError());
    double sig = 0.0;
    ILH_HIP(hipMemcpyAsync(&sig, F.sig.p, sizeof(double), hipMemcpyDeviceToHost, st));
    ILH_HIP(hipStreamSynchronize(st));
    if (!(sig != 0.0) || !std::isfinite(sig)) return false;
    F.ok = true;
    return true;
}

// Minv[q + ld p] = X[q + n p] / (c d[p]) for q < n, 0 for n <= q < ld: the
// row-major inverse (M diag(c d))^-1 = diag(1 / (c d)) M^-1 from X = M^-1
// (symmetric, so its row p is read along its column p, coalesced)
__global__ void k_rows_scaled(const double* __restrict__ X, int n, const double* __restrict__ d, double c,
                              double* __restrict__ Minv, int ld)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x, p = blockIdx.y;
    if (q >= ld) return;
    Minv[q + (size_t)ld * p] = q < n ? X[q + (size_t)n * p] / (c * d[p]) : 0.0;
}

// X[i + n j] /= c d[i]  (rows scaled: diag(1 / (c d)) X)
__global__ void k_scale_rows(double* __restrict__ X, int n, int ncols, const double* __restrict__ d, double c)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= n || j >= ncols) return;
    X[i + (size_t)n * j] /= c * d[i];
}

void scale_rows_inv(hipStream_t st, double* X, int n, int ncols, const double* d, double c)
{
    if (n <= 0 || ncols <= 0) return;
    k_scale_rows<<<dim3((n + 255) / 256, ncols), 256, 0, st>>>(X, n, ncols, d, c);
    ILH_HIP(hipGetLastError());
}

// Minv (row-major, leading dimension solve_ld(n)) = (M diag(c d))^-1 for the
// factor of sym_factor
__global__ void k_identity_cm(double* __restrict__ A, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= n) return;
    A[i + (size_t)n * j] = i == j ? 1.0 : 0.0;
}

void sym_inverse_rowmajor(ilh_plan* pl, SymFactor& F, const double* d, double c, double* Minv)
{
    hipStream_t st = pl->ctx->stream;
    const int n = F.n, ld = solve_ld(n);
    DevBuf X;
    X.ensure(sizeof(double) * (size_t)n * n);
    k_identity_cm<<<dim3((n + 255) / 256, n), 256, 0, st>>>(X.as<double>(), n);
    ILH_HIP(hipGetLastError());
    sym_solve(pl, F, X.as<double>(), n);
    k_rows_scaled<<<dim3((ld + 255) / 256, n), 256, 0, st>>>(X.as<double>(), n, d, c, Minv, ld);
    ILH_HIP(hipGetLastError());
    ILH_HIP(hipStreamSynchronize(st));   // X is released on return
}

// X <- M^-1 X (n x nrhs, leading dimension n) for the factor of sym_factor
void sym_solve(ilh_plan* pl, SymFactor& F, double* X, int nrhs)
{
    if (!F.ok) fail(ILH_ERR_STATE, "symmetric factor not built");
    hipStream_t st = pl->ctx->stream;
    const int n = F.n, m = n - 1;
    F.r.ensure(sizeof(double) * (size_t)std::max(nrhs, 1));
    double* r = F.r.as<double>();
    const double* w = F.w.as<double>();
    const dim3 g1((n + 255) / 256, nrhs);
    // X' = H X
    k_coldot<<<(nrhs + 3) / 4, 256, 0, st>>>(w, X, n, n, nrhs, r);
    k_rank1<<<g1, 256, 0, st>>>(X, n, n, -2.0, w, r);
    ILH_HIP(hipGetLastError());
    const double* A = F.A.as<double>();
    // rows 1.. : Mt^-1 X'[1:, :]
    if (m > 0) {
        F.T2.ensure(sizeof(double) * (size_t)m * nrhs);
        chol_solve(pl, F.W.as<double>(), m, X + 1, n, nrhs, F.T2.as<double>());
    }
    // first row: y1 = (X'[0, :] - beta^T Mt^-1 X'[1:, :]) / sigma
    k_coldot<<<(nrhs + 3) / 4, 256, 0, st>>>(A + 1, X + 1, m, n, nrhs, r);
    k_first_row<<<(nrhs + 255) / 256, 256, 0, st>>>(X, n, nrhs, r, F.sig.as<double>());
    // rows 1.. -= wt y1^T  (the first row of X now holds y1)
    if (m > 0) k_rank1_row<<<dim3((m + 255) / 256, nrhs), 256, 0, st>>>(X, n, m, F.wt.as<double>());
    // X = H X'
    k_coldot<<<(nrhs + 3) / 4, 256, 0, st>>>(w, X, n, n, nrhs, r);
    k_rank1<<<g1, 256, 0, st>>>(X, n, n, -2.0, w, r);
    ILH_HIP(hipGetLastError());
}

}  // namespace ilh
