// ilh_dense.hip -- Schur-complement matrices and their dense FP64 algebra.
//
// Direct assembly: the reference builds S = -E L^{-1} R one column at a time
// with N full-grid LGF solves (create_RTLinvR, src/matrix_operators.jl:9-30).
// Every operator in those products is a sum of separable 5x5 stencil terms
// (ilh_terms.hip), so each entry is
//   K[p,q] = coef_p coef_q dx^2 sum_{u,v=-4..4} cx(u) cy(v) G(dbx+u, dby+v),
// cx/cy the correlations of the two points' 1-D weights: a 9x9 window of the
// LGF table per term pair instead of a grid-sized FFT per column.
//
// LU with partial pivoting (what the reference's `S\s` runs, LAPACK getrf) is
// blocked right-looking; the trailing updates and the blocked triangular
// solves run on the FP64 matrix cores (v_mfma_f64_16x16x4f64).
#include "ilh_internal.hpp"

#include <algorithm>
#include <cstdlib>

namespace ilh {

typedef double double4_t __attribute__((ext_vector_type(4)));

// ------------------------------------------------------------ assembly
static constexpr int AT = 16;   // points per assembly tile edge

__global__ void __launch_bounds__(AT * AT) k_assemble(const Term* __restrict__ TA, int TAn, const Term* __restrict__ TB,
                                                      int TBn, int N, const double* __restrict__ G, int Tn,
                                                      double factor, int delta, double* __restrict__ K)
{
    __shared__ Term sa[AT * 2], sb[AT * 2];
    const int p0 = blockIdx.x * AT, q0 = blockIdx.y * AT;
    const int tid = threadIdx.x + AT * threadIdx.y;
    if (tid < AT * TAn) {
        const int p = p0 + tid / TAn;
        if (p < N) sa[tid] = TA[(size_t)p * TAn + tid % TAn];
    } else if (tid >= 64 && tid < 64 + AT * TBn) {
        const int k = tid - 64;
        const int q = q0 + k / TBn;
        if (q < N) sb[k] = TB[(size_t)q * TBn + k % TBn];
    }
    __syncthreads();
    const int p = p0 + threadIdx.x, q = q0 + threadIdx.y;
    if (p >= N || q >= N) return;
    double acc = 0.0;
    for (int ia = 0; ia < TAn; ++ia) {
        const Term& a = sa[threadIdx.x * TAn + ia];
        for (int ib = 0; ib < TBn; ++ib) {
            const Term& b = sb[threadIdx.y * TBn + ib];
            if (a.gk != b.gk) continue;
            double cx[9], cy[9];
#pragma unroll
            for (int u = 0; u < 9; ++u) { cx[u] = 0.0; cy[u] = 0.0; }
#pragma unroll
            for (int i = 0; i < 5; ++i)
#pragma unroll
                for (int k = 0; k < 5; ++k) {
                    cx[i - k + 4] += a.wx[i] * b.wx[k];
                    cy[i - k + 4] += a.wy[i] * b.wy[k];
                }
            const int dbx = a.bx - b.bx, dby = a.by - b.by;
            double s = 0.0;
            if (delta) {
                // identity kernel: only offset 0 contributes
                const int u = -dbx + 4, v = -dby + 4;
                if (u >= 0 && u < 9 && v >= 0 && v < 9) s = cx[u] * cy[v];
            } else {
                for (int v = 0; v < 9; ++v) {
                    if (cy[v] == 0.0) continue;
                    const int jj = abs(dby + v - 4);
                    const double* Gr = G + (size_t)Tn * jj;
                    double r = 0.0;
#pragma unroll
                    for (int u = 0; u < 9; ++u) r += cx[u] * Gr[abs(dbx + u - 4)];
                    s += cy[v] * r;
                }
            }
            acc += a.coef * b.coef * s;
        }
    }
    K[p + (size_t)N * q] = factor * acc;
}

void assemble(ilh_plan* pl, int opA, int opB, int gk, double factor, bool delta_kernel, double* K)
{
    (void)gk;
    const int N = pl->N;
    if (N == 0) return;
    const Term* A = get_terms(pl, opA);
    const Term* B = get_terms(pl, opB);
    if (!delta_kernel) ensure_lgf(pl);
    dim3 grid((N + AT - 1) / AT, (N + AT - 1) / AT);
    k_assemble<<<grid, dim3(AT, AT), 0, pl->ctx->stream>>>(A, op_terms_per_point(opA), B, op_terms_per_point(opB), N,
                                                           delta_kernel ? nullptr : pl->lgf.as<double>(), pl->lgf_mx,
                                                           factor, delta_kernel ? 1 : 0, K);
    ILH_HIP(hipGetLastError());
}

// ----------------------------------------------------------------- GEMM
// C[M x Nc] -= A[M x K] * B[K x Nc], column-major with leading dims.
// 64x64 tile per work-group of 4 waves; each wave owns a 32x32 quadrant as
// 2x2 MFMA 16x16 blocks; K staged through LDS in chunks of 16.
static constexpr int GT = 64, GK = 16;

__global__ void __launch_bounds__(256) k_gemm_sub(int M, int Nc, int K, const double* __restrict__ A, int lda,
                                                  const double* __restrict__ B, int ldb, double* __restrict__ C,
                                                  int ldc)
{
    __shared__ double As[GK][GT + 1];   // As[k][m]
    __shared__ double Bs[GK][GT + 1];   // Bs[k][n]
    const int m0 = blockIdx.x * GT, n0 = blockIdx.y * GT;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
    double4_t acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = (double4_t){0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < K; k0 += GK) {
        // A tile: 64 rows x 16 k, column-major source: consecutive threads read consecutive rows
        for (int e = tid; e < GT * GK; e += 256) {
            const int m = e % GT, k = e / GT;
            const int gm = m0 + m, gk = k0 + k;
            As[k][m] = (gm < M && gk < K) ? A[gm + (size_t)lda * gk] : 0.0;
        }
        for (int e = tid; e < GT * GK; e += 256) {
            const int k = e % GK, n = e / GK;
            const int gk = k0 + k, gn = n0 + n;
            Bs[k][n] = (gk < K && gn < Nc) ? B[gk + (size_t)ldb * gn] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < GK; kk += 4) {
            const int kr = kk + (lane >> 4);
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const double a = As[kr][wm + i * 16 + (lane & 15)];
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    const double b = Bs[kr][wn + j * 16 + (lane & 15)];
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[i][j], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }
    // D[r][c] of a 16x16 block held by v_mfma_f64_16x16x4f64:
    // c = lane % 16, r = lane / 16 + 4 * item
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int it = 0; it < 4; ++it) {
                const int r = m0 + wm + i * 16 + (lane >> 4) + 4 * it;
                const int c = n0 + wn + j * 16 + (lane & 15);
                if (r < M && c < Nc) C[r + (size_t)ldc * c] -= acc[i][j][it];
            }
}

// Plain FP64 FMA reference path for the same update (selected by ILH_GEMM_PLAIN=1).
__global__ void k_gemm_sub_plain(int M, int Nc, int K, const double* __restrict__ A, int lda,
                                 const double* __restrict__ B, int ldb, double* __restrict__ C, int ldc)
{
    const int r = blockIdx.x * 64 + (threadIdx.x & 63);
    const int c = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (r >= M || c >= Nc) return;
    double s = 0.0;
    for (int k = 0; k < K; ++k) s += A[r + (size_t)lda * k] * B[k + (size_t)ldb * c];
    C[r + (size_t)ldc * c] -= s;
}

static bool gemm_plain()
{
    static int v = -1;
    if (v < 0) {
        const char* e = getenv("ILH_GEMM_PLAIN");
        v = (e && e[0] == '1') ? 1 : 0;
    }
    return v == 1;
}

static void gemm_sub(hipStream_t st, int M, int Nc, int K, const double* A, int lda, const double* B, int ldb,
                     double* C, int ldc)
{
    if (M <= 0 || Nc <= 0 || K <= 0) return;
    if (gemm_plain()) {
        k_gemm_sub_plain<<<dim3((M + 63) / 64, (Nc + 3) / 4), 256, 0, st>>>(M, Nc, K, A, lda, B, ldb, C, ldc);
    } else {
        k_gemm_sub<<<dim3((M + GT - 1) / GT, (Nc + GT - 1) / GT), 256, 0, st>>>(M, Nc, K, A, lda, B, ldb, C, ldc);
    }
    ILH_HIP(hipGetLastError());
}

// ------------------------------------------------------------------- LU
static constexpr int NB = 32;

// Unblocked LU of the panel A[k0:n, k0:k0+nb] with partial pivoting (one
// work-group).  Row swaps are applied inside the panel only.
__global__ void __launch_bounds__(1024) k_panel(double* __restrict__ A, int n, int lda, int k0, int nb,
                                                int* __restrict__ piv, int* __restrict__ info)
{
    __shared__ double rv[1024];
    __shared__ int ri[1024];
    __shared__ double prow[NB];
    const int tid = threadIdx.x;
    for (int kk = 0; kk < nb; ++kk) {
        const int col = k0 + kk;
        double best = -1.0;
        int bi = n;
        for (int i = col + tid; i < n; i += 1024) {
            const double v = fabs(A[i + (size_t)lda * col]);
            if (v > best) { best = v; bi = i; }
        }
        rv[tid] = best;
        ri[tid] = bi;
        __syncthreads();
        for (int off = 512; off > 0; off >>= 1) {
            if (tid < off) {
                const double o = rv[tid + off];
                const int oi = ri[tid + off];
                if (o > rv[tid] || (o == rv[tid] && oi < ri[tid])) { rv[tid] = o; ri[tid] = oi; }
            }
            __syncthreads();
        }
        const int pr = ri[0];
        const double pv = rv[0];
        __syncthreads();
        if (tid == 0) {
            piv[col] = pr;
            if (pv == 0.0 && *info == 0) *info = col + 1;
        }
        if (pv != 0.0) {
            if (pr != col && tid < nb) {
                const int j = k0 + tid;
                const double t = A[col + (size_t)lda * j];
                A[col + (size_t)lda * j] = A[pr + (size_t)lda * j];
                A[pr + (size_t)lda * j] = t;
            }
            __syncthreads();
            if (tid < nb) prow[tid] = A[col + (size_t)lda * (k0 + tid)];
            __syncthreads();
            const double inv = 1.0 / prow[kk];
            for (int i = col + 1 + tid; i < n; i += 1024) {
                const double l = A[i + (size_t)lda * col] * inv;
                A[i + (size_t)lda * col] = l;
                for (int j = kk + 1; j < nb; ++j) A[i + (size_t)lda * (k0 + j)] -= l * prow[j];
            }
        }
        __syncthreads();
    }
}

// apply the panel's row interchanges to columns outside [k0, k0+nb)
__global__ void k_swap_rows(double* __restrict__ A, int n, int lda, int k0, int nb, const int* __restrict__ piv)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n - nb) return;
    if (j >= k0) j += nb;
    for (int kk = 0; kk < nb; ++kk) {
        const int r = k0 + kk, pr = piv[r];
        if (pr != r) {
            const double t = A[r + (size_t)lda * j];
            A[r + (size_t)lda * j] = A[pr + (size_t)lda * j];
            A[pr + (size_t)lda * j] = t;
        }
    }
}

// X[k0:k0+nb, j] = L11^{-1} X[k0:k0+nb, j] (unit lower, L11 at LU[k0,k0]) for each column j
__global__ void k_trsm_lower(const double* __restrict__ L, int ldl, int k0, int nb, double* __restrict__ X, int ldx,
                             int ncols, int col0)
{
    __shared__ double Ls[NB][NB + 1];
    for (int e = threadIdx.x; e < nb * nb; e += blockDim.x) {
        const int r = e % nb, c = e / nb;
        Ls[r][c] = L[(k0 + r) + (size_t)ldl * (k0 + c)];
    }
    __syncthreads();
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ncols) return;
    double* x = X + (size_t)ldx * (col0 + j) + k0;
    double v[NB];
    for (int r = 0; r < nb; ++r) v[r] = x[r];
    for (int r = 0; r < nb; ++r)
        for (int c = 0; c < r; ++c) v[r] -= Ls[r][c] * v[c];
    for (int r = 0; r < nb; ++r) x[r] = v[r];
}

// X[k0:k0+nb, j] = U11^{-1} X[k0:k0+nb, j] (upper, non-unit)
__global__ void k_trsm_upper(const double* __restrict__ U, int ldu, int k0, int nb, double* __restrict__ X, int ldx,
                             int ncols)
{
    __shared__ double Us[NB][NB + 1];
    for (int e = threadIdx.x; e < nb * nb; e += blockDim.x) {
        const int r = e % nb, c = e / nb;
        Us[r][c] = U[(k0 + r) + (size_t)ldu * (k0 + c)];
    }
    __syncthreads();
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ncols) return;
    double* x = X + (size_t)ldx * j + k0;
    double v[NB];
    for (int r = 0; r < nb; ++r) v[r] = x[r];
    for (int r = nb - 1; r >= 0; --r) {
        double s = v[r];
        for (int c = r + 1; c < nb; ++c) s -= Us[r][c] * v[c];
        v[r] = s / Us[r][r];
    }
    for (int r = 0; r < nb; ++r) x[r] = v[r];
}

int lu_factor(ilh_plan* pl, double* A, int n, int* piv, int* info)
{
    hipStream_t st = pl->ctx->stream;
    ILH_HIP(hipMemsetAsync(info, 0, sizeof(int), st));
    for (int k0 = 0; k0 < n; k0 += NB) {
        const int nb = std::min(NB, n - k0);
        k_panel<<<1, 1024, 0, st>>>(A, n, n, k0, nb, piv, info);
        if (n - nb > 0) k_swap_rows<<<(n - nb + 255) / 256, 256, 0, st>>>(A, n, n, k0, nb, piv);
        const int k1 = k0 + nb, rest = n - k1;
        if (rest > 0) {
            k_trsm_lower<<<(rest + 127) / 128, 128, 0, st>>>(A, n, k0, nb, A, n, rest, k1);
            gemm_sub(st, rest, rest, nb, A + k1 + (size_t)n * k0, n, A + k0 + (size_t)n * k1, n,
                     A + k1 + (size_t)n * k1, n);
        }
        ILH_HIP(hipGetLastError());
    }
    int h = 0;
    ILH_HIP(hipMemcpyAsync(&h, info, sizeof(int), hipMemcpyDeviceToHost, st));
    ILH_HIP(hipStreamSynchronize(st));
    return h;
}

__global__ void k_apply_piv(double* __restrict__ B, int n, int ldb, int ncols, const int* __restrict__ piv)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ncols) return;
    double* b = B + (size_t)ldb * j;
    for (int k = 0; k < n; ++k) {
        const int p = piv[k];
        if (p != k) {
            const double t = b[k];
            b[k] = b[p];
            b[p] = t;
        }
    }
}

// B <- A^{-1} B for the LU factors of A (n x n) and nrhs columns of B.
void lu_solve_multi(ilh_plan* pl, const double* LU, const int* piv, int n, double* B, int nrhs)
{
    hipStream_t st = pl->ctx->stream;
    k_apply_piv<<<(nrhs + 255) / 256, 256, 0, st>>>(B, n, n, nrhs, piv);
    for (int k0 = 0; k0 < n; k0 += NB) {
        const int nb = std::min(NB, n - k0);
        k_trsm_lower<<<(nrhs + 127) / 128, 128, 0, st>>>(LU, n, k0, nb, B, n, nrhs, 0);
        const int k1 = k0 + nb;
        gemm_sub(st, n - k1, nrhs, nb, LU + k1 + (size_t)n * k0, n, B + k0, n, B + k1, n);
    }
    const int last = ((n - 1) / NB) * NB;
    for (int k0 = last; k0 >= 0; k0 -= NB) {
        const int nb = std::min(NB, n - k0);
        k_trsm_upper<<<(nrhs + 127) / 128, 128, 0, st>>>(LU, n, k0, nb, B, n, nrhs);
        gemm_sub(st, k0, nrhs, nb, LU + (size_t)n * k0, n, B + k0, n, B, n);
    }
    ILH_HIP(hipGetLastError());
}

// ------------------------------------------------------------- helpers
__global__ void k_transpose(const double* __restrict__ A, double* __restrict__ At, int rows, int cols, int ldt)
{
    __shared__ double t[32][33];
    const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
    for (int k = threadIdx.y; k < 32; k += 8) {
        const int r = bx + threadIdx.x, c = by + k;
        if (r < rows && c < cols) t[k][threadIdx.x] = A[r + (size_t)rows * c];
    }
    __syncthreads();
    for (int k = threadIdx.y; k < 32; k += 8) {
        const int c = by + threadIdx.x, r = bx + k;
        if (r < rows && c < ldt) At[c + (size_t)ldt * r] = c < cols ? t[threadIdx.x][k] : 0.0;
    }
}

// A: rows x cols column-major -> At: row-major rows x cols with leading
// dimension ldt >= cols (padding columns zeroed)
void transpose(hipStream_t st, const double* A, double* At, int rows, int cols, int ldt)
{
    k_transpose<<<dim3((rows + 31) / 32, (ldt + 31) / 32), dim3(32, 8), 0, st>>>(A, At, rows, cols, ldt);
    ILH_HIP(hipGetLastError());
}

// One wave per row: y[i] = alpha * dot(M[i, :], x) (+ y[i]).  M row-major n x m
// with leading dimension ld (a multiple of 2, rows 16-byte aligned).
__device__ __forceinline__ double wave_dot(const double* __restrict__ row, const double* __restrict__ x, int m, int lane)
{
    if (((uintptr_t)x & 15) != 0) {   // caller's vector not 16-byte aligned: plain loads
        double s = 0.0;
        for (int k = lane; k < m; k += 64) s += row[k] * x[k];
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
        return s;
    }
    // 16-byte loads, four independent partial sums per lane
    const double2* r2 = (const double2*)row;
    const double2* x2 = (const double2*)x;
    const int m2 = m >> 1;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int k = lane;
    for (; k + 192 < m2; k += 256) {
        const double2 a0 = r2[k], a1 = r2[k + 64], a2 = r2[k + 128], a3 = r2[k + 192];
        const double2 b0 = x2[k], b1 = x2[k + 64], b2 = x2[k + 128], b3 = x2[k + 192];
        s0 += a0.x * b0.x + a0.y * b0.y;
        s1 += a1.x * b1.x + a1.y * b1.y;
        s2 += a2.x * b2.x + a2.y * b2.y;
        s3 += a3.x * b3.x + a3.y * b3.y;
    }
    for (; k < m2; k += 64) {
        const double2 a0 = r2[k], b0 = x2[k];
        s0 += a0.x * b0.x + a0.y * b0.y;
    }
    if ((m & 1) && lane == 0) s1 += row[m - 1] * x[m - 1];
    double s = (s0 + s1) + (s2 + s3);
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    return s;
}

__global__ void k_matvec_rows(const double* __restrict__ M, int ld, const double* __restrict__ x, double* __restrict__ y,
                              int n, int m, double alpha, int accumulate)
{
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (i >= n) return;
    const double s = wave_dot(M + (size_t)ld * i, x, m, lane);
    if (lane == 0) y[i] = alpha * s + (accumulate ? y[i] : 0.0);
}

void matvec_rows(hipStream_t st, const double* M, int ld, const double* x, double* y, int n, int m, double alpha,
                 int acc)
{
    if (n <= 0) return;
    k_matvec_rows<<<(n + 3) / 4, 256, 0, st>>>(M, ld, x, y, n, m, alpha, acc);
    ILH_HIP(hipGetLastError());
}

// Dirichlet surface step: s = M+ f+ (+ M- f-),  amp = [s; f+ - f-], s_out = s
__global__ void k_dir_amp(const double* __restrict__ Mp, const double* __restrict__ Mm, int ld,
                          const double* __restrict__ fp, const double* __restrict__ fm, double* __restrict__ amp,
                          double* __restrict__ s_out, int n)
{
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (i >= n) return;
    double s = wave_dot(Mp + (size_t)ld * i, fp, n, lane);
    if (fm) s += wave_dot(Mm + (size_t)ld * i, fm, n, lane);
    if (lane == 0) {
        amp[i] = s;
        amp[n + i] = fp[i] - (fm ? fm[i] : 0.0);
        if (s_out) s_out[i] = s;
    }
}

void dirichlet_amp(hipStream_t st, const double* Mp, const double* Mm, int ld, const double* fp, const double* fm,
                   double* amp, double* s_out, int n)
{
    if (n <= 0) return;
    k_dir_amp<<<(n + 3) / 4, 256, 0, st>>>(Mp, Mm, ld, fp, fm, amp, s_out, n);
    ILH_HIP(hipGetLastError());
}

// Neumann surface step: df = A+ v+ + A- v-, ds = B+ v+ + B- v-;
// primal amplitudes [df; v+ - v-], dual amplitudes [df; -ds]
__global__ void k_neu_amp(const double* __restrict__ Ap, const double* __restrict__ Am, const double* __restrict__ Bp,
                          const double* __restrict__ Bm, int ld, const double* __restrict__ vp,
                          const double* __restrict__ vm, double* __restrict__ af, double* __restrict__ as,
                          double* __restrict__ df, double* __restrict__ dsv, int n)
{
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (i >= n) return;
    double a = wave_dot(Ap + (size_t)ld * i, vp, n, lane);
    double b = wave_dot(Bp + (size_t)ld * i, vp, n, lane);
    if (vm) {
        a += wave_dot(Am + (size_t)ld * i, vm, n, lane);
        b += wave_dot(Bm + (size_t)ld * i, vm, n, lane);
    }
    if (lane == 0) {
        af[i] = a;
        af[n + i] = vp[i] - (vm ? vm[i] : 0.0);
        as[i] = a;
        as[n + i] = -b;
        df[i] = a;
        dsv[i] = b;
    }
}

void neumann_amp(hipStream_t st, const double* Ap, const double* Am, const double* Bp, const double* Bm, int ld,
                 const double* vp, const double* vm, double* af, double* as, double* df, double* dsv, int n)
{
    if (n <= 0) return;
    k_neu_amp<<<(n + 3) / 4, 256, 0, st>>>(Ap, Am, Bp, Bm, ld, vp, vm, af, as, df, dsv, n);
    ILH_HIP(hipGetLastError());
}

__global__ void k_finite(const double* __restrict__ a, size_t n, int* flag)
{
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
        if (!isfinite(a[k])) { atomicOr(flag, 1); return; }
}

void check_finite(hipStream_t st, const double* a, size_t n, int* flag)
{
    if (!n) return;
    size_t blocks = std::min<size_t>(1024, (n + 255) / 256);
    k_finite<<<(unsigned)blocks, 256, 0, st>>>(a, n, flag);
    ILH_HIP(hipGetLastError());
}

}  // namespace ilh
