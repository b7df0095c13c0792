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
// solves run on the FP64 matrix cores (dgemm, ilh_chol.hip).  It serves the
// Schur kinds that are not symmetrised (the Neumann and Stokes matrices) and
// the fallback of the Cholesky path.
#include "ilh_internal.hpp"

#include <algorithm>
#include <cstdlib>

namespace ilh {

typedef double double4_t __attribute__((ext_vector_type(4)));

// ------------------------------------------------------------ assembly
static constexpr int AT = 16;   // points per assembly tile edge

// doubles of the kernel table staged per tile (12 KB).  Tiles whose box is
// larger read the table from memory instead; on the 1024^2 headline bodies
// the occupancy a smaller box allows outweighs those reads (per N = 1885
// matrix under the profiler: 57.9 us at 4096, 50.2 at 2048, 48.6 at 1536,
// 58.0 at 1024)
static constexpr int GBOX = 1536;

// One 16x16 tile of (p, q) entries per work-group.  The table entries a tile
// touches lie in a box of |dbx + u|, |dby + v| whose sides are the two tiles'
// extents plus the 9-wide window (points of a tile are neighbours on a
// curve), independent of how far apart the two tiles are; the box is staged
// in LDS when it fits, and the table is read from memory otherwise.
// ROWMAJOR: K[q + ldk p] (lanes along q: coalesced stores), else K[p + ldk q].
// SYM: one term per point and identical windows in A and B (E and R of one
// grid kind): the kernel sum is symmetric in (p, q), so only tiles on or
// above the diagonal run and also write the mirrored entry
// K[q, p] = factor coefA_q coefB_p (same sum).
template <bool ROWMAJOR, bool SYM>
__global__ void __launch_bounds__(AT * AT) k_assemble(const Term* __restrict__ TA, int TAn, const Term* __restrict__ TB,
                                                      int TBn, int N, const double* __restrict__ G, int Tn,
                                                      double factor, int delta, double* __restrict__ K, int ldk)
{
    __shared__ Term sa[AT * 2], sb[AT * 2];
    __shared__ double sG[GBOX];
    __shared__ int sbox[5];
    __shared__ double cAq[AT], cBp[AT];   // SYM: coefficients of the mirrored entries
    if (SYM && blockIdx.y < blockIdx.x) return;
    const int p0 = blockIdx.x * AT, q0 = blockIdx.y * AT;
    const int tid = threadIdx.x + AT * threadIdx.y;
    const int na = AT * TAn, nb = AT * TBn;
    if (tid < na) {
        const int p = p0 + tid / TAn;
        if (p < N) sa[tid] = TA[(size_t)p * TAn + tid % TAn];
    } else if (tid >= 64 && tid < 64 + nb) {
        const int k = tid - 64;
        const int q = q0 + k / TBn;
        if (q < N) sb[k] = TB[(size_t)q * TBn + k % TBn];
    } else if (SYM && tid >= 128 && tid < 128 + AT) {
        const int k = tid - 128;
        if (q0 + k < N) cAq[k] = TA[q0 + k].coef;
    } else if (SYM && tid >= 192 && tid < 192 + AT) {
        const int k = tid - 192;
        if (p0 + k < N) cBp[k] = TB[p0 + k].coef;
    }
    __syncthreads();
    if (!delta && tid < 64) {
        // extents of the two tiles' window bases (wave 0: lanes < 32 the A
        // terms, lanes >= 32 the B terms)
        const int lane = tid;
        int ax0 = 1 << 30, ax1 = -(1 << 30), ay0 = 1 << 30, ay1 = -(1 << 30);
        int bx0 = 1 << 30, bx1 = -(1 << 30), by0 = 1 << 30, by1 = -(1 << 30);
        if (lane < 32) {
            if (lane < na && p0 + lane / TAn < N) { ax0 = ax1 = sa[lane].bx; ay0 = ay1 = sa[lane].by; }
        } else {
            const int k = lane - 32;
            if (k < nb && q0 + k / TBn < N) { bx0 = bx1 = sb[k].bx; by0 = by1 = sb[k].by; }
        }
        for (int off = 32; off > 0; off >>= 1) {
            ax0 = min(ax0, __shfl_xor(ax0, off, 64)); ax1 = max(ax1, __shfl_xor(ax1, off, 64));
            ay0 = min(ay0, __shfl_xor(ay0, off, 64)); ay1 = max(ay1, __shfl_xor(ay1, off, 64));
            bx0 = min(bx0, __shfl_xor(bx0, off, 64)); bx1 = max(bx1, __shfl_xor(bx1, off, 64));
            by0 = min(by0, __shfl_xor(by0, off, 64)); by1 = max(by1, __shfl_xor(by1, off, 64));
        }
        if (lane == 0) {
            // signed offsets d = dbx + u - 4 (and the same in y) the tile pair
            // needs: the box [x0, x1] x [y0, y1], staged unfolded, with the
            // (even) table mirrored into it and zero beyond its edge
            const int x0 = ax0 - bx1 - 4, x1 = ax1 - bx0 + 4;
            const int y0 = ay0 - by1 - 4, y1 = ay1 - by0 + 4;
            const int bw = x1 - x0 + 1, bh = y1 - y0 + 1;
            const bool fits = bw > 0 && bh > 0 && (long)bw * bh <= GBOX;
            sbox[0] = fits ? 1 : -1;
            sbox[1] = x0;
            sbox[2] = y0;
            sbox[3] = bw;
            sbox[4] = bh;
        }
    }
    __syncthreads();
    const bool boxed = !delta && sbox[0] > 0;
    const int bx0 = sbox[1], by0 = sbox[2], bw = sbox[3], bh = sbox[4];
    if (boxed) {
        for (int e = tid; e < bw * bh; e += AT * AT) {
            const int ix = abs(bx0 + e % bw), iy = abs(by0 + e / bw);
            sG[e] = (ix < Tn && iy < Tn) ? G[(size_t)Tn * iy + ix] : 0.0;
        }
    }
    __syncthreads();
    const int pi = ROWMAJOR ? threadIdx.y : threadIdx.x;
    const int qi = ROWMAJOR ? threadIdx.x : threadIdx.y;
    const int p = p0 + pi, q = q0 + qi;
    if (p >= N || q >= N) return;
    double acc = 0.0, wsym = 0.0;
    for (int ia = 0; ia < TAn; ++ia) {
        const Term& a = sa[pi * TAn + ia];
        for (int ib = 0; ib < TBn; ++ib) {
            const Term& b = sb[qi * TBn + ib];
            if (a.gk != b.gk) continue;
            // SYM kinds pair plain DDF windows, whose fifth tap is zero: the
            // correlations' end entries are then exact zeros and are left out
            // (the sums below skip only zero products, in the same order)
            constexpr int NW = SYM ? 4 : 5, U0 = 5 - NW, U1 = 4 + NW;
            double cx[9], cy[9];
#pragma unroll
            for (int u = 0; u < 9; ++u) { cx[u] = 0.0; cy[u] = 0.0; }
#pragma unroll
            for (int i = 0; i < NW; ++i)
#pragma unroll
                for (int k = 0; k < NW; ++k) {
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
                for (int v = U0; v < U1; ++v) {
                    const int jj = abs(dby + v - 4);
                    if (cy[v] == 0.0 || jj >= Tn) continue;   // kernel zero beyond the table
                    double r = 0.0;
                    if (boxed) {
                        // staged box: the row of offset dby + v - 4, from offset
                        // dbx - 4 on, holds the table values the u loop needs
                        // in order (zeros beyond the table add nothing).  Only
                        // the integer index is offset: a pointer before an LDS
                        // array, used as a generic pointer, can land outside
                        // the LDS aperture.
                        const int row = bw * (dby + v - 4 - by0) + (dbx - 4 - bx0);
#pragma unroll
                        for (int u = U0; u < U1; ++u) r += cx[u] * sG[row + u];
                    } else {
                        const double* Gr = G + (size_t)Tn * jj;
#pragma unroll
                        for (int u = U0; u < U1; ++u) {
                            const int ii = abs(dbx + u - 4);
                            if (ii < Tn) r += cx[u] * Gr[ii];
                        }
                    }
                    s += cy[v] * r;
                }
            }
            acc += a.coef * b.coef * s;
            wsym = s;
        }
    }
    if (ROWMAJOR)
        K[q + (size_t)ldk * p] = factor * acc;
    else
        K[p + (size_t)ldk * q] = factor * acc;
    if (SYM && blockIdx.x != blockIdx.y) {
        const double mirror = factor * (cAq[qi] * cBp[pi] * wsym);
        if (ROWMAJOR)
            K[p + (size_t)ldk * q] = mirror;
        else
            K[q + (size_t)ldk * p] = mirror;
    }
}

static void launch_assemble(ilh_plan* pl, int opA, int opB, const double* tab, int Tn, double factor, bool delta,
                            double* K, int ldk, bool rowmajor)
{
    const int N = pl->N;
    if (N == 0) return;
    const Term* A = get_terms(pl, opA);
    const Term* B = get_terms(pl, opB);
    dim3 grid((N + AT - 1) / AT, (N + AT - 1) / AT);
    const int ta = op_terms_per_point(opA), tb = op_terms_per_point(opB);
    if (ta > 2 || tb > 2) fail(ILH_ERR_ARG, "assembly supports at most two terms per point");
    // E and R of one grid kind share their windows (one term per point)
    // (E and E of one grid kind: the symmetrised Schur matrices, any kernel)
    const bool sym = (!delta && ((opA == OP_E_PRIMAL && opB == OP_R_PRIMAL) || (opA == OP_E_DUAL && opB == OP_R_DUAL))) ||
                     (opA == OP_E_PRIMAL && opB == OP_E_PRIMAL) || (opA == OP_E_DUAL && opB == OP_E_DUAL);
    const int dl = delta ? 1 : 0;
    hipStream_t st = pl->ctx->stream;
    if (rowmajor && sym)
        k_assemble<true, true><<<grid, dim3(AT, AT), 0, st>>>(A, ta, B, tb, N, tab, Tn, factor, dl, K, ldk);
    else if (rowmajor)
        k_assemble<true, false><<<grid, dim3(AT, AT), 0, st>>>(A, ta, B, tb, N, tab, Tn, factor, dl, K, ldk);
    else if (sym)
        k_assemble<false, true><<<grid, dim3(AT, AT), 0, st>>>(A, ta, B, tb, N, tab, Tn, factor, dl, K, ldk);
    else
        k_assemble<false, false><<<grid, dim3(AT, AT), 0, st>>>(A, ta, B, tb, N, tab, Tn, factor, dl, K, ldk);
    ILH_HIP(hipGetLastError());
}

void assemble(ilh_plan* pl, int opA, int opB, int gk, double factor, bool delta_kernel, double* K, int ldk,
              bool rowmajor)
{
    (void)gk;
    if (pl->N == 0) return;
    if (!delta_kernel) ensure_lgf(pl);
    launch_assemble(pl, opA, opB, delta_kernel ? nullptr : pl->lgf.as<double>(), delta_kernel ? 1 : pl->lgf_mx,
                    factor, delta_kernel, K, ldk > 0 ? ldk : pl->N, rowmajor);
}

// K[p,q] = factor * sum over term pairs of the kernel tab (Tn x Tn, even in
// both axes, zero beyond the table) correlated with the two points' windows
void assemble_table(ilh_plan* pl, int opA, int opB, const double* tab, int Tn, double factor, double* K)
{
    launch_assemble(pl, opA, opB, tab, Tn, factor, false, K, pl->N, false);
}

// C[M x Nc] -= A[M x K] B[K x Nc] on stream st (the trailing updates of the
// blocked LU and of its multi-right-hand-side solves), on the FP64 matrix cores
static void gemm_sub(hipStream_t st, int M, int Nc, int K, const double* A, int lda, const double* B, int ldb,
                     double* C, int ldc)
{
    if (M <= 0 || Nc <= 0 || K <= 0) return;
    dgemm(st, 0, 0, M, Nc, K, -1.0, A, lda, B, ldb, 1.0, C, ldc);
}

// ------------------------------------------------------------------- LU
static constexpr int NB = 32;

// Unblocked LU of the panel A[k0:n, k0:k0+nb] with partial pivoting (one
// work-group, global memory).  Used when the panel's sub-columns do not fit
// in LDS (n - k0 > 8192).  Row swaps are applied inside the panel only.
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

// Work-group barrier ordering LDS accesses only: outstanding global stores of
// the calling waves need not complete (a plain __syncthreads waits for them).
__device__ __forceinline__ void lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// max |value| over the work-group with the smallest row index among ties
__device__ __forceinline__ void argmax_block(double& best, int& bi, double* wv, int* wi)
{
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(best, off, 64);
        const int oi = __shfl_xor(bi, off, 64);
        if (o > best || (o == best && oi < bi)) { best = o; bi = oi; }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) { wv[wave] = best; wi[wave] = bi; }
    lds_sync();
    // every thread reduces the per-wave results itself (same order everywhere)
    const int nw = blockDim.x >> 6;
    best = wv[0];
    bi = wi[0];
    for (int w = 1; w < nw; ++w) {
        const double o = wv[w];
        const int oi = wi[w];
        if (o > best || (o == best && oi < bi)) { best = o; bi = oi; }
    }
    lds_sync();   // wv/wi may be overwritten by the next call
}

static constexpr int PT = 512;   // threads of the panel work-group

// Composite of the row interchanges (r_k <-> p_k, k = 0..cnt-1, in order) on
// the rows they touch: afterwards row key[i] holds the old row val[i].  Run by
// one thread with the table in registers (MAXK >= 2 cnt); returns the count.
template <int MAXK>
__device__ int swap_composite(const int* r, const int* pv, int cnt, int* key_out, int* val_out)
{
    int key[MAXK], val[MAXK];
    int nk = 0;
#pragma unroll
    for (int i = 0; i < MAXK; ++i) { key[i] = -1; val[i] = -1; }
    for (int k = 0; k < cnt; ++k) {
        const int a = r[k], b = pv[k];
        if (a == b) continue;
        int ia = -1, ib = -1;
#pragma unroll
        for (int i = 0; i < MAXK; ++i) {
            if (key[i] == a) ia = i;
            if (key[i] == b) ib = i;
        }
        if (ia < 0) {
#pragma unroll
            for (int i = 0; i < MAXK; ++i)
                if (i == nk) { key[i] = a; val[i] = a; }
            ia = nk++;
        }
        if (ib < 0) {
#pragma unroll
            for (int i = 0; i < MAXK; ++i)
                if (i == nk) { key[i] = b; val[i] = b; }
            ib = nk++;
        }
        int va = 0, vb = 0;
#pragma unroll
        for (int i = 0; i < MAXK; ++i) {
            if (i == ia) va = val[i];
            if (i == ib) vb = val[i];
        }
#pragma unroll
        for (int i = 0; i < MAXK; ++i) {
            if (i == ia) val[i] = vb;
            if (i == ib) val[i] = va;
        }
    }
#pragma unroll
    for (int i = 0; i < MAXK; ++i)
        if (i < nk) { key_out[i] = key[i]; val_out[i] = val[i]; }
    return nk;
}

// Panel LU with partial pivoting, W columns at a time in LDS (one work-group
// of PT threads).  Sub-panel s holds A[k0:n, k0+s:k0+s+W] as columns of m =
// n - k0 rows; after factoring it, its row swaps reach the rest of the panel
// and the rest of the panel is updated (unit-lower solve of its W rows, then a
// rank-W update below), one global read-modify-write per sub-panel.  At the
// end: pmap = {count, (row, source row) pairs} of the panel's interchanges for
// the columns outside the panel.
template <int W>
__global__ void __launch_bounds__(PT) k_panel_lds(double* __restrict__ A, int n, int lda, int k0, int nb,
                                                    int* __restrict__ piv, int* __restrict__ info,
                                                    int* __restrict__ pmap)
{
    extern __shared__ double sp[];                 // W x m, column-major
    __shared__ double wv[16];
    __shared__ int wi[16];
    __shared__ double Us[W][NB];
    __shared__ int spiv[NB], srow[NB];
    __shared__ double sbest[NB];
    __shared__ int skey[2 * NB], sval[2 * NB];
    __shared__ int snk;
    const int m = n - k0, tid = threadIdx.x;
    constexpr int RB = W >= 8 ? 2 : 4;   // rows per thread per update batch (register budget)
    for (int s0 = 0; s0 < nb; s0 += W) {
        const int w = min(W, nb - s0);
        // 16 loads in flight per thread, then their LDS stores
        for (int base = 0; base < w * m; base += 16 * PT) {
            double tmp[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const int e = base + u * PT + tid;
                const int c = e / m, r = e - c * m;
                tmp[u] = e < w * m ? A[(k0 + r) + (size_t)lda * (k0 + s0 + c)] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const int e = base + u * PT + tid;
                if (e < w * m) sp[e] = tmp[u];
            }
        }
        __syncthreads();
        for (int kk = 0; kk < w; ++kk) {
            const int col = s0 + kk;                // panel row of the diagonal
            double* x = sp + (size_t)kk * m;
            double best = -1.0;
            int bi = m;
            for (int r = col + tid; r < m; r += PT) {
                const double v = fabs(x[r]);
                if (v > best) { best = v; bi = r; }
            }
            argmax_block(best, bi, wv, wi);
            if (tid == 0) {
                spiv[col] = bi;
                srow[col] = col;
                sbest[col] = best;
            }
            if (best != 0.0) {
                if (bi != col && tid < w) {
                    double* y = sp + (size_t)tid * m;
                    const double t = y[col];
                    y[col] = y[bi];
                    y[bi] = t;
                }
                lds_sync();
                const double inv = 1.0 / x[col];
                for (int r = col + 1 + tid; r < m; r += PT) {
                    const double l = x[r] * inv;
                    x[r] = l;
                    for (int c = kk + 1; c < w; ++c) sp[(size_t)c * m + r] -= l * sp[(size_t)c * m + col];
                }
            }
            lds_sync();
        }
        if (tid == 0)
            for (int kk = 0; kk < w; ++kk) {
                piv[k0 + s0 + kk] = k0 + spiv[s0 + kk];
                if (sbest[s0 + kk] == 0.0 && *info == 0) *info = k0 + s0 + kk + 1;
            }
        for (int e = tid; e < w * m; e += PT) {
            const int c = e / m, r = e - c * m;
            A[(k0 + r) + (size_t)lda * (k0 + s0 + c)] = sp[(size_t)c * m + r];
        }
        if (tid == 0) snk = swap_composite<2 * W>(srow + s0, spiv + s0, w, skey, sval);
        __syncthreads();
        const int j0 = s0 + w, rest = nb - j0;
        // this sub-panel's interchanges on the panel columns before and after it
        if (tid < nb && (tid < s0 || tid >= j0)) {
            double* col = A + k0 + (size_t)lda * (k0 + tid);
            const int nk = snk;
            double v[2 * W];
#pragma unroll
            for (int i = 0; i < 2 * W; ++i) v[i] = i < nk ? col[sval[i]] : 0.0;
#pragma unroll
            for (int i = 0; i < 2 * W; ++i)
                if (i < nk) col[skey[i]] = v[i];
        }
        __syncthreads();
        if (rest > 0) {
            // U rows: unit-lower solve of the W x rest block
            if (tid < rest) {
                double* col = A + k0 + (size_t)lda * (k0 + j0 + tid);
                double u[W];
#pragma unroll
                for (int c = 0; c < W; ++c) u[c] = c < w ? col[s0 + c] : 0.0;
#pragma unroll
                for (int c = 0; c < W; ++c)
                    if (c < w) {
#pragma unroll
                        for (int k = 0; k < c; ++k) u[c] -= sp[(size_t)k * m + s0 + c] * u[k];
                    }
#pragma unroll
                for (int c = 0; c < W; ++c) {
                    if (c < w) col[s0 + c] = u[c];
                    Us[c][tid] = c < w ? u[c] : 0.0;
                }
            }
            __syncthreads();
            // rank-w update of the rows below: per thread RB rows x 8 columns of
            // loads in flight, then the arithmetic, then the stores
            for (int r0 = j0 + tid; r0 < m; r0 += RB * PT) {
#pragma unroll 1
                for (int jb = 0; jb < rest; jb += 8) {
                    double acc[RB][8];
#pragma unroll
                    for (int q = 0; q < RB; ++q) {
                        const int r = r0 + q * PT;
                        const double* a = A + (k0 + r) + (size_t)lda * (k0 + j0 + jb);
#pragma unroll
                        for (int j = 0; j < 8; ++j) acc[q][j] = (r < m && jb + j < rest) ? a[(size_t)lda * j] : 0.0;
                    }
#pragma unroll
                    for (int q = 0; q < RB; ++q) {
                        const int r = r0 + q * PT;
#pragma unroll
                        for (int c = 0; c < W; ++c) {
                            const double l = (r < m && c < w) ? sp[(size_t)c * m + r] : 0.0;
#pragma unroll
                            for (int j = 0; j < 8; ++j) acc[q][j] -= l * Us[c][jb + j];
                        }
                    }
#pragma unroll
                    for (int q = 0; q < RB; ++q) {
                        const int r = r0 + q * PT;
                        double* a = A + (k0 + r) + (size_t)lda * (k0 + j0 + jb);
#pragma unroll
                        for (int j = 0; j < 8; ++j)
                            if (r < m && jb + j < rest) a[(size_t)lda * j] = acc[q][j];
                    }
                }
            }
        }
        __syncthreads();
    }
    // the panel's interchange map for the columns outside it: the swaps
    // replayed on a row-index table in the (now free) LDS, then every moved
    // row listed
    int* pl_ = (int*)sp;
    __shared__ int cnt;
    for (int r = tid; r < m; r += PT) pl_[r] = r;
    if (tid == 0) cnt = 0;
    __syncthreads();
    if (tid == 0)
        for (int k = 0; k < nb; ++k) {
            const int q = spiv[k];
            if (q != k) {
                const int t = pl_[k];
                pl_[k] = pl_[q];
                pl_[q] = t;
            }
        }
    __syncthreads();
    for (int r = tid; r < m; r += PT)
        if (pl_[r] != r) {
            const int i = atomicAdd(&cnt, 1);
            pmap[1 + 2 * i] = k0 + r;
            pmap[2 + 2 * i] = k0 + pl_[r];
        }
    __syncthreads();
    if (tid == 0) pmap[0] = cnt;
}

// Rows key <- old rows val (the panel's interchange map) on the columns outside
// [k0, k0+nb): 64 columns per work-group, all reads before any write.
__global__ void __launch_bounds__(256) k_apply_pmap(double* __restrict__ A, int n, int lda, int k0, int nb,
                                                    const int* __restrict__ pmap)
{
    __shared__ double vals[64][2 * NB + 1];
    __shared__ int pk[2 * NB], pvv[2 * NB];
    const int cnt = pmap[0];
    for (int i = threadIdx.x; i < cnt; i += 256) {
        pk[i] = pmap[1 + 2 * i];
        pvv[i] = pmap[2 + 2 * i];
    }
    __syncthreads();
    const int nout = n - nb;
    const int jb = blockIdx.x * 64;
    for (int e = threadIdx.x; e < 64 * cnt; e += 256) {
        const int i = e % cnt, jj = e / cnt;
        int j = jb + jj;
        if (j < nout) {
            if (j >= k0) j += nb;
            vals[jj][i] = A[pvv[i] + (size_t)lda * j];
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < 64 * cnt; e += 256) {
        const int i = e % cnt, jj = e / cnt;
        int j = jb + jj;
        if (j < nout) {
            if (j >= k0) j += nb;
            A[pk[i] + (size_t)lda * j] = vals[jj][i];
        }
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

// Inverse of the nb x nb diagonal block at (k0, k0) of the LU factors, padded
// to NB x NB with the identity: UPPER = 0 the unit-lower L11, 1 the upper U11.
// Thread c forms column c by substitution on e_c.  With n > 0, block b
// handles the diagonal block at k0 = b NB of an n x n factor (all blocks in
// one launch) and writes Tinv + b NB^2.
template <int UPPER>
__global__ void __launch_bounds__(NB) k_tri_inv(const double* __restrict__ LU, int ld, int k0, int nb,
                                                double* __restrict__ Tinv, int n = 0)
{
    __shared__ double T[NB][NB + 1];
    if (n > 0) {
        k0 = blockIdx.x * NB;
        nb = min(NB, n - k0);
        Tinv += (size_t)blockIdx.x * NB * NB;
    }
    for (int e = threadIdx.x; e < NB * NB; e += NB) {
        const int r = e % NB, c = e / NB;
        T[r][c] = (r < nb && c < nb) ? LU[(k0 + r) + (size_t)ld * (k0 + c)] : (r == c ? 1.0 : 0.0);
    }
    __syncthreads();
    const int c = threadIdx.x;
    double x[NB];
#pragma unroll
    for (int r = 0; r < NB; ++r) x[r] = r == c ? 1.0 : 0.0;
    if (!UPPER) {
#pragma unroll
        for (int r = 0; r < NB; ++r)
#pragma unroll
            for (int k = 0; k < r; ++k) x[r] -= T[r][k] * x[k];
    } else {
#pragma unroll
        for (int r = NB - 1; r >= 0; --r) {
            double v = x[r];
#pragma unroll
            for (int k = r + 1; k < NB; ++k) v -= T[r][k] * x[k];
            x[r] = v / T[r][r];
        }
    }
#pragma unroll
    for (int r = 0; r < NB; ++r) Tinv[r + NB * c] = x[r];
}

// X[k0:k0+nb, col0+j] = Tinv X[k0:k0+nb, col0+j] for j < ncols: 64 columns per
// work-group, the block and the tile staged in LDS.
__global__ void __launch_bounds__(256) k_tri_apply(const double* __restrict__ Tinv, int k0, int nb,
                                                   double* __restrict__ X, int ldx, int ncols, int col0)
{
    __shared__ double Ts[NB][NB + 1];
    __shared__ double Xs[64][NB + 1];
    const int j0 = blockIdx.x * 64;
    for (int e = threadIdx.x; e < NB * NB; e += 256) Ts[e % NB][e / NB] = Tinv[e];
    for (int e = threadIdx.x; e < 64 * NB; e += 256) {
        const int r = e % NB, j = e / NB;
        Xs[j][r] = (r < nb && j0 + j < ncols) ? X[(k0 + r) + (size_t)ldx * (col0 + j0 + j)] : 0.0;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < 64 * NB; e += 256) {
        const int r = e % NB, j = e / NB;
        if (r >= nb || j0 + j >= ncols) continue;
        double acc = 0.0;
#pragma unroll 8
        for (int k = 0; k < NB; ++k) acc += Ts[r][k] * Xs[j][k];
        X[(k0 + r) + (size_t)ldx * (col0 + j0 + j)] = acc;
    }
}

// panel + its interchange map; false: the global fallback ran and the caller
// applies the swaps one by one
static bool panel_factor(hipStream_t st, double* A, int n, int k0, int nb, int* piv, int* info, int* pmap)
{
    const int m = n - k0;
    const void* fn = nullptr;
    int W = 0;
    if (m <= 2048) { fn = (const void*)k_panel_lds<8>; W = 8; }
    else if (m <= 4096) { fn = (const void*)k_panel_lds<4>; W = 4; }
    else if (m <= 8192) { fn = (const void*)k_panel_lds<2>; W = 2; }
    if (!W) {
        k_panel<<<1, 1024, 0, st>>>(A, n, n, k0, nb, piv, info);
        return false;
    }
    const size_t lds = sizeof(double) * (size_t)W * m;   // W * m <= 16384 doubles
    static bool set[3] = {false, false, false};
    const int slot = W == 8 ? 0 : W == 4 ? 1 : 2;
    if (!set[slot]) {
        ILH_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)(16384 * sizeof(double))));
        set[slot] = true;
    }
    if (W == 8) k_panel_lds<8><<<1, PT, lds, st>>>(A, n, n, k0, nb, piv, info, pmap);
    else if (W == 4) k_panel_lds<4><<<1, PT, lds, st>>>(A, n, n, k0, nb, piv, info, pmap);
    else k_panel_lds<2><<<1, PT, lds, st>>>(A, n, n, k0, nb, piv, info, pmap);
    return true;
}

int lu_factor(ilh_plan* pl, double* A, int n, int* piv, int* info)
{
    hipStream_t st = pl->ctx->stream;
    ILH_HIP(hipMemsetAsync(info, 0, sizeof(int), st));
    pl->lu_tinv.ensure(sizeof(double) * NB * NB);
    pl->lu_pmap.ensure(sizeof(int) * (1 + 4 * NB));
    double* Tinv = pl->lu_tinv.as<double>();
    int* pmap = pl->lu_pmap.as<int>();
    for (int k0 = 0; k0 < n; k0 += NB) {
        const int nb = std::min(NB, n - k0);
        const bool fused = panel_factor(st, A, n, k0, nb, piv, info, pmap);
        if (n - nb > 0) {
            if (fused) k_apply_pmap<<<(n - nb + 63) / 64, 256, 0, st>>>(A, n, n, k0, nb, pmap);
            else k_swap_rows<<<(n - nb + 255) / 256, 256, 0, st>>>(A, n, n, k0, nb, piv);
        }
        const int k1 = k0 + nb, rest = n - k1;
        if (rest > 0) {
            k_tri_inv<0><<<1, NB, 0, st>>>(A, n, k0, nb, Tinv);
            k_tri_apply<<<(rest + 63) / 64, 256, 0, st>>>(Tinv, k0, nb, A, n, rest, k1);
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

// perm[i] = the row of B that lands in row i after the interchanges piv[0..n)
// are applied in order (LAPACK getrs order); one work-group, LDS when n fits.
__global__ void __launch_bounds__(256) k_perm_from_piv(const int* __restrict__ piv, int n, int* __restrict__ perm)
{
    extern __shared__ int pl_[];
    const bool lds = n <= 16384;
    int* p = lds ? pl_ : perm;
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = i;
    __syncthreads();
    if (threadIdx.x == 0)
        for (int k = 0; k < n; ++k) {
            const int q = piv[k];
            if (q != k) {
                const int t = p[k];
                p[k] = p[q];
                p[q] = t;
            }
        }
    __syncthreads();
    if (lds)
        for (int i = threadIdx.x; i < n; i += blockDim.x) perm[i] = p[i];
}

// Bp[i, j] = B[perm[i], j]
__global__ void k_gather_rows(const double* __restrict__ B, int n, int ncols, const int* __restrict__ perm,
                              double* __restrict__ Bp)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= n || j >= ncols) return;
    Bp[i + (size_t)n * j] = B[perm[i] + (size_t)n * j];
}

// B <- A^{-1} B for the LU factors of A (n x n) and nrhs columns of B:
// permute, then blocked forward and backward substitution whose diagonal
// blocks are applied as explicit NB x NB inverses.
void lu_solve_multi(ilh_plan* pl, const double* LU, const int* piv, int n, double* B, int nrhs)
{
    hipStream_t st = pl->ctx->stream;
    pl->lu_perm.ensure(sizeof(int) * (size_t)n);
    pl->lu_tmp.ensure(sizeof(double) * (size_t)n * nrhs);
    const int nblk = (n + NB - 1) / NB;
    pl->lu_tinv.ensure(sizeof(double) * NB * NB * 2 * (size_t)nblk);
    int* perm = pl->lu_perm.as<int>();
    double* Bp = pl->lu_tmp.as<double>();
    double* Linv = pl->lu_tinv.as<double>();
    double* Uinv = Linv + (size_t)NB * NB * nblk;
    k_tri_inv<0><<<nblk, NB, 0, st>>>(LU, n, 0, 0, Linv, n);
    k_tri_inv<1><<<nblk, NB, 0, st>>>(LU, n, 0, 0, Uinv, n);
    k_perm_from_piv<<<1, 256, n <= 16384 ? sizeof(int) * (size_t)n : 0, st>>>(piv, n, perm);
    k_gather_rows<<<dim3((n + 255) / 256, nrhs), 256, 0, st>>>(B, n, nrhs, perm, Bp);
    ILH_HIP(hipMemcpyAsync(B, Bp, sizeof(double) * (size_t)n * nrhs, hipMemcpyDeviceToDevice, st));
    for (int k0 = 0; k0 < n; k0 += NB) {
        const int nb = std::min(NB, n - k0);
        k_tri_apply<<<(nrhs + 63) / 64, 256, 0, st>>>(Linv + (size_t)NB * NB * (k0 / NB), k0, nb, B, n, nrhs, 0);
        const int k1 = k0 + nb;
        gemm_sub(st, n - k1, nrhs, nb, LU + k1 + (size_t)n * k0, n, B + k0, n, B + k1, n);
    }
    const int last = ((n - 1) / NB) * NB;
    for (int k0 = last; k0 >= 0; k0 -= NB) {
        const int nb = std::min(NB, n - k0);
        k_tri_apply<<<(nrhs + 63) / 64, 256, 0, st>>>(Uinv + (size_t)NB * NB * (k0 / NB), k0, nb, B, n, nrhs, 0);
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
__global__ void __launch_bounds__(256) k_matvec_rows(const double* __restrict__ M, int ld, const double* __restrict__ x, double* __restrict__ y,
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
__global__ void __launch_bounds__(256) k_dir_amp(const double* __restrict__ Mp, const double* __restrict__ Mm, int ld,
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

// The surface step for nb right-hand sides (columns fp + k ldf, ldf even and
// fp 16-byte aligned; fm likewise or null): the matrix rows are read once
// for up to 8 right-hand sides, and every column's sum has wave_dot's order
// (each result equals the single-right-hand-side kernel's bit for bit).
// amp_k = amp + 2 n k, s_k = s_out + n k.
constexpr int AMP_GROUP = 8;

// the dots of one matrix row with NG columns x + ldf g (NG a compile-time
// count, so that the partial sums stay in registers)
template <int NG>
__device__ __forceinline__ void wave_dot_group(const double* __restrict__ row, const double* __restrict__ x, int ldf,
                                               int m, int lane, double* out)
{
    const double2* r2 = (const double2*)row;
    const int m2 = m >> 1;
    double s[NG][4];
#pragma unroll
    for (int g = 0; g < NG; ++g) s[g][0] = s[g][1] = s[g][2] = s[g][3] = 0.0;
    int k = lane;
    for (; k + 192 < m2; k += 256) {
        const double2 a0 = r2[k], a1 = r2[k + 64], a2 = r2[k + 128], a3 = r2[k + 192];
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            const double2* x2 = (const double2*)(x + (size_t)ldf * g);
            const double2 b0 = x2[k], b1 = x2[k + 64], b2 = x2[k + 128], b3 = x2[k + 192];
            s[g][0] += dot2(a0, b0);
            s[g][1] += dot2(a1, b1);
            s[g][2] += dot2(a2, b2);
            s[g][3] += dot2(a3, b3);
        }
    }
    for (; k < m2; k += 64) {
        const double2 a0 = r2[k];
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            const double2 b0 = ((const double2*)(x + (size_t)ldf * g))[k];
            s[g][0] += dot2(a0, b0);
        }
    }
#pragma unroll
    for (int g = 0; g < NG; ++g) {
        if ((m & 1) && lane == 0) s[g][1] = __builtin_fma(row[m - 1], x[(size_t)ldf * g + m - 1], s[g][1]);
        double t = (s[g][0] + s[g][1]) + (s[g][2] + s[g][3]);
        for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
        out[g] = t;
    }
}

template <int NG>
__device__ __forceinline__ void dir_amp_group(const double* __restrict__ Mp, const double* __restrict__ Mm, int ld,
                                              const double* __restrict__ fp, const double* __restrict__ fm, int ldf,
                                              int g0, double* __restrict__ amp, double* __restrict__ s_out, int n,
                                              int i, int lane)
{
    double sp[NG], sm[NG];
    wave_dot_group<NG>(Mp + (size_t)ld * i, fp + (size_t)ldf * g0, ldf, n, lane, sp);
    if (fm) wave_dot_group<NG>(Mm + (size_t)ld * i, fm + (size_t)ldf * g0, ldf, n, lane, sm);
    if (lane == 0)
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            const int k = g0 + g;
            const double sv = fm ? sp[g] + sm[g] : sp[g];
            amp[(size_t)2 * n * k + i] = sv;
            amp[(size_t)2 * n * k + n + i] = fp[(size_t)ldf * k + i] - (fm ? fm[(size_t)ldf * k + i] : 0.0);
            if (s_out) s_out[(size_t)n * k + i] = sv;
        }
}

__global__ void __launch_bounds__(256) k_dir_amp_multi(const double* __restrict__ Mp, const double* __restrict__ Mm, int ld,
                                const double* __restrict__ fp, const double* __restrict__ fm, int ldf, int nb,
                                double* __restrict__ amp, double* __restrict__ s_out, int n)
{
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (i >= n) return;
    int g0 = 0;
    for (; g0 + AMP_GROUP <= nb; g0 += AMP_GROUP)
        dir_amp_group<AMP_GROUP>(Mp, Mm, ld, fp, fm, ldf, g0, amp, s_out, n, i, lane);
    switch (nb - g0) {
    case 1: dir_amp_group<1>(Mp, Mm, ld, fp, fm, ldf, g0, amp, s_out, n, i, lane); break;
    case 2: dir_amp_group<2>(Mp, Mm, ld, fp, fm, ldf, g0, amp, s_out, n, i, lane); break;
    case 3: dir_amp_group<3>(Mp, Mm, ld, fp, fm, ldf, g0, amp, s_out, n, i, lane); break;
    case 4: dir_amp_group<4>(Mp, Mm, ld, fp, fm, ldf, g0, amp, s_out, n, i, lane); break;
    case 5: dir_amp_group<5>(Mp, Mm, ld, fp, fm, ldf, g0, amp, s_out, n, i, lane); break;
    case 6: dir_amp_group<6>(Mp, Mm, ld, fp, fm, ldf, g0, amp, s_out, n, i, lane); break;
    case 7: dir_amp_group<7>(Mp, Mm, ld, fp, fm, ldf, g0, amp, s_out, n, i, lane); break;
    default: break;
    }
}

void dirichlet_amp_multi(hipStream_t st, const double* Mp, const double* Mm, int ld, const double* fp,
                         const double* fm, int ldf, int nb, double* amp, double* s_out, int n)
{
    if (n <= 0 || nb <= 0) return;
    k_dir_amp_multi<<<(n + 3) / 4, 256, 0, st>>>(Mp, Mm, ld, fp, fm, ldf, nb, amp, s_out, n);
    ILH_HIP(hipGetLastError());
}

// Neumann surface step: df = A+ v+ + A- v-, ds = B+ v+ + B- v-;
// primal amplitudes [df; v+ - v-], dual amplitudes [df; -ds]
__global__ void __launch_bounds__(256) k_neu_amp(const double* __restrict__ Ap, const double* __restrict__ Am, const double* __restrict__ Bp,
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
