// ilh_terms.hip -- DDF stencil terms of the surface operators, their row
// binning, and the deterministic scatter (surface -> grid) / gather
// (grid -> surface) kernels.
//
// Every surface operator of the scalar cache (src/surface_operators.jl) is a
// sum over points of at most two separable 5x5 windows ("terms") on one of the
// four component grids; the symmetric tensor operators of the vector cache
// take eight (two amplitudes on each of four component grids).  The DDF weights and the MAC stencil of the composed
// operator are folded into the window weights once per geometry, so e.g.
// surface_divergence! = D Rf (n o f) / dx becomes a single scatter pass.
//
// Scatter is a gather over output cells: terms are binned by output row and
// sorted by window base, each output cell sums the terms covering it in that
// fixed order.  No float atomics; results are bitwise reproducible.
#include "ilh_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace ilh {

// ----------------------------------------------------------------- DDFs
// Yang3 (Yang, Zhang, Li & He 2009, the default Regularize kernel,
// src/ImmersedLayers.jl:83), Roma, M4', linear hat.  r in grid units.
__device__ __forceinline__ double ddf_eval(int kind, double r)
{
    const double a = fabs(r);
    if (kind == ILH_DDF_YANG3) {
        const double s3 = 1.7320508075688772935274463;
        const double pi = 3.14159265358979323846264338;
        if (a <= 1.0)
            return 17.0 / 48.0 + s3 * pi / 108.0 + a / 4.0 - a * a / 4.0 +
                   (1.0 - 2.0 * a) / 16.0 * sqrt(-12.0 * a * a + 12.0 * a + 1.0) -
                   s3 / 12.0 * asin(s3 / 2.0 * (2.0 * a - 1.0));
        if (a <= 2.0)
            return 55.0 / 48.0 - s3 * pi / 108.0 - 13.0 * a / 12.0 + a * a / 4.0 +
                   (2.0 * a - 3.0) / 48.0 * sqrt(-12.0 * a * a + 36.0 * a - 23.0) +
                   s3 / 36.0 * asin(s3 / 2.0 * (2.0 * a - 3.0));
        return 0.0;
    }
    if (kind == ILH_DDF_ROMA) {
        if (a <= 0.5) return (1.0 + sqrt(1.0 - 3.0 * a * a)) / 3.0;
        if (a <= 1.5) return (5.0 - 3.0 * a - sqrt(1.0 - 3.0 * (1.0 - a) * (1.0 - a))) / 6.0;
        return 0.0;
    }
    if (kind == ILH_DDF_M4PRIME) {
        if (a <= 1.0) return 1.0 - 2.5 * a * a + 1.5 * a * a * a;
        if (a <= 2.0) return 0.5 * (2.0 - a) * (2.0 - a) * (1.0 - a);
        return 0.0;
    }
    return a <= 1.0 ? 1.0 - a : 0.0;   // witch hat
}

// 4-tap window of a DDF centred at fractional index xi: base floor(xi)-1,
// weights w[e] = phi(base + e - xi) (w[4] = 0; the difference windows use it).
struct Win {
    int b;
    double w[5];
};

// first difference w'[e] = w[e] - w[e-1] (5 taps); the caller picks the base:
// same base for curl (out_i = a_i - a_{i-1}), base-1 for divergence
// (out_i = a_{i+1} - a_i).
__device__ __forceinline__ Win diff_win(const Win& a, int base_shift)
{
    Win r;
    r.b = a.b + base_shift;
    r.w[0] = a.w[0];
    for (int e = 1; e < 4; ++e) r.w[e] = a.w[e] - a.w[e - 1];
    r.w[4] = -a.w[3];
    return r;
}

__device__ __forceinline__ void put(Term& t, int ai, int gk, const Win& x, const Win& y, double coef)
{
    t.ai = ai;
    t.gk = gk;
    t.bx = x.b;
    t.by = y.b;
    t.coef = coef;
    for (int e = 0; e < 5; ++e) {
        t.wx[e] = x.w[e];
        t.wy[e] = y.w[e];
    }
}

// The four DDF windows of every point (primal x, primal y, dual x, dual y),
// one weight per thread: 16 threads per point, so the FP64 DDF evaluations
// (square roots and arcsines) run in parallel rather than in a chain; every
// operator's terms are then built from these.
__device__ __forceinline__ void k_ddf_windows_body(int N, int ddf, Grid g, const double* __restrict__ X, const double* __restrict__ Y,
                              Win* __restrict__ wins, int bx)
{
    const int k = bx * blockDim.x + threadIdx.x;
    if (k >= 16 * N) return;
    const int p = k >> 4, w = (k >> 2) & 3, e = k & 3;
    // fractional indices of the point on the primal (x_i = (i+1-I0) dx) and
    // dual (x_i = (i+1/2-I0) dx) axes, 0-based
    const double dx = g.dx;
    const double xp = X[p] / dx - (1.0 - g.I0x), yp = Y[p] / dx - (1.0 - g.I0y);
    const double xi = (w & 1 ? yp : xp) + (w & 2 ? 0.5 : 0.0);
    const int b = (int)floor(xi) - 1;
    Win& out = wins[4 * (size_t)p + w];
    out.w[e] = ddf_eval(ddf, (double)(b + e) - xi);
    if (e == 0) {
        out.b = b;
        out.w[4] = 0.0;
    }
}

__global__ void k_ddf_windows(int N, int ddf, Grid g, const double* __restrict__ X, const double* __restrict__ Y,
                              Win* __restrict__ wins)
{
    k_ddf_windows_body(N, ddf, g, X, Y, wins, blockIdx.x);
}

__device__ __forceinline__ void k_build_terms_body(int op, int N, Grid g, const Win* __restrict__ wins, const double* __restrict__ NXv,
                              const double* __restrict__ NYv, const double* __restrict__ DS, Term* __restrict__ out, int bx)
{
    const int p = bx * blockDim.x + threadIdx.x;
    if (p >= N) return;
    const double dx = g.dx;
    const Win Px = wins[4 * (size_t)p], Py = wins[4 * (size_t)p + 1];
    const Win Dx = wins[4 * (size_t)p + 2], Dy = wins[4 * (size_t)p + 3];
    const double n_x = NXv[p], n_y = NYv[p], ds = DS[p];
    const double r2 = ds / (dx * dx), r3 = ds / (dx * dx * dx), i1 = 1.0 / dx;
    Term a, b;
    switch (op) {
    case OP_R_PRIMAL: put(a, p, GK_PRIMAL, Px, Py, r2); out[p] = a; return;
    case OP_R_DUAL: put(a, p, GK_DUAL, Dx, Dy, r2); out[p] = a; return;
    case OP_E_PRIMAL: put(a, p, GK_PRIMAL, Px, Py, 1.0); out[p] = a; return;
    case OP_E_DUAL: put(a, p, GK_DUAL, Dx, Dy, 1.0); out[p] = a; return;
    case OP_R_EDGES:
        put(a, p, GK_EDGE_U, Dx, Py, r2);
        put(b, N + p, GK_EDGE_V, Px, Dy, r2);
        break;
    case OP_E_EDGES:
        put(a, p, GK_EDGE_U, Dx, Py, 1.0);
        put(b, N + p, GK_EDGE_V, Px, Dy, 1.0);
        break;
    case OP_RN:
        put(a, p, GK_EDGE_U, Dx, Py, n_x * r2);
        put(b, p, GK_EDGE_V, Px, Dy, n_y * r2);
        break;
    case OP_NI:
        put(a, p, GK_EDGE_U, Dx, Py, n_x);
        put(b, p, GK_EDGE_V, Px, Dy, n_y);
        break;
    case OP_RNX:   // n x (f e_z) = (n_y f, -n_x f)
        put(a, p, GK_EDGE_U, Dx, Py, n_y * r2);
        put(b, p, GK_EDGE_V, Px, Dy, -n_x * r2);
        break;
    case OP_NXI:   // e_z . (n x v) = n_x v_y - n_y v_x
        put(a, p, GK_EDGE_U, Dx, Py, -n_y);
        put(b, p, GK_EDGE_V, Px, Dy, n_x);
        break;
    case OP_DS:    // D Rf (n o f) / dx
        put(a, p, GK_PRIMAL, diff_win(Dx, -1), Py, n_x * r3);
        put(b, p, GK_PRIMAL, Px, diff_win(Dy, -1), n_y * r3);
        break;
    case OP_GS:    // n . Rf^T G phi / dx   (G = -D^T)
        put(a, p, GK_PRIMAL, diff_win(Dx, -1), Py, -n_x * i1);
        put(b, p, GK_PRIMAL, Px, diff_win(Dy, -1), -n_y * i1);
        break;
    case OP_GXS:   // e_z . (n x Rf^T G phi) / dx
        put(a, p, GK_PRIMAL, diff_win(Dx, -1), Py, n_y * i1);
        put(b, p, GK_PRIMAL, Px, diff_win(Dy, -1), -n_x * i1);
        break;
    case OP_CST:   // C^T Rf (n o f) / dx,  w = (v_i - v_{i-1}) - (u_j - u_{j-1})
        put(a, p, GK_DUAL, diff_win(Px, 0), Dy, n_y * r3);
        put(b, p, GK_DUAL, Dx, diff_win(Py, 0), -n_x * r3);
        break;
    case OP_CS:    // n . Rf^T C s / dx,  u = s_{j+1} - s_j, v = s_i - s_{i+1}
        put(a, p, GK_DUAL, diff_win(Px, 0), Dy, n_y * i1);
        put(b, p, GK_DUAL, Dx, diff_win(Py, 0), -n_x * i1);
        break;
    case OP_CXT:   // C^T Rf (n x f e_z) / dx
        put(a, p, GK_DUAL, diff_win(Px, 0), Dy, -n_x * r3);
        put(b, p, GK_DUAL, Dx, diff_win(Py, 0), -n_y * r3);
        break;
    case OP_CXS:   // e_z . (n x Rf^T C s) / dx
        put(a, p, GK_DUAL, diff_win(Px, 0), Dy, n_x * i1);
        put(b, p, GK_DUAL, Dx, diff_win(Py, 0), n_y * i1);
        break;
    // vector cache: VectorData f = (f_u at p, f_v at N + p)
    case OP_CVT:   // C^T Rf f / dx,  w = (v_i - v_{i-1}) - (u_j - u_{j-1})
        put(a, N + p, GK_DUAL, diff_win(Px, 0), Dy, r3);
        put(b, p, GK_DUAL, Dx, diff_win(Py, 0), -r3);
        break;
    case OP_CVS:   // Rf^T C s / dx,  u = s_{j+1} - s_j, v = s_i - s_{i+1}
        put(a, p, GK_DUAL, Dx, diff_win(Py, 0), -i1);
        put(b, N + p, GK_DUAL, diff_win(Px, 0), Dy, i1);
        break;
    case OP_DSYM: case OP_GSYM: case OP_RNSYM: case OP_NISYM: {
        // The symmetric surface tensor T = n o f + f o n - (n . f) I has
        // T_xx = -T_yy = n_x f_u - n_y f_v and T_xy = T_yx = n_y f_u + n_x f_v.
        // Differences: out_i = a_i - a_{i-1} maps nodes to the staggered
        // positions between them (diff_win(w, 0)), out_i = a_{i+1} - a_i the
        // staggered positions back to nodes (diff_win(w, -1)).
        Term* o = out + 8 * (size_t)p;
        const Win dPx = diff_win(Px, 0), dPy = diff_win(Py, 0), dDx = diff_win(Dx, -1), dDy = diff_win(Dy, -1);
        if (op == OP_DSYM) {
            // u edges: d/dx T_xx (primal -> x faces) + d/dy T_xy (dual -> x faces)
            put(o[0], p, GK_EDGE_U, dPx, Py, n_x * r3);
            put(o[1], N + p, GK_EDGE_U, dPx, Py, -n_y * r3);
            put(o[2], p, GK_EDGE_U, Dx, dDy, n_y * r3);
            put(o[3], N + p, GK_EDGE_U, Dx, dDy, n_x * r3);
            // v edges: d/dx T_yx (dual -> y faces) + d/dy T_yy (primal -> y faces)
            put(o[4], p, GK_EDGE_V, dDx, Dy, n_y * r3);
            put(o[5], N + p, GK_EDGE_V, dDx, Dy, n_x * r3);
            put(o[6], p, GK_EDGE_V, Px, dPy, -n_x * r3);
            put(o[7], N + p, GK_EDGE_V, Px, dPy, n_y * r3);
        } else if (op == OP_GSYM) {
            // A = G v: A_xx = (u_{i+1} - u_i), A_yy = (v_{j+1} - v_j) on primal
            // nodes, A_xy = (u_j - u_{j-1}), A_yx = (v_i - v_{i-1}) on dual
            // nodes; tau = n . (A + A^T - tr(A) I), the negative adjoint of DSYM
            put(o[0], p, GK_EDGE_U, dPx, Py, -n_x * i1);
            put(o[1], N + p, GK_EDGE_U, dPx, Py, n_y * i1);
            put(o[2], p, GK_EDGE_U, Dx, dDy, -n_y * i1);
            put(o[3], N + p, GK_EDGE_U, Dx, dDy, -n_x * i1);
            put(o[4], p, GK_EDGE_V, dDx, Dy, -n_y * i1);
            put(o[5], N + p, GK_EDGE_V, dDx, Dy, -n_x * i1);
            put(o[6], p, GK_EDGE_V, Px, dPy, n_x * i1);
            put(o[7], N + p, GK_EDGE_V, Px, dPy, -n_y * i1);
        } else if (op == OP_RNSYM) {   // R_t T
            put(o[0], p, GK_T_XX, Px, Py, n_x * r2);
            put(o[1], N + p, GK_T_XX, Px, Py, -n_y * r2);
            put(o[2], p, GK_T_XY, Dx, Dy, n_y * r2);
            put(o[3], N + p, GK_T_XY, Dx, Dy, n_x * r2);
            put(o[4], p, GK_T_YX, Dx, Dy, n_y * r2);
            put(o[5], N + p, GK_T_YX, Dx, Dy, n_x * r2);
            put(o[6], p, GK_T_YY, Px, Py, -n_x * r2);
            put(o[7], N + p, GK_T_YY, Px, Py, n_y * r2);
        } else {   // tau = n . R_t^T (A + A^T - tr(A) I)
            put(o[0], p, GK_T_XX, Px, Py, n_x);
            put(o[1], N + p, GK_T_XX, Px, Py, -n_y);
            put(o[2], p, GK_T_XY, Dx, Dy, n_y);
            put(o[3], N + p, GK_T_XY, Dx, Dy, n_x);
            put(o[4], p, GK_T_YX, Dx, Dy, n_y);
            put(o[5], N + p, GK_T_YX, Dx, Dy, n_x);
            put(o[6], p, GK_T_YY, Px, Py, -n_x);
            put(o[7], N + p, GK_T_YY, Px, Py, n_y);
        }
        return;
    }
    default: return;
    }
    out[2 * p] = a;
    out[2 * p + 1] = b;
}

__global__ void k_build_terms(int op, int N, Grid g, const Win* __restrict__ wins, const double* __restrict__ NXv,
                              const double* __restrict__ NYv, const double* __restrict__ DS, Term* __restrict__ out)
{
    k_build_terms_body(op, N, g, wins, NXv, NYv, DS, out, blockIdx.x);
}

int op_terms_per_point(int op)
{
    switch (op) {
    case OP_R_PRIMAL: case OP_R_DUAL: case OP_E_PRIMAL: case OP_E_DUAL: return 1;
    case OP_DSYM: case OP_GSYM: case OP_RNSYM: case OP_NISYM: return 8;
    default: return 2;
    }
}

bool op_is_scatter(int op)
{
    switch (op) {
    case OP_R_PRIMAL: case OP_R_DUAL: case OP_R_EDGES: case OP_RN: case OP_RNX:
    case OP_DS: case OP_CST: case OP_CXT: case OP_CVT: case OP_DSYM: case OP_RNSYM: return true;
    default: return false;
    }
}

void build_terms(ilh_plan* pl, int op)
{
    const int T = op_terms_per_point(op);
    pl->terms[op].ensure(sizeof(Term) * (size_t)std::max(1, pl->N * T));
    if (pl->N > 0) {
        if (!pl->wins_ok) {
            pl->wins.ensure(sizeof(Win) * 4 * (size_t)pl->N);
            k_ddf_windows<<<(16 * pl->N + 255) / 256, 256, 0, pl->ctx->stream>>>(
                pl->N, pl->ddf, pl->g, pl->X.as<double>(), pl->Y.as<double>(), pl->wins.as<Win>());
            ILH_HIP(hipGetLastError());
            pl->wins_ok = true;
        }
        k_build_terms<<<(pl->N + 127) / 128, 128, 0, pl->ctx->stream>>>(
            op, pl->N, pl->g, pl->wins.as<Win>(), pl->nx.as<double>(), pl->ny.as<double>(), pl->ds.as<double>(),
            pl->terms[op].as<Term>());
        ILH_HIP(hipGetLastError());
    }
    pl->terms_ok[op] = true;
}

const Term* get_terms(ilh_plan* pl, int op)
{
    if (!pl->terms_ok[op]) build_terms(pl, op);
    return pl->terms[op].as<Term>();
}

// ------------------------------------------------------------ row binning
// The binning kernels walk the terms with every thread's next BIN_BATCH
// terms (k0, k0 + blockDim.x, ...) loaded together: the window's first row
// and a mask of its rows with a nonzero weight (0 for terms of another grid
// kind or past the end), so a thread waits for one round of loads per batch
// rather than one per term.
constexpr int BIN_BATCH = 8;

__device__ __forceinline__ void bin_rows(const Term* __restrict__ t, int nt, int gk, int k0, int (&row0)[BIN_BATCH],
                                         unsigned (&msk)[BIN_BATCH])
{
#pragma unroll
    for (int u = 0; u < BIN_BATCH; ++u) {
        const int k = k0 + u * (int)blockDim.x;
        unsigned m = 0;
        int r = 0;
        if (k < nt && t[k].gk == gk) {
            r = t[k].by;
#pragma unroll
            for (int e = 0; e < 5; ++e) m |= (t[k].wy[e] != 0.0 ? 1u : 0u) << e;
        }
        row0[u] = r;
        msk[u] = m;
    }
}

// entries per row (terms of grid kind gk with a nonzero weight in the row):
// one work-group, counts in LDS (dynamic, ny ints), written to cnt[0..ny)
__device__ __forceinline__ void k_bin_count_body(const Term* __restrict__ t, int nt, int gk, int ny,
                                                    int* __restrict__ cnt, int bx)
{
    extern __shared__ int lc[];
    for (int r = threadIdx.x; r < ny; r += blockDim.x) lc[r] = 0;
    __syncthreads();
    for (int k0 = threadIdx.x; k0 < nt; k0 += blockDim.x * BIN_BATCH) {
        int row0[BIN_BATCH];
        unsigned msk[BIN_BATCH];
        bin_rows(t, nt, gk, k0, row0, msk);
#pragma unroll
        for (int u = 0; u < BIN_BATCH; ++u)
            for (int e = 0; e < 5; ++e) {
                const int row = row0[u] + e;
                if ((msk[u] >> e & 1u) && row >= 0 && row < ny) atomicAdd(&lc[row], 1);
            }
    }
    __syncthreads();
    for (int r = threadIdx.x; r < ny; r += blockDim.x) cnt[r] = lc[r];
}


// exclusive scan of cnt[0..n) into ptr[0..n], single work-group of 1024
// threads, 4096 values per round (4 per thread, wave shuffles, one LDS step).
// n is *n_dev when n_dev is not null (a count produced on the device, at most
// the n given).
__device__ __forceinline__ void k_scan_body(const int* __restrict__ cnt, int n, int* __restrict__ ptr,
                                               const int* __restrict__ n_dev, int bx)
{
    __shared__ int wsum[16];
    __shared__ int carry;
    if (n_dev) n = min(n, *n_dev);
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    if (t == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < n; base += 4096) {
        const int i0 = base + 4 * t;
        int v[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) v[e] = i0 + e < n ? cnt[i0 + e] : 0;
        const int own = v[0] + v[1] + v[2] + v[3];
        int inc = own;
        for (int off = 1; off < 64; off <<= 1) {
            const int u = __shfl_up(inc, off, 64);
            if (lane >= off) inc += u;
        }
        if (lane == 63) wsum[wave] = inc;
        __syncthreads();
        int before = carry;
        for (int w = 0; w < wave; ++w) before += wsum[w];
        int run = before + inc - own;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            if (i0 + e < n) ptr[i0 + e] = run;
            run += v[e];
        }
        __syncthreads();
        if (t == 1023) carry = run;
        __syncthreads();
    }
    if (t == 0) ptr[n] = carry;
}

__global__ void __launch_bounds__(1024) k_scan(const int* __restrict__ cnt, int n, int* __restrict__ ptr,
                                               const int* __restrict__ n_dev)
{
    k_scan_body(cnt, n, ptr, n_dev, blockIdx.x);
}

// the entries themselves, in arbitrary order within a row (k_bin_sort orders
// them); one work-group, row cursors in LDS (dynamic, ny ints)
__device__ __forceinline__ void k_bin_fill_body(const Term* __restrict__ t, int nt, int gk, int ny,
                                                   const int* __restrict__ ptr, int* __restrict__ ent, int bx)
{
    extern __shared__ int lc[];
    for (int r = threadIdx.x; r < ny; r += blockDim.x) lc[r] = 0;
    __syncthreads();
    for (int k0 = threadIdx.x; k0 < nt; k0 += blockDim.x * BIN_BATCH) {
        int row0[BIN_BATCH];
        unsigned msk[BIN_BATCH];
        bin_rows(t, nt, gk, k0, row0, msk);
#pragma unroll
        for (int u = 0; u < BIN_BATCH; ++u)
            for (int e = 0; e < 5; ++e) {
                const int row = row0[u] + e;
                if ((msk[u] >> e & 1u) && row >= 0 && row < ny) {
                    const int pos = atomicAdd(&lc[row], 1);
                    ent[ptr[row] + pos] = k0 + u * (int)blockDim.x;
                }
            }
    }
}


// The three steps of a row binning in one work-group of 1024: the entries per
// row counted in LDS, their exclusive scan (k_scan_body, LDS to LDS) copied to
// ptr[0..ny], and the entries placed with LDS cursors (in arbitrary order
// within a row; k_bin_sort orders them).  Dynamic LDS: 2 ny + 1 ints.  Two
// launches fewer than counting, scanning and filling apart (as the batched
// k_geo_bin_count, k_geo_scan and k_geo_bin_fill still do), with the same
// ptr and the same entries per row.
__global__ void __launch_bounds__(1024) k_bin_build(const Term* __restrict__ t, int nt, int gk, int ny,
                                                    int* __restrict__ ptr, int* __restrict__ ent)
{
    extern __shared__ int lc[];   // ny counts (then cursors), then ny + 1 offsets
    int* lp = lc + ny;
    for (int r = threadIdx.x; r < ny; r += blockDim.x) lc[r] = 0;
    __syncthreads();
    for (int k0 = threadIdx.x; k0 < nt; k0 += blockDim.x * BIN_BATCH) {
        int row0[BIN_BATCH];
        unsigned msk[BIN_BATCH];
        bin_rows(t, nt, gk, k0, row0, msk);
#pragma unroll
        for (int u = 0; u < BIN_BATCH; ++u)
            for (int e = 0; e < 5; ++e) {
                const int row = row0[u] + e;
                if ((msk[u] >> e & 1u) && row >= 0 && row < ny) atomicAdd(&lc[row], 1);
            }
    }
    __syncthreads();
    k_scan_body(lc, ny, lp, nullptr, 0);
    __syncthreads();
    for (int r = threadIdx.x; r <= ny; r += blockDim.x) ptr[r] = lp[r];
    for (int r = threadIdx.x; r < ny; r += blockDim.x) lc[r] = 0;
    __syncthreads();
    for (int k0 = threadIdx.x; k0 < nt; k0 += blockDim.x * BIN_BATCH) {
        int row0[BIN_BATCH];
        unsigned msk[BIN_BATCH];
        bin_rows(t, nt, gk, k0, row0, msk);
#pragma unroll
        for (int u = 0; u < BIN_BATCH; ++u)
            for (int e = 0; e < 5; ++e) {
                const int row = row0[u] + e;
                if ((msk[u] >> e & 1u) && row >= 0 && row < ny) {
                    const int pos = atomicAdd(&lc[row], 1);
                    ent[lp[row] + pos] = k0 + u * (int)blockDim.x;
                }
            }
    }
}

// Sort each row's entries by (bx, term index): the order in which a cell sums
// its contributions.  Bitonic sort in LDS for rows up to 4096 entries, an
// in-place insertion sort by one thread beyond that.
__device__ __forceinline__ void k_bin_sort_body(const Term* __restrict__ t, const int* __restrict__ ptr, int* __restrict__ ent, int bx)
{
    const int row = bx;
    const int lo = ptr[row], n = ptr[row + 1] - lo;
    if (n <= 1) return;
    __shared__ unsigned long long key[4096];
    auto mk = [&](int k) {
        return ((unsigned long long)(unsigned)(t[k].bx + (1 << 30)) << 32) | (unsigned)k;
    };
    if (n > 4096) {
        if (threadIdx.x == 0) {
            for (int i = 1; i < n; ++i) {
                const int v = ent[lo + i];
                const unsigned long long kv = mk(v);
                int j = i - 1;
                while (j >= 0 && mk(ent[lo + j]) > kv) {
                    ent[lo + j + 1] = ent[lo + j];
                    --j;
                }
                ent[lo + j + 1] = v;
            }
        }
        return;
    }
    int m = 1;
    while (m < n) m <<= 1;
    for (int i = threadIdx.x; i < m; i += blockDim.x) key[i] = i < n ? mk(ent[lo + i]) : ~0ull;
    __syncthreads();
    for (int size = 2; size <= m; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = threadIdx.x; i < m; i += blockDim.x) {
                const int j = i ^ stride;
                if (j > i) {
                    const bool up = (i & size) == 0;
                    const unsigned long long a = key[i], b = key[j];
                    if ((a > b) == up) {
                        key[i] = b;
                        key[j] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (int i = threadIdx.x; i < n; i += blockDim.x) ent[lo + i] = (int)(key[i] & 0xffffffffu);
}

__global__ void k_bin_sort(const Term* __restrict__ t, const int* __restrict__ ptr, int* __restrict__ ent)
{
    k_bin_sort_body(t, ptr, ent, blockIdx.x);
}

__device__ __forceinline__ void k_bin_bx_body(const Term* __restrict__ t, const int* __restrict__ ent, const int* __restrict__ n_dev,
                         int* __restrict__ ebx, int bx)
{
    const int k = bx * blockDim.x + threadIdx.x;
    if (k < *n_dev) ebx[k] = t[ent[k]].bx;
}

__global__ void k_bin_bx(const Term* __restrict__ t, const int* __restrict__ ent, const int* __restrict__ n_dev,
                         int* __restrict__ ebx)
{
    k_bin_bx_body(t, ent, n_dev, ebx, blockIdx.x);
}

// entry records of row j: amplitude index and coef * wy[j - by] * wx[e]
__device__ __forceinline__ void k_bin_rec_body(const Term* __restrict__ t, const int* __restrict__ ptr, const int* __restrict__ ent,
                          int* __restrict__ eai, double* __restrict__ erw, int bx)
{
    const int j = bx;
    for (int k = ptr[j] + threadIdx.x; k < ptr[j + 1]; k += blockDim.x) {
        const Term& tt = t[ent[k]];
        const double cw = tt.coef * tt.wy[j - tt.by];
        eai[k] = tt.ai;
#pragma unroll
        for (int e = 0; e < 5; ++e) erw[5 * (size_t)k + e] = cw * tt.wx[e];
    }
}

__global__ void k_bin_rec(const Term* __restrict__ t, const int* __restrict__ ptr, const int* __restrict__ ent,
                          int* __restrict__ eai, double* __restrict__ erw)
{
    k_bin_rec_body(t, ptr, ent, eai, erw, blockIdx.x);
}

// Cells of a row: the x positions some entry contributes to (nonzero weight).
// A work-group per row marks them in an LDS bitmap (x < 8192).
static constexpr int CELL_WORDS = 8192 / 32 + 1;

__device__ void mark_cells(const int* __restrict__ ptr, const int* __restrict__ ebx, const double* __restrict__ erw,
                           int j, int nx, unsigned* bm)
{
    for (int w = threadIdx.x; w < CELL_WORDS; w += blockDim.x) bm[w] = 0u;
    __syncthreads();
    for (int k = ptr[j] + threadIdx.x; k < ptr[j + 1]; k += blockDim.x)
        for (int e = 0; e < 5; ++e) {
            const int x = ebx[k] + e;
            if (x >= 0 && x < nx && erw[5 * (size_t)k + e] != 0.0) atomicOr(&bm[x >> 5], 1u << (x & 31));
        }
    __syncthreads();
}

__device__ __forceinline__ void k_cell_count_body(const int* __restrict__ ptr, const int* __restrict__ ebx, const double* __restrict__ erw,
                             int nx, int* __restrict__ cnt, int bx)
{
    __shared__ unsigned bm[CELL_WORDS];
    __shared__ int tot[256];
    const int j = bx;
    if (ptr[j] == ptr[j + 1]) {   // no entries: no cells
        if (threadIdx.x == 0) cnt[j] = 0;
        return;
    }
    mark_cells(ptr, ebx, erw, j, nx, bm);
    int c = 0;
    for (int w = threadIdx.x; w < CELL_WORDS; w += blockDim.x) c += __popc(bm[w]);
    tot[threadIdx.x] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        int s = 0;
        for (int i = 0; i < (int)blockDim.x; ++i) s += tot[i];
        cnt[j] = s;
    }
}

__global__ void k_cell_count(const int* __restrict__ ptr, const int* __restrict__ ebx, const double* __restrict__ erw,
                             int nx, int* __restrict__ cnt)
{
    k_cell_count_body(ptr, ebx, erw, nx, cnt, blockIdx.x);
}

static constexpr int CELL_LDS_ENTRIES = 4096;   // row entries whose bx are staged in LDS

__device__ __forceinline__ void k_cell_fill_body(const int* __restrict__ ptr, const int* __restrict__ ebx_g, const double* __restrict__ erw,
                            int nx, const int* __restrict__ cptr, int* __restrict__ cells3, int bx)
{
    __shared__ unsigned bm[CELL_WORDS];
    __shared__ int woff[CELL_WORDS];
    __shared__ int wsum[4];
    __shared__ int sbx[CELL_LDS_ENTRIES];
    const int j = bx;
    if (ptr[j] == ptr[j + 1]) return;   // no entries: no cells
    const int lo_g = ptr[j], nrow = ptr[j + 1] - lo_g;
    const bool staged = nrow <= CELL_LDS_ENTRIES;
    if (staged)
        for (int k = threadIdx.x; k < nrow; k += blockDim.x) sbx[k] = ebx_g[lo_g + k];
    // bx of the row's k-th entry (sorted), from LDS when staged.  Only indices
    // are offset, never the LDS base pointer: a pointer before an LDS array
    // that is used as a generic pointer can land outside the LDS aperture.
    auto bx_of = [&](int k) { return staged ? sbx[k] : ebx_g[lo_g + k]; };
    mark_cells(ptr, ebx_g, erw, j, nx, bm);   // its barriers also publish sbx
    // exclusive scan of the popcounts of words 0..255 (256 threads, 4 waves);
    // word 256 (x >= 8192) follows them
    {
        const int t = threadIdx.x, lane = t & 63;
        const int c = __popc(bm[t]);
        int v = c;
        for (int off = 1; off < 64; off <<= 1) {
            const int u = __shfl_up(v, off, 64);
            if (lane >= off) v += u;
        }
        if (lane == 63) wsum[t >> 6] = v;
        __syncthreads();
        int base = 0;
        for (int i = 0; i < (t >> 6); ++i) base += wsum[i];
        woff[t] = base + v - c;
        if (t == 255) woff[256] = base + v;
    }
    __syncthreads();
    // one thread per cell: the r-th set bit of the bitmap (x < nx <= 8192, so
    // word 256 is empty and woff[256] is the row's cell count)
    const int ncell = woff[256];
    for (int r = threadIdx.x; r < ncell; r += blockDim.x) {
        // the word holding rank r: the last w with woff[w] <= r
        int wl = 0, wh = 256;
        while (wh - wl > 1) {
            const int m = (wl + wh) >> 1;
            if (woff[m] <= r) wl = m; else wh = m;
        }
        unsigned bits = bm[wl];
        for (int k = r - woff[wl]; k > 0; --k) bits &= bits - 1;
        const int x = wl * 32 + __ffs(bits) - 1;
        // entries with bx in [x-4, x]: a contiguous run of the sorted row
        int a = 0, c = nrow;
        while (a < c) {
            const int m = (a + c) >> 1;
            if (bx_of(m) < x - 4) a = m + 1; else c = m;
        }
        int e = a;
        while (e < nrow && bx_of(e) <= x) ++e;
        int* cl = cells3 + 3 * (size_t)(cptr[j] + r);
        cl[0] = x;
        cl[1] = lo_g + a;
        cl[2] = lo_g + e;
    }
}

__global__ void k_cell_fill(const int* __restrict__ ptr, const int* __restrict__ ebx_g, const double* __restrict__ erw,
                            int nx, const int* __restrict__ cptr, int* __restrict__ cells3)
{
    k_cell_fill_body(ptr, ebx_g, erw, nx, cptr, cells3, blockIdx.x);
}

// contributions of cell c (x, entries [klo, khi)): the entries with a nonzero weight at x
__device__ __forceinline__ void k_contrib_count_body(const int* __restrict__ cells3, const int* __restrict__ n_dev,
                                const int* __restrict__ ebx, const double* __restrict__ erw, int* __restrict__ cnt, int bx)
{
    const int c = bx * blockDim.x + threadIdx.x;
    if (c >= *n_dev) return;
    const int x = cells3[3 * (size_t)c], klo = cells3[3 * (size_t)c + 1], khi = cells3[3 * (size_t)c + 2];
    int m = 0;
    for (int k = klo; k < khi; ++k) m += erw[5 * (size_t)k + (x - ebx[k])] != 0.0;
    cnt[c] = m;
}

__global__ void k_contrib_count(const int* __restrict__ cells3, const int* __restrict__ n_dev,
                                const int* __restrict__ ebx, const double* __restrict__ erw, int* __restrict__ cnt)
{
    k_contrib_count_body(cells3, n_dev, ebx, erw, cnt, blockIdx.x);
}

__device__ __forceinline__ void k_contrib_fill_body(const int* __restrict__ cells3, const int* __restrict__ n_dev,
                               const int* __restrict__ ebx, const int* __restrict__ eai,
                               const double* __restrict__ erw, const int* __restrict__ off, int4* __restrict__ cells,
                               int* __restrict__ cai, double* __restrict__ cw, int bx)
{
    const int c = bx * blockDim.x + threadIdx.x;
    if (c >= *n_dev) return;
    const int x = cells3[3 * (size_t)c], klo = cells3[3 * (size_t)c + 1], khi = cells3[3 * (size_t)c + 2];
    int e = off[c];
    for (int k = klo; k < khi; ++k) {
        const double w = erw[5 * (size_t)k + (x - ebx[k])];
        if (w != 0.0) {
            cai[e] = eai[k];
            cw[e] = w;
            ++e;
        }
    }
    cells[c] = make_int4(x, off[c], e, 0);
}

__global__ void k_contrib_fill(const int* __restrict__ cells3, const int* __restrict__ n_dev,
                               const int* __restrict__ ebx, const int* __restrict__ eai,
                               const double* __restrict__ erw, const int* __restrict__ off, int4* __restrict__ cells,
                               int* __restrict__ cai, double* __restrict__ cw)
{
    k_contrib_fill_body(cells3, n_dev, ebx, eai, erw, off, cells, cai, cw, blockIdx.x);
}

// Row bins of a term list for one output grid kind.  Buffers are sized by
// upper bounds (a term touches at most 5 rows, an entry at most 5 cells), so
// the whole build runs asynchronously on the stream.
namespace {
// The bins' sizes and buffers for nterms terms on grid kind gk (host side of
// build_bins: buffers grow, never shrink).
struct BinDims {
    int ny, nx;
    long c_max;
    size_t E, Cm;
};

BinDims bins_prepare(ilh_plan* pl, int nterms, int gk, RowBins& b)
{
    b.gk = gk;
    b.ny = pl->g.ny(gk);
    const int nx = pl->g.nx(gk);
    if (nx > 8192) fail(ILH_ERR_ARG, "grid too wide for the cell index (nx > 8192)");
    const long e_max = 5L * std::max(0, nterms), c_max = 5L * e_max;
    if (c_max > (1L << 30)) fail(ILH_ERR_ARG, "too many surface terms for the row bins");
    // buffers are kept from the previous build when large enough
    auto grow = [](auto*& p, size_t n) {
        if (p) (void)hipFree(p);
        p = nullptr;
        ILH_HIP(hipMalloc((void**)&p, sizeof(*p) * n));
    };
    if ((size_t)b.ny + 1 > b.cap_rows) {
        const size_t r = (size_t)b.ny + 1;
        grow(b.row_ptr, r);
        grow(b.cell_ptr, r);
        grow(b.cnt, 2 * r + 4);   // counts, cursors, three totals
        b.cap_rows = r;
    }
    if ((size_t)std::max(1L, e_max) > b.cap_e) {
        const size_t E = (size_t)std::max(1L, e_max) * 5 / 4;
        grow(b.entries, 2 * E);
        grow(b.eai, E);
        grow(b.erw, 5 * E);
        b.cap_e = E;
    }
    if ((size_t)std::max(1L, c_max) > b.cap_c) {
        const size_t Cm = (size_t)std::max(1L, c_max) * 5 / 4;
        grow(b.cells, Cm);
        grow(b.cai, Cm);
        grow(b.cw, Cm);
        grow(b.cells3, 3 * Cm);
        grow(b.ccnt, Cm);
        grow(b.coff, Cm + 1);
        b.cap_c = Cm;
    }
    if (b.ny > 16384) fail(ILH_ERR_ARG, "grid too tall for the row bins (more than 16384 rows)");
    return BinDims{b.ny, nx, c_max, b.cap_e, b.cap_c};
}

// the pair tables' buffers (host side of build_source_pairs)
void pairs_prepare(SourceSet& s)
{
    const int ny = s.bins.ny;
    const size_t np = (size_t)(ny + 1) / 2;
    // tables are kept from the previous geometry when large enough
    if (np + 1 > s.cap_pairs) {
        if (s.pinfo) (void)hipFree(s.pinfo);
        s.pinfo = nullptr;
        s.cap_pairs = np + 1;
        ILH_HIP(hipMalloc(&s.pinfo, sizeof(int4) * s.cap_pairs));
    }
    if ((size_t)ny + 1 > s.cap_rows) {
        if (s.slot_of_row) (void)hipFree(s.slot_of_row);
        s.slot_of_row = nullptr;
        s.cap_rows = (size_t)ny + 1;
        ILH_HIP(hipMalloc(&s.slot_of_row, sizeof(int) * s.cap_rows));   // rows, then the live count
    }
    s.npairs = (int)np;
    s.npairs_dev = s.slot_of_row + ny;
    s.inl_stride = 0;   // the inline copies describe pinned cells only
}
}  // namespace

void build_bins(ilh_plan* pl, const Term* terms, int nterms, int gk, RowBins& b)
{
    hipStream_t st = pl->ctx->stream;
    const BinDims d = bins_prepare(pl, nterms, gk, b);
    const int nx = d.nx;
    const long c_max = d.c_max;
    const size_t E = d.E, Cm = d.Cm;
    int* cnt = b.cnt;
    int* ebx = b.entries + E;
    const size_t lc_bytes = sizeof(int) * (size_t)b.ny;
    // entries per row, sorted by (bx, term)
    k_bin_build<<<1, 1024, 2 * lc_bytes + sizeof(int), st>>>(terms, nterms, gk, b.ny, b.row_ptr, b.entries);
    if (nterms > 0) {
        k_bin_sort<<<b.ny, 256, 0, st>>>(terms, b.row_ptr, b.entries);
        k_bin_bx<<<(unsigned)((E + 255) / 256), 256, 0, st>>>(terms, b.entries, b.row_ptr + b.ny, ebx);
        k_bin_rec<<<b.ny, 64, 0, st>>>(terms, b.row_ptr, b.entries, b.eai, b.erw);
    }
    ILH_HIP(hipGetLastError());
    // cells per row
    k_cell_count<<<b.ny, 256, 0, st>>>(b.row_ptr, ebx, b.erw, nx, cnt);
    k_scan<<<1, 1024, 0, st>>>(cnt, b.ny, b.cell_ptr, nullptr);
    // contributions per cell
    const unsigned g = (unsigned)((Cm + 255) / 256);
    if (nterms > 0) {
        k_cell_fill<<<b.ny, 256, 0, st>>>(b.row_ptr, ebx, b.erw, nx, b.cell_ptr, b.cells3);
        k_contrib_count<<<g, 256, 0, st>>>(b.cells3, b.cell_ptr + b.ny, ebx, b.erw, b.ccnt);
    }
    k_scan<<<1, 1024, 0, st>>>(b.ccnt, (int)c_max, b.coff, b.cell_ptr + b.ny);
    if (nterms > 0)
        k_contrib_fill<<<g, 256, 0, st>>>(b.cells3, b.cell_ptr + b.ny, ebx, b.eai, b.erw, b.coff, b.cells, b.cai,
                                          b.cw);
    ILH_HIP(hipGetLastError());
    // No read-back: every count stays on the device.  The buffers' bounds
    // hold by construction (entries <= 5 per term, cells and contributions
    // <= 5 per entry).
}

void free_bins(RowBins& b)
{
    for (void* p : {(void*)b.row_ptr, (void*)b.entries, (void*)b.eai, (void*)b.erw, (void*)b.cell_ptr, (void*)b.cells,
                    (void*)b.cai, (void*)b.cw, (void*)b.cnt, (void*)b.cells3, (void*)b.ccnt, (void*)b.coff})
        if (p) (void)hipFree(p);
    b = RowBins();
}

void free_sources(SourceSet& s)
{
    free_bins(s.bins);
    if (s.pinfo) (void)hipFree(s.pinfo);
    if (s.slot_of_row) (void)hipFree(s.slot_of_row);
    if (s.inl_ai) (void)hipFree(s.inl_ai);
    s = SourceSet();
}

// ----------------------------------------------------------- source pairs
// Row pairs (2p, 2p+1) with at least one entry, compacted in order: pinfo[k]
// = {2p, first cell of row 2p, first cell of row 2p+1, end of the pair's
// cells}, slot_of_row[2p] = 2k, slot_of_row[2p+1] = 2k+1 (-1 for rows of idle
// pairs), *npairs = live pairs.  One work-group of 1024 threads.
__device__ __forceinline__ void k_source_pairs_body(const int* __restrict__ row_ptr, const int* __restrict__ cell_ptr,
                                                    int ny, int4* __restrict__ pinfo, int* __restrict__ slot_of_row,
                                                    int* __restrict__ npairs)
{
    __shared__ int wsum[16];
    __shared__ int carry;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int np = (ny + 1) / 2;
    if (t == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < np; base += 1024) {
        const int p = base + t;
        const int j = 2 * p;
        bool on = false;
        if (p < np) {
            on = row_ptr[j + 1] > row_ptr[j];
            if (j + 1 < ny) on = on || row_ptr[j + 2] > row_ptr[j + 1];
        }
        const int f = on ? 1 : 0;
        int inc = f;
        for (int off = 1; off < 64; off <<= 1) {
            const int u = __shfl_up(inc, off, 64);
            if (lane >= off) inc += u;
        }
        if (lane == 63) wsum[wave] = inc;
        __syncthreads();
        int k = carry;
        for (int w = 0; w < wave; ++w) k += wsum[w];
        k += inc - f;
        if (p < np) {
            if (on) pinfo[k] = make_int4(j, cell_ptr[j], cell_ptr[j + 1], cell_ptr[min(j + 2, ny)]);
            slot_of_row[j] = on ? 2 * k : -1;
            if (j + 1 < ny) slot_of_row[j + 1] = on ? 2 * k + 1 : -1;
        }
        __syncthreads();
        if (t == 1023) carry = k + f;
        __syncthreads();
    }
    if (t == 0) *npairs = carry;
}

__global__ void __launch_bounds__(1024) k_source_pairs(const int* __restrict__ row_ptr,
                                                       const int* __restrict__ cell_ptr, int ny,
                                                       int4* __restrict__ pinfo, int* __restrict__ slot_of_row,
                                                       int* __restrict__ npairs)
{
    k_source_pairs_body(row_ptr, cell_ptr, ny, pinfo, slot_of_row, npairs);
}

// The source set's pair tables, built on the device from its row bins (no
// read-back: the row pass launches over the upper bound (ny + 1) / 2 and
// reads the live count).
void build_source_pairs(ilh_plan* pl, SourceSet& s)
{
    const int ny = s.bins.ny;
    pairs_prepare(s);
    k_source_pairs<<<1, 1024, 0, pl->ctx->stream>>>(s.bins.row_ptr, s.bins.cell_ptr, ny, s.pinfo, s.slot_of_row,
                                                    s.npairs_dev);
    ILH_HIP(hipGetLastError());
}

// The exact live pair count on the host (one read-back; for caches built
// once and used by many solves): the row pass then launches exactly.
// the first INL_K contributions of cells [0, ncells), by slot (layout of
// SourceSet::inl_ai)
__global__ void k_inline_contribs(const int4* __restrict__ cells, const int* __restrict__ cai,
                                  const double* __restrict__ cw, int ncells, int stride, int* __restrict__ ia,
                                  double* __restrict__ wa)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncells) return;
    const int4 cl = cells[c];
#pragma unroll
    for (int u = 0; u < INL_K; ++u) {
        const bool ok = cl.y + u < cl.z;
        ia[(size_t)u * stride + c] = ok ? cai[cl.y + u] : 0;
        wa[(size_t)u * stride + c] = ok ? cw[cl.y + u] : 0.0;
    }
}

void pin_source_pairs(ilh_plan* pl, SourceSet& s)
{
    if (!s.npairs_dev) return;
    hipStream_t st = pl->ctx->stream;
    int n = 0, ncells = 0;
    ILH_HIP(hipMemcpyAsync(&n, s.npairs_dev, sizeof(int), hipMemcpyDeviceToHost, st));
    ILH_HIP(hipMemcpyAsync(&ncells, s.bins.cell_ptr + s.bins.ny, sizeof(int), hipMemcpyDeviceToHost, st));
    ILH_HIP(hipStreamSynchronize(st));
    s.npairs = n;
    s.npairs_dev = nullptr;
    // the cells' first contributions copied by slot, for the row pass
    s.inl_stride = 0;
    if (ncells <= 0) return;
    const size_t cap = ((size_t)ncells + 255) / 256 * 256;
    if (cap > s.inl_cap) {
        if (s.inl_ai) (void)hipFree(s.inl_ai);
        s.inl_ai = nullptr;
        s.inl_w = nullptr;
        s.inl_cap = 0;
        void* p = nullptr;
        ILH_HIP(hipMalloc(&p, (sizeof(int) + sizeof(double)) * INL_K * cap));
        s.inl_ai = (int*)p;
        s.inl_w = (double*)((char*)p + sizeof(int) * INL_K * cap);
        s.inl_cap = cap;
    }
    k_inline_contribs<<<(ncells + 255) / 256, 256, 0, st>>>(s.bins.cells, s.bins.cai, s.bins.cw, ncells,
                                                            (int)s.inl_cap, s.inl_ai, s.inl_w);
    ILH_HIP(hipGetLastError());
    s.inl_stride = (int)s.inl_cap;
}


// --------------------------------------------------- batched source build
// The merged sources of many plans built together (the moving-body step of
// many trajectories): DDF windows, terms, the merge, the row bins and the
// pair tables, each kernel ONE launch over all plans (grid.y = plan; grid.x
// the largest plan's, blocks past a plan's own extent return at once).  The
// kernels are the single-plan ones' bodies, so the tables are identical.
namespace {
struct GeoJob {
    int N, ddf;
    Grid g;
    const double *X, *Y, *nxv, *nyv, *ds;
    Win* wins;
    Term* t[3];      // terms of op[k] (op[k] < 0: none)
    int op[3];
    Term* merged;    // [terms of op 0; terms of op 1, amplitude slots + N]
    int n0, n1;
    int nt, gk, ny, nx, c_max;
    int *cnt, *row_ptr, *entries, *ebx, *eai;
    double* erw;
    int *cell_ptr, *cells3, *ccnt, *coff;
    int4* cells;
    int* cai;
    double* cw;
    unsigned gE, gC;
    int4* pinfo;
    int *slot_of_row, *npairs_dev;
};

__global__ void k_geo_windows(const GeoJob* __restrict__ J)
{
    const GeoJob& j = J[blockIdx.y];
    k_ddf_windows_body(j.N, j.ddf, j.g, j.X, j.Y, j.wins, blockIdx.x);
}

__global__ void k_geo_terms(const GeoJob* __restrict__ J)
{
    const GeoJob& j = J[blockIdx.y];
    const int z = blockIdx.z;
    if (j.op[z] < 0) return;
    k_build_terms_body(j.op[z], j.N, j.g, j.wins, j.nxv, j.nyv, j.ds, j.t[z], blockIdx.x);
}

__global__ void k_geo_merge(const GeoJob* __restrict__ J)
{
    const GeoJob& j = J[blockIdx.y];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < j.n0) {
        j.merged[k] = j.t[0][k];
    } else if (k < j.n0 + j.n1) {
        Term t = j.t[1][k - j.n0];
        t.ai += j.N;
        j.merged[k] = t;
    }
}

__global__ void __launch_bounds__(1024) k_geo_bin_count(const GeoJob* __restrict__ J)
{
    const GeoJob& j = J[blockIdx.y];
    k_bin_count_body(j.merged, j.nt, j.gk, j.ny, j.cnt, 0);
}

// which: 0 entries per row, 1 cells per row, 2 contributions per cell
__global__ void __launch_bounds__(1024) k_geo_scan(const GeoJob* __restrict__ J, int which)
{
    const GeoJob& j = J[blockIdx.y];
    if (which == 0) k_scan_body(j.cnt, j.ny, j.row_ptr, nullptr, 0);
    else if (which == 1) k_scan_body(j.cnt, j.ny, j.cell_ptr, nullptr, 0);
    else k_scan_body(j.ccnt, j.c_max, j.coff, j.cell_ptr + j.ny, 0);
}

__global__ void __launch_bounds__(1024) k_geo_bin_fill(const GeoJob* __restrict__ J)
{
    const GeoJob& j = J[blockIdx.y];
    k_bin_fill_body(j.merged, j.nt, j.gk, j.ny, j.row_ptr, j.entries, 0);
}

__global__ void k_geo_bin_sort(const GeoJob* __restrict__ J)
{
    const GeoJob& j = J[blockIdx.y];
    if ((int)blockIdx.x >= j.ny) return;
    k_bin_sort_body(j.merged, j.row_ptr, j.entries, blockIdx.x);
}

__global__ void k_geo_bin_bx(const GeoJob* __restrict__ J)
{
    const GeoJob& j = J[blockIdx.y];
    if (blockIdx.x >= j.gE) return;
    k_bin_bx_body(j.merged, j.entries, j.row_ptr + j.ny, j.ebx, blockIdx.x);
}

__global__ void k_geo_bin_rec(const GeoJob* __restrict__ J)
{
    const GeoJob& j = J[blockIdx.y];
    if ((int)blockIdx.x >= j.ny) return;
    k_bin_rec_body(j.merged, j.row_ptr, j.entries, j.eai, j.erw, blockIdx.x);
}

__global__ void k_geo_cell_count(const GeoJob* __restrict__ J)
{
    const GeoJob& j = J[blockIdx.y];
    if ((int)blockIdx.x >= j.ny) return;
    k_cell_count_body(j.row_ptr, j.ebx, j.erw, j.nx, j.cnt, blockIdx.x);
}

__global__ void k_geo_cell_fill(const GeoJob* __restrict__ J)
{
    const GeoJob& j = J[blockIdx.y];
    if ((int)blockIdx.x >= j.ny) return;
    k_cell_fill_body(j.row_ptr, j.ebx, j.erw, j.nx, j.cell_ptr, j.cells3, blockIdx.x);
}

__global__ void k_geo_contrib_count(const GeoJob* __restrict__ J)
{
    const GeoJob& j = J[blockIdx.y];
    if (blockIdx.x >= j.gC) return;
    k_contrib_count_body(j.cells3, j.cell_ptr + j.ny, j.ebx, j.erw, j.ccnt, blockIdx.x);
}

__global__ void k_geo_contrib_fill(const GeoJob* __restrict__ J)
{
    const GeoJob& j = J[blockIdx.y];
    if (blockIdx.x >= j.gC) return;
    k_contrib_fill_body(j.cells3, j.cell_ptr + j.ny, j.ebx, j.eai, j.erw, j.coff, j.cells, j.cai, j.cw, blockIdx.x);
}

__global__ void __launch_bounds__(1024) k_geo_pairs(const GeoJob* __restrict__ J)
{
    const GeoJob& j = J[blockIdx.y];
    k_source_pairs_body(j.row_ptr, j.cell_ptr, j.ny, j.pinfo, j.slot_of_row, j.npairs_dev);
}
}  // namespace

void merge_sources_batched(ilh_ctx* ctx, const std::vector<SourceJob>& jobs, int op0, int op1, int extra_op, int gk)
{
    if (jobs.empty()) return;
    if (jobs.size() > 65535) fail(ILH_ERR_ARG, "too many plans for one batched source build");
    hipStream_t st = ctx->stream;
    // the previous call's job table may still be in flight from the host copy
    ILH_HIP(hipStreamSynchronize(st));
    std::vector<GeoJob> hj(jobs.size());
    unsigned gx_w = 1, gx_t = 1, gx_m = 1, gE = 1, gC = 1;
    int max_ny = 1;
    for (size_t i = 0; i < jobs.size(); ++i) {
        ilh_plan* pl = jobs[i].pl;
        const int N = pl->N;
        if (N <= 0) fail(ILH_ERR_ARG, "batched source build of a plan without points");
        GeoJob& j = hj[i];
        std::memset(&j, 0, sizeof(j));
        j.N = N;
        j.ddf = pl->ddf;
        j.g = pl->g;
        j.X = pl->X.as<double>();
        j.Y = pl->Y.as<double>();
        j.nxv = pl->nx.as<double>();
        j.nyv = pl->ny.as<double>();
        j.ds = pl->ds.as<double>();
        pl->wins.ensure(sizeof(Win) * 4 * (size_t)N);
        j.wins = pl->wins.as<Win>();
        const int ops[3] = {op0, op1, extra_op};
        for (int k = 0; k < 3; ++k) {
            j.op[k] = ops[k];
            if (ops[k] < 0) continue;
            pl->terms[ops[k]].ensure(sizeof(Term) * (size_t)N * op_terms_per_point(ops[k]));
            j.t[k] = pl->terms[ops[k]].as<Term>();
        }
        j.n0 = N * op_terms_per_point(op0);
        j.n1 = N * op_terms_per_point(op1);
        jobs[i].tbuf->ensure(sizeof(Term) * (size_t)(j.n0 + j.n1));
        j.merged = jobs[i].tbuf->as<Term>();
        SourceSet& src = *jobs[i].src;
        src.terms = j.merged;
        src.amp_len = 2 * N;
        const BinDims d = bins_prepare(pl, j.n0 + j.n1, gk, src.bins);
        RowBins& b = src.bins;
        j.nt = j.n0 + j.n1;
        j.gk = gk;
        j.ny = d.ny;
        j.nx = d.nx;
        j.c_max = (int)d.c_max;
        j.cnt = b.cnt;
        j.row_ptr = b.row_ptr;
        j.entries = b.entries;
        j.ebx = b.entries + d.E;
        j.eai = b.eai;
        j.erw = b.erw;
        j.cell_ptr = b.cell_ptr;
        j.cells3 = b.cells3;
        j.ccnt = b.ccnt;
        j.coff = b.coff;
        j.cells = b.cells;
        j.cai = b.cai;
        j.cw = b.cw;
        j.gE = (unsigned)((d.E + 255) / 256);
        j.gC = (unsigned)((d.Cm + 255) / 256);
        pairs_prepare(src);
        j.pinfo = src.pinfo;
        j.slot_of_row = src.slot_of_row;
        j.npairs_dev = src.npairs_dev;
        gx_w = std::max(gx_w, (unsigned)((16 * N + 255) / 256));
        gx_t = std::max(gx_t, (unsigned)((N + 127) / 128));
        gx_m = std::max(gx_m, (unsigned)((j.n0 + j.n1 + 255) / 256));
        gE = std::max(gE, j.gE);
        gC = std::max(gC, j.gC);
        max_ny = std::max(max_ny, d.ny);
    }
    const size_t bytes = sizeof(GeoJob) * hj.size();
    if (bytes > ctx->geo_cap) {
        if (ctx->geo_dev) (void)hipFree(ctx->geo_dev);
        ctx->geo_dev = nullptr;
        ctx->geo_cap = 0;
        ILH_HIP(hipMalloc(&ctx->geo_dev, bytes));
        ctx->geo_cap = bytes;
    }
    ctx->geo_host.assign((const char*)hj.data(), (const char*)hj.data() + bytes);
    const GeoJob* J = (const GeoJob*)ctx->geo_dev;
    ILH_HIP(hipMemcpyAsync(ctx->geo_dev, ctx->geo_host.data(), bytes, hipMemcpyHostToDevice, st));
    const unsigned np = (unsigned)hj.size();
    const size_t lc_bytes = sizeof(int) * (size_t)max_ny;
    k_geo_windows<<<dim3(gx_w, np), 256, 0, st>>>(J);
    k_geo_terms<<<dim3(gx_t, np, 3), 128, 0, st>>>(J);
    k_geo_merge<<<dim3(gx_m, np), 256, 0, st>>>(J);
    k_geo_bin_count<<<dim3(1, np), 1024, lc_bytes, st>>>(J);
    k_geo_scan<<<dim3(1, np), 1024, 0, st>>>(J, 0);
    k_geo_bin_fill<<<dim3(1, np), 1024, lc_bytes, st>>>(J);
    k_geo_bin_sort<<<dim3((unsigned)max_ny, np), 256, 0, st>>>(J);
    k_geo_bin_bx<<<dim3(gE, np), 256, 0, st>>>(J);
    k_geo_bin_rec<<<dim3((unsigned)max_ny, np), 64, 0, st>>>(J);
    k_geo_cell_count<<<dim3((unsigned)max_ny, np), 256, 0, st>>>(J);
    k_geo_scan<<<dim3(1, np), 1024, 0, st>>>(J, 1);
    k_geo_cell_fill<<<dim3((unsigned)max_ny, np), 256, 0, st>>>(J);
    k_geo_contrib_count<<<dim3(gC, np), 256, 0, st>>>(J);
    k_geo_scan<<<dim3(1, np), 1024, 0, st>>>(J, 2);
    k_geo_contrib_fill<<<dim3(gC, np), 256, 0, st>>>(J);
    k_geo_pairs<<<dim3(1, np), 1024, 0, st>>>(J);
    ILH_HIP(hipGetLastError());
    for (const SourceJob& sj : jobs) {
        sj.pl->wins_ok = true;
        for (int op : {op0, op1, extra_op})
            if (op >= 0) sj.pl->terms_ok[op] = true;
    }
}

// --------------------------------------------------------------- scatter
// out(x, j): zero, except at the row's cells, whose value is the sum of their
// contributions in (bx, term) order; with a base field, base(x, j) plus c
// times that.  A work-group per row: a coalesced zero (or base) sweep, then
// one thread per cell.
// FP64 (T = double) or FP32 (T = float) data; FP32 values are accumulated
// in FP64 with the FP64 weights and rounded once.
template <class T>
__global__ void k_scatter(const int* __restrict__ cell_ptr, const int4* __restrict__ cells,
                          const int* __restrict__ cai, const double* __restrict__ cw, int nx,
                          const T* __restrict__ amp, size_t amp_stride, T* __restrict__ out, size_t out_stride,
                          const T* __restrict__ base, double c)
{
    const int j = blockIdx.x;
    amp += amp_stride * blockIdx.y;
    T* row = out + out_stride * blockIdx.y + (size_t)nx * j;
    const T* brow = base ? base + out_stride * blockIdx.y + (size_t)nx * j : nullptr;
    for (int x = threadIdx.x; x < nx; x += blockDim.x) row[x] = brow ? brow[x] : T(0);
    const int c0 = cell_ptr[j], c1 = cell_ptr[j + 1];
    if (c0 == c1) return;   // uniform over the work-group
    __syncthreads();        // the zeros are ordered before the cell values
    for (int ci = c0 + threadIdx.x; ci < c1; ci += blockDim.x) {
        const int4 cr = cells[ci];
        double acc = 0.0;
        for (int e = cr.y; e < cr.z; ++e) acc += (double)amp[cai[e]] * cw[e];
        row[cr.x] = brow ? (T)((double)brow[cr.x] + c * acc) : (T)acc;
    }
}

template <class T>
void scatter(ilh_plan* pl, const RowBins& bins, const T* amp, size_t amp_stride, int nb, T* out, size_t out_stride,
             const T* base, double c)
{
    const int nx = pl->g.nx(bins.gk);
    if (bins.ny <= 0 || nb <= 0) return;
    k_scatter<T><<<dim3(bins.ny, nb), 256, 0, pl->ctx->stream>>>(bins.cell_ptr, bins.cells, bins.cai, bins.cw, nx,
                                                                 amp, amp_stride, out, out_stride, base, c);
    ILH_HIP(hipGetLastError());
}
template void scatter<double>(ilh_plan*, const RowBins&, const double*, size_t, int, double*, size_t, const double*,
                              double);
template void scatter<float>(ilh_plan*, const RowBins&, const float*, size_t, int, float*, size_t, const float*,
                             double);

// ---------------------------------------------------------------- gather
template <class T>
struct GatherFields {
    const T* f[GK_COUNT];   // by component grid
    int nx[GK_COUNT], ny[GK_COUNT];
};

// One wave per point: lane l takes the (term, row, column) triples l, l + 64,
// ... of the point's T terms (T x 5 x 5), each the field value times coef wy
// wx, into the partial sum of its term's amplitude slot; the wave then sums
// the two slots' partials in a fixed butterfly order (bitwise reproducible).
template <class D>
__global__ void __launch_bounds__(256) k_gather(const Term* __restrict__ t, int T, int N, GatherFields<D> F,
                                                size_t field_stride, D* __restrict__ out, size_t out_stride)
{
    const int lane = threadIdx.x & 63;
    const int p = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (p >= N) return;   // uniform over the wave
    out += out_stride * blockIdx.y;
    const size_t foff = field_stride * blockIdx.y;
    const Term* tp = t + (size_t)p * T;
    // the point's amplitude indices, term k on lane k (T <= 64, checked by the
    // launcher): ai0 of term 0 and the first that differs from it, in one load
    // round trip
    const int aik = lane < T ? tp[lane].ai : 0;
    const int ai0 = __shfl(aik, 0, 64);
    const unsigned long long dm = __ballot(lane < T && aik != ai0);
    const int ai1 = dm ? __shfl(aik, __ffsll((long long)dm) - 1, 64) : -1;
    double acc0 = 0.0, acc1 = 0.0;
    for (int e = lane; e < T * 25; e += 64) {
        const int k = e / 25, r = e - 25 * k, b = r / 5, a = r - 5 * b;
        const Term& tt = tp[k];
        const int gk = tt.gk, row = tt.by + b, col = tt.bx + a;
        const int nx = F.nx[gk], ny = F.ny[gk];
        double v = 0.0;
        const double w = tt.wy[b] * tt.wx[a];
        if (w != 0.0 && row >= 0 && row < ny && col >= 0 && col < nx)
            v = tt.coef * w * (double)F.f[gk][foff + col + (size_t)nx * row];
        if (tt.ai == ai0) acc0 += v;
        else acc1 += v;
    }
    for (int off = 32; off > 0; off >>= 1) {
        acc0 += __shfl_xor(acc0, off, 64);
        acc1 += __shfl_xor(acc1, off, 64);
    }
    if (lane == 0) {
        out[ai0] = (D)acc0;
        if (ai1 >= 0) out[ai1] = (D)acc1;
    }
}

template <class D>
void gather(ilh_plan* pl, const Term* terms, int T, const D* field, size_t field_stride, int nb, D* out,
            size_t out_stride)
{
    if (pl->N == 0 || nb <= 0) return;
    if (T < 1 || T > 64) fail(ILH_ERR_ARG, "gather: terms per point must fit one wave");
    GatherFields<D> F;
    const Grid& g = pl->g;
    for (int k = 0; k < GK_COUNT; ++k) {
        F.f[k] = field + g.offset(k);
        F.nx[k] = g.nx(k);
        F.ny[k] = g.ny(k);
    }
    k_gather<D><<<dim3((pl->N + 3) / 4, nb), 256, 0, pl->ctx->stream>>>(terms, T, pl->N, F, field_stride, out,
                                                                        out_stride);
    ILH_HIP(hipGetLastError());
}
template void gather<double>(ilh_plan*, const Term*, int, const double*, size_t, int, double*, size_t);
template void gather<float>(ilh_plan*, const Term*, int, const float*, size_t, int, float*, size_t);

}  // namespace ilh
