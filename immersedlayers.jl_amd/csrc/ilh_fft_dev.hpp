// ilh_fft_dev.hpp -- device side of the zero-padded LGF convolution: FFT
// plans (compile-time stage radices, strides and twiddle offsets), the
// Stockham radix-8 transform in registers with LDS exchanges, the row
// sources, and the bodies of the three passes (row forward, column
// forward + spectrum product + inverse, row inverse).  A body handles one
// logical work-group's transforms (`blk`, NB transforms of T = P/V threads
// each); the single-problem kernels (ilh_fft.hip) and the cross-problem
// batched kernels (ilh_batch.hip) both launch them.
#pragma once
#include "ilh_internal.hpp"

namespace ilh {

static constexpr int MAXP = 8192;
static constexpr int V = 8;   // complex values per thread; a length-P transform uses P/V threads

// --------------------------------------------------------------- plans
// P = 8^m * rem (rem in {1, 2, 4}).  "fwd" order: m radix-8 stages then rem;
// "rev" order: rem first, so that the inverse of the column pass starts from
// the registers the forward's last stage leaves.
__host__ __device__ constexpr int n_full(int P)
{
    int n = 0;
    while (P >= V) { P /= V; ++n; }
    return n;
}
__host__ __device__ constexpr int rem_of(int P)
{
    while (P >= V) P /= V;
    return P;
}
__host__ __device__ constexpr int n_stages(int P) { return n_full(P) + (rem_of(P) > 1 ? 1 : 0); }
__host__ __device__ constexpr int radix(int P, bool rev, int s)
{
    const int r = rem_of(P);
    if (r == 1) return V;
    if (!rev) return s < n_full(P) ? V : r;
    return s == 0 ? r : V;
}
// stride of stage s: the product of the radices before it
__host__ __device__ constexpr int nsub(int P, bool rev, int s)
{
    int n = 1;
    for (int k = 0; k < s; ++k) n *= radix(P, rev, k);
    return n;
}
// twiddle table: fwd stages then rev stages; stage s holds
// W_{Ns R}^{r k} = exp(-2 pi i r k / (Ns R)) at [(r-1) Ns + k], k < Ns, 1 <= r < R,
// so that the 64 lanes of a wave loading one r read consecutive k: 1 KB in
// 8 cache lines per load instead of 64 lanes 112 bytes apart
__host__ __device__ constexpr int tw_count(int P, bool rev)
{
    int c = 0;
    for (int k = 0; k < n_stages(P); ++k) c += nsub(P, rev, k) * (radix(P, rev, k) - 1);
    return c;
}
__host__ __device__ constexpr int tw_off(int P, bool rev, int s)
{
    int c = rev ? tw_count(P, false) : 0;
    for (int k = 0; k < s; ++k) c += nsub(P, rev, k) * (radix(P, rev, k) - 1);
    return c;
}
// threads per work-group: one transform, or several short ones
__host__ __device__ constexpr int tpb_of(int P) { return P / V > 64 ? P / V : 64; }

template <class R> struct Cplx;
template <> struct Cplx<double> {
    using C2 = double2;
    using Q4 = double4;
};
template <> struct Cplx<float> {
    using C2 = float2;
    using Q4 = float4;
};

template <class C2> __device__ __forceinline__ C2 mk2(decltype(C2().x) a, decltype(C2().x) b)
{
    C2 r;
    r.x = a;
    r.y = b;
    return r;
}

// ---------------------------------------------------------- butterflies
template <class C2> __device__ __forceinline__ C2 cmul(C2 a, C2 b)
{
    return mk2<C2>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <class C2> __device__ __forceinline__ C2 cadd(C2 a, C2 b) { return mk2<C2>(a.x + b.x, a.y + b.y); }
template <class C2> __device__ __forceinline__ C2 csub(C2 a, C2 b) { return mk2<C2>(a.x - b.x, a.y - b.y); }
template <int S, class C2>   // multiply by S*i
__device__ __forceinline__ C2 mul_si(C2 a) { return S > 0 ? mk2<C2>(-a.y, a.x) : mk2<C2>(a.y, -a.x); }

template <int S, class C2>
__device__ __forceinline__ void dft2(C2& a, C2& b)
{
    const C2 t = a;
    a = cadd(t, b);
    b = csub(t, b);
}

template <int S, class C2>
__device__ __forceinline__ void dft4(C2& a0, C2& a1, C2& a2, C2& a3)
{
    const C2 t0 = cadd(a0, a2), t1 = csub(a0, a2);
    const C2 t2 = cadd(a1, a3), t3 = mul_si<S>(csub(a1, a3));
    a0 = cadd(t0, t2);
    a2 = csub(t0, t2);
    a1 = cadd(t1, t3);
    a3 = csub(t1, t3);
}

template <int S, class C2>
__device__ __forceinline__ void dft8(C2* a)
{
    using R = decltype(C2().x);
    const R c = (R)0.70710678118654752440084436210485;
    C2 e0 = a[0], e1 = a[2], e2 = a[4], e3 = a[6];
    C2 o0 = a[1], o1 = a[3], o2 = a[5], o3 = a[7];
    dft4<S>(e0, e1, e2, e3);
    dft4<S>(o0, o1, o2, o3);
    o1 = cmul(o1, mk2<C2>(c, S * c));
    o2 = mul_si<S>(o2);
    o3 = cmul(o3, mk2<C2>(-c, S * c));
    a[0] = cadd(e0, o0);
    a[4] = csub(e0, o0);
    a[1] = cadd(e1, o1);
    a[5] = csub(e1, o1);
    a[2] = cadd(e2, o2);
    a[6] = csub(e2, o2);
    a[3] = cadd(e3, o3);
    a[7] = csub(e3, o3);
}

template <int R, int S, class C2>
__device__ __forceinline__ void dft(C2* a)
{
    if constexpr (R == 8) dft8<S>(a);
    else if constexpr (R == 4) dft4<S>(a[0], a[1], a[2], a[3]);
    else dft2<S>(a[0], a[1]);
}

// ------------------------------------------------------- stage helpers
// LDS exchange buffers hold P complex values; element i sits at slot
// i ^ ((i >> 3) & 7), which permutes each aligned run of 8 by its run index.
// Runs of consecutive elements (the loads of every stage) and the stride-8
// stores of a first stage then both fall in distinct banks: the column pass
// measures 128 bank-conflict cycles per launch against 786,560 with one pad
// slot after every 8 elements.
__device__ __forceinline__ int lx(int i) { return i ^ ((i >> 3) & 7); }
__host__ __device__ constexpr int lds_len(int P) { return P; }

// Thread t of a transform owns butterflies j = t + q T (q < V/R) of every
// stage; v[q R + r] is element r of butterfly q, read from / written to
// index j + r P/R (first / last stage: the natural order of input / output).
__device__ __forceinline__ int io_index(int R, int q, int r, int t, int T, int P)
{
    return t + q * T + r * (P / R);
}

template <int P, bool REV, int s, class C2>
__device__ __forceinline__ void tw_load(C2 (&w)[V - 1], int t, const C2* __restrict__ tw)
{
    constexpr int R = radix(P, REV, s), Ns = nsub(P, REV, s), T = P / V;
    if constexpr (Ns > 1) {
        const C2* base = tw + tw_off(P, REV, s);
#pragma unroll
        for (int q = 0; q < V / R; ++q) {
            const int k = (t + q * T) & (Ns - 1);
#pragma unroll
            for (int r = 1; r < R; ++r) w[q * (R - 1) + r - 1] = base[(r - 1) * Ns + k];
        }
    }
}

template <int P, int S, bool REV, int s, class C2>
__device__ __forceinline__ void st_compute(C2 (&v)[V], const C2 (&w)[V - 1])
{
    constexpr int R = radix(P, REV, s), Ns = nsub(P, REV, s);
#pragma unroll
    for (int q = 0; q < V / R; ++q) {
        if constexpr (Ns > 1) {
#pragma unroll
            for (int r = 1; r < R; ++r) {
                C2 ww = w[q * (R - 1) + r - 1];
                if (S > 0) ww.y = -ww.y;
                v[q * R + r] = cmul(v[q * R + r], ww);
            }
        }
        dft<R, S>(v + q * R);
    }
}

template <int P, bool REV, int s, class C2>
__device__ __forceinline__ void st_store(const C2 (&v)[V], C2* x, int t)
{
    constexpr int R = radix(P, REV, s), Ns = nsub(P, REV, s), T = P / V;
#pragma unroll
    for (int q = 0; q < V / R; ++q) {
        const int j = t + q * T;
        const int base = (j & ~(Ns - 1)) * R + (j & (Ns - 1));
#pragma unroll
        for (int r = 0; r < R; ++r) x[lx(base + r * Ns)] = v[q * R + r];
    }
}

template <int P, int R, class C2>
__device__ __forceinline__ void st_load(C2 (&v)[V], const C2* x, int t)
{
    constexpr int T = P / V;
#pragma unroll
    for (int q = 0; q < V / R; ++q)
#pragma unroll
        for (int r = 0; r < R; ++r) v[q * R + r] = x[lx(io_index(R, q, r, t, T, P))];
}

struct NoHook {
    __device__ void operator()() const {}
};

template <int P, int S, bool REV, int s, class Pre, class C2>
__device__ __forceinline__ void fft_from(C2 (&v)[V], C2 (&w)[V - 1], C2* x, int t, const C2* __restrict__ tw,
                                         Pre& pre)
{
    st_compute<P, S, REV, s>(v, w);
    if constexpr (s + 1 < n_stages(P)) {
        tw_load<P, REV, s + 1>(w, t, tw);   // in flight across the exchange
        __syncthreads();
        st_store<P, REV, s>(v, x, t);
        __syncthreads();
        st_load<P, radix(P, REV, s + 1)>(v, x, t);
        if constexpr (s + 2 == n_stages(P)) pre();
        fft_from<P, S, REV, s + 1>(v, w, x, t, tw, pre);
    }
}

// Length-P transform, sign S (-1 forward, +1 inverse, unnormalised), stage
// order REV.  On entry v holds the first stage's inputs (element (q, r) =
// x[io_index(R0, q, r)]), on exit the last stage's outputs (X[io_index(RL, q, r)]).
// `x` is the transform's LDS exchange buffer; the caller synchronises the
// work-group before the first exchange if it used the buffer.  `pre` runs
// before the last stage (loads issued there overlap its butterflies).
template <int P, int S, bool REV, class Pre = NoHook, class C2>
__device__ __forceinline__ void fft(C2 (&v)[V], C2* x, int t, const C2* __restrict__ tw, Pre pre = Pre())
{
    C2 w[V - 1];
    fft_from<P, S, REV, 0>(v, w, x, t, tw, pre);
}

template <int P> constexpr int R_first(bool rev) { return radix(P, rev, 0); }
template <int P> constexpr int R_last(bool rev) { return radix(P, rev, n_stages(P) - 1); }

// Work-groups are dispatched round-robin over the 8 XCDs (block b -> XCD b%8).
// Renumber blocks so that each XCD gets a run of consecutive logical blocks:
// neighbours that share 128-byte lines of the tiled spectra then meet in one L2.
__device__ __forceinline__ int xcd_block_of(int b, int nblk)
{
    const int full = (nblk / 8) * 8;
    return b < full ? (b % 8) * (nblk / 8) + b / 8 : b;
}

__device__ __forceinline__ int xcd_block() { return xcd_block_of(blockIdx.x, gridDim.x); }

// --------------------------------------------------------- row sources
enum RowMode { ROW_DENSE = 0, ROW_LGF = 1, ROW_TERMS = 2 };

struct RowSrc {
    int mode;
    const double* field;   // dense field (nx x ny)
    const float* field32;  // the same in FP32 (a Float32 cache), read instead of field when set
    int nx, ny;
    const double* tab;     // LGF table (circular rows of the Px x Py kernel)
    int tmx, Px, Py;
    const int4* pinfo;     // surface sources: per row pair {j0, cells of row j0, of row j0+1, end},
    const int4* cells;     // cells {x, c0, c1, 0}, contributions cw[e] * amp[cai[e]]
    const int* cai;
    const double* cw;
    const double* amp;
    int amp_len;
    int amp_lds;           // 1: stage amp[0..amp_len) in LDS after the transform buffers
    const int* npairs_dev; // surface sources: the live pair count on the device (the
                           // launch covers an upper bound), or null
    const int* inl_ai;     // surface sources: first contributions of each cell by slot
    const double* inl_w;   // (SourceSet::inl_ai), or null
    int inl_stride;
};

// values of two cells (a second with c0 == c1 is empty): each cell's
// contributions in order, loaded 8 per cell at a time.  Every amplitude of a
// round is read before the first is used (an empty slot reads amplitude 0,
// which exists whenever a cell does), so the reads are in flight together
// instead of one memory round trip per contribution.
template <bool LDS_AMP>
__device__ __forceinline__ void cell_values(const RowSrc& S, const double* a_lds, int4 ca, int4 cb, double& va,
                                            double& vb)
{
    double acc[2] = {0.0, 0.0};
    int e[2] = {ca.y, cb.y};
    const int end[2] = {ca.z, cb.z};
    while (e[0] < end[0] || e[1] < end[1]) {
        int a[2][8];
        double w[2][8], av[2][8];
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const bool ok = e[h] + u < end[h];
                a[h][u] = ok ? S.cai[e[h] + u] : 0;
                w[h][u] = ok ? S.cw[e[h] + u] : 0.0;
            }
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int u = 0; u < 8; ++u) av[h][u] = LDS_AMP ? a_lds[a[h][u]] : S.amp[a[h][u]];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (e[h] + u < end[h]) acc[h] = __builtin_fma(w[h][u], av[h][u], acc[h]);
            e[h] += 8;
        }
    }
    va = acc[0];
    vb = acc[1];
}

// cell_values with each cell's first 8 contributions already loaded (ia, wa:
// the inline copies, 0 past a cell's count): the same sums in the same order
template <bool LDS_AMP>
__device__ __forceinline__ void cell_values_inl(const RowSrc& S, const double* a_lds, int4 ca, int4 cb,
                                                const int (&ia)[2][8], const double (&wa)[2][8], double& va,
                                                double& vb)
{
    double acc[2] = {0.0, 0.0};
    const int c0[2] = {ca.y, cb.y};
    const int end[2] = {ca.z, cb.z};
    double av[2][8];
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int u = 0; u < 8; ++u) av[h][u] = LDS_AMP ? a_lds[ia[h][u]] : S.amp[ia[h][u]];
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (c0[h] + u < end[h]) acc[h] = __builtin_fma(wa[h][u], av[h][u], acc[h]);
    int e[2] = {c0[0] + 8, c0[1] + 8};
    while (e[0] < end[0] || e[1] < end[1]) {
        int a[2][8];
        double w[2][8];
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const bool ok = e[h] + u < end[h];
                a[h][u] = ok ? S.cai[e[h] + u] : 0;
                w[h][u] = ok ? S.cw[e[h] + u] : 0.0;
            }
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int u = 0; u < 8; ++u) av[h][u] = LDS_AMP ? a_lds[a[h][u]] : S.amp[a[h][u]];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (e[h] + u < end[h]) acc[h] = __builtin_fma(w[h][u], av[h][u], acc[h]);
            e[h] += 8;
        }
    }
    va = acc[0];
    vb = acc[1];
}

// max amplitudes staged in LDS beside NB transforms of length P
__host__ __device__ constexpr int amp_lds_cap(int P)
{
    const int room = (160 * 1024 - (tpb_of(P) / (P / V)) * lds_len(P) * 16) / 8;
    return room < 4096 ? room : 4096;
}

// Spectrum intermediates use an 8-row tiled layout: element (row, c) at
// (row/8)*8*H + c*8 + row%8, so a column's 8-row run is one 128-byte segment
// (column pass) and a row pair's value at c is 32 contiguous bytes (row passes).
__device__ __forceinline__ size_t tidx(int row, int c, int H)
{
    return (size_t)(row >> 3) * 8 * H + (size_t)c * 8 + (row & 7);
}
// The column pass output uses TR-row tiles instead (TR = 2 or 4): a row
// pair's value at c is 32 contiguous bytes (the inverse row pass's loads).
// TR = 2: consecutive c adjacent, so the inverse row pass reads whole lines
// and the column pass writes 32-byte pieces.  TR = 4: a column's run of 4
// rows is 64 bytes, written with streaming stores, and each line holds two
// columns.  One convolution at a time (1024^2) runs 37.9 us with TR = 4
// against 38.7 (TR = 2) and 40.4 (8-row tiles); launches over many data sets
// or problems, whose inverse row pass dominates, keep TR = 2 (8 data sets:
// 41k against 38.6k solves/s; the 64-problem batch: 77.5k against 75.4k).
template <int TR> __device__ __forceinline__ size_t pidx(int row, int c, int H)
{
    static_assert(TR == 2 || TR == 4, "T2 tile rows");
    return (size_t)(row / TR) * TR * H + (size_t)c * TR + (row % TR);
}

// ------------------------------------------------------------- pass 1
// Transform c of logical block blk handles row pair p = blk NB + c (field rows
// 2p, 2p+1, or pinfo[p].x, +1) and writes the two half spectra to rows 2p,
// 2p+1 of the tiled Tout (H = P/2 columns, column 0 = (DC, Nyquist)).  The
// work-group is NB transforms of T = P/V threads.  TERMS: the source is
// surface terms only (the other kinds compiled out).
template <int P, int NB, class R, bool TERMS = false>
__device__ __forceinline__ void rows_fwd_body(RowSrc S, const typename Cplx<R>::C2* __restrict__ tw, int npairs,
                                              int nrows_limit, typename Cplx<R>::C2* __restrict__ Tout, int blk,
                                              typename Cplx<R>::C2* lds, int tid)
{
    using C2 = typename Cplx<R>::C2;
    using Q4 = typename Cplx<R>::Q4;
    constexpr int T = P / V, H = P / 2;
    constexpr int R0 = R_first<P>(false), RL = R_last<P>(false);
    const int c = tid / T, t = tid - c * T;
    const int p = blk * NB + c;
    const bool live = p < npairs;
    C2* x = lds + (size_t)c * lds_len(P);
    C2 v[V];
    if (TERMS || S.mode == ROW_TERMS) {
        // the packed pair (row j0 -> real part, row j0+1 -> imaginary part) in
        // LDS: zeros, then the rows' nonzero cells
        double* a_lds = (double*)(lds + (size_t)NB * lds_len(P));   // 16-byte aligned (NB * P slots)
        const int4 pi = live ? S.pinfo[p] : make_int4(0, 0, 0, 0);
        // the first round of cells (and, with the inline copies, their first
        // contributions), in flight across the amplitude staging
        const int ci0 = pi.y + t;
        const int4 ca0 = ci0 < pi.w ? S.cells[ci0] : make_int4(0, 0, 0, 0);
        const int4 cb0 = ci0 + T < pi.w ? S.cells[ci0 + T] : make_int4(0, 0, 0, 0);
        int ia0[2][8];
        double wa0[2][8];
        if (S.inl_ai) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int ci = ci0 + h * T;
                const bool ok = ci < pi.w;
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    ia0[h][u] = ok ? S.inl_ai[(size_t)u * S.inl_stride + ci] : 0;
                    wa0[h][u] = ok ? S.inl_w[(size_t)u * S.inl_stride + ci] : 0.0;
                }
            }
        }
        if (S.amp_lds) {
            // 16 loads in flight per thread, then the LDS stores
            for (int base = 0; base < S.amp_len; base += 16 * (int)blockDim.x) {
                double tmp[16];
#pragma unroll
                for (int u = 0; u < 16; ++u) {
                    const int i = base + u * (int)blockDim.x + tid;
                    tmp[u] = i < S.amp_len ? S.amp[i] : 0.0;
                }
#pragma unroll
                for (int u = 0; u < 16; ++u) {
                    const int i = base + u * (int)blockDim.x + tid;
                    if (i < S.amp_len) a_lds[i] = tmp[u];
                }
            }
        }
        for (int i = t; i < P; i += T) x[lx(i)] = mk2<C2>(0, 0);
        __syncthreads();
        for (int ci = ci0; ci < pi.w; ci += 2 * T) {
            const int4 ca = ci == ci0 ? ca0 : S.cells[ci];
            const int4 cb = ci == ci0 ? cb0 : (ci + T < pi.w ? S.cells[ci + T] : make_int4(0, 0, 0, 0));
            double va, vb;
            if (S.inl_ai && ci == ci0) {
                if (S.amp_lds) cell_values_inl<true>(S, a_lds, ca, cb, ia0, wa0, va, vb);
                else cell_values_inl<false>(S, a_lds, ca, cb, ia0, wa0, va, vb);
            } else if (S.amp_lds) {
                cell_values<true>(S, a_lds, ca, cb, va, vb);
            } else {
                cell_values<false>(S, a_lds, ca, cb, va, vb);
            }
            if (ci < pi.z) x[lx(ca.x)].x = (R)va;
            else x[lx(ca.x)].y = (R)va;
            if (ci + T < pi.w) {
                if (ci + T < pi.z) x[lx(cb.x)].x = (R)vb;
                else x[lx(cb.x)].y = (R)vb;
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < V; ++q) v[q] = x[lx(io_index(R0, q / R0, q % R0, t, T, P))];
    } else {
        // one loop per source kind (the kind is uniform over the launch), so
        // that all of a thread's loads are in flight together
        const int j0 = 2 * p;
        const bool two = j0 + 1 < nrows_limit;
        if (S.mode == ROW_DENSE && S.field32) {
            // FP32 values loaded first, widened once all are in flight
            const float* ra = S.field32 + (size_t)S.nx * j0;
            float fa[V], fb[V];
#pragma unroll
            for (int q = 0; q < V; ++q) {
                const int xpos = io_index(R0, q / R0, q % R0, t, T, P);
                fa[q] = 0.0f;
                fb[q] = 0.0f;
                if (live && xpos < S.nx) {
                    fa[q] = ra[xpos];
                    if (two) fb[q] = ra[xpos + S.nx];
                }
            }
#pragma unroll
            for (int q = 0; q < V; ++q) v[q] = mk2<C2>((R)fa[q], (R)fb[q]);
        } else if (S.mode == ROW_DENSE) {
            const double* ra = S.field + (size_t)S.nx * j0;
#pragma unroll
            for (int q = 0; q < V; ++q) {
                const int xpos = io_index(R0, q / R0, q % R0, t, T, P);
                double a = 0.0, b = 0.0;
                if (live && xpos < S.nx) {
                    a = ra[xpos];
                    if (two) b = ra[xpos + S.nx];
                }
                v[q] = mk2<C2>((R)a, (R)b);
            }
        } else {
            const int ja = j0 <= S.Py / 2 ? j0 : S.Py - j0;
            const int jb = (j0 + 1) <= S.Py / 2 ? j0 + 1 : S.Py - (j0 + 1);
#pragma unroll
            for (int q = 0; q < V; ++q) {
                const int xpos = io_index(R0, q / R0, q % R0, t, T, P);
                double a = 0.0, b = 0.0;
                if (live) {
                    const int ia = xpos <= S.Px / 2 ? xpos : S.Px - xpos;
                    a = S.tab[ia + (size_t)S.tmx * ja];
                    if (two) b = S.tab[ia + (size_t)S.tmx * jb];
                }
                v[q] = mk2<C2>((R)a, (R)b);
            }
        }
    }
    fft<P, -1, false>(v, x, t, tw);
    // natural-order spectrum to LDS, then split into the two half spectra
    __syncthreads();
#pragma unroll
    for (int q = 0; q < V; ++q) x[lx(io_index(RL, q / RL, q % RL, t, T, P))] = v[q];
    __syncthreads();
    if (!live) return;
#pragma unroll
    for (int i = 0; i < H / T; ++i) {
        const int k = t + i * T;
        const C2 z = x[lx(k)];
        Q4 ab;
        if (k == 0) {
            const C2 zn = x[lx(P / 2)];
            ab.x = z.x;
            ab.y = zn.x;
            ab.z = z.y;
            ab.w = zn.y;
        } else {
            const C2 zc = x[lx(P - k)];
            const R h = (R)0.5;
            ab.x = h * (z.x + zc.x);
            ab.y = h * (z.y - zc.y);
            ab.z = h * (z.y + zc.y);
            ab.w = -h * (z.x - zc.x);
        }
        *(Q4*)(Tout + tidx(2 * p, k, H)) = ab;
    }
}

// ------------------------------------------------------------- pass 2
// Transform = one spectrum column c < H of length P = P_y.  Tiled T1 row
// `slot` holds field row r where slot = slot_of_row[r] (-1: the row is zero),
// or slot = r when slot_of_row is null and r < nrows_in.  The LGF spectrum is
// stored for ky <= P/2 (even in ky) as spec[c][ky], (P/2+1) per column, for
// kx = c < H and kx = H (the Nyquist column, at c = H).
// MODE 0: convolve, write rows 0..ny-1 of T2 (TR-row tiles).  MODE 1: forward only, write
// the spectrum * scale (the LGF spectrum build).
// R: the forward transform and the spectrum product; RI: the inverse
// transform and T2 (RI = float after R = double: the split-precision
// convolution of the mixed-precision solve).
// Logical block blk: columns c0 = blk NB .. c0 + NB - 1.
template <int P, int NB, int MODE, class R, class RI, int TR = 2>
__device__ __forceinline__ void cols_body(const typename Cplx<R>::C2* __restrict__ T1,
                                          const int* __restrict__ slot_of_row,
                                          const typename Cplx<R>::C2* __restrict__ tw,
                                          const typename Cplx<RI>::C2* __restrict__ tw_inv, int H, int ny,
                                          int nrows_in, const R* __restrict__ spec, double scale,
                                          typename Cplx<RI>::C2* __restrict__ T2, double* __restrict__ spec_out,
                                          int blk, typename Cplx<R>::C2* lds, int tid)
{
    using C2 = typename Cplx<R>::C2;
    using C2I = typename Cplx<RI>::C2;
    constexpr int T = P / V, SH = P / 2 + 1;
    constexpr int R0 = R_first<P>(false), RL = R_last<P>(false), RO = R_last<P>(true);
    static_assert(R_first<P>(true) == RL, "the inverse starts from the forward's registers");
    const int cl = tid / T, t = tid - cl * T;
    const int c0 = blk * NB;
    const int c = c0 + cl;
    const bool live = c < H;
    C2* x = lds + (size_t)cl * lds_len(P);
    C2 v[V];
    // all slot lookups first, then all row loads: loads return in order, so
    // a lookup issued after a row load would wait for that row load too
    int slot[V];
#pragma unroll
    for (int q = 0; q < V; ++q) {
        const int row = io_index(R0, q / R0, q % R0, t, T, P);
        slot[q] = -1;
        if (live && row < ny) slot[q] = slot_of_row ? slot_of_row[row] : (row < nrows_in ? row : -1);
    }
#pragma unroll
    for (int q = 0; q < V; ++q) {
        C2 val = mk2<C2>(0, 0);
        if (slot[q] >= 0) val = T1[tidx(slot[q], c, H)];
        v[q] = val;
    }
    // MODE 0: the spectrum values this thread multiplies by, loaded before the
    // forward's last stage
    R g[V];
    auto pre = [&]() {
        if (MODE == 0 && live) {
#pragma unroll
            for (int q = 0; q < V; ++q) {
                const int ky = io_index(RL, q / RL, q % RL, t, T, P);
                g[q] = spec[(size_t)c * SH + (ky <= P / 2 ? ky : P - ky)];
            }
        }
    };
    fft<P, -1, false>(v, x, t, tw, pre);
    if constexpr (MODE == 1) {
        if (!live) return;
#pragma unroll
        for (int q = 0; q < V; ++q) {
            const int ky = io_index(RL, q / RL, q % RL, t, T, P);
            if (ky <= P / 2) {
                spec_out[(size_t)c * SH + ky] = v[q].x * scale;
                if (c == 0) spec_out[(size_t)H * SH + ky] = v[q].y * scale;
            }
        }
        return;
    } else {
        // multiply by the real spectrum; the inverse's first stage reads exactly
        // these registers.  Column 0 packs the DC column a (real part) and the
        // Nyquist column b (imaginary part): with Z = A + iB (A, B Hermitian),
        // the product g_a A + i g_b B = (g_a+g_b)/2 Z + (g_a-g_b)/2 conj(Z(-ky)).
        if (c0 == 0) {
            __syncthreads();
#pragma unroll
            for (int q = 0; q < V; ++q) x[lx(io_index(RL, q / RL, q % RL, t, T, P))] = v[q];
            __syncthreads();
        }
        if (live) {
#pragma unroll
            for (int q = 0; q < V; ++q) {
                if (c == 0) {
                    const int ky = io_index(RL, q / RL, q % RL, t, T, P);
                    const R gn = spec[(size_t)H * SH + (ky <= P / 2 ? ky : P - ky)];
                    const C2 zc = x[lx((P - ky) & (P - 1))];
                    const R a = (R)0.5 * (g[q] + gn), b = (R)0.5 * (g[q] - gn);
                    v[q] = mk2<C2>(a * v[q].x + b * zc.x, a * v[q].y - b * zc.y);
                } else {
                    v[q].x *= g[q];
                    v[q].y *= g[q];
                }
            }
        }
        C2I vi[V];
#pragma unroll
        for (int q = 0; q < V; ++q) vi[q] = mk2<C2I>((RI)v[q].x, (RI)v[q].y);
        fft<P, 1, true>(vi, (C2I*)x, t, tw_inv);
        if (!live) return;
        // TR = 4: T2 is read once, by the next pass: streaming stores keep it
        // out of the L2 write-back at the kernel's end
        typedef RI v2 __attribute__((ext_vector_type(2)));
#pragma unroll
        for (int q = 0; q < V; ++q) {
            const int row = io_index(RO, q / RO, q % RO, t, T, P);
            if (row < ny) {
                if constexpr (TR == 4) {
                    const v2 tv = {vi[q].x, vi[q].y};
                    __builtin_nontemporal_store(tv, (v2*)&T2[pidx<TR>(row, c, H)]);
                } else {
                    T2[pidx<TR>(row, c, H)] = vi[q];
                }
            }
        }
    }
}

// ------------------------------------------------------------- pass 3
// Transform c of logical block blk makes output rows 2p, 2p+1 (p = blk NB +
// c), FP64 or (a Float32 cache) rounded to FP32.
template <int P, int NB, class TO, class R, int TR = 2>
__device__ __forceinline__ void rows_inv_body(const typename Cplx<R>::C2* __restrict__ T2,
                                              const typename Cplx<R>::C2* __restrict__ tw, int nx, int ny,
                                              TO* __restrict__ out, int blk, typename Cplx<R>::C2* lds, int tid)
{
    using C2 = typename Cplx<R>::C2;
    using Q4 = typename Cplx<R>::Q4;
    constexpr int T = P / V, H = P / 2;
    constexpr int R0 = R_first<P>(false), RL = R_last<P>(false);
    const int c = tid / T, t = tid - c * T;
    const int p = blk * NB + c;
    const int j0 = 2 * p;
    const bool live = j0 < ny;
    const bool two = j0 + 1 < ny;
    C2* x = lds + (size_t)c * lds_len(P);
    C2 v[V];
    // all loads first (in flight together), then the unpacking
    Q4 abq[V];
#pragma unroll
    for (int q = 0; q < V; ++q) {
        const int m = io_index(R0, q / R0, q % R0, t, T, P);
        const int k = m <= P / 2 ? m : P - m;
        if (live) abq[q] = *(const Q4*)(T2 + pidx<TR>(j0, k == H ? 0 : k, H));
        else abq[q].x = abq[q].y = abq[q].z = abq[q].w = 0;
    }
#pragma unroll
    for (int q = 0; q < V; ++q) {
        const int m = io_index(R0, q / R0, q % R0, t, T, P);
        C2 z = mk2<C2>(0, 0);
        if (live) {
            // Z(m) = A(m) + i B(m),  Z(P-m) = conj(A(m)) + i conj(B(m)); column 0
            // holds the real DC (x) and Nyquist (y) values of rows j0, j0+1
            const int k = m <= P / 2 ? m : P - m;
            const Q4 ab = abq[q];
            R Ax = ab.x, Ay = ab.y, Bx = ab.z, By = ab.w;
            if (k == 0) { Ay = 0; By = 0; }
            if (k == H) { Ax = ab.y; Bx = ab.w; Ay = 0; By = 0; }
            if (!two) { Bx = 0; By = 0; }
            z = m <= P / 2 ? mk2<C2>(Ax - By, Ay + Bx) : mk2<C2>(Ax + By, Bx - Ay);
        }
        v[q] = z;
    }
    fft<P, 1, false>(v, x, t, tw);
    if (!live) return;
    // the field is the pass's result, read by no later pass of the
    // convolution: streaming stores (whole lines per wave) leave it out of the
    // L2 write-back at the kernel's end
#pragma unroll
    for (int q = 0; q < V; ++q) {
        const int xo = io_index(RL, q / RL, q % RL, t, T, P);
        if (xo < nx) {
            __builtin_nontemporal_store((TO)v[q].x, &out[xo + (size_t)nx * j0]);
            if (two) __builtin_nontemporal_store((TO)v[q].y, &out[xo + (size_t)nx * (j0 + 1)]);
        }
    }
}

}  // namespace ilh
