// ilh_fft.hip -- unbounded-domain inverse Laplacian by zero-padded FFT
// convolution with the 5-point lattice Green's function (LGF), FP64.
//
// inverse_laplacian!(w, cache) = L \ w = dx^2 (G * w) over the field's own
// index set (src/grid_operators.jl:144-179, LGF plan of src/cache.jl:321-324),
// G(0,0) = 0, L G = delta.  G is even, so a circular period P >= 2n-2 per axis
// reproduces the linear convolution exactly; P is a power of two >= 16.
//
// Transforms are Stockham radix-8 stages (plus one radix-2 or -4 stage when P
// is not a power of 8), with each thread holding V = 8 complex values in
// registers; LDS is only the exchange between stages.  Every kernel is
// instantiated per P, so stage radices, strides and twiddle offsets are
// compile-time constants; each stage's twiddles are loaded before the LDS
// exchange that precedes it.  Three passes per convolution:
//   1. row pass:    two real rows packed as one complex P_x-point FFT, split
//                   into the two half spectra, P_x/2 columns each: column 0
//                   holds (DC, Nyquist) as (re, im) -- both are real.  Rows
//                   come from a field, from the LGF table, or are generated
//                   on the fly from the surface-point stencil terms (no
//                   grid-sized source array is written).  Only row pairs that
//                   hold sources are transformed.
//   2. column pass: one spectrum column per transform: forward P_y-point FFT,
//                   multiply by the real LGF spectrum in registers
//                   (dx^2/(P_x P_y) folded in; the packed column 0 applies the
//                   DC and Nyquist spectra to its Hermitian parts), inverse
//                   FFT whose first stage consumes those registers and whose
//                   last stage writes the n_y output rows.  Consecutive
//                   columns run on the same XCD and share its L2.
//   3. row pass:    the two half spectra of a row pair packed into one complex
//                   inverse FFT; real/imaginary parts are the two output rows.
// The LGF table is built here by Gauss-Legendre quadrature of
//   G(m,n) = (1/2pi) int_0^pi [1 - cos(mt) e^{-n s(t)}] / sinh s(t) dt,
//   cosh s = 2 - cos t,  n = max(|m|,|n|).
#include "ilh_internal.hpp"

#include <algorithm>
#include <cmath>
#include <map>

#include "ilh_fft_dev.hpp"

namespace ilh {

__global__ void k_stage_tw(double2* tw, int off, int Ns, int R)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= Ns * (R - 1)) return;
    const int r = e / Ns + 1, k = e % Ns;   // entry (r - 1) Ns + k: a stage's loads of one r are contiguous over k
    double s, c;
    sincospi(2.0 * (double)((long long)r * k) / (double)(Ns * R), &s, &c);
    tw[off + e] = make_double2(c, -s);
}

// Set while an allocation-free caller launches: tables are then looked up
// only, and a missing one fails the call instead of being built.
static thread_local bool t_lookup_only = false;
struct LookupOnly {
    bool prev;
    LookupOnly() : prev(t_lookup_only) { t_lookup_only = true; }
    ~LookupOnly() { t_lookup_only = prev; }
};

static const double2* get_twiddles(ilh_ctx* ctx, int P, bool build = true)
{
    build = build && !t_lookup_only;
    // one table per (device, P); immutable once built
    static std::map<std::pair<int, int>, DevBuf*> cache;
    auto key = std::make_pair(ctx->device, P);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second->as<double2>();
    if (!build) fail(ILH_ERR_STATE, "FFT twiddle table of this period is not built yet");
    if (P < 16 || P > MAXP || (P & (P - 1))) fail(ILH_ERR_ARG, "FFT length must be a power of two in [16, 8192]");
    DevBuf* b = new DevBuf();
    b->ensure(sizeof(double2) * (size_t)(tw_count(P, false) + tw_count(P, true)));
    for (bool rev : {false, true})
        for (int s = 0; s < n_stages(P); ++s) {
            const int Ns = nsub(P, rev, s), R = radix(P, rev, s);
            const int n = Ns * (R - 1);
            k_stage_tw<<<(n + 255) / 256, 256, 0, ctx->stream>>>(b->as<double2>(), tw_off(P, rev, s), Ns, R);
        }
    ILH_HIP(hipGetLastError());
    ILH_HIP(hipStreamSynchronize(ctx->stream));
    cache[key] = b;
    return b->as<double2>();
}

// The FP32 transforms' twiddles: the FP64 table rounded once.
__global__ void k_tw_to_f32(const double2* __restrict__ a, float2* __restrict__ b, int n)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) b[e] = make_float2((float)a[e].x, (float)a[e].y);
}

static const float2* get_twiddles_f32(ilh_ctx* ctx, int P, bool build = true)
{
    build = build && !t_lookup_only;
    static std::map<std::pair<int, int>, DevBuf*> cache;
    auto key = std::make_pair(ctx->device, P);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second->as<float2>();
    if (!build) fail(ILH_ERR_STATE, "FFT twiddle table of this period is not built yet");
    const double2* d = get_twiddles(ctx, P);
    const int n = tw_count(P, false) + tw_count(P, true);
    DevBuf* b = new DevBuf();
    b->ensure(sizeof(float2) * (size_t)n);
    k_tw_to_f32<<<(n + 255) / 256, 256, 0, ctx->stream>>>(d, b->as<float2>(), n);
    ILH_HIP(hipGetLastError());
    ILH_HIP(hipStreamSynchronize(ctx->stream));
    cache[key] = b;
    return b->as<float2>();
}

template <class R> static const typename Cplx<R>::C2* twiddles(ilh_ctx* ctx, int P);
template <> const double2* twiddles<double>(ilh_ctx* ctx, int P) { return get_twiddles(ctx, P); }
template <> const float2* twiddles<float>(ilh_ctx* ctx, int P) { return get_twiddles_f32(ctx, P); }
const double2* fft_twiddles(ilh_ctx* ctx, int P) { return get_twiddles(ctx, P); }

// ------------------------------------------------------------- kernels
// One problem per launch; grid.y = right-hand sides of a batch, each offset
// by its stride.
template <int P, class R>
__global__ void __launch_bounds__(tpb_of(P)) k_rows_fwd(RowSrc S, const typename Cplx<R>::C2* __restrict__ tw,
                                                        int npairs, int nrows_limit,
                                                        typename Cplx<R>::C2* __restrict__ Tout, size_t amp_stride,
                                                        size_t out_stride)
{
    using C2 = typename Cplx<R>::C2;
    extern __shared__ double2 lds_raw[];
    if (S.mode == ROW_DENSE) {
        if (S.field32) S.field32 += amp_stride * blockIdx.y;
        else S.field += amp_stride * blockIdx.y;
    } else if (S.mode == ROW_TERMS) {
        S.amp += amp_stride * blockIdx.y;
    }
    Tout += out_stride * blockIdx.y;
    constexpr int NB = tpb_of(P) / (P / V);
    // live pairs from the device when the launch covers an upper bound: the
    // first ceil(npairs / NB) blocks work (spread over the XCDs), the rest exit
    int nblk = gridDim.x;
    if (S.npairs_dev) {
        npairs = *S.npairs_dev;
        nblk = (npairs + NB - 1) / NB;
        if ((int)blockIdx.x >= nblk) return;
    }
    rows_fwd_body<P, NB, R>(S, tw, npairs, nrows_limit, Tout, xcd_block_of(blockIdx.x, nblk), (C2*)lds_raw,
                            threadIdx.x);
}

template <int P, int MODE, class R, class RI, int TR>
__global__ void __launch_bounds__(tpb_of(P)) __attribute__((amdgpu_waves_per_eu(4)))
k_cols(const typename Cplx<R>::C2* __restrict__ T1, const int* __restrict__ slot_of_row,
       const typename Cplx<R>::C2* __restrict__ tw, const typename Cplx<RI>::C2* __restrict__ tw_inv, int H, int ny,
       int nrows_in, const R* __restrict__ spec, double scale, typename Cplx<RI>::C2* __restrict__ T2,
       double* __restrict__ spec_out, size_t t1_stride, size_t t2_stride)
{
    using C2 = typename Cplx<R>::C2;
    extern __shared__ double2 lds_raw[];
    T1 += t1_stride * blockIdx.y;
    T2 += t2_stride * blockIdx.y;
    cols_body<P, tpb_of(P) / (P / V), MODE, R, RI, TR>(T1, slot_of_row, tw, tw_inv, H, ny, nrows_in, spec, scale, T2,
                                                    spec_out, xcd_block(), (C2*)lds_raw, threadIdx.x);
}

template <int P, class TO, class R, int TR>
__global__ void __launch_bounds__(tpb_of(P)) k_rows_inv(const typename Cplx<R>::C2* __restrict__ T2,
                                                        const typename Cplx<R>::C2* __restrict__ tw, int nx, int ny,
                                                        TO* __restrict__ out, size_t t2_stride, size_t out_stride)
{
    using C2 = typename Cplx<R>::C2;
    extern __shared__ double2 lds_raw[];
    T2 += t2_stride * blockIdx.y;
    out += out_stride * blockIdx.y;
    rows_inv_body<P, tpb_of(P) / (P / V), TO, R, TR>(T2, tw, nx, ny, out, xcd_block(), (C2*)lds_raw, threadIdx.x);
}

// ------------------------------------------------------------- launches
static void set_lds(const void* fn, size_t bytes)
{
    static std::map<const void*, size_t> lim;
    auto it = lim.find(fn);
    if (it != lim.end() && it->second >= bytes) return;
    ILH_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    lim[fn] = bytes;
}

#define ILH_FOR_P(P, BODY)                                                                               \
    switch (P) {                                                                                         \
    case 16: { constexpr int PP = 16; BODY; } break;                                                    \
    case 32: { constexpr int PP = 32; BODY; } break;                                                    \
    case 64: { constexpr int PP = 64; BODY; } break;                                                    \
    case 128: { constexpr int PP = 128; BODY; } break;                                                  \
    case 256: { constexpr int PP = 256; BODY; } break;                                                  \
    case 512: { constexpr int PP = 512; BODY; } break;                                                  \
    case 1024: { constexpr int PP = 1024; BODY; } break;                                                \
    case 2048: { constexpr int PP = 2048; BODY; } break;                                                \
    case 4096: { constexpr int PP = 4096; BODY; } break;                                                \
    case 8192: { constexpr int PP = 8192; BODY; } break;                                                \
    default: fail(ILH_ERR_ARG, "FFT length must be a power of two in [16, 8192]");                       \
    }

// Batched launches (nb right-hand sides, grid.y): operand k starts `stride`
// elements after operand k-1.
struct Batch {
    int nb = 1;
    size_t in = 0, out = 0;
};

template <int P, class C2>
static void launch_rows_fwd_p(ilh_ctx* ctx, RowSrc S, int npairs, int nrows_limit, C2* Tout, Batch bt)
{
    using R = decltype(C2().x);
    constexpr int NB = tpb_of(P) / (P / V);
    S.amp_lds = S.mode == ROW_TERMS && S.amp_len <= amp_lds_cap(P);
    const size_t lds = sizeof(C2) * (size_t)NB * lds_len(P) + (S.amp_lds ? sizeof(double) * S.amp_len : 0);
    set_lds((const void*)k_rows_fwd<P, R>, lds);
    k_rows_fwd<P, R><<<dim3((npairs + NB - 1) / NB, bt.nb), tpb_of(P), lds, ctx->stream>>>(
        S, twiddles<R>(ctx, P), npairs, nrows_limit, Tout, bt.in, bt.out);
    ILH_HIP(hipGetLastError());
}

template <class C2>
static void launch_rows_fwd(ilh_ctx* ctx, const RowSrc& S, int P, int npairs, int nrows_limit, C2* Tout,
                            Batch bt = Batch())
{
    if (npairs <= 0) return;
    ILH_FOR_P(P, (launch_rows_fwd_p<PP, C2>(ctx, S, npairs, nrows_limit, Tout, bt)));
}

template <int P, int MODE, class C2, class C2I>
static void launch_cols_p(ilh_ctx* ctx, const C2* T1, const int* slot_of_row, int H, int ny, int nrows_in,
                          const decltype(C2().x)* spec, double scale, C2I* T2, double* spec_out, Batch bt)
{
    using R = decltype(C2().x);
    using RI = decltype(C2I().x);
    constexpr int NB = tpb_of(P) / (P / V);
    const size_t lds = sizeof(C2) * (size_t)NB * lds_len(P);
    // one data set: 4-row T2 tiles with streaming stores; several: 2-row tiles
    // (the inverse row pass reads the same layout: launch_rows_inv)
    auto kern = bt.nb == 1 ? k_cols<P, MODE, R, RI, 4> : k_cols<P, MODE, R, RI, 2>;
    set_lds((const void*)kern, lds);
    kern<<<dim3((H + NB - 1) / NB, bt.nb), tpb_of(P), lds, ctx->stream>>>(
        T1, slot_of_row, twiddles<R>(ctx, P), twiddles<RI>(ctx, P), H, ny, nrows_in, spec, scale, T2, spec_out,
        bt.in, bt.out);
    ILH_HIP(hipGetLastError());
}

// T2 may hold a narrower type than T1 (C2I = float2 after C2 = double2)
template <int MODE, class C2, class C2I>
static void launch_cols(ilh_ctx* ctx, const C2* T1, const int* slot_of_row, int Px, int Py, int ny, int nrows_in,
                        const decltype(C2().x)* spec, double scale, C2I* T2, double* spec_out, Batch bt = Batch())
{
    const int H = Px / 2;
    ILH_FOR_P(Py, (launch_cols_p<PP, MODE, C2, C2I>(ctx, T1, slot_of_row, H, ny, nrows_in, spec, scale, T2,
                                                     spec_out, bt)));
}

template <int P, class TO, class C2>
static void launch_rows_inv_p(ilh_ctx* ctx, const C2* T2, int nx, int ny, TO* out, Batch bt)
{
    using R = decltype(C2().x);
    constexpr int NB = tpb_of(P) / (P / V);
    const int npairs = (ny + 1) / 2;
    const size_t lds = sizeof(C2) * (size_t)NB * lds_len(P);
    auto kern = bt.nb == 1 ? k_rows_inv<P, TO, R, 4> : k_rows_inv<P, TO, R, 2>;   // launch_cols' T2 layout
    set_lds((const void*)kern, lds);
    kern<<<dim3((npairs + NB - 1) / NB, bt.nb), tpb_of(P), lds, ctx->stream>>>(
        T2, twiddles<R>(ctx, P), nx, ny, out, bt.in, bt.out);
    ILH_HIP(hipGetLastError());
}

template <class TO, class C2>
static void launch_rows_inv(ilh_ctx* ctx, const C2* T2, int Px, int nx, int ny, TO* out, Batch bt = Batch())
{
    ILH_FOR_P(Px, (launch_rows_inv_p<PP, TO, C2>(ctx, T2, nx, ny, out, bt)));
}

// -------------------------------------------------------------- LGF table
__constant__ double c_glx[20];
__constant__ double c_glw[20];

__device__ __forceinline__ void lgf_sig(double t, double& sig, double& sh)
{
    const double h = sin(0.5 * t);
    const double u = 2.0 * h * h;
    sh = sqrt(u * (2.0 + u));
    sig = log1p(u + sh);
}

// Work-group per n (the larger index): G(m, n) for 0 <= m <= n, mirrored.
__global__ void __launch_bounds__(256) k_lgf(int Tn, double* __restrict__ tab)
{
    const int n = blockIdx.x;
    const double pi = 3.14159265358979323846;
    if (n == 0) {
        if (threadIdx.x == 0) tab[0] = 0.0;
        return;
    }
    __shared__ double st[20 * 16], sw[20 * 16];
    __shared__ double red[256];
    // A(n) = (1/2pi) int_0^pi (1 - e^{-n s}) / sinh s, panels [0,1/n],[1/n,2/n],[2/n,4/n],...
    double part = 0.0;
    {
        int npan = 1;
        double e = 1.0 / n;
        while (e < pi) { e *= 2.0; ++npan; }
        for (int q = threadIdx.x; q < npan * 20; q += 256) {
            const int pnl = q / 20, k = q % 20;
            double a = pnl == 0 ? 0.0 : ldexp(1.0 / n, pnl - 1);
            double b = ldexp(1.0 / n, pnl);
            if (b > pi) b = pi;
            const double t = 0.5 * (a + b) + 0.5 * (b - a) * c_glx[k];
            double sig, sh;
            lgf_sig(t, sig, sh);
            part += 0.5 * (b - a) * c_glw[k] * (-expm1(-n * sig)) / sh;
        }
    }
    red[threadIdx.x] = part;
    // B nodes: T = min(pi, 50/n), panels of width <= 4/n
    double T = 50.0 / n;
    if (T > pi) T = pi;
    int np = (int)ceil(T / (4.0 / n) - 1e-12);
    if (np < 1) np = 1;
    const double h = T / np;
    const int K = np * 20;
    for (int q = threadIdx.x; q < K; q += 256) {
        const int pnl = q / 20, k = q % 20;
        const double t = (pnl + 0.5) * h + 0.5 * h * c_glx[k];
        double sig, sh;
        lgf_sig(t, sig, sh);
        st[q] = t;
        sw[q] = 0.5 * h * c_glw[k] * exp(-n * sig) / sh;
    }
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    const double A = red[0] / (2.0 * pi);
    for (int m = threadIdx.x; m <= n; m += 256) {
        double B = 0.0;
        for (int q = 0; q < K; ++q) {
            const double s = sin(0.5 * m * st[q]);
            B += sw[q] * 2.0 * s * s;
        }
        const double G = A + B / (2.0 * pi);
        if (m < Tn && n < Tn) {
            tab[m + (size_t)Tn * n] = G;
            tab[n + (size_t)Tn * m] = G;
        }
    }
}

static void gauss_legendre20(double* x, double* w)
{
    const int K = 20;
    for (int i = 0; i < K; ++i) {
        double z = cos(M_PI * (i + 0.75) / (K + 0.5));
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = z;
            for (int k = 2; k <= K; ++k) { double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k; p0 = p1; p1 = p2; }
            const double dp = K * (z * p1 - p0) / (z * z - 1.0);
            const double d = p1 / dp;
            z -= d;
            if (fabs(d) < 1e-17) break;
        }
        double p0 = 1.0, p1 = z;
        for (int k = 2; k <= K; ++k) { double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k; p0 = p1; p1 = p2; }
        const double dp = K * (z * p1 - p0) / (z * z - 1.0);
        x[i] = z;
        w[i] = 2.0 / ((1.0 - z * z) * dp * dp);
    }
}

void ensure_lgf(ilh_plan* pl)
{
    const Grid& g = pl->g;
    int need = std::max(g.NX, g.NY) + 8;
    for (int gk = 0; gk < 2; ++gk) {
        int n = std::max(g.nx(gk), g.ny(gk));
        int P = 16;
        while (P < 2 * n - 2) P <<= 1;
        need = std::max(need, P / 2 + 1);
    }
    if (pl->lgf_mx >= need) return;
    static bool gl_set = false;
    if (!gl_set) {
        double x[20], w[20];
        gauss_legendre20(x, w);
        ILH_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_glx), x, sizeof(x)));
        ILH_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_glw), w, sizeof(w)));
        gl_set = true;
    }
    pl->lgf.ensure(sizeof(double) * (size_t)need * need);
    k_lgf<<<need, 256, 0, pl->ctx->stream>>>(need, pl->lgf.as<double>());
    ILH_HIP(hipGetLastError());
    pl->lgf_mx = pl->lgf_my = need;
}


// -------------------------------------------------------------- plans
static int period(int n)
{
    int P = 16;
    while (P < 2 * n - 2) P <<= 1;
    return P;
}

ConvPlan& get_conv(ilh_plan* pl, int gk)
{
    ConvPlan& C = pl->conv[gk];
    if (pl->conv_ok[gk]) return C;
    ilh_ctx* ctx = pl->ctx;
    ensure_lgf(pl);
    const Grid& g = pl->g;
    C.gk = gk;
    C.nx = g.nx(gk);
    C.ny = g.ny(gk);
    C.Px = period(C.nx);
    C.Py = period(C.ny);
    if (C.Px > MAXP || C.Py > MAXP) fail(ILH_ERR_ARG, "grid too large for the FFT convolution (period > 8192)");
    C.H = C.Px / 2;
    const int rows = (std::max(C.Py, C.ny + 1) + 7) / 8 * 8;   // tiled layout: whole 8-row tiles
    C.T1.ensure(sizeof(double2) * (size_t)rows * C.H);
    C.T2.ensure(sizeof(double2) * (size_t)((C.ny + 1 + 7) / 8 * 8) * C.H);
    C.spec.ensure(sizeof(double) * (size_t)(C.H + 1) * (C.Py / 2 + 1));
    // spectrum of the circular kernel: rows of G_circ, then forward columns
    RowSrc S{};
    S.mode = ROW_LGF;
    S.tab = pl->lgf.as<double>();
    S.tmx = pl->lgf_mx;
    S.Px = C.Px;
    S.Py = C.Py;
    launch_rows_fwd(ctx, S, C.Px, C.Py / 2, C.Py, C.T1.as<double2>());
    const double scale = g.dx * g.dx / ((double)C.Px * C.Py);
    launch_cols<1, double2, double2>(ctx, C.T1.as<double2>(), nullptr, C.Px, C.Py, C.Py, C.Py, nullptr, scale,
                                     (double2*)nullptr, C.spec.as<double>());
    pl->conv_ok[gk] = true;
    return C;
}

void conv_dense(ilh_plan* pl, int gk, const double* in, double* out)
{
    ConvPlan& C = get_conv(pl, gk);
    RowSrc S{};
    S.mode = ROW_DENSE;
    S.field = in;
    S.nx = C.nx;
    S.ny = C.ny;
    launch_rows_fwd(pl->ctx, S, C.Px, (C.ny + 1) / 2, C.ny, C.T1.as<double2>());
    launch_cols<0>(pl->ctx, C.T1.as<double2>(), nullptr, C.Px, C.Py, C.ny, C.ny, C.spec.as<double>(), 1.0,
                   C.T2.as<double2>(), nullptr);
    launch_rows_inv(pl->ctx, C.T2.as<double2>(), C.Px, C.nx, C.ny, out);
}

__global__ void k_d2f(const double* __restrict__ a, float* __restrict__ b, size_t n)
{
    const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (e < n) b[e] = (float)a[e];
}

// the LGF spectrum rounded to FP32 once (the FP32 transforms multiply by it)
static const float* spec_f32(ilh_plan* pl, ConvPlan& C)
{
    if (!C.spec32_ok) {
        const size_t n = (size_t)(C.H + 1) * (C.Py / 2 + 1);
        C.spec32.ensure(sizeof(float) * n);
        k_d2f<<<(unsigned)((n + 255) / 256), 256, 0, pl->ctx->stream>>>(C.spec.as<double>(), C.spec32.as<float>(), n);
        ILH_HIP(hipGetLastError());
        C.spec32_ok = true;
    }
    return C.spec32.as<float>();
}

// FP32 field -> FP32 field (a Float32 cache): the three passes in FP32
// (FP32 spectra intermediates, spectrum, twiddles and LDS exchanges), as an
// FP32 FFT plan of the same convolution
void conv_dense(ilh_plan* pl, int gk, const float* in, float* out)
{
    ConvPlan& C = get_conv(pl, gk);
    RowSrc S{};
    S.mode = ROW_DENSE;
    S.field32 = in;
    S.nx = C.nx;
    S.ny = C.ny;
    const float* sp = spec_f32(pl, C);
    launch_rows_fwd(pl->ctx, S, C.Px, (C.ny + 1) / 2, C.ny, C.T1.as<float2>());
    launch_cols<0>(pl->ctx, C.T1.as<float2>(), nullptr, C.Px, C.Py, C.ny, C.ny, sp, 1.0, C.T2.as<float2>(), nullptr);
    launch_rows_inv(pl->ctx, (const float2*)C.T2.as<float2>(), C.Px, C.nx, C.ny, out);
}

void conv_dense_multi(ilh_plan* pl, int gk, const double* in, int nb, double* out)
{
    if (nb <= 0) return;
    ConvPlan& C = get_conv(pl, gk);
    const size_t n = (size_t)C.nx * C.ny;
    const size_t s1 = C.T1.bytes / sizeof(double2), s2 = C.T2.bytes / sizeof(double2);
    C.T1m.ensure(C.T1.bytes * (size_t)nb);
    C.T2m.ensure(C.T2.bytes * (size_t)nb);
    RowSrc S{};
    S.mode = ROW_DENSE;
    S.field = in;
    S.nx = C.nx;
    S.ny = C.ny;
    launch_rows_fwd(pl->ctx, S, C.Px, (C.ny + 1) / 2, C.ny, C.T1m.as<double2>(), Batch{nb, n, s1});
    launch_cols<0>(pl->ctx, C.T1m.as<double2>(), nullptr, C.Px, C.Py, C.ny, C.ny, C.spec.as<double>(), 1.0,
                   C.T2m.as<double2>(), nullptr, Batch{nb, s1, s2});
    launch_rows_inv(pl->ctx, C.T2m.as<double2>(), C.Px, C.nx, C.ny, out, Batch{nb, s2, n});
}

static RowSrc terms_source(const ConvPlan& C, const SourceSet& src, const double* amp)
{
    RowSrc S{};
    S.mode = ROW_TERMS;
    S.nx = C.nx;
    S.ny = C.ny;
    S.pinfo = src.pinfo;
    S.cells = src.bins.cells;
    S.cai = src.bins.cai;
    S.cw = src.bins.cw;
    S.amp = amp;
    S.amp_len = src.amp_len;
    S.npairs_dev = src.npairs_dev;
    if (src.inl_stride > 0 && !src.npairs_dev) {
        S.inl_ai = src.inl_ai;
        S.inl_w = src.inl_w;
        S.inl_stride = src.inl_stride;
    }
    return S;
}

// one pass of conv_sources alone (0 = rows, 1 = columns, 2 = inverse rows), for timing
void conv_stage(ilh_plan* pl, int gk, const SourceSet& src, const double* amp, double* out, int stage)
{
    ConvPlan& C = get_conv(pl, gk);
    if (stage == 0) {
        launch_rows_fwd(pl->ctx, terms_source(C, src, amp), C.Px, src.npairs, C.ny, C.T1.as<double2>());
    } else if (stage == 1) {
        launch_cols<0>(pl->ctx, C.T1.as<double2>(), src.slot_of_row, C.Px, C.Py, C.ny, 0, C.spec.as<double>(), 1.0,
                       C.T2.as<double2>(), nullptr);
    } else {
        launch_rows_inv(pl->ctx, C.T2.as<double2>(), C.Px, C.nx, C.ny, out);
    }
}

// nb right-hand sides at once: amplitudes amp + k amp_stride, outputs
// out + k n(gk); the spectra in T1m / T2m (nb copies of T1 / T2, allocated here)
void conv_sources_multi(ilh_plan* pl, int gk, const SourceSet& src, const double* amp, size_t amp_stride, int nb,
                        double* out)
{
    ConvPlan& C = get_conv(pl, gk);
    const size_t s1 = C.T1.bytes / sizeof(double2), s2 = C.T2.bytes / sizeof(double2);
    C.T1m.ensure(C.T1.bytes * (size_t)nb);
    C.T2m.ensure(C.T2.bytes * (size_t)nb);
    launch_rows_fwd(pl->ctx, terms_source(C, src, amp), C.Px, src.npairs, C.ny, C.T1m.as<double2>(),
                    Batch{nb, amp_stride, s1});
    launch_cols<0>(pl->ctx, C.T1m.as<double2>(), src.slot_of_row, C.Px, C.Py, C.ny, 0, C.spec.as<double>(), 1.0,
                   C.T2m.as<double2>(), nullptr, Batch{nb, s1, s2});
    launch_rows_inv(pl->ctx, C.T2m.as<double2>(), C.Px, C.nx, C.ny, out,
                    Batch{nb, s2, (size_t)C.nx * C.ny});
}

// conv_sources with the two spectra in caller-supplied device buffers (T1:
// at least conv_ext_rows1(src) rows, T2: conv_ext_rows2 rows, H complex
// values per row, 8-row tiles) and no allocation: the plan's spectrum, its
// pinned source pairs and the twiddle table must exist (ILH_ERR_STATE).
int conv_ext_rows1(const SourceSet& src) { return (2 * src.npairs + 7) / 8 * 8; }
int conv_ext_rows2(const ConvPlan& C) { return (C.ny + 1 + 7) / 8 * 8; }

void conv_sources_ext(ilh_plan* pl, int gk, const SourceSet& src, const double* amp, double* out, double2* T1,
                      double2* T2)
{
    if (!pl->conv_ok[gk]) fail(ILH_ERR_STATE, "the plan's LGF spectrum is not built yet");
    if (src.npairs_dev) fail(ILH_ERR_STATE, "the source pairs are not pinned");
    LookupOnly guard;
    const ConvPlan& C = pl->conv[gk];
    launch_rows_fwd(pl->ctx, terms_source(C, src, amp), C.Px, src.npairs, C.ny, T1);
    launch_cols<0>(pl->ctx, T1, src.slot_of_row, C.Px, C.Py, C.ny, 0, C.spec.as<double>(), 1.0, T2, nullptr);
    launch_rows_inv(pl->ctx, (const double2*)T2, C.Px, C.nx, C.ny, out);
}

// conv_sources in split precision: the forward transforms and the product
// with the spectrum in FP64, the inverse column transform, the spectrum it
// writes and the inverse row transform in FP32.  (The surface sources have
// large high-wavenumber content -- the double layer's differences -- which an
// FP32 forward transform smears into the small low-wavenumber coefficients;
// after the product the spectrum is smooth and FP32 keeps ~1e-7.)
void conv_sources_f32(ilh_plan* pl, int gk, const SourceSet& src, const double* amp, double* out)
{
    ConvPlan& C = get_conv(pl, gk);
    launch_rows_fwd(pl->ctx, terms_source(C, src, amp), C.Px, src.npairs, C.ny, C.T1.as<double2>());
    launch_cols<0>(pl->ctx, (const double2*)C.T1.as<double2>(), src.slot_of_row, C.Px, C.Py, C.ny, 0,
                   C.spec.as<double>(), 1.0, C.T2.as<float2>(), nullptr);
    launch_rows_inv(pl->ctx, (const float2*)C.T2.as<float2>(), C.Px, C.nx, C.ny, out);
}

void conv_sources(ilh_plan* pl, int gk, const SourceSet& src, const double* amp, double* out, hipEvent_t* ev)
{
    ConvPlan& C = get_conv(pl, gk);
    hipStream_t st = pl->ctx->stream;
    launch_rows_fwd(pl->ctx, terms_source(C, src, amp), C.Px, src.npairs, C.ny, C.T1.as<double2>());
    if (ev) ILH_HIP(hipEventRecord(ev[0], st));
    launch_cols<0>(pl->ctx, C.T1.as<double2>(), src.slot_of_row, C.Px, C.Py, C.ny, 0, C.spec.as<double>(), 1.0,
                   C.T2.as<double2>(), nullptr);
    if (ev) ILH_HIP(hipEventRecord(ev[1], st));
    launch_rows_inv(pl->ctx, C.T2.as<double2>(), C.Px, C.nx, C.ny, out);
}

}  // namespace ilh
