// ilh_fft.hip -- unbounded-domain inverse Laplacian by zero-padded FFT
// convolution with the 5-point lattice Green's function (LGF), FP64.
//
// inverse_laplacian!(w, cache) = L \ w = dx^2 (G * w) over the field's own
// index set (src/grid_operators.jl:144-179, LGF plan of src/cache.jl:321-324),
// G(0,0) = 0, L G = delta.  G is even, so a circular period P >= 2n-2 per axis
// reproduces the linear convolution exactly; P is the next power of two.
//
// Three passes, all with hand-written Stockham FFTs staged in LDS:
//   1. row pass:    two real rows packed as one complex P_x-point FFT, split
//                   into the two half spectra (P_x/2+1 columns).  The rows are
//                   read from a field, generated from the LGF table, or built
//                   on the fly from surface-point stencil terms (no grid-sized
//                   source array is ever written).  Only row pairs that hold
//                   sources are transformed.
//   2. column pass: W adjacent spectrum columns per work-group; forward
//                   P_y-point FFT, multiply by the real LGF spectrum (dx^2 and
//                   1/(P_x P_y) folded in), inverse FFT, keep the n_y rows.
//                   Columns 0 and P_x/2 are real sequences and share one slot.
//   3. row pass:    two half spectra packed into one complex inverse FFT,
//                   real/imaginary parts are the two output rows.
// The LGF table itself is built here by Gauss-Legendre quadrature of
//   G(m,n) = (1/2pi) int_0^pi [1 - cos(mt) e^{-n s(t)}] / sinh s(t) dt,
//   cosh s = 2 - cos t,  n = max(|m|,|n|).
#include "ilh_internal.hpp"

#include <algorithm>
#include <cmath>

namespace ilh {

static constexpr int TWN = 8192;         // twiddle table length = largest period
static constexpr int NT = 256;           // threads per FFT work-group

// ------------------------------------------------------------- twiddles
__global__ void k_twiddle(double2* tw)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= TWN) return;
    double s, c;
    sincospi(2.0 * (double)k / TWN, &s, &c);
    tw[k] = make_double2(c, -s);         // exp(-2 pi i k / TWN)
}

void ensure_twiddles(ilh_ctx* ctx)
{
    if (ctx->tw_len == TWN) return;
    ctx->twiddle.ensure(sizeof(double2) * TWN);
    k_twiddle<<<TWN / 256, 256, 0, ctx->stream>>>(ctx->twiddle.as<double2>());
    ILH_HIP(hipGetLastError());
    ctx->tw_len = TWN;
}

// ------------------------------------------------------------ LDS FFT
__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
// multiply by s*i (s = +-1)
__device__ __forceinline__ double2 mul_si(double2 a, int s) { return s > 0 ? make_double2(-a.y, a.x) : make_double2(a.y, -a.x); }

template <int R>
__device__ __forceinline__ void dft(double2* a, int s);

template <>
__device__ __forceinline__ void dft<2>(double2* a, int)
{
    double2 t = a[0];
    a[0] = cadd(t, a[1]);
    a[1] = csub(t, a[1]);
}

template <>
__device__ __forceinline__ void dft<4>(double2* a, int s)
{
    double2 t0 = cadd(a[0], a[2]), t1 = csub(a[0], a[2]);
    double2 t2 = cadd(a[1], a[3]), t3 = mul_si(csub(a[1], a[3]), s);
    a[0] = cadd(t0, t2);
    a[2] = csub(t0, t2);
    a[1] = cadd(t1, t3);
    a[3] = csub(t1, t3);
}

template <>
__device__ __forceinline__ void dft<8>(double2* a, int s)
{
    const double c = 0.70710678118654752440084436210485;
    double2 e[4] = {a[0], a[2], a[4], a[6]};
    double2 o[4] = {a[1], a[3], a[5], a[7]};
    dft<4>(e, s);
    dft<4>(o, s);
    // W8^k, W8 = exp(s 2 pi i / 8)
    double2 o1 = cmul(o[1], make_double2(c, s * c));
    double2 o2 = mul_si(o[2], s);
    double2 o3 = cmul(o[3], make_double2(-c, s * c));
    a[0] = cadd(e[0], o[0]);
    a[4] = csub(e[0], o[0]);
    a[1] = cadd(e[1], o1);
    a[5] = csub(e[1], o1);
    a[2] = cadd(e[2], o2);
    a[6] = csub(e[2], o2);
    a[3] = cadd(e[3], o3);
    a[7] = csub(e[3], o3);
}

// One Stockham radix-R stage over `nb` contiguous transforms of length P held
// in LDS (in place: all inputs are read into registers before any output is
// written).  VALS = complex values per thread = nb*P/NT.
template <int R, int VALS>
__device__ __forceinline__ void stage(double2* x, int P, int nb, int Ns, const double2* __restrict__ tw, int s)
{
    constexpr int IT = VALS / R;
    const int PR = P / R, nbf = nb * PR;
    double2 v[VALS];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int b = threadIdx.x + it * NT;
        if (b < nbf) {
            const int c = b / PR, j = b - c * PR;
            const double2* xc = x + (size_t)c * P;
#pragma unroll
            for (int r = 0; r < R; ++r) v[it * R + r] = xc[j + r * PR];
        }
    }
    __syncthreads();
    const int tstep = TWN / (Ns * R);
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int b = threadIdx.x + it * NT;
        if (b < nbf) {
            const int c = b / PR, j = b - c * PR;
            const int k = j % Ns;
#pragma unroll
            for (int r = 1; r < R; ++r) {
                double2 w = tw[r * k * tstep];
                if (s > 0) w.y = -w.y;
                v[it * R + r] = cmul(v[it * R + r], w);
            }
            dft<R>(v + it * R, s);
            double2* xc = x + (size_t)c * P;
            const int base = (j / Ns) * Ns * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) xc[base + r * Ns] = v[it * R + r];
        }
    }
    __syncthreads();
}

// nb transforms of length P (power of two, 2 <= P <= TWN) in LDS, sign s.
template <int VALS>
__device__ void lds_fft(double2* x, int P, int nb, const double2* __restrict__ tw, int s)
{
    int Ns = 1;
    while (Ns * 8 <= P) {
        stage<8, VALS>(x, P, nb, Ns, tw, s);
        Ns *= 8;
    }
    if (Ns * 4 == P) stage<4, VALS>(x, P, nb, Ns, tw, s);
    else if (Ns * 2 == P) stage<2, VALS>(x, P, nb, Ns, tw, s);
}

// ------------------------------------------------------------ row sources
enum RowMode { ROW_DENSE = 0, ROW_LGF = 1, ROW_TERMS = 2 };

struct RowSrc {
    int mode;
    // dense field
    const double* field;
    int nx, ny;
    // LGF circular rows
    const double* tab;
    int tmx, Px, Py;
    // point terms
    const Term* terms;
    const int* ptr;
    const int* ent;
    const int* ebx;
    const double* amp;
    const int* pairs;      // compacted row pairs (ROW_TERMS), else null
};

__device__ __forceinline__ double row_value(const RowSrc& S, int x, int j)
{
    if (S.mode == ROW_LGF) {
        const int a = x <= S.Px / 2 ? x : S.Px - x;
        const int b = j <= S.Py / 2 ? j : S.Py - j;
        return S.tab[a + (size_t)S.tmx * b];
    }
    if (j >= S.ny || x >= S.nx) return 0.0;
    if (S.mode == ROW_DENSE) return S.field[x + (size_t)S.nx * j];
    const int lo = S.ptr[j], hi = S.ptr[j + 1];
    if (lo == hi) return 0.0;
    int a = lo, b = hi;
    while (a < b) {
        const int m = (a + b) >> 1;
        if (S.ebx[m] < x - 4) a = m + 1; else b = m;
    }
    double acc = 0.0;
    for (int k = a; k < hi; ++k) {
        const int bx = S.ebx[k];
        if (bx > x) break;
        const Term& tt = S.terms[S.ent[k]];
        acc += S.amp[tt.ai] * tt.coef * tt.wy[j - tt.by] * tt.wx[x - bx];
    }
    return acc;
}

// Pass 1: work-group b transforms rows (j0, j0+1) of the source, writes the two
// half spectra to T rows 2b, 2b+1 (row stride H = Px/2+1).
template <int VALS>
__global__ void __launch_bounds__(NT) k_rows_fwd(RowSrc S, int Px, int nrows_limit, double2* __restrict__ T,
                                                 const double2* __restrict__ tw)
{
    extern __shared__ double2 lds[];
    const int b = blockIdx.x;
    const int j0 = S.pairs ? 2 * S.pairs[b] : 2 * b;
    const int H = Px / 2 + 1;
    const int nx_src = S.mode == ROW_LGF ? Px : S.nx;
    for (int x = threadIdx.x; x < Px; x += NT) {
        double a = 0.0, c = 0.0;
        if (x < nx_src) {
            a = row_value(S, x, j0);
            if (j0 + 1 < nrows_limit) c = row_value(S, x, j0 + 1);
        }
        lds[x] = make_double2(a, c);
    }
    __syncthreads();
    lds_fft<VALS>(lds, Px, 1, tw, -1);
    double2* ra = T + (size_t)(2 * b) * H;
    double2* rb = ra + H;
    for (int k = threadIdx.x; k < H; k += NT) {
        const double2 z = lds[k];
        const double2 zc = lds[(Px - k) & (Px - 1)];
        // X_a = (Z + conj Zr)/2,  X_b = (Z - conj Zr)/(2i)
        ra[k] = make_double2(0.5 * (z.x + zc.x), 0.5 * (z.y - zc.y));
        rb[k] = make_double2(0.5 * (z.y + zc.y), -0.5 * (z.x - zc.x));
    }
}

// Pass 2: columns.  Work-group handles column slots [c0, c0+W); slot 0 packs
// the real columns kx = 0 and kx = Px/2.  Input rows: T1 row r holds field row
// rowmap(r); absent rows are zero.  Output: T2 rows 0..ny-1 (full).
template <int VALS>
__global__ void __launch_bounds__(NT) k_cols(const double2* __restrict__ T1, const int* __restrict__ pairs,
                                             int nin_rows, int Px, int Py, int ny, int W,
                                             const double* __restrict__ spec, double2* __restrict__ T2,
                                             const double2* __restrict__ tw)
{
    extern __shared__ double2 lds[];
    const int H = Px / 2 + 1;
    const int c0 = blockIdx.x * W;
    for (int k = threadIdx.x; k < W * Py; k += NT) lds[k] = make_double2(0.0, 0.0);
    __syncthreads();
    for (int k = threadIdx.x; k < W * nin_rows; k += NT) {
        const int w = k % W, r = k / W;
        const int row = pairs ? 2 * pairs[r >> 1] + (r & 1) : r;
        if (row >= ny) continue;
        const int c = c0 + w;
        if (c >= Px / 2) continue;
        const double2* src = T1 + (size_t)r * H;
        double2 v;
        if (c == 0) v = make_double2(src[0].x, src[Px / 2].x);
        else v = src[c];
        lds[w * Py + row] = v;
    }
    __syncthreads();
    lds_fft<VALS>(lds, Py, W, tw, -1);
    // multiply by the real LGF spectrum
    if (c0 == 0) {
        // packed slot: C' = ((G0+GN)/2) C(k) + ((G0-GN)/2) conj(C(Py-k))
        double2 nv[VALS];
        const int per = (Py + NT - 1) / NT;
        for (int q = 0; q < per && q < VALS; ++q) {
            const int ky = threadIdx.x + q * NT;
            if (ky < Py) {
                const double g0 = spec[ky], gn = spec[(size_t)(Px / 2) * Py + ky];
                const double2 cz = lds[ky], cr = lds[(Py - ky) & (Py - 1)];
                const double a = 0.5 * (g0 + gn), d = 0.5 * (g0 - gn);
                nv[q] = make_double2(a * cz.x + d * cr.x, a * cz.y - d * cr.y);
            }
        }
        __syncthreads();
        for (int q = 0; q < per && q < VALS; ++q) {
            const int ky = threadIdx.x + q * NT;
            if (ky < Py) lds[ky] = nv[q];
        }
        for (int k = threadIdx.x; k < (W - 1) * Py; k += NT) {
            const int w = 1 + k / Py, ky = k % Py;
            if (c0 + w >= Px / 2) continue;
            const double g = spec[(size_t)(c0 + w) * Py + ky];
            lds[w * Py + ky].x *= g;
            lds[w * Py + ky].y *= g;
        }
    } else {
        for (int k = threadIdx.x; k < W * Py; k += NT) {
            const int w = k / Py, ky = k - w * Py;
            if (c0 + w >= Px / 2) continue;
            const double g = spec[(size_t)(c0 + w) * Py + ky];
            lds[k].x *= g;
            lds[k].y *= g;
        }
    }
    __syncthreads();
    lds_fft<VALS>(lds, Py, W, tw, +1);
    for (int k = threadIdx.x; k < W * ny; k += NT) {
        const int w = k % W, r = k / W;
        const int c = c0 + w;
        if (c >= Px / 2) continue;
        const double2 v = lds[w * Py + r];
        double2* dst = T2 + (size_t)r * H;
        if (c == 0) {
            dst[0] = make_double2(v.x, 0.0);
            dst[Px / 2] = make_double2(v.y, 0.0);
        } else {
            dst[c] = v;
        }
    }
}

// Spectrum build: forward column FFT of T (all Py rows), keep the real part.
template <int VALS>
__global__ void __launch_bounds__(NT) k_cols_spec(const double2* __restrict__ T, int Px, int Py, int W, double scale,
                                                  double* __restrict__ spec, const double2* __restrict__ tw)
{
    extern __shared__ double2 lds[];
    const int H = Px / 2 + 1;
    const int c0 = blockIdx.x * W;
    for (int k = threadIdx.x; k < W * Py; k += NT) {
        const int w = k % W, r = k / W;
        const int c = c0 + w;
        lds[w * Py + r] = c < H ? T[(size_t)r * H + c] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    lds_fft<VALS>(lds, Py, W, tw, -1);
    for (int k = threadIdx.x; k < W * Py; k += NT) {
        const int w = k / Py, ky = k - w * Py;
        const int c = c0 + w;
        if (c < H) spec[(size_t)c * Py + ky] = lds[k].x * scale;
    }
}

// Pass 3: inverse rows.  Work-group b makes output rows 2b, 2b+1.
template <int VALS>
__global__ void __launch_bounds__(NT) k_rows_inv(const double2* __restrict__ T2, int Px, int nx, int ny,
                                                 double* __restrict__ out, const double2* __restrict__ tw)
{
    extern __shared__ double2 lds[];
    const int j0 = 2 * blockIdx.x;
    const int H = Px / 2 + 1;
    const bool two = j0 + 1 < ny;
    const double2* ra = T2 + (size_t)j0 * H;
    const double2* rb = ra + H;
    for (int k = threadIdx.x; k < H; k += NT) {
        const double2 A = ra[k];
        const double2 B = two ? rb[k] : make_double2(0.0, 0.0);
        // Z(k) = A + iB, Z(P-k) = conj(A) + i conj(B)
        lds[k] = make_double2(A.x - B.y, A.y + B.x);
        if (k > 0 && k < Px / 2) lds[Px - k] = make_double2(A.x + B.y, B.x - A.y);
    }
    __syncthreads();
    lds_fft<VALS>(lds, Px, 1, tw, +1);
    for (int x = threadIdx.x; x < nx; x += NT) {
        const double2 z = lds[x];
        out[x + (size_t)nx * j0] = z.x;
        if (two) out[x + (size_t)nx * (j0 + 1)] = z.y;
    }
}

// ------------------------------------------------------------- launches
static int vals_for(int total)   // complex values per thread
{
    const int v = (total + NT - 1) / NT;
    if (v <= 8) return 8;
    if (v <= 16) return 16;
    return 32;
}

static int cols_per_wg(int Py) { return std::max(1, std::min(8, 8192 / Py)); }

static void set_lds(const void* fn, size_t bytes)
{
    ILH_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
}

static void launch_rows_fwd(hipStream_t st, const RowSrc& S, int Px, int nrows_limit, int nblocks, double2* T,
                            const double2* tw)
{
    const size_t lds = sizeof(double2) * Px;
    switch (vals_for(Px)) {
    case 8: set_lds((const void*)k_rows_fwd<8>, lds); k_rows_fwd<8><<<nblocks, NT, lds, st>>>(S, Px, nrows_limit, T, tw); break;
    case 16: set_lds((const void*)k_rows_fwd<16>, lds); k_rows_fwd<16><<<nblocks, NT, lds, st>>>(S, Px, nrows_limit, T, tw); break;
    default: set_lds((const void*)k_rows_fwd<32>, lds); k_rows_fwd<32><<<nblocks, NT, lds, st>>>(S, Px, nrows_limit, T, tw); break;
    }
    ILH_HIP(hipGetLastError());
}

static void launch_cols(hipStream_t st, const double2* T1, const int* pairs, int nin_rows, int Px, int Py, int ny,
                        const double* spec, double2* T2, const double2* tw)
{
    const int W = cols_per_wg(Py);
    const int nslots = Px / 2;
    const int nblocks = (nslots + W - 1) / W;
    const size_t lds = sizeof(double2) * (size_t)W * Py;
    switch (vals_for(W * Py)) {
    case 8: set_lds((const void*)k_cols<8>, lds); k_cols<8><<<nblocks, NT, lds, st>>>(T1, pairs, nin_rows, Px, Py, ny, W, spec, T2, tw); break;
    case 16: set_lds((const void*)k_cols<16>, lds); k_cols<16><<<nblocks, NT, lds, st>>>(T1, pairs, nin_rows, Px, Py, ny, W, spec, T2, tw); break;
    default: set_lds((const void*)k_cols<32>, lds); k_cols<32><<<nblocks, NT, lds, st>>>(T1, pairs, nin_rows, Px, Py, ny, W, spec, T2, tw); break;
    }
    ILH_HIP(hipGetLastError());
}

static void launch_rows_inv(hipStream_t st, const double2* T2, int Px, int nx, int ny, double* out, const double2* tw)
{
    const size_t lds = sizeof(double2) * Px;
    const int nblocks = (ny + 1) / 2;
    switch (vals_for(Px)) {
    case 8: set_lds((const void*)k_rows_inv<8>, lds); k_rows_inv<8><<<nblocks, NT, lds, st>>>(T2, Px, nx, ny, out, tw); break;
    case 16: set_lds((const void*)k_rows_inv<16>, lds); k_rows_inv<16><<<nblocks, NT, lds, st>>>(T2, Px, nx, ny, out, tw); break;
    default: set_lds((const void*)k_rows_inv<32>, lds); k_rows_inv<32><<<nblocks, NT, lds, st>>>(T2, Px, nx, ny, out, tw); break;
    }
    ILH_HIP(hipGetLastError());
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
        int P = 2;
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
ConvPlan& get_conv(ilh_plan* pl, int gk)
{
    ConvPlan& C = pl->conv[gk];
    if (pl->conv_ok[gk]) return C;
    ilh_ctx* ctx = pl->ctx;
    ensure_twiddles(ctx);
    ensure_lgf(pl);
    const Grid& g = pl->g;
    C.gk = gk;
    C.nx = g.nx(gk);
    C.ny = g.ny(gk);
    C.Px = 2;
    while (C.Px < 2 * C.nx - 2) C.Px <<= 1;
    C.Py = 2;
    while (C.Py < 2 * C.ny - 2) C.Py <<= 1;
    if (C.Px > TWN || C.Py > TWN) fail(ILH_ERR_ARG, "grid too large for the FFT convolution (period > 8192)");
    C.H = C.Px / 2 + 1;
    const int rows = std::max(C.Py, C.ny + 1);
    C.T1.ensure(sizeof(double2) * (size_t)rows * C.H);
    C.T2.ensure(sizeof(double2) * (size_t)(C.ny + 1) * C.H);
    C.spec.ensure(sizeof(double) * (size_t)C.H * C.Py);
    // spectrum of the circular kernel: rows of G_circ, then forward columns
    RowSrc S{};
    S.mode = ROW_LGF;
    S.tab = pl->lgf.as<double>();
    S.tmx = pl->lgf_mx;
    S.Px = C.Px;
    S.Py = C.Py;
    hipStream_t st = ctx->stream;
    const double2* tw = ctx->twiddle.as<double2>();
    launch_rows_fwd(st, S, C.Px, C.Py, C.Py / 2, C.T1.as<double2>(), tw);
    const int W = cols_per_wg(C.Py);
    const size_t lds = sizeof(double2) * (size_t)W * C.Py;
    const int nblocks = (C.H + W - 1) / W;
    const double scale = g.dx * g.dx / ((double)C.Px * C.Py);
    switch (vals_for(W * C.Py)) {
    case 8: set_lds((const void*)k_cols_spec<8>, lds); k_cols_spec<8><<<nblocks, NT, lds, st>>>(C.T1.as<double2>(), C.Px, C.Py, W, scale, C.spec.as<double>(), tw); break;
    case 16: set_lds((const void*)k_cols_spec<16>, lds); k_cols_spec<16><<<nblocks, NT, lds, st>>>(C.T1.as<double2>(), C.Px, C.Py, W, scale, C.spec.as<double>(), tw); break;
    default: set_lds((const void*)k_cols_spec<32>, lds); k_cols_spec<32><<<nblocks, NT, lds, st>>>(C.T1.as<double2>(), C.Px, C.Py, W, scale, C.spec.as<double>(), tw); break;
    }
    ILH_HIP(hipGetLastError());
    pl->conv_ok[gk] = true;
    return C;
}

void conv_dense(ilh_plan* pl, int gk, const double* in, double* out)
{
    ConvPlan& C = get_conv(pl, gk);
    hipStream_t st = pl->ctx->stream;
    const double2* tw = pl->ctx->twiddle.as<double2>();
    RowSrc S{};
    S.mode = ROW_DENSE;
    S.field = in;
    S.nx = C.nx;
    S.ny = C.ny;
    const int npairs = (C.ny + 1) / 2;
    launch_rows_fwd(st, S, C.Px, C.ny, npairs, C.T1.as<double2>(), tw);
    launch_cols(st, C.T1.as<double2>(), nullptr, C.ny, C.Px, C.Py, C.ny, C.spec.as<double>(), C.T2.as<double2>(), tw);
    launch_rows_inv(st, C.T2.as<double2>(), C.Px, C.nx, C.ny, out, tw);
}

void build_source_pairs(ilh_plan* pl, SourceSet& s)
{
    const std::vector<int>& rp = s.bins.h_row_ptr;
    std::vector<int> pairs;
    const int ny = s.bins.ny;
    for (int j = 0; j < ny; j += 2) {
        const bool a = rp[j + 1] > rp[j];
        const bool b = (j + 1 < ny) && (rp[j + 2] > rp[j + 1]);
        if (a || b) pairs.push_back(j / 2);
    }
    s.npairs = (int)pairs.size();
    if (s.pairs) (void)hipFree(s.pairs);
    ILH_HIP(hipMalloc(&s.pairs, sizeof(int) * (size_t)std::max(1, s.npairs)));
    if (s.npairs)
        ILH_HIP(hipMemcpyAsync(s.pairs, pairs.data(), sizeof(int) * (size_t)s.npairs, hipMemcpyHostToDevice,
                               pl->ctx->stream));
    ILH_HIP(hipStreamSynchronize(pl->ctx->stream));
}

void conv_sources(ilh_plan* pl, int gk, const SourceSet& src, const double* amp, double* out, hipEvent_t* ev)
{
    ConvPlan& C = get_conv(pl, gk);
    hipStream_t st = pl->ctx->stream;
    const double2* tw = pl->ctx->twiddle.as<double2>();
    RowSrc S{};
    S.mode = ROW_TERMS;
    S.nx = C.nx;
    S.ny = C.ny;
    S.terms = src.terms;
    S.ptr = src.bins.row_ptr;
    S.ent = src.bins.entries;
    S.ebx = src.bins.entries + src.bins.n_entries;
    S.amp = amp;
    S.pairs = src.pairs;
    if (src.npairs > 0) launch_rows_fwd(st, S, C.Px, C.ny, src.npairs, C.T1.as<double2>(), tw);
    if (ev) ILH_HIP(hipEventRecord(ev[0], st));
    launch_cols(st, C.T1.as<double2>(), src.pairs, 2 * src.npairs, C.Px, C.Py, C.ny, C.spec.as<double>(),
                C.T2.as<double2>(), tw);
    if (ev) ILH_HIP(hipEventRecord(ev[1], st));
    launch_rows_inv(st, C.T2.as<double2>(), C.Px, C.nx, C.ny, out, tw);
}

}  // namespace ilh
