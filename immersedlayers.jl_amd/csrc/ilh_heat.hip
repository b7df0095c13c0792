// ilh_heat.hip -- unbounded heat conduction with Dirichlet bodies, advanced by
// the integrating-factor half-explicit Runge-Kutta method (IF-HERK).
//
// The reference poses the problem as the constrained ODE system
//   dT/dt = kappa L T + r1 + R sigma,   E T = r2
// (test/literate/heatconduction.jl:88-130: r1 = surface_divergence!(-kappa [T]),
// constraint force regularize!(-(-sigma)), bc_op interpolate!, bc_rhs the
// surface average) and marches it with ConstrainedSystems' LiskaIFHERK
// (src/timemarching.jl:5, integrating factor for lin_op = kappa L).  IF-HERK
// applies a half-explicit RK scheme to v = exp(-(t - t_n) kappa L) T, so the
// stiff diffusion is integrated exactly; the 3-stage tableau (Liska & Colonius
// 2016) is
//   c = (0, 1/3, 1),  a21 = 1/3,  a31 = -1, a32 = 2,  b = (0, 3/4, 1/4),
// and the multiplier of each stage solves a Schur system
//   E exp((c_i - c_{i-1}) dt kappa L) R sigma = E Yhat_i - r2
// with the factors exp(dt/3 kappa L), exp(2dt/3 kappa L) and the identity.
//
// exp(tau kappa L) on the unbounded lattice is separable:
//   [exp(tau kappa L) u](i,j) = sum_{p,q} K(p) K(q) u(i-p, j-q),
//   K(p) = e^{-2a} I_p(2a),  a = tau kappa / dx^2,
// applied here as an x then a y pass of a (2M+1)-tap filter (zero outside the
// grid).  The three Schur matrices are assembled directly from K(p)K(q) like
// the LGF ones, inverted once, and applied as dense matvecs.
#include "ilh_internal.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

struct ilh_heat {
    ilh_plan* pl = nullptr;
    unsigned gen = 0;
    double kappa = 0.0, dt = 0.0;
    int M13 = 0, M23 = 0;                 // filter half-widths
    ilh::DevBuf k13, k23;                 // K(0..M) of exp(dt/3 kL), exp(2dt/3 kL)
    ilh::DevBuf S13, S23, S0;             // row-major inverse Schur matrices
    ilh::DevBuf F, A, Bt, Y, W1, W2;  // grid fields
    ilh::DevBuf e, sg, r2;                // surface vectors
    bool have_bc = false;
};

namespace ilh {

// K(p) = e^{-x} I_p(x), x = 2a, p = 0..M, by backward recurrence normalised
// with e^{x} = I_0(x) + 2 sum_{p>=1} I_p(x) (so sum over all p of K(p) = 1);
// truncated after the last p with K(p) > 1e-17 K(0).
static std::vector<double> heat_kernel_1d(double a)
{
    if (a <= 0.0) return {1.0};
    const double x = 2.0 * a;
    const int top = (int)(2.0 * x + 60.0 + 10.0 * std::sqrt(x));
    std::vector<double> v(top + 2, 0.0);
    v[top + 1] = 0.0;
    v[top] = 1e-280;
    for (int p = top; p >= 1; --p) {
        v[p - 1] = (2.0 * p / x) * v[p] + v[p + 1];
        if (v[p - 1] > 1e250)   // rescale to stay in range
            for (int q = p - 1; q <= top + 1; ++q) v[q] *= 1e-250;
    }
    double sum = v[0];
    for (int p = 1; p <= top; ++p) sum += 2.0 * v[p];
    std::vector<double> K;
    for (int p = 0; p <= top; ++p) {
        const double k = v[p] / sum;
        if (p > 0 && k <= 1e-17 * (v[0] / sum)) break;
        K.push_back(k);
    }
    return K;
}

namespace {

constexpr int HX = 64, HY = 4;          // naive kernels (long filters)
constexpr int XT = 256, XR = 8;          // x pass: 256-wide row segments, 8 rows per work-group
constexpr int YC = 64, YR = 64;          // y pass: columns x rows of outputs per work-group
constexpr int MX_LDS = 64, MY_LDS = 32;  // longest filters the tiled passes take

// x pass: out(i,j) = sum_{|p|<=M} K(|p|) u(i-p, j), u = a + c b (b may be
// null), zero outside [0, nx).  Eight row segments and their halos are
// staged in LDS; every thread issues all of its loads before the first store.
__global__ void __launch_bounds__(XT) k_heat_x(const double* __restrict__ a, double c, const double* __restrict__ b,
                                               int nx, int ny, const double* __restrict__ K, int M,
                                               double* __restrict__ out)
{
    constexpr int W = XT + 2 * MX_LDS;
    __shared__ double buf[XR][W];
    const int j0 = blockIdx.y * XR, x0 = blockIdx.x * XT, t = threadIdx.x;
    const int w = XT + 2 * M;
    double va[XR][2], vb[XR][2];
#pragma unroll
    for (int r = 0; r < XR; ++r)
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int e = t + XT * q, x = x0 - M + e, j = j0 + r;
            const bool in = e < w && x >= 0 && x < nx && j < ny;
            va[r][q] = in ? a[x + (size_t)nx * j] : 0.0;
            vb[r][q] = in && b ? b[x + (size_t)nx * j] : 0.0;
        }
#pragma unroll
    for (int r = 0; r < XR; ++r)
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int e = t + XT * q;
            if (e < w) buf[r][e] = b ? va[r][q] + c * vb[r][q] : va[r][q];
        }
    __syncthreads();
    const int i = x0 + t;
    if (i >= nx) return;
#pragma unroll
    for (int r = 0; r < XR; ++r) {
        const int j = j0 + r;
        if (j >= ny) break;
        const double* row = buf[r] + t + M;
        double acc = K[0] * row[0];
        for (int p = 1; p <= M; ++p) acc += K[p] * (row[-p] + row[p]);
        out[i + (size_t)nx * j] = acc;
    }
}

// y pass: out(i,j) = sum_{|p|<=M} K(|p|) in(i, j-p) + c add(i,j) (add may be
// null), zero outside [0, ny).  A 64-column strip of YR + 2M rows is staged
// in LDS (dynamic, (YR + 2M) x 64 doubles); loads, including those of `add`,
// are all issued before the strip is stored.
__global__ void __launch_bounds__(256) k_heat_y(const double* __restrict__ in, int nx, int ny,
                                                const double* __restrict__ K, int M, double c,
                                                const double* __restrict__ add, double* __restrict__ out)
{
    extern __shared__ double strip[];   // [YR + 2M][YC]
    constexpr int QMAX = (YR + 2 * MY_LDS) / 4;
    const int tx = threadIdx.x & (YC - 1), ty = threadIdx.x / YC;   // 64 x 4
    const int i = blockIdx.x * YC + tx, y0 = blockIdx.y * YR;
    const int rows = YR + 2 * M;
    double v[QMAX], ad[YR / 4];
#pragma unroll
    for (int q = 0; q < QMAX; ++q) {
        const int r = ty + 4 * q, y = y0 - M + r;
        v[q] = (r < rows && i < nx && y >= 0 && y < ny) ? in[i + (size_t)nx * y] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < YR / 4; ++q) {
        const int y = y0 + ty + 4 * q;
        ad[q] = (add && i < nx && y < ny) ? add[i + (size_t)nx * y] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < QMAX; ++q) {
        const int r = ty + 4 * q;
        if (r < rows) strip[r * YC + tx] = v[q];
    }
    __syncthreads();
    if (i >= nx) return;
#pragma unroll
    for (int q = 0; q < YR / 4; ++q) {
        const int r = ty + 4 * q, y = y0 + r;
        if (y >= ny) break;
        const double* col = strip + (r + M) * YC + tx;
        double acc = K[0] * col[0];
        for (int p = 1; p <= M; ++p) acc += K[p] * (col[-p * YC] + col[p * YC]);
        out[i + (size_t)nx * y] = add ? acc + c * ad[q] : acc;
    }
}

// k_heat_y for filters of at most MB taps a side: the same staging, then each
// thread takes 16 consecutive rows of its column (thread (tx, ty): rows
// 16 ty .. 16 ty + 15) from a register window of 16 + 2 MB strip values, so
// the strip is read 16 + 2 MB times per thread instead of 16 (2M + 1).  Every
// output sums K[0] col[0] and then K[p] (col[-p] + col[p]) for p = 1, 2, ...
// in k_heat_y's order; the taps past M are zero and add nothing.
constexpr int YMB = 16;

template <int MB>
__global__ void __launch_bounds__(256) k_heat_y_w(const double* __restrict__ in, int nx, int ny,
                                                  const double* __restrict__ K, int M, double c,
                                                  const double* __restrict__ add, double* __restrict__ out)
{
    extern __shared__ double strip[];   // [YR + 2M][YC]
    constexpr int QMAX = (YR + 2 * MB) / 4, NO = YR / 4;
    const int tx = threadIdx.x & (YC - 1), ty = threadIdx.x / YC;   // 64 x 4
    const int i = blockIdx.x * YC + tx, y0 = blockIdx.y * YR;
    const int rows = YR + 2 * M;
    double v[QMAX], ad[NO];
#pragma unroll
    for (int q = 0; q < QMAX; ++q) {
        const int r = ty + 4 * q, y = y0 - M + r;
        v[q] = (r < rows && i < nx && y >= 0 && y < ny) ? in[i + (size_t)nx * y] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < NO; ++q) {
        const int y = y0 + NO * ty + q;
        ad[q] = (add && i < nx && y < ny) ? add[i + (size_t)nx * y] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < QMAX; ++q) {
        const int r = ty + 4 * q;
        if (r < rows) strip[r * YC + tx] = v[q];
    }
    __syncthreads();
    if (i >= nx) return;
    // window entry k is strip row NO ty + k - (MB - M): output q's centre is
    // entry q + MB
    double w[NO + 2 * MB], kk[MB + 1];
#pragma unroll
    for (int k = 0; k < NO + 2 * MB; ++k) {
        const int r = NO * ty + k - (MB - M);
        w[k] = r >= 0 && r < rows ? strip[r * YC + tx] : 0.0;
    }
#pragma unroll
    for (int p = 0; p <= MB; ++p) kk[p] = p <= M ? K[p] : 0.0;
#pragma unroll
    for (int q = 0; q < NO; ++q) {
        const int y = y0 + NO * ty + q;
        if (y >= ny) break;
        double acc = kk[0] * w[q + MB];
#pragma unroll
        for (int p = 1; p <= MB; ++p) acc += kk[p] * (w[q + MB - p] + w[q + MB + p]);
        out[i + (size_t)nx * y] = add ? acc + c * ad[q] : acc;
    }
}

// the same passes for filters longer than the tiled ones take
__global__ void k_heat_x_long(const double* __restrict__ a, double c, const double* __restrict__ b, int nx, int ny,
                              const double* __restrict__ K, int M, double* __restrict__ out)
{
    const int i = blockIdx.x * HX + threadIdx.x, j = blockIdx.y * HY + threadIdx.y;
    if (i >= nx || j >= ny) return;
    const double* ra = a + (size_t)nx * j;
    const double* rb = b ? b + (size_t)nx * j : nullptr;
    auto u = [&](int x) { return rb ? ra[x] + c * rb[x] : ra[x]; };
    double acc = K[0] * u(i);
    for (int p = 1; p <= M; ++p) {
        const double l = i - p >= 0 ? u(i - p) : 0.0;
        const double r = i + p < nx ? u(i + p) : 0.0;
        acc += K[p] * (l + r);
    }
    out[i + (size_t)nx * j] = acc;
}

__global__ void k_heat_y_long(const double* __restrict__ in, int nx, int ny, const double* __restrict__ K, int M,
                              double c, const double* __restrict__ add, double* __restrict__ out)
{
    const int i = blockIdx.x * HX + threadIdx.x, j = blockIdx.y * HY + threadIdx.y;
    if (i >= nx || j >= ny) return;
    double acc = K[0] * in[i + (size_t)nx * j];
    for (int p = 1; p <= M; ++p) {
        const double l = j - p >= 0 ? in[i + (size_t)nx * (j - p)] : 0.0;
        const double r = j + p < ny ? in[i + (size_t)nx * (j + p)] : 0.0;
        acc += K[p] * (l + r);
    }
    const size_t o = i + (size_t)nx * j;
    out[o] = add ? acc + c * add[o] : acc;
}

// d = Tb+ - Tb- (scaled by -kappa), r2 = (Tb+ + Tb-)/2
__global__ void k_heat_bc(const double* __restrict__ tp, const double* __restrict__ tm, int n, double kappa,
                          double* __restrict__ dneg, double* __restrict__ r2)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double a = tp[k], b = tm ? tm[k] : 0.0;
    dneg[k] = -kappa * (a - b);
    r2[k] = 0.5 * (a + b);
}

}  // namespace

// out = exp(tau kappa L) (a + ca b) + cp add   (b, add may be null; tmp: one
// grid field of scratch; out may alias a, b or add)
static void heat_apply(ilh_plan* pl, const double* K, int M, const double* a, double ca, const double* b,
                       double* tmp, double* out, double cp = 0.0, const double* add = nullptr)
{
    const int nx = pl->g.nx(GK_PRIMAL), ny = pl->g.ny(GK_PRIMAL);
    hipStream_t st = pl->ctx->stream;
    if (M <= MX_LDS)
        k_heat_x<<<dim3((nx + XT - 1) / XT, (ny + XR - 1) / XR), XT, 0, st>>>(a, ca, b, nx, ny, K, M, tmp);
    else
        k_heat_x_long<<<dim3((nx + HX - 1) / HX, (ny + HY - 1) / HY), dim3(HX, HY), 0, st>>>(a, ca, b, nx, ny, K, M,
                                                                                          tmp);
    if (M <= YMB)
        k_heat_y_w<YMB><<<dim3((nx + YC - 1) / YC, (ny + YR - 1) / YR), 256, sizeof(double) * (YR + 2 * M) * YC,
                           st>>>(tmp, nx, ny, K, M, cp, add, out);
    else if (M <= MY_LDS)
        k_heat_y<<<dim3((nx + YC - 1) / YC, (ny + YR - 1) / YR), 256, sizeof(double) * (YR + 2 * M) * YC, st>>>(
            tmp, nx, ny, K, M, cp, add, out);
    else
        k_heat_y_long<<<dim3((nx + HX - 1) / HX, (ny + HY - 1) / HY), dim3(HX, HY), 0, st>>>(tmp, nx, ny, K, M, cp,
                                                                                          add, out);
    ILH_HIP(hipGetLastError());
}

// 2-D kernel table K(m)K(n), (M+1)^2, for the Schur assembly
// The stage Schur matrix E H R = M diag(ds / dx^2) with M = E H E^T symmetric
// positive definite (H the heat kernel, or the identity): inverted through the
// deflated Cholesky of M, LU of E H R if that meets a non-positive pivot.
static void schur_inverse(ilh_plan* pl, const std::vector<double>& K, bool identity, DevBuf& inv)
{
    const int N = pl->N;
    hipStream_t st = pl->ctx->stream;
    DevBuf tabd;
    int T = 1;
    if (!identity) {
        T = (int)K.size();
        std::vector<double> tab((size_t)T * T);
        for (int n = 0; n < T; ++n)
            for (int m = 0; m < T; ++m) tab[m + (size_t)T * n] = K[m] * K[n];
        tabd.ensure(sizeof(double) * tab.size());
        ILH_HIP(hipMemcpyAsync(tabd.p, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice, st));
        ILH_HIP(hipStreamSynchronize(st));
    }
    auto build = [&](int opB, double* out) {
        if (identity) assemble(pl, OP_E_PRIMAL, opB, GK_PRIMAL, 1.0, true, out);
        else assemble_table(pl, OP_E_PRIMAL, opB, tabd.as<double>(), T, 1.0, out);
    };
    inv.ensure(sizeof(double) * (size_t)N * solve_ld(N));
    SymFactor F;
    F.A.ensure(sizeof(double) * (size_t)N * N);
    build(OP_E_PRIMAL, F.A.as<double>());
    if (sym_factor(pl, F, N)) {
        sym_inverse_rowmajor(pl, F, pl->ds.as<double>(), 1.0 / (pl->g.dx * pl->g.dx), inv.as<double>());
        return;
    }
    DevBuf S;
    S.ensure(sizeof(double) * (size_t)N * N);
    build(OP_R_PRIMAL, S.as<double>());
    inverse_rowmajor(pl, S.as<double>(), inv.as<double>());
}

}  // namespace ilh

using namespace ilh;

namespace {
constexpr double A21 = 1.0 / 3.0, A31 = -1.0, A32 = 2.0, B2 = 0.75, B3 = 0.25;
}

extern "C" int ilh_heat_create(ilh_plan* pl, double kappa, double dt, ilh_heat** out)
{
    return guarded_call([&] {
        if (!pl || !out) fail(ILH_ERR_ARG, "null argument");
        if (!(kappa > 0.0) || !(dt > 0.0) || !std::isfinite(kappa * dt)) fail(ILH_ERR_ARG, "kappa and dt must be positive");
        ILH_HIP(hipSetDevice(pl->ctx->device));
        if (pl->N > 0) need_margin(pl, "heat conduction");
        const double a13 = kappa * dt / 3.0 / (pl->g.dx * pl->g.dx);
        const std::vector<double> K13 = heat_kernel_1d(a13), K23 = heat_kernel_1d(2.0 * a13);
        if ((int)K23.size() > 256) fail(ILH_ERR_ARG, "time step too large for the integrating-factor filter (a > ~40)");
        ilh_heat* h = new ilh_heat();
        try {
            h->pl = pl;
            h->gen = pl->gen;
            h->kappa = kappa;
            h->dt = dt;
            h->M13 = (int)K13.size() - 1;
            h->M23 = (int)K23.size() - 1;
            hipStream_t st = pl->ctx->stream;
            h->k13.ensure(sizeof(double) * K13.size());
            h->k23.ensure(sizeof(double) * K23.size());
            ILH_HIP(hipMemcpyAsync(h->k13.p, K13.data(), sizeof(double) * K13.size(), hipMemcpyHostToDevice, st));
            ILH_HIP(hipMemcpyAsync(h->k23.p, K23.data(), sizeof(double) * K23.size(), hipMemcpyHostToDevice, st));
            const size_t ng = pl->g.n(GK_PRIMAL);
            for (DevBuf* b : {&h->F, &h->A, &h->Bt, &h->Y, &h->W1, &h->W2}) b->ensure(sizeof(double) * ng);
            const int N = pl->N;
            for (DevBuf* b : {&h->e, &h->sg, &h->r2}) b->ensure(sizeof(double) * std::max(1, N));
            if (N > 0) {
                schur_inverse(pl, K13, false, h->S13);
                schur_inverse(pl, K23, false, h->S23);
                schur_inverse(pl, {1.0}, true, h->S0);
            }
            ILH_HIP(hipStreamSynchronize(st));
        } catch (...) {
            delete h;
            throw;
        }
        plan_retain(pl);
        *out = h;
    });
}

extern "C" int ilh_heat_set_boundary(ilh_heat* h, const double* tb_plus, const double* tb_minus, int mem)
{
    return guarded_call([&] {
        if (!h) fail(ILH_ERR_ARG, "null heat plan");
        ilh_plan* pl = h->pl;
        if (pl->gen != h->gen) fail(ILH_ERR_STATE, "the plan's geometry changed since ilh_heat_create");
        Io io(pl, mem);
        const int N = pl->N;
        hipStream_t st = pl->ctx->stream;
        if (N > 0) {
            const double* tp = io.in(tb_plus, N, pl->io_a);
            const double* tm = tb_minus ? io.in(tb_minus, N, pl->io_c) : nullptr;
            // r1 = surface_divergence!(-kappa [T]) (heatconduction.jl:97-98), r2 = surface average
            k_heat_bc<<<(N + 255) / 256, 256, 0, st>>>(tp, tm, N, h->kappa, h->e.as<double>(), h->r2.as<double>());
            ILH_HIP(hipGetLastError());
            scatter_op(pl, OP_DS, h->e.as<double>(), h->F.as<double>());   // includes the 1/dx of GridScaling
        } else {
            fill_zero(st, h->F.as<double>(), pl->g.n(GK_PRIMAL));
        }
        h->have_bc = true;
        io.done();
    });
}

extern "C" int ilh_heat_advance(ilh_heat* h, double* T, int nsteps, int mem)
{
    return guarded_call([&] {
        if (!h) fail(ILH_ERR_ARG, "null heat plan");
        if (nsteps < 0) fail(ILH_ERR_ARG, "negative step count");
        ilh_plan* pl = h->pl;
        if (pl->gen != h->gen) fail(ILH_ERR_STATE, "the plan's geometry changed since ilh_heat_create");
        if (!h->have_bc) fail(ILH_ERR_STATE, "ilh_heat_set_boundary must come first");
        Io io(pl, mem);
        const int N = pl->N;
        const size_t ng = pl->g.n(GK_PRIMAL);
        hipStream_t st = pl->ctx->stream;
        double* d_T = io.out(T, ng, pl->io_b);
        if (mem == ILH_MEM_HOST)
            ILH_HIP(hipMemcpyAsync(d_T, T, sizeof(double) * ng, hipMemcpyHostToDevice, st));
        const double dt = h->dt;
        const int ld = solve_ld(std::max(1, N));
        double *F = h->F.as<double>(), *A = h->A.as<double>(), *Bt = h->Bt.as<double>(), *Y = h->Y.as<double>();
        double *W1 = h->W1.as<double>(), *W2 = h->W2.as<double>();
        double *e = h->e.as<double>(), *sg = h->sg.as<double>(), *r2 = h->r2.as<double>();
        const double* K13 = h->k13.as<double>();
        const double* K23 = h->k23.as<double>();
        // sigma = -(1/(dt c)) Sinv (E Y - r2);  W = base + cb R sigma (one scatter pass)
        auto multiplier = [&](const double* Yh, const double* Sinv, double c, const double* base, double cb,
                              double* W) {
            gather_op(pl, OP_E_PRIMAL, Yh, e);
            axpby(st, e, r2, -1.0, 1.0, N);                                  // e = E Y - r2
            matvec_rows(st, Sinv, ld, e, sg, N, N, -1.0 / (dt * c), 0);
            scatter_op_add(pl, OP_R_PRIMAL, sg, base, cb, W);
        };
        for (int n = 0; n < nsteps; ++n) {
            if (N == 0) {
                // no bodies: T <- H23 [H13 T + dt b2 F] + dt b3 F
                heat_apply(pl, K13, h->M13, d_T, 0.0, nullptr, Bt, A, dt * B2, F);
                heat_apply(pl, K23, h->M23, A, 0.0, nullptr, Bt, d_T, dt * B3, F);
                continue;
            }
            // stage 2: Yhat2 = H13 (T + dt a21 F)
            heat_apply(pl, K13, h->M13, d_T, dt * A21, F, Bt, Y);
            multiplier(Y, h->S13.as<double>(), A21, F, 1.0, W1);
            // stage 3: Yhat3 = H23 [H13 (T + dt a31 W1) + dt a32 F]
            heat_apply(pl, K13, h->M13, d_T, dt * A31, W1, Bt, A, dt * A32, F);
            heat_apply(pl, K23, h->M23, A, 0.0, nullptr, Bt, Y);
            multiplier(Y, h->S23.as<double>(), A32, F, 1.0, W2);
            // final: yhat = H23 [H13 T + dt b2 W2] + dt b3 F;  T = yhat + dt b3 R sigma3
            heat_apply(pl, K13, h->M13, d_T, 0.0, nullptr, Bt, A, dt * B2, W2);
            heat_apply(pl, K23, h->M23, A, 0.0, nullptr, Bt, Y, dt * B3, F);
            multiplier(Y, h->S0.as<double>(), B3, Y, dt * B3, d_T);
        }
        io.back(T, d_T, ng);
        io.done();
    });
}

extern "C" int ilh_heat_destroy(ilh_heat* h)
{
    return guarded_call([&] {
        if (!h) return;
        ilh_plan* pl = h->pl;
        (void)hipSetDevice(pl->ctx->device);
        (void)hipStreamSynchronize(pl->ctx->stream);
        delete h;
        plan_release(pl);
    });
}
