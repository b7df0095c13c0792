// ilh_api.hip -- the C ABI of libimmersedhip.so (include/immersedhip.h):
// context and plan lifetime, host/device staging, the per-op entry points and
// the Dirichlet / Neumann solve composites.
#include "ilh_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>

using namespace ilh;

// ------------------------------------------------ allocation-free scratch
// Device scratch of ilh_dirichlet_solve_noalloc (the surface amplitudes and
// the two convolution spectra).  Every GPU holds its own copy of the array
// until the process ends, at its own address, and every call on that GPU
// shares it.  `scratch` named in host code is a host-side stand-in, not a
// device address: host code only ever uses the per-device address that
// hipGetSymbolAddress returned (scratch_ptr), looked up once per device when
// the device's first context is created.
static constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)32 << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];
static constexpr int MAX_DEVICES = 64;
static void* g_scratch_ptr[MAX_DEVICES] = {};
static std::mutex g_scratch_mu;

// with `device` current
static void init_scratch(int device)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0 || device >= MAX_DEVICES) fail(ILH_ERR_ARG, "device index beyond the scratch table");
    if (g_scratch_ptr[device]) return;
    void* p = nullptr;
    ILH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) fail(ILH_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
    g_scratch_ptr[device] = p;
}

static void* scratch_ptr(int device)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    void* p = device >= 0 && device < MAX_DEVICES ? g_scratch_ptr[device] : nullptr;
    if (!p) fail(ILH_ERR_STATE, "no context was created on this device");
    return p;
}

// --------------------------------------------------------------- errors
static thread_local std::string g_err;

namespace ilh {
void set_error(const std::string& msg) { g_err = msg; }
[[noreturn]] void fail(int code, const std::string& msg) { throw Error{code, msg}; }
void check_hip(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw Error{ILH_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e)};
}
void DevBuf::ensure(size_t b)
{
    if (b <= bytes && p) return;
    release();
    if (b == 0) b = 8;
    if (hipMalloc(&p, b) != hipSuccess) {
        p = nullptr;
        bytes = 0;
        (void)hipGetLastError();
        throw Error{ILH_ERR_OOM, "device allocation of " + std::to_string(b) + " bytes failed"};
    }
    bytes = b;
}
void DevBuf::release()
{
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}
}  // namespace ilh

void ilh::set_last_error(const std::string& msg) { g_err = msg; }

template <class F>
static int guarded(F&& f)
{
    return guarded_call(static_cast<F&&>(f));
}

extern "C" const char* ilh_last_error(void) { return g_err.c_str(); }
extern "C" const char* ilh_version(void) { return "immersedhip 0.1 (gfx950, FP64)"; }

// -------------------------------------------------------- solve caches
struct SolveCache {
    DevBuf LU, piv, info;
    SymFactor sym;            // create_RTLinvR: the deflated Cholesky of M (S = M diag(ds)) instead of LU
    bool use_sym = false;
    bool minus_ok = false;    // Dirichlet: M[1] (the interior-data matrix) built
    DevBuf M[4];              // row-major N x N solve matrices
    DevBuf terms1, terms2;    // merged source terms
    SourceSet src1, src2;
    DevBuf amp1, amp2;
    DevBuf fin, fmin, ampm;   // batches of right-hand sides: inputs at an aligned stride, amplitudes
    ~SolveCache()
    {
        free_sources(src1);
        free_sources(src2);
    }
};

// Stokes flow (vector cache): inverses of create_CL2invCT and
// create_CLinvCT_scalar, and the solve's work buffers
struct StokesCache {
    DevBuf S2inv, Ssinv;          // row-major inverses, leading dimensions solve_ld(2N), solve_ld(N)
    DevBuf S2;                    // create_CL2invCT row-major (the residual of the refinement step)
    DevBuf Ka;                    // E C L^-2 C^T D_s^symm (2N x 2N, row-major): the slip of psi*
    DevBuf Cf;                    // the vector surface filter's two diagonal blocks (N x N each, row-major,
                                  // leading dimension solve_ld(N)), built on the first filtered solve
    bool cf_ok = false;
    DevBuf sv;                    // surface work: dv, vb, e_phi, w, vprime, sigma (2N each); dvn, dsv, ds (N each)
    DevBuf gE, gE2, gP;           // grid work: edges (2), primal
    DevBuf gDD;                   // two Nodes{Dual} fields back to back (one batched convolution)
};

// Moving bodies: the inverse Schur matrix of a reference geometry (row-major,
// leading dimension solve_ld(N)) preconditions GMRES on the current one; the
// current geometry's merged sources [R, D_s] are rebuilt after every
// update_points.
struct MovingCache {
    DevBuf P32;                // the reference inverse S_ref^-1 in FP32, row-major (leading dimension solve_ld(N))
    bool have_ref = false;
    int ref_age = 0;           // solves since the reference was factored
    DevBuf terms1, amp, V, w, z, b, fbar, grid;
    DevBuf S;                  // create_RTLinvR at the current geometry, row-major
    bool S_ok = false;
    DevBuf gm, gmf;            // GMRES state and flags on the device
    DevBuf arn;                // partial sums of an Arnoldi step
    int last_iters = 4;        // steps of the previous solve: the first chunk enqueued
    double* hpin = nullptr;    // pinned host buffer for the GMRES flags
    size_t hpin_n = 0;
    double* pinned(size_t n)
    {
        if (n > hpin_n) {
            if (hpin) (void)hipHostFree(hpin);
            hpin = nullptr;
            hpin_n = 0;
            ILH_HIP(hipHostMalloc((void**)&hpin, sizeof(double) * n, hipHostMallocDefault));
            hpin_n = n;
        }
        return hpin;
    }
    SourceSet src1;
    bool src_ok = false;
    // learned preconditioner: the band predicted for the current geometry
    DevBuf band;
    bool band_ok = false;
    bool use_band = false;     // this solve's preconditioner: the band (else the FP32 reference inverse)
    // the current geometry's matrix and sources are stale (their buffers are
    // kept for the next geometry)
    void drop_sources()
    {
        S_ok = false;
        src_ok = false;
        band_ok = false;
    }
    ~MovingCache()
    {
        free_sources(src1);
        if (hpin) (void)hipHostFree(hpin);
    }
};

// Geometry caches become stale; row-bin buffers are kept for the next
// geometry unless `release` (plan destruction).
static void drop_caches(ilh_plan* pl, bool release = false)
{
    ++pl->gen;
    if (pl->mov) pl->mov->drop_sources();
    delete pl->dir;
    delete pl->neu;
    delete pl->sto;
    pl->dir = pl->neu = nullptr;
    pl->sto = nullptr;
    for (auto& f : pl->fac) {
        delete f;
        f = nullptr;
    }
    for (int op = 0; op < OP_COUNT; ++op) {
        pl->terms_ok[op] = false;
        pl->wins_ok = false;
        pl->bins_ok[op] = false;
        if (release)
            for (auto& b : pl->bins[op]) free_bins(b);
    }
}

// ------------------------------------------------------------- context
extern "C" int ilh_ctx_create(int device, ilh_ctx** out)
{
    return guarded([&] {
        if (!out) fail(ILH_ERR_ARG, "null output pointer");
        int n = 0;
        ILH_HIP(hipGetDeviceCount(&n));
        if (device < 0 || device >= n) fail(ILH_ERR_ARG, "no such HIP device");
        ILH_HIP(hipSetDevice(device));
        init_scratch(device);
        ilh_ctx* c = new ilh_ctx();
        c->device = device;
        ILH_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        hipDeviceProp_t prop;
        ILH_HIP(hipGetDeviceProperties(&prop, device));
        c->num_cus = prop.multiProcessorCount;
        *out = c;
    });
}

void ilh::ctx_retain(ilh_ctx* c) { c->refs.fetch_add(1); }

void ilh::ctx_release(ilh_ctx* c)
{
    if (c->refs.fetch_sub(1) > 1) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (int k = 0; k < ilh_ctx::POOL; ++k)
        if (c->pool[k]) {
            (void)hipStreamSynchronize(c->pool[k]);
            (void)hipStreamDestroy(c->pool[k]);
        }
    for (hipEvent_t e : c->ev)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : c->tev)
        if (e) (void)hipEventDestroy(e);
    if (c->fused) fused_destroy(c->fused);
    for (FusedBatch* fb : c->fused_mv)
        if (fb) fused_destroy(fb);
    if (c->mv_dev) (void)hipFree(c->mv_dev);
    if (c->geo_dev) (void)hipFree(c->geo_dev);
    if (c->mv_host) (void)hipHostFree(c->mv_host);
    (void)hipStreamDestroy(c->stream);
    delete c;
}

// the caller's handle; plans and batches still using the context keep it
// alive until they are destroyed
extern "C" int ilh_ctx_destroy(ilh_ctx* c)
{
    return guarded([&] {
        if (c) ctx_release(c);
    });
}

extern "C" int ilh_ctx_synchronize(ilh_ctx* c)
{
    return guarded([&] {
        if (!c) fail(ILH_ERR_ARG, "null ctx");
        // a stream with nothing in flight returns at once (the blocking wait
        // costs ~35 us even on an idle stream)
        const hipError_t q = hipStreamQuery(c->stream);
        if (q == hipSuccess) return;
        if (q != hipErrorNotReady) ILH_HIP(q);
        ILH_HIP(hipStreamSynchronize(c->stream));
    });
}

extern "C" int ilh_malloc(ilh_ctx* c, size_t bytes, void** p)
{
    return guarded([&] {
        if (!c || !p) fail(ILH_ERR_ARG, "null argument");
        ILH_HIP(hipSetDevice(c->device));
        if (hipMalloc(p, bytes ? bytes : 8) != hipSuccess) {
            (void)hipGetLastError();
            fail(ILH_ERR_OOM, "hipMalloc failed");
        }
    });
}

extern "C" int ilh_free(ilh_ctx* c, void* p)
{
    return guarded([&] {
        if (!c) fail(ILH_ERR_ARG, "null ctx");
        if (p) ILH_HIP(hipFree(p));
    });
}

extern "C" int ilh_memcpy_h2d(ilh_ctx* c, void* dst, const void* src, size_t bytes)
{
    return guarded([&] {
        ILH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
        ILH_HIP(hipStreamSynchronize(c->stream));
    });
}

extern "C" int ilh_memcpy_d2h(ilh_ctx* c, void* dst, const void* src, size_t bytes)
{
    return guarded([&] {
        ILH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
        ILH_HIP(hipStreamSynchronize(c->stream));
    });
}

// ---------------------------------------------------------------- plan
static void load_points(ilh_plan* pl, const double* X, const double* Y, const double* nx, const double* ny,
                        const double* ds)
{
    const int N = pl->N;
    if (N > 0 && (!X || !Y || !nx || !ny || !ds)) fail(ILH_ERR_ARG, "null geometry array");
    double mm = 1e300;
    for (int p = 0; p < N; ++p) {
        if (!std::isfinite(X[p]) || !std::isfinite(Y[p]) || !std::isfinite(nx[p]) || !std::isfinite(ny[p]) ||
            !std::isfinite(ds[p]))
            fail(ILH_ERR_NONFINITE, "non-finite surface geometry");
        const double xp = X[p] / pl->g.dx - (1.0 - pl->g.I0x), yp = Y[p] / pl->g.dx - (1.0 - pl->g.I0y);
        mm = std::min({mm, xp, (pl->g.NX - 2) - xp, yp, (pl->g.NY - 2) - yp});
    }
    pl->min_margin = N > 0 ? (int)std::floor(mm) : 1 << 20;
    const size_t b = sizeof(double) * (size_t)std::max(1, N);
    pl->X.ensure(b);
    pl->Y.ensure(b);
    pl->nx.ensure(b);
    pl->ny.ensure(b);
    pl->ds.ensure(b);
    hipStream_t st = pl->ctx->stream;
    if (N > 0) {
        ILH_HIP(hipMemcpyAsync(pl->X.p, X, b, hipMemcpyHostToDevice, st));
        ILH_HIP(hipMemcpyAsync(pl->Y.p, Y, b, hipMemcpyHostToDevice, st));
        ILH_HIP(hipMemcpyAsync(pl->nx.p, nx, b, hipMemcpyHostToDevice, st));
        ILH_HIP(hipMemcpyAsync(pl->ny.p, ny, b, hipMemcpyHostToDevice, st));
        ILH_HIP(hipMemcpyAsync(pl->ds.p, ds, b, hipMemcpyHostToDevice, st));
    }
    ILH_HIP(hipStreamSynchronize(st));
}

extern "C" int ilh_plan_create(ilh_ctx* ctx, const ilh_grid* grid, int npts, const double* X, const double* Y,
                               const double* nx, const double* ny, const double* ds, int ddf, int scaling,
                               ilh_plan** out)
{
    ilh_plan* pl = nullptr;
    int rc = guarded([&] {
        if (!ctx || !grid || !out) fail(ILH_ERR_ARG, "null argument");
        if (grid->NX < 4 || grid->NY < 4) fail(ILH_ERR_DIM, "grid must have NX, NY >= 4");
        if (!(grid->dx > 0.0) || !std::isfinite(grid->dx)) fail(ILH_ERR_ARG, "cell size must be positive");
        if (npts < 0) fail(ILH_ERR_DIM, "negative point count");
        if (ddf < ILH_DDF_YANG3 || ddf > ILH_DDF_WITCHHAT) fail(ILH_ERR_ARG, "unsupported DDF kind");
        if (scaling != ILH_SCALING_GRID) fail(ILH_ERR_ARG, "only GridScaling is supported");
        ILH_HIP(hipSetDevice(ctx->device));
        pl = new ilh_plan();
        pl->ctx = ctx;
        pl->g = Grid{grid->NX, grid->NY, grid->I0x, grid->I0y, grid->dx};
        pl->N = npts;
        pl->ddf = ddf;
        load_points(pl, X, Y, nx, ny, ds);
        ctx_retain(ctx);
        *out = pl;
    });
    if (rc != ILH_OK && pl) {
        drop_caches(pl, true);
        delete pl;
    }
    return rc;
}

extern "C" int ilh_plan_update_points(ilh_plan* pl, const double* X, const double* Y, const double* nx,
                                      const double* ny, const double* ds)
{
    return guarded([&] {
        if (!pl) fail(ILH_ERR_ARG, "null plan");
        ILH_HIP(hipSetDevice(pl->ctx->device));
        ILH_HIP(hipStreamSynchronize(pl->ctx->stream));
        drop_caches(pl);
        load_points(pl, X, Y, nx, ny, ds);
    });
}

void ilh::plan_retain(ilh_plan* pl) { pl->refs.fetch_add(1); }

void ilh::plan_release(ilh_plan* pl)
{
    if (pl->refs.fetch_sub(1) > 1) return;
    ilh_ctx* ctx = pl->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    drop_caches(pl, true);
    delete pl->mov;
    delete pl;
    ctx_release(ctx);
}

// the caller's handle; batches and heat integrators still using the plan
// keep it alive until they are destroyed
extern "C" int ilh_plan_destroy(ilh_plan* pl)
{
    return guarded([&] {
        if (pl) plan_release(pl);
    });
}

extern "C" int ilh_plan_npts(const ilh_plan* pl) { return pl ? pl->N : -1; }

// ------------------------------------------------------------- staging
namespace {
size_t grid_len(const Grid& g, int kind)
{
    switch (kind) {
    case ILH_NODES_PRIMAL: return g.n(GK_PRIMAL);
    case ILH_NODES_DUAL: return g.n(GK_DUAL);
    case ILH_EDGES_PRIMAL: return g.n_edges();
    }
    fail(ILH_ERR_ARG, "unknown grid kind");
}

}  // namespace

namespace ilh {

void need_margin(ilh_plan* pl, const char* what)
{
    if (pl->N > 0 && pl->min_margin < 5)
        fail(ILH_ERR_RANGE, std::string(what) + ": surface points must lie at least 5 cells inside the grid");
}

// component grids touched by a scatter op, in storage order of its output
int op_components(int op, int* gks)
{
    switch (op) {
    case OP_R_PRIMAL: case OP_DS: gks[0] = GK_PRIMAL; return 1;
    case OP_R_DUAL: case OP_CST: case OP_CXT: case OP_CVT: gks[0] = GK_DUAL; return 1;
    case OP_RNSYM: gks[0] = GK_T_XX; gks[1] = GK_T_XY; gks[2] = GK_T_YX; gks[3] = GK_T_YY; return 4;
    default: gks[0] = GK_EDGE_U; gks[1] = GK_EDGE_V; return 2;
    }
}

template <class T>
static void scatter_op_t(ilh_plan* pl, int op, const T* amp, size_t amp_stride, int nb, T* out, size_t out_stride,
                         const T* base = nullptr, double cb = 1.0)
{
    const Term* t = get_terms(pl, op);
    int gks[4];
    const int nc = op_components(op, gks);
    if (!pl->bins_ok[op]) {
        for (int c = 0; c < nc; ++c) build_bins(pl, t, pl->N * op_terms_per_point(op), gks[c], pl->bins[op][c]);
        pl->bins_ok[op] = true;
    }
    size_t off = 0;
    for (int c = 0; c < nc; ++c) {
        scatter(pl, pl->bins[op][c], amp, amp_stride, nb, out + off, out_stride, base ? base + off : nullptr, cb);
        off += pl->g.n(gks[c]);
    }
}

void scatter_op_multi(ilh_plan* pl, int op, const double* amp, size_t amp_stride, int nb, double* out,
                      size_t out_stride)
{
    scatter_op_t(pl, op, amp, amp_stride, nb, out, out_stride);
}

void scatter_op(ilh_plan* pl, int op, const double* amp, double* out) { scatter_op_t(pl, op, amp, 0, 1, out, 0); }
void scatter_op_add(ilh_plan* pl, int op, const double* amp, const double* base, double c, double* out)
{
    scatter_op_t(pl, op, amp, 0, 1, out, 0, base, c);
}
void scatter_op(ilh_plan* pl, int op, const float* amp, float* out) { scatter_op_t(pl, op, amp, 0, 1, out, 0); }

void gather_op_multi(ilh_plan* pl, int op, const double* field, size_t field_stride, int nb, double* out,
                     size_t out_stride)
{
    gather(pl, get_terms(pl, op), op_terms_per_point(op), field, field_stride, nb, out, out_stride);
}

void gather_op(ilh_plan* pl, int op, const double* field, double* out)
{
    gather_op_multi(pl, op, field, 0, 1, out, 0);
}

void gather_op(ilh_plan* pl, int op, const float* field, float* out)
{
    gather(pl, get_terms(pl, op), op_terms_per_point(op), field, 0, 1, out, 0);
}

}  // namespace ilh

namespace {

int kind_to_gk(int kind)
{
    if (kind == ILH_NODES_PRIMAL) return GK_PRIMAL;
    if (kind == ILH_NODES_DUAL) return GK_DUAL;
    fail(ILH_ERR_ARG, "grid kind must be Nodes{Primal} or Nodes{Dual}");
}

// generic surface -> grid and grid -> surface entry points (FP64, or FP32 data
// of a Float32 cache)
template <class T>
int surf_to_grid(ilh_plan* pl, int op, size_t nin, size_t nout, T* out, const T* f, int mem, bool derivative,
                 const char* name)
{
    return guarded([&] {
        Io io(pl, mem);
        if (derivative) need_margin(pl, name);
        const T* d_in = io.in(f, nin, pl->io_a);
        T* d_out = io.out(out, nout, pl->io_b);
        if (pl->N == 0) fill_zero(pl->ctx->stream, d_out, nout);
        else scatter_op(pl, op, d_in, d_out);
        io.back(out, d_out, nout);
        io.done();
    });
}

template <class T>
int grid_to_surf(ilh_plan* pl, int op, size_t nin, size_t nout, T* out, const T* gfield, int mem, bool derivative,
                 const char* name)
{
    return guarded([&] {
        Io io(pl, mem);
        if (derivative) need_margin(pl, name);
        const T* d_in = io.in(gfield, nin, pl->io_a);
        T* d_out = io.out(out, nout, pl->io_b);
        if (pl->N > 0) gather_op(pl, op, d_in, d_out);
        io.back(out, d_out, nout);
        io.done();
    });
}
}  // namespace

// ------------------------------------------------- surface <-> grid ops
extern "C" int ilh_regularize(ilh_plan* pl, int kind, double* out, const double* f, int mem)
{
    if (!pl) { g_err = "null plan"; return ILH_ERR_ARG; }
    if (kind == ILH_NODES_PRIMAL)
        return surf_to_grid(pl, OP_R_PRIMAL, pl->N, pl->g.n(GK_PRIMAL), out, f, mem, false, "regularize!");
    if (kind == ILH_NODES_DUAL)
        return surf_to_grid(pl, OP_R_DUAL, pl->N, pl->g.n(GK_DUAL), out, f, mem, false, "regularize!");
    if (kind == ILH_EDGES_PRIMAL)
        return surf_to_grid(pl, OP_R_EDGES, 2 * (size_t)pl->N, pl->g.n_edges(), out, f, mem, false, "regularize!");
    g_err = "unknown grid kind";
    return ILH_ERR_ARG;
}

extern "C" int ilh_interpolate(ilh_plan* pl, int kind, double* out, const double* gf, int mem)
{
    if (!pl) { g_err = "null plan"; return ILH_ERR_ARG; }
    if (kind == ILH_NODES_PRIMAL)
        return grid_to_surf(pl, OP_E_PRIMAL, pl->g.n(GK_PRIMAL), pl->N, out, gf, mem, false, "interpolate!");
    if (kind == ILH_NODES_DUAL)
        return grid_to_surf(pl, OP_E_DUAL, pl->g.n(GK_DUAL), pl->N, out, gf, mem, false, "interpolate!");
    if (kind == ILH_EDGES_PRIMAL)
        return grid_to_surf(pl, OP_E_EDGES, pl->g.n_edges(), 2 * (size_t)pl->N, out, gf, mem, false, "interpolate!");
    g_err = "unknown grid kind";
    return ILH_ERR_ARG;
}

#define PL_CHECK if (!pl) { g_err = "null plan"; return ILH_ERR_ARG; }

extern "C" int ilh_regularize_normal(ilh_plan* pl, double* q, const double* f, int mem)
{
    PL_CHECK
    return surf_to_grid(pl, OP_RN, pl->N, pl->g.n_edges(), q, f, mem, false, "regularize_normal!");
}
extern "C" int ilh_normal_interpolate(ilh_plan* pl, double* vn, const double* q, int mem)
{
    PL_CHECK
    return grid_to_surf(pl, OP_NI, pl->g.n_edges(), pl->N, vn, q, mem, false, "normal_interpolate!");
}
extern "C" int ilh_regularize_normal_cross(ilh_plan* pl, double* q, const double* f, int mem)
{
    PL_CHECK
    return surf_to_grid(pl, OP_RNX, pl->N, pl->g.n_edges(), q, f, mem, false, "regularize_normal_cross!");
}
extern "C" int ilh_normal_cross_interpolate(ilh_plan* pl, double* vn, const double* q, int mem)
{
    PL_CHECK
    return grid_to_surf(pl, OP_NXI, pl->g.n_edges(), pl->N, vn, q, mem, false, "normal_cross_interpolate!");
}
extern "C" int ilh_surface_divergence(ilh_plan* pl, double* th, const double* f, int mem)
{
    PL_CHECK
    return surf_to_grid(pl, OP_DS, pl->N, pl->g.n(GK_PRIMAL), th, f, mem, true, "surface_divergence!");
}
extern "C" int ilh_surface_grad(ilh_plan* pl, double* vn, const double* phi, int mem)
{
    PL_CHECK
    return grid_to_surf(pl, OP_GS, pl->g.n(GK_PRIMAL), pl->N, vn, phi, mem, true, "surface_grad!");
}
extern "C" int ilh_surface_curl(ilh_plan* pl, double* w, const double* f, int mem)
{
    PL_CHECK
    return surf_to_grid(pl, OP_CST, pl->N, pl->g.n(GK_DUAL), w, f, mem, true, "surface_curl!");
}
extern "C" int ilh_surface_curl_adj(ilh_plan* pl, double* vn, const double* s, int mem)
{
    PL_CHECK
    return grid_to_surf(pl, OP_CS, pl->g.n(GK_DUAL), pl->N, vn, s, mem, true, "surface_curl!");
}
extern "C" int ilh_surface_curl_cross(ilh_plan* pl, double* w, const double* f, int mem)
{
    PL_CHECK
    return surf_to_grid(pl, OP_CXT, pl->N, pl->g.n(GK_DUAL), w, f, mem, true, "surface_curl_cross!");
}
extern "C" int ilh_surface_curl_cross_adj(ilh_plan* pl, double* vn, const double* s, int mem)
{
    PL_CHECK
    return grid_to_surf(pl, OP_CXS, pl->g.n(GK_DUAL), pl->N, vn, s, mem, true, "surface_curl_cross!");
}
extern "C" int ilh_surface_grad_cross(ilh_plan* pl, double* vn, const double* phi, int mem)
{
    PL_CHECK
    return grid_to_surf(pl, OP_GXS, pl->g.n(GK_PRIMAL), pl->N, vn, phi, mem, true, "surface_grad_cross!");
}

// vector cache (SurfaceVectorCache, src/cache.jl:184-201): VectorData is
// [u; v] (2N), EdgeGradient [dudx; dudy; dvdx; dvdy]
extern "C" int ilh_surface_curl_vector(ilh_plan* pl, double* w, const double* v, int mem)
{
    PL_CHECK
    return surf_to_grid(pl, OP_CVT, 2 * (size_t)pl->N, pl->g.n(GK_DUAL), w, v, mem, true, "surface_curl!");
}
extern "C" int ilh_surface_curl_vector_adj(ilh_plan* pl, double* v, const double* s, int mem)
{
    PL_CHECK
    return grid_to_surf(pl, OP_CVS, pl->g.n(GK_DUAL), 2 * (size_t)pl->N, v, s, mem, true, "surface_curl!");
}
extern "C" int ilh_surface_divergence_symm(ilh_plan* pl, double* q, const double* dv, int mem)
{
    PL_CHECK
    return surf_to_grid(pl, OP_DSYM, 2 * (size_t)pl->N, pl->g.n_edges(), q, dv, mem, true,
                        "surface_divergence_symm!");
}
extern "C" int ilh_surface_grad_symm(ilh_plan* pl, double* tau, const double* q, int mem)
{
    PL_CHECK
    return grid_to_surf(pl, OP_GSYM, pl->g.n_edges(), 2 * (size_t)pl->N, tau, q, mem, true, "surface_grad_symm!");
}
extern "C" int ilh_regularize_normal_symm(ilh_plan* pl, double* A, const double* v, int mem)
{
    PL_CHECK
    return surf_to_grid(pl, OP_RNSYM, 2 * (size_t)pl->N, pl->g.n_tensor(), A, v, mem, false,
                        "regularize_normal_symm!");
}
extern "C" int ilh_normal_interpolate_symm(ilh_plan* pl, double* tau, const double* A, int mem)
{
    PL_CHECK
    return grid_to_surf(pl, OP_NISYM, pl->g.n_tensor(), 2 * (size_t)pl->N, tau, A, mem, false,
                        "normal_interpolate_symm!");
}

// -------------------------------------------------------------- stencils
template <class T, class K>
static int grid_op(ilh_plan* pl, size_t nin, size_t nout, T* out, const T* in, int mem, K&& kern)
{
    return guarded([&] {
        Io io(pl, mem);
        const T* d_in = io.in(in, nin, pl->io_a);
        T* d_out = io.out(out, nout, pl->io_b);
        if (mem == ILH_MEM_DEVICE && (const void*)d_in == (const void*)d_out)
            fail(ILH_ERR_ARG, "stencil output must not alias its input");
        kern(d_out, d_in);
        io.back(out, d_out, nout);
        io.done();
    });
}

extern "C" int ilh_divergence(ilh_plan* pl, double* p, const double* q, int mem)
{
    PL_CHECK
    return grid_op(pl, pl->g.n_edges(), pl->g.n(GK_PRIMAL), p, q, mem,
                   [&](double* o, const double* i) { k_divergence(pl, o, i); });
}
extern "C" int ilh_grad(ilh_plan* pl, double* q, const double* p, int mem)
{
    PL_CHECK
    return grid_op(pl, pl->g.n(GK_PRIMAL), pl->g.n_edges(), q, p, mem,
                   [&](double* o, const double* i) { k_grad(pl, o, i); });
}
extern "C" int ilh_curl_nodes_to_edges(ilh_plan* pl, double* q, const double* s, int mem)
{
    PL_CHECK
    return grid_op(pl, pl->g.n(GK_DUAL), pl->g.n_edges(), q, s, mem,
                   [&](double* o, const double* i) { k_curl_n2e(pl, o, i); });
}
extern "C" int ilh_curl_edges_to_nodes(ilh_plan* pl, double* w, const double* q, int mem)
{
    PL_CHECK
    return grid_op(pl, pl->g.n_edges(), pl->g.n(GK_DUAL), w, q, mem,
                   [&](double* o, const double* i) { k_curl_e2n(pl, o, i); });
}
extern "C" int ilh_laplacian(ilh_plan* pl, int kind, double* out, const double* in, int mem)
{
    PL_CHECK
    int gk;
    if (int rc = guarded([&] { gk = kind_to_gk(kind); })) return rc;
    return grid_op(pl, pl->g.n(gk), pl->g.n(gk), out, in, mem,
                   [&](double* o, const double* i) { k_laplacian(pl, gk, o, i); });
}

extern "C" int ilh_inverse_laplacian(ilh_plan* pl, int kind, double* out, const double* in, int mem)
{
    PL_CHECK
    return guarded([&] {
        Io io(pl, mem);
        if (kind == ILH_EDGES_PRIMAL) {
            const size_t nu = pl->g.n(GK_EDGE_U), n = pl->g.n_edges();
            const double* d_in = io.in(in, n, pl->io_a);
            double* d_out = io.out(out, n, pl->io_b);
            conv_dense(pl, GK_EDGE_U, d_in, d_out);
            conv_dense(pl, GK_EDGE_V, d_in + nu, d_out + nu);
            io.back(out, d_out, n);
            io.done();
            return;
        }
        const int gk = kind_to_gk(kind);
        const size_t n = pl->g.n(gk);
        const double* d_in = io.in(in, n, pl->io_a);
        double* d_out = io.out(out, n, pl->io_b);
        conv_dense(pl, gk, d_in, d_out);
        io.back(out, d_out, n);
        io.done();
    });
}

// ------------------------------------------------- FP32 (Float32 caches)
// The reference caches take dtype = Float32 (src/cache.jl:164-201,
// `dtype` keyword; the Laplacian plan of src/cache.jl:321-324 follows it):
// the same operators on Float32 grid and surface data.  Stencils compute in
// FP32; regularize/interpolate accumulate the FP64 DDF weights and round once;
// inverse_laplacian! reads and writes FP32 fields around FP64 transforms.
extern "C" int ilh_regularize_f32(ilh_plan* pl, int kind, float* out, const float* f, int mem)
{
    if (!pl) { g_err = "null plan"; return ILH_ERR_ARG; }
    if (kind == ILH_NODES_PRIMAL)
        return surf_to_grid(pl, OP_R_PRIMAL, pl->N, pl->g.n(GK_PRIMAL), out, f, mem, false, "regularize!");
    if (kind == ILH_NODES_DUAL)
        return surf_to_grid(pl, OP_R_DUAL, pl->N, pl->g.n(GK_DUAL), out, f, mem, false, "regularize!");
    if (kind == ILH_EDGES_PRIMAL)
        return surf_to_grid(pl, OP_R_EDGES, 2 * (size_t)pl->N, pl->g.n_edges(), out, f, mem, false, "regularize!");
    g_err = "unknown grid kind";
    return ILH_ERR_ARG;
}

extern "C" int ilh_interpolate_f32(ilh_plan* pl, int kind, float* out, const float* gf, int mem)
{
    if (!pl) { g_err = "null plan"; return ILH_ERR_ARG; }
    if (kind == ILH_NODES_PRIMAL)
        return grid_to_surf(pl, OP_E_PRIMAL, pl->g.n(GK_PRIMAL), pl->N, out, gf, mem, false, "interpolate!");
    if (kind == ILH_NODES_DUAL)
        return grid_to_surf(pl, OP_E_DUAL, pl->g.n(GK_DUAL), pl->N, out, gf, mem, false, "interpolate!");
    if (kind == ILH_EDGES_PRIMAL)
        return grid_to_surf(pl, OP_E_EDGES, pl->g.n_edges(), 2 * (size_t)pl->N, out, gf, mem, false, "interpolate!");
    g_err = "unknown grid kind";
    return ILH_ERR_ARG;
}

extern "C" int ilh_divergence_f32(ilh_plan* pl, float* p, const float* q, int mem)
{
    PL_CHECK
    return grid_op(pl, pl->g.n_edges(), pl->g.n(GK_PRIMAL), p, q, mem,
                   [&](float* o, const float* i) { k_divergence(pl, o, i); });
}
extern "C" int ilh_grad_f32(ilh_plan* pl, float* q, const float* p, int mem)
{
    PL_CHECK
    return grid_op(pl, pl->g.n(GK_PRIMAL), pl->g.n_edges(), q, p, mem,
                   [&](float* o, const float* i) { k_grad(pl, o, i); });
}
extern "C" int ilh_curl_nodes_to_edges_f32(ilh_plan* pl, float* q, const float* s, int mem)
{
    PL_CHECK
    return grid_op(pl, pl->g.n(GK_DUAL), pl->g.n_edges(), q, s, mem,
                   [&](float* o, const float* i) { k_curl_n2e(pl, o, i); });
}
extern "C" int ilh_curl_edges_to_nodes_f32(ilh_plan* pl, float* w, const float* q, int mem)
{
    PL_CHECK
    return grid_op(pl, pl->g.n_edges(), pl->g.n(GK_DUAL), w, q, mem,
                   [&](float* o, const float* i) { k_curl_e2n(pl, o, i); });
}
extern "C" int ilh_laplacian_f32(ilh_plan* pl, int kind, float* out, const float* in, int mem)
{
    PL_CHECK
    int gk;
    if (int rc = guarded([&] { gk = kind_to_gk(kind); })) return rc;
    return grid_op(pl, pl->g.n(gk), pl->g.n(gk), out, in, mem,
                   [&](float* o, const float* i) { k_laplacian(pl, gk, o, i); });
}
extern "C" int ilh_inverse_laplacian_f32(ilh_plan* pl, int kind, float* out, const float* in, int mem)
{
    PL_CHECK
    return guarded([&] {
        const int gk = kind_to_gk(kind);
        Io io(pl, mem);
        const size_t n = pl->g.n(gk);
        const float* d_in = io.in(in, n, pl->io_a);
        float* d_out = io.out(out, n, pl->io_b);
        conv_dense(pl, gk, d_in, d_out);
        io.back(out, d_out, n);
        io.done();
    });
}

// --------------------------------------------------------------- Schur
namespace {
__global__ void k_rhs_sign(double* B, int n, double sign)   // B = sign K - I/2, K in B (n x n)
{
    const size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (k >= (size_t)n * n) return;
    const int r = (int)(k % n), c = (int)(k / n);
    B[k] = sign * B[k] - (r == c ? 0.5 : 0.0);
}

__global__ void k_rhs_dir(double* B, int n)   // B = [K - I/2, -K - I/2], K in B[:, 0:n]
{
    const size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (k >= (size_t)n * n) return;
    const int r = (int)(k % n), c = (int)(k / n);
    const double v = B[k];
    const double h = (r == c) ? 0.5 : 0.0;
    B[k] = v - h;
    B[k + (size_t)n * n] = -v - h;
}

__global__ void k_neg(const double* a, double* b, size_t n)
{
    const size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (k < n) b[k] = -a[k];
}

__global__ void k_term_shift(const Term* src, int n, int ai_off, Term* dst)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    Term t = src[k];
    t.ai += ai_off;
    dst[k] = t;
}

// create_CL2invCT (src/matrix_operators.jl:95-125): -C_s L^-2 C_s^T on
// VectorData (2N x 2N, column-major), column by column as the reference
// builds it (the twice-applied inverse Laplacian of the grid is not a
// translation-invariant kernel, so no table assembly): batches of unit
// columns, C_s^T by one scatter, L^-1 twice by batched FFT convolutions on
// Nodes{Dual}, C_s by one gather.
__global__ void k_unit_columns(double* __restrict__ a, int n, int c0, int m)
{
    const size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (k >= (size_t)n * m) return;
    const int col = (int)(k / n), row = (int)(k - (size_t)col * n);
    a[k] = row == c0 + col ? 1.0 : 0.0;
}

void build_cl2invct(ilh_plan* pl, double scale, double* K)
{
    const int n2 = 2 * pl->N;
    const size_t nd = pl->g.n(GK_DUAL);
    hipStream_t st = pl->ctx->stream;
    // batch width: at most 32 columns and ~1 GB of fields and spectra
    const size_t per_col = 2 * nd * sizeof(double) + 4 * nd * sizeof(double2);
    const int nb = (int)std::max<size_t>(1, std::min<size_t>({(size_t)32, (size_t)n2, ((size_t)1 << 30) / per_col}));
    DevBuf amp, f1, f2;
    amp.ensure(sizeof(double) * (size_t)n2 * nb);
    f1.ensure(sizeof(double) * nd * nb);
    f2.ensure(sizeof(double) * nd * nb);
    for (int c0 = 0; c0 < n2; c0 += nb) {
        const int m = std::min(nb, n2 - c0);
        const size_t na = (size_t)n2 * m;
        k_unit_columns<<<(unsigned)((na + 255) / 256), 256, 0, st>>>(amp.as<double>(), n2, c0, m);
        ILH_HIP(hipGetLastError());
        scatter_op_multi(pl, OP_CVT, amp.as<double>(), n2, m, f1.as<double>(), nd);
        conv_dense_multi(pl, GK_DUAL, f1.as<double>(), m, f2.as<double>());
        conv_dense_multi(pl, GK_DUAL, f2.as<double>(), m, f1.as<double>());
        gather_op_multi(pl, OP_CVS, f1.as<double>(), nd, m, K + (size_t)n2 * c0, n2);
    }
    axpby(st, K, K, -scale, 0.0, (size_t)n2 * n2);
    ILH_HIP(hipStreamSynchronize(st));   // the batch buffers are released on return
}

// Ka = E C L^-2 C^T D_s^symm (2N x 2N, column-major): the surface velocity
// of the streamfunction L^-2 C^T D_s^symm [v] that the Stokes solve's psi*
// is (with the sign), column by column like create_CL2invCT (the two curls
// are grid stencils, applied per column).
void build_stokes_ka(ilh_plan* pl, double* K)
{
    const int n2 = 2 * pl->N;
    const Grid& g = pl->g;
    const size_t nd = g.n(GK_DUAL), ne = g.n_edges();
    hipStream_t st = pl->ctx->stream;
    const size_t per_col = (2 * nd + ne) * sizeof(double) + 4 * nd * sizeof(double2);
    const int nb = (int)std::max<size_t>(1, std::min<size_t>({(size_t)32, (size_t)n2, ((size_t)1 << 30) / per_col}));
    DevBuf amp, fe, f1, f2;
    amp.ensure(sizeof(double) * (size_t)n2 * nb);
    fe.ensure(sizeof(double) * ne * nb);
    f1.ensure(sizeof(double) * nd * nb);
    f2.ensure(sizeof(double) * nd * nb);
    for (int c0 = 0; c0 < n2; c0 += nb) {
        const int m = std::min(nb, n2 - c0);
        const size_t na = (size_t)n2 * m;
        k_unit_columns<<<(unsigned)((na + 255) / 256), 256, 0, st>>>(amp.as<double>(), n2, c0, m);
        ILH_HIP(hipGetLastError());
        scatter_op_multi(pl, OP_DSYM, amp.as<double>(), n2, m, fe.as<double>(), ne);
        for (int k = 0; k < m; ++k) k_curl_e2n(pl, f1.as<double>() + nd * k, fe.as<double>() + ne * k);
        conv_dense_multi(pl, GK_DUAL, f1.as<double>(), m, f2.as<double>());
        conv_dense_multi(pl, GK_DUAL, f2.as<double>(), m, f1.as<double>());
        for (int k = 0; k < m; ++k) k_curl_n2e(pl, fe.as<double>() + ne * k, f1.as<double>() + nd * k);
        gather_op_multi(pl, OP_E_EDGES, fe.as<double>(), ne, m, K + (size_t)n2 * c0, n2);
    }
    ILH_HIP(hipStreamSynchronize(st));   // the batch buffers are released on return
}

// create_GLinvD_symm = -G_s^symm L^-1 D_s^symm (2N x 2N, column-major), column
// batches like create_CL2invCT: unit VectorData columns, D_s^symm onto
// Edges{Primal}, the inverse Laplacian of each edge component on its own grid
// (the components are gathered into contiguous batches for the batched
// convolution), G_s^symm back to VectorData.
void build_glinvd_symm(ilh_plan* pl, double scale, double* K)
{
    const int n2 = 2 * pl->N;
    const Grid& g = pl->g;
    const size_t nu = g.n(GK_EDGE_U), nv = g.n(GK_EDGE_V), ne = g.n_edges();
    hipStream_t st = pl->ctx->stream;
    const size_t per_col = 3 * ne * sizeof(double) + 4 * std::max(nu, nv) * sizeof(double2);
    const int nb = (int)std::max<size_t>(1, std::min<size_t>({(size_t)32, (size_t)n2, ((size_t)1 << 30) / per_col}));
    DevBuf amp, fe, c1, c2;
    amp.ensure(sizeof(double) * (size_t)n2 * nb);
    fe.ensure(sizeof(double) * ne * nb);
    c1.ensure(sizeof(double) * std::max(nu, nv) * nb);
    c2.ensure(sizeof(double) * std::max(nu, nv) * nb);
    for (int c0 = 0; c0 < n2; c0 += nb) {
        const int m = std::min(nb, n2 - c0);
        const size_t na = (size_t)n2 * m;
        k_unit_columns<<<(unsigned)((na + 255) / 256), 256, 0, st>>>(amp.as<double>(), n2, c0, m);
        ILH_HIP(hipGetLastError());
        scatter_op_multi(pl, OP_DSYM, amp.as<double>(), n2, m, fe.as<double>(), ne);
        for (int comp = 0; comp < 2; ++comp) {
            const size_t off = comp ? nu : 0, n = comp ? nv : nu;
            const int gk = comp ? GK_EDGE_V : GK_EDGE_U;
            ILH_HIP(hipMemcpy2DAsync(c1.as<double>(), sizeof(double) * n, fe.as<double>() + off, sizeof(double) * ne,
                                     sizeof(double) * n, m, hipMemcpyDeviceToDevice, st));
            conv_dense_multi(pl, gk, c1.as<double>(), m, c2.as<double>());
            ILH_HIP(hipMemcpy2DAsync(fe.as<double>() + off, sizeof(double) * ne, c2.as<double>(), sizeof(double) * n,
                                     sizeof(double) * n, m, hipMemcpyDeviceToDevice, st));
        }
        gather_op_multi(pl, OP_GSYM, fe.as<double>(), ne, m, K + (size_t)n2 * c0, n2);
    }
    axpby(st, K, K, -scale, 0.0, (size_t)n2 * n2);
    ILH_HIP(hipStreamSynchronize(st));   // the batch buffers are released on return
}

// order of the matrix a Schur kind builds
int schur_dim(const ilh_plan* pl, int kind)
{
    return kind == ILH_SCHUR_CL2INVCT || kind == ILH_SCHUR_GLINVD_SYMM ? 2 * pl->N : pl->N;
}

void build_schur(ilh_plan* pl, int kind, double scale, double* K)
{
    const double dx2 = pl->g.dx * pl->g.dx;
    switch (kind) {
    case ILH_SCHUR_CL2INVCT: build_cl2invct(pl, scale, K); break;
    case ILH_SCHUR_GLINVD_SYMM: build_glinvd_symm(pl, scale, K); break;
    case ILH_SCHUR_RTLINVR: assemble(pl, OP_E_PRIMAL, OP_R_PRIMAL, GK_PRIMAL, -scale * dx2, false, K); break;
    case ILH_SCHUR_CLINVCT: assemble(pl, OP_CS, OP_CST, GK_DUAL, -scale * dx2, false, K); break;
    case ILH_SCHUR_GLINVD: assemble(pl, OP_GS, OP_DS, GK_PRIMAL, -scale * dx2, false, K); break;
    case ILH_SCHUR_NRTRN: assemble(pl, OP_NI, OP_RN, GK_EDGE_U, scale, true, K); break;
    default: fail(ILH_ERR_ARG, "unknown Schur kind");
    }
}

// create_RTLinvR = M diag(ds) with M = -E L^-1 E^T (the (E, E) assembly of
// the LGF, exactly symmetric): M factored by the deflated Cholesky; false
// when it meets a non-positive pivot
bool sym_rtlinvr(ilh_plan* pl, SymFactor& F)
{
    const int N = pl->N;
    F.A.ensure(sizeof(double) * (size_t)N * N);
    assemble(pl, OP_E_PRIMAL, OP_E_PRIMAL, GK_PRIMAL, -1.0, false, F.A.as<double>());
    return sym_factor(pl, F, N);
}

void factor_into(ilh_plan* pl, SolveCache* sc, int kind)
{
    const int N = schur_dim(pl, kind);
    if (kind == ILH_SCHUR_RTLINVR && sym_rtlinvr(pl, sc->sym)) {
        sc->use_sym = true;
        return;
    }
    sc->sym = SymFactor();
    sc->use_sym = false;
    sc->LU.ensure(sizeof(double) * (size_t)N * N);
    sc->piv.ensure(sizeof(int) * (size_t)std::max(1, N));
    sc->info.ensure(sizeof(int));
    build_schur(pl, kind, 1.0, sc->LU.as<double>());
    const int info = lu_factor(pl, sc->LU.as<double>(), N, sc->piv.as<int>(), sc->info.as<int>());
    if (info) fail(ILH_ERR_SINGULAR, "Schur complement is singular (zero pivot at column " + std::to_string(info) + ")");
}

// merged source terms: first op with amplitude slot 0, second with slot 1
void merge_sources(ilh_plan* pl, int op0, int op1, int gk, DevBuf& tbuf, SourceSet& src)
{
    const int N = pl->N;
    const int n0 = N * op_terms_per_point(op0), n1 = N * op_terms_per_point(op1);
    tbuf.ensure(sizeof(Term) * (size_t)std::max(1, n0 + n1));
    hipStream_t st = pl->ctx->stream;
    Term* t = tbuf.as<Term>();
    if (n0) k_term_shift<<<(n0 + 255) / 256, 256, 0, st>>>(get_terms(pl, op0), n0, 0, t);
    if (n1) k_term_shift<<<(n1 + 255) / 256, 256, 0, st>>>(get_terms(pl, op1), n1, N, t + n0);
    ILH_HIP(hipGetLastError());
    src.terms = t;
    src.amp_len = 2 * N;
    build_bins(pl, t, n0 + n1, gk, src.bins);
    build_source_pairs(pl, src);
}

// M[k] = S^-1 (sign K_D - I/2) row-major, K_D = E L^-1 D_s (sign +1 for k = 0,
// the exterior data; -1 for k = 1, the interior data), with the kept factor
static void dirichlet_matrix(ilh_plan* pl, SolveCache* sc, int k)
{
    const int N = pl->N;
    hipStream_t st = pl->ctx->stream;
    const double dx2 = pl->g.dx * pl->g.dx;
    DevBuf B;
    B.ensure(sizeof(double) * (size_t)N * N);
    assemble(pl, OP_E_PRIMAL, OP_DS, GK_PRIMAL, dx2, false, B.as<double>());
    k_rhs_sign<<<(unsigned)(((size_t)N * N + 255) / 256), 256, 0, st>>>(B.as<double>(), N, k ? -1.0 : 1.0);
    ILH_HIP(hipGetLastError());
    if (sc->use_sym) {
        // S^-1 B = diag(1 / ds) M^-1 B
        sym_solve(pl, sc->sym, B.as<double>(), N);
        scale_rows_inv(st, B.as<double>(), N, N, pl->ds.as<double>(), 1.0);
    } else {
        lu_solve_multi(pl, sc->LU.as<double>(), sc->piv.as<int>(), N, B.as<double>(), N);
    }
    const int ld = solve_ld(N);
    sc->M[k].ensure(sizeof(double) * (size_t)N * ld);
    transpose(st, B.as<double>(), sc->M[k].as<double>(), N, N, ld);
    ILH_HIP(hipStreamSynchronize(st));   // B is released on return
}

// The plan's Dirichlet solve data: the factor of S, the exterior-data matrix
// M+ and the merged sources; the interior-data matrix M- is built on the first
// call that has interior data (need_minus), from the kept factor.
SolveCache* prepare_dirichlet(ilh_plan* pl, bool need_minus = true)
{
    SolveCache* sc = pl->dir;
    if (!sc) {
        need_margin(pl, "Dirichlet solve");
        const int N = pl->N;
        sc = new SolveCache();
        try {
            if (N > 0) {
                factor_into(pl, sc, ILH_SCHUR_RTLINVR);
                dirichlet_matrix(pl, sc, 0);
                merge_sources(pl, OP_R_PRIMAL, OP_DS, GK_PRIMAL, sc->terms1, sc->src1);
                pin_source_pairs(pl, sc->src1);
                ILH_HIP(hipStreamSynchronize(pl->ctx->stream));
            }
            sc->amp1.ensure(sizeof(double) * (size_t)std::max(1, 2 * N));
        } catch (...) {
            delete sc;
            throw;
        }
        pl->dir = sc;
    }
    if (need_minus && !sc->minus_ok) {
        if (pl->N > 0) dirichlet_matrix(pl, sc, 1);
        sc->minus_ok = true;
    }
    return sc;
}

SolveCache* prepare_neumann(ilh_plan* pl)
{
    if (pl->neu) return pl->neu;
    need_margin(pl, "Neumann solve");
    const int N = pl->N;
    SolveCache* sc = new SolveCache();
    try {
        hipStream_t st = pl->ctx->stream;
        const double dx2 = pl->g.dx * pl->g.dx;
        if (N > 0) {
            factor_into(pl, sc, ILH_SCHUR_CLINVCT);
            DevBuf B;
            const size_t NN = (size_t)N * N;
            B.ensure(sizeof(double) * NN * 3);
            assemble(pl, OP_GS, OP_R_PRIMAL, GK_PRIMAL, dx2, false, B.as<double>());        // K_GR
            k_rhs_dir<<<(unsigned)((NN + 255) / 256), 256, 0, st>>>(B.as<double>(), N);
            assemble(pl, OP_GXS, OP_R_PRIMAL, GK_PRIMAL, dx2, false, B.as<double>() + 2 * NN);  // K_XR
            ILH_HIP(hipGetLastError());
            lu_solve_multi(pl, sc->LU.as<double>(), sc->piv.as<int>(), N, B.as<double>(), 3 * N);
            const int ld = solve_ld(N);
            const size_t NL = (size_t)N * ld;
            for (int k = 0; k < 3; ++k) {
                sc->M[k].ensure(sizeof(double) * NL);
                transpose(st, B.as<double>() + (size_t)k * NN, sc->M[k].as<double>(), N, N, ld);
            }
            sc->M[3].ensure(sizeof(double) * NL);
            k_neg<<<(unsigned)((NL + 255) / 256), 256, 0, st>>>(sc->M[2].as<double>(), sc->M[3].as<double>(), NL);
            ILH_HIP(hipGetLastError());
            merge_sources(pl, OP_DS, OP_R_PRIMAL, GK_PRIMAL, sc->terms1, sc->src1);
            merge_sources(pl, OP_CST, OP_CXT, GK_DUAL, sc->terms2, sc->src2);
            pin_source_pairs(pl, sc->src1);
            pin_source_pairs(pl, sc->src2);
            ILH_HIP(hipStreamSynchronize(st));
        }
        sc->amp1.ensure(sizeof(double) * (size_t)std::max(1, 2 * N));
        sc->amp2.ensure(sizeof(double) * (size_t)std::max(1, 2 * N));
    } catch (...) {
        delete sc;
        throw;
    }
    pl->neu = sc;
    return sc;
}
}  // namespace

// S^-1 of create_RTLinvR at the plan's geometry (the deflated Cholesky of
// M, or LU when that meets a non-positive pivot), row-major
void ilh::exact_inverse_rowmajor(ilh_plan* pl, double* Pd)
{
    const int N = pl->N;
    SymFactor F;
    if (sym_rtlinvr(pl, F)) {
        sym_inverse_rowmajor(pl, F, pl->ds.as<double>(), 1.0, Pd);
        return;
    }
    DevBuf LU;
    LU.ensure(sizeof(double) * (size_t)N * N);
    build_schur(pl, ILH_SCHUR_RTLINVR, 1.0, LU.as<double>());
    inverse_rowmajor(pl, LU.as<double>(), Pd);
}

extern "C" int ilh_schur_build(ilh_plan* pl, int kind, double scale, double* out, int mem)
{
    return guarded([&] {
        Io io(pl, mem);
        if (kind < 0 || kind > ILH_SCHUR_GLINVD_SYMM) fail(ILH_ERR_ARG, "unknown Schur kind");
        if (kind != ILH_SCHUR_NRTRN) need_margin(pl, "Schur build");
        const int N = schur_dim(pl, kind);
        const size_t NN = (size_t)N * N;
        double* d_out = io.out(out, NN, pl->io_b);
        if (N > 0) build_schur(pl, kind, scale, d_out);
        io.back(out, d_out, NN);
        io.done();
    });
}

extern "C" int ilh_schur_factor(ilh_plan* pl, int kind)
{
    return guarded([&] {
        Io io(pl, ILH_MEM_DEVICE);
        if (kind < 0 || kind > ILH_SCHUR_GLINVD_SYMM) fail(ILH_ERR_ARG, "unknown Schur kind");
        if (kind != ILH_SCHUR_NRTRN) need_margin(pl, "Schur factor");
        delete pl->fac[kind];
        pl->fac[kind] = nullptr;
        SolveCache* sc = new SolveCache();
        try {
            if (pl->N > 0) factor_into(pl, sc, kind);
        } catch (...) {
            delete sc;
            throw;
        }
        pl->fac[kind] = sc;
    });
}

extern "C" int ilh_schur_solve(ilh_plan* pl, int kind, double* x, const double* b, int mem)
{
    return guarded([&] {
        Io io(pl, mem);
        if (kind < 0 || kind > ILH_SCHUR_GLINVD_SYMM || !pl->fac[kind])
            fail(ILH_ERR_STATE, "ilh_schur_factor must be called first");
        const int N = schur_dim(pl, kind);
        const double* d_b = io.in(b, N, pl->io_a);
        double* d_x = io.out(x, N, pl->io_b);
        if (N > 0) {
            if (d_x != d_b)
                ILH_HIP(hipMemcpyAsync(d_x, d_b, sizeof(double) * N, hipMemcpyDeviceToDevice, pl->ctx->stream));
            SolveCache* sc = pl->fac[kind];
            if (sc->use_sym) {
                sym_solve(pl, sc->sym, d_x, 1);
                scale_rows_inv(pl->ctx->stream, d_x, N, 1, pl->ds.as<double>(), 1.0);
            } else {
                lu_solve_multi(pl, sc->LU.as<double>(), sc->piv.as<int>(), N, d_x, 1);
            }
        }
        io.back(x, d_x, N);
        io.done();
    });
}

// ---------------------------------------------------------- composites
extern "C" int ilh_dirichlet_solve(ilh_plan* pl, const double* fp, const double* fm, double* f, double* s, int mem)
{
    return guarded([&] {
        Io io(pl, mem);
        SolveCache* sc = prepare_dirichlet(pl, fm != nullptr);
        const int N = pl->N;
        const size_t ng = pl->g.n(GK_PRIMAL);
        hipStream_t st = pl->ctx->stream;
        const double* d_fp = io.in(fp, N, pl->io_a);
        const double* d_fm = fm ? io.in(fm, N, pl->io_c) : nullptr;
        double* d_f = io.out(f, ng, pl->io_b);
        double* d_s = io.out(s, N, pl->io_d);
        if (N == 0) {
            fill_zero(st, d_f, ng);
        } else {
            dirichlet_amp(st, sc->M[0].as<double>(), sc->M[1].as<double>(), solve_ld(N), d_fp, d_fm,
                          sc->amp1.as<double>(), d_s, N);
            conv_sources(pl, GK_PRIMAL, sc->src1, sc->amp1.as<double>(), d_f);
        }
        io.back(f, d_f, ng);
        io.back(s, d_s, N);
        io.done();
    });
}

// Mixed-precision form: the surface step (solve matrices, matvec) in FP64,
// the convolution that forms f in FP32
extern "C" int ilh_dirichlet_solve_mixed(ilh_plan* pl, const double* fp, const double* fm, double* f, double* s,
                                         int mem)
{
    return guarded([&] {
        Io io(pl, mem);
        SolveCache* sc = prepare_dirichlet(pl, fm != nullptr);
        const int N = pl->N;
        const size_t ng = pl->g.n(GK_PRIMAL);
        hipStream_t st = pl->ctx->stream;
        const double* d_fp = io.in(fp, N, pl->io_a);
        const double* d_fm = fm ? io.in(fm, N, pl->io_c) : nullptr;
        double* d_f = io.out(f, ng, pl->io_b);
        double* d_s = io.out(s, N, pl->io_d);
        if (N == 0) {
            fill_zero(st, d_f, ng);
        } else {
            dirichlet_amp(st, sc->M[0].as<double>(), sc->M[1].as<double>(), solve_ld(N), d_fp, d_fm,
                          sc->amp1.as<double>(), d_s, N);
            conv_sources_f32(pl, GK_PRIMAL, sc->src1, sc->amp1.as<double>(), d_f);
        }
        io.back(f, d_f, ng);
        io.back(s, d_s, N);
        io.done();
    });
}

// ilh_dirichlet_solve on device pointers with every work buffer in the
// device's static scratch array: no allocation of any kind.  The plan's solve
// data (solve matrices, source tables, LGF spectrum, twiddles) must exist,
// built by an earlier ilh_dirichlet_solve on this geometry.
extern "C" int ilh_dirichlet_solve_noalloc(ilh_plan* pl, const double* fp, const double* fm, double* f, double* s)
{
    return guarded([&] {
        if (!pl) fail(ILH_ERR_ARG, "null plan");
        ILH_HIP(hipSetDevice(pl->ctx->device));
        SolveCache* sc = pl->dir;
        if (!sc) fail(ILH_ERR_STATE, "the plan's Dirichlet solve data are not built: call ilh_dirichlet_solve once first");
        if (fm && !sc->minus_ok)
            fail(ILH_ERR_STATE, "the interior-data solve matrix is not built: call ilh_dirichlet_solve once with fb_minus");
        const int N = pl->N;
        const size_t ng = pl->g.n(GK_PRIMAL);
        if (!f || (N > 0 && (!fp || !s))) fail(ILH_ERR_ARG, "null array argument");
        hipStream_t st = pl->ctx->stream;
        if (N == 0) {
            fill_zero(st, f, ng);
            return;
        }
        unsigned char* base = (unsigned char*)scratch_ptr(pl->ctx->device);
        if (!pl->conv_ok[GK_PRIMAL]) fail(ILH_ERR_STATE, "the plan's LGF spectrum is not built yet");
        const ConvPlan& C = pl->conv[GK_PRIMAL];
        auto up = [](size_t b) { return (b + 255) / 256 * 256; };
        const size_t b_amp = up(sizeof(double) * 2 * (size_t)N);
        const size_t b_t1 = up(sizeof(double2) * (size_t)conv_ext_rows1(sc->src1) * C.H);
        const size_t b_t2 = up(sizeof(double2) * (size_t)conv_ext_rows2(C) * C.H);
        if (b_amp + b_t1 + b_t2 > NOALLOC_SCRATCH_BYTES)
            fail(ILH_ERR_ARG, "the allocation-free solve needs " + std::to_string(b_amp + b_t1 + b_t2) +
                                  " bytes of scratch; it holds at most " + std::to_string(NOALLOC_SCRATCH_BYTES));
        double* amp = (double*)base;
        double2* T1 = (double2*)(base + b_amp);
        double2* T2 = (double2*)(base + b_amp + b_t1);
        dirichlet_amp(st, sc->M[0].as<double>(), sc->M[1].as<double>(), solve_ld(N), fp, fm, amp, s, N);
        conv_sources_ext(pl, GK_PRIMAL, sc->src1, amp, f, T1, T2);
    });
}

// nrhs Dirichlet problems on one geometry (the boundary data of each a column
// of fb_plus / fb_minus, N x nrhs column-major; f (n_g x nrhs), s (N x nrhs)):
// the solve matrices are read once per group of 8 right-hand sides and each
// FFT pass runs once over all of them.  Each column equals ilh_dirichlet_solve
// on that column bit for bit.
extern "C" int ilh_dirichlet_solve_multi(ilh_plan* pl, int nrhs, const double* fp, const double* fm, double* f,
                                         double* s, int mem)
{
    return guarded([&] {
        if (!pl) fail(ILH_ERR_ARG, "null plan");
        if (nrhs < 1 || nrhs > 256) fail(ILH_ERR_ARG, "nrhs must be in [1, 256]");
        Io io(pl, mem);
        SolveCache* sc = prepare_dirichlet(pl, fm != nullptr);
        const int N = pl->N;
        const size_t ng = pl->g.n(GK_PRIMAL);
        hipStream_t st = pl->ctx->stream;
        double* d_f = io.out(f, ng * nrhs, pl->io_b);
        double* d_s = io.out(s, (size_t)N * nrhs, pl->io_d);
        if (N == 0) {
            fill_zero(st, d_f, ng * nrhs);
        } else {
            // inputs at the aligned stride ldf (the surface step's 16-byte loads)
            const int ldf = solve_ld(N);
            const hipMemcpyKind kind = mem == ILH_MEM_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
            sc->fin.ensure(sizeof(double) * (size_t)ldf * nrhs);
            ILH_HIP(hipMemcpy2DAsync(sc->fin.p, sizeof(double) * ldf, fp, sizeof(double) * N, sizeof(double) * N,
                                     nrhs, kind, st));
            const double* d_fm = nullptr;
            if (fm) {
                sc->fmin.ensure(sizeof(double) * (size_t)ldf * nrhs);
                ILH_HIP(hipMemcpy2DAsync(sc->fmin.p, sizeof(double) * ldf, fm, sizeof(double) * N,
                                         sizeof(double) * N, nrhs, kind, st));
                d_fm = sc->fmin.as<double>();
            }
            sc->ampm.ensure(sizeof(double) * 2 * (size_t)N * nrhs);
            dirichlet_amp_multi(st, sc->M[0].as<double>(), sc->M[1].as<double>(), solve_ld(N), sc->fin.as<double>(),
                                d_fm, ldf, nrhs, sc->ampm.as<double>(), d_s, N);
            conv_sources_multi(pl, GK_PRIMAL, sc->src1, sc->ampm.as<double>(), 2 * (size_t)N, nrhs, d_f);
        }
        io.back(f, d_f, ng * nrhs);
        io.back(s, d_s, (size_t)N * nrhs);
        io.done();
    });
}

extern "C" int ilh_neumann_solve(ilh_plan* pl, const double* vp, const double* vm, double* f, double* df, double* s,
                                 double* dsv, int mem)
{
    return guarded([&] {
        Io io(pl, mem);
        SolveCache* sc = prepare_neumann(pl);
        const int N = pl->N;
        const size_t ng = pl->g.n(GK_PRIMAL), nd = pl->g.n(GK_DUAL);
        hipStream_t st = pl->ctx->stream;
        const double* d_vp = io.in(vp, N, pl->io_a);
        const double* d_vm = vm ? io.in(vm, N, pl->io_c) : nullptr;
        double* d_f = io.out(f, ng, pl->io_b);
        double* d_s = io.out(s, nd, pl->scratch_a);
        double* d_df = io.out(df, N, pl->io_d);
        double* d_ds = io.out(dsv, N, pl->scratch_b);
        if (N == 0) {
            fill_zero(st, d_f, ng);
            fill_zero(st, d_s, nd);
        } else {
            neumann_amp(st, sc->M[0].as<double>(), sc->M[1].as<double>(), sc->M[2].as<double>(), sc->M[3].as<double>(),
                        solve_ld(N), d_vp, d_vm, sc->amp1.as<double>(), sc->amp2.as<double>(), d_df, d_ds, N);
            conv_sources(pl, GK_PRIMAL, sc->src1, sc->amp1.as<double>(), d_f);
            conv_sources(pl, GK_DUAL, sc->src2, sc->amp2.as<double>(), d_s);
        }
        io.back(f, d_f, ng);
        io.back(df, d_df, N);
        io.back(s, d_s, nd);
        io.back(dsv, d_ds, N);
        io.done();
    });
}

// ----------------------------------------------------------------- masks
namespace {

// Nodes{Primal} -> other grid kinds by averaging the neighbouring primal cells
// (CartesianGrids grid_interpolate!): dual node (i,j) from cells (i-1..i,
// j-1..j), u-edge (i,j) from cells (i-1..i, j), v-edge (i,j) from (i, j-1..j);
// values on the outer ring (missing cells) are 0.  comp: the primal values
// are 1 - m (the complementary mask is averaged, as the reference averages
// its primal complementary mask).
__global__ void k_mask_to(const double* __restrict__ m, int NX, int NY, int kind, int comp, double* __restrict__ out,
                          int multiply)
{
    const int np = NX - 1;
    int nx, ny;
    if (kind == ILH_NODES_PRIMAL) { nx = NX - 1; ny = NY - 1; }
    else if (kind == ILH_NODES_DUAL) { nx = NX; ny = NY; }
    else if (kind == 2) { nx = NX; ny = NY - 1; }     // u edges
    else { nx = NX - 1; ny = NY; }                    // v edges
    const int i = blockIdx.x * 64 + threadIdx.x % 64, j = blockIdx.y * 4 + threadIdx.x / 64;
    if (i >= nx || j >= ny) return;
    auto mp = [&](int a, int b) {
        const double x = m[a + (size_t)np * b];
        return comp ? 1.0 - x : x;
    };
    double v = 0.0;
    if (kind == ILH_NODES_PRIMAL) {
        v = mp(i, j);
    } else if (kind == ILH_NODES_DUAL) {
        if (i >= 1 && i <= NX - 2 && j >= 1 && j <= NY - 2)
            v = 0.25 * (mp(i - 1, j - 1) + mp(i, j - 1) + mp(i - 1, j) + mp(i, j));
    } else if (kind == 2) {
        if (i >= 1 && i <= NX - 2) v = 0.5 * (mp(i - 1, j) + mp(i, j));
    } else {
        if (j >= 1 && j <= NY - 2) v = 0.5 * (mp(i, j - 1) + mp(i, j));
    }
    double* o = out + i + (size_t)nx * j;
    *o = multiply ? *o * v : v;
}

__global__ void k_fill_value(double* a, size_t n, double v)
{
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) a[k] = v;
}

// mask(cache) = -L^-1 D_s 1 on Nodes{Primal} (src/surface_operators.jl:863-871);
// 1 everywhere without surface points
const double* primal_mask(ilh_plan* pl)
{
    if (pl->mask_gen == pl->gen) return pl->mask.as<double>();
    const size_t ng = pl->g.n(GK_PRIMAL);
    hipStream_t st = pl->ctx->stream;
    pl->mask.ensure(sizeof(double) * ng);
    double* m = pl->mask.as<double>();
    if (pl->N == 0) {
        k_fill_value<<<(unsigned)std::min<size_t>(4096, (ng + 255) / 256), 256, 0, st>>>(m, ng, 1.0);
    } else {
        need_margin(pl, "mask");
        DevBuf ones, th;
        ones.ensure(sizeof(double) * pl->N);
        th.ensure(sizeof(double) * ng);
        k_fill_value<<<(pl->N + 255) / 256, 256, 0, st>>>(ones.as<double>(), pl->N, 1.0);
        scatter_op(pl, OP_DS, ones.as<double>(), th.as<double>());
        conv_dense(pl, GK_PRIMAL, th.as<double>(), m);
        axpby(st, m, m, 0.0, -1.0, ng);
        ILH_HIP(hipStreamSynchronize(st));
    }
    pl->mask_gen = pl->gen;
    return m;
}

size_t kind_len(const Grid& g, int kind)
{
    switch (kind) {
    case ILH_NODES_PRIMAL: return g.n(GK_PRIMAL);
    case ILH_NODES_DUAL: return g.n(GK_DUAL);
    case 2: return g.n(GK_EDGE_U);
    case 3: return g.n(GK_EDGE_V);
    }
    fail(ILH_ERR_ARG, "unknown grid kind");
}

int mask_call(ilh_plan* pl, int grid_kind, int complementary, double* w, int mem, bool multiply)
{
    return guarded_call([&] {
        Io io(pl, mem);
        if (grid_kind < ILH_NODES_PRIMAL || grid_kind > ILH_EDGES_PRIMAL) fail(ILH_ERR_ARG, "unknown grid kind");
        const size_t n = grid_len(pl->g, grid_kind);
        hipStream_t st = pl->ctx->stream;
        double* d_w = io.out(w, n, pl->io_b);
        if (multiply && mem == ILH_MEM_HOST)
            ILH_HIP(hipMemcpyAsync(d_w, w, sizeof(double) * n, hipMemcpyHostToDevice, st));
        const double* m = primal_mask(pl);
        const int NX = pl->g.NX, NY = pl->g.NY;
        auto run = [&](int k, double* out) {
            int nx, ny;
            if (k == ILH_NODES_PRIMAL) { nx = NX - 1; ny = NY - 1; }
            else if (k == ILH_NODES_DUAL) { nx = NX; ny = NY; }
            else if (k == 2) { nx = NX; ny = NY - 1; }
            else { nx = NX - 1; ny = NY; }
            dim3 grid((nx + 63) / 64, (ny + 3) / 4);
            k_mask_to<<<grid, 256, 0, st>>>(m, NX, NY, k, complementary ? 1 : 0, out, multiply ? 1 : 0);
            ILH_HIP(hipGetLastError());
        };
        if (grid_kind == ILH_EDGES_PRIMAL) {
            run(2, d_w);
            run(3, d_w + kind_len(pl->g, 2));
        } else {
            run(grid_kind, d_w);
        }
        io.back(w, d_w, n);
        io.done();
    });
}

}  // namespace

extern "C" int ilh_mask(ilh_plan* pl, int grid_kind, int complementary, double* out, int mem)
{
    if (!pl) { g_err = "null plan"; return ILH_ERR_ARG; }
    return mask_call(pl, grid_kind, complementary, out, mem, false);
}

extern "C" int ilh_mask_multiply(ilh_plan* pl, int grid_kind, int complementary, double* w, int mem)
{
    if (!pl) { g_err = "null plan"; return ILH_ERR_ARG; }
    return mask_call(pl, grid_kind, complementary, w, mem, true);
}

// ----------------------------------------------------------- surface filter
namespace {

// C = Et R (create_surface_filter, src/matrix_operators.jl:254-267): Et is the
// interpolation with each grid value divided by (R 1) there, so that C keeps
// constants and the surface integral (sum_p ds_p (C s)_p = sum_q ds_q s_q).
// One thread per point p and its interpolation term TE[p tstride + toff] on
// the component grid of the bins: its 4 x 4 support cells in (row, x) order,
// each cell's contributions (amplitude slots q with R weight cw) in the bins'
// order, into row t.ai - off of C (column-major n x n, zeroed by the caller;
// slots are taken relative to off too: one diagonal block of a larger C).
__global__ void k_filter_build(const Term* __restrict__ TE, int tstride, int toff, int N,
                               const int* __restrict__ cell_ptr, const int4* __restrict__ cells,
                               const int* __restrict__ cai, const double* __restrict__ cw, int ny, int n, int off,
                               double* __restrict__ C)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N) return;
    const Term& t = TE[(size_t)p * tstride + toff];
    for (int b = 0; b < 5; ++b) {
        const int j = t.by + b;
        if (t.wy[b] == 0.0 || j < 0 || j >= ny) continue;
        const int lo = cell_ptr[j], hi = cell_ptr[j + 1];
        for (int a = 0; a < 5; ++a) {
            if (t.wx[a] == 0.0) continue;
            const int x = t.bx + a;
            int l = lo, h = hi;   // the cell at x (cells of a row ascend in x)
            while (l < h) {
                const int m = (l + h) >> 1;
                if (cells[m].x < x) l = m + 1; else h = m;
            }
            if (l >= hi || cells[l].x != x) continue;
            const int4 cl = cells[l];
            double r1 = 0.0;
            for (int e = cl.y; e < cl.z; ++e) r1 += cw[e];
            if (r1 == 0.0) continue;
            const double phi = t.coef * t.wx[a] * t.wy[b] / r1;
            for (int e = cl.y; e < cl.z; ++e) C[(t.ai - off) + (size_t)n * (cai[e] - off)] += phi * cw[e];
        }
    }
}

// g <- g / (R 1) where R 1 != 0, else 0
__global__ void k_filter_weight(double* __restrict__ g, const double* __restrict__ r1, size_t n)
{
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
        g[k] = r1[k] != 0.0 ? g[k] / r1[k] : 0.0;
}

const double* filter_r1(ilh_plan* pl, DevBuf& ones)
{
    // R 1 on Nodes{Primal}, cached per geometry in the mask buffer's sibling
    const size_t ng = pl->g.n(GK_PRIMAL);
    if (pl->filt_gen != pl->gen) {
        pl->filt_r1.ensure(sizeof(double) * ng);
        ones.ensure(sizeof(double) * std::max(1, pl->N));
        k_fill_value<<<(std::max(1, pl->N) + 255) / 256, 256, 0, pl->ctx->stream>>>(ones.as<double>(), pl->N, 1.0);
        scatter_op(pl, OP_R_PRIMAL, ones.as<double>(), pl->filt_r1.as<double>());
        pl->filt_gen = pl->gen;
    }
    return pl->filt_r1.as<double>();
}

}  // namespace

extern "C" int ilh_surface_filter_build(ilh_plan* pl, double* C, int mem)
{
    return guarded([&] {
        if (!pl) fail(ILH_ERR_ARG, "null plan");
        Io io(pl, mem);
        const int N = pl->N;
        const size_t nn = (size_t)N * N;
        double* d_C = io.out(C, nn, pl->io_b);
        hipStream_t st = pl->ctx->stream;
        if (N > 0) {
            fill_zero(st, d_C, nn);
            const Term* te = get_terms(pl, OP_E_PRIMAL);
            const Term* tr = get_terms(pl, OP_R_PRIMAL);
            if (!pl->bins_ok[OP_R_PRIMAL]) {
                build_bins(pl, tr, N, GK_PRIMAL, pl->bins[OP_R_PRIMAL][0]);
                pl->bins_ok[OP_R_PRIMAL] = true;
            }
            const RowBins& b = pl->bins[OP_R_PRIMAL][0];
            k_filter_build<<<(N + 127) / 128, 128, 0, st>>>(te, 1, 0, N, b.cell_ptr, b.cells, b.cai, b.cw, b.ny, N, 0,
                                                            d_C);
            ILH_HIP(hipGetLastError());
        }
        io.back(C, d_C, nn);
        io.done();
    });
}

extern "C" int ilh_surface_filter_apply(ilh_plan* pl, int power, const double* s_in, double* s_out, int mem)
{
    return guarded([&] {
        if (!pl) fail(ILH_ERR_ARG, "null plan");
        if (power < 0 || power > 1000) fail(ILH_ERR_ARG, "power must be in [0, 1000]");
        Io io(pl, mem);
        const int N = pl->N;
        const size_t ng = pl->g.n(GK_PRIMAL);
        hipStream_t st = pl->ctx->stream;
        const double* d_in = io.in(s_in, N, pl->io_a);
        double* d_out = io.out(s_out, N, pl->io_d);
        if (N > 0) {
            if (d_out != d_in) ILH_HIP(hipMemcpyAsync(d_out, d_in, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
            DevBuf ones;
            const double* r1 = filter_r1(pl, ones);
            pl->scratch_a.ensure(sizeof(double) * ng);
            double* g = pl->scratch_a.as<double>();
            const unsigned gb = (unsigned)std::min<size_t>(4096, (ng + 255) / 256);
            for (int k = 0; k < power; ++k) {
                scatter_op(pl, OP_R_PRIMAL, d_out, g);
                k_filter_weight<<<gb, 256, 0, st>>>(g, r1, ng);
                gather_op(pl, OP_E_PRIMAL, g, d_out);
            }
            ILH_HIP(hipGetLastError());
            ILH_HIP(hipStreamSynchronize(st));
        }
        io.back(s_out, d_out, N);
        io.done();
    });
}

// ------------------------------------------- vector cache: surface filter
namespace {
// R 1 on Edges{Primal} (both components), cached per geometry
const double* filter_r1e(ilh_plan* pl, DevBuf& ones)
{
    const size_t ne = pl->g.n_edges();
    if (pl->filt_gen_e != pl->gen) {
        const int n2 = std::max(1, 2 * pl->N);
        pl->filt_r1e.ensure(sizeof(double) * ne);
        ones.ensure(sizeof(double) * n2);
        k_fill_value<<<(n2 + 255) / 256, 256, 0, pl->ctx->stream>>>(ones.as<double>(), (size_t)2 * pl->N, 1.0);
        scatter_op(pl, OP_R_EDGES, ones.as<double>(), pl->filt_r1e.as<double>());
        ILH_HIP(hipStreamSynchronize(pl->ctx->stream));   // `ones` is released by the caller
        pl->filt_gen_e = pl->gen;
    }
    return pl->filt_r1e.as<double>();
}

// v <- C^power v on VectorData (device, 2N), C applied matrix-free: scatter
// onto the edges, divide by R 1, interpolate back; g: edge-sized scratch
void filter_vector(ilh_plan* pl, int power, double* v, double* g)
{
    DevBuf ones;
    const double* r1 = filter_r1e(pl, ones);
    const size_t ne = pl->g.n_edges();
    const unsigned gb = (unsigned)std::min<size_t>(4096, (ne + 255) / 256);
    for (int k = 0; k < power; ++k) {
        scatter_op(pl, OP_R_EDGES, v, g);
        k_filter_weight<<<gb, 256, 0, pl->ctx->stream>>>(g, r1, ne);
        gather_op(pl, OP_E_EDGES, g, v);
    }
    ILH_HIP(hipGetLastError());
}
}  // namespace

namespace {
// create_surface_filter of the vector cache into d_C (2N x 2N, column-major,
// N > 0): block-diagonal, the u block from the u edges, the v block from the
// v edges.  block = 0 or 1: that diagonal block alone (N x N, column-major).
void build_filter_vector(ilh_plan* pl, double* d_C, int block = -1)
{
    const int N = pl->N, n2 = 2 * N;
    hipStream_t st = pl->ctx->stream;
    const int n = block < 0 ? n2 : N;
    fill_zero(st, d_C, (size_t)n * n);
    const Term* te = get_terms(pl, OP_E_EDGES);
    const Term* tr = get_terms(pl, OP_R_EDGES);
    if (!pl->bins_ok[OP_R_EDGES]) {
        build_bins(pl, tr, 2 * N, GK_EDGE_U, pl->bins[OP_R_EDGES][0]);
        build_bins(pl, tr, 2 * N, GK_EDGE_V, pl->bins[OP_R_EDGES][1]);
        pl->bins_ok[OP_R_EDGES] = true;
    }
    for (int c = 0; c < 2; ++c) {
        if (block >= 0 && c != block) continue;
        const RowBins& b = pl->bins[OP_R_EDGES][c];
        k_filter_build<<<(N + 127) / 128, 128, 0, st>>>(te, 2, c, N, b.cell_ptr, b.cells, b.cai, b.cw, b.ny, n,
                                                        block < 0 ? 0 : c * N, d_C);
    }
    ILH_HIP(hipGetLastError());
}
}  // namespace

extern "C" int ilh_surface_filter_build_vector(ilh_plan* pl, double* C, int mem)
{
    return guarded([&] {
        if (!pl) fail(ILH_ERR_ARG, "null plan");
        Io io(pl, mem);
        const int N = pl->N, n2 = 2 * N;
        const size_t nn = (size_t)n2 * n2;
        double* d_C = io.out(C, nn, pl->io_b);
        if (N > 0) build_filter_vector(pl, d_C);
        io.back(C, d_C, nn);
        io.done();
    });
}

extern "C" int ilh_surface_filter_apply_vector(ilh_plan* pl, int power, const double* v_in, double* v_out, int mem)
{
    return guarded([&] {
        if (!pl) fail(ILH_ERR_ARG, "null plan");
        if (power < 0 || power > 1000) fail(ILH_ERR_ARG, "power must be in [0, 1000]");
        Io io(pl, mem);
        const int n2 = 2 * pl->N;
        hipStream_t st = pl->ctx->stream;
        const double* d_in = io.in(v_in, n2, pl->io_a);
        double* d_out = io.out(v_out, n2, pl->io_d);
        if (n2 > 0) {
            if (d_out != d_in) ILH_HIP(hipMemcpyAsync(d_out, d_in, sizeof(double) * n2, hipMemcpyDeviceToDevice, st));
            pl->scratch_a.ensure(sizeof(double) * pl->g.n_edges());
            filter_vector(pl, power, d_out, pl->scratch_a.as<double>());
        }
        io.back(v_out, d_out, n2);
        io.done();
    });
}

// ------------------------------------------------------------ Stokes flow
// test/literate/stokes.jl:101-166 on the vector cache.  Per geometry: S =
// create_CL2invCT (2N x 2N) and its inverse, the inverse of Ss =
// create_CLinvCT_scalar (N x N), and Ka = E C L^-2 C^T D_s^symm (2N x 2N),
// row-major.  Per solve, with a = L^-2 C^T D_s^symm [v] (the reference's psi*
// is -a), phi = L^-1 R (n . [v]), v_phi = G phi:
//   sigma = S^-1 (v_bar - E v_phi + Ka [v]),   b = L^-2 C_s^T sigma,
//   t = a + b = L^-1 (L^-1 (C^T D_s^symm [v] + C_s^T sigma)),
//   v = v_phi - C t,   s = -(t + L^-1 C^x_s^T Ss^-1 G^x_s phi).
// The reference's six convolutions per solve become three passes: phi on
// Nodes{Primal}, L^-1 of the summed sources on Nodes{Dual}, then L^-1 of
// that field and of the potential's sources together as one batched pass.
namespace {
__global__ void k_stokes_rhs(int N, const double* __restrict__ vp, const double* __restrict__ vm,
                             const double* __restrict__ nx, const double* __restrict__ ny, double* __restrict__ dv,
                             double* __restrict__ vb, double* __restrict__ dvn)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N) return;
    double d[2];
    for (int c = 0; c < 2; ++c) {
        const int k = p + c * N;
        const double a = vp[k], b = vm ? vm[k] : 0.0;
        d[c] = a - b;                 // prescribed_surface_jump!
        dv[k] = d[c];
        vb[k] = 0.5 * a + 0.5 * b;    // prescribed_surface_average!
    }
    dvn[p] = nx[p] * d[0] + ny[p] * d[1];   // pointwise_dot!(dvn, nrm, dv)
}

// vprime = (vb - e_phi) + w
__global__ void k_stokes_vprime(int n, const double* __restrict__ vb, const double* __restrict__ ephi,
                                const double* __restrict__ w, double* __restrict__ out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) out[k] = (vb[k] - ephi[k]) + w[k];
}

// the vector filter's diagonal blocks (C is block-diagonal: u and v
// components filter independently)
void stokes_filter(ilh_plan* pl, StokesCache* sc)
{
    if (sc->cf_ok || pl->N == 0) return;
    const int N = pl->N, ld = solve_ld(N);
    DevBuf K;
    K.ensure(sizeof(double) * (size_t)N * N);
    sc->Cf.ensure(sizeof(double) * 2 * (size_t)N * ld);
    for (int c = 0; c < 2; ++c) {
        build_filter_vector(pl, K.as<double>(), c);
        transpose(pl->ctx->stream, K.as<double>(), sc->Cf.as<double>() + (size_t)c * N * ld, N, N, ld);
    }
    ILH_HIP(hipStreamSynchronize(pl->ctx->stream));   // K is released on return
    sc->cf_ok = true;
}

StokesCache* prepare_stokes(ilh_plan* pl)
{
    if (pl->sto) return pl->sto;
    need_margin(pl, "Stokes solve");
    StokesCache* sc = new StokesCache();
    try {
        const int N = pl->N, n2 = 2 * N;
        const Grid& g = pl->g;
        sc->sv.ensure(sizeof(double) * (size_t)std::max(1, 6 * n2 + 3 * N));
        sc->gE.ensure(sizeof(double) * std::max(g.n_edges(), (size_t)n2));   // also the filter's second iterate
        sc->gE2.ensure(sizeof(double) * g.n_edges());
        sc->gDD.ensure(2 * sizeof(double) * g.n(GK_DUAL));
        sc->gP.ensure(sizeof(double) * g.n(GK_PRIMAL));
        if (N > 0) {
            DevBuf K;
            K.ensure(sizeof(double) * (size_t)n2 * n2);
            build_cl2invct(pl, 1.0, K.as<double>());
            sc->S2.ensure(sizeof(double) * (size_t)n2 * solve_ld(n2));
            transpose(pl->ctx->stream, K.as<double>(), sc->S2.as<double>(), n2, n2, solve_ld(n2));
            sc->S2inv.ensure(sizeof(double) * (size_t)n2 * solve_ld(n2));
            inverse_rowmajor(pl, K.as<double>(), sc->S2inv.as<double>(), n2);
            build_schur(pl, ILH_SCHUR_CLINVCT, 1.0, K.as<double>());
            sc->Ssinv.ensure(sizeof(double) * (size_t)N * solve_ld(N));
            inverse_rowmajor(pl, K.as<double>(), sc->Ssinv.as<double>(), N);
            build_stokes_ka(pl, K.as<double>());
            sc->Ka.ensure(sizeof(double) * (size_t)n2 * solve_ld(n2));
            transpose(pl->ctx->stream, K.as<double>(), sc->Ka.as<double>(), n2, n2, solve_ld(n2));
            ILH_HIP(hipStreamSynchronize(pl->ctx->stream));
        }
    } catch (...) {
        delete sc;
        throw;
    }
    pl->sto = sc;
    return sc;
}
}  // namespace

extern "C" int ilh_stokes_solve(ilh_plan* pl, const double* v_plus, const double* v_minus, double* v, double* s,
                                double* sigma, int filter_power, int mem)
{
    return guarded([&] {
        if (!pl) fail(ILH_ERR_ARG, "null plan");
        if (filter_power < 0 || filter_power > 1000) fail(ILH_ERR_ARG, "filter_power must be in [0, 1000]");
        Io io(pl, mem);
        StokesCache* sc = prepare_stokes(pl);
        const int N = pl->N, n2 = 2 * N;
        const Grid& g = pl->g;
        const size_t ne = g.n_edges(), nd = g.n(GK_DUAL);
        hipStream_t st = pl->ctx->stream;
        const double* d_vp = io.in(v_plus, n2, pl->io_a);
        const double* d_vm = v_minus ? io.in(v_minus, n2, pl->io_c) : nullptr;
        double* d_v = io.out(v, ne, pl->io_b);
        double* d_s = io.out(s, nd, pl->io_d);
        if (n2 > 0 && !sigma) fail(ILH_ERR_ARG, "null output array");
        // sigma is staged in the surface work area when it lives on the host
        double* sv = sc->sv.as<double>();
        double *dv = sv, *vb = sv + n2, *ephi = sv + 2 * n2, *w = sv + 3 * n2, *vpr = sv + 4 * n2;
        double* sig = mem == ILH_MEM_DEVICE ? sigma : sv + 5 * n2;
        double *dvn = sv + 6 * n2, *dsv = dvn + N, *dsn = dsv + N;
        double *gE = sc->gE.as<double>(), *gE2 = sc->gE2.as<double>(), *gP = sc->gP.as<double>();
        double *gD = sc->gDD.as<double>(), *gD2 = gD + nd;
        if (N == 0) {
            fill_zero(st, d_v, ne);
            fill_zero(st, d_s, nd);
        } else {
            const unsigned bN = (unsigned)((N + 127) / 128), b2 = (unsigned)((n2 + 127) / 128);
            const int ld2 = solve_ld(n2);
            k_stokes_rhs<<<bN, 128, 0, st>>>(N, d_vp, d_vm, pl->nx.as<double>(), pl->ny.as<double>(), dv, vb, dvn);
            // phi = L^-1 R (n . [v]),  v_phi = G phi
            scatter_op(pl, OP_R_PRIMAL, dvn, gP);
            conv_dense(pl, GK_PRIMAL, gP, gP);
            k_grad(pl, gE2, gP);
            gather_op(pl, OP_E_EDGES, gE2, ephi);
            // w = Ka [v] = E C a (the reference's -E C psi*)
            matvec_rows(st, sc->Ka.as<double>(), ld2, dv, w, n2, n2, 1.0, 0);
            k_stokes_vprime<<<b2, 128, 0, st>>>(n2, vb, ephi, w, vpr);
            ILH_HIP(hipGetLastError());
            // sigma = S^-1 vprime with one refinement step, so that the residual
            // (the constraint E v = v_bar) is at roundoff as with the
            // reference's LU solve, not at cond(S) eps
            matvec_rows(st, sc->S2inv.as<double>(), ld2, vpr, sig, n2, n2, 1.0, 0);
            axpby(st, ephi, vpr, 1.0, 0.0, n2);
            matvec_rows(st, sc->S2.as<double>(), ld2, sig, ephi, n2, n2, -1.0, 1);
            matvec_rows(st, sc->S2inv.as<double>(), ld2, ephi, sig, n2, n2, 1.0, 1);
            // y = L^-1 (C^T D_s^symm [v] + C_s^T sigma)
            scatter_op(pl, OP_DSYM, dv, gE);
            k_curl_e2n(pl, gD, gE);
            scatter_op(pl, OP_CVT, sig, gD2);
            axpby(st, gD, gD2, 1.0, 1.0, nd);
            conv_dense(pl, GK_DUAL, gD, gD);
            // the potential's streamfunction sources C^x_s^T Ss^-1 G^x_s phi
            gather_op(pl, OP_GXS, gP, dsv);
            matvec_rows(st, sc->Ssinv.as<double>(), solve_ld(N), dsv, dsn, N, N, 1.0, 0);
            scatter_op(pl, OP_CXT, dsn, gD2);
            // [t, c] = L^-1 [y, C^x_s^T ds] in one batched pass
            conv_dense_multi(pl, GK_DUAL, gD, 2, gD);
            // v = v_phi - C t,  s = -(t + c)
            k_curl_n2e(pl, d_v, gD);
            axpby(st, d_v, gE2, 1.0, -1.0, ne);
            axpby(st, gD2, gD, -1.0, -1.0, nd);
            ILH_HIP(hipMemcpyAsync(d_s, gD2, sizeof(double) * nd, hipMemcpyDeviceToDevice, st));
            // sigma <- C^k sigma with the stored diagonal blocks of the filter
            // (gE holds the other iterate)
            if (filter_power > 0) stokes_filter(pl, sc);
            double *fa = sig, *fb = gE;
            const int ld1 = solve_ld(N);
            for (int k = 0; k < filter_power; ++k) {
                matvec_rows(st, sc->Cf.as<double>(), ld1, fa, fb, N, N, 1.0, 0);
                matvec_rows(st, sc->Cf.as<double>() + (size_t)N * ld1, ld1, fa + N, fb + N, N, N, 1.0, 0);
                std::swap(fa, fb);
            }
            if (fa != sig) ILH_HIP(hipMemcpyAsync(sig, fa, sizeof(double) * n2, hipMemcpyDeviceToDevice, st));
        }
        io.back(v, d_v, ne);
        io.back(s, d_s, nd);
        if (mem != ILH_MEM_DEVICE && n2 > 0)
            ILH_HIP(hipMemcpyAsync(sigma, sig, sizeof(double) * n2, hipMemcpyDeviceToHost, st));
        io.done();
    });
}

// ------------------------------------------------------ batched solves
// Repeat Dirichlet solves of many independent plans of one context: the
// plans' caches are prepared first (in order, on the context stream), then
// their solves are issued round-robin on POOL side streams that wait for the
// context stream's prior work, and the context stream waits for all of them.
// Kernels of different problems overlap on the GPU.
namespace {

ilh_ctx* batch_ctx(int nprob, ilh_plan* const* plans)
{
    if (nprob <= 0) fail(ILH_ERR_ARG, "batch needs at least one problem");
    if (!plans) fail(ILH_ERR_ARG, "null array of plans");
    ilh_ctx* ctx = nullptr;
    for (int i = 0; i < nprob; ++i) {
        if (!plans[i]) fail(ILH_ERR_ARG, "null plan in batch");
        if (!ctx) ctx = plans[i]->ctx;
        if (plans[i]->ctx != ctx) fail(ILH_ERR_ARG, "all plans of a batch must share one context");
        for (int k = 0; k < i; ++k)
            if (plans[k] == plans[i]) fail(ILH_ERR_ARG, "a plan appears twice in one batch");
    }
    ILH_HIP(hipSetDevice(ctx->device));
    for (int k = 0; k < ilh_ctx::POOL; ++k)
        if (!ctx->pool[k]) ILH_HIP(hipStreamCreateWithFlags(&ctx->pool[k], hipStreamNonBlocking));
    for (hipEvent_t& e : ctx->ev)
        if (!e) ILH_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    return ctx;
}

// Issue the batch's solves round-robin on the side streams, forked from and
// joined back into the context stream.  Caches are prepared by the caller.
void batch_issue(ilh_ctx* ctx, int nprob, ilh_plan* const* plans, SolveCache* const* sc,
                 const double* const* fb_plus, const double* const* fb_minus, double* const* f, double* const* s,
                 int mem)
{
    const hipStream_t main = ctx->stream;
    ILH_HIP(hipEventRecord(ctx->ev[ilh_ctx::POOL], main));
    for (int k = 0; k < ilh_ctx::POOL; ++k) ILH_HIP(hipStreamWaitEvent(ctx->pool[k], ctx->ev[ilh_ctx::POOL], 0));
    struct Restore {
        ilh_ctx* c;
        hipStream_t m;
        ~Restore() { c->stream = m; }
    } restore{ctx, main};
    for (int i = 0; i < nprob; ++i) {
        ilh_plan* pl = plans[i];
        ctx->stream = ctx->pool[i % ilh_ctx::POOL];
        Io io(pl, mem);
        const int N = pl->N;
        const size_t ng = pl->g.n(GK_PRIMAL);
        const double* d_fp = io.in(fb_plus[i], N, pl->io_a);
        const double* fm_i = fb_minus ? fb_minus[i] : nullptr;
        const double* d_fm = fm_i ? io.in(fm_i, N, pl->io_c) : nullptr;
        double* d_f = io.out(f[i], ng, pl->io_b);
        double* d_s = io.out(s[i], N, pl->io_d);
        if (N == 0) {
            fill_zero(ctx->stream, d_f, ng);
        } else {
            dirichlet_amp(ctx->stream, sc[i]->M[0].as<double>(), sc[i]->M[1].as<double>(), solve_ld(N), d_fp, d_fm,
                          sc[i]->amp1.as<double>(), d_s, N);
            conv_sources(pl, GK_PRIMAL, sc[i]->src1, sc[i]->amp1.as<double>(), d_f);
        }
        io.back(f[i], d_f, ng);
        io.back(s[i], d_s, N);
    }
    ctx->stream = main;
    for (int k = 0; k < ilh_ctx::POOL; ++k) {
        ILH_HIP(hipEventRecord(ctx->ev[k], ctx->pool[k]));
        ILH_HIP(hipStreamWaitEvent(main, ctx->ev[k], 0));
    }
}

// The fused form (one launch per pass over every problem, ilh_batch.hip)
// serves a batch whose problems' FFT periods it supports; problems without
// surface points only get their field zeroed.
bool batch_fusable(int nprob, ilh_plan* const* plans)
{
    for (int i = 0; i < nprob; ++i) {
        if (plans[i]->N == 0) continue;
        const ConvPlan& C = get_conv(plans[i], GK_PRIMAL);
        if (!fused_supported(C.Px, C.Py)) return false;
    }
    return true;
}

// the problems' operands (device pointers) for the fused passes; zero: the
// fields of problems without surface points
void batch_probs(ilh_ctx* ctx, int nprob, ilh_plan* const* plans, SolveCache* const* sc,
                 const double* const* fp, const double* const* fm, double* const* f, double* const* s,
                 std::vector<BatchProb>& probs, std::vector<std::pair<double*, size_t>>& zero)
{
    probs.clear();
    zero.clear();
    for (int i = 0; i < nprob; ++i) {
        ilh_plan* pl = plans[i];
        const int N = pl->N;
        if (N == 0) {
            zero.emplace_back(f[i], pl->g.n(GK_PRIMAL));
            continue;
        }
        const SourceSet& src = sc[i]->src1;
        if (src.npairs_dev) fail(ILH_ERR_STATE, "the source pairs are not pinned");
        ConvPlan& C = get_conv(pl, GK_PRIMAL);
        BatchProb b;
        std::memset(&b, 0, sizeof(b));   // padding too: batches are compared bytewise
        b.Mp = sc[i]->M[0].as<double>();
        b.Mm = fm[i] ? sc[i]->M[1].as<double>() : nullptr;
        b.ld = solve_ld(N);
        b.fp = fp[i];
        b.fm = fm[i];
        b.amp = sc[i]->amp1.as<double>();
        b.s = s[i];
        b.n = N;
        b.pinfo = src.pinfo;
        b.cells = src.bins.cells;
        b.cai = src.bins.cai;
        b.cw = src.bins.cw;
        b.npairs = src.npairs;
        b.nx = C.nx;
        b.ny = C.ny;
        b.Px = C.Px;
        b.Py = C.Py;
        b.T1 = C.T1.as<double2>();
        b.slot = src.slot_of_row;
        b.spec = C.spec.as<double>();
        b.T2 = C.T2.as<double2>();
        b.out = f[i];
        b.twx = fft_twiddles(ctx, C.Px);
        b.twy = fft_twiddles(ctx, C.Py);
        probs.push_back(b);
    }
}

bool same_probs(const std::vector<BatchProb>& a, const std::vector<BatchProb>& b)
{
    return a.size() == b.size() && std::memcmp(a.data(), b.data(), sizeof(BatchProb) * a.size()) == 0;
}

}  // namespace

extern "C" int ilh_dirichlet_solve_batched(int nprob, ilh_plan* const* plans, const double* const* fb_plus,
                                           const double* const* fb_minus, double* const* f, double* const* s,
                                           int mem)
{
    return guarded([&] {
        if (nprob == 0) return;
        if (nprob < 0) fail(ILH_ERR_ARG, "negative problem count");
        if (!fb_plus || !f || !s) fail(ILH_ERR_ARG, "null array of pointers");
        if (mem != ILH_MEM_HOST && mem != ILH_MEM_DEVICE) fail(ILH_ERR_ARG, "mem must be ILH_MEM_HOST or ILH_MEM_DEVICE");
        ilh_ctx* ctx = batch_ctx(nprob, plans);
        std::vector<SolveCache*> sc(nprob);
        for (int i = 0; i < nprob; ++i) sc[i] = prepare_dirichlet(plans[i], fb_minus && fb_minus[i]);
        if (!batch_fusable(nprob, plans)) {
            batch_issue(ctx, nprob, plans, sc.data(), fb_plus, fb_minus, f, s, mem);
            if (mem == ILH_MEM_HOST) ILH_HIP(hipStreamSynchronize(ctx->stream));
            return;
        }
        // device operands: the caller's, or each plan's staging buffers
        std::vector<const double*> dp(nprob), dm(nprob);
        std::vector<double*> df(nprob), ds(nprob);
        for (int i = 0; i < nprob; ++i) {
            ilh_plan* pl = plans[i];
            Io io(pl, mem);
            const int N = pl->N;
            dp[i] = io.in(fb_plus[i], N, pl->io_a);
            const double* fm_i = fb_minus ? fb_minus[i] : nullptr;
            dm[i] = fm_i ? io.in(fm_i, N, pl->io_c) : nullptr;
            df[i] = io.out(f[i], pl->g.n(GK_PRIMAL), pl->io_b);
            ds[i] = io.out(s[i], N, pl->io_d);
        }
        std::vector<BatchProb> probs;
        std::vector<std::pair<double*, size_t>> zero;
        batch_probs(ctx, nprob, plans, sc.data(), dp.data(), dm.data(), df.data(), ds.data(), probs, zero);
        if (!probs.empty()) {
            if (!ctx->fused || !same_probs(fused_probs(ctx->fused), probs)) {
                if (ctx->fused) {
                    ILH_HIP(hipStreamSynchronize(ctx->stream));
                    fused_destroy(ctx->fused);
                    ctx->fused = nullptr;
                }
                ctx->fused = fused_create(ctx, probs);
            }
            fused_run(ctx->fused);
        }
        for (auto& z : zero) fill_zero(ctx->stream, z.first, z.second);
        if (mem == ILH_MEM_HOST) {
            for (int i = 0; i < nprob; ++i) {
                Io io(plans[i], mem);
                io.back(f[i], df[i], plans[i]->g.n(GK_PRIMAL));
                io.back(s[i], ds[i], plans[i]->N);
            }
            ILH_HIP(hipStreamSynchronize(ctx->stream));
        }
    });
}

// A batch captured once as a HIP graph (device buffers fixed at capture) and
// replayed with one launch per step.
struct ilh_batch {
    ilh_ctx* ctx = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<ilh_plan*> plans;
    std::vector<unsigned> gens;   // plan generations the graph's buffers belong to
    FusedBatch* fused = nullptr;  // the fused passes (autotuned), when the batch's periods allow them
    ~ilh_batch()
    {
        if (fused) fused_destroy(fused);
    }
};

extern "C" int ilh_batch_create(int nprob, ilh_plan* const* plans, const double* const* fb_plus,
                                const double* const* fb_minus, double* const* f, double* const* s, ilh_batch** out)
{
    return guarded([&] {
        if (!out) fail(ILH_ERR_ARG, "null output pointer");
        if (!fb_plus || !f || !s) fail(ILH_ERR_ARG, "null array of pointers");
        ilh_ctx* ctx = batch_ctx(nprob, plans);
        std::vector<SolveCache*> sc(nprob);
        for (int i = 0; i < nprob; ++i) sc[i] = prepare_dirichlet(plans[i], fb_minus && fb_minus[i]);
        std::unique_ptr<ilh_batch> own(new ilh_batch());
        ilh_batch* b = own.get();
        b->ctx = ctx;
        b->plans.assign(plans, plans + nprob);
        for (int i = 0; i < nprob; ++i) b->gens.push_back(plans[i]->gen);
        std::vector<std::pair<double*, size_t>> zero;
        if (batch_fusable(nprob, plans)) {
            std::vector<const double*> fm(nprob, nullptr);
            if (fb_minus) fm.assign(fb_minus, fb_minus + nprob);
            std::vector<BatchProb> probs;
            batch_probs(ctx, nprob, plans, sc.data(), fb_plus, fm.data(), f, s, probs, zero);
            if (!probs.empty()) {
                b->fused = fused_create(ctx, probs);
                fused_tune(b->fused);   // also the one uncaptured pass below needs
            }
        }
        auto issue = [&] {
            if (b->fused || !zero.empty()) {
                if (b->fused) fused_run(b->fused);
                for (auto& z : zero) fill_zero(ctx->stream, z.first, z.second);
            } else {
                batch_issue(ctx, nprob, plans, sc.data(), fb_plus, fb_minus, f, s, ILH_MEM_DEVICE);
            }
        };
        // one uncaptured pass: every lazily built table and attribute exists
        // before capture (capture must record kernels only)
        issue();
        ILH_HIP(hipStreamSynchronize(ctx->stream));
        hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) fail(ILH_ERR_HIP, std::string("hipStreamBeginCapture: ") + hipGetErrorString(e));
        try {
            issue();
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(ctx->stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        e = hipStreamEndCapture(ctx->stream, &b->graph);
        if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            if (b->graph) (void)hipGraphDestroy(b->graph);
            fail(ILH_ERR_HIP, std::string("batch graph capture: ") + hipGetErrorString(e));
        }
        ctx_retain(ctx);
        for (ilh_plan* p : b->plans) plan_retain(p);
        *out = own.release();
    });
}

extern "C" int ilh_batch_tuning(const ilh_batch* b, int* config, float* pass_ms)
{
    return guarded([&] {
        if (!b) fail(ILH_ERR_ARG, "null batch");
        if (!b->fused) fail(ILH_ERR_STATE, "the batch runs per-problem kernels (an FFT period the fused passes lack)");
        fused_config(b->fused, config, pass_ms);
    });
}

extern "C" int ilh_batch_launch(ilh_batch* b)
{
    return guarded([&] {
        if (!b) fail(ILH_ERR_ARG, "null batch");
        for (size_t i = 0; i < b->plans.size(); ++i)
            if (b->plans[i]->gen != b->gens[i] || !b->plans[i]->dir)
                fail(ILH_ERR_STATE, "a plan of the batch changed geometry since the batch was captured");
        ILH_HIP(hipSetDevice(b->ctx->device));
        ILH_HIP(hipGraphLaunch(b->exec, b->ctx->stream));
    });
}

extern "C" int ilh_batch_destroy(ilh_batch* b)
{
    return guarded([&] {
        if (!b) return;
        (void)hipSetDevice(b->ctx->device);
        (void)hipStreamSynchronize(b->ctx->stream);
        if (b->exec) (void)hipGraphExecDestroy(b->exec);
        if (b->graph) (void)hipGraphDestroy(b->graph);
        if (b->fused) fused_destroy(b->fused);
        b->fused = nullptr;
        for (ilh_plan* p : b->plans) plan_release(p);
        ctx_release(b->ctx);
        delete b;
    });
}

// ------------------------------------------------- moving-body Dirichlet
namespace {

// GMRES state on the device (m = restart length):
//   H (m+1) x m column-major, cs[m], sn[m], g[m+1], then bnorm and tol;
//   flags[0]: 0 running, 1 converged, 2 not converged within m; flags[1]: steps
struct GmLayout {
    int m;
    __host__ __device__ size_t H() const { return 0; }
    __host__ __device__ size_t cs() const { return (size_t)(m + 1) * m; }
    __host__ __device__ size_t sn() const { return cs() + m; }
    __host__ __device__ size_t g() const { return sn() + m; }
    __host__ __device__ size_t bnorm() const { return g() + m + 1; }
    __host__ __device__ size_t tol() const { return bnorm() + 1; }
    __host__ __device__ size_t size() const { return tol() + 1; }
};

// block-wide sum over 1024 threads (fixed order)
__device__ __forceinline__ double block_sum_1024(double a, double* red)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, 64);
    __syncthreads();
    if (lane == 0) red[wave] = a;
    __syncthreads();
    double t = 0.0;
    for (int i = 0; i < 16; ++i) t += red[i];
    return t;
}

// r0 = w: beta = |r0|, bnorm = |b|, v_0 = r0 / beta, g = (beta, 0, ...);
// converged at once when b = 0 or beta <= tol |b|
__device__ __forceinline__ void gm_init_body(const double* __restrict__ w, const double* __restrict__ b, int n,
                                             double* __restrict__ V, double* __restrict__ st, int m, double tol,
                                             int* __restrict__ flags)   // v_0 = V[0, n)
{
    __shared__ double red[16];
    const GmLayout L{m};
    double a = 0.0, c = 0.0;
    for (int k = threadIdx.x; k < n; k += 1024) {
        a += w[k] * w[k];
        c += b[k] * b[k];
    }
    const double beta = sqrt(block_sum_1024(a, red));
    const double bn = sqrt(block_sum_1024(c, red));
    const bool done = bn == 0.0 || beta <= tol * bn;
    if (threadIdx.x == 0) {
        st[L.g()] = beta;
        for (int i = 1; i <= m; ++i) st[L.g() + i] = 0.0;
        st[L.bnorm()] = bn;
        st[L.tol()] = tol;
        flags[0] = done ? 1 : 0;
        flags[1] = 0;
    }
    if (!done) {
        const double inv = 1.0 / beta;
        for (int k = threadIdx.x; k < n; k += 1024) V[k] = w[k] * inv;
    }
}

__global__ void __launch_bounds__(1024) k_gm_init(const double* __restrict__ w, const double* __restrict__ b, int n,
                                                  double* __restrict__ V, double* __restrict__ st, int m, double tol,
                                                  int* __restrict__ flags)
{
    gm_init_body(w, b, n, V, st, m, tol, flags);
}

// y = M x (M row-major, one wave per row) unless the iteration has stopped
__global__ void __launch_bounds__(256) k_gm_matvec(const double* __restrict__ M, int ld, const double* __restrict__ x,
                                                   double* __restrict__ y, int n, const int* __restrict__ flags)
{
    if (flags[0]) return;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (i >= n) return;
    const double s = wave_dot(M + (size_t)ld * i, x, n, lane);
    if (lane == 0) y[i] = s;
}

// dot(row, x) by one wave with the row in FP32 (the preconditioner), widened
// to FP64: 16-byte row loads (4 values) when the row and x allow them, two in
// flight per lane, fixed order
// (double)a0 b.x + (double)a1 b.y with one explicit fused multiply-add, so
// that the rounding is the same wherever the compiler places it
__device__ __forceinline__ double dot2f(float a0, float a1, double2 b)
{
    return __builtin_fma((double)a0, b.x, (double)a1 * b.y);
}

__device__ __forceinline__ double wave_dot_f32(const float* __restrict__ row, const double* __restrict__ x, int m,
                                               int lane)
{
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int k0 = 0;
    if ((((uintptr_t)row | (uintptr_t)x) & 15) == 0) {
        // lane l takes the float4 entries l + 64 r, round r even into s0, s1
        // and odd into s2, s3, in increasing r; the loads of four rounds go
        // out before the first is used (a row of a few thousand entries is
        // two memory round trips)
        const float4* r4 = (const float4*)row;
        const double2* x2 = (const double2*)x;
        const int m4 = m >> 2;
        for (int k = lane; k < m4; k += 64 * 4) {
            float4 a[4];
            double2 b0[4], b1[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int e = k + 64 * r;
                const bool ok = e < m4;
                a[r] = ok ? r4[e] : make_float4(0.f, 0.f, 0.f, 0.f);
                b0[r] = ok ? x2[2 * e] : make_double2(0.0, 0.0);
                b1[r] = ok ? x2[2 * e + 1] : make_double2(0.0, 0.0);
            }
#pragma unroll
            for (int r = 0; r < 4; r += 2) {
                if (k + 64 * r < m4) {
                    s0 += dot2f(a[r].x, a[r].y, b0[r]);
                    s1 += dot2f(a[r].z, a[r].w, b1[r]);
                }
                if (k + 64 * (r + 1) < m4) {
                    s2 += dot2f(a[r + 1].x, a[r + 1].y, b0[r + 1]);
                    s3 += dot2f(a[r + 1].z, a[r + 1].w, b1[r + 1]);
                }
            }
        }
        k0 = 4 * m4;
    }
    for (int k = k0 + lane; k < m; k += 64) s0 = __builtin_fma((double)row[k], x[k], s0);
    double t = (s0 + s1) + (s2 + s3);
    for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
    return t;
}

// wave_dot_f32 of two FP32 rows with one x by one wave: x is loaded once for
// both, and each row's sum is wave_dot_f32's, bit for bit (rows whose
// alignment differs take wave_dot_f32 one by one)
__device__ __forceinline__ double2 wave_dot_f32_rows2(const float* __restrict__ ra, const float* __restrict__ rb,
                                                      const double* __restrict__ x, int m, int lane)
{
    const bool va = (((uintptr_t)ra | (uintptr_t)x) & 15) == 0, vb = (((uintptr_t)rb | (uintptr_t)x) & 15) == 0;
    if (va != vb) return make_double2(wave_dot_f32(ra, x, m, lane), wave_dot_f32(rb, x, m, lane));
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0, t0 = 0.0, t1 = 0.0, t2 = 0.0, t3 = 0.0;
    int k0 = 0;
    if (va) {
        const float4* a4 = (const float4*)ra;
        const float4* c4 = (const float4*)rb;
        const double2* x2 = (const double2*)x;
        const int m4 = m >> 2;
        for (int k = lane; k < m4; k += 64 * 4) {
            float4 a[4], c[4];
            double2 b0[4], b1[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int e = k + 64 * r;
                const bool ok = e < m4;
                a[r] = ok ? a4[e] : make_float4(0.f, 0.f, 0.f, 0.f);
                c[r] = ok ? c4[e] : make_float4(0.f, 0.f, 0.f, 0.f);
                b0[r] = ok ? x2[2 * e] : make_double2(0.0, 0.0);
                b1[r] = ok ? x2[2 * e + 1] : make_double2(0.0, 0.0);
            }
#pragma unroll
            for (int r = 0; r < 4; r += 2) {
                if (k + 64 * r < m4) {
                    s0 += dot2f(a[r].x, a[r].y, b0[r]);
                    s1 += dot2f(a[r].z, a[r].w, b1[r]);
                    t0 += dot2f(c[r].x, c[r].y, b0[r]);
                    t1 += dot2f(c[r].z, c[r].w, b1[r]);
                }
                if (k + 64 * (r + 1) < m4) {
                    s2 += dot2f(a[r + 1].x, a[r + 1].y, b0[r + 1]);
                    s3 += dot2f(a[r + 1].z, a[r + 1].w, b1[r + 1]);
                    t2 += dot2f(c[r + 1].x, c[r + 1].y, b0[r + 1]);
                    t3 += dot2f(c[r + 1].z, c[r + 1].w, b1[r + 1]);
                }
            }
        }
        k0 = 4 * m4;
    }
    for (int k = k0 + lane; k < m; k += 64) {
        s0 = __builtin_fma((double)ra[k], x[k], s0);
        t0 = __builtin_fma((double)rb[k], x[k], t0);
    }
    double s = (s0 + s1) + (s2 + s3), t = (t0 + t1) + (t2 + t3);
    for (int off = 32; off > 0; off >>= 1) {
        s += __shfl_xor(s, off, 64);
        t += __shfl_xor(t, off, 64);
    }
    return make_double2(s, t);
}

// y = M x (M in FP32, row-major, one wave per row), or y += M x when
// accumulate; stops once the iteration has (flags non-null and set)
__global__ void __launch_bounds__(256) k_gm_matvec_f32(const float* __restrict__ M, int ld, const double* __restrict__ x,
                                                       double* __restrict__ y, int n, const int* __restrict__ flags,
                                                       int accumulate)
{
    if (flags && flags[0]) return;
    const int i = blockIdx.x * 4 + (int)(threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (i >= n) return;
    const double s = wave_dot_f32(M + (size_t)ld * i, x, n, lane);
    if (lane == 0) y[i] = s + (accumulate ? y[i] : 0.0);
}

__global__ void k_to_f32(const double* __restrict__ a, float* __restrict__ b, size_t n)
{
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
        b[k] = (float)a[k];
}

// ---- the GMRES steps of many plans in lockstep, one launch per phase
// One plan's operands of a batched GMRES step (device pointers).
struct BGm {
    const float* P32;          // FP32 reference inverse (row-major, leading dimension ld), or null
    const double* band;        // the learned band (n x bw) and its neighbours, when P32 is null
    const int* nbr;
    int bw;
    const double* S;           // the current geometry's matrix (row-major, ld)
    int ld, n;
    double *V, *w, *z, *st;    // basis (stride ld), work vectors, GMRES state
    int* flags;
    double* part;              // the Arnoldi step's partial sums (ArnLayout)
    // the step around the GMRES steps: boundary data, amplitudes [s; d]
    // (s the solution), fbar, the right-hand side b, the grid field f*, E's
    // terms on the primal grid (nx x ny), the strength out
    const double *fp, *fm;
    double *amp, *fbar, *b, *grid, *s_out;
    const Term* tE;
    int gnx, gny;
};

// z = P v_j for every plan still running; tiles (plan << 16 | 4-row block),
// two rows per wave (work-groups of 128), each row in the single solve's order
// (k_gm_matvec_f32, k_band_apply)
__global__ void __launch_bounds__(128) k_bgm_prec(const BGm* __restrict__ g, const unsigned* __restrict__ tiles, int j)
{
    const unsigned code = tiles[blockIdx.x];
    const BGm& G = g[code >> 16];
    if (G.flags[0]) return;
    const int i = (int)(code & 0xFFFF) * 4 + 2 * (int)(threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (i >= G.n) return;
    const bool two = i + 1 < G.n;
    const double* v = G.V + (size_t)G.ld * j;
    if (G.P32) {
        const double2 s =
            wave_dot_f32_rows2(G.P32 + (size_t)G.ld * i, G.P32 + (size_t)G.ld * (two ? i + 1 : i), v, G.n, lane);
        if (lane == 0) {
            G.z[i] = s.x;
            if (two) G.z[i + 1] = s.y;
        }
    } else {
        for (int r = 0; r < (two ? 2 : 1); ++r) {
            const double s =
                band_row(G.band + (size_t)(i + r) * G.bw, G.nbr + (size_t)(i + r) * G.bw, G.bw, v, lane);
            if (lane == 0) G.z[i + r] = s;
        }
    }
}

// w = S z for every plan still running (k_gm_matvec's arithmetic)
__global__ void __launch_bounds__(256) k_bgm_matvec(const BGm* __restrict__ g, const unsigned* __restrict__ tiles)
{
    const unsigned code = tiles[blockIdx.x];
    const BGm& G = g[code >> 16];
    if (G.flags[0]) return;
    const int i = (int)(code & 0xFFFF) * 4 + (int)(threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (i >= G.n) return;
    const double s = wave_dot(G.S + (size_t)G.ld * i, G.z, G.n, lane);
    if (lane == 0) G.w[i] = s;
}

// Arnoldi step j of plan blockIdx.x
// ---- Arnoldi step j, unless the iteration has stopped: w is orthogonalised
// against v_0..v_j (columns of V, stride ld) by classical Gram-Schmidt,
// repeated once when the first pass removed more than 1 - 1/sqrt(2) of |w|
// (the Daniel-Gragg-Kaufman-Stewart test), and normalised into v_{j+1};
// column j of H gets the summed coefficients and |w|; the previous Givens
// rotations are applied, a new one zeroes H[j+1, j], and |g[j+1]| (the
// residual norm) is tested against tol |b|.
//
// Each phase is a launch over slices of 64 entries (one work-group each), so
// the basis streams through every CU's loads rather than one:
//   k_arn_dots    slice dots v_i . w and w . w
//   k_arn_update  h = the slice dots summed, w1 = w - V h on the slice, its
//                 slice |w1|^2, and the second pass's slice dots v_i . w1
//                 (slice-local: they need only this slice of w1)
//   k_arn_final   the second pass (when the test asks for it): h' = its dots
//                 summed, w2 = w1 - V h'; |w2| = sqrt(|w1|^2 - |h'|^2)
//                 (Pythagoras, V orthonormal: h' is of the order of the first
//                 pass's rounding, eps |w|, so the difference is exact to
//                 working precision unless w lies in span V to that
//                 precision, where the clamp at 0 ends the iteration as a
//                 breakdown); v_{j+1} = w2 / |w2| on the slice; slice 0
//                 finishes column j of H
// Partial sums go to `part` (ArnLayout) and every block that needs a total
// adds them in slice order, so that all of them see the same value; every sum
// has a fixed order (deterministic).  The batched form runs the same kernels
// over (slice, plan) grids: each plan's arithmetic is the single solve's.

// partial sums of one Arnoldi step (doubles): the slice dots of each pass
// [G][m + 2], the first pass's slice norms [G], its h [m + 2]
struct ArnLayout {
    int G, m;   // slices of 64 entries, restart length
    __host__ __device__ size_t dots(int pass) const { return (size_t)pass * G * (m + 2); }
    __host__ __device__ size_t norms() const { return (size_t)2 * G * (m + 2); }
    __host__ __device__ size_t h() const { return (size_t)2 * G * (m + 2) + G; }
    __host__ __device__ size_t size() const { return h() + (size_t)m + 2; }
};
__host__ __device__ inline int arn_slices(int n) { return (n + 63) / 64; }

// one plan's Arnoldi operands (device pointers)
struct ArnOp {
    double* V;
    int ld;
    double* w;
    int n;
    double* st;
    int* flags;
    double* part;
};

__device__ __forceinline__ ArnOp arn_of(const BGm& G) { return ArnOp{G.V, G.ld, G.w, G.n, G.st, G.flags, G.part}; }

// x[0] + x[s] + ... (cnt terms) by one wave: lane l adds terms l, l + 64, ...
// in order, then a fixed tree over the lanes (every lane gets the sum)
__device__ __forceinline__ double wave_sum_strided(const double* __restrict__ x, size_t s, int cnt, int lane)
{
    double a = 0.0;
    for (int b = lane; b < cnt; b += 64) a += x[(size_t)b * s];
    for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, 64);
    return a;
}

// the second pass runs: |w| after the first pass < |w| before it / sqrt(2)
// (ww = sum of the first pass's slice dots w . w, q0 = sum of its slice norms)
__device__ __forceinline__ bool arn_again(double ww, double q0)
{
    return sqrt(q0) < 0.70710678118654752 * sqrt(ww);
}

// the operands of this block: the single solve's (BAT false, slice =
// blockIdx.x) or plan blockIdx.y's of the batch
template <bool BAT>
__device__ __forceinline__ ArnOp arn_block(const ArnOp& one, const BGm* __restrict__ g)
{
    return BAT ? arn_of(g[blockIdx.y]) : one;
}

// the slice dots v_i . w, i < nd (index nv: w . w), of entry k = 64 b + lane
// (wk = w[k], 0 past the end): wave u takes v_u, v_{u+4}, ..., eight of them
// per batch of loads, each a fixed tree over the slice's 64 entries
__device__ __forceinline__ void arn_slice_dots(const ArnOp& A, double* __restrict__ out, int nd, int nv, int k,
                                               bool in, double wk, int wave, int lane)
{
    for (int i0 = wave; i0 < nd; i0 += 32) {
        double p[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int i = i0 + 4 * u;
            const double x = in && i < nd ? (i < nv ? A.V[(size_t)A.ld * i + k] : wk) : 0.0;
            p[u] = x * wk;
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1)
#pragma unroll
            for (int u = 0; u < 8; ++u) p[u] += __shfl_xor(p[u], off, 64);
        if (lane == 0)
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (i0 + 4 * u < nd) out[i0 + 4 * u] = p[u];
    }
}

// The kernels below issue every load that does not depend on a test (the
// flags, the second-pass test) before the test: data written by the previous
// launch may come from another XCD's writes, and each dependent round trip
// costs about a microsecond.  The loads stay inside the buffers whatever the
// flags say.

// the first pass's slice dots (and w . w)
template <bool BAT>
__global__ void __launch_bounds__(256) k_arn_dots(ArnOp one, const BGm* __restrict__ g, int j, int m)
{
    const ArnOp A = arn_block<BAT>(one, g);
    const ArnLayout L{arn_slices(A.n), m};
    const int b = blockIdx.x;
    if (b >= L.G) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nv = j + 1;
    const int k = 64 * b + lane;
    const bool in = k < A.n;
    const double wk = in ? A.w[k] : 0.0;
    if (A.flags[0]) return;
    arn_slice_dots(A, A.part + L.dots(0) + (size_t)(m + 2) * b, nv + 1, nv, k, in, wk, wave, lane);
}

// the first pass's h_t = its slice dots summed in slice order, 32 slices per
// batch of loads (slice 0 keeps h for k_arn_final); every thread t < nv
__device__ __forceinline__ double arn_h_sum(const double* __restrict__ d, int m, int G, const double (&hp)[32])
{
    double a = 0.0;
#pragma unroll
    for (int u = 0; u < 32; ++u)
        if (u < G) a += hp[u];
    for (int s0 = 32; s0 < G; s0 += 8) {
        double y[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) y[u] = s0 + u < G ? d[(size_t)(m + 2) * (s0 + u)] : 0.0;
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (s0 + u < G) a += y[u];
    }
    return a;
}

// acc = sum_{i = wave mod 4, i < nv} x_i h_i at entry k, the first eight
// basis entries already in x
__device__ __forceinline__ double arn_vh(const ArnOp& A, const double* __restrict__ h, double (&x)[8], int nv, int k,
                                        bool in, int wave)
{
    double acc = 0.0;
#pragma unroll
    for (int u = 0; u < 8; ++u)
        if (wave + 4 * u < nv) acc += x[u] * h[wave + 4 * u];
    for (int i0 = wave + 32; i0 < nv; i0 += 32) {
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int i = i0 + 4 * u;
            x[u] = in && i < nv ? A.V[(size_t)A.ld * i + k] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (i0 + 4 * u < nv) acc += x[u] * h[i0 + 4 * u];
    }
    return acc;
}

// this wave's first eight basis entries at entry k
__device__ __forceinline__ void arn_load_x(const ArnOp& A, double (&x)[8], int nv, int k, bool in, int wave)
{
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        const int i = wave + 4 * u;
        x[u] = in && i < nv ? A.V[(size_t)A.ld * i + k] : 0.0;
    }
}

// first pass: h = the slice dots summed; the slice's w1 = w - (sum over waves
// u of sum_{i = u mod 4} h_i v_i), its |w1|^2, and the second pass's slice dots
template <bool BAT>
__global__ void __launch_bounds__(256) k_arn_update(ArnOp one, const BGm* __restrict__ g, int j, int m)
{
    __shared__ double h[256];
    __shared__ double su[4][64];
    __shared__ double sw[64];
    const ArnOp A = arn_block<BAT>(one, g);
    const ArnLayout L{arn_slices(A.n), m};
    const int b = blockIdx.x;
    if (b >= L.G) return;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6, nv = j + 1;
    const int k = 64 * b + lane;
    const bool in = k < A.n;
    double x[8];
    arn_load_x(A, x, nv, k, in, wave);
    const double wk = in && wave == 0 ? A.w[k] : 0.0;
    const double* d = A.part + L.dots(0) + t;
    const bool ht = t < nv;
    double hp[32];
#pragma unroll
    for (int u = 0; u < 32; ++u) hp[u] = ht && u < L.G ? d[(size_t)(m + 2) * u] : 0.0;
    if (A.flags[0]) return;
    if (ht) {
        const double a = arn_h_sum(d, m, L.G, hp);
        h[t] = a;
        if (b == 0) A.part[L.h() + t] = a;
    }
    __syncthreads();
    su[wave][lane] = arn_vh(A, h, x, nv, k, in, wave);
    __syncthreads();
    if (wave == 0) {
        double q = 0.0, r = 0.0;
        if (in) {
            r = wk - ((su[0][lane] + su[1][lane]) + (su[2][lane] + su[3][lane]));
            A.w[k] = r;
            q = r * r;
        }
        sw[lane] = r;
        for (int off = 32; off > 0; off >>= 1) q += __shfl_xor(q, off, 64);
        if (lane == 0) A.part[L.norms() + b] = q;
    }
    __syncthreads();
    arn_slice_dots(A, A.part + L.dots(1) + (size_t)(m + 2) * b, nv, nv, k, in, sw[lane], wave, lane);
}

// the second pass if the test asks for it, |w|, v_{j+1} on the slice; slice 0
// (one thread for the serial part, the column and rotations staged in LDS)
// finishes column j of H.  Slices that start after slice 0 has set the flags
// skip v_{j+1}, which a stopped iteration never reads.
template <bool BAT>
__global__ void __launch_bounds__(256) k_arn_final(ArnOp one, const BGm* __restrict__ g, int j, int m)
{
    __shared__ double h[256];
    __shared__ double su[4][64];
    __shared__ double col[256];
    __shared__ double rc[256], rs[256];
    const ArnOp A = arn_block<BAT>(one, g);
    const ArnLayout L{arn_slices(A.n), m};
    const int b = blockIdx.x;
    if (b >= L.G) return;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6, nv = j + 1;
    const int k = 64 * b + lane;
    const bool in = k < A.n;
    const GmLayout G{m};
    double* st = A.st;
    // everything this block reads, before the tests
    double x[8];
    arn_load_x(A, x, nv, k, in, wave);
    const double wk = in && wave == 0 ? A.w[k] : 0.0;
    const double* d = A.part + L.dots(1) + t;
    const bool ht = t < nv;
    double hp[32];
#pragma unroll
    for (int u = 0; u < 32; ++u) hp[u] = ht && u < L.G ? d[(size_t)(m + 2) * u] : 0.0;
    const double ww = wave_sum_strided(A.part + L.dots(0) + nv, (size_t)m + 2, L.G, lane);
    const double q0 = wave_sum_strided(A.part + L.norms(), 1, L.G, lane);
    double c0 = 0.0, gj = 0.0, tl = 0.0, bn = 0.0;
    if (b == 0) {
        if (ht) c0 = A.part[L.h() + t];
        if (t < j) {
            rc[t] = st[G.cs() + t];
            rs[t] = st[G.sn() + t];
        }
        if (t == 0) {
            gj = st[G.g() + j];
            tl = st[G.tol()];
            bn = st[G.bnorm()];
        }
    }
    if (A.flags[0]) return;
    const bool again = arn_again(ww, q0);
    double hn, r = wk;
    if (again) {
        if (ht) h[t] = arn_h_sum(d, m, L.G, hp);
        __syncthreads();
        double hh = 0.0;   // |h'|^2, lanes i = lane mod 64, then a fixed tree
        for (int i = lane; i < nv; i += 64) hh += h[i] * h[i];
        for (int off = 32; off > 0; off >>= 1) hh += __shfl_xor(hh, off, 64);
        hn = sqrt(fmax(q0 - hh, 0.0));
        su[wave][lane] = arn_vh(A, h, x, nv, k, in, wave);
        __syncthreads();
        if (wave == 0) r = wk - ((su[0][lane] + su[1][lane]) + (su[2][lane] + su[3][lane]));
    } else {
        hn = sqrt(q0);
    }
    if (wave == 0 && hn > 0.0 && in) {
        const double inv = 1.0 / hn;
        A.V[(size_t)A.ld * nv + k] = r * inv;
    }
    if (b != 0) return;
    if (ht) col[t] = c0 + (again ? h[t] : 0.0);
    __syncthreads();
    if (t == 0) {
        col[j + 1] = hn;
        double a = col[0];   // entry i of the column, rotated by rotations < i
        for (int i = 0; i < j; ++i) {
            const double c = col[i + 1];
            col[i] = rc[i] * a + rs[i] * c;
            a = -rs[i] * a + rc[i] * c;
        }
        col[j] = a;
        const double rr = hypot(col[j], col[j + 1]);
        const double cj = rr == 0.0 ? 1.0 : col[j] / rr, sj = rr == 0.0 ? 0.0 : col[j + 1] / rr;
        col[j] = rr;
        col[j + 1] = 0.0;
        double* H = st + G.H() + (size_t)(m + 1) * j;
        for (int i = 0; i <= j + 1; ++i) H[i] = col[i];
        st[G.cs() + j] = cj;
        st[G.sn() + j] = sj;
        st[G.g() + j + 1] = -sj * gj;
        st[G.g() + j] = cj * gj;
        int* flags = A.flags;
        if (fabs(-sj * gj) <= tl * bn || hn == 0.0) {
            flags[0] = 1;
            flags[1] = j + 1;
        } else if (j + 1 == m) {
            flags[0] = 2;
            flags[1] = m;
        }
    }
}

// the three launches of Arnoldi step j (grid: slices, plans)
template <bool BAT>
static void arn_step(hipStream_t st, const ArnOp& one, const BGm* g, dim3 grid, int j, int m)
{
    k_arn_dots<BAT><<<grid, 256, 0, st>>>(one, g, j, m);
    k_arn_update<BAT><<<grid, 256, 0, st>>>(one, g, j, m);
    k_arn_final<BAT><<<grid, 256, 0, st>>>(one, g, j, m);
}

// every plan's two flags into one array (for one read-back)
__global__ void k_bgm_flags(const BGm* __restrict__ g, int nprob, int* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nprob) return;
    out[2 * i] = g[i].flags[0];
    out[2 * i + 1] = g[i].flags[1];
}

// y = H^-1 g (upper triangular, j x j), then w = V y on one 64-entry slice b
// per work-group of 256.  Every work-group solves for y itself (wave 0, the
// column sweep: the right-hand side in registers, lane l holding rows
// l + 64 q, y_i broadcast by v_readlane, H staged in LDS when j <= 64), the
// same arithmetic in each; its slice's V entries for the first 32 basis
// vectors are loaded before the sweep.  The waves split the basis vectors
// mod 4 and their sums are added in a fixed order (deterministic).
__device__ __forceinline__ void gm_finish_slice(const double* __restrict__ st, int m, int j,
                                                const double* __restrict__ V, int ldv, int n,
                                                double* __restrict__ w, int b)
{
    __shared__ double Hs[64 * 64];   // the leading j x j block of H when j <= 64
    __shared__ double y[256];
    __shared__ double su[4][64];
    const GmLayout L{m};
    const double* H = st + L.H();
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int k = 64 * b + lane;
    const bool in = k < n;
    double x[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        const int i = wave + 4 * u;
        x[u] = in && i < j ? V[k + (size_t)ldv * i] : 0.0;
    }
    const bool staged = j <= 64;
    if (staged)
        for (int e = t; e < j * j; e += 256) Hs[e] = H[e % j + (size_t)(m + 1) * (e / j)];
    __syncthreads();
    if (wave == 0) {
        double r[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int i = lane + 64 * q;
            r[q] = i < j ? st[L.g() + i] : 0.0;
        }
        for (int i = j - 1; i >= 0; --i) {
            const int q = i >> 6;
            const double ri = readlane_f64(q == 0 ? r[0] : q == 1 ? r[1] : q == 2 ? r[2] : r[3], i & 63);
            const double yi = ri / (staged ? Hs[i + j * i] : H[i + (size_t)(m + 1) * i]);
            if (lane == 0) y[i] = yi;
#pragma unroll
            for (int q2 = 0; q2 < 4; ++q2) {
                const int kk = lane + 64 * q2;
                if (kk < i)
                    r[q2] = __builtin_fma(-(staged ? Hs[kk + j * i] : H[kk + (size_t)(m + 1) * i]), yi, r[q2]);
            }
        }
    }
    __syncthreads();
    double acc = 0.0;
#pragma unroll
    for (int u = 0; u < 8; ++u)
        if (wave + 4 * u < j) acc = __builtin_fma(x[u], y[wave + 4 * u], acc);
    for (int i0 = wave + 32; i0 < j; i0 += 32) {
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int i = i0 + 4 * u;
            x[u] = in && i < j ? V[k + (size_t)ldv * i] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (i0 + 4 * u < j) acc = __builtin_fma(x[u], y[i0 + 4 * u], acc);
    }
    su[wave][lane] = acc;
    __syncthreads();
    if (wave == 0 && in) w[k] = (su[0][lane] + su[1][lane]) + (su[2][lane] + su[3][lane]);
}

// grid: 64-entry slices of n
__global__ void __launch_bounds__(256) k_gm_finish(const double* __restrict__ st, int m, int j,
                                                   const double* __restrict__ V, int ldv, int n,
                                                   double* __restrict__ w)
{
    gm_finish_slice(st, m, j, V, ldv, n, w, blockIdx.x);
}

// ---- the rest of a moving-body step over every plan (grid.y = plan; the
// single-plan kernels' arithmetic, element for element)
// amp = [0; f+ - f-], fbar = (f+ + f-)/2   (k_jump_avg)
__global__ void k_bgm_jump(const BGm* __restrict__ g)
{
    const BGm& G = g[blockIdx.y];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= G.n) return;
    const double a = G.fp[k], b = G.fm ? G.fm[k] : 0.0;
    G.amp[k] = 0.0;
    G.amp[G.n + k] = a - b;
    G.fbar[k] = 0.5 * (a + b);
}

// b = E f* (k_gather with one term per point on Nodes{Primal}), then
// b = -fbar + b (axpby)
__global__ void __launch_bounds__(256) k_bgm_rhs(const BGm* __restrict__ g)
{
    const BGm& G = g[blockIdx.y];
    const int lane = threadIdx.x & 63;
    const int p = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (p >= G.n) return;
    const Term& tt = G.tE[p];
    double acc0 = 0.0;
    if (lane < 25) {
        const int b = lane / 5, a = lane - 5 * b;
        const int row = tt.by + b, col = tt.bx + a;
        const double w = tt.wy[b] * tt.wx[a];
        if (w != 0.0 && row >= 0 && row < G.gny && col >= 0 && col < G.gnx)
            acc0 += tt.coef * w * G.grid[col + (size_t)G.gnx * row];
    }
    for (int off = 32; off > 0; off >>= 1) acc0 += __shfl_xor(acc0, off, 64);
    if (lane == 0) {
        const double bv = acc0;
        G.b[tt.ai] = -1.0 * G.fbar[tt.ai] + 1.0 * bv;
    }
}

// mode 0: x = P b (x0 of GMRES); mode 1: x += P w for plans that converged
// (the update after the steps)
__global__ void __launch_bounds__(256) k_bgm_prec_xb(const BGm* __restrict__ g, const unsigned* __restrict__ tiles,
                                                     int mode)
{
    const unsigned code = tiles[blockIdx.x];
    const BGm& G = g[code >> 16];
    if (mode == 1 && (G.flags[0] != 1 || G.flags[1] <= 0)) return;   // no steps: x stays P b
    const int i = (int)(code & 0xFFFF) * 4 + (int)(threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (i >= G.n) return;
    const double* v = mode == 0 ? G.b : G.w;
    double* x = G.amp;
    if (G.P32) {
        const double s = wave_dot_f32(G.P32 + (size_t)G.ld * i, v, G.n, lane);
        if (lane == 0) x[i] = s + (mode == 1 ? x[i] : 0.0);
    } else {
        const double s = band_row(G.band + (size_t)i * G.bw, G.nbr + (size_t)i * G.bw, G.bw, v, lane);
        if (lane == 0) x[i] = mode == 1 ? x[i] + s : s;
    }
}

// w = S x, then w = b - w (matvec_rows and axpby of gm_begin)
__global__ void __launch_bounds__(256) k_bgm_resid(const BGm* __restrict__ g, const unsigned* __restrict__ tiles)
{
    const unsigned code = tiles[blockIdx.x];
    const BGm& G = g[code >> 16];
    const int i = (int)(code & 0xFFFF) * 4 + (int)(threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (i >= G.n) return;
    const double s = wave_dot(G.S + (size_t)G.ld * i, G.amp, G.n, lane);
    if (lane == 0) {
        const double sx = 1.0 * s + 0.0;
        G.w[i] = 1.0 * G.b[i] + -1.0 * sx;
    }
}

__global__ void __launch_bounds__(1024) k_bgm_init(const BGm* __restrict__ g, int m, double tol)
{
    const BGm& G = g[blockIdx.x];
    gm_init_body(G.w, G.b, G.n, G.V, G.st, m, tol, G.flags);
}

// y = H^-1 g, w = V y for the plans that converged (grid: slices, plans)
__global__ void __launch_bounds__(256) k_bgm_finish(const BGm* __restrict__ g, int m)
{
    const BGm& G = g[blockIdx.y];
    if (G.flags[0] != 1) return;
    const int j = G.flags[1];
    if (j <= 0 || (int)blockIdx.x >= arn_slices(G.n)) return;
    gm_finish_slice(G.st, m, j, G.V, G.ld, G.n, G.w, blockIdx.x);
}

// s = amp[0, n)
__global__ void k_bgm_copy_s(const BGm* __restrict__ g)
{
    const BGm& G = g[blockIdx.y];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < G.n) G.s_out[k] = G.amp[k];
}

// amp = [0; f+ - f-] (2n),  fbar = (f+ + f-)/2
__global__ void k_jump_avg(const double* __restrict__ fp, const double* __restrict__ fm, int n,
                           double* __restrict__ amp, double* __restrict__ fbar)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double a = fp[k], b = fm ? fm[k] : 0.0;
    amp[k] = 0.0;
    amp[n + k] = a - b;
    fbar[k] = 0.5 * (a + b);
}

__global__ void k_identity(double* __restrict__ A, int n)   // A (zeroed) += I, column-major n x n
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) A[k + (size_t)n * k] = 1.0;
}

}  // namespace

// S v with S = -E L^-1 R (create_RTLinvR, src/matrix_operators.jl:9-30) at
// the current geometry, assembled row-major from the LGF table
static void schur_apply(ilh_plan* pl, MovingCache* mc, const double* v, double* out)
{
    matvec_rows(pl->ctx->stream, mc->S.as<double>(), solve_ld(pl->N), v, out, pl->N, pl->N, 1.0, 0);
}

// K (N x N column-major, overwritten by its LU factors; N = n, or the point
// count when n = 0) -> Minv = K^-1, row-major with leading dimension
// solve_ld(N) (allocated by the caller)
void ilh::inverse_rowmajor(ilh_plan* pl, double* K, double* Minv, int n)
{
    const int N = n > 0 ? n : pl->N;
    hipStream_t st = pl->ctx->stream;
    DevBuf piv, info, Sinv;
    piv.ensure(sizeof(int) * (size_t)N);
    info.ensure(sizeof(int));
    const int bad = lu_factor(pl, K, N, piv.as<int>(), info.as<int>());
    if (bad) fail(ILH_ERR_SINGULAR, "Schur complement is singular (zero pivot at column " + std::to_string(bad) + ")");
    Sinv.ensure(sizeof(double) * (size_t)N * N);
    fill_zero(st, Sinv.as<double>(), (size_t)N * N);
    k_identity<<<(N + 255) / 256, 256, 0, st>>>(Sinv.as<double>(), N);
    lu_solve_multi(pl, K, piv.as<int>(), N, Sinv.as<double>(), N);
    transpose(st, Sinv.as<double>(), Minv, N, N, solve_ld(N));
    ILH_HIP(hipStreamSynchronize(st));
}

// reference factorisation at the current geometry: P = S^-1, row-major, kept
// in FP32.  The Krylov method (vectors, Gram-Schmidt, the residual test and S
// itself) stays FP64, so the FP32 preconditioner changes the path but not the
// accuracy of the solution: S P differs from S S_ref^-1 by about
// cond(S) 2^-24 (~0.02 at the benchmark's cond 3e5), far less than the
// geometry change it already absorbs, and it halves the bytes of one of the
// two dense matvecs per GMRES step.
static void moving_reference(ilh_plan* pl, MovingCache* mc)
{
    const int N = pl->N;
    const size_t nl = (size_t)N * solve_ld(N);
    DevBuf Pd;
    Pd.ensure(sizeof(double) * nl);
    exact_inverse_rowmajor(pl, Pd.as<double>());
    mc->P32.ensure(sizeof(float) * nl);
    k_to_f32<<<(unsigned)std::min<size_t>(4096, (nl + 255) / 256), 256, 0, pl->ctx->stream>>>(
        Pd.as<double>(), mc->P32.as<float>(), nl);
    ILH_HIP(hipGetLastError());
    ILH_HIP(hipStreamSynchronize(pl->ctx->stream));   // Pd is released on return
    mc->have_ref = true;
    mc->ref_age = 0;
}

// GMRES(m) on S x = b, right-preconditioned by P (the FP32 reference
// inverse), classical Gram-Schmidt with a conditional re-orthogonalisation;
// x0 = P b.  The Arnoldi steps,
// Givens rotations and the residual test run on the device; the host enqueues
// steps in chunks (the previous solve's count first) and reads back two
// flags after each chunk, steps after convergence returning at once.  The
// phases below issue work on the context's current stream without waiting,
// so that the batched form can run many plans' solves side by side.

// z (+)= P v with this solve's preconditioner: the learned band or the FP32
// reference inverse; skipped once flags[0] is set (flags may be null)
static void prec_apply(ilh_plan* pl, MovingCache* mc, const double* v, double* z, const int* flags, int accumulate)
{
    const int N = pl->N;
    hipStream_t st = pl->ctx->stream;
    if (mc->use_band) {
        band_apply(st, mc->band.as<double>(), pl->nbr.as<int>(), lpc_width(pl->lpc), v, z, N, flags, accumulate);
    } else {
        k_gm_matvec_f32<<<(unsigned)((N + 3) / 4), 256, 0, st>>>(mc->P32.as<float>(), solve_ld(N), v, z, N, flags,
                                                                 accumulate);
        ILH_HIP(hipGetLastError());
    }
}

// x0 = P b, r0 = b - S x0, v_0, g
static void gm_begin(ilh_plan* pl, MovingCache* mc, const double* b, double* x, double tol, int m)
{
    const int N = pl->N, ld = solve_ld(N);
    hipStream_t st = pl->ctx->stream;
    const GmLayout L{m};
    mc->gm.ensure(sizeof(double) * L.size());
    mc->gmf.ensure(sizeof(int) * 2);
    mc->arn.ensure(sizeof(double) * ArnLayout{arn_slices(N), m}.size());
    double* w = mc->w.as<double>();
    (void)ld;
    prec_apply(pl, mc, b, x, nullptr, 0);
    schur_apply(pl, mc, x, w);
    axpby(st, w, b, 1.0, -1.0, N);
    k_gm_init<<<1, 1024, 0, st>>>(w, b, N, mc->V.as<double>(), mc->gm.as<double>(), m, tol, mc->gmf.as<int>());
    ILH_HIP(hipGetLastError());
}

// Arnoldi steps j0 .. j1-1, then the two flags into pinned host memory
static void gm_issue(ilh_plan* pl, MovingCache* mc, int j0, int j1, int m)
{
    const int N = pl->N, ld = solve_ld(N);
    hipStream_t st = pl->ctx->stream;
    const unsigned gR = (unsigned)((N + 3) / 4);
    double* V = mc->V.as<double>();
    int* flags = mc->gmf.as<int>();
    const ArnOp A{V, ld, mc->w.as<double>(), N, mc->gm.as<double>(), flags, mc->arn.as<double>()};
    for (int j = j0; j < j1; ++j) {
        prec_apply(pl, mc, V + (size_t)ld * j, mc->z.as<double>(), flags, 0);
        k_gm_matvec<<<gR, 256, 0, st>>>(mc->S.as<double>(), ld, mc->z.as<double>(), mc->w.as<double>(), N, flags);
        arn_step<false>(st, A, nullptr, dim3((unsigned)arn_slices(N)), j, m);
    }
    ILH_HIP(hipGetLastError());
    ILH_HIP(hipMemcpyAsync(mc->pinned(1), flags, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
}

// x += P (V y) after `steps` converged steps
static void gm_end(ilh_plan* pl, MovingCache* mc, double* x, int m, int steps)
{
    const int N = pl->N, ld = solve_ld(N);
    hipStream_t st = pl->ctx->stream;
    if (steps <= 0) return;
    mc->last_iters = steps;
    k_gm_finish<<<(unsigned)arn_slices(N), 256, 0, st>>>(mc->gm.as<double>(), m, steps, mc->V.as<double>(), ld, N,
                                                         mc->w.as<double>());
    ILH_HIP(hipGetLastError());
    prec_apply(pl, mc, mc->w.as<double>(), x, nullptr, 1);
}

// the whole solve on one plan: the steps taken, or -1 without convergence
static int moving_gmres(ilh_plan* pl, MovingCache* mc, const double* b, double* x, double tol, int m)
{
    hipStream_t st = pl->ctx->stream;
    gm_begin(pl, mc, b, x, tol, m);
    int j = 0, state = 0;
    int chunk = std::max(1, std::min(m, mc->last_iters));
    const int* hflags = (const int*)mc->pinned(1);
    while (j < m) {
        const int j1 = std::min(m, j + chunk);
        gm_issue(pl, mc, j, j1, m);
        j = j1;
        ILH_HIP(hipStreamSynchronize(st));
        state = hflags[0];
        if (state) break;
        chunk = 2;
    }
    if (state != 1) return -1;
    const int steps = hflags[1];
    gm_end(pl, mc, x, m, steps);
    ILH_HIP(hipStreamSynchronize(st));
    return steps;
}

static constexpr int REFACTOR_STEPS = 8;   // step budget floor of a re-solve after refactoring

// Buffers and, on the first call, the reference inverse (which waits for the
// device); nothing else of the current geometry.
static MovingCache* moving_setup(ilh_plan* pl, int m)
{
    const int N = pl->N;
    const size_t ng = pl->g.n(GK_PRIMAL);
    need_margin(pl, "Dirichlet solve");
    if (!pl->mov) pl->mov = new MovingCache();
    MovingCache* mc = pl->mov;
    mc->amp.ensure(sizeof(double) * 2 * (size_t)N);
    // basis columns, stride solve_ld(N): m + 1, and at least the 9 a re-solve
    // after a refactorisation may take (moving_refactor_solve)
    mc->V.ensure(sizeof(double) * (size_t)solve_ld(N) * (std::max(m, REFACTOR_STEPS) + 1));
    for (DevBuf* bb : {&mc->w, &mc->z, &mc->b, &mc->fbar}) bb->ensure(sizeof(double) * (size_t)N);
    mc->grid.ensure(sizeof(double) * ng);
    mc->S.ensure(sizeof(double) * (size_t)N * solve_ld(N));
    (void)mc->pinned(1);
    // with a learned preconditioner attached no factorisation is needed
    mc->use_band = pl->lpc != nullptr;
    if (mc->use_band) {
        plan_neighbours(pl, lpc_half_band(pl->lpc));
        mc->band.ensure(sizeof(double) * (size_t)N * lpc_width(pl->lpc));
    } else if (!mc->have_ref) {
        moving_reference(pl, mc);
    }
    return mc;
}

// The current geometry's merged sources [R, D_s] and S, issued on the
// context's current stream without waiting.
static void moving_geometry(ilh_plan* pl, MovingCache* mc)
{
    if (!mc->src_ok) {
        merge_sources(pl, OP_R_PRIMAL, OP_DS, GK_PRIMAL, mc->terms1, mc->src1);
        mc->src_ok = true;
    }
    if (!mc->S_ok) {
        // create_RTLinvR (scale 1), stored row-major for the matvecs
        assemble(pl, OP_E_PRIMAL, OP_R_PRIMAL, GK_PRIMAL, -pl->g.dx * pl->g.dx, false, mc->S.as<double>(),
                 solve_ld(pl->N), true);
        mc->S_ok = true;
    }
    if (mc->use_band && !mc->band_ok) {
        // the network's band for this geometry (features, BF16 forward pass)
        lpc_band(pl->lpc, pl, mc->band.as<double>());
        mc->band_ok = true;
    }
}

// amp = [0; d], f* = L^-1 D_s d, b = E f* - fbar  (test/literate/problems.jl:281-291);
// the convolution left out when `conv` is false (the batched form runs all
// plans' convolutions in one fused call between the two halves)
static void moving_rhs_jump(ilh_plan* pl, MovingCache* mc, const double* d_fp, const double* d_fm)
{
    const int N = pl->N;
    k_jump_avg<<<(N + 255) / 256, 256, 0, pl->ctx->stream>>>(d_fp, d_fm, N, mc->amp.as<double>(),
                                                            mc->fbar.as<double>());
}

static void moving_rhs_finish(ilh_plan* pl, MovingCache* mc)
{
    gather_op(pl, OP_E_PRIMAL, mc->grid.as<double>(), mc->b.as<double>());
    axpby(pl->ctx->stream, mc->b.as<double>(), mc->fbar.as<double>(), -1.0, 1.0, pl->N);
}

static void moving_rhs(ilh_plan* pl, MovingCache* mc, const double* d_fp, const double* d_fm)
{
    moving_rhs_jump(pl, mc, d_fp, d_fm);
    conv_sources(pl, GK_PRIMAL, mc->src1, mc->amp.as<double>(), mc->grid.as<double>());
    moving_rhs_finish(pl, mc);
}

// f = L^-1 (R s + D_s d) with s in amp[0, N) (the convolution left out when
// `conv` is false: the batched form runs it for all plans at once)
static void moving_final(ilh_plan* pl, MovingCache* mc, double* d_f, double* d_s, bool conv = true)
{
    const int N = pl->N;
    double* amp = mc->amp.as<double>();
    if (conv) conv_sources(pl, GK_PRIMAL, mc->src1, amp, d_f);
    ILH_HIP(hipMemcpyAsync(d_s, amp, sizeof(double) * N, hipMemcpyDeviceToDevice, pl->ctx->stream));
    ++mc->ref_age;
}

// The convolutions L^-1 (sources of amp) of many moving-body plans as one
// fused call (four launches in all), out[i] the grid field of plan i; the
// plans' row-pair counts stay on the device.  False when a plan's FFT periods
// are outside the fused passes' set (the caller convolves plan by plan).
static bool moving_conv_fused(ilh_ctx* ctx, int slot, const std::vector<ilh_plan*>& pls,
                              const std::vector<MovingCache*>& mcs, const std::vector<double*>& out)
{
    std::vector<BatchProb> probs;
    for (size_t k = 0; k < pls.size(); ++k) {
        ilh_plan* pl = pls[k];
        const ConvPlan& C = get_conv(pl, GK_PRIMAL);
        if (!fused_supported(C.Px, C.Py)) return false;
        const SourceSet& src = mcs[k]->src1;
        BatchProb b;
        std::memset(&b, 0, sizeof(b));
        b.amp = mcs[k]->amp.as<double>();
        b.n = pl->N;
        b.pinfo = src.pinfo;
        b.cells = src.bins.cells;
        b.cai = src.bins.cai;
        b.cw = src.bins.cw;
        b.npairs = src.npairs;
        b.npairs_dev = src.npairs_dev;
        b.nx = C.nx;
        b.ny = C.ny;
        b.Px = C.Px;
        b.Py = C.Py;
        b.T1 = C.T1.as<double2>();
        b.slot = src.slot_of_row;
        b.spec = C.spec.as<double>();
        b.T2 = C.T2.as<double2>();
        b.out = out[k];
        b.twx = fft_twiddles(ctx, C.Px);
        b.twy = fft_twiddles(ctx, C.Py);
        probs.push_back(b);
    }
    if (probs.empty()) return true;
    FusedBatch*& fb = ctx->fused_mv[slot];
    if (!fb || !same_probs(fused_probs(fb), probs)) {
        if (fb) {
            ILH_HIP(hipStreamSynchronize(ctx->stream));
            fused_destroy(fb);
            fb = nullptr;
        }
        fb = fused_create(ctx, probs);
    }
    fused_run(fb, 1);
    return true;
}

// the solve after a failed GMRES: refactor at the current geometry, solve again
// The re-solve after refactoring at the current geometry may take up to
// max(m, REFACTOR_STEPS) steps (the basis is sized for that in
// moving_setup): its preconditioner is this geometry's inverse rounded to
// FP32, with which GMRES converges in two or three.
static int moving_refactor_solve(ilh_plan* pl, MovingCache* mc, double tol, int m)
{
    // the exact fallback: this geometry's own factorisation (also when the
    // learned band was the preconditioner; the next solve uses the band again)
    moving_reference(pl, mc);
    const bool band = mc->use_band;
    mc->use_band = false;
    const int m2 = std::max(m, REFACTOR_STEPS);
    if (mc->V.bytes < sizeof(double) * (size_t)solve_ld(pl->N) * (m2 + 1))
        fail(ILH_ERR_STATE, "GMRES basis smaller than the re-solve needs");
    const int it2 = moving_gmres(pl, mc, mc->b.as<double>(), mc->amp.as<double>(), tol, m2);
    mc->use_band = band;
    if (it2 < 0) fail(ILH_ERR_SINGULAR, "GMRES did not converge with the current geometry's factor");
    return m + it2;
}

static void check_moving_args(double tol, int max_iter)
{
    if (!(tol > 0.0) || tol >= 1.0) fail(ILH_ERR_ARG, "tol must be in (0, 1)");
    if (max_iter < 1 || max_iter > 200) fail(ILH_ERR_ARG, "max_iter must be in [1, 200]");
}

// The moving-body solves of this plan use the learned band of `model` as
// their GMRES preconditioner (null: the FP32 reference inverse again).
extern "C" int ilh_plan_set_preconditioner(ilh_plan* pl, ilh_lpc* model)
{
    return guarded([&] {
        if (!pl) fail(ILH_ERR_ARG, "null plan");
        pl->lpc = model;
        if (pl->mov) {
            pl->mov->band_ok = false;
            pl->mov->use_band = false;
        }
    });
}

extern "C" int ilh_dirichlet_solve_moving(ilh_plan* pl, const double* fp, const double* fm, double* f, double* s,
                                          int mem, double tol, int max_iter, int* iters)
{
    return guarded([&] {
        Io io(pl, mem);
        check_moving_args(tol, max_iter);
        const int N = pl->N;
        const size_t ng = pl->g.n(GK_PRIMAL);
        hipStream_t st = pl->ctx->stream;
        const double* d_fp = io.in(fp, N, pl->io_a);
        const double* d_fm = fm ? io.in(fm, N, pl->io_c) : nullptr;
        double* d_f = io.out(f, ng, pl->io_b);
        double* d_s = io.out(s, N, pl->io_d);
        int it = 0;
        if (N == 0) {
            fill_zero(st, d_f, ng);
        } else {
            const int m = max_iter;
            MovingCache* mc = moving_setup(pl, m);
            moving_geometry(pl, mc);
            moving_rhs(pl, mc, d_fp, d_fm);
            // s lands in amp[0, N): the sources of the final convolution
            it = moving_gmres(pl, mc, mc->b.as<double>(), mc->amp.as<double>(), tol, m);
            if (it < 0) it = moving_refactor_solve(pl, mc, tol, m);
            moving_final(pl, mc, d_f, d_s);
        }
        if (iters) *iters = it;
        io.back(f, d_f, ng);
        io.back(s, d_s, N);
        io.done();
    });
}

// Batched moving-body solves: many plans of one context (independent bodies
// or trajectories, each after its own ilh_plan_update_points) solved side by
// side.  Reference factorisations (first calls) run plan by plan on the
// context stream; then each plan's geometry (sources, S), right-hand side,
// GMRES steps and final convolution run on the side streams, the GMRES steps of all plans advancing in lockstep chunks
// with one wait per chunk for the whole batch.  Device pointers only.
extern "C" int ilh_dirichlet_solve_moving_batched(int nprob, ilh_plan* const* plans, const double* const* fb_plus,
                                                  const double* const* fb_minus, double* const* f, double* const* s,
                                                  double tol, int max_iter, int* iters)
{
    return guarded([&] {
        check_moving_args(tol, max_iter);
        if (!fb_plus || !f || !s) fail(ILH_ERR_ARG, "null array of pointers");
        ilh_ctx* ctx = batch_ctx(nprob, plans);
        const int m = max_iter;
        std::vector<MovingCache*> mc(nprob, nullptr);
        for (int i = 0; i < nprob; ++i)
            if (plans[i]->N > 0) mc[i] = moving_setup(plans[i], m);
        // learned bands first, plan by plan on the context stream (the model's
        // work buffers are shared by its plans)
        for (int i = 0; i < nprob; ++i)
            if (mc[i] && mc[i]->use_band && !mc[i]->band_ok) {
                lpc_band(plans[i]->lpc, plans[i], mc[i]->band.as<double>());
                mc[i]->band_ok = true;
            }
        const hipStream_t main = ctx->stream;
        bool joined = false;
        // on any exit: the context stream back, and (if the step ended before
        // its join) the side streams' work ordered before later work on it
        struct Restore {
            ilh_ctx* c;
            hipStream_t m;
            const bool* joined;
            ~Restore()
            {
                c->stream = m;
                if (!*joined)
                    for (int k = 0; k < ilh_ctx::POOL; ++k) {
                        (void)hipEventRecord(c->ev[k], c->pool[k]);
                        (void)hipStreamWaitEvent(m, c->ev[k], 0);
                    }
            }
        } restore{ctx, main, &joined};
        // fork the side streams from / join them into the context stream
        auto fork = [&] {
            ILH_HIP(hipEventRecord(ctx->ev[ilh_ctx::POOL], main));
            for (int k = 0; k < ilh_ctx::POOL; ++k)
                ILH_HIP(hipStreamWaitEvent(ctx->pool[k], ctx->ev[ilh_ctx::POOL], 0));
        };
        auto join = [&] {
            ctx->stream = main;
            for (int k = 0; k < ilh_ctx::POOL; ++k) {
                ILH_HIP(hipEventRecord(ctx->ev[k], ctx->pool[k]));
                ILH_HIP(hipStreamWaitEvent(main, ctx->ev[k], 0));
            }
            joined = true;
        };
        // the merged sources [R, D_s] (and E's terms, for S) of every plan whose
        // geometry changed, built together on the context stream
        std::vector<SourceJob> sj;
        for (int i = 0; i < nprob; ++i)
            if (mc[i] && !mc[i]->src_ok) sj.push_back(SourceJob{plans[i], &mc[i]->terms1, &mc[i]->src1});
        merge_sources_batched(ctx, sj, OP_R_PRIMAL, OP_DS, OP_E_PRIMAL, GK_PRIMAL);
        for (const SourceJob& j : sj) j.pl->mov->src_ok = true;
        // every plan's LGF table and spectrum exist before the fork: the side
        // streams' assemblies read the table, and the first convolution runs
        // on the context stream beside them
        for (int i = 0; i < nprob; ++i)
            if (mc[i]) (void)get_conv(plans[i], GK_PRIMAL);
        fork();
        auto on = [&](int i) { ctx->stream = ctx->pool[i % ilh_ctx::POOL]; };
        std::vector<int> state(nprob, 0), it(nprob, 0);
        std::vector<ilh_plan*> live_pl;
        std::vector<MovingCache*> live_mc;
        std::vector<double*> grid_out, f_out;
        for (int i = 0; i < nprob; ++i) {
            on(i);
            ilh_plan* pl = plans[i];
            if (!mc[i]) {
                fill_zero(ctx->stream, f[i], pl->g.n(GK_PRIMAL));
                state[i] = 1;
                continue;
            }
            moving_geometry(pl, mc[i]);   // S of the new geometry (the sources are built above)
            live_pl.push_back(pl);
            live_mc.push_back(mc[i]);
            grid_out.push_back(mc[i]->grid.as<double>());
            f_out.push_back(f[i]);
        }
        // the side streams (S and the learned band of every plan) are joined
        // where the step first needs them, before the preconditioner: the
        // first convolution and the right-hand side run beside them
        ctx->stream = main;
        // every plan's step around the GMRES steps as batched launches over all
        // plans on the context stream; the GMRES steps in lockstep, one launch
        // per phase, one flag read-back per chunk of steps
        std::vector<BGm> hg;
        std::vector<int> gi;   // plan of each descriptor
        std::vector<unsigned> tiles;
        int chunk = 1, maxn = 1;
        for (int i = 0; i < nprob; ++i) {
            if (!mc[i]) continue;
            ilh_plan* pl = plans[i];
            MovingCache* c = mc[i];
            c->gm.ensure(sizeof(double) * GmLayout{m}.size());
            c->gmf.ensure(sizeof(int) * 2);
            c->arn.ensure(sizeof(double) * ArnLayout{arn_slices(pl->N), m}.size());
            BGm G{};
            G.P32 = c->use_band ? nullptr : c->P32.as<float>();
            G.band = c->use_band ? c->band.as<double>() : nullptr;
            G.nbr = c->use_band ? pl->nbr.as<int>() : nullptr;
            G.bw = c->use_band ? lpc_width(pl->lpc) : 0;
            G.S = c->S.as<double>();
            G.ld = solve_ld(pl->N);
            G.n = pl->N;
            G.V = c->V.as<double>();
            G.w = c->w.as<double>();
            G.z = c->z.as<double>();
            G.st = c->gm.as<double>();
            G.flags = c->gmf.as<int>();
            G.part = c->arn.as<double>();
            G.fp = fb_plus[i];
            G.fm = fb_minus ? fb_minus[i] : nullptr;
            G.amp = c->amp.as<double>();
            G.fbar = c->fbar.as<double>();
            G.b = c->b.as<double>();
            G.grid = c->grid.as<double>();
            G.s_out = s[i];
            G.tE = get_terms(pl, OP_E_PRIMAL);
            G.gnx = pl->g.nx(GK_PRIMAL);
            G.gny = pl->g.ny(GK_PRIMAL);
            if ((pl->N + 3) / 4 > 65535 || hg.size() >= 65535) fail(ILH_ERR_ARG, "batched GMRES too large");
            for (int b = 0; b < (pl->N + 3) / 4; ++b) tiles.push_back(((unsigned)hg.size() << 16) | (unsigned)b);
            hg.push_back(G);
            gi.push_back(i);
            chunk = std::max(chunk, std::min(m, c->last_iters));
            maxn = std::max(maxn, pl->N);
        }
        const int nd = (int)hg.size();
        if (nd > 0) {
            const size_t bytes_g = sizeof(BGm) * nd, off_t = (bytes_g + 255) / 256 * 256;
            const size_t off_f = off_t + (sizeof(unsigned) * tiles.size() + 255) / 256 * 256;
            const size_t need = off_f + sizeof(int) * 2 * nd;
            if (need > ctx->mv_cap) {
                ILH_HIP(hipStreamSynchronize(main));
                if (ctx->mv_dev) (void)hipFree(ctx->mv_dev);
                ctx->mv_dev = nullptr;
                ctx->mv_cap = 0;
                ILH_HIP(hipMalloc(&ctx->mv_dev, need));
                ctx->mv_cap = need;
            }
            if ((size_t)(2 * nd) > ctx->mv_host_cap) {
                if (ctx->mv_host) (void)hipHostFree(ctx->mv_host);
                ctx->mv_host = nullptr;
                ctx->mv_host_cap = 0;
                ILH_HIP(hipHostMalloc((void**)&ctx->mv_host, sizeof(int) * 2 * nd, hipHostMallocDefault));
                ctx->mv_host_cap = 2 * nd;
            }
            char* base = (char*)ctx->mv_dev;
            BGm* dg = (BGm*)base;
            unsigned* dt = (unsigned*)(base + off_t);
            int* df = (int*)(base + off_f);
            ILH_HIP(hipMemcpyAsync(dg, hg.data(), bytes_g, hipMemcpyHostToDevice, main));
            ILH_HIP(hipMemcpyAsync(dt, tiles.data(), sizeof(unsigned) * tiles.size(), hipMemcpyHostToDevice, main));
            const unsigned nt = (unsigned)tiles.size();
            const dim3 gel((unsigned)((maxn + 255) / 256), (unsigned)nd);
            // amp = [0; d], fbar; f* = L^-1 D_s d of every plan (one fused call);
            // b = E f* - fbar; x0 = P b, r0 = b - S x0, v_0
            k_bgm_jump<<<gel, 256, 0, main>>>(dg);
            if (!moving_conv_fused(ctx, 0, live_pl, live_mc, grid_out))
                for (size_t k = 0; k < live_pl.size(); ++k)
                    conv_sources(live_pl[k], GK_PRIMAL, live_mc[k]->src1, live_mc[k]->amp.as<double>(), grid_out[k]);
            k_bgm_rhs<<<dim3((unsigned)((maxn + 3) / 4), (unsigned)nd), 256, 0, main>>>(dg);
            join();
            k_bgm_prec_xb<<<nt, 256, 0, main>>>(dg, dt, 0);
            k_bgm_resid<<<nt, 256, 0, main>>>(dg, dt);
            k_bgm_init<<<(unsigned)nd, 1024, 0, main>>>(dg, m, tol);
            ILH_HIP(hipGetLastError());
            int j = 0;
            for (;;) {
                const int j1 = std::min(m, j + chunk);
                for (int jj = j; jj < j1; ++jj) {
                    k_bgm_prec<<<nt, 128, 0, main>>>(dg, dt, jj);
                    k_bgm_matvec<<<nt, 256, 0, main>>>(dg, dt);
                    arn_step<true>(main, ArnOp{}, dg, dim3((unsigned)arn_slices(maxn), (unsigned)nd), jj, m);
                }
                k_bgm_flags<<<(unsigned)((nd + 255) / 256), 256, 0, main>>>(dg, nd, df);
                ILH_HIP(hipGetLastError());
                ILH_HIP(hipMemcpyAsync(ctx->mv_host, df, sizeof(int) * 2 * nd, hipMemcpyDeviceToHost, main));
                ILH_HIP(hipStreamSynchronize(main));   // also keeps hg and tiles alive for their copies
                j = j1;
                bool running = false;
                for (int d = 0; d < nd; ++d) {
                    const int i = gi[d];
                    if (state[i]) continue;
                    state[i] = ctx->mv_host[2 * d];
                    if (state[i] == 1) it[i] = ctx->mv_host[2 * d + 1];
                    else if (!state[i] && j >= m) state[i] = 2;
                    running = running || !state[i];
                }
                if (!running) break;
                chunk = 2;
            }
            // x += P (V y) for the plans that converged
            k_bgm_finish<<<dim3((unsigned)arn_slices(maxn), (unsigned)nd), 256, 0, main>>>(dg, m);
            k_bgm_prec_xb<<<nt, 256, 0, main>>>(dg, dt, 1);
            ILH_HIP(hipGetLastError());
            for (int i = 0; i < nprob; ++i)
                if (mc[i] && state[i] == 1 && it[i] > 0) mc[i]->last_iters = it[i];
            // plans whose reference no longer preconditions are refactored and
            // solved one by one
            for (int i = 0; i < nprob; ++i) {
                if (!mc[i] || state[i] == 1) continue;
                it[i] = moving_refactor_solve(plans[i], mc[i], tol, m);
            }
            // f = L^-1 (R s + D_s d) of every plan in one fused call; s out
            if (!moving_conv_fused(ctx, 1, live_pl, live_mc, f_out))
                for (size_t k = 0; k < live_pl.size(); ++k)
                    conv_sources(live_pl[k], GK_PRIMAL, live_mc[k]->src1, live_mc[k]->amp.as<double>(), f_out[k]);
            k_bgm_copy_s<<<gel, 256, 0, main>>>(dg);
            ILH_HIP(hipGetLastError());
            for (int i = 0; i < nprob; ++i)
                if (mc[i]) ++mc[i]->ref_age;
        }
        if (!joined) join();
        if (iters)
            for (int i = 0; i < nprob; ++i) iters[i] = it[i];
    });
}

extern "C" int ilh_profile_dirichlet(ilh_plan* pl, const double* fp, const double* fm, double* f, double* s, int reps,
                                     double* stage_ms, double* total_ms)
{
    return guarded([&] {
        Io io(pl, ILH_MEM_DEVICE);
        if (reps < 1 || reps > 1000000) fail(ILH_ERR_ARG, "reps must be in [1, 1000000]");
        SolveCache* sc = prepare_dirichlet(pl, fm != nullptr);
        const int N = pl->N;
        if (N == 0) fail(ILH_ERR_ARG, "no surface points");
        hipStream_t st = pl->ctx->stream;
        for (hipEvent_t& e : pl->ctx->tev)
            if (!e) ILH_HIP(hipEventCreate(&e));
        const hipEvent_t e0 = pl->ctx->tev[0], e1 = pl->ctx->tev[1];
        float ms = 0.f;
        auto timed = [&](auto&& body) {
            ILH_HIP(hipEventRecord(e0, st));
            for (int r = 0; r < reps; ++r) body();
            ILH_HIP(hipEventRecord(e1, st));
            // polled rather than waited on: the blocking wait's wake-up would
            // land in the caller's wall-clock measurement
            hipError_t q;
            while ((q = hipEventQuery(e1)) == hipErrorNotReady) {
            }
            ILH_HIP(q);
            ILH_HIP(hipEventElapsedTime(&ms, e0, e1));
            return (double)ms / reps;
        };
        const int ld = solve_ld(N);
        auto amp = [&] {
            dirichlet_amp(st, sc->M[0].as<double>(), sc->M[1].as<double>(), ld, fp, fm, sc->amp1.as<double>(), s, N);
        };
        // whole solves back to back, then each stage alone (stages are idempotent)
        const double tot = timed([&] {
            amp();
            conv_sources(pl, GK_PRIMAL, sc->src1, sc->amp1.as<double>(), f);
        });
        if (stage_ms) {
            stage_ms[0] = timed(amp);
            for (int k = 1; k < 4; ++k)
                stage_ms[k] = timed([&] { conv_stage(pl, GK_PRIMAL, sc->src1, sc->amp1.as<double>(), f, k - 1); });
        }
        if (total_ms) *total_ms = tot;
    });
}
