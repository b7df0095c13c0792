/*
 * immersedhip.h -- C ABI of libimmersedhip.so, the MI355X (gfx950) immersed-layer
 * hot path: DDF regularization/interpolation, staggered-grid stencils, the
 * lattice-Green's-function inverse Laplacian and the constrained (Schur
 * complement) Dirichlet/Neumann Poisson solves of ImmersedLayers.jl.
 *
 * Every reference citation below is a path inside JuliaIBPM/ImmersedLayers.jl.
 *
 * Conventions
 *  - Grid: (NX, NY) = size(g) (dual-node counts), (I0x, I0y) = origin(g) (1-based,
 *    as CartesianGrids reports it), dx = cellsize(g).
 *  - Layouts are the reference's, column-major (x fastest):
 *      Nodes{Primal} (NX-1)x(NY-1), Nodes{Dual} NX x NY,
 *      Edges{Primal} = [u (NX x (NY-1)); v ((NX-1) x NY)] in one buffer,
 *      ScalarData N, VectorData [u(N); v(N)].
 *  - In-place, output first, as the reference's `op!(out, in, cache)`.
 *  - `mem` says where array pointers live: ILH_MEM_HOST (synchronous copy in and
 *    out on the plan's stream) or ILH_MEM_DEVICE (device pointers, the call is
 *    asynchronous on the ctx stream; synchronize with ilh_ctx_synchronize).
 *  - Scaling is GridScaling (the reference default, src/ImmersedLayers.jl:81).
 *  - Calls on one plan are not re-entrant (the reference's cache owns scratch,
 *    src/cache.jl:170-178); distinct ctxs are independent.
 *  - Return value: 0 = ILH_OK, negative = error; ilh_last_error() gives the
 *    message of the last failing call on this thread.
 */
#ifndef IMMERSEDHIP_H
#define IMMERSEDHIP_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct ilh_ctx ilh_ctx;
typedef struct ilh_plan ilh_plan;

typedef struct {
    int NX, NY;     /* size(g) */
    int I0x, I0y;   /* origin(g) */
    double dx;      /* cellsize(g) */
} ilh_grid;

enum {
    ILH_OK = 0,
    ILH_ERR_ARG = -1,          /* bad argument / unsupported option */
    ILH_ERR_DIM = -2,          /* DimensionMismatch (test/surface_ops.jl:524) */
    ILH_ERR_NONFINITE = -3,    /* NaN/Inf in an input */
    ILH_ERR_SINGULAR = -4,     /* zero pivot in the Schur factorisation */
    ILH_ERR_HIP = -5,          /* HIP runtime error */
    ILH_ERR_OOM = -6,          /* device allocation failed */
    ILH_ERR_RANGE = -7,        /* surface point support leaves the grid interior */
    ILH_ERR_STATE = -8         /* call order (e.g. solve before factor) */
};

enum { ILH_MEM_HOST = 0, ILH_MEM_DEVICE = 1 };
/* DDF kinds (CartesianGrids DDF types; default Yang3, src/ImmersedLayers.jl:83) */
enum { ILH_DDF_YANG3 = 0, ILH_DDF_ROMA = 1, ILH_DDF_M4PRIME = 2, ILH_DDF_WITCHHAT = 3 };
enum { ILH_SCALING_GRID = 0 };
/* grid data kinds */
enum { ILH_NODES_PRIMAL = 0, ILH_NODES_DUAL = 1, ILH_EDGES_PRIMAL = 2 };
/* Schur-complement kinds (src/matrix_operators.jl) */
enum { ILH_SCHUR_RTLINVR = 0, ILH_SCHUR_CLINVCT = 1, ILH_SCHUR_GLINVD = 2, ILH_SCHUR_NRTRN = 3,
       ILH_SCHUR_CL2INVCT = 4,
       /* create_GLinvD_symm of the vector cache (src/matrix_operators.jl:165-185):
          -G_s^symm L^-1 D_s^symm on VectorData, 2N x 2N */
       ILH_SCHUR_GLINVD_SYMM = 5,
       /* create_CLinvCT_scalar of the vector cache (src/matrix_operators.jl:71-91) is
          the same matrix as create_CLinvCT of the scalar cache: its scalar curl
          regularizes n f onto the edges exactly as the scalar cache does */
       ILH_SCHUR_CLINVCT_SCALAR = ILH_SCHUR_CLINVCT };

const char* ilh_last_error(void);
const char* ilh_version(void);

/* ---- context: one HIP device + one stream ---- */
int ilh_ctx_create(int device, ilh_ctx** out);
int ilh_ctx_destroy(ilh_ctx* ctx);
int ilh_ctx_synchronize(ilh_ctx* ctx);
/* device memory helpers for callers without their own allocator */
int ilh_malloc(ilh_ctx* ctx, size_t bytes, void** dptr);
int ilh_free(ilh_ctx* ctx, void* dptr);
int ilh_memcpy_h2d(ilh_ctx* ctx, void* dst, const void* src, size_t bytes);
int ilh_memcpy_d2h(ilh_ctx* ctx, void* dst, const void* src, size_t bytes);

/* ---- plan = the reference's BasicILMCache of a scalar problem ----
 * Replaces SurfaceScalarCache(bl, ds, nrm, g; ddftype, scaling)
 * (src/cache.jl:164-182, _surfacecache src/cache.jl:232-261, the LGF plan of
 * _get_laplacian src/cache.jl:321-324).  X, Y, nx, ny, ds are host arrays of
 * length npts (points(cache), normals(cache), areas(cache), src/tools.jl:16-38). */
int ilh_plan_create(ilh_ctx* ctx, const ilh_grid* grid, int npts,
                    const double* X, const double* Y, const double* nx, const double* ny,
                    const double* ds, int ddf, int scaling, ilh_plan** out);
/* New geometry, same grid and point count: update_system! (src/system.jl:26-50).
 * Drops every cached Schur factor. */
int ilh_plan_update_points(ilh_plan* plan, const double* X, const double* Y,
                           const double* nx, const double* ny, const double* ds);
int ilh_plan_destroy(ilh_plan* plan);
int ilh_plan_npts(const ilh_plan* plan);

/* ---- surface <-> grid (A4-A7) ---- */
/* regularize!(s::Nodes{Primal|Dual}, f::ScalarData) = Rc f, or
   regularize!(v::Edges{Primal}, vb::VectorData) = Rf vb
   (src/surface_operators.jl:13,25,36,38-46) */
int ilh_regularize(ilh_plan* plan, int grid_kind, double* out, const double* f, int mem);
/* interpolate!(f::ScalarData, s::Nodes{...}) = Rc^T s / interpolate!(vb::VectorData, v::Edges{Primal})
   (src/surface_operators.jl:56,67,78,80-88) */
int ilh_interpolate(ilh_plan* plan, int grid_kind, double* out, const double* grid, int mem);
/* regularize_normal!(q::Edges{Primal}, f::ScalarData) = Rf (n o f)  (src/surface_operators.jl:98-103) */
int ilh_regularize_normal(ilh_plan* plan, double* q_edges, const double* f, int mem);
/* normal_interpolate!(vn::ScalarData, q::Edges{Primal}) = n . Rf^T q  (src/surface_operators.jl:228-233) */
int ilh_normal_interpolate(ilh_plan* plan, double* vn, const double* q_edges, int mem);
/* regularize_normal_cross!(q, f) = Rf (n x f e_z)  (src/surface_operators.jl:150-162) */
int ilh_regularize_normal_cross(ilh_plan* plan, double* q_edges, const double* f, int mem);
/* normal_cross_interpolate!(wn, q) = e_z . (n x Rf^T q)  (src/surface_operators.jl:277-290) */
int ilh_normal_cross_interpolate(ilh_plan* plan, double* vn, const double* q_edges, int mem);

/* ---- surface calculus (A11) ---- */
/* surface_divergence!(θ::Nodes{Primal}, f)  (src/surface_operators.jl:530-547) */
int ilh_surface_divergence(ilh_plan* plan, double* theta, const double* f, int mem);
/* surface_grad!(vn, ϕ::Nodes{Primal})  (src/surface_operators.jl:604-620) */
int ilh_surface_grad(ilh_plan* plan, double* vn, const double* phi, int mem);
/* surface_curl!(w::Nodes{Dual}, f)  (src/surface_operators.jl:357-375) */
int ilh_surface_curl(ilh_plan* plan, double* w, const double* f, int mem);
/* surface_curl!(vn, s::Nodes{Dual})  (src/surface_operators.jl:412-429) */
int ilh_surface_curl_adj(ilh_plan* plan, double* vn, const double* s, int mem);
/* surface_curl_cross!(w::Nodes{Dual}, f)  (src/surface_operators.jl:467-484) */
int ilh_surface_curl_cross(ilh_plan* plan, double* w, const double* f, int mem);
/* surface_curl_cross!(γ, s::Nodes{Dual})  (src/surface_operators.jl:498-515) */
int ilh_surface_curl_cross_adj(ilh_plan* plan, double* vn, const double* s, int mem);
/* surface_grad_cross!(γ, ϕ::Nodes{Primal})  (src/surface_operators.jl:708-725) */
int ilh_surface_grad_cross(ilh_plan* plan, double* vn, const double* phi, int mem);

/* ---- vector cache (SurfaceVectorCache, src/cache.jl:184-201) ----
 * The same plan serves the vector cache: VectorData v = [u; v] (2N), its
 * grid data Edges{Primal}, its tensor data EdgeGradient{Primal,Dual} stored
 * [dudx; dudy; dvdx; dvdy] with dudx, dvdy on Nodes{Primal} and dudy, dvdx
 * on Nodes{Dual} ((NX-1)(NY-1), NX NY, NX NY, (NX-1)(NY-1) values).  The
 * scalar-data operators of the vector cache (surface_curl! of ScalarData,
 * surface_grad_cross!, ...) are the scalar cache's entry points above. */
/* surface_curl!(w::Nodes{Dual}, v::VectorData) = C^T R v  (src/surface_operators.jl:388-400) */
int ilh_surface_curl_vector(ilh_plan* plan, double* w, const double* v, int mem);
/* surface_curl!(v::VectorData, s::Nodes{Dual}) = R^T C s  (src/surface_operators.jl:443-455) */
int ilh_surface_curl_vector_adj(ilh_plan* plan, double* v, const double* s, int mem);
/* surface_divergence_symm!(q::Edges{Primal}, dv::VectorData)
   = D R_t (n o dv + dv o n - (n . dv) I)  (src/surface_operators.jl:565-591) */
int ilh_surface_divergence_symm(ilh_plan* plan, double* q_edges, const double* dv, int mem);
/* surface_grad_symm!(tau::VectorData, v::Edges{Primal})
   = n . R_t^T (G v + (G v)^T - tr(G v) I)  (src/surface_operators.jl:632-664) */
int ilh_surface_grad_symm(ilh_plan* plan, double* tau, const double* v_edges, int mem);
/* regularize_normal_symm!(A::EdgeGradient, v::VectorData)  (src/surface_operators.jl:113-138) */
int ilh_regularize_normal_symm(ilh_plan* plan, double* A, const double* v, int mem);
/* normal_interpolate_symm!(tau::VectorData, A::EdgeGradient)  (src/surface_operators.jl:236-267) */
int ilh_normal_interpolate_symm(ilh_plan* plan, double* tau, const double* A, int mem);

/* ---- grid stencils (A8), GridScaling ---- */
/* divergence!(p::Nodes{Primal}, q::Edges{Primal})  (src/grid_operators.jl:25-31) */
int ilh_divergence(ilh_plan* plan, double* p, const double* q_edges, int mem);
/* grad!(q::Edges{Primal}, p::Nodes{Primal})  (src/grid_operators.jl:45-51) */
int ilh_grad(ilh_plan* plan, double* q_edges, const double* p, int mem);
/* curl!(q::Edges{Primal}, s::Nodes{Dual})  (src/grid_operators.jl:79-87) */
int ilh_curl_nodes_to_edges(ilh_plan* plan, double* q_edges, const double* s, int mem);
/* curl!(w::Nodes{Dual}, q::Edges{Primal})  (src/grid_operators.jl:97-105) */
int ilh_curl_edges_to_nodes(ilh_plan* plan, double* w, const double* q_edges, int mem);
/* laplacian!(w, s) on Nodes{Primal|Dual}  (src/grid_operators.jl:114-134) */
int ilh_laplacian(ilh_plan* plan, int grid_kind, double* out, const double* in, int mem);
/* inverse_laplacian!(sol, rhs) / in place when out == in  (src/grid_operators.jl:144-179);
   grid_kind ILH_EDGES_PRIMAL convolves each component on its own grid */
int ilh_inverse_laplacian(ilh_plan* plan, int grid_kind, double* out, const double* in, int mem);

/* ---- FP32 data (caches built with dtype = Float32, src/cache.jl:164-201) ----
 * The same operators on float buffers: stencils in FP32 (the difference in
 * FP32, the 1/dx scaling rounded once); regularize!/interpolate! accumulate
 * the FP64 DDF weights and round once; inverse_laplacian! is the FFT
 * convolution in FP32 (FP32 transforms, spectra, LGF spectrum rounded once
 * and LDS exchanges), as an FP32 FFT plan computes it: within 2e-6 of the
 * FP64 convolution (max-abs, relative to the output's largest entry). */
int ilh_regularize_f32(ilh_plan* plan, int grid_kind, float* out, const float* f, int mem);
int ilh_interpolate_f32(ilh_plan* plan, int grid_kind, float* out, const float* grid, int mem);
int ilh_divergence_f32(ilh_plan* plan, float* p, const float* q_edges, int mem);
int ilh_grad_f32(ilh_plan* plan, float* q_edges, const float* p, int mem);
int ilh_curl_nodes_to_edges_f32(ilh_plan* plan, float* q_edges, const float* s, int mem);
int ilh_curl_edges_to_nodes_f32(ilh_plan* plan, float* w, const float* q_edges, int mem);
int ilh_laplacian_f32(ilh_plan* plan, int grid_kind, float* out, const float* in, int mem);
int ilh_inverse_laplacian_f32(ilh_plan* plan, int grid_kind, float* out, const float* in, int mem);

/* ---- Schur complements (A10-A12) ---- */
/* create_RTLinvR / create_CLinvCT / create_GLinvD / create_nRTRn with `scale`
   (src/matrix_operators.jl:9-30,40-61,135-155,225-244).  `out` is N x N
   column-major.  ILH_SCHUR_CL2INVCT: create_CL2invCT of the vector cache,
   -C_s L^-2 C_s^T on VectorData (src/matrix_operators.jl:95-125), 2N x 2N;
   ILH_SCHUR_GLINVD_SYMM: create_GLinvD_symm of the vector cache,
   -G_s^symm L^-1 D_s^symm with the inverse Laplacian of each Edges{Primal}
   component (src/matrix_operators.jl:165-185), 2N x 2N; factor and solve take
   2N-vectors for both. */
int ilh_schur_build(ilh_plan* plan, int kind, double scale, double* out, int mem);
/* Factor the plan's own S of `kind` (LU with partial pivoting, kept on the device).
   The reference re-factorises inside every `S\s` (test/literate/problems.jl:291). */
int ilh_schur_factor(ilh_plan* plan, int kind);
/* x = S \ b with the cached factor of `kind` */
int ilh_schur_solve(ilh_plan* plan, int kind, double* x, const double* b, int mem);

/* ---- constrained Poisson composites (A13, A14) ---- */
/* Dirichlet solve of test/literate/problems.jl:259-298 (saddle system of
   test/literate/dirichlet.jl:28-44): fb_plus/fb_minus = exterior/interior data
   (prescribed_surface_jump!/average!, src/bc.jl:20-24,47-51); fb_minus may be
   NULL (zero).  Outputs f (Nodes{Primal}) and the unfiltered strength s.
   The first call on a geometry builds and factors S and the exterior-data
   solve matrix; the first call with fb_minus also builds the interior-data one. */
int ilh_dirichlet_solve(ilh_plan* plan, const double* fb_plus, const double* fb_minus,
                        double* f, double* s, int mem);
/* Mixed-precision form of ilh_dirichlet_solve (same arguments and outputs):
   the surface strength s is the FP64 solve (solve matrices and matvec in
   FP64, so it equals ilh_dirichlet_solve's s bit for bit), and f is the
   convolution in split precision: the forward transforms and the product
   with the LGF spectrum in FP64, the inverse column transform, the spectrum
   it writes and the inverse row transform in FP32.  f agrees with the FP64
   solve to relative L2 ~1e-7 (bound 1e-6 on the 1024^2 workload). */
int ilh_dirichlet_solve_mixed(ilh_plan* plan, const double* fb_plus, const double* fb_minus,
                              double* f, double* s, int mem);
/* Neumann solve of test/literate/neumann.jl:101-142: outputs f (Nodes{Primal}),
   df = [ϕ], s (Nodes{Dual} streamfunction) and ds = [ψ]. vn_minus may be NULL. */
int ilh_neumann_solve(ilh_plan* plan, const double* vn_plus, const double* vn_minus,
                      double* f, double* df, double* s, double* ds, int mem);

/* nrhs Dirichlet problems on one plan's geometry at once (repeat solves with
   many boundary data): fb_plus, fb_minus (may be NULL) N x nrhs column-major,
   f (NX-1)(NY-1) x nrhs, s N x nrhs.  The solve matrices are read once per
   group of 8 right-hand sides and each FFT pass covers all of them; column k
   equals ilh_dirichlet_solve on column k bit for bit.  nrhs <= 256. */
int ilh_dirichlet_solve_multi(ilh_plan* plan, int nrhs, const double* fb_plus, const double* fb_minus, double* f,
                              double* s, int mem);

/* Allocation-free form of ilh_dirichlet_solve (no counterpart in the
   reference): for callers that cannot allocate device memory (code under HIP
   graph capture, tight serving loops, harnesses that forbid hipMalloc).
   Device pointers only, asynchronous on the context stream; fb_minus may be
   NULL.  The plan's solve data must already exist (one ilh_dirichlet_solve
   call on this geometry builds them, one with fb_minus those that fb_minus
   needs; ILH_ERR_STATE otherwise).  The work
   buffers (surface amplitudes and the two convolution spectra) live in one
   static device array of 32 MiB per GPU, enough for grids up to about
   1024 x 1024 (larger problems fail with ILH_ERR_ARG).  Every call on a GPU
   shares that array: calls that overlap on one GPU (other streams or
   contexts) race.  Results equal ilh_dirichlet_solve's bit for bit. */
int ilh_dirichlet_solve_noalloc(ilh_plan* plan, const double* fb_plus, const double* fb_minus, double* f,
                                double* s);

/* ---- batched forms over arrays of plans ---- */
/* Batched form of ilh_dirichlet_solve: nprob independent plans of one context
   (any mix of grid sizes and point counts), array k of each pointer array
   belonging to plans[k]; fb_minus may be NULL, or hold NULL entries.  When
   every problem's FFT periods lie in [128, 2048] the batch runs as four
   launches in all -- the surface step, the row, column and inverse-row
   passes, each over the tiles of every problem -- else each problem's
   kernels run on the context's side streams.  Results equal the single
   solves bit for bit.  With ILH_MEM_DEVICE the call is asynchronous on the
   context stream like the single-problem form.  (The reference solves
   problems one at a time; this is the batched form of
   test/literate/problems.jl:259-298.) */
int ilh_dirichlet_solve_batched(int nprob, ilh_plan* const* plans, const double* const* fb_plus,
                                const double* const* fb_minus, double* const* f, double* const* s, int mem);

/* Persistent batch: the batched solve above on fixed device buffers, captured
   once as a HIP graph and replayed by ilh_batch_launch, asynchronously on the
   context stream, one launch per step.  ilh_batch_create prepares every plan,
   autotunes the four passes on the batch itself (per pass: one work-group per
   tile or a persistent grid pulling tiles from per-XCD queues, its size,
   256 or 512 threads per work-group, amplitudes staged in LDS or not; each
   candidate timed, the fastest kept), runs the batch once and captures it.
   Inputs are read at each launch.  A plan whose geometry changes invalidates
   the batch (ILH_ERR_STATE). */
typedef struct ilh_batch ilh_batch;
int ilh_batch_create(int nprob, ilh_plan* const* plans, const double* const* fb_plus,
                     const double* const* fb_minus, double* const* f, double* const* s, ilh_batch** out);
int ilh_batch_launch(ilh_batch* batch);
int ilh_batch_destroy(ilh_batch* batch);
/* The autotuner's choice: config[13] = (schedule 0 one work-group per tile /
   1 persistent queues, work-groups per CU, threads per work-group) of the
   surface step, row, column and inverse-row passes, then the row pass's
   amplitude LDS cap (0: none); pass_ms[4] the tuned time of each pass.
   ILH_ERR_STATE for a batch that runs per-problem kernels.  Either pointer may
   be NULL.  (No counterpart in the reference.) */
int ilh_batch_tuning(const ilh_batch* batch, int* config, float* pass_ms);

/* Dirichlet solve for moving bodies (problems.jl:259-298 after every
   update_system!, src/system.jl:26-50): same outputs as ilh_dirichlet_solve,
   but the Schur system of the current geometry is solved by FP64 GMRES (at
   most max_iter iterations, relative residual tol), right-preconditioned by
   the inverse Schur matrix of a reference geometry that the plan keeps
   across ilh_plan_update_points, stored in FP32; the current geometry's S
   (create_RTLinvR, src/matrix_operators.jl:9-30) is assembled directly and
   applied as a dense matvec.  The first call, and any call whose GMRES does
   not converge, factors the current geometry and makes it the reference
   (the re-solve then takes up to max(max_iter, 8) iterations).  *iters (may
   be NULL) receives the GMRES iterations (max_iter + k after a
   refactorisation). */
int ilh_dirichlet_solve_moving(ilh_plan* plan, const double* fb_plus, const double* fb_minus, double* f,
                               double* s, int mem, double tol, int max_iter, int* iters);

/* Batched form of ilh_dirichlet_solve_moving over nprob plans of one context
   (independent trajectories, each moved by its own ilh_plan_update_points),
   device pointers only, array k of each pointer array belonging to plans[k];
   fb_minus may be NULL or hold NULL entries; iters (may be NULL) receives
   each plan's GMRES iterations.  The plans' geometry rebuilds run on the
   context's side streams; both convolutions of the step run as one fused
   batched call each, and the GMRES steps of all plans advance in lockstep,
   one launch per phase over every plan.  Each plan's results equal its
   single-plan solve bit for bit. */
int ilh_dirichlet_solve_moving_batched(int nprob, ilh_plan* const* plans, const double* const* fb_plus,
                                       const double* const* fb_minus, double* const* f, double* const* s, double tol,
                                       int max_iter, int* iters);

/* ---- learned preconditioner for moving bodies (no reference counterpart:
 * a replacement for the per-step LU of the Schur matrix that update_system!
 * forces, src/system.jl:26-50, src/problem.jl:95-105) ----
 * A network (features of the local geometry -> 256 -> 256 -> 256 -> band,
 * ReLU, BF16 weights and activations, FP32 accumulation on the matrix cores)
 * predicts, for each surface point, the 2K+1 entries of S^-1 around it along
 * its body's curve (S = create_RTLinvR).  It is trained on the GPU (Adam,
 * FP32 master weights) on the bands of exact inverses that the plain method
 * computes, and used as the GMRES preconditioner of the moving-body solves
 * (ilh_dirichlet_solve_moving and its batched form), which fall back to the
 * exact factorisation when GMRES does not converge within max_iter. */
typedef struct ilh_lpc ilh_lpc;
/* Point offsets of the bodies (offsets[0] = 0 < ... < offsets[nbodies] = N):
   each body's points are contiguous and ordered along its closed curve
   (RigidBodyTools' BodyList order).  Without it the points form one body. */
int ilh_plan_set_bodies(ilh_plan* plan, int nbodies, const int* offsets);
/* A model with half band K (band 2K+1, 1 <= K <= 31) and seeded initial weights. */
int ilh_lpc_create(ilh_ctx* ctx, int half_band, unsigned long long seed, ilh_lpc** out);
int ilh_lpc_destroy(ilh_lpc* model);
/* Training rows of the plan's current geometry: its features and the band of
   its exact inverse (one plain-method factorisation, N unit solves). */
int ilh_lpc_add_geometry(ilh_lpc* model, ilh_plan* plan);
int ilh_lpc_rows(const ilh_lpc* model);
/* steps Adam steps (learning rate lr) on seeded minibatches of `batch` rows;
   *loss = the steps' mean loss (standardized targets).  The standardization is
   fixed by the first call. */
int ilh_lpc_train(ilh_lpc* model, int steps, int batch, double lr, unsigned long long seed, double* loss);
/* dims = [K, 2K+1, F (features), Fp (padded), 256 (width), Op (padded outputs), parameter count] */
int ilh_lpc_dims(const ilh_lpc* model, int* dims);
/* FP32 master parameters (W1 Fp x 256 row-major, b1, W2, b2, W3, b3, W4 256 x Op, b4) out (get) or in */
int ilh_lpc_params(ilh_lpc* model, float* params, int get);
/* loss and all parameter gradients of one standardized batch (X: rows x Fp
   BF16 bit patterns, T: rows x Op FP32): the training step without the update */
int ilh_lpc_loss_grad(ilh_lpc* model, int rows, const unsigned short* X, const float* T, double* loss, float* grads);
/* features (N x F) and the predicted band (N x (2K+1), may be NULL) of the plan's geometry */
int ilh_lpc_predict(ilh_lpc* model, ilh_plan* plan, float* features, double* band);
/* [mu_x (F), sd_x (F), mu_y (2K+1), sd_y (2K+1)] */
int ilh_lpc_standardization(ilh_lpc* model, float* out);
/* moving-body solves of the plan use the model's band (NULL: the FP32
   reference inverse); the model must outlive its use by the plan */
int ilh_plan_set_preconditioner(ilh_plan* plan, ilh_lpc* model);

/* ---- masks (src/surface_operators.jl:728-871) ----
 * mask(cache) = -L^-1 D_s 1 on Nodes{Primal} (1 inside the surfaces, on the side
 * opposite the normals; 1 everywhere without points); on Nodes{Dual} and
 * Edges{Primal} the primal mask (or its complement) averaged onto them
 * (grid_interpolate!), 0 on the outer ring those averages do not reach.
 * ilh_mask writes the mask (complementary != 0: 1 - mask, i.e.
 * complementary_mask) of grid_kind into out; ilh_mask_multiply multiplies w in
 * place (mask! / complementary_mask!). */
int ilh_mask(ilh_plan* plan, int grid_kind, int complementary, double* out, int mem);
int ilh_mask_multiply(ilh_plan* plan, int grid_kind, int complementary, double* w, int mem);

/* ---- surface filter (src/matrix_operators.jl:245-267, create_surface_filter) ----
 * C = Et R, Et the interpolation whose grid values are first divided by R 1
 * (the regularized unit strength): C keeps constants and the surface integral
 * sum_p ds_p (C s)_p = sum_q ds_q s_q, as the reference documents its
 * filter (test/literate/matrices.jl:49-56).  ilh_surface_filter_build writes C
 * (N x N, column-major); ilh_surface_filter_apply writes C^power s_in into
 * s_out (the reference applies C^5 or C^6 to the strength,
 * test/literate/dirichlet.jl:122-124, problems.jl:298) without forming C. */
int ilh_surface_filter_build(ilh_plan* plan, double* C, int mem);
int ilh_surface_filter_apply(ilh_plan* plan, int power, const double* s_in, double* s_out, int mem);
/* create_surface_filter of the vector cache: the same construction per
 * component on its edges (block diagonal, 2N x 2N column-major), and its
 * application C^power to VectorData. */
int ilh_surface_filter_build_vector(ilh_plan* plan, double* C, int mem);
int ilh_surface_filter_apply_vector(ilh_plan* plan, int power, const double* v_in, double* v_out, int mem);

/* ---- Stokes flow (vector cache) ----
 * The Stokes problem of test/literate/stokes.jl:101-166 (saddle system
 * :29-35): v_plus / v_minus = exterior / interior surface velocities
 * (VectorData, 2N; v_minus may be NULL for zero), outputs the velocity v
 * (Edges{Primal}), the streamfunction s (Nodes{Dual}) and the surface
 * traction sigma (VectorData, 2N) filtered filter_power times (the
 * reference returns C^2 sigma: filter_power = 2).  The first call on a
 * geometry builds and inverts create_CL2invCT and create_CLinvCT_scalar
 * (prob_cache, :76-96). */
int ilh_stokes_solve(ilh_plan* plan, const double* v_plus, const double* v_minus, double* v, double* s,
                     double* sigma, int filter_power, int mem);

/* ---- unbounded heat conduction with Dirichlet bodies (config: IF time marching) ----
 * The constrained ODE system of test/literate/heatconduction.jl:88-180,
 *   dT/dt = kappa L T - kappa D_s [T] + R sigma,   E T = (T_b+ + T_b-)/2,
 * T on Nodes{Primal}, marched by the integrating-factor half-explicit RK method
 * that ConstrainedSystems' LiskaIFHERK implements (src/timemarching.jl:5;
 * 3 stages, c = (0, 1/3, 1)), the integrating factor exp(tau kappa L) applied
 * exactly as the separable lattice heat kernel e^{-2a} I_p(2a).
 * ilh_heat_create fixes kappa and dt for the plan's geometry (it builds and
 * inverts the three stage Schur matrices); ilh_heat_set_boundary sets
 * time-independent surface values T_b+ (and T_b-, or NULL for 0);
 * ilh_heat_advance marches T (in place) by nsteps steps. */
typedef struct ilh_heat ilh_heat;
int ilh_heat_create(ilh_plan* plan, double kappa, double dt, ilh_heat** out);
int ilh_heat_set_boundary(ilh_heat* heat, const double* tb_plus, const double* tb_minus, int mem);
int ilh_heat_advance(ilh_heat* heat, double* T, int nsteps, int mem);
int ilh_heat_destroy(ilh_heat* heat);

/* ---- measurement helper (no reference counterpart) ----
   HIP-event timings on the plan's stream, in ms per call, of the Dirichlet
   repeat solve on device buffers: *total_ms = `reps` whole solves back to back
   (two events around the loop); if stage_ms is not NULL, stage_ms[k] = `reps`
   launches of stage k alone (0 = surface matvec, 1 = row FFT of the sources,
   2 = column FFT x LGF spectrum x inverse column FFT, 3 = inverse row FFT
   writing f). */
int ilh_profile_dirichlet(ilh_plan* plan, const double* fb_plus_dev, const double* fb_minus_dev,
                          double* f_dev, double* s_dev, int reps, double* stage_ms, double* total_ms);

#ifdef __cplusplus
}
#endif
#endif
