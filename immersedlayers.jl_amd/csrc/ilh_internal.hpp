// ilh_internal.hpp -- shared types of libimmersedhip.so (not part of the ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <atomic>
#include <cstddef>
#include <cstdint>
#include <new>
#include <string>
#include <vector>

#include "../../include/immersedhip.h"

namespace ilh {

// ----------------------------------------------------------------- errors
void set_error(const std::string& msg);
struct Error {
    int code;
    std::string msg;
};
[[noreturn]] void fail(int code, const std::string& msg);
void check_hip(hipError_t e, const char* what);
#define ILH_HIP(call) ::ilh::check_hip((call), #call)

// -------------------------------------------------------------- grid kinds
// Component grids: Nodes{Primal}, Nodes{Dual}, Edges{Primal}.u, Edges{Primal}.v,
// and the four components of EdgeGradient{Primal,Dual} in storage order
// [dudx; dudy; dvdx; dvdy]: the diagonal ones on Nodes{Primal}, the
// off-diagonal ones on Nodes{Dual} (where the MAC differences of the edge
// components land).
enum GK { GK_PRIMAL = 0, GK_DUAL = 1, GK_EDGE_U = 2, GK_EDGE_V = 3, GK_T_XX = 4, GK_T_XY = 5, GK_T_YX = 6,
          GK_T_YY = 7, GK_COUNT = 8 };

struct Grid {
    int NX, NY, I0x, I0y;
    double dx;
    static constexpr bool primal_shaped(int gk) { return gk == GK_PRIMAL || gk == GK_T_XX || gk == GK_T_YY; }
    int nx(int gk) const { return (primal_shaped(gk) || gk == GK_EDGE_V) ? NX - 1 : NX; }
    int ny(int gk) const { return (primal_shaped(gk) || gk == GK_EDGE_U) ? NY - 1 : NY; }
    size_t n(int gk) const { return (size_t)nx(gk) * ny(gk); }
    size_t n_edges() const { return n(GK_EDGE_U) + n(GK_EDGE_V); }
    size_t n_tensor() const { return 2 * (n(GK_PRIMAL) + n(GK_DUAL)); }
    // offset of a component grid inside its field (Edges or EdgeGradient)
    size_t offset(int gk) const
    {
        switch (gk) {
        case GK_EDGE_V: return n(GK_EDGE_U);
        case GK_T_XY: return n(GK_PRIMAL);
        case GK_T_YX: return n(GK_PRIMAL) + n(GK_DUAL);
        case GK_T_YY: return n(GK_PRIMAL) + 2 * n(GK_DUAL);
        default: return 0;
        }
    }
};

// -------------------------------------------------------------------- terms
// One separable piece of a point's stencil on one component grid:
//   contribution(i, j) = coef * wx[i - bx] * wy[j - by],  0 <= i-bx, j-by < 5.
// Surface->grid operators scatter amp[ai] * contribution; grid->surface
// operators gather sum_ij contribution(i, j) * field(i, j) into out[ai].
struct Term {
    int ai;        // amplitude / output index
    int gk;        // component grid
    int bx, by;    // window base
    double coef;
    double wx[5];
    double wy[5];
};

// Surface operators, each a list of terms per point (T terms per point,
// stored point-major: term index = p * T + t).
enum OpKind {
    OP_R_PRIMAL = 0,   // regularize!  ScalarData -> Nodes{Primal}
    OP_R_DUAL,         // regularize!  ScalarData -> Nodes{Dual}
    OP_R_EDGES,        // regularize!  VectorData -> Edges{Primal}
    OP_E_PRIMAL,       // interpolate! Nodes{Primal} -> ScalarData
    OP_E_DUAL,         // interpolate! Nodes{Dual} -> ScalarData
    OP_E_EDGES,        // interpolate! Edges{Primal} -> VectorData
    OP_RN,             // regularize_normal!
    OP_NI,             // normal_interpolate!
    OP_RNX,            // regularize_normal_cross!
    OP_NXI,            // normal_cross_interpolate!
    OP_DS,             // surface_divergence! ScalarData -> Nodes{Primal}
    OP_GS,             // surface_grad!       Nodes{Primal} -> ScalarData
    OP_CST,            // surface_curl!       ScalarData -> Nodes{Dual}
    OP_CS,             // surface_curl!       Nodes{Dual} -> ScalarData
    OP_CXT,            // surface_curl_cross! ScalarData -> Nodes{Dual}
    OP_CXS,            // surface_curl_cross! Nodes{Dual} -> ScalarData
    OP_GXS,            // surface_grad_cross! Nodes{Primal} -> ScalarData
    // vector cache (VectorData on Edges{Primal}, tensors on EdgeGradient)
    OP_CVT,            // surface_curl!       VectorData -> Nodes{Dual}
    OP_CVS,            // surface_curl!       Nodes{Dual} -> VectorData
    OP_DSYM,           // surface_divergence_symm! VectorData -> Edges{Primal}
    OP_GSYM,           // surface_grad_symm!       Edges{Primal} -> VectorData
    OP_RNSYM,          // regularize_normal_symm!  VectorData -> EdgeGradient
    OP_NISYM,          // normal_interpolate_symm! EdgeGradient -> VectorData
    OP_COUNT
};
int op_terms_per_point(int op);
bool op_is_scatter(int op);

// Row-binned index of scatter terms on one component grid:
// entries[row_ptr[j] .. row_ptr[j+1]) are term indices whose window touches
// row j, sorted by (bx, term index).
// Each entry also carries its record: amplitude index and the five x-weights
// with the term's coefficient and its row weight folded in (erw[5k + e]).
struct RowBins {
    int gk = 0;
    int ny = 0;
    int* row_ptr = nullptr;     // device, ny + 1
    int* entries = nullptr;     // device: term indices, then their bx
    int* eai = nullptr;         // device: amplitude index per entry
    double* erw = nullptr;      // device: 5 weights per entry
    // nonzero cells of each row, ascending x: cells[c] = {x, c0, c1, 0}; the
    // cell's value is sum_{c0 <= e < c1} cw[e] * amp[cai[e]] in that order (the
    // row's entries k with bx in [x-4, x] and a nonzero weight erw[5k + x - bx])
    int* cell_ptr = nullptr;    // device, ny + 1
    int4* cells = nullptr;      // device, one per cell
    int* cai = nullptr;         // device: amplitude index per contribution
    double* cw = nullptr;       // device: weight per contribution
    // build scratch: row counts; per-cell records, counts and offsets
    int* cnt = nullptr;
    int* cells3 = nullptr;
    int* ccnt = nullptr;
    int* coff = nullptr;
    // allocated capacities (buffers are kept across geometry updates and only
    // grown): rows, entries, cells
    size_t cap_rows = 0, cap_e = 0, cap_c = 0;
};

// ---------------------------------------------------------- device buffers
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    void ensure(size_t b);
    void release();
    template <class T> T* as() const { return (T*)p; }
    ~DevBuf() { release(); }
};

// Zero-padded FFT convolution with the LGF on one grid kind.
struct ConvPlan {
    int gk = 0;
    int nx = 0, ny = 0;       // field dims
    int Px = 0, Py = 0;       // padded periods (powers of two, >= 2n-2)
    int H = 0;                // Px/2: half-spectrum columns (column 0 packs DC and Nyquist)
    DevBuf spec;              // real spectrum * dx^2/(Px*Py): [kx][ky], kx <= H, ky <= Py/2 (even in ky)
    DevBuf spec32;            // the same rounded to FP32 (the FP32 passes), built on first use
    bool spec32_ok = false;
    DevBuf T1, T2;            // intermediate spectra, rows x H complex, 8-row tiles
    DevBuf T1m, T2m;          // the same for batches of right-hand sides (copies back to back)
};

struct FusedBatch;            // fused batched solves (ilh_batch.hip)

}  // namespace ilh

struct ilh_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int num_cus = 0;
    static constexpr int POOL = 4;            // side streams of batched calls (one per hardware queue)
    hipStream_t pool[POOL] = {};
    hipEvent_t ev[POOL + 1] = {};
    hipEvent_t tev[2] = {};                   // timing events of the profiling entry points (made on first use)
    ilh::FusedBatch* fused = nullptr;         // the last eager fused batch (reused while its problems are unchanged)
    ilh::FusedBatch* fused_mv[2] = {};        // the batched moving solve's two convolutions (right-hand side, final)
    // batched moving-body GMRES: device descriptors and tiles, pinned flags
    void* mv_dev = nullptr;
    size_t mv_cap = 0;
    int* mv_host = nullptr;
    size_t mv_host_cap = 0;
    // batched source builds: device job table and its host copy
    void* geo_dev = nullptr;
    size_t geo_cap = 0;
    std::vector<char> geo_host;
    // owners: the caller's handle, every plan and batch of the context; the
    // context is torn down when the last one lets go (any destruction order)
    std::atomic<int> refs{1};
};

struct SolveCache;   // Schur factor and solve matrices of one composite
struct ilh_lpc;      // learned preconditioner model (ilh_learned.hip)

struct MovingCache;
struct StokesCache;
struct ilh_plan {
    ilh_ctx* ctx = nullptr;
    ilh::Grid g{};
    int N = 0;
    int ddf = 0;
    int min_margin = 0;       // cells between the point supports and the grid edge
    ilh::DevBuf X, Y, nx, ny, ds;            // device geometry
    ilh::DevBuf terms[ilh::OP_COUNT];        // per-op term arrays (built lazily)
    bool terms_ok[ilh::OP_COUNT] = {};
    ilh::DevBuf wins;                         // DDF windows of each point (primal x, y, dual x, y)
    bool wins_ok = false;
    ilh::RowBins bins[ilh::OP_COUNT][4];     // row bins of scatter ops per component grid
    bool bins_ok[ilh::OP_COUNT] = {};
    ilh::ConvPlan conv[4];                    // primal, dual, u edges, v edges
    bool conv_ok[4] = {};
    ilh::DevBuf lgf;                          // LGF table tmx x tmy (covers both kinds)
    int lgf_mx = 0, lgf_my = 0;
    ilh::DevBuf io_a, io_b, io_c, io_d;       // staging for host-memory calls
    ilh::DevBuf scratch_a, scratch_b;         // grid scratch
    ilh::DevBuf lu_tinv, lu_perm, lu_tmp, lu_pmap;   // LU scratch (block inverses, row permutations)
    SolveCache* dir = nullptr;                // Dirichlet
    SolveCache* neu = nullptr;                // Neumann
    SolveCache* fac[6] = {};                  // ilh_schur_factor caches per kind
    StokesCache* sto = nullptr;               // Stokes flow (vector cache) solve matrices
    MovingCache* mov = nullptr;               // moving-body Dirichlet solves (kept across update_points)
    unsigned gen = 0;                         // incremented whenever the geometry caches are dropped
    ilh::DevBuf mask;                         // mask(cache) on Nodes{Primal} (built lazily)
    unsigned mask_gen = ~0u;                  // generation the mask belongs to
    ilh::DevBuf filt_r1;                      // R 1 on Nodes{Primal} for the surface filter
    unsigned filt_gen = ~0u;
    ilh::DevBuf filt_r1e;                     // R 1 on Edges{Primal} for the vector surface filter
    unsigned filt_gen_e = ~0u;
    std::atomic<int> refs{1};                 // the caller's handle, batches and heat integrators using it
    // bodies: point offsets of each body's contiguous points (empty: one body)
    std::vector<int> body_off;
    unsigned body_gen = 0;
    ilh::DevBuf nbr;                          // N x (2K+1) neighbours along each body's curve (learned preconditioner)
    int nbr_K = -1;
    unsigned nbr_gen = ~0u;
    ilh_lpc* lpc = nullptr;                   // learned preconditioner of the moving-body solves (not owned)
};

namespace ilh {
// shared ownership of contexts and plans (ilh_api.hip)
void ctx_retain(ilh_ctx* c);
void ctx_release(ilh_ctx* c);
void plan_retain(ilh_plan* pl);
void plan_release(ilh_plan* pl);

// ---------------------------------------------------------------- launchers
// terms / binning (ilh_terms.hip)
void build_terms(ilh_plan* pl, int op);
const Term* get_terms(ilh_plan* pl, int op);
void build_bins(ilh_plan* pl, const Term* terms, int nterms, int gk, RowBins& bins);
void free_bins(RowBins& b);
// batched over nb operands (grid.y): operand k reads amp + k amp_stride /
// field + k field_stride and writes out + k out_stride
template <class T>
void scatter(ilh_plan* pl, const RowBins& bins, const T* amp, size_t amp_stride, int nb, T* out, size_t out_stride,
             const T* base = nullptr, double c = 1.0);
// field: the operator's whole input (Nodes, Edges or EdgeGradient); each term
// reads its component grid at Grid::offset(gk)
template <class T>
void gather(ilh_plan* pl, const Term* terms, int nterms_per_point, const T* field, size_t field_stride, int nb,
            T* out, size_t out_stride);

// stencils (ilh_stencil.hip)
void k_divergence(ilh_plan* pl, double* p, const double* q);
void k_grad(ilh_plan* pl, double* q, const double* p);
void k_curl_n2e(ilh_plan* pl, double* q, const double* s);
void k_curl_e2n(ilh_plan* pl, double* w, const double* q);
void k_laplacian(ilh_plan* pl, int gk, double* w, const double* s);
// FP32 grid data (a Float32 cache): the same stencils
void k_divergence(ilh_plan* pl, float* p, const float* q);
void k_grad(ilh_plan* pl, float* q, const float* p);
void k_curl_n2e(ilh_plan* pl, float* q, const float* s);
void k_curl_e2n(ilh_plan* pl, float* w, const float* q);
void k_laplacian(ilh_plan* pl, int gk, float* w, const float* s);
void fill_zero(hipStream_t st, double* a, size_t n);
void fill_zero(hipStream_t st, float* a, size_t n);
void axpby(hipStream_t st, double* y, const double* x, double a, double b, size_t n);  // y = a x + b y

// FFT convolution (ilh_fft.hip)
void ensure_lgf(ilh_plan* pl);
ConvPlan& get_conv(ilh_plan* pl, int gk);
// field -> field:  out = dx^2 (G * in)
void conv_dense(ilh_plan* pl, int gk, const double* in, double* out);
void conv_dense(ilh_plan* pl, int gk, const float* in, float* out);   // FP32 fields, FP32 transforms
// nb fields at once: field k at in + k n(gk), out + k n(gk) (in == out allowed)
void conv_dense_multi(ilh_plan* pl, int gk, const double* in, int nb, double* out);
// point sources -> field: out = dx^2 (G * sum_terms amp*term), rows listed in pairs
struct SourceSet {
    const Term* terms = nullptr;
    int amp_len = 0;           // amplitude slots the terms index
    RowBins bins;
    int4* pinfo = nullptr;     // device, per nonzero row pair p: {j0, first cell, first cell of row j0+1, end}
    int* slot_of_row = nullptr;  // device: field row -> row of the compacted spectrum, or -1
    int npairs = 0;            // upper bound of the live row pairs ((ny + 1) / 2)
    int* npairs_dev = nullptr; // device: live row pairs (after slot_of_row's rows)
    size_t cap_pairs = 0, cap_rows = 0;   // allocated sizes of pinfo and slot_of_row
    // The first INL_K contributions of every cell, copied by slot (built by
    // pin_source_pairs; null otherwise): inl_ai[u inl_stride + c] and
    // inl_w[u inl_stride + c] are bins.cai / bins.cw[c0 + u] of cell c (0 past
    // the cell's count).  The row pass loads them with the cell itself instead
    // of after it.
    int* inl_ai = nullptr;
    double* inl_w = nullptr;
    int inl_stride = 0;
    size_t inl_cap = 0;        // cells the inline copies have room for
};
constexpr int INL_K = 8;
void free_sources(SourceSet& s);   // releases the bins and pair tables
void build_source_pairs(ilh_plan* pl, SourceSet& s);
// the merged sources [op0; op1] of many plans (and the terms of extra_op, or
// -1), every kernel one launch over all plans, on the context stream
struct SourceJob {
    ilh_plan* pl;
    DevBuf* tbuf;
    SourceSet* src;
};
void merge_sources_batched(ilh_ctx* ctx, const std::vector<SourceJob>& jobs, int op0, int op1, int extra_op, int gk);
void pin_source_pairs(ilh_plan* pl, SourceSet& s);
void conv_sources(ilh_plan* pl, int gk, const SourceSet& src, const double* amp, double* out,
                  hipEvent_t* ev = nullptr);
void conv_stage(ilh_plan* pl, int gk, const SourceSet& src, const double* amp, double* out, int stage);
// the same convolution in split precision: forward transforms and the spectrum
// product in FP64, the inverse transforms in FP32 (FP64 output)
void conv_sources_f32(ilh_plan* pl, int gk, const SourceSet& src, const double* amp, double* out);
// conv_sources into caller-supplied spectra buffers, allocating nothing
int conv_ext_rows1(const SourceSet& src);
int conv_ext_rows2(const ConvPlan& C);
void conv_sources_ext(ilh_plan* pl, int gk, const SourceSet& src, const double* amp, double* out, double2* T1,
                      double2* T2);
void conv_sources_multi(ilh_plan* pl, int gk, const SourceSet& src, const double* amp, size_t amp_stride, int nb,
                        double* out);
// the FP64 twiddle table of period P (ilh_fft.hip), built on first use
const double2* fft_twiddles(ilh_ctx* ctx, int P);

// ---------------------------------------- fused batched solves (ilh_batch.hip)
// One problem of a fused batch of Dirichlet repeat solves: device operands.
struct BatchProb {
    const double* Mp;          // exterior-data solve matrix (row-major, leading dimension ld)
    const double* Mm;          // interior-data one, or null
    const double* fp;          // boundary data f+ (n), f- (n or null)
    const double* fm;
    double* amp;               // amplitudes [s; f+ - f-] (2n)
    double* s;                 // surface strength out (n)
    int n, ld;
    const int4* pinfo;         // merged source pairs (pinned count npairs)
    const int4* cells;
    const int* cai;
    const double* cw;
    int npairs;
    int nx, ny, Px, Py;        // primal grid and periods
    double2* T1;               // the problem's spectra buffers
    const int* slot;
    const double* spec;
    double2* T2;
    double* out;               // grid field out (nx ny)
    const double2* twx;        // twiddles of Px and Py
    const double2* twy;
    const int* npairs_dev;     // the live pair count on the device (npairs then an upper bound), or null
};
// Per pass (surface step, rows, columns, inverse rows): schedule (0 one
// work-group per tile, 1 persistent queues), work-groups per CU of the
// queue schedule, threads per work-group; the row pass's amplitude LDS cap.
struct FusedCfg {
    int sched[4], mult[4], tpb[4];
    int amp_cap;
};
struct FusedBatch;
bool fused_supported(int Px, int Py);
FusedBatch* fused_create(ilh_ctx* ctx, const std::vector<BatchProb>& probs);
void fused_destroy(FusedBatch* fb);
// the passes from first_pass on (1: the convolution only, of amplitudes
// already in place)
void fused_run(FusedBatch* fb, int first_pass = 0);
void fused_tune(FusedBatch* fb);
// cfg: 13 ints (sched, mult, threads of each pass; the amplitude cap), ms:
// the tuned time of each pass (0 before tuning)
void fused_config(const FusedBatch* fb, int* cfg, float* ms);
const std::vector<BatchProb>& fused_probs(const FusedBatch* fb);

// operators on the plan's geometry (ilh_api.hip): scatter R/D_s-type ops to the
// grid, gather E-type ops to the surface; inverse of a dense N x N matrix
// (overwritten) into row-major Minv with leading dimension solve_ld(N)
void scatter_op(ilh_plan* pl, int op, const double* amp, double* out);
// out = base + c (op amp), one pass (base and out grid-sized, distinct)
void scatter_op_add(ilh_plan* pl, int op, const double* amp, const double* base, double c, double* out);
void gather_op(ilh_plan* pl, int op, const double* field, double* out);
// nb operands at once: amplitudes / fields k at a stride, outputs k at a stride
void scatter_op_multi(ilh_plan* pl, int op, const double* amp, size_t amp_stride, int nb, double* out,
                      size_t out_stride);
void gather_op_multi(ilh_plan* pl, int op, const double* field, size_t field_stride, int nb, double* out,
                     size_t out_stride);
// FP32 data of a Float32 cache
void scatter_op(ilh_plan* pl, int op, const float* amp, float* out);
void gather_op(ilh_plan* pl, int op, const float* field, float* out);
void inverse_rowmajor(ilh_plan* pl, double* K, double* Minv, int n = 0);

// dense linear algebra (ilh_dense.hip)
// K = scale * (opA . kernel . opB): column-major N x N with leading dimension
// ldk (0: N), or row-major (K[q + ldk p]) when rowmajor
void assemble(ilh_plan* pl, int opA, int opB, int gk, double scale, bool delta_kernel, double* K, int ldk = 0,
              bool rowmajor = false);
void assemble_table(ilh_plan* pl, int opA, int opB, const double* tab, int Tn, double factor,
                    double* K /* N x N col-major */);
int  lu_factor(ilh_plan* pl, double* A, int n, int* piv_dev, int* info_dev);
// C (M x Nc) = beta C + alpha opA(A) opB(B) on the FP64 matrix cores (ilh_chol.hip),
// column-major, ta/tb = 1 for a transposed operand, beta 0 or 1; lower: only
// the lower triangle of a square C is needed (ta = 0, tb = 1, beta = 1)
// tri: the K range of each tile restricted to a triangular operand's nonzero half
enum { TRI_NONE = 0, TRI_A_LOWER = 1, TRI_A_UPPER = 2, TRI_B_LOWER = 3, TRI_B_UPPER = 4 };
void dgemm(hipStream_t st, int ta, int tb, int M, int Nc, int K, double alpha, const double* A, int lda,
           const double* B, int ldb, double beta, double* C, int ldc, bool lower = false, int tri = TRI_NONE);
// Recursive Cholesky (ilh_chol.hip): W (n x n) = L^-1 for A = L L^T; 0 or the
// first non-positive pivot.  chol_solve: X <- W^T W X.
int  chol_factor(ilh_plan* pl, double* A, int n, int lda, double* W, double* T, int* info_dev);
void chol_solve(ilh_plan* pl, const double* W, int n, double* X, int ldx, int nrhs, double* T);
// A symmetric matrix M that is positive definite on the vectors orthogonal to
// the constant one (the symmetrised Schur matrices E K E^T), factored as a
// Householder-deflated Cholesky (ilh_chol.hip).  The caller fills A (n x n,
// both triangles) and calls sym_factor; false: a non-positive pivot (the
// caller falls back to LU).  sym_solve: X <- M^-1 X (n x nrhs, leading dim n).
struct SymFactor {
    int n = 0;
    bool ok = false;
    DevBuf A, W, T, T2, w, wt, vec, sig, r, info;
};
bool sym_factor(ilh_plan* pl, SymFactor& F, int n);
void sym_solve(ilh_plan* pl, SymFactor& F, double* X, int nrhs);
// with S = M diag(c d): Minv = S^-1 row-major (leading dimension solve_ld(n)),
// and X <- diag(1 / (c d)) X (n x ncols)
void sym_inverse_rowmajor(ilh_plan* pl, SymFactor& F, const double* d, double c, double* Minv);
void scale_rows_inv(hipStream_t st, double* X, int n, int ncols, const double* d, double c);
void lu_solve_multi(ilh_plan* pl, const double* LU, const int* piv, int n, double* B, int nrhs);
// row-major solve matrices use a leading dimension padded to a multiple of 32
inline int solve_ld(int n) { return (n + 31) / 32 * 32; }
void transpose(hipStream_t st, const double* A, double* At, int rows, int cols, int ldt);
// a.x b.x + a.y b.y with one explicit fused multiply-add, so that the rounding
// is the same wherever the compiler places it
__device__ __forceinline__ double dot2(double2 a, double2 b) { return __builtin_fma(a.x, b.x, a.y * b.y); }

// dot(row, x) over m entries by one wave of 64 lanes (16-byte loads when x is
// 16-byte aligned), the same fixed order on every call
template <int CH = 4>
__device__ __forceinline__ double wave_dot(const double* __restrict__ row, const double* __restrict__ x, int m, int lane)
{
    if (((uintptr_t)x & 15) != 0) {   // caller's vector not 16-byte aligned: plain loads
        double s = 0.0;
        for (int k = lane; k < m; k += 64) s = __builtin_fma(row[k], x[k], s);
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
        return s;
    }
    // 16-byte loads, four independent partial sums per lane: lane l adds
    // entries l + 256 j + 64 u into s_u while l + 256 j + 192 < m/2, then the
    // remaining l + 256 J + 64 u into s0.  Loads go out CH of those rounds
    // (8 CH loads) at a time, and the remainder's with the last of them, all
    // before the first is used: with CH = 4 a row of a few thousand entries is
    // one or two memory round trips.  CH only sets how many loads are in
    // flight (and the registers they hold: kernels with many waves per SIMD
    // use CH = 1); every CH sums in the same order.
    const double2* r2 = (const double2*)row;
    const double2* x2 = (const double2*)x;
    const int m2 = m >> 1;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int k = lane;
    for (;;) {
        double2 a[4 * CH], b[4 * CH], ra[3], rb[3];
        // rounds j < 4 of this chunk, and where the remainder starts if the
        // chunk holds the last round
        int nr = 0;
#pragma unroll
        for (int j = 0; j < CH; ++j) nr += k + 256 * j + 192 < m2 ? 1 : 0;
        const int kr = k + 256 * nr;
#pragma unroll
        for (int j = 0; j < CH; ++j)
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const bool ok = j < nr;
                a[4 * j + u] = ok ? r2[k + 256 * j + 64 * u] : make_double2(0.0, 0.0);
                b[4 * j + u] = ok ? x2[k + 256 * j + 64 * u] : make_double2(0.0, 0.0);
            }
#pragma unroll
        for (int u = 0; u < 3; ++u) {
            const bool ok = nr < CH && kr + 64 * u < m2;
            ra[u] = ok ? r2[kr + 64 * u] : make_double2(0.0, 0.0);
            rb[u] = ok ? x2[kr + 64 * u] : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int j = 0; j < CH; ++j)
            if (j < nr) {
                s0 += dot2(a[4 * j], b[4 * j]);
                s1 += dot2(a[4 * j + 1], b[4 * j + 1]);
                s2 += dot2(a[4 * j + 2], b[4 * j + 2]);
                s3 += dot2(a[4 * j + 3], b[4 * j + 3]);
            }
        k = kr;
        if (nr < CH) {
#pragma unroll
            for (int u = 0; u < 3; ++u)
                if (kr + 64 * u < m2) s0 += dot2(ra[u], rb[u]);
            break;
        }
    }
    if ((m & 1) && lane == 0) s1 = __builtin_fma(row[m - 1], x[m - 1], s1);
    double s = (s0 + s1) + (s2 + s3);
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    return s;
}

void matvec_rows(hipStream_t st, const double* Mrow, int ld, const double* x, double* y, int n, int m,
                 double alpha, int accumulate);
void dirichlet_amp(hipStream_t st, const double* Mp, const double* Mm, int ld, const double* fp,
                   const double* fm, double* amp, double* s_out, int n);
void dirichlet_amp_multi(hipStream_t st, const double* Mp, const double* Mm, int ld, const double* fp,
                         const double* fm, int ldf, int nb, double* amp, double* s_out, int n);
void neumann_amp(hipStream_t st, const double* Mdf_p, const double* Mdf_m, const double* Mds_p,
                 const double* Mds_m, int ld, const double* vp, const double* vm, double* amp_f, double* amp_s,
                 double* df, double* dsv, int n);
void check_finite(hipStream_t st, const double* a, size_t n, int* flag_dev);

// learned preconditioner (ilh_learned.hip): neighbour table of the plan,
// predicted band coefficients (N x (2K+1)) of its current geometry, and the
// band's product z (+)= B v (skipped once flags[0] is set)
void plan_neighbours(ilh_plan* pl, int K);
void lpc_band(ilh_lpc* m, ilh_plan* pl, double* coef);
int lpc_width(const ilh_lpc* m);
int lpc_half_band(const ilh_lpc* m);
void band_apply(hipStream_t st, const double* coef, const int* nbr, int W, const double* v, double* z, int n,
                const int* flags, int accumulate);

// sum_j coef[j] v[nbr[j]] over one band row of W entries by one wave, in the
// order j = 0 .. W-1 (one fused multiply-add per entry, every lane gets the
// sum): lane l loads coef[j0 + l] and nbr[j0 + l], then gathers v, 64 entries
// per round of two dependent loads; the chain then reads the lanes' values in
// order (v_readlane), the same arithmetic as a single thread's loop
__device__ __forceinline__ double readlane_f64(double x, int l)
{
    const long long b = __double_as_longlong(x);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)b, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

__device__ __forceinline__ double band_row(const double* __restrict__ coef, const int* __restrict__ nbr, int W,
                                           const double* __restrict__ v, int lane)
{
    double s = 0.0;
    for (int j0 = 0; j0 < W; j0 += 64) {
        const bool ok = j0 + lane < W;
        const int ix = ok ? nbr[j0 + lane] : 0;
        const double c = ok ? coef[j0 + lane] : 0.0;
        const double x = ok ? v[ix] : 0.0;
        const int cnt = W - j0 < 64 ? W - j0 : 64;
        for (int u = 0; u < cnt; ++u) s = __builtin_fma(readlane_f64(c, u), readlane_f64(x, u), s);
    }
    return s;
}
// S^-1 of the plan's create_RTLinvR, row-major with leading dimension solve_ld(N) (ilh_api.hip)
void exact_inverse_rowmajor(ilh_plan* pl, double* Minv);

}  // namespace ilh

// ---------------------------------------------------- C-ABI call helpers
namespace ilh {
void set_last_error(const std::string& msg);   // thread-local message of ilh_last_error()

// run f, turning exceptions into the C ABI's status codes
template <class F>
int guarded_call(F&& f)
{
    try {
        f();
        return ILH_OK;
    } catch (const Error& e) {
        set_last_error(e.msg);
        return e.code;
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed");
        return ILH_ERR_OOM;
    } catch (...) {
        set_last_error("unknown error");
        return ILH_ERR_ARG;
    }
}

void need_margin(ilh_plan* pl, const char* what);

// host or device array arguments of one call: host arrays are staged through
// the plan's buffers on its stream
struct Io {
    ilh_plan* pl;
    int mem;
    Io(ilh_plan* p, int m) : pl(p), mem(m)
    {
        if (!p) fail(ILH_ERR_ARG, "null plan");
        if (m != ILH_MEM_HOST && m != ILH_MEM_DEVICE) fail(ILH_ERR_ARG, "mem must be ILH_MEM_HOST or ILH_MEM_DEVICE");
        ILH_HIP(hipSetDevice(p->ctx->device));
    }
    template <class T> const T* in(const T* h, size_t n, DevBuf& buf)
    {
        if (n > 0 && !h) fail(ILH_ERR_ARG, "null input array");
        if (mem == ILH_MEM_DEVICE) return h;
        buf.ensure(sizeof(T) * std::max<size_t>(n, 1));
        if (n) ILH_HIP(hipMemcpyAsync(buf.p, h, sizeof(T) * n, hipMemcpyHostToDevice, pl->ctx->stream));
        return buf.as<T>();
    }
    template <class T> T* out(T* h, size_t n, DevBuf& buf)
    {
        if (n > 0 && !h) fail(ILH_ERR_ARG, "null output array");
        if (mem == ILH_MEM_DEVICE) return h;
        buf.ensure(sizeof(T) * std::max<size_t>(n, 1));
        return buf.as<T>();
    }
    template <class T> void back(T* h, const T* d, size_t n)
    {
        if (mem == ILH_MEM_DEVICE || n == 0) return;
        ILH_HIP(hipMemcpyAsync(h, d, sizeof(T) * n, hipMemcpyDeviceToHost, pl->ctx->stream));
    }
    void done()
    {
        if (mem == ILH_MEM_HOST) ILH_HIP(hipStreamSynchronize(pl->ctx->stream));
    }
};

}  // namespace ilh
