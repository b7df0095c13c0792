/*
 * ilh_oracle.h -- CPU FP64 restatement of the immersed-layer hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path
 * (libimmersedhip.so): only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product never links or calls it.
 *
 * Conventions (restating CartesianGrids.jl as used by ImmersedLayers.jl):
 *   grid   = (NX, NY) dual-node counts, I0 = origin index (1-based, Julia),
 *            dx = cell size.
 *   kinds  = ORC_PRIMAL  Nodes{Primal}  (NX-1)x(NY-1), x_i=(i+1-I0x)dx
 *            ORC_DUAL    Nodes{Dual}    NX x NY,       x_i=(i+1-I0x-1/2)dx
 *            ORC_EDGE_U  Edges{Primal}.u NX x (NY-1)  (dual x, primal y)
 *            ORC_EDGE_V  Edges{Primal}.v (NX-1) x NY  (primal x, dual y)
 *   arrays are column-major (x fastest), 0-based indices.
 *   Edges{Primal} arrays are stored as [u; v] (u block then v block).
 *   VectorData is [u(N); v(N)].
 * Only GridScaling is restated (the reference default, ImmersedLayers.jl:81).
 */
#ifndef ILH_ORACLE_H
#define ILH_ORACLE_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct { int NX, NY, I0x, I0y; double dx; } orc_grid;

enum { ORC_PRIMAL = 0, ORC_DUAL = 1, ORC_EDGE_U = 2, ORC_EDGE_V = 3 };
enum { ORC_YANG3 = 0, ORC_ROMA = 1, ORC_M4PRIME = 2, ORC_WITCHHAT = 3 };

double orc_ddf(int ddf, double r);
void   orc_dims(const orc_grid* g, int kind, int* nx, int* ny);

/* surface <-> grid (A4-A7) */
void orc_regularize(const orc_grid* g, int kind, int ddf, int N, const double* X, const double* Y,
                    const double* ds, const double* f, double* out);
void orc_interpolate(const orc_grid* g, int kind, int ddf, int N, const double* X, const double* Y,
                     const double* grid, double* out);
void orc_regularize_normal(const orc_grid* g, int ddf, int N, const double* X, const double* Y,
                           const double* nx, const double* ny, const double* ds, const double* f,
                           double* q_edges);
void orc_normal_interpolate(const orc_grid* g, int ddf, int N, const double* X, const double* Y,
                            const double* nx, const double* ny, const double* q_edges, double* vn);
void orc_regularize_normal_cross(const orc_grid* g, int ddf, int N, const double* X, const double* Y,
                                 const double* nx, const double* ny, const double* ds, const double* f,
                                 double* q_edges);
void orc_normal_cross_interpolate(const orc_grid* g, int ddf, int N, const double* X, const double* Y,
                                  const double* nx, const double* ny, const double* q_edges, double* vn);

/* grid stencils (A8), GridScaling */
void orc_divergence(const orc_grid* g, const double* q_edges, double* p);
void orc_grad(const orc_grid* g, const double* p, double* q_edges);
void orc_curl_n2e(const orc_grid* g, const double* s_dual, double* q_edges);
void orc_curl_e2n(const orc_grid* g, const double* q_edges, double* w_dual);
void orc_laplacian(const orc_grid* g, int kind, const double* in, double* out);

/* lattice Green's function (A9) */
double orc_lgf(int m, int n);
void   orc_lgf_table(int mx, int my, double* tab);   /* tab[m + mx*n], 0<=m<mx, 0<=n<my */
void   orc_conv_direct(int nx, int ny, const double* tab, int tmx, const double* in, double* out);

typedef struct orc_lap orc_lap;
orc_lap* orc_lap_create(const orc_grid* g, int kind);
void     orc_lap_destroy(orc_lap* L);
void     orc_inverse_laplacian(orc_lap* L, const double* in, double* out);  /* in may == out */
const double* orc_lap_table(orc_lap* L, int* tmx);

/* surface calculus composites (A11) */
void orc_surface_divergence(const orc_grid* g, int ddf, int N, const double* X, const double* Y,
                            const double* nx, const double* ny, const double* ds, const double* f, double* theta);
void orc_surface_grad(const orc_grid* g, int ddf, int N, const double* X, const double* Y,
                      const double* nx, const double* ny, const double* phi, double* vn);
void orc_surface_curl(const orc_grid* g, int ddf, int N, const double* X, const double* Y,
                      const double* nx, const double* ny, const double* ds, const double* f, double* w);
void orc_surface_curl_adj(const orc_grid* g, int ddf, int N, const double* X, const double* Y,
                          const double* nx, const double* ny, const double* s, double* vn);
void orc_surface_curl_cross(const orc_grid* g, int ddf, int N, const double* X, const double* Y,
                            const double* nx, const double* ny, const double* ds, const double* f, double* w);
void orc_surface_curl_cross_adj(const orc_grid* g, int ddf, int N, const double* X, const double* Y,
                                const double* nx, const double* ny, const double* s, double* vn);
void orc_surface_grad_cross(const orc_grid* g, int ddf, int N, const double* X, const double* Y,
                            const double* nx, const double* ny, const double* phi, double* vn);

/* Schur builders (A10, A11); A is N x N column-major */
enum { ORC_S_RTLINVR = 0, ORC_S_CLINVCT = 1, ORC_S_GLINVD = 2, ORC_S_NRTRN = 3 };
void orc_schur_columns(const orc_grid* g, int ddf, int N, const double* X, const double* Y,
                       const double* nx, const double* ny, const double* ds, int kind, double scale, double* A);
void orc_schur_direct(const orc_grid* g, int ddf, int N, const double* X, const double* Y,
                      const double* nx, const double* ny, const double* ds, int kind, double scale, double* A);

/* vector cache: VectorData [u; v] (2N); EdgeGradient [dudx; dudy; dvdx; dvdy]
   with dudx, dvdy on Nodes{Primal} and dudy, dvdx on Nodes{Dual} */
size_t orc_tensor_size(const orc_grid* g);
void orc_regularize_normal_symm(const orc_grid* g, int ddf, int N, const double* X, const double* Y,
                                const double* nx, const double* ny, const double* ds, const double* f, double* A);
void orc_normal_interpolate_symm(const orc_grid* g, int ddf, int N, const double* X, const double* Y,
                                 const double* nx, const double* ny, const double* A, double* tau);
void orc_surface_divergence_symm(const orc_grid* g, int ddf, int N, const double* X, const double* Y,
                                 const double* nx, const double* ny, const double* ds, const double* f, double* q);
void orc_surface_grad_symm(const orc_grid* g, int ddf, int N, const double* X, const double* Y,
                           const double* nx, const double* ny, const double* q, double* tau);
void orc_surface_curl_vector(const orc_grid* g, int ddf, int N, const double* X, const double* Y,
                             const double* ds, const double* f, double* w);
void orc_surface_curl_vector_adj(const orc_grid* g, int ddf, int N, const double* X, const double* Y,
                                 const double* s, double* v);
void orc_cl2invct(const orc_grid* g, int ddf, int N, const double* X, const double* Y, const double* ds,
                  double scale, double* A);
void orc_glinvd_symm(const orc_grid* g, int ddf, int N, const double* X, const double* Y, const double* nx,
                     const double* ny, const double* ds, double scale, double* A);
int orc_stokes_solve(const orc_grid* g, int ddf, int N, const double* X, const double* Y,
                     const double* nx, const double* ny, const double* ds, const double* S2, const double* Ss,
                     const double* vplus, const double* vminus, double* v, double* s, double* sigma);

/* dense LU with partial pivoting (A12) */
/* unbounded heat conduction with Dirichlet bodies, IF-HERK (the constrained
   system of test/literate/heatconduction.jl:88-180 marched by ConstrainedSystems'
   LiskaIFHERK, src/timemarching.jl:5): lattice heat kernel K(p) = e^{-2a} I_p(2a)
   (p = 0..M, returns M+1 <= maxlen), its separable application on an nx x ny
   node field (zero outside), and nsteps steps of T (Nodes{Primal}, in place). */
int  orc_heat_kernel(double a, double* K, int maxlen);
void orc_heat_apply(int nx, int ny, const double* K, int M, const double* in, double* out);
/* the stage matrix E H R of kernel K (K(|m|) K(|n|), |m|,|n| <= M; NULL: the
   identity), method 1 the column loop of the reference's create_RTLinvR shape,
   0 the direct assembly */
void orc_heat_schur(const orc_grid* g, int ddf, int N, const double* X, const double* Y, const double* ds,
                    const double* K, int M, int method, double* S);
int  orc_heat_advance(const orc_grid* g, int ddf, int N, const double* X, const double* Y, const double* nx,
                      const double* ny, const double* ds, double kappa, double dt, int nsteps, double* T,
                      const double* tb_plus, const double* tb_minus);

int  orc_lu_factor(int n, double* A, int* piv);
void orc_lu_solve(int n, const double* LU, const int* piv, double* b);

/* composites (A13, A14).  S may be NULL (built by direct assembly). */
int orc_dirichlet_solve(const orc_grid* g, int ddf, int N, const double* X, const double* Y,
                        const double* nx, const double* ny, const double* ds, const double* S,
                        const double* fplus, const double* fminus, double* f, double* s);
int orc_neumann_solve(const orc_grid* g, int ddf, int N, const double* X, const double* Y,
                      const double* nx, const double* ny, const double* ds, const double* S,
                      const double* vplus, const double* vminus,
                      double* f, double* df, double* s, double* dsv);

void orc_set_threads(int n);
int  orc_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
