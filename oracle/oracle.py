"""ctypes wrapper of the CPU FP64 oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module.
The restated reference functions and their citations are in ilh_oracle.c.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")

PRIMAL, DUAL, EDGE_U, EDGE_V = 0, 1, 2, 3
YANG3, ROMA, M4PRIME, WITCHHAT = 0, 1, 2, 3
S_RTLINVR, S_CLINVCT, S_GLINVD, S_NRTRN = 0, 1, 2, 3


class Grid(C.Structure):
    _fields_ = [("NX", C.c_int), ("NY", C.c_int), ("I0x", C.c_int), ("I0y", C.c_int), ("dx", C.c_double)]


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def _load():
    if not os.path.exists(_LIB):
        build()
    return C.CDLL(_LIB)


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
        dp = C.POINTER(C.c_double)
        _lib.orc_ddf.restype = C.c_double
        _lib.orc_ddf.argtypes = [C.c_int, C.c_double]
        _lib.orc_lgf.restype = C.c_double
        _lib.orc_lgf.argtypes = [C.c_int, C.c_int]
        _lib.orc_lap_create.restype = C.c_void_p
        _lib.orc_lap_create.argtypes = [C.POINTER(Grid), C.c_int]
        _lib.orc_lap_destroy.argtypes = [C.c_void_p]
        _lib.orc_inverse_laplacian.argtypes = [C.c_void_p, dp, dp]
        _lib.orc_lap_table.restype = dp
        _lib.orc_lap_table.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def grid(NX, NY, I0x, I0y, dx):
    return Grid(NX, NY, I0x, I0y, dx)


def dims(g, kind):
    return {PRIMAL: (g.NX - 1, g.NY - 1), DUAL: (g.NX, g.NY),
            EDGE_U: (g.NX, g.NY - 1), EDGE_V: (g.NX - 1, g.NY)}[kind]


def nel(g, kind):
    nx, ny = dims(g, kind)
    return nx * ny


def edge_len(g):
    return nel(g, EDGE_U) + nel(g, EDGE_V)


def set_threads(n):
    lib().orc_set_threads(int(n))


def get_threads():
    return lib().orc_get_threads()


def ddf(kind, r):
    return lib().orc_ddf(kind, float(r))


def lgf(m, n):
    return lib().orc_lgf(int(m), int(n))


def lgf_table(mx, my):
    t = np.zeros(mx * my)
    lib().orc_lgf_table(mx, my, _p(t))
    return t.reshape(my, mx).T.copy()     # [m, n]


def conv_direct(tab, w):
    """tab: G table indexed [m, n] (Fortran-ordered copy made here); w: (nx, ny)."""
    nx, ny = w.shape
    t = np.asfortranarray(tab)
    tflat = t.ravel(order="F").copy()
    out = np.zeros(nx * ny)
    wf = w.ravel(order="F").copy()
    lib().orc_conv_direct(nx, ny, _p(tflat), tab.shape[0], _p(wf), _p(out))
    return out.reshape((nx, ny), order="F")


class Geometry:
    """Surface points, unit normals and panel lengths as plain FP64 arrays."""

    def __init__(self, X, Y, nx, ny, ds):
        self.X = np.ascontiguousarray(X, dtype=np.float64)
        self.Y = np.ascontiguousarray(Y, dtype=np.float64)
        self.nx = np.ascontiguousarray(nx, dtype=np.float64)
        self.ny = np.ascontiguousarray(ny, dtype=np.float64)
        self.ds = np.ascontiguousarray(ds, dtype=np.float64)
        self.N = len(self.X)

    def args(self):
        return self.N, _p(self.X), _p(self.Y)


def regularize(g, kind, geo, f, ddf_kind=YANG3):
    out = np.zeros(nel(g, kind))
    lib().orc_regularize(C.byref(g), kind, ddf_kind, *geo.args(), _p(geo.ds), _p(np.ascontiguousarray(f, float)), _p(out))
    return out


def interpolate(g, kind, geo, field, ddf_kind=YANG3):
    out = np.zeros(geo.N)
    lib().orc_interpolate(C.byref(g), kind, ddf_kind, *geo.args(), _p(np.ascontiguousarray(field, float)), _p(out))
    return out


def regularize_normal(g, geo, f, ddf_kind=YANG3):
    out = np.zeros(edge_len(g))
    lib().orc_regularize_normal(C.byref(g), ddf_kind, *geo.args(), _p(geo.nx), _p(geo.ny), _p(geo.ds),
                                _p(np.ascontiguousarray(f, float)), _p(out))
    return out


def normal_interpolate(g, geo, q, ddf_kind=YANG3):
    out = np.zeros(geo.N)
    lib().orc_normal_interpolate(C.byref(g), ddf_kind, *geo.args(), _p(geo.nx), _p(geo.ny),
                                 _p(np.ascontiguousarray(q, float)), _p(out))
    return out


def regularize_normal_cross(g, geo, f, ddf_kind=YANG3):
    out = np.zeros(edge_len(g))
    lib().orc_regularize_normal_cross(C.byref(g), ddf_kind, *geo.args(), _p(geo.nx), _p(geo.ny), _p(geo.ds),
                                      _p(np.ascontiguousarray(f, float)), _p(out))
    return out


def normal_cross_interpolate(g, geo, q, ddf_kind=YANG3):
    out = np.zeros(geo.N)
    lib().orc_normal_cross_interpolate(C.byref(g), ddf_kind, *geo.args(), _p(geo.nx), _p(geo.ny),
                                       _p(np.ascontiguousarray(q, float)), _p(out))
    return out


def _grid_to_grid(fn, g, n_out, a):
    out = np.zeros(n_out)
    fn(C.byref(g), _p(np.ascontiguousarray(a, float)), _p(out))
    return out


def divergence(g, q):
    return _grid_to_grid(lib().orc_divergence, g, nel(g, PRIMAL), q)


def grad(g, p):
    return _grid_to_grid(lib().orc_grad, g, edge_len(g), p)


def curl_n2e(g, s):
    return _grid_to_grid(lib().orc_curl_n2e, g, edge_len(g), s)


def curl_e2n(g, q):
    return _grid_to_grid(lib().orc_curl_e2n, g, nel(g, DUAL), q)


def laplacian(g, kind, s):
    out = np.zeros(nel(g, kind))
    lib().orc_laplacian(C.byref(g), kind, _p(np.ascontiguousarray(s, float)), _p(out))
    return out


class InverseLaplacian:
    """Zero-padded FFT convolution with the lattice Green's function."""

    def __init__(self, g, kind):
        self.g, self.kind = g, kind
        self.h = lib().orc_lap_create(C.byref(g), kind)

    def __call__(self, w):
        w = np.ascontiguousarray(w, float)
        out = np.zeros(nel(self.g, self.kind))
        lib().orc_inverse_laplacian(self.h, _p(w), _p(out))
        return out

    def table(self):
        tmx = C.c_int()
        ptr = lib().orc_lap_table(self.h, C.byref(tmx))
        nx, ny = dims(self.g, self.kind)
        return tmx.value, ptr

    def __del__(self):
        try:
            lib().orc_lap_destroy(self.h)
        except Exception:
            pass


def _surf_to_grid(name, g, geo, f, kind, ddf_kind=YANG3):
    out = np.zeros(nel(g, kind))
    getattr(lib(), name)(C.byref(g), ddf_kind, *geo.args(), _p(geo.nx), _p(geo.ny), _p(geo.ds),
                         _p(np.ascontiguousarray(f, float)), _p(out))
    return out


def _grid_to_surf(name, g, geo, a, ddf_kind=YANG3):
    out = np.zeros(geo.N)
    getattr(lib(), name)(C.byref(g), ddf_kind, *geo.args(), _p(geo.nx), _p(geo.ny),
                         _p(np.ascontiguousarray(a, float)), _p(out))
    return out


def surface_divergence(g, geo, f, ddf_kind=YANG3):
    return _surf_to_grid("orc_surface_divergence", g, geo, f, PRIMAL, ddf_kind)


def surface_curl(g, geo, f, ddf_kind=YANG3):
    return _surf_to_grid("orc_surface_curl", g, geo, f, DUAL, ddf_kind)


def surface_curl_cross(g, geo, f, ddf_kind=YANG3):
    return _surf_to_grid("orc_surface_curl_cross", g, geo, f, DUAL, ddf_kind)


def surface_grad(g, geo, phi, ddf_kind=YANG3):
    return _grid_to_surf("orc_surface_grad", g, geo, phi, ddf_kind)


def surface_curl_adj(g, geo, s, ddf_kind=YANG3):
    return _grid_to_surf("orc_surface_curl_adj", g, geo, s, ddf_kind)


def surface_curl_cross_adj(g, geo, s, ddf_kind=YANG3):
    return _grid_to_surf("orc_surface_curl_cross_adj", g, geo, s, ddf_kind)


def surface_grad_cross(g, geo, phi, ddf_kind=YANG3):
    return _grid_to_surf("orc_surface_grad_cross", g, geo, phi, ddf_kind)


def schur(g, geo, kind, scale=1.0, method="direct", ddf_kind=YANG3):
    """N x N matrix (numpy, [row, col]) of create_RTLinvR / create_CLinvCT / create_GLinvD / create_nRTRn."""
    A = np.zeros(geo.N * geo.N)
    fn = lib().orc_schur_direct if method == "direct" else lib().orc_schur_columns
    fn(C.byref(g), ddf_kind, *geo.args(), _p(geo.nx), _p(geo.ny), _p(geo.ds), kind, C.c_double(scale), _p(A))
    return A.reshape(geo.N, geo.N).T.copy()


def lu_solve(S, b):
    n = S.shape[0]
    A = np.asfortranarray(S).ravel(order="F").copy()
    piv = np.zeros(n, dtype=np.int32)
    info = lib().orc_lu_factor(n, _p(A), piv.ctypes.data_as(C.POINTER(C.c_int)))
    x = np.ascontiguousarray(b, float).copy()
    lib().orc_lu_solve(n, _p(A), piv.ctypes.data_as(C.POINTER(C.c_int)), _p(x))
    return info, x


def dirichlet_solve(g, geo, fplus, fminus=None, S=None, ddf_kind=YANG3):
    """Returns (f, s) of test/literate/problems.jl:259-298; S in [row, col] or None."""
    f = np.zeros(nel(g, PRIMAL))
    s = np.zeros(geo.N)
    Sf = None if S is None else np.asfortranarray(S).ravel(order="F").copy()
    fm = None if fminus is None else np.ascontiguousarray(fminus, float)
    info = lib().orc_dirichlet_solve(C.byref(g), ddf_kind, *geo.args(), _p(geo.nx), _p(geo.ny), _p(geo.ds),
                                     _p(Sf), _p(np.ascontiguousarray(fplus, float)), _p(fm), _p(f), _p(s))
    if info:
        raise np.linalg.LinAlgError("singular Schur complement (info=%d)" % info)
    return f, s


def neumann_solve(g, geo, vplus, vminus=None, S=None, ddf_kind=YANG3):
    """Returns (f, df, s, ds) of test/literate/neumann.jl:101-142."""
    f = np.zeros(nel(g, PRIMAL))
    df = np.zeros(geo.N)
    s = np.zeros(nel(g, DUAL))
    dsv = np.zeros(geo.N)
    Sf = None if S is None else np.asfortranarray(S).ravel(order="F").copy()
    vm = None if vminus is None else np.ascontiguousarray(vminus, float)
    info = lib().orc_neumann_solve(C.byref(g), ddf_kind, *geo.args(), _p(geo.nx), _p(geo.ny), _p(geo.ds),
                                   _p(Sf), _p(np.ascontiguousarray(vplus, float)), _p(vm),
                                   _p(f), _p(df), _p(s), _p(dsv))
    if info:
        raise np.linalg.LinAlgError("singular Schur complement (info=%d)" % info)
    return f, df, s, dsv


# ------------------------------------------------------------------ heat
def heat_kernel(a, maxlen=512):
    """K(p) = e^{-2a} I_p(2a), p = 0..M (the 1-D lattice heat kernel of exp(a L1))."""
    K = np.zeros(maxlen)
    n = lib().orc_heat_kernel(C.c_double(a), _p(K), maxlen)
    return K[:n].copy()


def heat_apply(nx, ny, K, field):
    """Separable application of the kernel K to an (nx, ny) node field (zero outside)."""
    K = np.ascontiguousarray(K, float)
    a = np.ascontiguousarray(np.asarray(field, float).ravel(order="F"))
    out = np.zeros(nx * ny)
    lib().orc_heat_apply(nx, ny, _p(K), len(K) - 1, _p(a), _p(out))
    return out


def heat_schur(g, geo, K, method="direct", ddf_kind=YANG3):
    """Stage matrix E H R (numpy [row, col]) of the separable kernel K (None: identity)."""
    A = np.zeros(geo.N * geo.N)
    Kp = None if K is None else np.ascontiguousarray(K, float)
    lib().orc_heat_schur(C.byref(g), ddf_kind, *geo.args(), _p(geo.ds), _p(Kp), 0 if K is None else len(K) - 1,
                         1 if method == "columns" else 0, _p(A))
    return A.reshape(geo.N, geo.N).T.copy()


def heat_advance(g, geo, kappa, dt, nsteps, T0, tb_plus, tb_minus=None, ddf_kind=YANG3):
    """nsteps IF-HERK steps of the Dirichlet heat conduction problem (Nodes{Primal} T)."""
    T = np.ascontiguousarray(np.asarray(T0, float).ravel(order="F")).copy()
    tp = np.ascontiguousarray(tb_plus, float)
    tm = None if tb_minus is None else np.ascontiguousarray(tb_minus, float)
    info = lib().orc_heat_advance(C.byref(g), ddf_kind, *geo.args(), _p(geo.nx), _p(geo.ny), _p(geo.ds),
                                  C.c_double(kappa), C.c_double(dt), int(nsteps), _p(T), _p(tp), _p(tm))
    if info:
        raise np.linalg.LinAlgError("singular stage Schur matrix")
    return T


def mask(g, geo, complementary=False, ddf_kind=YANG3):
    """mask(cache) on Nodes{Primal}: -L^-1 D_s 1 (src/surface_operators.jl:863-868);
    ones without points (:870-873); complementary_mask is 1 - mask."""
    if geo is None or geo.N == 0:
        m = np.ones(nel(g, PRIMAL))
    else:
        th = surface_divergence(g, geo, np.ones(geo.N), ddf_kind)
        m = -InverseLaplacian(g, PRIMAL)(th)
    return 1.0 - m if complementary else m


def primal_to(g, m, kind):
    """Nodes{Primal} data averaged onto another grid kind, as CartesianGrids'
    grid_interpolate! (used by src/surface_operators.jl:882-900): a dual node from
    its four surrounding primal nodes, an x edge (EDGE_U) from the two primal nodes
    beside it in x, a y edge (EDGE_V) from the two in y.  The outer ring, which
    has no surrounding nodes, is 0."""
    NX, NY = g.NX, g.NY
    p = np.asarray(m, float).reshape((NX - 1, NY - 1), order="F")
    if kind == PRIMAL:
        return p.ravel(order="F").copy()
    if kind == DUAL:
        out = np.zeros((NX, NY))
        out[1:-1, 1:-1] = 0.25 * (p[:-1, :-1] + p[1:, :-1] + p[:-1, 1:] + p[1:, 1:])
    elif kind == EDGE_U:
        out = np.zeros((NX, NY - 1))
        out[1:-1, :] = 0.5 * (p[:-1, :] + p[1:, :])
    else:
        out = np.zeros((NX - 1, NY))
        out[:, 1:-1] = 0.5 * (p[:, :-1] + p[:, 1:])
    return out.ravel(order="F")


def surface_filter(g, geo, ddf_kind=YANG3):
    """C = Et R (create_surface_filter, src/matrix_operators.jl:254-267), dense
    [row, col].  The reference builds Et from CartesianGrids' regularization
    with filter=true (a dependency not in the reference tree; Project.toml
    pins CartesianGrids 0.1.29 / 0.2), the filtered interpolation of Goza et
    al. (2016, J. Comput. Phys. 321, 860-873) that test/literate/matrices.jl:55,73
    cites: the interpolation weight of point p at grid node j divided by the
    regularized unit strength there,
        Et[p, j] = E[p, j] / (R 1)_j    (0 where (R 1)_j = 0),
    so that C keeps constants and the surface integral, sum_p ds_p (C s)_p =
    sum_q ds_q s_q (the property matrices.jl:49-56 states), and ds C is
    symmetric positive semidefinite.  E is formed from R by adjointness
    (E = dx^2 diag(1/ds) R^T)."""
    N = geo.N
    R = np.zeros((nel(g, PRIMAL), N))
    for q in range(N):
        e = np.zeros(N)
        e[q] = 1.0
        R[:, q] = regularize(g, PRIMAL, geo, e, ddf_kind)
    r1 = R.sum(axis=1)
    w = np.zeros_like(r1)
    w[r1 != 0] = 1.0 / r1[r1 != 0]
    E = (R * (g.dx ** 2 / geo.ds)[None, :]).T
    return E @ (w[:, None] * R)


# ------------------------------------------------------------ vector cache
def tensor_len(g):
    """EdgeGradient{Primal,Dual} length: [dudx; dudy; dvdx; dvdy]."""
    return 2 * (nel(g, PRIMAL) + nel(g, DUAL))


def regularize_normal_symm(g, geo, f, ddf_kind=YANG3):
    out = np.zeros(tensor_len(g))
    lib().orc_regularize_normal_symm(C.byref(g), ddf_kind, *geo.args(), _p(geo.nx), _p(geo.ny), _p(geo.ds),
                                     _p(np.ascontiguousarray(f, float)), _p(out))
    return out


def normal_interpolate_symm(g, geo, A, ddf_kind=YANG3):
    out = np.zeros(2 * geo.N)
    lib().orc_normal_interpolate_symm(C.byref(g), ddf_kind, *geo.args(), _p(geo.nx), _p(geo.ny),
                                      _p(np.ascontiguousarray(A, float)), _p(out))
    return out


def surface_divergence_symm(g, geo, f, ddf_kind=YANG3):
    out = np.zeros(edge_len(g))
    lib().orc_surface_divergence_symm(C.byref(g), ddf_kind, *geo.args(), _p(geo.nx), _p(geo.ny), _p(geo.ds),
                                      _p(np.ascontiguousarray(f, float)), _p(out))
    return out


def surface_grad_symm(g, geo, q, ddf_kind=YANG3):
    out = np.zeros(2 * geo.N)
    lib().orc_surface_grad_symm(C.byref(g), ddf_kind, *geo.args(), _p(geo.nx), _p(geo.ny),
                                _p(np.ascontiguousarray(q, float)), _p(out))
    return out


def surface_curl_vector(g, geo, f, ddf_kind=YANG3):
    out = np.zeros(nel(g, DUAL))
    lib().orc_surface_curl_vector(C.byref(g), ddf_kind, *geo.args(), _p(geo.ds),
                                  _p(np.ascontiguousarray(f, float)), _p(out))
    return out


def surface_curl_vector_adj(g, geo, s, ddf_kind=YANG3):
    out = np.zeros(2 * geo.N)
    lib().orc_surface_curl_vector_adj(C.byref(g), ddf_kind, *geo.args(), _p(np.ascontiguousarray(s, float)),
                                      _p(out))
    return out


def cl2invct(g, geo, scale=1.0, ddf_kind=YANG3):
    """create_CL2invCT (src/matrix_operators.jl:95-125): 2N x 2N [row, col]."""
    n = 2 * geo.N
    A = np.zeros(n * n)
    lib().orc_cl2invct(C.byref(g), ddf_kind, *geo.args(), _p(geo.ds), C.c_double(scale), _p(A))
    return A.reshape(n, n).T.copy()


def glinvd_symm(g, geo, scale=1.0, ddf_kind=YANG3):
    """create_GLinvD_symm (src/matrix_operators.jl:165-185): 2N x 2N [row, col]."""
    n = 2 * geo.N
    A = np.zeros(n * n)
    lib().orc_glinvd_symm(C.byref(g), ddf_kind, *geo.args(), _p(geo.nx), _p(geo.ny), _p(geo.ds),
                          C.c_double(scale), _p(A))
    return A.reshape(n, n).T.copy()


def stokes_solve(g, geo, vplus, vminus=None, S2=None, Ss=None, ddf_kind=YANG3):
    """(v, s, sigma) of test/literate/stokes.jl:101-166 with sigma unfiltered."""
    v = np.zeros(edge_len(g))
    s = np.zeros(nel(g, DUAL))
    sig = np.zeros(2 * geo.N)
    S2f = None if S2 is None else np.asfortranarray(S2).ravel(order="F").copy()
    Ssf = None if Ss is None else np.asfortranarray(Ss).ravel(order="F").copy()
    vm = None if vminus is None else np.ascontiguousarray(vminus, float)
    info = lib().orc_stokes_solve(C.byref(g), ddf_kind, *geo.args(), _p(geo.nx), _p(geo.ny), _p(geo.ds),
                                  _p(S2f), _p(Ssf), _p(np.ascontiguousarray(vplus, float)), _p(vm),
                                  _p(v), _p(s), _p(sig))
    if info:
        raise np.linalg.LinAlgError("singular Stokes Schur complement (info=%d)" % info)
    return v, s, sig


def surface_filter_vector(g, geo, ddf_kind=YANG3):
    """The vector cache's surface filter: surface_filter's construction on each
    edge component (u from the u edges, v from the v edges), block diagonal
    2N x 2N [row, col]."""
    N = geo.N
    Cm = np.zeros((2 * N, 2 * N))
    for c, kind in enumerate((EDGE_U, EDGE_V)):
        R = np.zeros((nel(g, kind), N))
        for q in range(N):
            e = np.zeros(N)
            e[q] = 1.0
            R[:, q] = regularize(g, kind, geo, e, ddf_kind)
        r1 = R.sum(axis=1)
        w = np.zeros_like(r1)
        w[r1 != 0] = 1.0 / r1[r1 != 0]
        E = (R * (g.dx ** 2 / geo.ds)[None, :]).T
        Cm[c * N:(c + 1) * N, c * N:(c + 1) * N] = E @ (w[:, None] * R)
    return Cm
