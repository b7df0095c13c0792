"""Host-side mirror of ImmersedLayers.jl's scalar-cache API over libimmersedhip.so.

Same names, argument order (output first) and meaning as the reference's
operators; Julia's trailing ``!`` is dropped.  Grid data live in numpy arrays
with the reference's column-major layouts (Nodes are ``(nx, ny)`` Fortran-ordered
arrays, Edges{Primal} is one buffer ``[u; v]``, VectorData is ``[u; v]``), or in
device buffers (:class:`DeviceArray`) for chained calls without PCIe traffic.

Reference entry points mirrored here (paths in the reference checkout):
  SurfaceScalarCache          src/cache.jl:164-182, 205-224
  regularize / interpolate    src/surface_operators.jl:13-88
  regularize_normal(_cross)   src/surface_operators.jl:98-103, 150-162
  normal_(cross_)interpolate  src/surface_operators.jl:228-233, 277-290
  surface_divergence/grad/curl/curl_cross/grad_cross
                              src/surface_operators.jl:357-725
  divergence/grad/curl/laplacian/inverse_laplacian
                              src/grid_operators.jl:25-179
  create_RTLinvR/CLinvCT/GLinvD/nRTRn
                              src/matrix_operators.jl:9-61, 135-155, 225-244
"""
import atexit
import ctypes as C
import weakref

import numpy as np

from . import _lib as L
from .geometry import GridSpec

Primal, Dual = "Primal", "Dual"
GridScaling = "GridScaling"
Yang3, Roma, M4prime, Witchhat = L.DDF_YANG3, L.DDF_ROMA, L.DDF_M4PRIME, L.DDF_WITCHHAT


# Objects that own library handles, closed at interpreter exit in dependency
# order (buffers, heat integrators, batches, caches, contexts) while the library
# and the HIP runtime are still fully loaded, instead of in whatever order
# finalization reaches them.
_LIVE_ORDER = ("buffer", "heat", "batch", "cache", "ctx")
_live = {k: weakref.WeakSet() for k in _LIVE_ORDER}


@atexit.register
def _close_all():
    for kind in _LIVE_ORDER:
        for obj in list(_live[kind]):
            try:
                obj.free() if kind == "buffer" else obj.close()
            except Exception:
                pass


# ----------------------------------------------------------------- context
class Context:
    """One HIP device and one stream (``ilh_ctx``)."""

    def __init__(self, device=0):
        L.load()
        h = C.c_void_p()
        L.call("ilh_ctx_create", device, C.byref(h))
        self.h = h
        self.device = device
        _live["ctx"].add(self)

    def synchronize(self):
        L.call("ilh_ctx_synchronize", self.h)

    def close(self):
        if self.h:
            L.load().ilh_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


class DeviceArray:
    """FP64 (or FP32) device buffer owned by a context (``ilh_malloc``)."""

    def __init__(self, ctx, n, dtype=np.float64):
        self.ctx, self.n, self.dtype = ctx, int(n), np.dtype(dtype)
        self.nbytes = self.dtype.itemsize * self.n
        p = C.c_void_p()
        L.call("ilh_malloc", ctx.h, C.c_size_t(max(self.nbytes, 8)), C.byref(p))
        self.ptr = p
        _live["buffer"].add(self)

    @classmethod
    def from_host(cls, ctx, a, dtype=np.float64):
        a = np.ascontiguousarray(np.asarray(a, dtype=dtype).ravel(order="F"))
        d = cls(ctx, a.size, dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(np.asarray(a, dtype=self.dtype).ravel(order="F"))
        if a.size != self.n:
            raise L.DimensionMismatch("upload of %d values into %d" % (a.size, self.n))
        L.call("ilh_memcpy_h2d", self.ctx.h, self.ptr, a.ctypes.data_as(C.c_void_p), C.c_size_t(self.nbytes))

    def download(self):
        out = np.empty(self.n, dtype=self.dtype)
        L.call("ilh_memcpy_d2h", self.ctx.h, out.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(self.nbytes))
        return out

    def free(self):
        if self.ptr:
            L.load().ilh_free(self.ctx.h, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ------------------------------------------------------------ grid data
def _dt(data, dtype):
    """Element type of grid/surface data: FP64, or FP32 (a Float32 cache) when
    asked for or when the given data are float32."""
    if dtype is not None:
        return np.dtype(dtype)
    if data is not None and np.asarray(data).dtype == np.float32:
        return np.dtype(np.float32)
    return np.dtype(np.float64)


class Nodes:
    """Nodes{Primal|Dual} grid data: ``data`` is an (nx, ny) Fortran-ordered array."""

    def __init__(self, kind, g, data=None, dtype=None):
        self.kind = kind
        shape = g.primal_shape() if kind == Primal else g.dual_shape()
        dt = _dt(data, dtype)
        self.data = np.zeros(shape, dtype=dt, order="F") if data is None else np.asfortranarray(data, dtype=dt)
        if self.data.shape != shape:
            raise L.DimensionMismatch("Nodes{%s} of size %s, expected %s" % (kind, self.data.shape, shape))

    @property
    def ilh_kind(self):
        return L.NODES_PRIMAL if self.kind == Primal else L.NODES_DUAL

    def flat(self):
        return self.data.reshape(-1, order="F")


class Edges:
    """Edges{Primal} grid data: one buffer [u; v], u (NX, NY-1), v (NX-1, NY)."""
    ilh_kind = L.EDGES_PRIMAL

    def __init__(self, g, data=None, dtype=None):
        nu, nv = g.NX * (g.NY - 1), (g.NX - 1) * g.NY
        dt = _dt(data, dtype)
        self.buf = np.zeros(nu + nv, dtype=dt) if data is None else np.ascontiguousarray(data, dtype=dt)
        if self.buf.size != nu + nv:
            raise L.DimensionMismatch("Edges{Primal} of length %d, expected %d" % (self.buf.size, nu + nv))
        self.u = self.buf[:nu].reshape((g.NX, g.NY - 1), order="F")
        self.v = self.buf[nu:].reshape((g.NX - 1, g.NY), order="F")

    def flat(self):
        return self.buf


class ScalarData:
    def __init__(self, n, data=None, dtype=None):
        dt = _dt(data, dtype)
        self.data = np.zeros(n, dtype=dt) if data is None else np.ascontiguousarray(data, dtype=dt).copy()
        if self.data.size != n:
            raise L.DimensionMismatch("ScalarData of length %d, expected %d" % (self.data.size, n))

    def flat(self):
        return self.data


class VectorData:
    def __init__(self, n, data=None, dtype=None):
        dt = _dt(data, dtype)
        self.buf = np.zeros(2 * n, dtype=dt) if data is None else np.ascontiguousarray(data, dtype=dt).copy()
        if self.buf.size != 2 * n:
            raise L.DimensionMismatch("VectorData of length %d, expected %d" % (self.buf.size, 2 * n))
        self.u = self.buf[:n]
        self.v = self.buf[n:]

    def flat(self):
        return self.buf


class EdgeGradient:
    """EdgeGradient{Primal,Dual} tensor grid data: one buffer [dudx; dudy; dvdx; dvdy],
    dudx and dvdy on Nodes{Primal} ((NX-1, NY-1)), dudy and dvdx on Nodes{Dual} ((NX, NY))."""

    def __init__(self, g, data=None):
        npr, ndu = (g.NX - 1) * (g.NY - 1), g.NX * g.NY
        n = 2 * (npr + ndu)
        self.buf = np.zeros(n) if data is None else np.ascontiguousarray(data, dtype=np.float64)
        if self.buf.size != n:
            raise L.DimensionMismatch("EdgeGradient of length %d, expected %d" % (self.buf.size, n))
        cuts = np.cumsum([0, npr, ndu, ndu, npr])
        shapes = [(g.NX - 1, g.NY - 1), (g.NX, g.NY), (g.NX, g.NY), (g.NX - 1, g.NY - 1)]
        self.dudx, self.dudy, self.dvdx, self.dvdy = (
            self.buf[cuts[k]:cuts[k + 1]].reshape(shapes[k], order="F") for k in range(4))

    def flat(self):
        return self.buf


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


# -------------------------------------------------------------- the cache
class SurfaceScalarCache:
    """The reference's SurfaceScalarCache (src/cache.jl:164-182) backed by an
    ``ilh_plan``: grid, surface points, normals, areas, DDF and scaling."""

    def __init__(self, surface, g, ddftype=Yang3, scaling=GridScaling, ctx=None, dtype=np.float64):
        if scaling != GridScaling:
            raise L.ArgumentError("only GridScaling is supported")
        self.dtype = np.dtype(dtype)
        if self.dtype not in (np.dtype(np.float64), np.dtype(np.float32)):
            raise L.ArgumentError("dtype must be float64 or float32")
        self.ctx = ctx or default_context()
        self.g = g if isinstance(g, GridSpec) else GridSpec(*g)
        self.ddftype = ddftype
        self._set_points(surface)
        grid = L.Grid(self.g.NX, self.g.NY, self.g.I0x, self.g.I0y, self.g.dx)
        h = C.c_void_p()
        L.call("ilh_plan_create", self.ctx.h, C.byref(grid), self.N, _ptr(self._X), _ptr(self._Y),
               _ptr(self._nx), _ptr(self._ny), _ptr(self._ds), ddftype, 0, C.byref(h))
        self.h = h
        _live["cache"].add(self)
        self.body_offsets = _body_offsets(getattr(surface, "body", None), self.N)
        if len(self.body_offsets) > 2:
            off = np.ascontiguousarray(self.body_offsets, dtype=np.int32)
            L.call("ilh_plan_set_bodies", self.h, len(off) - 1, off.ctypes.data_as(C.c_void_p))

    def set_preconditioner(self, model):
        """Moving-body solves use the learned band of `model` (None: the FP32
        reference inverse)."""
        self._lpc = model
        L.call("ilh_plan_set_preconditioner", self.h, None if model is None else model.h)

    def _set_points(self, s):
        self._X = np.ascontiguousarray(s.X, dtype=np.float64)
        self._Y = np.ascontiguousarray(s.Y, dtype=np.float64)
        self._nx = np.ascontiguousarray(s.nx, dtype=np.float64)
        self._ny = np.ascontiguousarray(s.ny, dtype=np.float64)
        self._ds = np.ascontiguousarray(s.ds, dtype=np.float64)
        self.N = len(self._X)
        for a in (self._Y, self._nx, self._ny, self._ds):
            if len(a) != self.N:
                raise L.DimensionMismatch("geometry arrays of unequal length")

    def update_points(self, surface):
        """update_system! for moving bodies (src/system.jl:26-50): same N, new geometry."""
        n_old = self.N
        self._set_points(surface)
        if self.N != n_old:
            raise L.DimensionMismatch("point count changed from %d to %d" % (n_old, self.N))
        L.call("ilh_plan_update_points", self.h, _ptr(self._X), _ptr(self._Y), _ptr(self._nx), _ptr(self._ny),
               _ptr(self._ds))

    # reference accessors (src/tools.jl:16-38, src/cache.jl:399-595)
    def points(self):
        return VectorData(self.N, np.concatenate([self._X, self._Y]))

    def normals(self):
        return VectorData(self.N, np.concatenate([self._nx, self._ny]))

    def areas(self):
        return ScalarData(self.N, self._ds)

    def zeros_grid(self):
        return Nodes(Primal, self.g, dtype=self.dtype)

    def zeros_gridcurl(self):
        return Nodes(Dual, self.g, dtype=self.dtype)

    def zeros_gridgrad(self):
        return Edges(self.g, dtype=self.dtype)

    def zeros_surface(self):
        return ScalarData(self.N, dtype=self.dtype)

    def close(self):
        if getattr(self, "h", None):
            L.load().ilh_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _body_offsets(body, N):
    """Point offsets of the bodies from the per-point body labels (each body's
    points contiguous), [0, N] for one body."""
    if body is None or N == 0:
        return [0, N]
    b = np.asarray(body)
    starts = [0] + [i for i in range(1, N) if b[i] != b[i - 1]]
    if len(set(b[starts].tolist())) != len(starts):
        raise L.ArgumentError("each body's points must be contiguous")
    return starts + [N]


class LearnedPreconditioner:
    """The learned Schur preconditioner for moving bodies (ilh_lpc): a network
    predicting the band of S^-1 along each body's curve from the local
    geometry, trained on the GPU on exact inverses the plain method computes."""

    def __init__(self, ctx=None, half_band=16, seed=0):
        self.ctx = ctx or default_context()
        h = C.c_void_p()
        L.call("ilh_lpc_create", self.ctx.h, int(half_band), C.c_ulonglong(seed), C.byref(h))
        self.h = h
        d = (C.c_int * 7)()
        L.call("ilh_lpc_dims", self.h, d)
        self.dims = tuple(d)
        self.K, self.W, self.F, self.Fp, self.H, self.Op, self.nparam = self.dims
        _live["heat"].add(self)

    def add_geometry(self, cache):
        L.call("ilh_lpc_add_geometry", self.h, cache.h)

    @property
    def rows(self):
        return L.load().ilh_lpc_rows(self.h)

    def train(self, steps, batch=2048, lr=1e-3, seed=0):
        loss = C.c_double()
        L.call("ilh_lpc_train", self.h, int(steps), int(batch), float(lr), C.c_ulonglong(seed), C.byref(loss))
        return loss.value

    def params(self):
        p = np.zeros(self.nparam, dtype=np.float32)
        L.call("ilh_lpc_params", self.h, p.ctypes.data_as(C.c_void_p), 1)
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.nparam:
            raise L.DimensionMismatch("expected %d parameters" % self.nparam)
        L.call("ilh_lpc_params", self.h, p.ctypes.data_as(C.c_void_p), 0)

    def loss_grad(self, X_bits, T):
        """Loss and gradients of one standardized batch (X: rows x Fp BF16 bit
        patterns as uint16, T: rows x Op float32)."""
        X_bits = np.ascontiguousarray(X_bits, dtype=np.uint16)
        T = np.ascontiguousarray(T, dtype=np.float32)
        rows = X_bits.shape[0]
        if X_bits.shape != (rows, self.Fp) or T.shape != (rows, self.Op):
            raise L.DimensionMismatch("batch shapes")
        g = np.zeros(self.nparam, dtype=np.float32)
        loss = C.c_double()
        L.call("ilh_lpc_loss_grad", self.h, rows, X_bits.ctypes.data_as(C.c_void_p), T.ctypes.data_as(C.c_void_p),
               C.byref(loss), g.ctypes.data_as(C.c_void_p))
        return loss.value, g

    def predict(self, cache, band=True):
        """(features N x F, band N x (2K+1) or None) of the cache's geometry."""
        F = np.zeros((cache.N, self.F), dtype=np.float32)
        B = np.zeros((cache.N, self.W)) if band else None
        L.call("ilh_lpc_predict", self.h, cache.h, F.ctypes.data_as(C.c_void_p),
               None if B is None else B.ctypes.data_as(C.c_void_p))
        return F, B

    def standardization(self):
        out = np.zeros(2 * self.F + 2 * self.W, dtype=np.float32)
        L.call("ilh_lpc_standardization", self.h, out.ctypes.data_as(C.c_void_p))
        F, W = self.F, self.W
        return out[:F], out[F:2 * F], out[2 * F:2 * F + W], out[2 * F + W:]

    def close(self):
        if getattr(self, "h", None):
            L.load().ilh_lpc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SurfaceVectorCache(SurfaceScalarCache):
    """The reference's SurfaceVectorCache (src/cache.jl:184-201): the same plan,
    with VectorData as its surface data, Edges{Primal} as its grid data and
    EdgeGradient as its tensor grid data."""

    def zeros_grid(self):
        return Edges(self.g, dtype=self.dtype)

    def zeros_gridgrad(self):
        return EdgeGradient(self.g)

    def zeros_surface(self):
        return VectorData(self.N, dtype=self.dtype)

    def zeros_surfacescalar(self):
        return ScalarData(self.N, dtype=self.dtype)


# -------------------------------------------------------------- dispatch
def _check_len(a, n, what):
    if a.flat().size != n:
        raise L.DimensionMismatch("%s has %d values, expected %d" % (what, a.flat().size, n))


# operators with an FP32 entry point (Float32 caches)
_F32 = {"ilh_regularize", "ilh_interpolate", "ilh_divergence", "ilh_grad", "ilh_curl_nodes_to_edges",
        "ilh_curl_edges_to_nodes", "ilh_laplacian", "ilh_inverse_laplacian"}


def _run(name, cache, out, inp, *lead):
    o, i = out.flat(), inp.flat()
    if not o.flags.c_contiguous:
        raise L.ArgumentError("output must be contiguous")
    if o.dtype != i.dtype:
        raise L.ArgumentError("input and output element types differ (%s, %s)" % (i.dtype, o.dtype))
    if o.dtype == np.float32:
        if name not in _F32:
            raise L.ArgumentError("%s has no FP32 form" % name)
        name += "_f32"
    elif o.dtype != np.float64:
        raise L.ArgumentError("grid and surface data must be float64 or float32")
    i = np.ascontiguousarray(i)
    L.call(name, cache.h, *lead, _ptr(o), _ptr(i), L.MEM_HOST)
    return out


def regularize(out, f, cache):
    """s = Rc f (Nodes) or v = Rf vb (Edges{Primal} from VectorData)."""
    if isinstance(out, Nodes) and isinstance(f, ScalarData):
        _check_len(f, cache.N, "ScalarData")
        return _run("ilh_regularize", cache, out, f, out.ilh_kind)
    if isinstance(out, Edges) and isinstance(f, VectorData):
        _check_len(f, 2 * cache.N, "VectorData")
        return _run("ilh_regularize", cache, out, f, L.EDGES_PRIMAL)
    raise TypeError("regularize: no method for (%s, %s)" % (type(out).__name__, type(f).__name__))


def interpolate(out, field, cache):
    """f = Rc^T s (ScalarData from Nodes) or vb = Rf^T v (VectorData from Edges{Primal})."""
    if isinstance(out, ScalarData) and isinstance(field, Nodes):
        _check_len(out, cache.N, "ScalarData")
        return _run("ilh_interpolate", cache, out, field, field.ilh_kind)
    if isinstance(out, VectorData) and isinstance(field, Edges):
        _check_len(out, 2 * cache.N, "VectorData")
        return _run("ilh_interpolate", cache, out, field, L.EDGES_PRIMAL)
    raise TypeError("interpolate: no method for (%s, %s)" % (type(out).__name__, type(field).__name__))


def regularize_normal(q, f, cache):
    _check_len(f, cache.N, "ScalarData")
    return _run("ilh_regularize_normal", cache, q, f)


def normal_interpolate(vn, q, cache):
    _check_len(vn, cache.N, "ScalarData")
    return _run("ilh_normal_interpolate", cache, vn, q)


def regularize_normal_cross(q, f, cache):
    _check_len(f, cache.N, "ScalarData")
    return _run("ilh_regularize_normal_cross", cache, q, f)


def normal_cross_interpolate(vn, q, cache):
    _check_len(vn, cache.N, "ScalarData")
    return _run("ilh_normal_cross_interpolate", cache, vn, q)


def surface_divergence(theta, f, cache):
    _check_len(f, cache.N, "ScalarData")
    return _run("ilh_surface_divergence", cache, theta, f)


def surface_grad(vn, phi, cache):
    _check_len(vn, cache.N, "ScalarData")
    return _run("ilh_surface_grad", cache, vn, phi)


def surface_curl(out, inp, cache):
    """w = C_s^T f (Nodes{Dual} from ScalarData or VectorData) or C_s s (ScalarData
    or VectorData from Nodes{Dual}) (src/surface_operators.jl:357-455)."""
    if isinstance(out, Nodes) and out.kind == Dual and isinstance(inp, VectorData):
        _check_len(inp, 2 * cache.N, "VectorData")
        return _run("ilh_surface_curl_vector", cache, out, inp)
    if isinstance(out, VectorData) and isinstance(inp, Nodes) and inp.kind == Dual:
        _check_len(out, 2 * cache.N, "VectorData")
        return _run("ilh_surface_curl_vector_adj", cache, out, inp)
    if isinstance(out, Nodes) and out.kind == Dual:
        return _run("ilh_surface_curl", cache, out, inp)
    if isinstance(out, ScalarData) and isinstance(inp, Nodes) and inp.kind == Dual:
        return _run("ilh_surface_curl_adj", cache, out, inp)
    raise TypeError("surface_curl: no method for (%s, %s)" % (type(out).__name__, type(inp).__name__))


def surface_divergence_symm(q, dv, cache):
    """q = D R_t (n o dv + dv o n - (n . dv) I) on Edges{Primal} (src/surface_operators.jl:574-580)."""
    _check_len(dv, 2 * cache.N, "VectorData")
    _check_len(q, cache.g.NX * (cache.g.NY - 1) + (cache.g.NX - 1) * cache.g.NY, "Edges{Primal}")
    return _run("ilh_surface_divergence_symm", cache, q, dv)


def surface_grad_symm(tau, v, cache):
    """tau = n . R_t^T (G v + (G v)^T - tr(G v) I) (src/surface_operators.jl:641-644)."""
    _check_len(tau, 2 * cache.N, "VectorData")
    return _run("ilh_surface_grad_symm", cache, tau, v)


def regularize_normal_symm(A, v, cache):
    """A = R_t (n o v + v o n - (n . v) I) on EdgeGradient (src/surface_operators.jl:121)."""
    _check_len(v, 2 * cache.N, "VectorData")
    return _run("ilh_regularize_normal_symm", cache, A, v)


def normal_interpolate_symm(tau, A, cache):
    """tau = n . R_t^T (A + A^T - tr(A) I) (src/surface_operators.jl:245)."""
    _check_len(tau, 2 * cache.N, "VectorData")
    return _run("ilh_normal_interpolate_symm", cache, tau, A)


def surface_curl_cross(out, inp, cache):
    if isinstance(out, Nodes) and out.kind == Dual:
        return _run("ilh_surface_curl_cross", cache, out, inp)
    if isinstance(out, ScalarData) and isinstance(inp, Nodes) and inp.kind == Dual:
        return _run("ilh_surface_curl_cross_adj", cache, out, inp)
    raise TypeError("surface_curl_cross: no method for (%s, %s)" % (type(out).__name__, type(inp).__name__))


def surface_grad_cross(vn, phi, cache):
    _check_len(vn, cache.N, "ScalarData")
    return _run("ilh_surface_grad_cross", cache, vn, phi)


def divergence(p, q, cache):
    return _run("ilh_divergence", cache, p, q)


def grad(q, p, cache):
    return _run("ilh_grad", cache, q, p)


def curl(out, inp, cache):
    if isinstance(out, Edges) and isinstance(inp, Nodes) and inp.kind == Dual:
        return _run("ilh_curl_nodes_to_edges", cache, out, inp)
    if isinstance(out, Nodes) and out.kind == Dual and isinstance(inp, Edges):
        return _run("ilh_curl_edges_to_nodes", cache, out, inp)
    raise TypeError("curl: no method for (%s, %s)" % (type(out).__name__, type(inp).__name__))


def laplacian(w, s, cache):
    if w.kind != s.kind:
        raise TypeError("laplacian: grid types differ")
    return _run("ilh_laplacian", cache, w, s, w.ilh_kind)


def inverse_laplacian(w, cache, rhs=None):
    """inverse_laplacian!(w, cache) in place, or inverse_laplacian!(sol, rhs, cache) when rhs is given."""
    src = w if rhs is None else rhs
    return _run("ilh_inverse_laplacian", cache, w, src, w.ilh_kind)


def _grid_kind_of(w):
    if isinstance(w, Edges):
        return L.EDGES_PRIMAL
    if isinstance(w, Nodes):
        return w.ilh_kind
    raise TypeError("masks act on Nodes or Edges{Primal}")


def _blank_like(cache, like):
    if like is None:
        return Nodes(Primal, cache.g)
    return Nodes(like.kind, cache.g) if isinstance(like, Nodes) else Edges(cache.g)


def mask(cache, like=None):
    """mask(cache): 1 inside the surfaces (opposite the normals), 0 outside, as
    Nodes{Primal}, or on the grid kind of `like` (src/surface_operators.jl:737)."""
    w = _blank_like(cache, like)
    L.call("ilh_mask", cache.h, _grid_kind_of(w), 0, _ptr(w.flat()), L.MEM_HOST)
    return w


def complementary_mask(cache, like=None):
    """complementary_mask(cache) = 1 - mask(cache) (src/surface_operators.jl:749)."""
    w = _blank_like(cache, like)
    L.call("ilh_mask", cache.h, _grid_kind_of(w), 1, _ptr(w.flat()), L.MEM_HOST)
    return w


def mask_(w, cache):
    """mask!(w, cache): w times the mask, in place (src/surface_operators.jl:785)."""
    L.call("ilh_mask_multiply", cache.h, _grid_kind_of(w), 0, _ptr(w.flat()), L.MEM_HOST)
    return w


def complementary_mask_(w, cache):
    """complementary_mask!(w, cache): w times 1 - mask, in place."""
    L.call("ilh_mask_multiply", cache.h, _grid_kind_of(w), 1, _ptr(w.flat()), L.MEM_HOST)
    return w


def create_surface_filter(cache):
    """C = Et R (src/matrix_operators.jl:254-267) as an N x N array (2N x 2N,
    block diagonal, for a SurfaceVectorCache): the interpolation divides the
    regularized field by R 1 first, so C keeps constants and the surface integral."""
    n = 2 * cache.N if isinstance(cache, SurfaceVectorCache) else cache.N
    Cm = np.zeros(n * n)
    name = "ilh_surface_filter_build_vector" if isinstance(cache, SurfaceVectorCache) else "ilh_surface_filter_build"
    L.call(name, cache.h, _ptr(Cm), L.MEM_HOST)
    return Cm.reshape((n, n), order="F")


def filter_surface(s, cache, power=1):
    """C^power s without forming C (the C^5 s / C^6 s of test/literate/dirichlet.jl:122-124,
    the C^2 sigma of test/literate/stokes.jl:162 for VectorData)."""
    if isinstance(s, VectorData):
        x = np.ascontiguousarray(s.flat(), dtype=np.float64)
        _check_len(s, 2 * cache.N, "VectorData")
        out = VectorData(cache.N)
        L.call("ilh_surface_filter_apply_vector", cache.h, int(power), _ptr(x), _ptr(out.flat()), L.MEM_HOST)
        return out
    x = np.ascontiguousarray(s.flat() if isinstance(s, ScalarData) else s, dtype=np.float64)
    out = ScalarData(cache.N)
    L.call("ilh_surface_filter_apply", cache.h, int(power), _ptr(x), _ptr(out.flat()), L.MEM_HOST)
    return out


def _schur_dim(cache, kind):
    return 2 * cache.N if kind in (L.SCHUR_CL2INVCT, L.SCHUR_GLINVD_SYMM) else cache.N


def _schur(cache, kind, scale):
    n = _schur_dim(cache, kind)
    A = np.zeros(n * n)
    L.call("ilh_schur_build", cache.h, kind, C.c_double(scale), _ptr(A), L.MEM_HOST)
    return A.reshape((n, n), order="F")


def create_RTLinvR(cache, scale=1.0):
    return _schur(cache, L.SCHUR_RTLINVR, scale)


def create_CLinvCT(cache, scale=1.0):
    return _schur(cache, L.SCHUR_CLINVCT, scale)


def create_GLinvD(cache, scale=1.0):
    return _schur(cache, L.SCHUR_GLINVD, scale)


def create_nRTRn(cache, scale=1.0):
    return _schur(cache, L.SCHUR_NRTRN, scale)


def create_CL2invCT(cache, scale=1.0):
    """-C_s L^-2 C_s^T on VectorData (src/matrix_operators.jl:95-125), 2N x 2N."""
    return _schur(cache, L.SCHUR_CL2INVCT, scale)


def create_GLinvD_symm(cache, scale=1.0):
    """-G_s^symm L^-1 D_s^symm on VectorData (src/matrix_operators.jl:165-185), 2N x 2N."""
    return _schur(cache, L.SCHUR_GLINVD_SYMM, scale)


def create_CLinvCT_scalar(cache, scale=1.0):
    """-C_s L^-1 C_s^T on the vector cache's ScalarData (src/matrix_operators.jl:71-91)."""
    return _schur(cache, L.SCHUR_CLINVCT_SCALAR, scale)


# -------------------------------------------------------------- problems
def _vec(a, n):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel())
    if a.size != n:
        raise L.DimensionMismatch("surface data of length %d, expected %d" % (a.size, n))
    return a


class DirichletPoisson:
    """The Dirichlet Poisson problem of test/literate/problems.jl:243-298:
    S = create_RTLinvR is built and factored once per geometry; ``solve``
    returns the grid field f and the surface strength s, filtered as C^6 s
    by default like the reference's solve."""

    def __init__(self, cache):
        self.cache = cache

    def solve(self, fbplus, fbminus=None, filter_power=6):
        """(f, s) with s filtered as C^k s, k = filter_power (6, as the
        reference's solve returns C^6 s, test/literate/problems.jl:298;
        0: the unfiltered strength, which the device-buffer, batched and
        moving-body forms return)."""
        c = self.cache
        fp = _vec(fbplus, c.N)
        fm = None if fbminus is None else _vec(fbminus, c.N)
        f = Nodes(Primal, c.g)
        s = ScalarData(c.N)
        L.call("ilh_dirichlet_solve", c.h, _ptr(fp), None if fm is None else _ptr(fm), _ptr(f.flat()),
               _ptr(s.flat()), L.MEM_HOST)
        if filter_power:
            s = filter_surface(s, c, filter_power)
        return f, s

    def solve_mixed(self, fbplus, fbminus=None):
        """(f, s) with the convolution that forms f in FP32 and s in FP64
        (ilh_dirichlet_solve_mixed): f to the FP32 rounding of the convolution."""
        c = self.cache
        fp = _vec(fbplus, c.N)
        fm = None if fbminus is None else _vec(fbminus, c.N)
        f = Nodes(Primal, c.g)
        s = ScalarData(c.N)
        L.call("ilh_dirichlet_solve_mixed", c.h, _ptr(fp), None if fm is None else _ptr(fm), _ptr(f.flat()),
               _ptr(s.flat()), L.MEM_HOST)
        return f, s

    def solve_mixed_device(self, fp_dev, fm_dev, f_dev, s_dev):
        """Device-resident form of solve_mixed (asynchronous)."""
        L.call("ilh_dirichlet_solve_mixed", self.cache.h, fp_dev.ptr, None if fm_dev is None else fm_dev.ptr,
               f_dev.ptr, s_dev.ptr, L.MEM_DEVICE)

    def solve_multi(self, fbplus, fbminus=None):
        """Many boundary data on this geometry at once: fbplus (and fbminus)
        N x K; returns f (K Nodes{Primal}) and s (N x K array)."""
        c = self.cache
        fp = np.asarray(fbplus, dtype=np.float64)
        if fp.ndim == 1:
            fp = fp.reshape(-1, 1)
        if fp.ndim != 2 or fp.shape[0] != c.N:
            raise L.DimensionMismatch("boundary data of shape %s, expected (%d, K)" % (fp.shape, c.N))
        fp = np.asfortranarray(fp)
        k = fp.shape[1]
        fm = None
        if fbminus is not None:
            fm = np.asfortranarray(np.asarray(fbminus, dtype=np.float64).reshape(fp.shape))
        nx, ny = c.g.primal_shape()
        f = np.zeros((nx * ny, k), order="F")
        s = np.zeros((c.N, k), order="F")
        L.call("ilh_dirichlet_solve_multi", c.h, k, _ptr(fp), None if fm is None else _ptr(fm), _ptr(f), _ptr(s),
               L.MEM_HOST)
        return [Nodes(Primal, c.g, f[:, i].reshape((nx, ny), order="F")) for i in range(k)], s

    def solve_multi_device(self, nrhs, fp_dev, fm_dev, f_dev, s_dev):
        """Device-buffer form of solve_multi (asynchronous)."""
        L.call("ilh_dirichlet_solve_multi", self.cache.h, int(nrhs), fp_dev.ptr,
               None if fm_dev is None else fm_dev.ptr, f_dev.ptr, s_dev.ptr, L.MEM_DEVICE)

    def solve_moving(self, fbplus, fbminus=None, tol=1e-12, max_iter=30):
        """The same problem after cache.update_points (moving bodies): GMRES on
        the current Schur system preconditioned by a reference geometry's
        inverse.  Returns (f, s, gmres_iterations)."""
        c = self.cache
        fp = _vec(fbplus, c.N)
        fm = None if fbminus is None else _vec(fbminus, c.N)
        f = Nodes(Primal, c.g)
        s = ScalarData(c.N)
        it = C.c_int()
        L.call("ilh_dirichlet_solve_moving", c.h, _ptr(fp), None if fm is None else _ptr(fm), _ptr(f.flat()),
               _ptr(s.flat()), L.MEM_HOST, C.c_double(tol), int(max_iter), C.byref(it))
        return f, s, it.value

    def solve_moving_device(self, fp_dev, fm_dev, f_dev, s_dev, tol=1e-12, max_iter=30):
        """Device-buffer form of solve_moving; returns the GMRES iterations."""
        it = C.c_int()
        L.call("ilh_dirichlet_solve_moving", self.cache.h, fp_dev.ptr, None if fm_dev is None else fm_dev.ptr,
               f_dev.ptr, s_dev.ptr, L.MEM_DEVICE, C.c_double(tol), int(max_iter), C.byref(it))
        return it.value

    def solve_device(self, fp_dev, fm_dev, f_dev, s_dev):
        """Device-resident form: all four arguments are DeviceArrays (fm may be None); asynchronous."""
        L.call("ilh_dirichlet_solve", self.cache.h, fp_dev.ptr, None if fm_dev is None else fm_dev.ptr,
               f_dev.ptr, s_dev.ptr, L.MEM_DEVICE)

    def profile(self, fp_dev, fm_dev, f_dev, s_dev, reps, stages=True):
        """Device-buffer repeat solves timed with HIP events: returns
        (per-stage ms of each stage launched alone, or None; ms per whole solve)."""
        st = (C.c_double * 4)() if stages else None
        tot = C.c_double()
        L.call("ilh_profile_dirichlet", self.cache.h, fp_dev.ptr, None if fm_dev is None else fm_dev.ptr,
               f_dev.ptr, s_dev.ptr, int(reps), st, C.byref(tot))
        return (list(st) if stages else None), tot.value


def dirichlet_solve_batched(problems, fbplus, fbminus=None):
    """Batched form of DirichletPoisson.solve over independent caches of one
    context (ilh_dirichlet_solve_batched): returns [(f, s), ...]."""
    n = len(problems)
    caches = [p.cache if isinstance(p, DirichletPoisson) else p for p in problems]
    fps = [_vec(fbplus[k], caches[k].N) for k in range(n)]
    fms = None if fbminus is None else [None if fbminus[k] is None else _vec(fbminus[k], caches[k].N)
                                        for k in range(n)]
    fs = [Nodes(Primal, c.g) for c in caches]
    ss = [ScalarData(c.N) for c in caches]
    arr = lambda xs: (C.c_void_p * n)(*[None if x is None else x.ctypes.data for x in xs])
    plans = (C.c_void_p * n)(*[c.h.value for c in caches])
    L.call("ilh_dirichlet_solve_batched", n, plans, arr(fps), None if fms is None else arr(fms),
           arr([f.flat() for f in fs]), arr([s.flat() for s in ss]), L.MEM_HOST)
    return list(zip(fs, ss))


def dirichlet_solve_batched_device(caches, fp_devs, f_devs, s_devs):
    """Device-buffer form (DeviceArrays, asynchronous on the context stream)."""
    n = len(caches)
    ptrs = lambda xs: (C.c_void_p * n)(*[x.ptr.value for x in xs])
    plans = (C.c_void_p * n)(*[c.h.value for c in caches])
    L.call("ilh_dirichlet_solve_batched", n, plans, ptrs(fp_devs), None, ptrs(f_devs), ptrs(s_devs), L.MEM_DEVICE)


def dirichlet_solve_moving_batched(caches, fp_devs, f_devs, s_devs, fm_devs=None, tol=1e-12, max_iter=30):
    """Moving-body solves of many caches side by side (each after its own
    update_points), device buffers only; returns the GMRES iterations per cache."""
    n = len(caches)
    ptrs = lambda xs: (C.c_void_p * n)(*[None if x is None else x.ptr.value for x in xs])
    plans = (C.c_void_p * n)(*[c.h.value for c in caches])
    its = (C.c_int * n)()
    L.call("ilh_dirichlet_solve_moving_batched", n, plans, ptrs(fp_devs), None if fm_devs is None else ptrs(fm_devs),
           ptrs(f_devs), ptrs(s_devs), C.c_double(tol), int(max_iter), its)
    return list(its)


class DirichletBatch:
    """A batch of independent Dirichlet problems on fixed device buffers,
    captured as one HIP graph and replayed with `launch()`."""

    def __init__(self, caches, fp_devs, f_devs, s_devs, fm_devs=None):
        n = len(caches)
        ptrs = lambda xs: (C.c_void_p * n)(*[None if x is None else x.ptr.value for x in xs])
        plans = (C.c_void_p * n)(*[c.h.value for c in caches])
        h = C.c_void_p()
        L.call("ilh_batch_create", n, plans, ptrs(fp_devs), None if fm_devs is None else ptrs(fm_devs),
               ptrs(f_devs), ptrs(s_devs), C.byref(h))
        self.h, self._keep = h, (caches, fp_devs, f_devs, s_devs, fm_devs)
        _live["batch"].add(self)

    def launch(self):
        L.call("ilh_batch_launch", self.h)

    PASSES = ("surface", "rows", "columns", "inverse_rows")

    def tuning(self):
        """The autotuner's choice per pass ({schedule, workgroups_per_cu,
        threads, ms}) and the row pass's amplitude LDS cap."""
        cfg = (C.c_int * 13)()
        ms = (C.c_float * 4)()
        L.call("ilh_batch_tuning", self.h, cfg, ms)
        out = {}
        for k, name in enumerate(self.PASSES):
            out[name] = {"schedule": ("dispatch", "queue")[cfg[3 * k]], "workgroups_per_cu": cfg[3 * k + 1],
                         "threads": cfg[3 * k + 2], "ms": float(ms[k])}
        out["amp_lds_cap"] = cfg[12]
        return out

    def close(self):
        if self.h:
            L.load().ilh_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DirichletHeatConduction:
    """Unbounded heat conduction with Dirichlet bodies (test/literate/heatconduction.jl)
    marched by IF-HERK: ``advance(T, nsteps)`` updates the Nodes{Primal}
    temperature in place; boundary values are time-independent."""

    def __init__(self, cache, kappa, dt, tb_plus, tb_minus=None):
        h = C.c_void_p()
        L.call("ilh_heat_create", cache.h, C.c_double(kappa), C.c_double(dt), C.byref(h))
        self.h, self.cache, self.kappa, self.dt = h, cache, kappa, dt
        _live["heat"].add(self)
        self.set_boundary(tb_plus, tb_minus)

    def set_boundary(self, tb_plus, tb_minus=None):
        tp = _vec(tb_plus, self.cache.N)
        tm = None if tb_minus is None else _vec(tb_minus, self.cache.N)
        L.call("ilh_heat_set_boundary", self.h, _ptr(tp), None if tm is None else _ptr(tm), L.MEM_HOST)

    def advance(self, T, nsteps):
        if not isinstance(T, Nodes) or T.kind != Primal:
            raise TypeError("T must be Nodes{Primal}")
        L.call("ilh_heat_advance", self.h, _ptr(T.flat()), int(nsteps), L.MEM_HOST)
        return T

    def advance_device(self, T_dev, nsteps):
        L.call("ilh_heat_advance", self.h, T_dev.ptr, int(nsteps), L.MEM_DEVICE)

    def close(self):
        if self.h:
            L.load().ilh_heat_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class NeumannPoisson:
    """The Neumann Poisson problem of test/literate/neumann.jl:80-142: returns
    (f, [ϕ], s, [ψ]) with S = create_CLinvCT factored once per geometry."""

    def __init__(self, cache):
        self.cache = cache

    def solve(self, vnplus, vnminus=None):
        c = self.cache
        vp = _vec(vnplus, c.N)
        vm = None if vnminus is None else _vec(vnminus, c.N)
        f = Nodes(Primal, c.g)
        s = Nodes(Dual, c.g)
        df = ScalarData(c.N)
        ds = ScalarData(c.N)
        L.call("ilh_neumann_solve", c.h, _ptr(vp), None if vm is None else _ptr(vm), _ptr(f.flat()),
               _ptr(df.flat()), _ptr(s.flat()), _ptr(ds.flat()), L.MEM_HOST)
        return f, df, s, ds


class SchurFactor:
    """Cached LU factor of one Schur complement kind; ``solve(b)`` = S \\ b."""

    def __init__(self, cache, kind=L.SCHUR_RTLINVR):
        self.cache, self.kind = cache, kind
        L.call("ilh_schur_factor", cache.h, kind)

    def solve(self, b):
        b = _vec(b, _schur_dim(self.cache, self.kind))
        x = np.zeros_like(b)
        L.call("ilh_schur_solve", self.cache.h, self.kind, _ptr(x), _ptr(b), L.MEM_HOST)
        return x


class StokesFlow:
    """The Stokes flow problem of test/literate/stokes.jl:44-166 on a
    SurfaceVectorCache: create_CL2invCT and create_CLinvCT_scalar are built and
    inverted once per geometry (prob_cache); ``solve`` returns the velocity
    (Edges{Primal}), the streamfunction (Nodes{Dual}) and the surface traction
    (VectorData) filtered ``filter_power`` times (2, as the reference)."""

    def __init__(self, cache):
        if not isinstance(cache, SurfaceVectorCache):
            raise TypeError("StokesFlow needs a SurfaceVectorCache")
        self.cache = cache

    def solve(self, vplus, vminus=None, filter_power=2):
        c = self.cache
        vp = _vec(vplus.flat() if isinstance(vplus, VectorData) else vplus, 2 * c.N)
        vm = None if vminus is None else _vec(vminus.flat() if isinstance(vminus, VectorData) else vminus, 2 * c.N)
        v, s, sigma = Edges(c.g), Nodes(Dual, c.g), VectorData(c.N)
        L.call("ilh_stokes_solve", c.h, _ptr(vp), None if vm is None else _ptr(vm), _ptr(v.flat()), _ptr(s.flat()),
               _ptr(sigma.flat()), int(filter_power), L.MEM_HOST)
        return v, s, sigma
