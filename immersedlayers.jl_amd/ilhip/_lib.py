"""ctypes binding of libimmersedhip.so (the C ABI of include/immersedhip.h).

The library is the product: every compute call below goes to a HIP kernel on
the GPU.  There is no CPU fallback; loading fails loudly when the library has
not been built.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.normpath(os.path.join(_HERE, "..", "lib", "libimmersedhip.so"))

OK = 0
ERR_ARG, ERR_DIM, ERR_NONFINITE, ERR_SINGULAR, ERR_HIP, ERR_OOM, ERR_RANGE, ERR_STATE = -1, -2, -3, -4, -5, -6, -7, -8
MEM_HOST, MEM_DEVICE = 0, 1
NODES_PRIMAL, NODES_DUAL, EDGES_PRIMAL = 0, 1, 2
SCHUR_RTLINVR, SCHUR_CLINVCT, SCHUR_GLINVD, SCHUR_NRTRN, SCHUR_CL2INVCT, SCHUR_GLINVD_SYMM = 0, 1, 2, 3, 4, 5
SCHUR_CLINVCT_SCALAR = SCHUR_CLINVCT
DDF_YANG3, DDF_ROMA, DDF_M4PRIME, DDF_WITCHHAT = 0, 1, 2, 3


class ILHError(RuntimeError):
    code = None


class DimensionMismatch(ILHError, ValueError):
    """Size mismatch (the reference throws DimensionMismatch)."""


class ArgumentError(ILHError, ValueError):
    pass


class NonFiniteError(ILHError, ValueError):
    pass


class SingularError(ILHError):
    pass


class HIPError(ILHError):
    pass


class RangeError(ILHError, ValueError):
    pass


class StateError(ILHError):
    pass


_ERRORS = {ERR_ARG: ArgumentError, ERR_DIM: DimensionMismatch, ERR_NONFINITE: NonFiniteError,
           ERR_SINGULAR: SingularError, ERR_HIP: HIPError, ERR_OOM: HIPError, ERR_RANGE: RangeError,
           ERR_STATE: StateError}

# (name, argtypes) of every exported entry point, in header order
_vp, _dp, _ip, _i, _d, _sz = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_int, C.c_double, C.c_size_t


class Grid(C.Structure):
    _fields_ = [("NX", C.c_int), ("NY", C.c_int), ("I0x", C.c_int), ("I0y", C.c_int), ("dx", C.c_double)]


SIGNATURES = {
    "ilh_last_error": ([], C.c_char_p),
    "ilh_version": ([], C.c_char_p),
    "ilh_ctx_create": ([_i, C.POINTER(_vp)], _i),
    "ilh_ctx_destroy": ([_vp], _i),
    "ilh_ctx_synchronize": ([_vp], _i),
    "ilh_malloc": ([_vp, _sz, C.POINTER(_vp)], _i),
    "ilh_free": ([_vp, _vp], _i),
    "ilh_memcpy_h2d": ([_vp, _vp, _vp, _sz], _i),
    "ilh_memcpy_d2h": ([_vp, _vp, _vp, _sz], _i),
    "ilh_plan_create": ([_vp, C.POINTER(Grid), _i, _vp, _vp, _vp, _vp, _vp, _i, _i, C.POINTER(_vp)], _i),
    "ilh_plan_update_points": ([_vp, _vp, _vp, _vp, _vp, _vp], _i),
    "ilh_plan_destroy": ([_vp], _i),
    "ilh_plan_npts": ([_vp], _i),
    "ilh_regularize": ([_vp, _i, _vp, _vp, _i], _i),
    "ilh_interpolate": ([_vp, _i, _vp, _vp, _i], _i),
    "ilh_regularize_normal": ([_vp, _vp, _vp, _i], _i),
    "ilh_normal_interpolate": ([_vp, _vp, _vp, _i], _i),
    "ilh_regularize_normal_cross": ([_vp, _vp, _vp, _i], _i),
    "ilh_normal_cross_interpolate": ([_vp, _vp, _vp, _i], _i),
    "ilh_surface_divergence": ([_vp, _vp, _vp, _i], _i),
    "ilh_surface_grad": ([_vp, _vp, _vp, _i], _i),
    "ilh_surface_curl": ([_vp, _vp, _vp, _i], _i),
    "ilh_surface_curl_adj": ([_vp, _vp, _vp, _i], _i),
    "ilh_surface_curl_cross": ([_vp, _vp, _vp, _i], _i),
    "ilh_surface_curl_cross_adj": ([_vp, _vp, _vp, _i], _i),
    "ilh_surface_grad_cross": ([_vp, _vp, _vp, _i], _i),
    "ilh_regularize_f32": ([_vp, _i, _vp, _vp, _i], _i),
    "ilh_interpolate_f32": ([_vp, _i, _vp, _vp, _i], _i),
    "ilh_divergence_f32": ([_vp, _vp, _vp, _i], _i),
    "ilh_grad_f32": ([_vp, _vp, _vp, _i], _i),
    "ilh_curl_nodes_to_edges_f32": ([_vp, _vp, _vp, _i], _i),
    "ilh_curl_edges_to_nodes_f32": ([_vp, _vp, _vp, _i], _i),
    "ilh_laplacian_f32": ([_vp, _i, _vp, _vp, _i], _i),
    "ilh_inverse_laplacian_f32": ([_vp, _i, _vp, _vp, _i], _i),
    "ilh_surface_curl_vector": ([_vp, _vp, _vp, _i], _i),
    "ilh_surface_curl_vector_adj": ([_vp, _vp, _vp, _i], _i),
    "ilh_surface_divergence_symm": ([_vp, _vp, _vp, _i], _i),
    "ilh_surface_grad_symm": ([_vp, _vp, _vp, _i], _i),
    "ilh_regularize_normal_symm": ([_vp, _vp, _vp, _i], _i),
    "ilh_normal_interpolate_symm": ([_vp, _vp, _vp, _i], _i),
    "ilh_divergence": ([_vp, _vp, _vp, _i], _i),
    "ilh_grad": ([_vp, _vp, _vp, _i], _i),
    "ilh_curl_nodes_to_edges": ([_vp, _vp, _vp, _i], _i),
    "ilh_curl_edges_to_nodes": ([_vp, _vp, _vp, _i], _i),
    "ilh_laplacian": ([_vp, _i, _vp, _vp, _i], _i),
    "ilh_inverse_laplacian": ([_vp, _i, _vp, _vp, _i], _i),
    "ilh_schur_build": ([_vp, _i, _d, _vp, _i], _i),
    "ilh_schur_factor": ([_vp, _i], _i),
    "ilh_schur_solve": ([_vp, _i, _vp, _vp, _i], _i),
    "ilh_dirichlet_solve": ([_vp, _vp, _vp, _vp, _vp, _i], _i),
    "ilh_neumann_solve": ([_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i], _i),
    "ilh_profile_dirichlet": ([_vp, _vp, _vp, _vp, _vp, _i, _dp, _dp], _i),
    "ilh_dirichlet_solve_moving": ([_vp, _vp, _vp, _vp, _vp, _i, C.c_double, _i, _vp], _i),
    "ilh_dirichlet_solve_multi": ([_vp, _i, _vp, _vp, _vp, _vp, _i], _i),
    "ilh_dirichlet_solve_noalloc": ([_vp, _vp, _vp, _vp, _vp], _i),
    "ilh_dirichlet_solve_mixed": ([_vp, _vp, _vp, _vp, _vp, _i], _i),
    "ilh_dirichlet_solve_moving_batched": ([_i, _vp, _vp, _vp, _vp, _vp, C.c_double, _i, _vp], _i),
    "ilh_dirichlet_solve_batched": ([_i, _vp, _vp, _vp, _vp, _vp, _i], _i),
    "ilh_batch_create": ([_i, _vp, _vp, _vp, _vp, _vp, _vp], _i),
    "ilh_batch_launch": ([_vp], _i),
    "ilh_batch_tuning": ([_vp, _vp, _vp], _i),
    "ilh_batch_destroy": ([_vp], _i),
    "ilh_plan_set_bodies": ([_vp, _i, _vp], _i),
    "ilh_lpc_create": ([_vp, _i, C.c_ulonglong, C.POINTER(_vp)], _i),
    "ilh_lpc_destroy": ([_vp], _i),
    "ilh_lpc_add_geometry": ([_vp, _vp], _i),
    "ilh_lpc_rows": ([_vp], _i),
    "ilh_lpc_train": ([_vp, _i, _i, _d, C.c_ulonglong, _dp], _i),
    "ilh_lpc_dims": ([_vp, _ip], _i),
    "ilh_lpc_params": ([_vp, _vp, _i], _i),
    "ilh_lpc_loss_grad": ([_vp, _i, _vp, _vp, _dp, _vp], _i),
    "ilh_lpc_predict": ([_vp, _vp, _vp, _vp], _i),
    "ilh_lpc_standardization": ([_vp, _vp], _i),
    "ilh_plan_set_preconditioner": ([_vp, _vp], _i),
    "ilh_mask": ([_vp, _i, _i, _vp, _i], _i),
    "ilh_surface_filter_build": ([_vp, _vp, _i], _i),
    "ilh_surface_filter_apply": ([_vp, _i, _vp, _vp, _i], _i),
    "ilh_surface_filter_build_vector": ([_vp, _vp, _i], _i),
    "ilh_surface_filter_apply_vector": ([_vp, _i, _vp, _vp, _i], _i),
    "ilh_stokes_solve": ([_vp, _vp, _vp, _vp, _vp, _vp, _i, _i], _i),
    "ilh_mask_multiply": ([_vp, _i, _i, _vp, _i], _i),
    "ilh_heat_create": ([_vp, C.c_double, C.c_double, _vp], _i),
    "ilh_heat_set_boundary": ([_vp, _vp, _vp, _i], _i),
    "ilh_heat_advance": ([_vp, _vp, _i, _i], _i),
    "ilh_heat_destroy": ([_vp], _i),
}

_lib = None


def load(path=LIB_PATH):
    """Load libimmersedhip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError("libimmersedhip.so is not built (%s); run __graft_entry__.build()" % path)
    lib = C.CDLL(path)
    for name, (args, res) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def check(rc):
    if rc != OK:
        msg = load().ilh_last_error().decode(errors="replace")
        exc = _ERRORS.get(rc, ILHError)(msg)
        exc.code = rc
        raise exc
    return rc


def call(name, *args):
    return check(getattr(load(), name)(*args))
