"""ilhip -- MI355X-native immersed-layer hot path (HIP/gfx950) with the
operator API of ImmersedLayers.jl's scalar surface cache.

All compute runs in libimmersedhip.so (lib/libimmersedhip.so next to this
package); see include/immersedhip.h for the C ABI.
"""
from . import geometry
from ._lib import (ILHError, DimensionMismatch, ArgumentError, NonFiniteError, SingularError, HIPError,
                   RangeError, StateError, LIB_PATH, load as load_library)
from .core import (Context, DeviceArray, Nodes, Edges, ScalarData, VectorData, SurfaceScalarCache,
                   Primal, Dual, GridScaling, Yang3, Roma, M4prime, Witchhat,
                   regularize, interpolate, regularize_normal, normal_interpolate,
                   regularize_normal_cross, normal_cross_interpolate,
                   surface_divergence, surface_grad, surface_curl, surface_curl_cross, surface_grad_cross,
                   divergence, grad, curl, laplacian, inverse_laplacian,
                   create_RTLinvR, create_CLinvCT, create_GLinvD, create_nRTRn,
                   DirichletPoisson, NeumannPoisson, SchurFactor, default_context,
                   dirichlet_solve_batched, dirichlet_solve_batched_device, DirichletBatch,
                   DirichletHeatConduction, mask, complementary_mask, mask_, complementary_mask_,
                   create_surface_filter, filter_surface,
                   dirichlet_solve_moving_batched,
                   EdgeGradient, SurfaceVectorCache, surface_divergence_symm, surface_grad_symm,
                   regularize_normal_symm, normal_interpolate_symm, create_CL2invCT, create_CLinvCT_scalar,
                   create_GLinvD_symm,
                   StokesFlow, LearnedPreconditioner)
from .geometry import GridSpec as PhysicalGrid

__all__ = [n for n in dir() if not n.startswith("_")]
