"""Parity of the vector-cache operators, create_CL2invCT, the vector surface
filter and the Stokes flow solve (test/literate/stokes.jl:101-166) with the
CPU FP64 oracle.  Operators to 1e-12 (max-abs relative), the 2N x 2N Schur
matrix to 1e-12, the Stokes solve to relative L2 1e-10 for the velocity, 1e-9
for the traction, and 1e-10 for the streamfunction up to an additive
constant.  create_CL2invCT has condition number ~1e6 here; its smallest
singular direction is the traction that adds a uniform streamfunction (which
carries no velocity), so two FP64 solves fix that constant only to ~kappa eps
(observed 1e-9 relative to s, uniform to 5%) while the rest of s agrees to
roundoff.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from conftest import max_rel, rel_err

pytestmark = pytest.mark.gpu

TOL_OP = 1e-12


def _cases():
    from ilhip import geometry as G
    g1 = G.GridSpec(66, 66, 33, 33, 4.0 / 64)
    s1 = G.circle(0.5, 1.4 * g1.dx, cx=0.013, cy=-0.021)
    g2 = G.GridSpec(101, 101, 51, 51, 0.04)
    s2 = G.rectangle(0.5, 0.25, 1.4 * g2.dx)
    g3 = G.GridSpec(90, 71, 44, 37, 0.03)
    rng = np.random.default_rng(5)
    s3 = G.star(0.55, rng.uniform(-0.1, 0.1, 5), 1.3 * g3.dx, cx=0.05, cy=-0.02)
    return {"circle": (g1, s1), "rectangle": (g2, s2), "star": (g3, s3)}


CASES = _cases()


@pytest.fixture(scope="module", params=sorted(CASES))
def setup(request, ctx):
    import ilhip
    gs, s = CASES[request.param]
    print("case", request.param, "grid", gs, "N", s.N, flush=True)
    cache = ilhip.SurfaceVectorCache(s, gs, ctx=ctx)
    og = O.grid(gs.NX, gs.NY, gs.I0x, gs.I0y, gs.dx)
    geo = O.Geometry(s.X, s.Y, s.nx, s.ny, s.ds)
    rng = np.random.default_rng(17)
    return ilhip, cache, gs, og, geo, rng


def test_vector_operators(setup):
    il, cache, gs, og, geo, rng = setup
    N = geo.N
    f = rng.standard_normal(2 * N)
    w = il.Nodes(il.Dual, gs)
    il.surface_curl(w, il.VectorData(N, f), cache)
    assert max_rel(w.flat(), O.surface_curl_vector(og, geo, f)) < TOL_OP
    psi = rng.standard_normal(gs.dual_shape())
    vd = il.VectorData(N)
    il.surface_curl(vd, il.Nodes(il.Dual, gs, psi), cache)
    assert max_rel(vd.flat(), O.surface_curl_vector_adj(og, geo, psi.ravel(order="F"))) < TOL_OP
    q = il.Edges(gs)
    il.surface_divergence_symm(q, il.VectorData(N, f), cache)
    assert max_rel(q.flat(), O.surface_divergence_symm(og, geo, f)) < TOL_OP
    qe = rng.standard_normal(O.edge_len(og))
    il.surface_grad_symm(vd, il.Edges(gs, qe), cache)
    assert max_rel(vd.flat(), O.surface_grad_symm(og, geo, qe)) < TOL_OP
    A = il.EdgeGradient(gs)
    il.regularize_normal_symm(A, il.VectorData(N, f), cache)
    assert max_rel(A.flat(), O.regularize_normal_symm(og, geo, f)) < TOL_OP
    At = rng.standard_normal(O.tensor_len(og))
    il.normal_interpolate_symm(vd, il.EdgeGradient(gs, At), cache)
    assert max_rel(vd.flat(), O.normal_interpolate_symm(og, geo, At)) < TOL_OP


def test_cl2invct(setup):
    il, cache, gs, og, geo, rng = setup
    S2 = il.create_CL2invCT(cache)
    ref = O.cl2invct(og, geo)
    assert S2.shape == (2 * geo.N, 2 * geo.N)
    assert max_rel(S2, ref) < TOL_OP
    assert max_rel(il.create_CL2invCT(cache, scale=-0.5), -0.5 * ref) < TOL_OP
    # create_CLinvCT_scalar is the scalar cache's create_CLinvCT
    assert np.array_equal(il.create_CLinvCT_scalar(cache), il.create_CLinvCT(cache))
    # factor and solve with 2N-vectors
    b = rng.standard_normal(2 * geo.N)
    fac = il.SchurFactor(cache, kind=il._lib.SCHUR_CL2INVCT)
    x = fac.solve(b)
    assert rel_err(ref @ x, b) < 1e-9


def test_glinvd_symm(setup):
    il, cache, gs, og, geo, rng = setup
    A = il.create_GLinvD_symm(cache)
    ref = O.glinvd_symm(og, geo)
    assert A.shape == (2 * geo.N, 2 * geo.N)
    assert max_rel(A, ref) < TOL_OP
    assert max_rel(il.create_GLinvD_symm(cache, scale=2.0), 2.0 * ref) < TOL_OP
    # the inverse Laplacian of Edges{Primal}, each component on its own grid
    q = rng.standard_normal(O.edge_len(og))
    out = il.Edges(gs)
    il.inverse_laplacian(out, cache, il.Edges(gs, q))
    nu = O.nel(og, O.EDGE_U)
    ref_q = np.concatenate([O.InverseLaplacian(og, O.EDGE_U)(q[:nu]), O.InverseLaplacian(og, O.EDGE_V)(q[nu:])])
    assert max_rel(out.flat(), ref_q) < TOL_OP


def test_vector_surface_filter(setup):
    il, cache, gs, og, geo, rng = setup
    Cr = O.surface_filter_vector(og, geo)
    Cm = il.create_surface_filter(cache)
    assert Cm.shape == (2 * geo.N, 2 * geo.N)
    assert max_rel(Cm, Cr) < 1e-12
    v = rng.standard_normal(2 * geo.N)
    for k in (0, 1, 2, 5):
        got = il.filter_surface(il.VectorData(geo.N, v), cache, k).flat()
        assert max_rel(got, np.linalg.matrix_power(Cr, k) @ v) < 1e-12, k


def test_stokes_solve(setup, ctx):
    il, cache, gs, og, geo, rng = setup
    import ilhip._lib as L
    N = geo.N
    vp = rng.standard_normal(2 * N)
    vm = rng.standard_normal(2 * N)
    prob = il.StokesFlow(cache)
    for minus in (None, vm):
        v, s, sig = prob.solve(vp, minus)
        rv, rs, rsig = O.stokes_solve(og, geo, vp, minus)
        Cf = O.surface_filter_vector(og, geo)
        rsig2 = Cf @ (Cf @ rsig)
        assert rel_err(v.flat(), rv) < 1e-10
        ds = s.flat() - rs
        assert rel_err(s.flat() - ds.mean(), rs) < 1e-10
        assert rel_err(s.flat(), rs) < 1e-8
        assert rel_err(sig.flat(), rsig2) < 1e-9
        v0, s0, sig0 = prob.solve(vp, minus, filter_power=0)
        assert rel_err(sig0.flat(), rsig) < 1e-9
        assert np.array_equal(v0.flat(), v.flat()) and np.array_equal(s0.flat(), s.flat())
    # bitwise repeatable, and the device-memory form equals the host form
    v2, s2, sig2 = prob.solve(vp, vm)
    assert np.array_equal(v2.flat(), v.flat()) and np.array_equal(sig2.flat(), sig.flat())
    dv_p, dv_m = il.DeviceArray.from_host(ctx, vp), il.DeviceArray.from_host(ctx, vm)
    d_v = il.DeviceArray(ctx, v.flat().size)
    d_s = il.DeviceArray(ctx, s.flat().size)
    d_g = il.DeviceArray(ctx, 2 * N)
    L.call("ilh_stokes_solve", cache.h, dv_p.ptr, dv_m.ptr, d_v.ptr, d_s.ptr, d_g.ptr, 2, L.MEM_DEVICE)
    ctx.synchronize()
    assert np.array_equal(d_v.download(), v.flat())
    assert np.array_equal(d_s.download(), s.flat())
    assert np.array_equal(d_g.download(), sig.flat())
    # odd and even filter powers (the last iterate lands in either buffer), on
    # host and device memory, against C^k of the unfiltered traction
    rv, rs, rsig = O.stokes_solve(og, geo, vp, vm)
    Cf = O.surface_filter_vector(og, geo)
    for k in (1, 3, 4):
        _, _, sigk = prob.solve(vp, vm, filter_power=k)
        assert rel_err(sigk.flat(), np.linalg.matrix_power(Cf, k) @ rsig) < 1e-9, k
        L.call("ilh_stokes_solve", cache.h, dv_p.ptr, dv_m.ptr, d_v.ptr, d_s.ptr, d_g.ptr, k, L.MEM_DEVICE)
        ctx.synchronize()
        assert np.array_equal(d_g.download(), sigk.flat()), k
        assert np.array_equal(d_v.download(), v.flat())


def test_stokes_translating_rectangle_constraint(ctx):
    """The reference's example problem (exterior velocity (1, 0), interior 0)
    on the 101^2 grid: E v equals the average surface velocity."""
    import ilhip as il
    from ilhip import geometry as G
    gs = G.GridSpec(101, 101, 51, 51, 0.04)
    s = G.rectangle(0.5, 0.25, 1.4 * gs.dx)
    cache = il.SurfaceVectorCache(s, gs, ctx=ctx)
    vp = np.concatenate([np.ones(s.N), np.zeros(s.N)])
    v, st, sig = il.StokesFlow(cache).solve(vp)
    Ev = il.VectorData(s.N)
    il.interpolate(Ev, v, cache)
    assert np.abs(Ev.flat() - 0.5 * vp).max() < 1e-12


def test_stokes_edge_cases(ctx):
    import ilhip as il
    import ilhip._lib as L
    from ilhip import geometry as G
    gs = G.GridSpec(40, 40, 20, 20, 0.05)
    empty = G.Surface(*(np.zeros(0) for _ in range(5)), np.zeros(0, int))
    cache = il.SurfaceVectorCache(empty, gs, ctx=ctx)
    v, s, sig = il.StokesFlow(cache).solve(np.zeros(0))
    assert not v.flat().any() and not s.flat().any() and sig.flat().size == 0
    s1 = G.circle(0.4, 1.4 * gs.dx)
    c1 = il.SurfaceVectorCache(s1, gs, ctx=ctx)
    prob = il.StokesFlow(c1)
    with pytest.raises(L.DimensionMismatch):
        prob.solve(np.zeros(s1.N))
    with pytest.raises(L.ArgumentError):
        prob.solve(np.zeros(2 * s1.N), filter_power=-1)
    out = np.zeros(O.edge_len(O.grid(gs.NX, gs.NY, gs.I0x, gs.I0y, gs.dx)))
    vp = np.zeros(2 * s1.N)
    with pytest.raises(L.ArgumentError):   # null traction output
        L.call("ilh_stokes_solve", c1.h, vp.ctypes.data_as(C.c_void_p), None, out.ctypes.data_as(C.c_void_p),
               np.zeros(gs.NX * gs.NY).ctypes.data_as(C.c_void_p), None, 2, L.MEM_HOST)
    # points closer than 5 cells to the edge are refused by the derivative operators
    near = G.circle(0.8, 1.4 * gs.dx)
    c2 = il.SurfaceVectorCache(near, gs, ctx=ctx)
    with pytest.raises(L.RangeError):
        il.StokesFlow(c2).solve(np.zeros(2 * near.N))
    # geometry updates drop the cached Stokes matrices
    v1, _, _ = prob.solve(np.concatenate([np.ones(s1.N), np.zeros(s1.N)]))
    moved = G.circle(0.4, 1.4 * gs.dx, cx=0.1)
    c1.update_points(moved)
    v2, _, _ = prob.solve(np.concatenate([np.ones(s1.N), np.zeros(s1.N)]))
    og = O.grid(gs.NX, gs.NY, gs.I0x, gs.I0y, gs.dx)
    geo = O.Geometry(moved.X, moved.Y, moved.nx, moved.ny, moved.ds)
    rv, _, _ = O.stokes_solve(og, geo, np.concatenate([np.ones(s1.N), np.zeros(s1.N)]))
    assert rel_err(v2.flat(), rv) < 1e-10 and not np.array_equal(v1.flat(), v2.flat())
