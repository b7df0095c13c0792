"""FP32 operators (caches with dtype = Float32, src/cache.jl:164-201) against
the FP64 oracle and the FP64 HIP path.

regularize!/interpolate! accumulate the FP64 DDF weights and round once, so
on the same (FP32) inputs their FP32 outputs equal the FP64 outputs rounded
to FP32 bit for bit.  The stencils compute in FP32 (the difference in FP32,
the 1/dx scaling rounded once): against the oracle to 4 FP32 ulps of the
output's magnitude (max-abs relative 5e-7).  inverse_laplacian! is the FFT
convolution in FP32 arithmetic: against the oracle within TOL_CONV32 = 2e-6
(max-abs relative), the accumulated rounding of FP32 transforms of these
lengths (an FP32 FFT of the same zero-padded convolution lands at 3e-7 to
8e-7).
"""
import numpy as np
import pytest

import oracle as O
from conftest import max_rel

pytestmark = pytest.mark.gpu

TOL_F32 = 5e-7
TOL_CONV32 = 2e-6


@pytest.fixture(scope="module")
def setup(ctx):
    import ilhip
    from ilhip import geometry as G
    gs = G.GridSpec(150, 131, 70, 60, 0.03)
    rng = np.random.default_rng(11)
    s = G.star(0.9, rng.uniform(-0.1, 0.1, 5), 1.3 * gs.dx, cx=0.1, cy=0.05)
    c32 = ilhip.SurfaceScalarCache(s, gs, ctx=ctx, dtype=np.float32)
    c64 = ilhip.SurfaceScalarCache(s, gs, ctx=ctx)
    og = O.grid(gs.NX, gs.NY, gs.I0x, gs.I0y, gs.dx)
    geo = O.Geometry(s.X, s.Y, s.nx, s.ny, s.ds)
    return ilhip, c32, c64, gs, og, geo, np.random.default_rng(3)


def test_regularize_interpolate_f32(setup):
    il, c32, c64, gs, og, geo, rng = setup
    N = geo.N
    f = rng.standard_normal(N).astype(np.float32)
    for kind in (il.Primal, il.Dual):
        o32 = il.Nodes(kind, gs, dtype=np.float32)
        o64 = il.Nodes(kind, gs)
        il.regularize(o32, il.ScalarData(N, f), c32)
        il.regularize(o64, il.ScalarData(N, f.astype(np.float64)), c64)
        assert o32.data.dtype == np.float32
        assert np.array_equal(o32.flat(), o64.flat().astype(np.float32)), kind
        fld = rng.standard_normal(o32.data.shape).astype(np.float32)
        s32, s64 = il.ScalarData(N, dtype=np.float32), il.ScalarData(N)
        il.interpolate(s32, il.Nodes(kind, gs, fld), c32)
        il.interpolate(s64, il.Nodes(kind, gs, fld.astype(np.float64)), c64)
        assert np.array_equal(s32.flat(), s64.flat().astype(np.float32)), kind
        okind = O.PRIMAL if kind == il.Primal else O.DUAL
        assert max_rel(s32.flat(), O.interpolate(og, okind, geo, fld.ravel(order="F").astype(np.float64))) < TOL_F32
    vb = rng.standard_normal(2 * N).astype(np.float32)
    q32, q64 = il.Edges(gs, dtype=np.float32), il.Edges(gs)
    il.regularize(q32, il.VectorData(N, vb), c32)
    il.regularize(q64, il.VectorData(N, vb.astype(np.float64)), c64)
    assert np.array_equal(q32.flat(), q64.flat().astype(np.float32))


def test_stencils_f32(setup):
    il, c32, c64, gs, og, geo, rng = setup
    q = rng.standard_normal(O.edge_len(og)).astype(np.float32)
    p = rng.standard_normal(gs.primal_shape()).astype(np.float32)
    s = rng.standard_normal(gs.dual_shape()).astype(np.float32)
    q64, p64, s64 = q.astype(np.float64), p.ravel(order="F").astype(np.float64), s.ravel(order="F").astype(np.float64)
    out = il.Nodes(il.Primal, gs, dtype=np.float32)
    il.divergence(out, il.Edges(gs, q), c32)
    assert out.data.dtype == np.float32 and max_rel(out.flat(), O.divergence(og, q64)) < TOL_F32
    e = il.Edges(gs, dtype=np.float32)
    il.grad(e, il.Nodes(il.Primal, gs, p), c32)
    assert max_rel(e.flat(), O.grad(og, p64)) < TOL_F32
    il.curl(e, il.Nodes(il.Dual, gs, s), c32)
    assert max_rel(e.flat(), O.curl_n2e(og, s64)) < TOL_F32
    w = il.Nodes(il.Dual, gs, dtype=np.float32)
    il.curl(w, il.Edges(gs, q), c32)
    assert max_rel(w.flat(), O.curl_e2n(og, q64)) < TOL_F32
    il.laplacian(out, il.Nodes(il.Primal, gs, p), c32)
    assert max_rel(out.flat(), O.laplacian(og, O.PRIMAL, p64)) < 2 * TOL_F32
    il.laplacian(w, il.Nodes(il.Dual, gs, s), c32)
    assert max_rel(w.flat(), O.laplacian(og, O.DUAL, s64)) < 2 * TOL_F32


@pytest.mark.parametrize("kind", ["Primal", "Dual"])
def test_inverse_laplacian_f32(setup, kind):
    il, c32, c64, gs, og, geo, rng = setup
    k = il.Primal if kind == "Primal" else il.Dual
    shape = gs.primal_shape() if kind == "Primal" else gs.dual_shape()
    a = rng.standard_normal(shape).astype(np.float32)
    o32 = il.Nodes(k, gs, dtype=np.float32)
    o64 = il.Nodes(k, gs)
    il.inverse_laplacian(o32, c32, rhs=il.Nodes(k, gs, a))
    il.inverse_laplacian(o64, c64, rhs=il.Nodes(k, gs, a.astype(np.float64)))
    okind = O.PRIMAL if kind == "Primal" else O.DUAL
    ref = O.InverseLaplacian(og, okind)(a.ravel(order="F").astype(np.float64))
    err = max_rel(o32.flat(), ref)
    print("FP32 inverse_laplacian", kind, "max rel", err, flush=True)
    assert err < TOL_CONV32
    assert max_rel(o64.flat(), ref) < 1e-13


def test_f32_type_checks(setup):
    il, c32, c64, gs, og, geo, rng = setup
    import ilhip._lib as L
    with pytest.raises(L.ArgumentError):   # mixed element types
        il.divergence(il.Nodes(il.Primal, gs, dtype=np.float32), il.Edges(gs), c32)
    with pytest.raises(L.ArgumentError):   # no FP32 form of the surface calculus
        il.surface_divergence(il.Nodes(il.Primal, gs, dtype=np.float32),
                              il.ScalarData(geo.N, dtype=np.float32), c32)
    with pytest.raises(L.ArgumentError):
        il.SurfaceScalarCache(G_star(gs), gs, ctx=c32.ctx, dtype=np.float16)


def G_star(gs):
    from ilhip import geometry as G
    return G.circle(0.5, 1.4 * gs.dx)


def test_mixed_precision_dirichlet_solve(ctx):
    """ilh_dirichlet_solve_mixed on the 1024^2 workload: s equals the FP64
    solve bit for bit (the surface step is FP64); f, formed by the FP32
    convolution, is within 1e-6 of the oracle (relative L2)."""
    import ilhip
    from ilhip import geometry as G
    from conftest import rel_err
    g, s, fb = G.config_headline()
    cache = ilhip.SurfaceScalarCache(s, g, ctx=ctx)
    prob = ilhip.DirichletPoisson(cache)
    f64, s64 = prob.solve(fb, filter_power=0)
    f32, s32 = prob.solve_mixed(fb)
    assert np.array_equal(s32.flat(), s64.flat())
    og = O.grid(g.NX, g.NY, g.I0x, g.I0y, g.dx)
    fr, sr = O.dirichlet_solve(og, O.Geometry(s.X, s.Y, s.nx, s.ny, s.ds), fb)
    ef = rel_err(f32.flat(), fr)
    print("mixed solve: f rel L2 vs oracle", ef, "(FP64 path", rel_err(f64.flat(), fr), ")", flush=True)
    assert ef < 1e-6
    assert rel_err(s32.flat(), sr) < 1e-9
    cache.close()
