"""Parity of every HIP operator with the CPU FP64 oracle on seeded inputs.

Cases: the plain CPU-runnable workload (206x206 grid, circle of 112 points)
and a non-square grid with odd/even sizes and a star-shaped body.  Direct
operators are compared to 1e-12 (max-abs relative to the oracle's max), solves
to relative L2 <= 1e-10 (the FP64 bar).
"""
import numpy as np
import pytest

import oracle as O
from conftest import max_rel, rel_err

pytestmark = pytest.mark.gpu

TOL_OP = 1e-12
TOL_SOLVE = 1e-10


def _cases():
    from ilhip import geometry as G
    g1, s1, _ = G.config_plain()
    g2 = G.GridSpec(150, 131, 70, 60, 0.03)
    rng = np.random.default_rng(11)
    s2 = G.star(0.9, rng.uniform(-0.1, 0.1, 5), 1.3 * g2.dx, cx=0.1, cy=0.05)
    return {"plain": (g1, s1), "nonsquare": (g2, s2)}


CASES = _cases()


@pytest.fixture(scope="module", params=sorted(CASES))
def setup(request, ctx):
    import ilhip
    gs, s = CASES[request.param]
    print("case", request.param, "grid", gs, "N", s.N, flush=True)
    cache = ilhip.SurfaceScalarCache(s, gs, ctx=ctx)
    og = O.grid(gs.NX, gs.NY, gs.I0x, gs.I0y, gs.dx)
    geo = O.Geometry(s.X, s.Y, s.nx, s.ny, s.ds)
    rng = np.random.default_rng(7)
    return ilhip, cache, gs, og, geo, rng


def test_regularize_interpolate(setup):
    il, cache, gs, og, geo, rng = setup
    f = rng.standard_normal(geo.N)
    for kind, okind in ((il.Primal, O.PRIMAL), (il.Dual, O.DUAL)):
        out = il.Nodes(kind, gs)
        il.regularize(out, il.ScalarData(geo.N, f), cache)
        ref = O.regularize(og, okind, geo, f)
        assert max_rel(out.flat(), ref) < TOL_OP, kind
        fld = rng.standard_normal(out.flat().size)
        s = il.ScalarData(geo.N)
        il.interpolate(s, il.Nodes(kind, gs, fld.reshape(out.data.shape, order="F")), cache)
        assert max_rel(s.flat(), O.interpolate(og, okind, geo, fld)) < TOL_OP, kind
    # Edges{Primal} <-> VectorData
    vb = rng.standard_normal(2 * geo.N)
    q = il.Edges(gs)
    il.regularize(q, il.VectorData(geo.N, vb), cache)
    ref = np.concatenate([O.regularize(og, O.EDGE_U, geo, vb[:geo.N]), O.regularize(og, O.EDGE_V, geo, vb[geo.N:])])
    assert max_rel(q.flat(), ref) < TOL_OP
    qe = rng.standard_normal(q.flat().size)
    out = il.VectorData(geo.N)
    il.interpolate(out, il.Edges(gs, qe), cache)
    nu = gs.NX * (gs.NY - 1)
    ref = np.concatenate([O.interpolate(og, O.EDGE_U, geo, qe[:nu]), O.interpolate(og, O.EDGE_V, geo, qe[nu:])])
    assert max_rel(out.flat(), ref) < TOL_OP


def test_normal_ops(setup):
    il, cache, gs, og, geo, rng = setup
    f = rng.standard_normal(geo.N)
    q = il.Edges(gs)
    il.regularize_normal(q, il.ScalarData(geo.N, f), cache)
    assert max_rel(q.flat(), O.regularize_normal(og, geo, f)) < TOL_OP
    il.regularize_normal_cross(q, il.ScalarData(geo.N, f), cache)
    assert max_rel(q.flat(), O.regularize_normal_cross(og, geo, f)) < TOL_OP
    qe = rng.standard_normal(q.flat().size)
    vn = il.ScalarData(geo.N)
    il.normal_interpolate(vn, il.Edges(gs, qe), cache)
    assert max_rel(vn.flat(), O.normal_interpolate(og, geo, qe)) < TOL_OP
    il.normal_cross_interpolate(vn, il.Edges(gs, qe), cache)
    assert max_rel(vn.flat(), O.normal_cross_interpolate(og, geo, qe)) < TOL_OP


def test_surface_calculus(setup):
    il, cache, gs, og, geo, rng = setup
    f = rng.standard_normal(geo.N)
    th = il.Nodes(il.Primal, gs)
    il.surface_divergence(th, il.ScalarData(geo.N, f), cache)
    assert max_rel(th.flat(), O.surface_divergence(og, geo, f)) < TOL_OP
    w = il.Nodes(il.Dual, gs)
    il.surface_curl(w, il.ScalarData(geo.N, f), cache)
    assert max_rel(w.flat(), O.surface_curl(og, geo, f)) < TOL_OP
    il.surface_curl_cross(w, il.ScalarData(geo.N, f), cache)
    assert max_rel(w.flat(), O.surface_curl_cross(og, geo, f)) < TOL_OP
    phi = rng.standard_normal(gs.primal_shape())
    psi = rng.standard_normal(gs.dual_shape())
    vn = il.ScalarData(geo.N)
    il.surface_grad(vn, il.Nodes(il.Primal, gs, phi), cache)
    assert max_rel(vn.flat(), O.surface_grad(og, geo, phi.ravel(order="F"))) < TOL_OP
    il.surface_grad_cross(vn, il.Nodes(il.Primal, gs, phi), cache)
    assert max_rel(vn.flat(), O.surface_grad_cross(og, geo, phi.ravel(order="F"))) < TOL_OP
    il.surface_curl(vn, il.Nodes(il.Dual, gs, psi), cache)
    assert max_rel(vn.flat(), O.surface_curl_adj(og, geo, psi.ravel(order="F"))) < TOL_OP
    il.surface_curl_cross(vn, il.Nodes(il.Dual, gs, psi), cache)
    assert max_rel(vn.flat(), O.surface_curl_cross_adj(og, geo, psi.ravel(order="F"))) < TOL_OP


def test_stencils(setup):
    il, cache, gs, og, geo, rng = setup
    q = rng.standard_normal(O.edge_len(og))
    p = rng.standard_normal(gs.primal_shape())
    s = rng.standard_normal(gs.dual_shape())
    out = il.Nodes(il.Primal, gs)
    il.divergence(out, il.Edges(gs, q), cache)
    assert max_rel(out.flat(), O.divergence(og, q)) < TOL_OP
    e = il.Edges(gs)
    il.grad(e, il.Nodes(il.Primal, gs, p), cache)
    assert max_rel(e.flat(), O.grad(og, p.ravel(order="F"))) < TOL_OP
    il.curl(e, il.Nodes(il.Dual, gs, s), cache)
    assert max_rel(e.flat(), O.curl_n2e(og, s.ravel(order="F"))) < TOL_OP
    w = il.Nodes(il.Dual, gs)
    il.curl(w, il.Edges(gs, q), cache)
    assert max_rel(w.flat(), O.curl_e2n(og, q)) < TOL_OP
    il.laplacian(out, il.Nodes(il.Primal, gs, p), cache)
    assert max_rel(out.flat(), O.laplacian(og, O.PRIMAL, p.ravel(order="F"))) < TOL_OP
    il.laplacian(w, il.Nodes(il.Dual, gs, s), cache)
    assert max_rel(w.flat(), O.laplacian(og, O.DUAL, s.ravel(order="F"))) < TOL_OP


def test_inverse_laplacian(setup):
    il, cache, gs, og, geo, rng = setup
    for kind, okind, shape in ((il.Primal, O.PRIMAL, gs.primal_shape()), (il.Dual, O.DUAL, gs.dual_shape())):
        w = rng.standard_normal(shape)
        out = il.Nodes(kind, gs)
        il.inverse_laplacian(out, cache, rhs=il.Nodes(kind, gs, w))
        ref = O.InverseLaplacian(og, okind)(w.ravel(order="F"))
        assert rel_err(out.flat(), ref) < 1e-13, kind
        # in place
        x = il.Nodes(kind, gs, w)
        il.inverse_laplacian(x, cache)
        assert rel_err(x.flat(), ref) < 1e-13, kind


@pytest.mark.parametrize("kind", ["RTLinvR", "CLinvCT", "GLinvD", "nRTRn"])
def test_schur_build(setup, kind):
    il, cache, gs, og, geo, rng = setup
    ok = {"RTLinvR": O.S_RTLINVR, "CLinvCT": O.S_CLINVCT, "GLinvD": O.S_GLINVD, "nRTRn": O.S_NRTRN}[kind]
    A = getattr(il, "create_" + kind)(cache, scale=gs.dx)
    ref = O.schur(og, geo, ok, scale=gs.dx, method="direct")
    assert max_rel(A, ref) < 1e-12


def test_schur_factor_solve(setup):
    il, cache, gs, og, geo, rng = setup
    fac = il.SchurFactor(cache)
    b = rng.standard_normal(geo.N)
    x = fac.solve(b)
    S = O.schur(og, geo, O.S_RTLINVR)
    info, xr = O.lu_solve(S, b)
    assert info == 0
    assert rel_err(x, xr) < TOL_SOLVE


def test_dirichlet_solve(setup):
    il, cache, gs, og, geo, rng = setup
    fp = rng.standard_normal(geo.N)
    fm = rng.standard_normal(geo.N)
    prob = il.DirichletPoisson(cache)
    S = O.schur(og, geo, O.S_RTLINVR)
    for minus in (None, fm):
        f, s = prob.solve(fp, minus, filter_power=0)
        fr, sr = O.dirichlet_solve(og, geo, fp, minus, S)
        assert rel_err(f.flat(), fr) < TOL_SOLVE
        assert rel_err(s.flat(), sr) < TOL_SOLVE


def test_neumann_solve(setup):
    il, cache, gs, og, geo, rng = setup
    vp = rng.standard_normal(geo.N)
    vm = rng.standard_normal(geo.N)
    prob = il.NeumannPoisson(cache)
    S = O.schur(og, geo, O.S_CLINVCT)
    for minus in (None, vm):
        f, df, s, ds = prob.solve(vp, minus)
        fr, dfr, sr, dsr = O.neumann_solve(og, geo, vp, minus, S)
        assert rel_err(df.flat(), dfr) < TOL_SOLVE
        assert rel_err(ds.flat(), dsr) < TOL_SOLVE
        assert rel_err(f.flat(), fr) < TOL_SOLVE
        assert rel_err(s.flat(), sr) < TOL_SOLVE


@pytest.mark.parametrize("NX,NY", [(12, 9), (9, 40), (40, 200), (300, 70), (1100, 20), (2100, 12), (20, 1100),
                                   (12, 2100)])
def test_inverse_laplacian_periods(ctx, NX, NY):
    """Every FFT length 16..8192 for the row passes and for the column pass, and
    unequal row/column periods (no surface points)."""
    import ilhip as il
    from ilhip.geometry import GridSpec, Surface
    gs = GridSpec(NX, NY, NX // 2, NY // 2, 0.05)
    print("inverse_laplacian grid", NX, NY, flush=True)
    e = np.zeros(0)
    cache = il.SurfaceScalarCache(Surface(e, e, e, e, e, np.zeros(0, int)), gs, ctx=ctx)
    og = O.grid(NX, NY, NX // 2, NY // 2, 0.05)
    rng = np.random.default_rng(NX * 1000 + NY)
    for kind, okind, shape in ((il.Primal, O.PRIMAL, gs.primal_shape()), (il.Dual, O.DUAL, gs.dual_shape())):
        w = rng.standard_normal(shape)
        out = il.Nodes(kind, gs)
        il.inverse_laplacian(out, cache, rhs=il.Nodes(kind, gs, w))
        ref = O.InverseLaplacian(og, okind)(w.ravel(order="F"))
        assert rel_err(out.flat(), ref) < 1e-13, (kind, NX, NY)


def test_masks(setup):
    """mask / complementary_mask / mask! (src/surface_operators.jl:728-900) on all
    grid kinds against the oracle's -L^-1 D_s 1 and its averaging."""
    il, cache, gs, og, geo, rng = setup
    m = O.mask(og, geo)
    likes = [(None, O.PRIMAL), (il.Nodes(il.Dual, gs), O.DUAL), (il.Edges(gs), None)]
    to = lambda v, okind: (np.concatenate([O.primal_to(og, v, O.EDGE_U), O.primal_to(og, v, O.EDGE_V)])
                           if okind is None else O.primal_to(og, v, okind))
    for like, okind in likes:
        ref = to(m, okind)
        cref = to(1.0 - m, okind)     # the primal complement, averaged
        got = il.mask(cache, like)
        assert max_rel(got.flat(), ref) < 1e-12, okind
        assert max_rel(il.complementary_mask(cache, like).flat(), cref) < 1e-12, okind
        w = rng.standard_normal(ref.size)
        x = il.mask(cache, like)
        x.flat()[:] = w
        il.mask_(x, cache)
        assert max_rel(x.flat(), w * ref) < 1e-12, okind
        x.flat()[:] = w
        il.complementary_mask_(x, cache)
        assert max_rel(x.flat(), w * cref) < 1e-12, okind


def test_surface_filter(setup):
    """create_surface_filter and C^k s (src/matrix_operators.jl:254-267,
    test/literate/dirichlet.jl:122-124) against the oracle's dense filter."""
    il, cache, gs, og, geo, rng = setup
    Cr = O.surface_filter(og, geo)
    C = il.create_surface_filter(cache)
    assert max_rel(C, Cr) < 1e-12
    sv = rng.standard_normal(geo.N)
    for k in (0, 1, 6):
        got = il.filter_surface(il.ScalarData(geo.N, sv), cache, k).flat()
        ref = np.linalg.matrix_power(Cr, k) @ sv
        assert max_rel(got, ref) < 1e-12, k


@pytest.mark.parametrize("ddf", ["Roma", "M4prime", "Witchhat"])
def test_other_ddfs(ctx, ddf):
    """The non-default DDF kinds (CartesianGrids' Roma, M4prime, Witchhat) on
    the plain workload: regularize/interpolate to 1e-12, create_RTLinvR to
    1e-12 and the Dirichlet solve to the FP64 bar against the oracle with the
    same kernel."""
    import ilhip as il
    gs, s = CASES["plain"]
    kind = getattr(il, ddf)
    okind = getattr(O, ddf.upper())
    cache = il.SurfaceScalarCache(s, gs, ddftype=kind, ctx=ctx)
    og = O.grid(gs.NX, gs.NY, gs.I0x, gs.I0y, gs.dx)
    geo = O.Geometry(s.X, s.Y, s.nx, s.ny, s.ds)
    rng = np.random.default_rng(23)
    f = rng.standard_normal(geo.N)
    out = il.Nodes(il.Primal, gs)
    il.regularize(out, il.ScalarData(geo.N, f), cache)
    assert max_rel(out.flat(), O.regularize(og, O.PRIMAL, geo, f, ddf_kind=okind)) < TOL_OP
    fld = rng.standard_normal(out.flat().size)
    sv = il.ScalarData(geo.N)
    il.interpolate(sv, il.Nodes(il.Primal, gs, fld.reshape(out.data.shape, order="F")), cache)
    assert max_rel(sv.flat(), O.interpolate(og, O.PRIMAL, geo, fld, ddf_kind=okind)) < TOL_OP
    S = O.schur(og, geo, O.S_RTLINVR, ddf_kind=okind)
    assert max_rel(il.create_RTLinvR(cache), S) < TOL_OP
    fp = rng.standard_normal(geo.N)
    fg, sg = il.DirichletPoisson(cache).solve(fp, filter_power=0)
    fr, sr = O.dirichlet_solve(og, geo, fp, None, S, ddf_kind=okind)
    print(ddf, "f", rel_err(fg.flat(), fr), "s", rel_err(sg.flat(), sr), flush=True)
    assert rel_err(fg.flat(), fr) < TOL_SOLVE and rel_err(sg.flat(), sr) < TOL_SOLVE
