"""GPU tests against the golden fixtures, the reference's published Neumann
added mass, headline-size parity, determinism and error paths.

Tolerances: FP64 solves agree with the oracle to relative L2 <= 1e-10,
including the surface strength s of the 1024^2 headline (a Schur system of
condition number ~3e5: two FP64 computations in different summation orders
agree to ~kappa*eps; observed 6.6e-12).  Two larger or worse-conditioned
systems below state their own bars and why.
"""
import os

import numpy as np
import pytest

import oracle as O
from conftest import rel_err

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _load(name):
    import ilhip
    d = np.load(os.path.join(ROOT, "tests", "golden", name))
    NX, NY, I0x, I0y = (int(v) for v in d["grid"])
    g = ilhip.PhysicalGrid(NX, NY, I0x, I0y, float(d["dx"]))
    from ilhip.geometry import Surface
    s = Surface(d["X"], d["Y"], d["nx"], d["ny"], d["ds"], d["body"] if "body" in d else np.zeros(len(d["X"]), int))
    return d, g, s


def test_dirichlet_golden(ctx):
    import ilhip
    d, g, s = _load("dirichlet_plain.npz")
    print("dirichlet golden N", s.N, flush=True)
    cache = ilhip.SurfaceScalarCache(s, g, ctx=ctx)
    f, sv = ilhip.DirichletPoisson(cache).solve(d["fb"], filter_power=0)
    assert rel_err(f.flat(), d["f"]) < 1e-10
    assert rel_err(sv.flat(), d["s"]) < 1e-10
    # the default, like the reference's solve (problems.jl:298): C^6 s
    f6, s6 = ilhip.DirichletPoisson(cache).solve(d["fb"])
    assert np.array_equal(f6.flat(), f.flat())
    Cm = O.surface_filter(O.grid(g.NX, g.NY, g.I0x, g.I0y, g.dx), O.Geometry(s.X, s.Y, s.nx, s.ny, s.ds))
    assert rel_err(s6.flat(), np.linalg.matrix_power(Cm, 6) @ d["s"]) < 1e-10


def test_neumann_added_mass(ctx):
    """examples/neumann.ipynb cell 37: M = [0.22649277527914002, ~0]."""
    import ilhip
    d, g, s = _load("neumann_square_circle.npz")
    print("neumann added mass N", s.N, flush=True)
    cache = ilhip.SurfaceScalarCache(s, g, ctx=ctx)
    f, df, st, ds = ilhip.NeumannPoisson(cache).solve(d["vp"])
    m = s.body == 1
    Mx = -np.sum(df.flat()[m] * s.nx[m] * s.ds[m])
    My = -np.sum(df.flat()[m] * s.ny[m] * s.ds[m])
    assert abs(Mx - 0.22649277527914002) < 1e-11
    assert abs(My) < 1e-9
    assert rel_err(df.flat(), d["df"]) < 1e-10
    assert rel_err(ds.flat(), d["dsv"]) < 1e-10
    assert abs(np.linalg.norm(f.flat()) - d["f_norm"]) < 1e-10 * d["f_norm"]
    assert abs(np.linalg.norm(st.flat()) - d["s_norm"]) < 1e-10 * d["s_norm"]
    assert np.abs(f.flat()[d["probe"]] - d["f_probe"]).max() < 1e-10 * np.abs(d["f_probe"]).max()


def test_stokes_golden(ctx):
    """The reference's Stokes example (test/literate/stokes.jl, 214 points on
    406^2): traction (C^2-filtered), velocity and streamfunction probes (the
    streamfunction up to its additive constant) and the total force."""
    import ilhip
    d, g, s = _load("stokes_rectangle.npz")
    print("stokes golden N", s.N, flush=True)
    cache = ilhip.SurfaceVectorCache(s, g, ctx=ctx)
    v, st, sig = ilhip.StokesFlow(cache).solve(d["vp"])
    # traction: to 1e-10 off create_CL2invCT's smallest singular direction
    # (condition number 1.7e7), to cond * eps along it
    dsig = sig.flat() - d["sigma2"]
    u = d["null2"]
    assert np.linalg.norm(dsig - (u @ dsig) * u) < 1e-10 * np.linalg.norm(d["sigma2"])
    assert rel_err(sig.flat(), d["sigma2"]) < 4e-9
    assert np.abs(v.flat()[d["probe_v"]] - d["v_probe"]).max() < 1e-10 * np.abs(d["v_probe"]).max()
    assert abs(np.linalg.norm(v.flat()) - d["v_norm"]) < 1e-10 * d["v_norm"]
    ds = st.flat()[d["probe_s"]] - d["s_probe"]
    assert np.abs(ds - ds.mean()).max() < 1e-10 * np.abs(d["s_probe"]).max()
    F = np.array([np.sum(sig.flat()[:s.N] * s.ds), np.sum(sig.flat()[s.N:] * s.ds)])
    assert abs(F[0] - d["force"][0]) < 1e-9 * abs(d["force"][0]) and abs(F[1]) < 1e-9 * abs(F[0])


def test_headline_parity_and_determinism(ctx):
    import ilhip
    from ilhip import geometry as G
    g, s, fb = G.config_headline()
    print("headline parity N", s.N, flush=True)
    cache = ilhip.SurfaceScalarCache(s, g, ctx=ctx)
    prob = ilhip.DirichletPoisson(cache)
    f1, s1 = prob.solve(fb, filter_power=0)
    f2, s2 = prob.solve(fb, filter_power=0)
    assert np.array_equal(f1.flat(), f2.flat()) and np.array_equal(s1.flat(), s2.flat())   # bitwise
    og = O.grid(g.NX, g.NY, g.I0x, g.I0y, g.dx)
    geo = O.Geometry(s.X, s.Y, s.nx, s.ny, s.ds)
    fr, sr = O.dirichlet_solve(og, geo, fb)
    assert rel_err(f1.flat(), fr) < 1e-10
    assert rel_err(s1.flat(), sr) < 1e-10


def test_dirichlet_many_points_reads_amplitudes_from_memory(ctx):
    """About 5400 surface points (eight circles of radius 0.5 at the headline's
    1.2 dx spacing) on the 1024^2 grid: more amplitudes than the forward row
    pass stages in LDS (4096), so the sources read them from memory.  f, which
    that path produces, is held to 1e-10.  s is held to 2e-8: at this size a
    random one-ulp perturbation of the oracle's Schur matrix moves the oracle's
    own s by 5.5e-10 (and f by 1.3e-13), and the GPU assembles and factors S in
    another summation order (observed 4.7e-9)."""
    import ilhip
    from ilhip import geometry as G
    dx = 4.0 / 1024
    g = G.GridSpec(1025, 1025, 513, 513, dx)
    centres = [(cx, cy) for cx in (-1.3, 0.0, 1.3) for cy in (-1.3, 0.0, 1.3) if cx or cy]
    s = G.concat([G.circle(0.5, 1.2 * dx, cx=cx, cy=cy) for cx, cy in centres])
    fb, _, _ = G.fourier_data(s.theta(), np.random.default_rng(11))
    print("many points N", s.N, flush=True)
    assert s.N > 4096
    cache = ilhip.SurfaceScalarCache(s, g, ctx=ctx)
    f1, s1 = ilhip.DirichletPoisson(cache).solve(fb, filter_power=0)
    og = O.grid(g.NX, g.NY, g.I0x, g.I0y, g.dx)
    fr, sr = O.dirichlet_solve(og, O.Geometry(s.X, s.Y, s.nx, s.ny, s.ds), fb)
    assert rel_err(f1.flat(), fr) < 1e-10
    assert rel_err(s1.flat(), sr) < 2e-8


def test_moving_geometry_update(ctx):
    import ilhip
    from ilhip import geometry as G
    g, s, fb = G.config_plain()
    cache = ilhip.SurfaceScalarCache(s, g, ctx=ctx)
    prob = ilhip.DirichletPoisson(cache)
    prob.solve(fb, filter_power=0)
    s2 = G.circle(0.5, 1.4 * g.dx, cx=0.13, cy=-0.07)
    assert s2.N == s.N
    cache.update_points(s2)
    f, sv = prob.solve(fb, filter_power=0)
    og = O.grid(g.NX, g.NY, g.I0x, g.I0y, g.dx)
    fr, sr = O.dirichlet_solve(og, O.Geometry(s2.X, s2.Y, s2.nx, s2.ny, s2.ds), fb)
    assert rel_err(f.flat(), fr) < 1e-10 and rel_err(sv.flat(), sr) < 1e-10


def test_error_paths(ctx):
    import ilhip
    from ilhip import geometry as G
    g, s, fb = G.config_plain()
    cache = ilhip.SurfaceScalarCache(s, g, ctx=ctx)
    with pytest.raises(ilhip.DimensionMismatch):
        ilhip.DirichletPoisson(cache).solve(fb[:-1])
    with pytest.raises(TypeError):
        ilhip.regularize(ilhip.ScalarData(s.N), ilhip.ScalarData(s.N), cache)
    bad = G.circle(0.5, 1.4 * g.dx, cx=1.98)          # support leaves the grid interior
    edge = ilhip.SurfaceScalarCache(bad, g, ctx=ctx)
    out = ilhip.Nodes(ilhip.Primal, g)
    ilhip.regularize(out, ilhip.ScalarData(bad.N, np.ones(bad.N)), edge)   # clipped, allowed
    with pytest.raises(ilhip.RangeError):
        ilhip.DirichletPoisson(edge).solve(np.ones(bad.N))
    nan = G.circle(0.5, 1.4 * g.dx)
    nan.X[3] = np.nan
    with pytest.raises(ilhip.NonFiniteError):
        ilhip.SurfaceScalarCache(nan, g, ctx=ctx)
    with pytest.raises(ilhip.StateError):
        from ilhip import _lib
        _lib.call("ilh_schur_solve", cache.h, 0, np.zeros(s.N).ctypes.data_as(__import__("ctypes").c_void_p),
                  np.zeros(s.N).ctypes.data_as(__import__("ctypes").c_void_p), 0)


def test_empty_body_list(ctx):
    """No surface points: operators give exact zeros (test/surface_ops.jl:355-371)."""
    import ilhip
    from ilhip.geometry import GridSpec, Surface
    g = GridSpec(40, 40, 20, 20, 0.1)
    e = np.zeros(0)
    cache = ilhip.SurfaceScalarCache(Surface(e, e, e, e, e, np.zeros(0, int)), g, ctx=ctx)
    out = ilhip.Nodes(ilhip.Primal, g, np.ones(g.primal_shape()))
    ilhip.regularize(out, ilhip.ScalarData(0), cache)
    assert np.all(out.data == 0.0)
    f, s = ilhip.DirichletPoisson(cache).solve(np.zeros(0), filter_power=0)
    assert np.all(f.data == 0.0) and s.flat().size == 0
    # masks revert to ones (complement zeros) without points (src/surface_operators.jl:870-873)
    assert np.all(ilhip.mask(cache).data == 1.0)
    assert np.all(ilhip.complementary_mask(cache, ilhip.Nodes(ilhip.Dual, g)).data == 0.0)
    fs, ss = ilhip.DirichletPoisson(cache).solve_multi(np.zeros((0, 3)))
    assert all(np.all(x.data == 0.0) for x in fs) and ss.size == 0
    f, s, it = ilhip.DirichletPoisson(cache).solve_moving(np.zeros(0))
    assert np.all(f.data == 0.0) and it == 0
    assert ilhip.create_surface_filter(cache).size == 0


def test_error_paths_of_the_extended_entry_points(ctx):
    import ctypes as C
    import ilhip
    from ilhip import _lib as L
    from ilhip import geometry as G
    g, s, fb = G.config_plain()
    cache = ilhip.SurfaceScalarCache(s, g, ctx=ctx)
    buf = np.zeros(g.NX * g.NY * 4)
    p = buf.ctypes.data_as(C.c_void_p)
    with pytest.raises(ilhip.ArgumentError):
        L.call("ilh_mask", cache.h, 7, 0, p, L.MEM_HOST)                  # no such grid kind
    with pytest.raises(ilhip.ArgumentError):
        L.call("ilh_surface_filter_apply", cache.h, -1, p, p, L.MEM_HOST)
    with pytest.raises(ilhip.ArgumentError):
        L.call("ilh_dirichlet_solve_multi", cache.h, 0, p, None, p, p, L.MEM_HOST)
    prob = ilhip.DirichletPoisson(cache)
    for tol, it in ((0.0, 10), (1e-10, 0), (1e-10, 201)):
        with pytest.raises(ilhip.ArgumentError):
            prob.solve_moving(fb, tol=tol, max_iter=it)
    with pytest.raises(ilhip.ArgumentError):
        ilhip.DirichletHeatConduction(cache, -1.0, 1e-4, np.ones(s.N))
    # the plan still works after the refused calls
    f, sv = prob.solve(fb, filter_power=0)
    assert np.isfinite(f.data).all()


def _moved(s, angle, dx, dy):
    from ilhip.geometry import Surface
    c, n = np.cos(angle), np.sin(angle)
    return Surface(c * s.X - n * s.Y + dx, n * s.X + c * s.Y + dy, c * s.nx - n * s.ny, n * s.nx + c * s.ny,
                   s.ds, s.body)


def test_moving_body_gmres(ctx):
    """Moving bodies (update_system!, src/system.jl:26-50): GMRES preconditioned
    by a reference geometry's inverse reproduces the oracle's direct solve at
    every step; a motion too large for the preconditioner refactors."""
    import ilhip
    from ilhip import geometry as G
    g, s0, fb = G.config_plain()
    rng = np.random.default_rng(21)
    fm = 0.3 * rng.standard_normal(s0.N)
    cache = ilhip.SurfaceScalarCache(s0, g, ctx=ctx)
    prob = ilhip.DirichletPoisson(cache)
    og = O.grid(g.NX, g.NY, g.I0x, g.I0y, g.dx)
    its = []
    for k, (ang, ox, oy) in enumerate([(0.0, 0.0, 0.0), (0.01, 0.2 * g.dx, 0.0), (0.02, 0.5 * g.dx, -0.3 * g.dx),
                                       (0.05, 1.5 * g.dx, 0.7 * g.dx)]):
        sk = _moved(s0, ang, ox, oy)
        if k:
            cache.update_points(sk)
        f, sv, it = prob.solve_moving(fb, fm, tol=1e-13, max_iter=40)
        its.append(it)
        fr, sr = O.dirichlet_solve(og, O.Geometry(sk.X, sk.Y, sk.nx, sk.ny, sk.ds), fb, fm)
        print("moving step", k, "iterations", it, rel_err(f.flat(), fr), rel_err(sv.flat(), sr), flush=True)
        assert rel_err(f.flat(), fr) < 1e-10
        assert rel_err(sv.flat(), sr) < 1e-10
    # the first call factors this geometry: its inverse, rounded to FP32 for
    # the preconditioner, leaves a perturbation of ~cond(S) 2^-24 for GMRES
    assert its[0] <= 2
    # a large jump with a one-iteration budget falls back to a refactorisation
    sk = _moved(s0, 0.4, 6 * g.dx, -4 * g.dx)
    cache.update_points(sk)
    f, sv, it = prob.solve_moving(fb, fm, tol=1e-13, max_iter=1)
    fr, sr = O.dirichlet_solve(og, O.Geometry(sk.X, sk.Y, sk.nx, sk.ny, sk.ds), fb, fm)
    assert it >= 1 and rel_err(f.flat(), fr) < 1e-10 and rel_err(sv.flat(), sr) < 1e-10


def test_batched_solves(ctx):
    """ilh_dirichlet_solve_batched: every problem of a mixed batch equals its
    own single solve bit for bit, and the oracle to the FP64 bar."""
    import ilhip
    from ilhip import geometry as G
    probs = G.config_batched(seed=9, nprob=7, sizes=(64, 96, 128), dx=0.02)
    caches = [ilhip.SurfaceScalarCache(s, g, ctx=ctx) for g, s, _ in probs]
    print("batched N", [c.N for c in caches], flush=True)
    rng = np.random.default_rng(5)
    fms = [None if k % 2 else 0.2 * rng.standard_normal(c.N) for k, c in enumerate(caches)]
    out = ilhip.dirichlet_solve_batched(caches, [fb for _, _, fb in probs], fms)
    for k, ((g, s, fb), c) in enumerate(zip(probs, caches)):
        f1, s1 = ilhip.DirichletPoisson(c).solve(fb, fms[k], filter_power=0)
        assert np.array_equal(out[k][0].flat(), f1.flat()) and np.array_equal(out[k][1].flat(), s1.flat())
        og = O.grid(g.NX, g.NY, g.I0x, g.I0y, g.dx)
        fr, sr = O.dirichlet_solve(og, O.Geometry(s.X, s.Y, s.nx, s.ny, s.ds), fb, fms[k])
        assert rel_err(f1.flat(), fr) < 1e-10 and rel_err(s1.flat(), sr) < 1e-10
    with pytest.raises(ilhip.ArgumentError):
        ilhip.dirichlet_solve_batched([caches[0], caches[0]], [probs[0][2], probs[0][2]])


def test_batch_graph_replays(ctx):
    """DirichletBatch: a captured batch replayed several times, with new inputs
    and eager solves in between, matches the eager solves bit for bit; a
    geometry change invalidates it."""
    import ilhip
    from ilhip import geometry as G
    probs = G.config_batched(seed=4, nprob=5, sizes=(64, 96, 128), dx=0.02)
    caches = [ilhip.SurfaceScalarCache(s, g, ctx=ctx) for g, s, _ in probs]
    rng = np.random.default_rng(8)
    fps = [ilhip.DeviceArray.from_host(ctx, fb) for _, _, fb in probs]
    fds = [ilhip.DeviceArray(ctx, (g.NX - 1) * (g.NY - 1)) for g, _, _ in probs]
    sds = [ilhip.DeviceArray(ctx, s.N) for _, s, _ in probs]
    batch = ilhip.DirichletBatch(caches, fps, fds, sds)
    for rep in range(3):
        fbs = [rng.standard_normal(c.N) for c in caches]
        for d, fb in zip(fps, fbs):
            d.upload(fb)
        batch.launch()
        ctx.synchronize()
        print("batch replay", rep, flush=True)
        for k, c in enumerate(caches):
            f1, s1 = ilhip.DirichletPoisson(c).solve(fbs[k], filter_power=0)      # eager work between replays
            assert np.array_equal(fds[k].download(), f1.flat())
            assert np.array_equal(sds[k].download(), s1.flat())
    g, s, fb = probs[0]
    caches[0].update_points(G.Surface(s.X + 0.3 * g.dx, s.Y, s.nx, s.ny, s.ds, s.body))
    with pytest.raises(ilhip.StateError):
        batch.launch()
    batch.close()


@pytest.mark.parametrize("fo,steps,tol_body", [(0.25, 12, 1e-10), (50.0, 3, 1e-3)])
def test_heat_if_herk_matches_oracle(ctx, fo, steps, tol_body):
    """IF-HERK heat conduction (test/literate/heatconduction.jl): the GPU march
    equals the oracle's, with and without a body, at Fo = 0.25 (12 steps) and
    Fo = 50 (3 steps: the stage filters' half-widths, 54 and 75 points, take
    the long x and y passes; the bodiless march pins them to 1e-10).  At
    Fo = 50 the stage Schur matrices E exp(c dt kappa L) R have condition
    numbers 4e11, 9e13 and 4e15 (c = 1/3, 2/3, 1: the heat kernel is ~10 cells
    wide against 1.4-cell point spacing), so the body case agrees with the
    oracle's LU solves only to ~kappa eps (1e-3 stated, 5e-5 observed), and the
    march overshoots (|T| up to ~7, the oracle's alike): the scheme's accuracy
    limit, not a kernel error."""
    import ilhip
    from ilhip import geometry as G
    n, dx = 96, 0.03
    g = G.GridSpec(n + 1, n + 1, n // 2 + 1, n // 2 + 1, dx)
    xs = (np.arange(n) + 1 - (n // 2 + 1)) * dx
    Xg, Yg = np.meshgrid(xs, xs, indexing="ij")
    T0 = np.exp(-((Xg - 0.5) ** 2 + (Yg + 0.2) ** 2) / (2 * (5 * dx) ** 2))
    og = O.grid(g.NX, g.NY, g.I0x, g.I0y, g.dx)
    kappa, dt = 1.0, fo * dx * dx
    for s in (G.circle(0.4, 1.4 * dx, -0.2, 0.1), None):
        if s is None:
            e = np.zeros(0)
            s = G.Surface(e, e, e, e, e, np.zeros(0, int))
        tb = np.ones(s.N)
        print("heat Fo", fo, "N", s.N, flush=True)
        cache = ilhip.SurfaceScalarCache(s, g, ctx=ctx)
        heat = ilhip.DirichletHeatConduction(cache, kappa, dt, tb)
        T = heat.advance(ilhip.Nodes(ilhip.Primal, g, T0), steps)
        Tr = O.heat_advance(og, O.Geometry(s.X, s.Y, s.nx, s.ny, s.ds), kappa, dt, steps, T0.ravel(order="F"), tb)
        err = rel_err(T.flat(), Tr)
        print("  rel err", err, flush=True)
        assert err < (tol_body if s.N else 1e-10)
        if s.N and fo < 1:
            Eb = ilhip.ScalarData(s.N)
            ilhip.interpolate(Eb, T, cache)
            assert np.abs(Eb.flat() - 0.5).max() < 1e-10      # E T = (T_b+ + T_b-)/2, T_b- = 0
        heat.close()


def test_handles_outlive_in_any_order():
    """A plan keeps its context alive and a heat integrator its plan: closing
    the caller's handles in any order is safe (the library tears a context
    down when its last plan lets go)."""
    import ilhip
    from ilhip import geometry as G
    g, s, fb = G.config_plain()
    ctx = ilhip.Context(0)
    cache = ilhip.SurfaceScalarCache(s, g, ctx=ctx)
    heat = ilhip.DirichletHeatConduction(cache, 1.0, 0.25 * g.dx ** 2, np.ones(s.N))
    ctx.close()                               # the plan still holds the context
    f1, s1 = ilhip.DirichletPoisson(cache).solve(fb, filter_power=0)
    cache.close()                             # the heat integrator still holds the plan
    T = ilhip.Nodes(ilhip.Primal, g)
    heat.advance(T, 1)
    heat.close()                              # last owner: plan and context go
    ctx2 = ilhip.Context(0)
    cache2 = ilhip.SurfaceScalarCache(s, g, ctx=ctx2)
    f2, s2 = ilhip.DirichletPoisson(cache2).solve(fb, filter_power=0)
    assert np.array_equal(f1.flat(), f2.flat()) and np.array_equal(s1.flat(), s2.flat())
    assert np.isfinite(T.flat()).all() and T.flat().max() > 0
    cache2.close()
    ctx2.close()


def test_moving_batched_matches_single(ctx):
    """ilh_dirichlet_solve_moving_batched: three trajectories solved side by
    side equal their single-plan solves bit for bit at the reference geometry
    and after a move, and the oracle's direct solve to the FP64 bar."""
    import ilhip
    from ilhip import geometry as G
    g, s0, fb = G.config_plain()
    og = O.grid(g.NX, g.NY, g.I0x, g.I0y, g.dx)
    motions = [(0.01, 0.2 * g.dx, 0.0), (0.02, 0.5 * g.dx, -0.3 * g.dx), (0.015, -0.4 * g.dx, 0.2 * g.dx)]
    nb = len(motions)
    cb = [ilhip.SurfaceScalarCache(s0, g, ctx=ctx) for _ in range(nb)]
    cs = [ilhip.SurfaceScalarCache(s0, g, ctx=ctx) for _ in range(nb)]
    rng = np.random.default_rng(31)
    fbs = [fb * (1.0 + 0.1 * k) + 0.05 * rng.standard_normal(s0.N) for k in range(nb)]
    ng = g.primal_shape()[0] * g.primal_shape()[1]
    fp = [ilhip.DeviceArray.from_host(ctx, x) for x in fbs]
    fd = [ilhip.DeviceArray(ctx, ng) for _ in range(nb)]
    sd = [ilhip.DeviceArray(ctx, s0.N) for _ in range(nb)]
    for step in range(2):
        if step:
            for k, (ang, ox, oy) in enumerate(motions):
                sk = _moved(s0, ang, ox, oy)
                cb[k].update_points(sk)
                cs[k].update_points(sk)
        print("batched moving step", step, flush=True)
        its = ilhip.dirichlet_solve_moving_batched(cb, fp, fd, sd, tol=1e-13, max_iter=40)
        ctx.synchronize()
        for k in range(nb):
            f1, s1, it1 = ilhip.DirichletPoisson(cs[k]).solve_moving(fbs[k], tol=1e-13, max_iter=40)
            assert its[k] == it1
            assert np.array_equal(fd[k].download(), f1.flat()) and np.array_equal(sd[k].download(), s1.flat())
            sk = _moved(s0, *motions[k]) if step else s0
            fr, sr = O.dirichlet_solve(og, O.Geometry(sk.X, sk.Y, sk.nx, sk.ny, sk.ds), fbs[k])
            assert rel_err(f1.flat(), fr) < 1e-10 and rel_err(s1.flat(), sr) < 1e-10


def test_moving_gmres_many_slices(ctx):
    """The Arnoldi and finishing launches work on slices of 64 points and sum
    the slices' partial dots 32 at a time: the headline bodies sampled at
    1.0 dx (N = 2263 > 2048, 36 slices) take the second batch; the FP32
    reference inverse of this finer (worse conditioned) sampling leaves GMRES
    more than 64 steps (67; at 0.9 dx it does not converge within 200).  After a move the moving solve equals the oracle's
    direct solve, and two trajectories solved batched equal their single
    solves bit for bit."""
    import ilhip
    from ilhip import geometry as G
    g, s0, fb = G.config_headline(spacing=1.0)
    print("many slices: N", s0.N, "slices", (s0.N + 63) // 64, flush=True)
    assert s0.N > 2048
    og = O.grid(g.NX, g.NY, g.I0x, g.I0y, g.dx)
    ng = g.primal_shape()[0] * g.primal_shape()[1]
    fbs = [fb, 0.7 * fb + 0.2]
    cb = [ilhip.SurfaceScalarCache(s0, g, ctx=ctx) for _ in fbs]
    cs = [ilhip.SurfaceScalarCache(s0, g, ctx=ctx) for _ in fbs]
    fp = [ilhip.DeviceArray.from_host(ctx, x) for x in fbs]
    fd = [ilhip.DeviceArray(ctx, ng) for _ in fbs]
    sd = [ilhip.DeviceArray(ctx, s0.N) for _ in fbs]
    ilhip.dirichlet_solve_moving_batched(cb, fp, fd, sd, tol=1e-12, max_iter=200)   # factors the references
    for k in range(len(fbs)):
        ilhip.DirichletPoisson(cs[k]).solve_moving(fbs[k], tol=1e-12, max_iter=200)
    sk = _moved(s0, 0.02, 0.4 * g.dx, -0.3 * g.dx)
    for c in cb + cs:
        c.update_points(sk)
    its = ilhip.dirichlet_solve_moving_batched(cb, fp, fd, sd, tol=1e-12, max_iter=200)
    ctx.synchronize()
    fr, sr = O.dirichlet_solve(og, O.Geometry(sk.X, sk.Y, sk.nx, sk.ny, sk.ds), fbs[0])
    for k in range(len(fbs)):
        f1, s1, it1 = ilhip.DirichletPoisson(cs[k]).solve_moving(fbs[k], tol=1e-12, max_iter=200)
        print("many slices: trajectory", k, "GMRES steps", it1, flush=True)
        assert 64 < it1 <= 200 and its[k] == it1
        assert np.array_equal(fd[k].download(), f1.flat()) and np.array_equal(sd[k].download(), s1.flat())
        if k == 0:
            print("many slices: f", rel_err(f1.flat(), fr), "s", rel_err(s1.flat(), sr), flush=True)
            # GMRES stops at a 1e-12 relative residual of S s = b; S of this
            # finer sampling is much worse conditioned than at 1.2 dx, which
            # shows in s (1.5e-8) but not in f (7e-13)
            assert rel_err(f1.flat(), fr) < 1e-10 and rel_err(s1.flat(), sr) < 1e-7
    for c in cb + cs:
        c.close()


def test_multi_rhs_matches_single(ctx):
    """ilh_dirichlet_solve_multi: eleven boundary data on one geometry (two
    groups of the surface step, with and without f-) equal eleven single
    solves bit for bit."""
    import ilhip
    from ilhip import geometry as G
    g, s, fb = G.config_plain()
    cache = ilhip.SurfaceScalarCache(s, g, ctx=ctx)
    prob = ilhip.DirichletPoisson(cache)
    rng = np.random.default_rng(41)
    K = 11
    fps = np.stack([fb * (1 + 0.1 * k) + 0.1 * rng.standard_normal(s.N) for k in range(K)], axis=1)
    fms = 0.2 * rng.standard_normal((s.N, K))
    for fm in (None, fms):
        print("multi rhs", K, "with f-" if fm is not None else "", flush=True)
        fs, ss = prob.solve_multi(fps, fm)
        for k in range(K):
            f1, s1 = prob.solve(fps[:, k], None if fm is None else fm[:, k], filter_power=0)
            assert np.array_equal(fs[k].flat(), f1.flat()), k
            assert np.array_equal(ss[:, k], s1.flat()), k
