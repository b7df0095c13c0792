#!/usr/bin/env python3
"""Benchmark: Dirichlet Poisson repeat solves on one MI355X.

Workload: the constrained (Schur-complement) Dirichlet Poisson solve of
ImmersedLayers.jl (test/literate/problems.jl:259-298) on a 1024x1024
Nodes{Primal} grid over [-2,2]^2 (dx = 4/1024) with an ellipse of aspect 2 and a
random star curve (about 1000 surface points each, Yang3 DDF, GridScaling),
FP64.  One step = one solve (exterior data f+ -> grid field f and surface
strength s) with the geometry fixed and the Schur factor cached; inputs and
outputs stay in HBM.  Synthetic seeded data (numpy default_rng).

The JSON line carries the whole-job solves/s (wall clock around K solves),
the roofline of the dominant kernel (its stage launched alone between HIP
events), the relative error against the CPU FP64 oracle, and the oracle's own
solves/s on the host.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "immersedlayers.jl_amd"))

HBM_PEAK_GBS = 8000.0
METRIC = "Dirichlet Poisson solves/s, 1024² grid + 2k surf pts, rel err≤1e-6; HBM GB/s"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(g, s, fb, budget_s):
    """The oracle (a C restatement of the reference algorithm: two zero-padded
    FFT convolutions, sparse regularize/interpolate, LU factorisation inside
    every S\\s) timed on the host; S is built once as construct_system would."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    threads = min(16, os.cpu_count() or 1)
    O.set_threads(threads)
    og = O.grid(g.NX, g.NY, g.I0x, g.I0y, g.dx)
    geo = O.Geometry(s.X, s.Y, s.nx, s.ny, s.ds)
    t0 = time.perf_counter()
    S = O.schur(og, geo, O.S_RTLINVR)
    t_build = time.perf_counter() - t0
    f_ref, s_ref = O.dirichlet_solve(og, geo, fb, None, S)   # warm-up (plans cached)
    n, t0 = 0, time.perf_counter()
    while True:
        O.dirichlet_solve(og, geo, fb, None, S)
        n += 1
        el = time.perf_counter() - t0
        if el >= budget_s or n >= 20:
            break
    # the same on one core (a shorter sample: at least one solve, about a third of the budget)
    O.set_threads(1)
    n1, t1 = 0, time.perf_counter()
    while True:
        O.dirichlet_solve(og, geo, fb, None, S)
        n1 += 1
        el1 = time.perf_counter() - t1
        if el1 >= budget_s / 3 or n1 >= 10:
            break
    O.set_threads(threads)
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            model = next(l.split(":", 1)[1].strip() for l in fh if l.startswith("model name"))
    except (OSError, StopIteration):
        pass
    return {"value": n / el, "unit": "solves/s", "cores": threads, "kind": "port",
            "value_1core": n1 / el1, "cpu_model": model,
            "sample": "%d oracle Dirichlet solves on %d threads and %d on one (LU inside each S\\s, 2 FFT "
                      "convolutions of 2048^2) on the same 1024^2 grid and %d points; S built once in %.1f s "
                      "(not timed)" % (n, threads, n1, s.N, t_build)}, f_ref, s_ref


def moved(s, angle, heave):
    """The surface rotated by `angle` about the origin and shifted by `heave`
    in y: same points, panel lengths and count (rigid-body motion)."""
    from ilhip.geometry import Surface
    c, n = np.cos(angle), np.sin(angle)
    return Surface(c * s.X - n * s.Y, n * s.X + c * s.Y + heave, c * s.nx - n * s.ny, n * s.nx + c * s.ny, s.ds,
                   s.body)


def measured_traffic(stage):
    """HBM bytes per launch of a solve stage from the committed rocprofv3 PMC
    summary (FETCH_SIZE + WRITE_SIZE, calibrated on a copy kernel), or None."""
    path = os.path.join(ROOT, "profiles", "dirichlet_1024_n1885_hbm_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes[%d]" % stage):
            return float(line.split("=")[1])
    return None


def bench_ops(args):
    """Grid and surface operators on 2048^2 Nodes{Primal} (4 star curves of
    2048 points, i.i.d. N(0,1) fields), device buffers, FP64.  Each op is
    launched `steps` times back to back between two synchronisations;
    algorithmic bytes per call follow the reference's byte model (each argument
    read once, the output written once)."""
    import ilhip
    from ilhip import _lib as L
    from ilhip import geometry as G
    g, s = G.config_microbench()
    ctx = ilhip.Context(0)
    cache = ilhip.SurfaceScalarCache(s, g, ctx=ctx)
    rng = np.random.default_rng(1)
    ng, nd, ne, N = (g.NX - 1) * (g.NY - 1), g.NX * g.NY, g.NX * (g.NY - 1) + (g.NX - 1) * g.NY, s.N
    P = 4096
    results = {}
    # FP64 (the reference's default) and FP32 (a Float32 cache: b = 4 bytes per value; the
    # LGF spectrum stays FP64 and the weights of the 16-point stencils FP64)
    for dt, b, sfx in ((np.float64, 8, ""), (np.float32, 4, "_f32")):
        dev = {k: ilhip.DeviceArray.from_host(ctx, rng.standard_normal(n), dtype=dt)
               for k, n in (("p", ng), ("p2", ng), ("d", nd), ("d2", nd), ("e", ne), ("e2", ne), ("f", N))}
        ops = [
            ("laplacian (Nodes{Primal})", "ilh_laplacian", (L.NODES_PRIMAL, "p2", "p"), 2 * b * ng),
            ("divergence (Edges -> Nodes{Primal})", "ilh_divergence", ("p", "e"), b * (ne + ng)),
            ("grad (Nodes{Primal} -> Edges)", "ilh_grad", ("e2", "p"), b * (ne + ng)),
            ("curl (Nodes{Dual} -> Edges)", "ilh_curl_nodes_to_edges", ("e2", "d"), b * (ne + nd)),
            ("curl (Edges -> Nodes{Dual})", "ilh_curl_edges_to_nodes", ("d2", "e"), b * (ne + nd)),
            ("regularize Rc (ScalarData -> Nodes{Primal})", "ilh_regularize", (L.NODES_PRIMAL, "p2", "f"),
             b * ng + 32 * N),
            ("interpolate Rc^T (Nodes{Primal} -> ScalarData)", "ilh_interpolate", (L.NODES_PRIMAL, "f", "p"),
             N * (16 * b + b + 16)),
            ("inverse_laplacian (Nodes{Primal}, LGF period 4096^2)", "ilh_inverse_laplacian",
             (L.NODES_PRIMAL, "p2", "p"), 2 * b * ng + 8 * P * (P // 2 + 1)),
        ]
        res = {}
        for name, fn, a, nbytes in ops:
            fn += sfx
            ptrs = [dev[x].ptr if isinstance(x, str) else x for x in a]
            for _ in range(max(1, args.warmup // 10)):
                L.call(fn, cache.h, *ptrs, L.MEM_DEVICE)
            ctx.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                L.call(fn, cache.h, *ptrs, L.MEM_DEVICE)
            ctx.synchronize()
            us = 1e6 * (time.perf_counter() - t0) / args.steps
            res[name] = {"us_per_call": us, "algorithmic_bytes": nbytes, "GB_s": nbytes / us * 1e-3,
                         "frac_hbm": nbytes / us * 1e-3 / HBM_PEAK_GBS}
            if fn.startswith("ilh_interpolate"):
                res[name]["points_per_s"] = N / (us * 1e-6)
            log("%-55s %-4s %9.2f us  %8.1f GB/s" % (name, "f32" if sfx else "f64", us, nbytes / us * 1e-3))
        results["f32" if sfx else "f64"] = res
        for d in dev.values():
            d.free()
    res = results["f64"]
    best = max(res, key=lambda k: res[k]["GB_s"])
    out = {"metric": "grid/surface operator sweeps on 2048^2, HBM GB/s", "value": res[best]["GB_s"], "unit": "GB/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic (i.i.d. N(0,1) fields, 4 seeded star curves)",
           "config": {"workload": "operator microbenchmark", "grid_primal": [g.NX - 1, g.NY - 1], "npts": N},
           "ops": res, "ops_f32": results["f32"]}
    print(json.dumps(out), flush=True)


def bench_batched(args):
    """64 independent Dirichlet problems (grids 256^2..1024^2, 1-3 bodies,
    dx = 0.01) solved per step; repeat solves with each problem's Schur factor
    cached.  The step is four launches over the tiles of every problem
    (surface step, row, column and inverse-row passes).  value: the step
    captured once as a HIP graph (after the autotuner picked each pass's
    schedule and work-group shape) and replayed; also the eager batched call
    and the 64 solves issued one after another on one stream."""
    import ilhip
    from ilhip import geometry as G
    probs = G.config_batched()
    ctx = ilhip.Context(0)
    caches = [ilhip.SurfaceScalarCache(s, g, ctx=ctx) for g, s, _ in probs]
    t0 = time.perf_counter()
    ilhip.dirichlet_solve_batched(caches, [fb for _, _, fb in probs])      # prepares every problem
    t_prep = time.perf_counter() - t0
    fps = [ilhip.DeviceArray.from_host(ctx, fb) for _, _, fb in probs]
    fds = [ilhip.DeviceArray(ctx, (g.NX - 1) * (g.NY - 1)) for g, _, _ in probs]
    sds = [ilhip.DeviceArray(ctx, s.N) for _, s, _ in probs]
    nprob = len(probs)
    t0 = time.perf_counter()
    batch = ilhip.DirichletBatch(caches, fps, fds, sds)     # autotunes, then captures the step as one graph
    ctx.synchronize()
    t_tune = time.perf_counter() - t0
    tuning = batch.tuning()
    for _ in range(args.warmup):
        batch.launch()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.launch()
    ctx.synchronize()
    graphed = nprob * args.steps / (time.perf_counter() - t0)
    batch.close()
    for _ in range(args.warmup):
        ilhip.dirichlet_solve_batched_device(caches, fps, fds, sds)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ilhip.dirichlet_solve_batched_device(caches, fps, fds, sds)
    ctx.synchronize()
    batched = nprob * args.steps / (time.perf_counter() - t0)
    probs_dp = [ilhip.DirichletPoisson(c) for c in caches]
    t0 = time.perf_counter()
    for _ in range(args.steps):
        for k in range(nprob):
            probs_dp[k].solve_device(fps[k], None, fds[k], sds[k])
    ctx.synchronize()
    sequential = nprob * args.steps / (time.perf_counter() - t0)
    sizes = [g.NX - 1 for g, _, _ in probs]
    # the whole step with the headline's byte model per problem (two
    # convolutions' fields and spectra, the N x N solve matrix)
    step_bytes = 0
    for g, s, _ in probs:
        nx, ny = g.primal_shape()
        P = 16
        while P < 2 * max(nx, ny) - 2:
            P *= 2
        step_bytes += 9 * 8 * nx * ny + 2 * 8 * P * (P // 2 + 1) + 8 * s.N * s.N
    step_gbs = step_bytes * graphed / nprob / 1e9
    # value: the captured batch replayed (one graph launch per step of 64 solves)
    out = {"metric": "batched Dirichlet solves/s (64 mixed problems)", "value": graphed, "unit": "solves/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * nprob / graphed,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic (seeded bodies and Fourier surface data)",
           "config": {"workload": "64 independent problems per step, Schur factors cached",
                      "grid_primal_sizes": sorted(set(sizes)), "npts_total": int(sum(s.N for _, s, _ in probs)),
                      "dx": 0.01},
           "roofline": {"bound": "hbm", "kernel": "whole batched step (4 launches)",
                        "achieved": step_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": step_gbs / HBM_PEAK_GBS,
                        "traffic": None, "algorithmic_bytes_per_step": step_bytes},
           "tuning": tuning, "tune_and_capture_s": t_tune,
           "eager_batched_solves_per_s": batched, "sequential_solves_per_s": sequential, "prep_s": t_prep}
    print(json.dumps(out), flush=True)


def bench_heat(args):
    """Unbounded heat conduction with 3 Dirichlet bodies (T_b = 1) on 2048^2,
    IF-HERK at Fo = 0.25: `steps` time steps (200 by the workload's definition
    when run with --steps 200) after `warmup` untimed ones."""
    import ilhip
    from ilhip import geometry as G
    g, s, T0 = G.config_heat()
    ctx = ilhip.Context(0)
    cache = ilhip.SurfaceScalarCache(s, g, ctx=ctx)
    kappa, dt = 1.0, 0.25 * g.dx ** 2
    t0 = time.perf_counter()
    heat = ilhip.DirichletHeatConduction(cache, kappa, dt, np.ones(s.N))
    t_prep = time.perf_counter() - t0
    Td = ilhip.DeviceArray.from_host(ctx, T0)
    heat.advance_device(Td, args.warmup)
    ctx.synchronize()
    t0 = time.perf_counter()
    heat.advance_device(Td, args.steps)
    ctx.synchronize()
    el = time.perf_counter() - t0
    T = ilhip.Nodes(ilhip.Primal, g, Td.download().reshape(g.primal_shape(), order="F"))
    Eb = ilhip.ScalarData(s.N)
    ilhip.interpolate(Eb, T, cache)
    out = {"metric": "IF-HERK heat-conduction steps/s (2048^2, 3 bodies)", "value": args.steps / el, "unit": "steps/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * el / args.steps,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic (3 seeded star bodies at T_b = 1, 4 Gaussian bumps)",
           "config": {"workload": "Dirichlet heat conduction, Fo = 0.25, kappa = 1",
                      "grid_primal": list(g.primal_shape()), "npts": s.N, "dt": dt},
           "prep_s": t_prep, "constraint_residual": float(np.abs(Eb.flat() - 0.5).max())}
    heat.close()
    print(json.dumps(out), flush=True)


def bench_stokes(args):
    """The Stokes flow example of test/literate/stokes.jl (vector cache): 406^2
    dual grid at dx = 0.01, Rectangle(0.5, 0.25) with 214 points, translated at
    unit velocity.  One step = one solve (surface velocities -> velocity field,
    streamfunction and the C^2-filtered traction) with the geometry fixed;
    the reference notebook times the same call at 0.0963 s (examples/stokes.ipynb)."""
    import ilhip
    import ilhip._lib as L
    from ilhip import geometry as G
    g, s, vplus = G.config_stokes()
    ctx = ilhip.Context(0)
    cache = ilhip.SurfaceVectorCache(s, g, ctx=ctx)
    prob = ilhip.StokesFlow(cache)
    t0 = time.perf_counter()
    v, st, sig = prob.solve(vplus)          # builds and inverts the two Schur matrices
    t_prep = time.perf_counter() - t0
    steps = min(args.steps, 500)
    ne, nd = v.flat().size, st.flat().size
    d_vp = ilhip.DeviceArray.from_host(ctx, vplus)
    d_v, d_s, d_g = ilhip.DeviceArray(ctx, ne), ilhip.DeviceArray(ctx, nd), ilhip.DeviceArray(ctx, 2 * s.N)

    def run(k):
        for _ in range(k):
            L.call("ilh_stokes_solve", cache.h, d_vp.ptr, None, d_v.ptr, d_s.ptr, d_g.ptr, 2, L.MEM_DEVICE)

    run(args.warmup)
    ctx.synchronize()
    t0 = time.perf_counter()
    run(steps)
    ctx.synchronize()
    el = time.perf_counter() - t0
    assert np.array_equal(d_v.download(), v.flat()) and np.array_equal(d_g.download(), sig.flat())
    t0 = time.perf_counter()
    nh = 50
    for _ in range(nh):
        prob.solve(vplus)
    t_host = (time.perf_counter() - t0) / nh
    Ev = ilhip.VectorData(s.N)
    ilhip.interpolate(Ev, v, cache)
    out = {"metric": "Stokes flow solves/s (406^2 grid, rectangle of 214 points)", "value": steps / el,
           "unit": "solves/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup, "ms_per_step": 1e3 * el / steps,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "the reference example's geometry and boundary data (exterior velocity (1, 0), interior 0)",
           "config": {"workload": "Stokes flow, test/literate/stokes.jl", "grid_dual": [g.NX, g.NY], "dx": g.dx,
                      "npts": s.N},
           "host_io_ms_per_solve": 1e3 * t_host, "reference_published_s_per_solve": 0.0963,
           "speedup_vs_published_host_io": 0.0963 / t_host, "prep_s": t_prep,
           "constraint_residual": float(np.abs(Ev.flat() - 0.5 * vplus).max())}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        threads = min(16, os.cpu_count() or 1)
        O.set_threads(threads)
        og = O.grid(g.NX, g.NY, g.I0x, g.I0y, g.dx)
        geo = O.Geometry(s.X, s.Y, s.nx, s.ny, s.ds)
        t0 = time.perf_counter()
        S2 = O.cl2invct(og, geo)
        Ss = O.schur(og, geo, O.S_CLINVCT)
        t_build = time.perf_counter() - t0
        n, t0 = 0, time.perf_counter()
        while True:
            rv, rs, rsig = O.stokes_solve(og, geo, vplus, None, S2, Ss)
            n += 1
            e = time.perf_counter() - t0
            if e >= args.cpu_budget or n >= 20:
                break
        Cf = O.surface_filter_vector(og, geo)
        out["rel_err_v"] = float(np.linalg.norm(v.flat() - rv) / np.linalg.norm(rv))
        out["rel_err_sigma"] = float(np.linalg.norm(sig.flat() - Cf @ (Cf @ rsig)) / np.linalg.norm(Cf @ (Cf @ rsig)))
        out["cpu_baseline"] = {"value": n / e, "unit": "solves/s", "cores": threads, "kind": "port",
                               "sample": "%d oracle Stokes solves (6 FFT convolutions of 1024^2, LU inside each "
                                         "S\\v' and Ss\\ds) on the same grid and points; S by the reference's "
                                         "column loop and Ss by direct assembly, once, in %.1f s (not timed)"
                                         % (n, t_build)}
    print(json.dumps(out), flush=True)


def moving_measurements(ilhip, G, ctx, cache, prob, g, s, fp, fd, sd, ng):
    """Measurement (ii) on the headline geometry: the plain new-geometry solve,
    moving-body steps (GMRES with the recycled FP32 inverse; with the learned
    band), and 32 trajectories batched.  Leaves `cache` at geometry `s`."""
    # new-geometry solve (a moving-body step): update_system! with the bodies
    # pitched and heaved, then Schur build + factorisation + solve matrices + one solve
    new_geo_ms = []
    for k in range(1, 4):
        cache.update_points(moved(s, 0.02 * k, 0.01 * k))
        ctx.synchronize()
        t0 = time.perf_counter()
        prob.solve_device(fp, None, fd, sd)
        ctx.synchronize()
        new_geo_ms.append(1e3 * (time.perf_counter() - t0))

    # moving bodies (the learned-preconditioner workload's motion, solved by
    # GMRES preconditioned with a reference geometry's inverse): heave
    # 0.2 sin(2 pi t) and pitch 0.05 sin(2 pi t) about the origin, dt = 0.004
    cache.update_points(s)
    mv_ms, mv_it = [], []
    for k in range(11):
        t = 0.004 * k
        cache.update_points(moved(s, 0.05 * np.sin(2 * np.pi * t), 0.2 * np.sin(2 * np.pi * t)))
        ctx.synchronize()
        t0 = time.perf_counter()
        it = prob.solve_moving_device(fp, None, fd, sd, tol=1e-12, max_iter=30)
        ctx.synchronize()
        if k:   # step 0 factors the reference geometry
            mv_ms.append(1e3 * (time.perf_counter() - t0))
            mv_it.append(it)

    # the learned preconditioner: trained on the exact inverses of ten
    # geometries (the bodies pitched and heaved at seeded amplitudes, and two
    # other star shapes), then the moving steps above solved with its band
    # (no factorisation at any geometry)
    t0 = time.perf_counter()
    model = ilhip.LearnedPreconditioner(ctx, half_band=16, seed=7)
    rng_l = np.random.default_rng(17)
    geoms = [moved(s, rng_l.uniform(-0.3, 0.3), rng_l.uniform(-0.2, 0.2)) for _ in range(8)]
    geoms += [G.config_headline(seed=sd_)[1] for sd_ in (6, 7)]
    for sk in geoms:
        ck = ilhip.SurfaceScalarCache(sk, g, ctx=ctx)
        model.add_geometry(ck)
        ck.close()
    t_data = time.perf_counter() - t0
    n_rows = model.rows
    t0 = time.perf_counter()
    model.train(2000, batch=2048, lr=1e-3, seed=1)
    loss = model.train(1000, batch=2048, lr=2e-4, seed=2)
    t_train = time.perf_counter() - t0
    lcache = ilhip.SurfaceScalarCache(s, g, ctx=ctx)
    lcache.set_preconditioner(model)
    lprob = ilhip.DirichletPoisson(lcache)
    ln_ms, ln_it = [], []
    for k in range(11):
        t = 0.004 * k
        lcache.update_points(moved(s, 0.05 * np.sin(2 * np.pi * t), 0.2 * np.sin(2 * np.pi * t)))
        ctx.synchronize()
        t0 = time.perf_counter()
        it = lprob.solve_moving_device(fp, None, fd, sd, tol=1e-12, max_iter=80)
        ctx.synchronize()
        if k:
            ln_ms.append(1e3 * (time.perf_counter() - t0))
            ln_it.append(it)
    lcache.close()
    model.close()

    # batched moving bodies: 32 trajectories of the same bodies, each heaving
    # 0.2 sin(2 pi f t) and pitching 0.05 sin(2 pi f t) with its own seeded
    # frequency f ~ U(0.5, 2), dt = 0.004; one batched call solves a step of
    # all 32 (update_points of every trajectory outside the timing)
    ntraj = 32
    rng = np.random.default_rng(7)
    freq = rng.uniform(0.5, 2.0, ntraj)
    tcaches = [ilhip.SurfaceScalarCache(s, g, ctx=ctx) for _ in range(ntraj)]
    tf = [ilhip.DeviceArray(ctx, ng) for _ in range(ntraj)]
    ts = [ilhip.DeviceArray(ctx, s.N) for _ in range(ntraj)]
    bat_ms, bat_it = [], []
    for k in range(5):
        t = 0.004 * k
        if k:
            for c, fr in zip(tcaches, freq):
                ph = np.sin(2 * np.pi * fr * t)
                c.update_points(moved(s, 0.05 * ph, 0.2 * ph))
        ctx.synchronize()
        t0 = time.perf_counter()
        its = ilhip.dirichlet_solve_moving_batched(tcaches, [fp] * ntraj, tf, ts, tol=1e-12, max_iter=30)
        ctx.synchronize()
        if k:   # step 0 factors each trajectory's reference geometry
            bat_ms.append(1e3 * (time.perf_counter() - t0) / ntraj)
            bat_it.append(float(np.mean(its)))
    for c in tcaches:
        c.close()
    del tf, ts

    cache.update_points(s)
    return {
        "new_geometry_solve_ms": float(np.median(new_geo_ms)),
        "moving_step_ms": float(np.mean(mv_ms)),
        "moving_gmres_iterations": mv_it,
        "moving_batched_ms_per_solve": float(np.mean(bat_ms)),
        "moving_batched_trajectories": ntraj,
        "moving_batched_mean_iterations": bat_it,
        "learned_ms_per_solve": float(np.mean(ln_ms)),
        "learned_gmres_iterations": ln_it,
        "learned_training": {"rows": n_rows, "data_s": t_data, "train_s": t_train, "steps": 3000,
                             "final_loss": loss},
        "preconditioned_speedup": {
            "single": float(np.mean(new_geo_ms) / np.mean(mv_ms)),
            "batched": float(np.mean(new_geo_ms) / np.mean(bat_ms)),
            "learned": float(np.mean(new_geo_ms) / np.mean(ln_ms))},
    }


def bench_moving(args):
    """Measurement (ii) alone: new-geometry, moving-body and batched
    moving-body solves on the headline geometry (32 trajectories)."""
    import ilhip
    from ilhip import geometry as G
    g, s, fb = G.config_headline()
    ng = (g.NX - 1) * (g.NY - 1)
    ctx = ilhip.Context(0)
    cache = ilhip.SurfaceScalarCache(s, g, ctx=ctx)
    prob = ilhip.DirichletPoisson(cache)
    prob.solve(fb, filter_power=0)
    fp = ilhip.DeviceArray.from_host(ctx, fb)
    fd = ilhip.DeviceArray(ctx, ng)
    sd = ilhip.DeviceArray(ctx, s.N)
    ex = moving_measurements(ilhip, G, ctx, cache, prob, g, s, fp, fd, sd, ng)
    value = 1e3 / ex["moving_batched_ms_per_solve"]
    out = {"metric": "moving-body Dirichlet solves/s (32 trajectories batched, new geometry each step)",
           "value": value, "unit": "solves/s", "n_gpus": 1, "steps": 4, "warmup": 1,
           "ms_per_step": ex["moving_batched_ms_per_solve"] * ex["moving_batched_trajectories"],
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic (seeded pitch/heave trajectories of the headline bodies)",
           "config": {"workload": "moving bodies, 1024^2 grid, N = %d, 32 trajectories" % s.N}}
    out.update(ex)
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of oracle CPU time")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the measurements after the headline's (new geometry, moving bodies, learned "
                         "preconditioner), e.g. for counter runs")
    ap.add_argument("--workload", choices=["dirichlet", "ops", "batched", "heat", "stokes", "moving"],
                    default="dirichlet",
                    help="dirichlet: the repeat-solve metric; ops: operator sweeps on the 2048^2 "
                         "microbenchmark grid (GB/s against HBM peak)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run 5 device copies of 256 MiB (torch copy_) as the reference "
                         "kernel for scaling the FETCH_SIZE/WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        sys.exit(2)
    if os.environ.get("RANK", "0") != "0":
        return

    if args.copy_calibration:
        import torch
        a = torch.ones(32 << 20, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        log("copy calibration: 5 x %d bytes read + written" % (a.numel() * 8))

    if args.workload == "ops":
        return bench_ops(args)
    if args.workload == "batched":
        return bench_batched(args)
    if args.workload == "heat":
        return bench_heat(args)
    if args.workload == "stokes":
        return bench_stokes(args)
    if args.workload == "moving":
        return bench_moving(args)

    import ilhip
    from ilhip import geometry as G

    g, s, fb = G.config_headline()
    nx, ny = g.primal_shape()
    ng = nx * ny
    log("workload: grid %dx%d, N=%d" % (nx, ny, s.N))
    ctx = ilhip.Context(0)
    cache = ilhip.SurfaceScalarCache(s, g, ctx=ctx)
    prob = ilhip.DirichletPoisson(cache)
    t0 = time.perf_counter()
    f_gpu, s_gpu = prob.solve(fb, filter_power=0)   # builds S, its factor and the solve matrices
    t_prep = time.perf_counter() - t0
    fp = ilhip.DeviceArray.from_host(ctx, fb)
    fd = ilhip.DeviceArray(ctx, ng)
    sd = ilhip.DeviceArray(ctx, s.N)
    # per-kernel launch time, before the timed region: each stage launched
    # alone, back to back, between two HIP events on the plan's stream
    stage_ms, _ = prob.profile(fp, None, fd, sd, max(200, min(args.steps, 2000)))
    for _ in range(args.warmup):
        prob.solve_device(fp, None, fd, sd)
    ctx.synchronize()
    # timed region: exactly K solves launched back to back from one host call
    # (HIP events only around the whole loop), synchronized on both sides
    t0 = time.perf_counter()
    _, tot_ms = prob.profile(fp, None, fd, sd, args.steps, stages=False)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    sps = args.steps / wall
    f_dev = fd.download()
    assert np.allclose(f_dev, f_gpu.flat(), rtol=0, atol=0)

    # repeat solves with many boundary data at once (8 right-hand sides per
    # call, seeded scalings of the same data): solves/s of the batched form
    nrhs = 8
    fps = np.stack([fb * (1.0 + 0.05 * k) for k in range(nrhs)], axis=1)
    fp_m = ilhip.DeviceArray.from_host(ctx, fps)
    fd_m = ilhip.DeviceArray(ctx, ng * nrhs)
    sd_m = ilhip.DeviceArray(ctx, s.N * nrhs)
    for _ in range(5):
        prob.solve_multi_device(nrhs, fp_m, None, fd_m, sd_m)
    ctx.synchronize()
    reps = max(20, args.steps // (4 * nrhs))
    t0 = time.perf_counter()
    for _ in range(reps):
        prob.solve_multi_device(nrhs, fp_m, None, fd_m, sd_m)
    ctx.synchronize()
    multi_sps = reps * nrhs / (time.perf_counter() - t0)
    f_multi0 = fd_m.download()[:ng]
    assert np.array_equal(f_multi0, f_gpu.flat())   # column 0 is the headline solve, bit for bit
    del fp_m, fd_m, sd_m

    # the split-precision solve (FP64 forward transforms and s, FP32 inverse
    # transforms): solves/s and f against the FP64 solve
    fd32 = ilhip.DeviceArray(ctx, ng)
    for _ in range(5):
        prob.solve_mixed_device(fp, None, fd32, sd)
    ctx.synchronize()
    reps = max(50, min(args.steps, 1000))
    t0 = time.perf_counter()
    for _ in range(reps):
        prob.solve_mixed_device(fp, None, fd32, sd)
    ctx.synchronize()
    mixed_sps = reps / (time.perf_counter() - t0)
    f32 = fd32.download()
    mixed_err = float(np.linalg.norm(f32 - f_gpu.flat()) / np.linalg.norm(f_gpu.flat()))
    del fd32

    # the same solve with host arrays in and out (PCIe-inclusive, not the metric)
    t0 = time.perf_counter()
    for _ in range(20):
        prob.solve(fb, filter_power=0)
    host_sps = 20 / (time.perf_counter() - t0)
    extra = moving_measurements(ilhip, G, ctx, cache, prob, g, s, fp, fd, sd, ng) if not args.no_extras else {}

    # algorithmic bytes per launch (FP64) of each kernel of the solve
    N, b = s.N, 8
    Px = Py = 2048
    spec = b * Px * (Py // 2 + 1)
    kernels = [
        ("surface solve matvec (Schur solve with cached factor)", b * N * N + 24 * N),
        ("row FFT of the regularized sources", b * ng),
        ("column FFT x LGF spectrum x inverse column FFT", spec),
        ("inverse row FFT -> f", b * ng),
    ]
    dom = int(np.argmax(stage_ms))
    name, abytes = kernels[dom]
    achieved = abytes / (stage_ms[dom] * 1e-3) / 1e9
    # whole-solve view with the reference's own byte model (two convolutions)
    solve_bytes = 9 * b * ng + 2 * spec + 8 * N * N
    solve_gbs = solve_bytes * sps / 1e9

    out = {
        "metric": METRIC,
        "value": sps,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 / sps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded ellipse + star geometry, Fourier surface data)",
        "config": {"workload": "Dirichlet repeat solve, geometry fixed, Schur factor cached",
                   "grid_primal": [nx, ny], "dx": g.dx, "npts": N, "ddf": "Yang3", "scaling": "GridScaling",
                   "fft_period": [Px, Py]},
        "roofline": {"bound": "hbm", "kernel": name, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(dom),
                     "algorithmic_bytes_per_launch": abytes, "avg_launch_ms": stage_ms[dom]},
        "stage_ms": dict(zip(["matvec", "row_fft", "col_fft", "inv_row_fft"], stage_ms)),
        "event_ms_per_solve": tot_ms,
        "solve_model_gbs": solve_gbs,
        "solve_model_frac": solve_gbs / HBM_PEAK_GBS,
        "prep_s": t_prep,
        "host_io_solves_per_s": host_sps,
        "multi_rhs_solves_per_s": multi_sps,
        "multi_rhs_per_call": nrhs,
        "mixed_solves_per_s": mixed_sps,
        "mixed_f_rel_err_vs_fp64": mixed_err,
    }
    out.update(extra)
    if not args.no_cpu:
        cb, f_ref, s_ref = cpu_baseline(g, s, fb, args.cpu_budget)
        out["cpu_baseline"] = cb
        ef = float(np.linalg.norm(f_gpu.flat() - f_ref) / np.linalg.norm(f_ref))
        es = float(np.linalg.norm(s_gpu.flat() - s_ref) / np.linalg.norm(s_ref))
        out["rel_err_f"], out["rel_err_s"] = ef, es
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
