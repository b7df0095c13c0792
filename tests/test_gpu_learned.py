"""The learned Schur preconditioner for moving bodies on the GPU.

* features of a two-body geometry equal the CPU restatement's (FP32 rounding);
* the BF16-MFMA forward and backward kernels give the loss and every
  parameter gradient of the CPU FP64 restatement (oracle/learned.py, itself
  pinned by finite differences) evaluated with the same BF16 roundings:
  loss to 1e-5, each gradient block to 2e-3 (relative L2; the FP32
  accumulation can move an activation across a BF16 rounding boundary);
* trained on the GPU on the exact inverses of ten geometries (about 19k
  plain-method solves), it preconditions GMRES on S at moved geometries it
  was not trained on: far fewer iterations than no preconditioner, no
  factorisation, and solves that match the oracle's direct solve;
* the batched moving-body solve with the learned band equals the single
  solves bit for bit;
* a briefly trained band leaves GMRES more than 64 steps (the Arnoldi
  launches' later batches of basis vectors, the final triangular solve read
  from memory rather than LDS): the solve still equals the oracle's, and the
  batched form the single one bit for bit.
"""
import time

import numpy as np
import pytest

import oracle as O
from conftest import rel_err

pytestmark = pytest.mark.gpu


def _moved(s, angle, heave):
    from ilhip.geometry import Surface
    c, n = np.cos(angle), np.sin(angle)
    return Surface(c * s.X - n * s.Y, n * s.X + c * s.Y + heave, c * s.nx - n * s.ny, n * s.nx + c * s.ny, s.ds,
                   s.body)


@pytest.fixture(scope="module")
def trained(ctx):
    """A model trained on ten geometries of the 1024^2 workload: the ellipse +
    star pitched and heaved at seeded amplitudes, and two other star shapes."""
    import ilhip
    from ilhip import geometry as G
    g, s0, fb = G.config_headline()
    model = ilhip.LearnedPreconditioner(ctx, half_band=16, seed=7)
    rng = np.random.default_rng(17)
    geoms = [_moved(s0, rng.uniform(-0.3, 0.3), rng.uniform(-0.2, 0.2)) for _ in range(8)]
    geoms += [G.config_headline(seed=sd)[1] for sd in (6, 7)]
    t0 = time.perf_counter()
    for sk in geoms:
        c = ilhip.SurfaceScalarCache(sk, g, ctx=ctx)
        model.add_geometry(c)
        c.close()
    t_data = time.perf_counter() - t0
    t0 = time.perf_counter()
    l1 = model.train(2000, batch=2048, lr=1e-3, seed=1)
    l2 = model.train(1000, batch=2048, lr=2e-4, seed=2)
    t_train = time.perf_counter() - t0
    print("learned: %d training rows in %.2f s, 3000 Adam steps in %.2f s, loss %.4g -> %.4g"
          % (model.rows, t_data, t_train, l1, l2), flush=True)
    return ilhip, G, g, s0, fb, model, l2


def test_learned_features_match_oracle(ctx):
    import ilhip
    import learned as LN
    from ilhip import geometry as G
    g, s, fb = G.config_headline()
    c = ilhip.SurfaceScalarCache(s, g, ctx=ctx)
    model = ilhip.LearnedPreconditioner(ctx, half_band=16, seed=1)
    F, _ = model.predict(c, band=False)
    ref = LN.features(s.X, s.Y, s.nx, s.ny, s.ds, g.dx, g.I0x, g.I0y, c.body_offsets, 16)
    assert F.shape == ref.shape
    assert np.abs(F - ref).max() < 2e-5 * max(1.0, np.abs(ref).max())
    model.close()
    c.close()


def test_learned_gradients_match_cpu_fp64(ctx):
    import ilhip
    import learned as LN
    model = ilhip.LearnedPreconditioner(ctx, half_band=16, seed=3)
    K, W, F, Fp, H, Op, n = model.dims
    rng = np.random.default_rng(4)
    rows = 96
    X = np.zeros((rows, Fp))
    X[:, :F] = rng.standard_normal((rows, F))
    T = np.zeros((rows, Op), dtype=np.float32)
    T[:, :W] = rng.standard_normal((rows, W))
    bits = LN.bf16_bits(X)
    loss, g = model.loss_grad(bits, T)
    p = model.params()
    loss_r, g_r = LN.loss_grad(p, model.dims, LN.bf16(X), T.astype(np.float64), bf16_round=True)
    print("gradient check: loss", loss, loss_r, flush=True)
    assert abs(loss - loss_r) < 1e-5 * abs(loss_r)
    o = 0
    shapes = [(Fp, H), (H,), (H, H), (H,), (H, H), (H,), (H, Op), (Op,)]
    for k, sh in enumerate(shapes):
        sz = int(np.prod(sh))
        e = rel_err(g[o:o + sz], g_r[o:o + sz])
        print("  block", k, sh, "rel L2", e, flush=True)
        assert e < 2e-3, (k, e)
        o += sz
    model.close()


def test_learned_preconditioner_gain(ctx, trained):
    """Pitched/heaved geometries of the workload, none of them a training
    geometry: GMRES with the learned band against no preconditioner (the same
    S, scipy on the host) and against the FP32 inverse of the unmoved geometry;
    every solve matches the oracle's direct solve."""
    import scipy.sparse.linalg as spla
    ilhip, G, g, s0, fb, model, loss = trained
    assert loss < 0.05
    og = O.grid(g.NX, g.NY, g.I0x, g.I0y, g.dx)
    cl = ilhip.SurfaceScalarCache(s0, g, ctx=ctx)   # learned band
    cl.set_preconditioner(model)
    cr = ilhip.SurfaceScalarCache(s0, g, ctx=ctx)   # recycled FP32 inverse of s0
    pl, pr = ilhip.DirichletPoisson(cl), ilhip.DirichletPoisson(cr)
    pr.solve_moving(fb, tol=1e-12, max_iter=60)    # factors s0: the reference
    its_l, its_r = [], []
    for k in range(1, 5):
        sk = _moved(s0, 0.05 * k, 0.03 * k)
        cl.update_points(sk)
        cr.update_points(sk)
        f1, s1, it1 = pl.solve_moving(fb, tol=1e-12, max_iter=80)
        f2, s2, it2 = pr.solve_moving(fb, tol=1e-12, max_iter=80)
        its_l.append(it1)
        its_r.append(it2)
        fr, sr = O.dirichlet_solve(og, O.Geometry(sk.X, sk.Y, sk.nx, sk.ny, sk.ds), fb)
        print("step", k, "learned", it1, "recycled", it2, "f", rel_err(f1.flat(), fr), "s", rel_err(s1.flat(), sr),
              flush=True)
        assert it1 <= 80, "the learned band fell back to a factorisation"
        assert rel_err(f1.flat(), fr) < 1e-10 and rel_err(s1.flat(), sr) < 1e-9
        assert rel_err(f2.flat(), fr) < 1e-10
    # no preconditioner on the last geometry: GMRES on the same S on the host
    S = O.schur(og, O.Geometry(sk.X, sk.Y, sk.nx, sk.ny, sk.ds), O.S_RTLINVR)
    cnt = [0]
    b = np.random.default_rng(0).standard_normal(S.shape[0])
    spla.gmres(S, b, rtol=1e-12, restart=600, maxiter=1, callback=lambda r: cnt.__setitem__(0, cnt[0] + 1),
               callback_type="pr_norm")
    print("iterations: none", cnt[0], "learned", its_l, "recycled FP32 inverse", its_r, flush=True)
    assert max(its_l) * 4 < cnt[0]
    # wall time of one step with each preconditioner (same geometry, warm)
    for name, prob in (("learned", pl), ("recycled", pr)):
        t0 = time.perf_counter()
        for _ in range(5):
            prob.solve_moving(fb, tol=1e-12, max_iter=80)
        print(name, "ms per solve %.3f" % ((time.perf_counter() - t0) / 5 * 1e3), flush=True)
    cl.close()
    cr.close()


def test_learned_batched_matches_single(ctx, trained):
    ilhip, G, g, s0, fb, model, loss = trained
    nb = 4
    ng = (g.NX - 1) * (g.NY - 1)
    cb = [ilhip.SurfaceScalarCache(s0, g, ctx=ctx) for _ in range(nb)]
    cs = [ilhip.SurfaceScalarCache(s0, g, ctx=ctx) for _ in range(nb)]
    for c in cb + cs:
        c.set_preconditioner(model)
    fp = [ilhip.DeviceArray.from_host(ctx, fb * (1 + 0.1 * k)) for k in range(nb)]
    fd = [ilhip.DeviceArray(ctx, ng) for _ in range(nb)]
    sd = [ilhip.DeviceArray(ctx, s0.N) for _ in range(nb)]
    for k in range(nb):
        sk = _moved(s0, 0.07 * (k + 1), -0.02 * k)
        cb[k].update_points(sk)
        cs[k].update_points(sk)
    its = ilhip.dirichlet_solve_moving_batched(cb, fp, fd, sd, tol=1e-12, max_iter=80)
    ctx.synchronize()
    for k in range(nb):
        f1, s1, it1 = ilhip.DirichletPoisson(cs[k]).solve_moving(fb * (1 + 0.1 * k), tol=1e-12, max_iter=80)
        assert its[k] == it1
        assert np.array_equal(fd[k].download(), f1.flat()) and np.array_equal(sd[k].download(), s1.flat())
    for c in cb + cs:
        c.close()


def test_learned_long_gmres(ctx):
    """A band trained for 300 Adam steps on three geometries: more than 64 but
    at most 200 GMRES steps at a moved geometry; the solve equals the oracle's
    direct solve, and two batched trajectories their single solves bit for
    bit."""
    import ilhip
    from ilhip import geometry as G
    g, s0, fb = G.config_headline()
    og = O.grid(g.NX, g.NY, g.I0x, g.I0y, g.dx)
    model = ilhip.LearnedPreconditioner(ctx, half_band=16, seed=7)
    rng = np.random.default_rng(17)
    for _ in range(3):
        c = ilhip.SurfaceScalarCache(_moved(s0, rng.uniform(-0.3, 0.3), rng.uniform(-0.2, 0.2)), g, ctx=ctx)
        model.add_geometry(c)
        c.close()
    model.train(300, batch=2048, lr=1e-3, seed=1)
    sk = _moved(s0, 0.006, 0.025)
    fbs = [fb, 0.5 * fb - 0.1]
    ng = (g.NX - 1) * (g.NY - 1)
    cb = [ilhip.SurfaceScalarCache(s0, g, ctx=ctx) for _ in fbs]
    cs = [ilhip.SurfaceScalarCache(s0, g, ctx=ctx) for _ in fbs]
    for c in cb + cs:
        c.set_preconditioner(model)
        c.update_points(sk)
    fp = [ilhip.DeviceArray.from_host(ctx, x) for x in fbs]
    fd = [ilhip.DeviceArray(ctx, ng) for _ in fbs]
    sd = [ilhip.DeviceArray(ctx, s0.N) for _ in fbs]
    its = ilhip.dirichlet_solve_moving_batched(cb, fp, fd, sd, tol=1e-12, max_iter=200)
    ctx.synchronize()
    fr, sr = O.dirichlet_solve(og, O.Geometry(sk.X, sk.Y, sk.nx, sk.ny, sk.ds), fbs[0])
    for k in range(len(fbs)):
        f1, s1, it1 = ilhip.DirichletPoisson(cs[k]).solve_moving(fbs[k], tol=1e-12, max_iter=200)
        print("long GMRES: trajectory", k, "steps", it1, "batched", its[k], flush=True)
        assert 64 < it1 <= 200 and its[k] == it1
        assert np.array_equal(fd[k].download(), f1.flat()) and np.array_equal(sd[k].download(), s1.flat())
        if k == 0:
            print("long GMRES: f", rel_err(f1.flat(), fr), "s", rel_err(s1.flat(), sr), flush=True)
            assert rel_err(f1.flat(), fr) < 1e-10 and rel_err(s1.flat(), sr) < 1e-9
    for c in cb + cs:
        c.close()
    model.close()
