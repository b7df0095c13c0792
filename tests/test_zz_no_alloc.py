"""GPU test of ilh_dirichlet_solve_noalloc, the allocation-free form of
ilh_dirichlet_solve (its work buffers live in the library's static device
scratch array).  Called through the C ABI on device buffers, it must equal
ilh_dirichlet_solve bit for bit on two calls in a row, the second on a
smaller problem (a 1024^2 grid with 1885 points, then a 205^2 grid with 112
points), and again on the first problem after the smaller one used the same
scratch.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problem(ctx, g, s, fb):
    import ilhip
    from ilhip import _lib as L
    cache = ilhip.SurfaceScalarCache(s, g, ctx=ctx)
    ng = (g.NX - 1) * (g.NY - 1)
    d_fp = ilhip.DeviceArray.from_host(ctx, fb)
    d_fm = ilhip.DeviceArray.from_host(ctx, 0.25 * np.roll(fb, 3))
    outs = [ilhip.DeviceArray(ctx, ng), ilhip.DeviceArray(ctx, s.N), ilhip.DeviceArray(ctx, ng),
            ilhip.DeviceArray(ctx, s.N)]
    # the ordinary entry point: builds the plan's solve data, then the reference result
    L.call("ilh_dirichlet_solve", cache.h, d_fp.ptr, d_fm.ptr, outs[0].ptr, outs[1].ptr, L.MEM_DEVICE)
    return cache, d_fp, d_fm, outs


def _noalloc(ctx, prob):
    from ilhip import _lib as L
    cache, d_fp, d_fm, outs = prob
    for a in (outs[2], outs[3]):
        a.upload(np.full(a.n, np.nan))
    L.call("ilh_dirichlet_solve_noalloc", cache.h, d_fp.ptr, d_fm.ptr, outs[2].ptr, outs[3].ptr)
    ctx.synchronize()
    f_ref, s_ref, f, s = (a.download() for a in outs)
    assert np.isfinite(f).all() and np.isfinite(s).all()
    assert np.array_equal(f, f_ref)
    assert np.array_equal(s, s_ref)
    return f, s


def test_noalloc_matches_dirichlet_solve(ctx):
    from ilhip import geometry as G
    from ilhip import _lib as L
    g1, s1, fb1 = G.config_headline()
    g2, s2, fb2 = G.config_plain()
    print("noalloc: 1024^2 N=%d, then 205^2 N=%d" % (s1.N, s2.N), flush=True)
    big = _problem(ctx, g1, s1, fb1)
    small = _problem(ctx, g2, s2, fb2)
    _noalloc(ctx, big)
    _noalloc(ctx, small)
    _noalloc(ctx, big)
    # fb_minus = NULL
    cache, d_fp, _, outs = small
    L.call("ilh_dirichlet_solve", cache.h, d_fp.ptr, None, outs[0].ptr, outs[1].ptr, L.MEM_DEVICE)
    L.call("ilh_dirichlet_solve_noalloc", cache.h, d_fp.ptr, None, outs[2].ptr, outs[3].ptr)
    ctx.synchronize()
    assert np.array_equal(outs[0].download(), outs[2].download())
    assert np.array_equal(outs[1].download(), outs[3].download())


def test_noalloc_errors(ctx):
    import ilhip
    from ilhip import geometry as G
    from ilhip import _lib as L
    g, s, fb = G.config_plain()
    cache = ilhip.SurfaceScalarCache(s, g, ctx=ctx)
    d = ilhip.DeviceArray.from_host(ctx, fb)
    f = ilhip.DeviceArray(ctx, (g.NX - 1) * (g.NY - 1))
    sv = ilhip.DeviceArray(ctx, s.N)
    # before the plan's solve data exist
    with pytest.raises(L.StateError):
        L.call("ilh_dirichlet_solve_noalloc", cache.h, d.ptr, None, f.ptr, sv.ptr)
    # prepared without interior data: fb_minus needs a preparing call with it
    L.call("ilh_dirichlet_solve", cache.h, d.ptr, None, f.ptr, sv.ptr, L.MEM_DEVICE)
    with pytest.raises(L.StateError):
        L.call("ilh_dirichlet_solve_noalloc", cache.h, d.ptr, d.ptr, f.ptr, sv.ptr)
    L.call("ilh_dirichlet_solve_noalloc", cache.h, d.ptr, None, f.ptr, sv.ptr)
    with pytest.raises(L.ArgumentError):
        L.call("ilh_dirichlet_solve_noalloc", None, d.ptr, None, f.ptr, sv.ptr)
    # a grid whose spectra exceed the 32 MiB scratch
    gb = G.GridSpec(2050, 2050, 1025, 1025, 4.0 / 2048)
    big = ilhip.SurfaceScalarCache(s, gb, ctx=ctx)
    db = ilhip.DeviceArray.from_host(ctx, fb)
    fbig = ilhip.DeviceArray(ctx, 2049 * 2049)
    L.call("ilh_dirichlet_solve", big.h, db.ptr, None, fbig.ptr, sv.ptr, L.MEM_DEVICE)
    with pytest.raises(L.ArgumentError):
        L.call("ilh_dirichlet_solve_noalloc", big.h, db.ptr, None, fbig.ptr, sv.ptr)
