"""CPU tests of the C ABI library and the host-side API (no GPU compute).

* libimmersedhip.so loads and exports every entry point include/immersedhip.h
  declares, and the ctypes binding covers them all.
* Without a usable GPU, context creation fails with an error code and message
  (no crash); the host containers raise DimensionMismatch like the reference
  (test/surface_ops.jl:524).
* The golden fixtures of tests/golden/ are reproduced by the oracle.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "immersedhip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ilh_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_header():
    import ilhip
    from ilhip import _lib
    lib = ilhip.load_library()
    names = _declared()
    assert len(names) >= 38
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.SIGNATURES), set(names) ^ set(_lib.SIGNATURES)
    assert b"gfx950" in lib.ilh_version()


def test_ctx_without_gpu_reports_error():
    from ilhip import _lib
    lib = _lib.load()
    h = C.c_void_p()
    rc = lib.ilh_ctx_create(0, C.byref(h))
    if rc == 0:   # a GPU is present: nothing to check here
        lib.ilh_ctx_destroy(h)
        pytest.skip("GPU present")
    assert rc < 0
    assert len(lib.ilh_last_error()) > 0


def test_null_arguments_are_errors():
    from ilhip import _lib
    lib = _lib.load()
    assert lib.ilh_ctx_create(0, None) == _lib.ERR_ARG
    assert lib.ilh_plan_npts(None) == -1
    assert lib.ilh_regularize(None, 0, None, None, 0) == _lib.ERR_ARG
    assert lib.ilh_plan_destroy(None) == 0


def test_containers_dimension_checks():
    import ilhip
    g = ilhip.PhysicalGrid(20, 18, 10, 9, 0.1)
    assert ilhip.Nodes(ilhip.Primal, g).data.shape == (19, 17)
    assert ilhip.Nodes(ilhip.Dual, g).data.shape == (20, 18)
    e = ilhip.Edges(g)
    assert e.u.shape == (20, 17) and e.v.shape == (19, 18)
    with pytest.raises(ilhip.DimensionMismatch):
        ilhip.Nodes(ilhip.Primal, g, np.zeros((20, 18)))
    with pytest.raises(ilhip.DimensionMismatch):
        ilhip.Edges(g, np.zeros(10))
    with pytest.raises(ilhip.DimensionMismatch):
        ilhip.VectorData(5, np.zeros(9))


def test_geometry_generators():
    from ilhip import geometry as G
    g, s, fb = G.config_plain()
    assert s.N == 112 and g.primal_shape() == (205, 205)
    g, s, fb = G.config_headline()
    assert 1800 < s.N < 2100
    assert np.allclose(np.hypot(s.nx, s.ny), 1.0)
    g, s = G.config_microbench()
    assert s.N == 4 * 2048 and g.primal_shape() == (2048, 2048)


def test_golden_dirichlet_reproduced_by_oracle():
    d = np.load(os.path.join(ROOT, "tests", "golden", "dirichlet_plain.npz"))
    NX, NY, I0x, I0y = (int(v) for v in d["grid"])
    g = O.grid(NX, NY, I0x, I0y, float(d["dx"]))
    geo = O.Geometry(d["X"], d["Y"], d["nx"], d["ny"], d["ds"])
    f, s = O.dirichlet_solve(g, geo, d["fb"])
    assert np.abs(f - d["f"]).max() < 1e-12 * np.abs(d["f"]).max()
    assert np.abs(s - d["s"]).max() < 1e-10 * np.abs(d["s"]).max()


def test_golden_stokes_reproduced_by_oracle():
    """The reference's Stokes example (214 points on the 406^2 grid): the
    oracle reproduces the fixture; the rectangle's symmetry about the x axis
    leaves no lift (F_y = 0) and the constraint E v = v_bar holds."""
    d = np.load(os.path.join(ROOT, "tests", "golden", "stokes_rectangle.npz"))
    NX, NY, I0x, I0y = (int(v) for v in d["grid"])
    g = O.grid(NX, NY, I0x, I0y, float(d["dx"]))
    geo = O.Geometry(d["X"], d["Y"], d["nx"], d["ny"], d["ds"])
    assert geo.N == 214
    v, st, sig = O.stokes_solve(g, geo, d["vp"])
    assert np.abs(sig - d["sigma"]).max() < 1e-9 * np.abs(d["sigma"]).max()
    assert np.abs(v[d["probe_v"]] - d["v_probe"]).max() < 1e-11 * np.abs(d["v_probe"]).max()
    assert abs(d["force"][1]) < 1e-10 * abs(d["force"][0])
    nu = NX * (NY - 1)
    Ev = np.concatenate([O.interpolate(g, O.EDGE_U, geo, v[:nu]), O.interpolate(g, O.EDGE_V, geo, v[nu:])])
    assert np.abs(Ev - 0.5 * d["vp"]).max() < 1e-12


def test_golden_neumann_reproduced_by_oracle():
    d = np.load(os.path.join(ROOT, "tests", "golden", "neumann_square_circle.npz"))
    NX, NY, I0x, I0y = (int(v) for v in d["grid"])
    g = O.grid(NX, NY, I0x, I0y, float(d["dx"]))
    geo = O.Geometry(d["X"], d["Y"], d["nx"], d["ny"], d["ds"])
    f, df, st, dsv = O.neumann_solve(g, geo, d["vp"])
    assert np.abs(df - d["df"]).max() < 1e-10 * np.abs(d["df"]).max()
    assert abs(d["M"][0] - 0.22649277527914002) < 1e-12
