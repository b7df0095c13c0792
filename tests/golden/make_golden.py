"""Write the golden fixtures of tests/golden/ from the CPU FP64 oracle.

The reference stores no arrays, so these fixtures record the oracle's results
on seeded synthetic inputs (self-consistency of oracle and HIP path); the
Neumann case is the reference's own published example (Circle(0.25) inside
Square(1), dx = 0.01 on [-2,2]^2, examples/neumann.ipynb cell 37), whose
added mass M_x = 0.22649277527914002 the oracle reproduces; the Stokes case
is the reference's Stokes example (examples/stokes.ipynb: 214 points), whose
outputs are published only as plots (traction, velocity and streamfunction
probes are recorded).

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "immersedlayers.jl_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as O  # noqa: E402
from ilhip import geometry as G  # noqa: E402


def square(a, dst):
    n = int(round(2 * a / dst))
    h = 2 * a / n
    s = -a + h * (np.arange(n) + 0.5)
    X = np.concatenate([s, np.full(n, a), s[::-1], np.full(n, -a)])
    Y = np.concatenate([np.full(n, -a), s, np.full(n, a), s[::-1]])
    nx = np.concatenate([np.zeros(n), np.ones(n), np.zeros(n), -np.ones(n)])
    ny = np.concatenate([-np.ones(n), np.zeros(n), np.ones(n), np.zeros(n)])
    return G.Surface(X, Y, nx, ny, np.full(4 * n, h), np.zeros(4 * n, int))


def neumann_case():
    dx = 0.01
    g = G.GridSpec(406, 406, 203, 203, dx)
    s = G.concat([square(1.0, 1.4 * dx), G.circle(0.25, 1.4 * dx)])
    vp = np.where(s.body == 1, s.nx, 0.0)
    return g, s, vp


def main():
    gs, s, fb = G.config_plain()
    og = O.grid(gs.NX, gs.NY, gs.I0x, gs.I0y, gs.dx)
    geo = O.Geometry(s.X, s.Y, s.nx, s.ny, s.ds)
    f, sv = O.dirichlet_solve(og, geo, fb)
    np.savez_compressed(os.path.join(HERE, "dirichlet_plain.npz"), grid=np.array([gs.NX, gs.NY, gs.I0x, gs.I0y]),
                        dx=gs.dx, X=s.X, Y=s.Y, nx=s.nx, ny=s.ny, ds=s.ds, fb=fb, f=f, s=sv)
    g, s, vp = neumann_case()
    og = O.grid(g.NX, g.NY, g.I0x, g.I0y, g.dx)
    geo = O.Geometry(s.X, s.Y, s.nx, s.ny, s.ds)
    f, df, st, dsv = O.neumann_solve(og, geo, vp)
    m = s.body == 1
    M = -np.array([np.sum(df[m] * s.nx[m] * s.ds[m]), np.sum(df[m] * s.ny[m] * s.ds[m])])
    probe = np.arange(0, f.size, 997)
    np.savez_compressed(os.path.join(HERE, "neumann_square_circle.npz"), grid=np.array([g.NX, g.NY, g.I0x, g.I0y]),
                        dx=g.dx, X=s.X, Y=s.Y, nx=s.nx, ny=s.ny, ds=s.ds, body=s.body, vp=vp, df=df, dsv=dsv, M=M,
                        f_norm=np.linalg.norm(f), s_norm=np.linalg.norm(st), probe=probe, f_probe=f[probe],
                        s_probe=st[probe])
    print("M =", M)
    # the Stokes example of test/literate/stokes.jl (Rectangle(0.5, 0.25) at
    # 1.4 dx, 214 points on the 406^2 grid, translated at unit speed)
    g, s, vp = G.config_stokes()
    og = O.grid(g.NX, g.NY, g.I0x, g.I0y, g.dx)
    geo = O.Geometry(s.X, s.Y, s.nx, s.ny, s.ds)
    v, st, sig = O.stokes_solve(og, geo, vp)
    Cf = O.surface_filter_vector(og, geo)
    sig2 = Cf @ (Cf @ sig)
    pv = np.arange(0, v.size, 1009)
    ps = np.arange(0, st.size, 503)
    # create_CL2invCT's smallest singular direction (the traction of a uniform
    # streamfunction; condition number 1.7e7 here), after the C^2 filter: the
    # direction in which two FP64 solves may differ by cond * eps
    _, _, Vt = np.linalg.svd(O.cl2invct(og, geo))
    null2 = Cf @ (Cf @ Vt[-1])
    np.savez_compressed(os.path.join(HERE, "stokes_rectangle.npz"), grid=np.array([g.NX, g.NY, g.I0x, g.I0y]),
                        dx=g.dx, X=s.X, Y=s.Y, nx=s.nx, ny=s.ny, ds=s.ds, vp=vp, sigma=sig, sigma2=sig2,
                        null2=null2 / np.linalg.norm(null2),
                        v_norm=np.linalg.norm(v), probe_v=pv, v_probe=v[pv], probe_s=ps, s_probe=st[ps],
                        force=np.array([np.sum(sig2[:s.N] * s.ds), np.sum(sig2[s.N:] * s.ds)]))
    print("Stokes force (C^2-filtered traction) =", np.sum(sig2[:s.N] * s.ds), np.sum(sig2[s.N:] * s.ds))


if __name__ == "__main__":
    main()
