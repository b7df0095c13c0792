"""Seeded synthetic geometry and grids for the immersed-layer workloads.

Surface geometry is an explicit input of both the HIP library and its checker:
point coordinates, unit normals and panel lengths ``ds`` (the reference takes
them from RigidBodyTools' ``points``/``normals``/``areas``, src/tools.jl:16-38).
Circles follow RigidBodyTools' ``Circle(a, Δs)`` convention: ``N = ⌊2πa/Δs⌋``
points at ``θ_p = 2π(p-1)/N``, radial normals and ``ds = 2a·tan(π/N)``.
Other curves are sampled at equal arclength from their analytic form.
"""
from dataclasses import dataclass

import numpy as np


@dataclass
class GridSpec:
    """CartesianGrids ``PhysicalGrid`` essentials: ``size(g) = (NX, NY)`` dual-node
    counts, ``origin(g) = (I0x, I0y)`` (1-based) and the cell size ``dx``."""
    NX: int
    NY: int
    I0x: int
    I0y: int
    dx: float

    def primal_shape(self):
        return (self.NX - 1, self.NY - 1)

    def dual_shape(self):
        return (self.NX, self.NY)

    def x_primal(self):
        return (np.arange(self.NX - 1) + 1 - self.I0x) * self.dx

    def y_primal(self):
        return (np.arange(self.NY - 1) + 1 - self.I0y) * self.dx

    def x_dual(self):
        return (np.arange(self.NX) + 0.5 - self.I0x) * self.dx

    def y_dual(self):
        return (np.arange(self.NY) + 0.5 - self.I0y) * self.dx


@dataclass
class Surface:
    X: np.ndarray
    Y: np.ndarray
    nx: np.ndarray
    ny: np.ndarray
    ds: np.ndarray
    body: np.ndarray          # body index of each point

    @property
    def N(self):
        return len(self.X)

    def theta(self, cx=0.0, cy=0.0):
        return np.arctan2(self.Y - cy, self.X - cx)


def concat(surfs):
    cat = lambda k: np.concatenate([getattr(s, k) for s in surfs])
    body = np.concatenate([np.full(s.N, i) for i, s in enumerate(surfs)])
    return Surface(cat("X"), cat("Y"), cat("nx"), cat("ny"), cat("ds"), body)


def circle(a, ds_target, cx=0.0, cy=0.0):
    N = int(np.floor(2 * np.pi * a / ds_target))
    th = 2 * np.pi * np.arange(N) / N
    ds = np.full(N, 2 * a * np.tan(np.pi / N))
    return Surface(cx + a * np.cos(th), cy + a * np.sin(th), np.cos(th), np.sin(th), ds, np.zeros(N, int))


def rectangle(a, b, ds_target, cx=0.0, cy=0.0):
    """Rectangle of half-lengths a (x) and b (y), as RigidBodyTools'
    ``Rectangle(a, b, Δs)`` polygon: each side split into round(length/Δs)
    equal panels, one point at each panel midpoint with the side's outward
    normal (214 points for Rectangle(0.5, 0.25, 1.4·0.01), the count printed
    in examples/stokes.ipynb)."""
    corners = [(a, -b), (a, b), (-a, b), (-a, -b)]
    normals = [(1.0, 0.0), (0.0, 1.0), (-1.0, 0.0), (0.0, -1.0)]
    X, Y, NX, NY, DS = [], [], [], [], []
    for k in range(4):
        (x0, y0), (x1, y1) = corners[k], corners[(k + 1) % 4]
        length = np.hypot(x1 - x0, y1 - y0)
        n = max(1, int(round(length / ds_target)))
        t = (np.arange(n) + 0.5) / n
        X.append(cx + x0 + t * (x1 - x0))
        Y.append(cy + y0 + t * (y1 - y0))
        NX.append(np.full(n, normals[k][0]))
        NY.append(np.full(n, normals[k][1]))
        DS.append(np.full(n, length / n))
    cat = np.concatenate
    return Surface(cat(X), cat(Y), cat(NX), cat(NY), cat(DS), np.zeros(sum(len(x) for x in X), int))


def _sample_curve(rfun, drfun, N, cx, cy, fine=200000):
    """Equal-arclength sampling of a closed curve given as r(t) = (x(t), y(t))."""
    t = np.linspace(0, 2 * np.pi, fine + 1)
    dx, dy = drfun(t)
    sp = np.hypot(dx, dy)
    s = np.concatenate([[0], np.cumsum(0.5 * (sp[1:] + sp[:-1]) * np.diff(t))])
    L = s[-1]
    s_pts = L * np.arange(N) / N
    s_mid = L * (np.arange(N) + 0.5) / N
    tp = np.interp(s_pts, s, t)
    x, y = rfun(tp)
    dxp, dyp = drfun(tp)
    nrm = np.hypot(dxp, dyp)
    # outward normal for a counter-clockwise curve: (dy, -dx)/|r'|
    nx, ny = dyp / nrm, -dxp / nrm
    ds = np.full(N, L / N)
    del s_mid
    return Surface(cx + x, cy + y, nx, ny, ds, np.zeros(N, int)), L


def ellipse(a, b, ds_target, cx=0.0, cy=0.0):
    rf = lambda t: (a * np.cos(t), b * np.sin(t))
    drf = lambda t: (-a * np.sin(t), b * np.cos(t))
    _, L = _sample_curve(rf, drf, 16, 0, 0)
    N = max(8, int(np.floor(L / ds_target)))
    return _sample_curve(rf, drf, N, cx, cy)[0]


def star(r0, bk, ds_target, cx=0.0, cy=0.0, N=None):
    """r(θ) = r0 (1 + Σ_{k=2..} b_k cos kθ)."""
    ks = np.arange(2, 2 + len(bk))
    bk = np.asarray(bk)

    def R(t):
        return r0 * (1 + np.sum(bk[:, None] * np.cos(np.outer(ks, t)), axis=0))

    def dR(t):
        return r0 * (-np.sum((bk * ks)[:, None] * np.sin(np.outer(ks, t)), axis=0))

    rf = lambda t: (R(t) * np.cos(t), R(t) * np.sin(t))
    drf = lambda t: (dR(t) * np.cos(t) - R(t) * np.sin(t), dR(t) * np.sin(t) + R(t) * np.cos(t))
    if N is None:
        _, L = _sample_curve(rf, drf, 16, 0, 0)
        N = max(8, int(np.floor(L / ds_target)))
    return _sample_curve(rf, drf, N, cx, cy)[0]


def fourier_data(theta, rng, K=8):
    """f_b(θ) = Σ_{k=1..K} a_k cos(kθ + φ_k), a_k ~ N(0, 1/k²), φ_k ~ U(0, 2π)."""
    k = np.arange(1, K + 1)
    a = rng.standard_normal(K) / k
    ph = rng.uniform(0, 2 * np.pi, K)
    return np.sum(a[:, None] * np.cos(np.outer(k, theta) + ph[:, None]), axis=0), a, ph


# ----------------------------------------------------------------- workloads
def config_plain(seed=1):
    """Plain FP64 CPU-runnable case: Δx=0.02 on [-2,2]² (NX=NY=206, I0=103),
    circle r=0.5 at the origin, points at 1.4Δx, Dirichlet f_b on the exterior."""
    g = GridSpec(206, 206, 103, 103, 0.02)
    s = circle(0.5, 1.4 * g.dx)
    rng = np.random.default_rng(seed)
    fb, a, ph = fourier_data(s.theta(), rng)
    return g, s, fb


def config_microbench(seed=2, npts_per_curve=2048):
    """2048² Nodes{Primal} on [-2,2]² with 4 random star curves of 2048 points."""
    dx = 4.0 / 2048
    g = GridSpec(2049, 2049, 1025, 1025, dx)
    rng = np.random.default_rng(seed)
    surfs = []
    for _ in range(4):
        r0 = rng.uniform(0.2, 0.4)
        bk = rng.uniform(-0.1, 0.1, 5)
        c = rng.uniform(-1.3, 1.3, 2)
        surfs.append(star(r0, bk, None, c[0], c[1], N=npts_per_curve))
    return g, concat(surfs)


def config_batched(seed=3, nprob=64, sizes=(256, 384, 512, 768, 1024), dx=0.01, spacing=1.4):
    """Independent Dirichlet problems for batched solves: Nodes{Primal} n² with
    n drawn from `sizes`, dx fixed (the domain is n·dx), 1-3 bodies (circle,
    ellipse of aspect 1-4, star) placed with >= 10 dx clearance from each other
    and the grid edge, 200-3000 points in total at `spacing`·dx (fewer where
    the target does not fit).  Returns a list of (grid, surface, f_b)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nprob):
        n = int(rng.choice(sizes))
        g = GridSpec(n + 1, n + 1, n // 2 + 1, n // 2 + 1, dx)
        half = 0.5 * n * dx
        clear = 10 * dx
        nb = int(rng.integers(1, 4))
        target = rng.uniform(200, 3000)
        # body radius from its share of the points (perimeter ~ 2 pi r)
        r_share = target * spacing * dx / (2 * np.pi * nb)
        bodies, placed = [], []
        for b in range(nb):
            kind = int(rng.integers(0, 3))
            r = min(r_share, 0.45 * (half - clear))
            for _try in range(200):
                c = rng.uniform(-half + clear + 1.1 * r, half - clear - 1.1 * r, 2)
                if all(np.hypot(*(c - pc)) > 1.1 * (r + pr) + clear for pc, pr in placed):
                    break
                r *= 0.97
            else:
                continue
            placed.append((c, r))
            if kind == 0:
                bodies.append(circle(r, spacing * dx, c[0], c[1]))
            elif kind == 1:
                asp = rng.uniform(1, 4)
                bodies.append(ellipse(r, r / asp, spacing * dx, c[0], c[1]))
            else:
                bodies.append(star(r / 1.1, rng.uniform(-0.1, 0.1, 5), spacing * dx, c[0], c[1]))
        surf = concat(bodies)
        fb, _, _ = fourier_data(np.arctan2(surf.Y, surf.X), rng)
        out.append((g, surf, fb))
    return out


def config_heat(seed=4, n=2048, spacing=1.3):
    """Unbounded heat conduction: n² Nodes{Primal} on [-2,2]² with 3 seeded star
    bodies (about 2000 points each at `spacing`·dx) and an initial field of 4
    Gaussian bumps (widths >= 4 dx) away from the bodies.  Returns (grid,
    surface, T0 as an (n, n) array)."""
    dx = 4.0 / n
    g = GridSpec(n + 1, n + 1, n // 2 + 1, n // 2 + 1, dx)
    rng = np.random.default_rng(seed)
    centres = [(-1.0, -0.95), (1.0, -0.95), (0.0, 0.95)]
    bodies = []
    for cx, cy in centres:
        r0 = rng.uniform(0.6, 0.7)
        bodies.append(star(r0, rng.uniform(-0.1, 0.1, 5), spacing * dx, cx + rng.uniform(-0.05, 0.05),
                           cy + rng.uniform(-0.05, 0.05)))
    s = concat(bodies)
    xs = (np.arange(n) + 1 - g.I0x) * dx
    Xg, Yg = np.meshgrid(xs, xs, indexing="ij")
    T0 = np.zeros((n, n))
    for bx, by in [(-1.0, 0.6), (1.0, 0.6), (0.0, -0.6), (0.0, -1.7)]:
        w = rng.uniform(4, 12) * dx
        T0 += np.exp(-((Xg - bx - rng.uniform(-0.1, 0.1)) ** 2 + (Yg - by - rng.uniform(-0.1, 0.1)) ** 2) / (2 * w * w))
    return g, s, T0


def config_headline(seed=5, spacing=1.2):
    """1024² Nodes{Primal} on [-2,2]² (Δx=4/1024): an ellipse of aspect 2 and a
    random star curve, about 1000 points each at ``spacing``·Δx."""
    dx = 4.0 / 1024
    g = GridSpec(1025, 1025, 513, 513, dx)
    rng = np.random.default_rng(seed)
    e = ellipse(0.95, 0.475, spacing * dx, -1.0, 0.0)
    bk = rng.uniform(-0.1, 0.1, 5)
    st = star(0.62, bk, spacing * dx, 1.0, 0.0)
    s = concat([e, st])
    fb, _, _ = fourier_data(s.theta(), rng)
    return g, s, fb


def config_stokes():
    """The Stokes flow example of test/literate/stokes.jl:168-185: Δx = 0.01 on
    [-2,2]² (NX = NY = 406 dual nodes, origin 203, the x_grid of -2.02..2.02
    the reference's notebooks print for this grid), Rectangle(0.5, 0.25) at
    Δs = 1.4Δx (214 points), translated to the right at unit velocity:
    exterior surface velocity (1, 0), interior (0, 0)."""
    dx = 0.01
    g = GridSpec(406, 406, 203, 203, dx)
    s = rectangle(0.5, 0.25, 1.4 * dx)
    vplus = np.concatenate([np.ones(s.N), np.zeros(s.N)])
    return g, s, vplus


def square_grid(n_primal, L=4.0, margin=0):
    """Grid whose Nodes{Primal} is n_primal² covering about [-L/2, L/2]²."""
    dx = L / n_primal
    I0 = n_primal // 2 + 1
    return GridSpec(n_primal + 1, n_primal + 1, I0, I0, dx)
