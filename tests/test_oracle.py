"""CPU tests: the FP64 oracle pinned against the reference's own checks,
its published numbers, closed forms and brute force.

Pins (reference paths relative to the reference checkout):
  * Neumann added mass of Circle(0.25) in Square(1), dx = 0.01 on [-2,2]^2:
    M_x = 0.22649277527914002 (examples/neumann.ipynb cell 37, from
    test/literate/neumann.jl:286-287).
  * sum(ds) of Circle(1, 0.014) = 6.283288300933757 (examples/caches.ipynb cell 28);
    integral of x.n over a Circle(0.5, 0.014) = 1.5708993266236526
    (examples/multbodies.ipynb cell 16).
  * partition of unity and adjointness of Rc/Ec, Rf integral (test/tools.jl:100-124).
  * RnTRn, create_CLinvCT, create_GLinvD, create_nRTRn magnitudes and the mask
    area (test/surface_ops.jl:60-77, 184-188).
  * Yang3 moments, LGF closed values G(1,0)=1/4, G(1,1)=1/pi, G(2,0)=1-2/pi,
    G(n,n) = (1/pi) sum_{k<=n} 1/(2k-1), L G = delta, far field.
  * FFT convolution against brute-force convolution; L(L^-1 w) = w.
  * MAC identities D G = L5, D C = 0, C G = 0, D = -G^T.
  * Schur direct assembly against the reference's column loop.
  * Dirichlet constraint E f = fbar and convergence to the exact exterior field.
"""
import numpy as np
import pytest

import oracle as O
from conftest import max_rel, rel_err


def _geo(s):
    return O.Geometry(s.X, s.Y, s.nx, s.ny, s.ds)


def _G():
    from ilhip import geometry as G
    return G


# ------------------------------------------------------------------- DDFs
def test_yang3_values():
    assert abs(O.ddf(O.YANG3, 0.0) - 0.618199929359) < 1e-11
    assert abs(O.ddf(O.YANG3, 1.0) - 0.190900035320) < 1e-11
    assert abs(O.ddf(O.YANG3, 2.0)) < 1e-15
    assert O.ddf(O.YANG3, 2.5) == 0.0


@pytest.mark.parametrize("kind", [O.YANG3, O.ROMA, O.M4PRIME, O.WITCHHAT])
def test_ddf_moments(kind):
    for r in np.linspace(-0.5, 0.5, 11):
        w = np.array([O.ddf(kind, r - j) for j in range(-4, 5)])
        d = np.array([r - j for j in range(-4, 5)])
        assert abs(w.sum() - 1.0) < 1e-14
        assert abs((d * w).sum()) < 1e-14


# -------------------------------------------------------------------- LGF
def test_lgf_closed_values():
    assert O.lgf(0, 0) == 0.0
    assert abs(O.lgf(1, 0) - 0.25) < 1e-15
    assert abs(O.lgf(0, 1) - 0.25) < 1e-15
    assert abs(O.lgf(1, 1) - 1 / np.pi) < 1e-15
    assert abs(O.lgf(2, 0) - (1 - 2 / np.pi)) < 1e-15
    for n in (3, 17, 100, 1000, 2047):
        exact = sum(1.0 / (2 * k - 1) for k in range(1, n + 1)) / np.pi
        assert abs(O.lgf(n, n) - exact) < 4e-15 * exact


def test_lgf_is_discrete_green_function():
    T = O.lgf_table(160, 160)
    a = np.abs(np.arange(-159, 160))
    G = T[a[:, None], a[None, :]]
    LG = G[:-2, 1:-1] + G[2:, 1:-1] + G[1:-1, :-2] + G[1:-1, 2:] - 4 * G[1:-1, 1:-1]
    c = LG.shape[0] // 2
    assert abs(LG[c, c] - 1.0) < 1e-13
    LG[c, c] = 0.0
    assert np.abs(LG).max() < 1e-13


def test_lgf_far_field():
    # G = (1/2pi)(ln r + gamma + 3/2 ln 2) - cos(4 theta)/(24 pi r^2) + O(r^-4)
    for m, n in [(300, 0), (200, 200), (310, 170), (1000, 37)]:
        r2 = m * m + n * n
        th = np.arctan2(n, m)
        ff = (0.5 * np.log(r2) + np.euler_gamma + 1.5 * np.log(2)) / (2 * np.pi) - np.cos(4 * th) / (24 * np.pi * r2)
        assert abs(O.lgf(m, n) - ff) < 5.0 / r2 ** 2


@pytest.mark.parametrize("kind", [O.PRIMAL, O.DUAL])
def test_inverse_laplacian_vs_direct(kind):
    g = O.grid(23, 18, 11, 9, 0.1)
    nx, ny = O.dims(g, kind)
    rng = np.random.default_rng(3)
    w = rng.standard_normal((nx, ny))
    out = O.InverseLaplacian(g, kind)(w.ravel(order="F"))
    ref = O.conv_direct(O.lgf_table(40, 40), w).ravel(order="F") * g.dx ** 2
    assert max_rel(out, ref) < 2e-15 * 10
    lap = O.laplacian(g, kind, out).reshape((nx, ny), order="F")
    assert np.abs(lap[1:-1, 1:-1] - w[1:-1, 1:-1]).max() < 1e-12


# ---------------------------------------------------- regularize/interpolate
def _tools_setup():
    # test/tools.jl: dx = 0.02 on +-5.98, Circle(1, 1.5 dx), Nodes{Dual}
    G = _G()
    g = O.grid(602, 602, 301, 301, 0.02)
    s = G.circle(1.0, 1.5 * 0.02)
    return g, s


def test_partition_of_unity_and_adjointness():
    g, s = _tools_setup()
    geo = _geo(s)
    ones = np.ones(s.N)
    Rc1 = O.regularize(g, O.DUAL, geo, ones)
    assert abs(g.dx ** 2 * Rc1.sum() - s.ds.sum()) < 1e-14            # tools.jl:107
    rng = np.random.default_rng(5)
    u = rng.standard_normal(O.nel(g, O.DUAL))
    phi = rng.standard_normal(s.N)
    lhs = g.dx ** 2 * np.dot(u, O.regularize(g, O.DUAL, geo, phi))
    rhs = np.dot(O.interpolate(g, O.DUAL, geo, u) * phi, s.ds)
    assert abs(lhs - rhs) < 1e-14 * max(1.0, abs(lhs)) * 10          # tools.jl:112
    # Rf integral (tools.jl:119-120): u-component of Rf with f.u = 1
    Rfu = O.regularize(g, O.EDGE_U, geo, ones)
    assert abs(g.dx ** 2 * Rfu.sum() - s.ds.sum()) < 1e-12
    assert abs(g.dx ** 2 * Rfu.sum() - 2 * np.pi) < 1e-3


def test_interpolation_reproduces_linear_fields():
    G = _G()
    g = O.grid(120, 110, 60, 55, 0.03)
    s = G.star(0.8, [0.05, -0.03, 0.02, 0.0, 0.01], 1.4 * 0.03)
    geo = _geo(s)
    gs = G.GridSpec(120, 110, 60, 55, 0.03)
    for kind, xs, ys in ((O.PRIMAL, gs.x_primal(), gs.y_primal()), (O.DUAL, gs.x_dual(), gs.y_dual())):
        X, Y = np.meshgrid(xs, ys, indexing="ij")
        for fld, exp in ((np.ones_like(X), np.ones(s.N)), (X, s.X), (Y, s.Y), (2 * X - 3 * Y, 2 * s.X - 3 * s.Y)):
            out = O.interpolate(g, kind, geo, fld.ravel(order="F"))
            assert np.abs(out - exp).max() < 1e-13


def test_regularize_brute_force():
    # dense R from the DDF on a tiny grid (R_ij,p = ds_p/dx^2 phi(.)phi(.))
    G = _G()
    gs = G.GridSpec(24, 22, 12, 11, 0.1)
    g = O.grid(gs.NX, gs.NY, gs.I0x, gs.I0y, gs.dx)
    s = G.circle(0.5, 0.13)
    f = np.random.default_rng(2).standard_normal(s.N)
    xs, ys = gs.x_primal(), gs.y_primal()
    ref = np.zeros((len(xs), len(ys)))
    for p in range(s.N):
        wx = np.array([O.ddf(O.YANG3, (x - s.X[p]) / gs.dx) for x in xs])
        wy = np.array([O.ddf(O.YANG3, (y - s.Y[p]) / gs.dx) for y in ys])
        ref += s.ds[p] / gs.dx ** 2 * f[p] * np.outer(wx, wy)
    out = O.regularize(g, O.PRIMAL, _geo(s), f).reshape(ref.shape, order="F")
    assert np.abs(out - ref).max() < 1e-13 * np.abs(ref).max()


# ------------------------------------------------------------------ stencils
def test_mac_identities():
    g = O.grid(31, 27, 15, 13, 0.05)
    rng = np.random.default_rng(9)
    npr, nd, ne = O.nel(g, O.PRIMAL), O.nel(g, O.DUAL), O.edge_len(g)
    p = rng.standard_normal(npr)
    s = rng.standard_normal(nd)
    # D G = 5-point Laplacian on interior primal nodes
    DG = O.divergence(g, O.grad(g, p)).reshape((30, 26), order="F")
    L5 = O.laplacian(g, O.PRIMAL, p).reshape((30, 26), order="F")
    assert np.abs(DG[1:-1, 1:-1] - L5[1:-1, 1:-1]).max() < 1e-10
    # D C = 0 and C G = 0 (interior)
    assert np.abs(O.divergence(g, O.curl_n2e(g, s))).max() < 1e-10
    CG = O.curl_e2n(g, O.grad(g, p)).reshape((31, 27), order="F")
    assert np.abs(CG).max() < 1e-10
    # D = -G^T for edge data vanishing on boundary faces
    q = rng.standard_normal(ne)
    Gp = O.grad(g, p)
    q[Gp == 0.0] = 0.0
    assert abs(np.dot(O.divergence(g, q), p) + np.dot(q, Gp)) < 1e-9


# --------------------------------------------------------------------- Schur
def _circle_setup(dx=0.04, a=1.0):
    # test/surface_ops.jl: dx = 0.04 on [-2,2]^2, Circle(1, 1.4 dx)
    G = _G()
    n = int(round(2.0 / dx)) + 2
    g = O.grid(2 * n, 2 * n, n, n, dx)
    s = G.circle(a, 1.4 * dx)
    return g, s


@pytest.mark.parametrize("kind", [O.S_RTLINVR, O.S_CLINVCT, O.S_GLINVD, O.S_NRTRN])
def test_schur_direct_matches_column_loop(kind):
    G = _G()
    g = O.grid(60, 56, 30, 28, 0.05)
    s = G.star(0.6, [0.08, -0.05, 0.03, 0.0, 0.02], 1.4 * 0.05)
    A1 = O.schur(g, _geo(s), kind, method="columns")
    A2 = O.schur(g, _geo(s), kind, method="direct")
    assert max_rel(A2, A1) < 1e-13


def test_surface_ops_magnitudes():
    g, s = _circle_setup()
    geo = _geo(s)
    th = 2 * np.pi * np.arange(s.N) / s.N
    f = np.sin(th - np.pi / 4)
    v = O.normal_interpolate(g, geo, O.regularize_normal(g, geo, f))
    assert abs(v.max() - 11) < 1 and abs(v.min() + 11) < 1                      # surface_ops.jl:66-67
    A = O.schur(g, geo, O.S_CLINVCT, scale=g.dx)
    assert abs(np.abs(np.linalg.eigvals(A)).max() - 0.2) < 0.1                    # :71
    A = O.schur(g, geo, O.S_GLINVD, scale=g.dx)
    assert abs(np.abs(np.linalg.eigvals(A)).max() - 0.45) < 0.1                   # :74
    A = O.schur(g, geo, O.S_NRTRN)
    assert abs(np.linalg.svd(A, compute_uv=False).max() - 11) < 1                 # :77


def test_cross_ops_are_rotated_normal_ops():
    """n x (f e_z) = (n_y f, -n_x f) and e_z . (n x v) = n_x v_y - n_y v_x
    (src/cartesian_extensions.jl:19-23): every cross operator equals its
    plain counterpart with the normals rotated by -90 degrees (regularizing
    side) or +90 degrees (interpolating side)."""
    g = O.grid(60, 56, 30, 28, 0.05)
    G = _G()
    s = G.star(0.6, [0.08, -0.05, 0.03, 0.0, 0.02], 1.4 * 0.05)
    geo = _geo(s)
    rm = O.Geometry(s.X, s.Y, s.ny, -s.nx, s.ds)     # (n_y, -n_x)
    rp = O.Geometry(s.X, s.Y, -s.ny, s.nx, s.ds)     # (-n_y, n_x)
    rng = np.random.default_rng(3)
    f = rng.standard_normal(s.N)
    q = rng.standard_normal(O.edge_len(g))
    sd = rng.standard_normal(O.nel(g, O.DUAL))
    ph = rng.standard_normal(O.nel(g, O.PRIMAL))
    assert np.array_equal(O.regularize_normal_cross(g, geo, f), O.regularize_normal(g, rm, f))
    assert np.array_equal(O.normal_cross_interpolate(g, geo, q), O.normal_interpolate(g, rp, q))
    assert np.array_equal(O.surface_curl_cross(g, geo, f), O.surface_curl(g, rm, f))
    assert np.array_equal(O.surface_curl_cross_adj(g, geo, sd), O.surface_curl_adj(g, rp, sd))
    assert np.array_equal(O.surface_grad_cross(g, geo, ph), O.surface_grad(g, rp, ph))


def test_surface_operator_adjoints():
    """GridScaling: R carries ds/dx^2, E does not, so for the surface
    calculus pairs  dx^2 <C_s^T f, s> = <f, C_s s>_ds  and, from D = -G^T on
    fields whose support stays inside the grid,  dx^2 <D_s f, phi> = -<f, G_s phi>_ds."""
    g, s = _circle_setup()
    geo = _geo(s)
    rng = np.random.default_rng(4)
    f = rng.standard_normal(s.N)
    sd = rng.standard_normal(O.nel(g, O.DUAL))
    ph = rng.standard_normal(O.nel(g, O.PRIMAL))
    lhs = g.dx ** 2 * np.dot(O.surface_curl(g, geo, f), sd)
    rhs = np.dot(f * s.ds, O.surface_curl_adj(g, geo, sd))
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    lhs = g.dx ** 2 * np.dot(O.surface_divergence(g, geo, f), ph)
    rhs = -np.dot(f * s.ds, O.surface_grad(g, geo, ph))
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    # regularize_normal / normal_interpolate: <Rf(n f), q> dx^2 = <f, n.Rf^T q>_ds
    q = rng.standard_normal(O.edge_len(g))
    lhs = g.dx ** 2 * np.dot(O.regularize_normal(g, geo, f), q)
    rhs = np.dot(f * s.ds, O.normal_interpolate(g, geo, q))
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


def test_mask_area():
    # mask = -L^-1 D_s 1 (src/surface_operators.jl:865-871), area of the unit circle
    g, s = _circle_setup()
    msk = -O.InverseLaplacian(g, O.PRIMAL)(O.surface_divergence(g, _geo(s), np.ones(s.N)))
    assert abs(g.dx ** 2 * msk.sum() - np.pi) < 2e-3                              # surface_ops.jl:188


# ------------------------------------------------------------ geometry anchors
def test_circle_geometry_anchors():
    G = _G()
    c = G.circle(1.0, 0.014)
    assert c.N == 448
    # RigidBodyTools measures the panels numerically: 5.7e-13 relative off the closed form 2aN tan(pi/N)
    assert abs(c.ds.sum() - 6.283288300933757) < 1e-11                           # caches.ipynb cell 28
    c = G.circle(0.5, 0.014)
    assert c.N == 224
    xdotn = np.sum((c.X * c.nx + c.Y * c.ny) * c.ds)
    assert abs(xdotn - 1.5708993266236526) < 1e-11                               # multbodies.ipynb cell 16


# ----------------------------------------------------------------- composites
def test_dirichlet_constraint_and_exact_field():
    G = _G()
    gs, s, fb = G.config_plain()
    g = O.grid(gs.NX, gs.NY, gs.I0x, gs.I0y, gs.dx)
    geo = _geo(s)
    f, sv = O.dirichlet_solve(g, geo, fb)
    assert np.abs(O.interpolate(g, O.PRIMAL, geo, f) - 0.5 * fb).max() < 1e-12   # E f = fbar
    # exterior data f_b(theta) = sum a_k cos(k theta + phi_k): harmonic continuation
    rng = np.random.default_rng(1)
    k = np.arange(1, 9)
    a = rng.standard_normal(8) / k
    ph = rng.uniform(0, 2 * np.pi, 8)
    X, Y = np.meshgrid(gs.x_primal(), gs.y_primal(), indexing="ij")
    R, TH = np.hypot(X, Y), np.arctan2(Y, X)
    F = f.reshape(gs.primal_shape(), order="F")
    out = (R > 0.65) & (R < 1.8)
    exact = sum(a[i] * (0.5 / R[out]) ** k[i] * np.cos(k[i] * TH[out] + ph[i]) for i in range(8))
    assert np.abs(F[out] - exact).max() < 0.02 * np.abs(fb).max()
    assert np.abs(F[R < 0.35]).max() < 0.05 * np.abs(fb).max()                   # interior ~ 0


def _square(a, dst):
    """Square of half side a: points at equal spacing along each side."""
    G = _G()
    n = int(round(2 * a / dst))
    h = 2 * a / n
    s = -a + h * (np.arange(n) + 0.5)
    X = np.concatenate([s, np.full(n, a), s[::-1], np.full(n, -a)])
    Y = np.concatenate([np.full(n, -a), s, np.full(n, a), s[::-1]])
    nx = np.concatenate([np.zeros(n), np.ones(n), np.zeros(n), -np.ones(n)])
    ny = np.concatenate([-np.ones(n), np.zeros(n), np.ones(n), np.zeros(n)])
    return G.Surface(X, Y, nx, ny, np.full(4 * n, h), np.zeros(4 * n, int))


def test_neumann_added_mass_matches_reference():
    """The reference's published added mass (examples/neumann.ipynb cell 37)."""
    G = _G()
    dx = 0.01
    g = O.grid(406, 406, 203, 203, dx)
    s = G.concat([_square(1.0, 1.4 * dx), G.circle(0.25, 1.4 * dx)])
    vp = np.where(s.body == 1, s.nx, 0.0)
    f, df, st, dsv = O.neumann_solve(g, _geo(s), vp)
    m = s.body == 1
    Mx = -np.sum(df[m] * s.nx[m] * s.ds[m])
    My = -np.sum(df[m] * s.ny[m] * s.ds[m])
    assert abs(Mx - 0.22649277527914002) < 1e-12
    assert abs(My) < 1e-9


def test_lu_solve():
    rng = np.random.default_rng(4)
    A = rng.standard_normal((50, 50))
    b = rng.standard_normal(50)
    info, x = O.lu_solve(A, b)
    assert info == 0
    assert rel_err(x, np.linalg.solve(A, b)) < 1e-12


# ------------------------------------------------------------------- heat
def test_heat_kernel_pins():
    """K(p) = e^{-2a} I_p(2a): against scipy's scaled Bessel function, unit mass,
    the semigroup K_a * K_b = K_{a+b}, and dK/da = K(p-1) - 2K(p) + K(p+1)."""
    from scipy.special import ive
    for a in (0.01, 0.125, 1.0 / 3.0, 2.0, 9.0):
        K = O.heat_kernel(a)
        assert np.abs(K - ive(np.arange(len(K)), 2 * a)).max() < 1e-15
        assert abs(K[0] + 2 * K[1:].sum() - 1.0) < 1e-14
    full = lambda K: np.concatenate([K[::-1], K[1:]])
    c = np.convolve(full(O.heat_kernel(0.2)), full(O.heat_kernel(0.3)))
    K5 = O.heat_kernel(0.5)
    m = len(c) // 2
    assert np.abs(c[m:m + len(K5)] - K5).max() < 1e-15
    a, h = 0.7, 1e-4
    Kp, Km, K0 = O.heat_kernel(a + h), O.heat_kernel(a - h), O.heat_kernel(a)
    n = min(len(Kp), len(Km), len(K0)) - 1
    dK = (Kp[:n] - Km[:n]) / (2 * h)
    lap = np.concatenate([[2 * K0[1] - 2 * K0[0]], K0[:n - 1] - 2 * K0[1:n] + K0[2:n + 1]])
    assert np.abs(dK[:n - 1] - lap[:n - 1]).max() < 1e-8


def _heat_setup(n=64, dx=0.04, width=4.0):
    G = _G()
    g = O.grid(n + 1, n + 1, n // 2 + 1, n // 2 + 1, dx)
    s = G.circle(0.5, 1.4 * dx)
    xs = (np.arange(n) + 1 - (n // 2 + 1)) * dx
    Xg, Yg = np.meshgrid(xs, xs, indexing="ij")
    T0 = np.exp(-((Xg - 0.4) ** 2 + (Yg + 0.3) ** 2) / (2 * (width * dx) ** 2)).ravel(order="F")
    return g, s, T0


def test_heat_stage_matrix_direct_matches_column_loop():
    """The direct assembly of the stage matrices E H R equals the column loop
    (regularize a unit strength, apply the heat kernel on the whole grid,
    interpolate) for the kernels of both stages and for the identity."""
    g, s, _ = _heat_setup()
    geo = _geo(s)
    for a in (0.25 / 3.0, 0.5 / 3.0, 3.0):
        K = O.heat_kernel(a)
        A = O.heat_schur(g, geo, K, "direct")
        B = O.heat_schur(g, geo, K, "columns")
        assert np.abs(A - B).max() < 1e-13 * np.abs(B).max(), a
    A = O.heat_schur(g, geo, None, "direct")
    B = O.heat_schur(g, geo, None, "columns")
    assert np.abs(A - B).max() < 1e-13 * np.abs(B).max()


def test_heat_no_bodies_is_the_exact_semigroup():
    """Without bodies IF-HERK integrates diffusion exactly: k steps of dt equal
    one application of exp(k dt kappa L)."""
    g, s, T0 = _heat_setup(width=2.0)   # negligible at the grid edge: no truncation there
    e = np.zeros(0)
    geo0 = O.Geometry(e, e, e, e, e)
    kappa, dt, k = 1.0, 0.25 * g.dx ** 2, 7
    T = O.heat_advance(g, geo0, kappa, dt, k, T0, e)
    K = O.heat_kernel(k * kappa * dt / g.dx ** 2)
    assert np.abs(T - O.heat_apply(g.NX - 1, g.NY - 1, K, T0)).max() < 1e-13


def test_heat_constraint_and_time_convergence():
    """With a body at T_b = 1: E T = T_b/2 + T_b/2 holds after every step, and
    halving dt shrinks the error at fixed time at better than second order."""
    g, s, T0 = _heat_setup()
    geo = _geo(s)
    tb = np.ones(s.N)
    kappa, tend = 1.0, 0.5 * g.dx ** 2 * 8
    T = O.heat_advance(g, geo, kappa, 0.5 * g.dx ** 2, 3, T0, tb, tb)
    assert np.abs(O.interpolate(g, O.PRIMAL, geo, T) - 1.0).max() < 1e-11
    sols = [O.heat_advance(g, geo, kappa, tend / k, k, T0, tb) for k in (2, 4, 8)]
    ratio = np.linalg.norm(sols[0] - sols[2]) / np.linalg.norm(sols[1] - sols[2])
    assert ratio > 4.0, ratio


# ------------------------------------------------------------------- masks
def _mask_setup():
    from ilhip import geometry as G
    # test/surface_ops.jl:4-12: dx = 0.04 on [-2,2]^2, unit circle, points at 1.4 dx
    gs = G.GridSpec(106, 106, 53, 53, 0.04)
    s = G.circle(1.0, 1.4 * gs.dx)
    return gs, O.grid(gs.NX, gs.NY, gs.I0x, gs.I0y, gs.dx), O.Geometry(s.X, s.Y, s.nx, s.ny, s.ds)


def test_mask_area_pins():
    """test/surface_ops.jl:180-206: the mask integrates to pi R^2 (2e-3) and the
    masked ones on primal nodes, dual nodes and both edge families to pi (1e-3);
    the complementary mask is 1 - mask; no points gives ones."""
    gs, og, geo = _mask_setup()
    m = O.mask(og, geo)
    dA = gs.dx ** 2
    assert abs(m.sum() * dA - np.pi) < 2e-3
    for kind in (O.PRIMAL, O.DUAL, O.EDGE_U, O.EDGE_V):
        assert abs(O.primal_to(og, m, kind).sum() * dA - np.pi) < 1e-3, kind
    # 1 deep inside, 0 far outside
    P = m.reshape(gs.primal_shape(), order="F")
    assert abs(P[52, 52] - 1) < 1e-3 and abs(P[5, 5]) < 1e-3
    assert np.array_equal(O.mask(og, geo, complementary=True), 1.0 - m)
    assert np.array_equal(O.mask(og, None), np.ones(O.nel(og, O.PRIMAL)))


def test_primal_to_is_the_averaging_of_a_linear_field():
    """grid_interpolate! of a linear field is that field at the target points."""
    from ilhip import geometry as G
    gs = G.GridSpec(23, 17, 11, 8, 0.1)
    og = O.grid(gs.NX, gs.NY, gs.I0x, gs.I0y, gs.dx)
    lin = lambda x, y: 0.3 + 2.0 * x - 1.5 * y
    xp, yp = np.meshgrid(gs.x_primal(), gs.y_primal(), indexing="ij")
    m = lin(xp, yp).ravel(order="F")
    xd, yd = np.meshgrid(gs.x_dual(), gs.y_dual(), indexing="ij")
    d = O.primal_to(og, m, O.DUAL).reshape(gs.dual_shape(), order="F")
    assert np.allclose(d[1:-1, 1:-1], lin(xd, yd)[1:-1, 1:-1], atol=1e-13)
    xu, yu = np.meshgrid(gs.x_dual(), gs.y_primal(), indexing="ij")
    u = O.primal_to(og, m, O.EDGE_U).reshape((gs.NX, gs.NY - 1), order="F")
    assert np.allclose(u[1:-1], lin(xu, yu)[1:-1], atol=1e-13)
    xv, yv = np.meshgrid(gs.x_primal(), gs.y_dual(), indexing="ij")
    v = O.primal_to(og, m, O.EDGE_V).reshape((gs.NX - 1, gs.NY), order="F")
    assert np.allclose(v[:, 1:-1], lin(xv, yv)[:, 1:-1], atol=1e-13)


# ---------------------------------------------------------- surface filter
def test_surface_filter_pins():
    """create_surface_filter (src/matrix_operators.jl:254-267): C keeps the
    surface integral and constants (test/literate/matrices.jl:49-56), ds C is
    symmetric with eigenvalues of C in [0, 1], and a column equals the oracle's
    own interpolate of the regularized unit strength divided by R 1."""
    from ilhip import geometry as G
    g, s, _ = G.config_plain()
    og = O.grid(g.NX, g.NY, g.I0x, g.I0y, g.dx)
    geo = O.Geometry(s.X, s.Y, s.nx, s.ny, s.ds)
    C = O.surface_filter(og, geo)
    ds = geo.ds
    assert np.abs(ds @ C - ds).max() < 1e-13 * ds.max()
    assert np.abs(C @ np.ones(geo.N) - 1.0).max() < 1e-13
    M = ds[:, None] * C
    assert np.abs(M - M.T).max() < 1e-13 * np.abs(M).max()
    ev = np.linalg.eigvals(C).real
    assert ev.min() > -1e-12 and ev.max() < 1.0 + 1e-12
    r1 = O.regularize(og, O.PRIMAL, geo, np.ones(geo.N))
    for q in (0, 37, geo.N - 1):
        e = np.zeros(geo.N)
        e[q] = 1.0
        f = O.regularize(og, O.PRIMAL, geo, e)
        f = np.where(r1 != 0, f / np.where(r1 != 0, r1, 1.0), 0.0)
        assert np.abs(O.interpolate(og, O.PRIMAL, geo, f) - C[:, q]).max() < 1e-14
    # the pins tell the grid-side normalization from the point-side one
    # (diag(1 / (E R 1)) E R, which keeps constants but not the integral)
    R = np.stack([O.regularize(og, O.PRIMAL, geo, np.eye(geo.N)[q]) for q in range(geo.N)], axis=1)
    E = (R * (og.dx ** 2 / ds)[None, :]).T
    ER = E @ R
    Cp = ER / (ER @ np.ones(geo.N))[:, None]
    assert np.abs(Cp @ np.ones(geo.N) - 1.0).max() < 1e-13
    assert np.abs(ds @ Cp - ds).max() > 1e-6 * ds.max()
    # filtering smooths: the filtered strength of a noisy signal is closer to its smooth part
    rng = np.random.default_rng(5)
    th = s.theta()
    smooth = np.cos(2 * th)
    noisy = smooth + 0.3 * rng.standard_normal(geo.N)
    f6 = np.linalg.matrix_power(C, 6) @ noisy
    assert np.linalg.norm(f6 - smooth) < 0.5 * np.linalg.norm(noisy - smooth)


# ------------------------------------------------------- learned network
def test_learned_network_backward_matches_finite_differences():
    """The FP64 restatement's backward pass (no BF16 rounding) against central
    differences of its loss, on a small network of the same structure."""
    import learned as LN
    rng = np.random.default_rng(12)
    K, H, Op = 2, 16, 16
    W = 2 * K + 1
    F = 6 * W + 3
    Fp = 64
    shapes = [(Fp, H), (H,), (H, H), (H,), (H, H), (H,), (H, Op), (Op,)]
    n = sum(int(np.prod(s)) for s in shapes)
    dims = (K, W, F, Fp, H, Op, n)
    p = rng.standard_normal(n) * 0.3
    X = rng.standard_normal((23, Fp))
    T = rng.standard_normal((23, Op))
    loss, g = LN.loss_grad(p, dims, X, T, bf16_round=False)
    h = 1e-6
    for i in rng.choice(n, 40, replace=False):
        e = np.zeros(n)
        e[i] = h
        fd = (LN.loss_grad(p + e, dims, X, T, False)[0] - LN.loss_grad(p - e, dims, X, T, False)[0]) / (2 * h)
        assert abs(fd - g[i]) < 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_learned_features_geometry():
    """Features of a circle: tangent/normal offsets of the neighbours follow
    the chord geometry (the j-th neighbour at angle j dtheta sits at
    (r sin(j dtheta), r (cos(j dtheta) - 1)) / dx in the point's frame, the
    normal pointing outward), and the periodic index wraps within each body."""
    import learned as LN
    from ilhip import geometry as G
    dx = 0.02
    c1 = G.circle(0.5, 1.4 * dx)
    c2 = G.circle(0.3, 1.4 * dx, 1.0, 0.0)
    s = G.concat([c1, c2])
    off = [0, c1.N, c1.N + c2.N]
    K = 3
    Fm = LN.features(s.X, s.Y, s.nx, s.ny, s.ds, dx, 103, 103, off, K)
    W = 2 * K + 1
    dth = 2 * np.pi / c1.N
    j = np.arange(-K, K + 1)
    # tangent (-ny, nx) is counter-clockwise for the outward normal
    assert np.allclose(Fm[0, :W], 0.5 * np.sin(j * dth) / dx, atol=1e-9)
    assert np.allclose(Fm[0, W:2 * W], 0.5 * (np.cos(j * dth) - 1) / dx, atol=1e-9)
    nb = LN.neighbours(off, s.N, K)
    assert nb[c1.N].min() >= c1.N and nb[0].max() < c1.N
