"""Pins the CPU oracle with closed forms, invariants and brute force.

The reference ships no operator-level golden vectors and cannot be built
here (mpi.h / hdf5.h absent), so every oracle function is checked against
what mathematics fixes: textbook B-spline values, exact reproduction of
uniform/linear fields, the Boris map's invariants, the discrete continuity
equation of Esirkepov's scheme, the exact discrete Yee plane-wave mode,
decomposition invariance of the per-patch machinery, conservation of
particles and of the Gauss residual.  Each check is chosen so that a
plausible slip (wrong sign, index, staggering, dropped term) fails it.
"""
import numpy as np
import pytest

from oracle import pyoracle as O
from tests.plasma import cfl_dt, thermal

# --------------------------------------------------------------- B-splines

def bspline(order, t):
    """Centred cardinal B-spline of degree `order` (textbook piecewise form)."""
    t = np.abs(t)
    if order == 2:
        return np.where(t <= 0.5, 0.75 - t * t, np.where(t <= 1.5, 0.5 * (1.5 - t) ** 2, 0.0))
    return np.where(t <= 0.5, 115 / 192 - 5 / 8 * t ** 2 + t ** 4 / 4,
                    np.where(t <= 1.5, (55 + 20 * t - 120 * t ** 2 + 80 * t ** 3 - 16 * t ** 4) / 96,
                             np.where(t <= 2.5, (5 - 2 * t) ** 4 / 384, 0.0)))


def patch_arrays(np3, rng, nd=3):
    """Random per-patch field arrays with the reference staggering."""
    dual = {"Ex": (1, 0, 0), "Ey": (0, 1, 0), "Ez": (0, 0, 1),
            "Bx_m": (0, 1, 1), "By_m": (1, 0, 1), "Bz_m": (1, 1, 0)}
    out = {}
    for k, dd in dual.items():
        shp = tuple(np3[d] + dd[d] for d in range(nd))
        out[k] = rng.standard_normal(shp)
    return out


def brute_gather(order, f, dual, dom, dinv, pos):
    """Independent gather: sum over all nodes of f * prod B(x/dx - node)."""
    h = order // 2
    val = 0.0
    nd = f.ndim
    grids = []
    for d in range(nd):
        xn = pos[d] * dinv[d]
        # node l of a primal axis sits at dom+l, of a dual axis at dom+l-1/2
        nodes = dom[d] + np.arange(f.shape[d]) - 0.5 * dual[d]
        grids.append(bspline(order, xn - nodes))
    if nd == 3:
        val = np.einsum("i,j,k,ijk->", grids[0], grids[1], grids[2], f)
    else:
        val = np.einsum("i,j,ij->", grids[0], grids[1], f)
    assert h in (1, 2)
    return val


@pytest.mark.parametrize("order", [2, 4])
def test_gather3d_matches_textbook_splines(order):
    rng = np.random.default_rng(10 + order)
    n_patch, os_ = 8, order
    np3 = [n_patch + 2 * os_ + 1] * 3
    dom = [16 - os_] * 3  # patch starting at cell 16
    dx = [0.3, 0.2, 0.25]
    dinv = [1.0 / d for d in dx]
    f = patch_arrays(np3, rng)
    n = 40
    pos = [(16 + rng.random(n) * n_patch) * dx[d] for d in range(3)]
    Ep, Bp, iold, dold = O.interp3d(order, f, np3, dom, dinv, *pos)
    duals = {"Ex": (1, 0, 0), "Ey": (0, 1, 0), "Ez": (0, 0, 1), "Bx_m": (0, 1, 1), "By_m": (1, 0, 1), "Bz_m": (1, 1, 0)}
    for i in range(n):
        p = [pos[d][i] for d in range(3)]
        for c, k in enumerate(("Ex", "Ey", "Ez")):
            ref = brute_gather(order, f[k], duals[k], dom, dinv, p)
            assert abs(Ep[c * n + i] - ref) <= 1e-12 * (1 + abs(ref))
        for c, k in enumerate(("Bx_m", "By_m", "Bz_m")):
            ref = brute_gather(order, f[k], duals[k], dom, dinv, p)
            assert abs(Bp[c * n + i] - ref) <= 1e-12 * (1 + abs(ref))
        for d in range(3):
            xn = p[d] * dinv[d]
            assert iold[d * n + i] == int(np.floor(xn + 0.5)) - dom[d]
            assert dold[d * n + i] == xn - np.floor(xn + 0.5)


def test_gather2d_matches_textbook_splines():
    rng = np.random.default_rng(3)
    np2 = [8 + 5, 8 + 5]
    dom = [6, 14]
    dx = [0.125, 0.2]
    dinv = [1 / d for d in dx]
    f = patch_arrays(np2, rng, nd=2)
    n = 50
    pos = [(dom[d] + 2 + rng.random(n) * 8) * dx[d] for d in range(2)]
    Ep, Bp, iold, dold = O.interp2d(2, f, np2, dom, dinv, *pos)
    duals = {"Ex": (1, 0), "Ey": (0, 1), "Ez": (0, 0), "Bx_m": (0, 1), "By_m": (1, 0), "Bz_m": (1, 1)}
    for i in range(n):
        p = [pos[d][i] for d in range(2)]
        for c, k in enumerate(("Ex", "Ey", "Ez")):
            ref = brute_gather(2, f[k], duals[k], dom, dinv, p)
            assert abs(Ep[c * n + i] - ref) <= 1e-12 * (1 + abs(ref))
        for c, k in enumerate(("Bx_m", "By_m", "Bz_m")):
            ref = brute_gather(2, f[k], duals[k], dom, dinv, p)
            assert abs(Bp[c * n + i] - ref) <= 1e-12 * (1 + abs(ref))


@pytest.mark.parametrize("order", [2, 4])
def test_gather_reproduces_linear_fields(order):
    """Momentum-conserving splines reproduce affine fields exactly."""
    np3 = [8 + 2 * order + 1] * 3
    dom = [-order] * 3
    dx = [0.2, 0.2, 0.2]
    f = {}
    for k, dd in {"Ex": (1, 0, 0), "Ey": (0, 1, 0), "Ez": (0, 0, 1),
                  "Bx_m": (0, 1, 1), "By_m": (1, 0, 1), "Bz_m": (1, 1, 0)}.items():
        shp = tuple(np3[d] + dd[d] for d in range(3))
        axes = [(dom[d] + np.arange(shp[d]) - 0.5 * dd[d]) * dx[d] for d in range(3)]
        X, Y, Z = np.meshgrid(*axes, indexing="ij")
        f[k] = 1.5 + 2.0 * X - 0.5 * Y + 0.25 * Z
    rng = np.random.default_rng(0)
    n = 100
    pos = [rng.random(n) * 8 * dx[d] for d in range(3)]
    Ep, Bp, _, _ = O.interp3d(order, f, np3, dom, [5.0] * 3, *pos)
    ref = 1.5 + 2.0 * pos[0] - 0.5 * pos[1] + 0.25 * pos[2]
    for c in range(3):
        np.testing.assert_allclose(Ep[c * n:(c + 1) * n], ref, rtol=0, atol=1e-13)
        np.testing.assert_allclose(Bp[c * n:(c + 1) * n], ref, rtol=0, atol=1e-13)


# --------------------------------------------------------------- Boris

def test_boris_pure_magnetic_rotation():
    rng = np.random.default_rng(1)
    n = 200
    u = rng.standard_normal((3, n)) * 3
    B = rng.standard_normal((3, n))
    E = np.zeros((3, n))
    q, m, dt = -1.0, 1.0, 0.3
    x = np.zeros((3, n))
    res = O.push_boris(3, q, m, dt, x[0], x[1], x[2], u[0], u[1], u[2], E.ravel(), B.ravel())
    un = np.array(res[3:6])
    np.testing.assert_allclose(np.linalg.norm(un, axis=0), np.linalg.norm(u, axis=0), rtol=1e-14)
    # the component along B is unchanged; the rotation angle is 2 atan(|t|)
    g = np.sqrt(1 + (u ** 2).sum(0))
    t = (q / m * dt / 2) / g * B
    bh = B / np.linalg.norm(B, axis=0)
    np.testing.assert_allclose((un * bh).sum(0), (u * bh).sum(0), atol=1e-12)
    up = u - (u * bh).sum(0) * bh
    unp = un - (un * bh).sum(0) * bh
    cosang = (up * unp).sum(0) / (np.linalg.norm(up, axis=0) * np.linalg.norm(unp, axis=0))
    np.testing.assert_allclose(np.arccos(np.clip(cosang, -1, 1)), 2 * np.arctan(np.linalg.norm(t, axis=0)),
                               atol=1e-7)
    # sense of rotation: for q<0 the rotation is counter-clockwise about B
    cross = np.cross(up.T, unp.T).T
    assert np.all((cross * bh).sum(0) > 0)
    # position advance uses the new velocity
    np.testing.assert_allclose(res[0], dt * un[0] / np.sqrt(1 + (un ** 2).sum(0)), rtol=1e-14)


def test_boris_pure_electric_kick():
    rng = np.random.default_rng(2)
    n = 100
    u = rng.standard_normal((3, n))
    E = rng.standard_normal((3, n))
    q, m, dt = 2.0, 3.0, 0.1
    z = np.zeros(n)
    res = O.push_boris(3, q, m, dt, z, z, z, u[0], u[1], u[2], E.ravel(), np.zeros(3 * n))
    un = np.array(res[3:6])
    np.testing.assert_allclose(un, u + q / m * dt * E, rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(res[6], 1 / np.sqrt(1 + (un ** 2).sum(0)), rtol=1e-15)


def test_boris_exb_drift():
    """Crossed fields: a particle launched at v = E x B / B^2 keeps it (the
    Boris scheme has no secular drift error); a wrong sign or a misplaced
    half-kick sends it gyrating away at O(v_d)."""
    E0, B0 = 0.05, 1.0
    q, m, dt = -1.0, 1.0, 0.05
    vd = E0 / B0
    gd = 1 / np.sqrt(1 - vd ** 2)
    x, y, z = np.zeros(1), np.zeros(1), np.zeros(1)
    ux, uy, uz = np.array([gd * vd]), np.zeros(1), np.zeros(1)
    E = np.array([0.0, E0, 0.0])
    B = np.array([0.0, 0.0, B0])
    nsteps = 2000
    for _ in range(nsteps):
        x, y, z, ux, uy, uz, _g = O.push_boris(3, q, m, dt, x, y, z, ux, uy, uz, E, B)
    assert abs(x[0] / (nsteps * dt) - vd) < 1e-6 * vd
    assert abs(ux[0] - gd * vd) < 1e-6 * vd and abs(uy[0]) < 1e-6 * vd and abs(y[0]) < 1e-6


# --------------------------------------------------------------- Esirkepov

def _cont_setup(order, rng, n):
    os_ = order
    npatch = 8
    np3 = [npatch + 2 * os_ + 1] * 3
    dom = [-os_] * 3
    dx = [0.3, 0.2, 0.25]
    dt = 0.9 * cfl_dt(dx)
    icv = 1.0 / (dx[0] * dx[1] * dx[2])
    # old positions inside the patch, moves below one cell
    old = [(2 + rng.random(n) * (npatch - 4)) * dx[d] for d in range(3)]
    new = [old[d] + (rng.random(n) - 0.5) * 1.6 * dx[d] for d in range(3)]
    return np3, dom, dx, dt, icv, old, new


@pytest.mark.parametrize("order", [2, 4])
def test_esirkepov_discrete_continuity_3d(order):
    rng = np.random.default_rng(20 + order)
    n = 30
    np3, dom, dx, dt, icv, old, new = _cont_setup(order, rng, n)
    dinv = [1 / d for d in dx]
    q = -1.0
    w = rng.random(n) + 0.5
    # iold / deltaold as the interpolator leaves them
    iold = np.empty(3 * n, np.int32)
    dold = np.empty(3 * n)
    for d in range(3):
        xn = old[d] * dinv[d]
        r = np.floor(xn + 0.5)
        iold[d * n:(d + 1) * n] = r.astype(np.int32) - dom[d]
        dold[d * n:(d + 1) * n] = xn - r
    Jx = np.zeros((np3[0] + 1, np3[1], np3[2]))
    Jy = np.zeros((np3[0], np3[1] + 1, np3[2]))
    Jz = np.zeros((np3[0], np3[1], np3[2] + 1))
    O.proj3d(order, [Jx, Jy, Jz], np3, dom, dx, dt, icv, q, *new, w, iold, dold)
    r0 = np.zeros(np3)
    r1 = np.zeros(np3)
    O.rho3d(order, r0, np3, dom, dinv, icv, q, *old, w)
    O.rho3d(order, r1, np3, dom, dinv, icv, q, *new, w)
    # dual index l+1 sits at (l+1/2)dx: div J at primal node l
    divJ = ((Jx[1:] - Jx[:-1]) / dx[0] + (Jy[:, 1:] - Jy[:, :-1]) / dx[1] + (Jz[:, :, 1:] - Jz[:, :, :-1]) / dx[2])
    resid = (r1 - r0) / dt + divJ
    scale = np.abs(r1).max() / dt
    assert np.abs(resid).max() < 1e-13 * scale
    # total current = sum q w v (first moment of the displacement)
    V = dx[0] * dx[1] * dx[2]
    tot = [Jx.sum() * V, Jy.sum() * V, Jz.sum() * V]
    for d in range(3):
        np.testing.assert_allclose(tot[d], (q * w * (new[d] - old[d]) / dt).sum(), rtol=1e-12)


@pytest.mark.parametrize("order", [2, 4])
def test_esirkepov_particle_at_rest_gives_zero_current(order):
    rng = np.random.default_rng(5)
    n = 10
    np3, dom, dx, dt, icv, old, _ = _cont_setup(order, rng, n)
    dinv = [1 / d for d in dx]
    iold = np.empty(3 * n, np.int32)
    dold = np.empty(3 * n)
    for d in range(3):
        xn = old[d] * dinv[d]
        r = np.floor(xn + 0.5)
        iold[d * n:(d + 1) * n] = r.astype(np.int32) - dom[d]
        dold[d * n:(d + 1) * n] = xn - r
    J = [np.zeros((np3[0] + 1, np3[1], np3[2])), np.zeros((np3[0], np3[1] + 1, np3[2])),
         np.zeros((np3[0], np3[1], np3[2] + 1))]
    O.proj3d(order, J, np3, dom, dx, dt, icv, 1.0, *old, np.ones(n), iold, dold)
    for a in J:
        assert np.abs(a).max() < 1e-15


def test_esirkepov_discrete_continuity_2d():
    rng = np.random.default_rng(7)
    n = 40
    npatch = 8
    np2 = [npatch + 5, npatch + 5]
    dom = [-2, -2]
    dx = [0.125, 0.2]
    dt = 0.9 * cfl_dt(dx)
    icv = 1.0 / (dx[0] * dx[1])
    dinv = [1 / d for d in dx]
    old = [(2 + rng.random(n) * (npatch - 4)) * dx[d] for d in range(2)]
    new = [old[d] + (rng.random(n) - 0.5) * 1.6 * dx[d] for d in range(2)]
    pz = rng.standard_normal(n)
    invgf = 1 / np.sqrt(1 + pz ** 2)
    w = rng.random(n) + 0.5
    q = -1.0
    iold = np.empty(2 * n, np.int32)
    dold = np.empty(2 * n)
    for d in range(2):
        xn = old[d] * dinv[d]
        r = np.floor(xn + 0.5)
        iold[d * n:(d + 1) * n] = r.astype(np.int32) - dom[d]
        dold[d * n:(d + 1) * n] = xn - r
    Jx = np.zeros((np2[0] + 1, np2[1]))
    Jy = np.zeros((np2[0], np2[1] + 1))
    Jz = np.zeros((np2[0], np2[1]))
    O.proj2d(2, [Jx, Jy, Jz], np2, dom, dx, dt, icv, q, *new, pz, w, invgf, iold, dold)
    r0 = np.zeros(np2)
    r1 = np.zeros(np2)
    O.rho2d(2, r0, np2, dom, dinv, icv, q, *old, w)
    O.rho2d(2, r1, np2, dom, dinv, icv, q, *new, w)
    divJ = (Jx[1:] - Jx[:-1]) / dx[0] + (Jy[:, 1:] - Jy[:, :-1]) / dx[1]
    resid = (r1 - r0) / dt + divJ
    assert np.abs(resid).max() < 1e-13 * np.abs(r1).max() / dt
    # Jz carries q w v_z with the time-averaged shape (weights sum to one)
    V = dx[0] * dx[1]
    np.testing.assert_allclose(Jz.sum() * V, (q * w * pz * invgf).sum(), rtol=1e-12)


# --------------------------------------------------------------- Yee

def _wave_sim(shape, patch, dx, dt):
    return O.OracleSim(3, 2, shape, patch, dx, dt)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_yee_exact_discrete_plane_wave(axis):
    """A plane wave obeying sin(w dt/2)/dt = sin(k dx/2)/dx is an exact
    eigenmode of the Yee scheme: after N steps only rounding remains."""
    shape = (16, 16, 16)
    dx = (0.2, 0.25, 0.3)
    dt = cfl_dt(dx)
    S = _wave_sim(shape, (8, 8, 8), dx, dt)
    k = 2 * np.pi * 2 / (shape[axis] * dx[axis])
    w = 2 / dt * np.arcsin(dt / dx[axis] * np.sin(k * dx[axis] / 2))
    # E transverse component e, B component b with E x B along +axis
    e, b = {0: (1, 2), 1: (2, 0), 2: (0, 1)}[axis]
    ename, bname = "E" + "xyz"[e], "B" + "xyz"[b]
    idx = np.indices(shape).astype(float)
    # primal along `axis` for E_e, dual (offset -1/2) for B_b
    xe = idx[axis] * dx[axis]
    xb = (idx[axis] - 0.5) * dx[axis]
    S.set_field(ename, np.cos(k * xe))
    # leapfrog: E at t = 0, B at t = +dt/2 (MA advances E with B^{n+1/2})
    S.set_field(bname, np.cos(k * xb - w * dt / 2))
    n = 25
    S.step(n)
    np.testing.assert_allclose(S.get_field(ename), np.cos(k * xe - w * n * dt), atol=5e-13)
    np.testing.assert_allclose(S.get_field(bname), np.cos(k * xb - w * (n + 0.5) * dt), atol=5e-13)
    # B_m is the centred average (B^{n-1/2} + B^{n+1/2}) / 2
    np.testing.assert_allclose(S.get_field(bname + "_m"), 0.5 * (np.cos(k * xb - w * (n + 0.5) * dt) +
                                                                 np.cos(k * xb - w * (n - 0.5) * dt)), atol=5e-13)


def _yee2d_mode(family, N, dx, m, dt):
    """Exact discrete plane wave of the 2D Yee scheme with wave vector
    k = 2 pi (m0 / L0, m1 / L1).  The Yee differences turn k into
    K_d = 2 sin(k_d dx_d / 2) / dx_d and w into W = 2 sin(w dt / 2) / dt, so
    the mode obeys W^2 = K0^2 + K1^2.  family "Ez": Ez = cos(phi) with
    Bx = (K1/W) cos, By = -(K0/W) cos; family "Bz": Bz = cos(phi) with
    Ex = -(K1/W) cos, Ey = (K0/W) cos (signs from dBx/dt = -dEz/dy,
    dBy/dt = dEz/dx, dEx/dt = dBz/dy, dEy/dt = -dBz/dx).  Returns
    f(name, t) giving each component at its staggered points at time t."""
    k = [2 * np.pi * m[d] / (N[d] * dx[d]) for d in range(2)]
    K = [2 * np.sin(k[d] * dx[d] / 2) / dx[d] for d in range(2)]
    W = np.sqrt(K[0] ** 2 + K[1] ** 2)
    w = 2 / dt * np.arcsin(W * dt / 2)
    dual = {"Ex": (1, 0), "Ey": (0, 1), "Ez": (0, 0), "Bx": (0, 1), "By": (1, 0), "Bz": (1, 1)}
    amp = ({"Ez": 1.0, "Bx": K[1] / W, "By": -K[0] / W} if family == "Ez" else
           {"Bz": 1.0, "Ex": -K[1] / W, "Ey": K[0] / W})
    idx = np.indices(tuple(N)).astype(float)

    def f(name, t):
        if name not in amp:
            return np.zeros(tuple(N))
        pos = [(idx[d] - 0.5 * dual[name][d]) * dx[d] for d in range(2)]
        return amp[name] * np.cos(k[0] * pos[0] + k[1] * pos[1] - w * t)
    return f


@pytest.mark.parametrize("family", ["Ez", "Bz"])
@pytest.mark.parametrize("m", [(2, 0), (0, 1), (1, 2)])
def test_yee2d_exact_discrete_plane_wave(family, m):
    """2D Yee (MA_Solver2D_norm, MF_Solver2D_Yee): both polarisations,
    propagating along x, along y and obliquely, on cells with dx != dy, are
    exact eigenmodes; a swapped dt/dx, dt/dy or a flipped curl sign changes
    the dispersion or the amplitudes and fails at O(1)."""
    N, dx = (16, 12), (0.125, 0.2)
    dt = 0.95 * cfl_dt(dx)
    S = O.OracleSim(2, 2, N, (8, 4), dx, dt)
    f = _yee2d_mode(family, N, dx, m, dt)
    for name in ("Ex", "Ey", "Ez"):
        S.set_field(name, f(name, 0.0))
    for name in ("Bx", "By", "Bz"):
        S.set_field(name, f(name, 0.5 * dt))
    n = 30
    S.step(n)
    for name in ("Ex", "Ey", "Ez"):
        np.testing.assert_allclose(S.get_field(name).reshape(N), f(name, n * dt), atol=5e-13, err_msg=name)
    for name in ("Bx", "By", "Bz"):
        np.testing.assert_allclose(S.get_field(name).reshape(N), f(name, (n + 0.5) * dt), atol=5e-13, err_msg=name)
        np.testing.assert_allclose(S.get_field(name + "_m").reshape(N),
                                   0.5 * (f(name, (n + 0.5) * dt) + f(name, (n - 0.5) * dt)), atol=5e-13)


def test_yee2d_preserves_div_b():
    rng = np.random.default_rng(19)
    N, dx = (16, 8), (0.125, 0.2)
    S = O.OracleSim(2, 2, N, (8, 4), dx, 0.95 * cfl_dt(dx))
    for k in ("Ex", "Ey", "Ez", "Bx", "By", "Bz"):
        S.set_field(k, rng.standard_normal(N))

    def divB():
        Bx, By = S.get_field("Bx").reshape(N), S.get_field("By").reshape(N)
        return (Bx - np.roll(Bx, 1, 0)) / dx[0] + (By - np.roll(By, 1, 1)) / dx[1]

    d0 = divB()
    S.step(10)
    np.testing.assert_allclose(divB(), d0, atol=1e-11)


def test_yee_preserves_div_b():
    rng = np.random.default_rng(9)
    shape = (16, 8, 16)
    dx = (0.2, 0.2, 0.2)
    S = _wave_sim(shape, (8, 8, 8), dx, cfl_dt(dx))
    for k in ("Ex", "Ey", "Ez", "Bx", "By", "Bz"):
        S.set_field(k, rng.standard_normal(shape))

    def divB():
        Bx, By, Bz = S.get_field("Bx"), S.get_field("By"), S.get_field("Bz")
        # divergence at the dual-dual-dual node (I,J,K): Bx is primal in x,
        # so its nodes I-1 and I straddle the dual node I
        return ((Bx - np.roll(Bx, 1, 0)) / dx[0] + (By - np.roll(By, 1, 1)) / dx[1] +
                (Bz - np.roll(Bz, 1, 2)) / dx[2])

    d0 = divB()
    S.step(10)
    np.testing.assert_allclose(divB(), d0, atol=1e-11)


# --------------------------------------------------------------- whole step

def _sim_with_plasma(nd, order, shape, patch, ppc, seed, nspec=2):
    dx = (0.2, 0.2, 0.2)[:nd] if nd == 3 else (0.125, 0.125)
    dt = cfl_dt(dx)
    S = O.OracleSim(nd, order, shape, patch, dx, dt)
    parts = []
    for s, (m, q, sig) in enumerate([(1.0, -1.0, 0.1), (1836.0, 1.0, 0.1 / np.sqrt(1836))][:nspec]):
        sid = S.add_species(m, q)
        p = thermal(shape, dx, ppc, sig, seed + s)
        S.set_particles(sid, p["x"], p["y"], p["z"], p["px"], p["py"], p["pz"], p["w"])
        parts.append(p)
    return S, dx, parts


def _patch_of(pos, dx, shape, patch, nd):
    pc = []
    for d in range(3):
        if d < nd:
            L = patch[d] * dx[d]
            c = np.clip(np.floor(pos[d] / L).astype(int), 0, shape[d] // patch[d] - 1)
            c = np.where(pos[d] < c * L, c - 1, c)
            c = np.where(pos[d] >= (c + 1) * L, c + 1, c)
            pc.append(c)
        else:
            pc.append(np.zeros(len(pos[0]), int))
    np_ = [shape[d] // patch[d] if d < nd else 1 for d in range(3)]
    return (pc[0] * np_[1] + pc[1]) * np_[2] + pc[2]


def _key_table(patch, nd):
    """In-patch cell key of every cell by enumeration, in the order the
    reference's Cluster3D::CellKey (nvidiaParticles.cu:639-660) ranks cells:
    clusters of 4 cells per axis walked x-major (x, then y, then z), and
    inside a cluster its cells walked x-major.  The table is built by
    counting, not by the key's arithmetic, so a transposed or mis-scaled
    key formula in the oracle disagrees with it."""
    ext = list(patch[:nd]) + [1] * (3 - nd)
    w = [4 if d < nd else 1 for d in range(3)]
    table = np.full(ext, -1, np.int64)
    n = 0
    for cx in range(ext[0] // w[0]):
        for cy in range(ext[1] // w[1]):
            for cz in range(ext[2] // w[2]):
                for ix in range(w[0]):
                    for iy in range(w[1]):
                        for iz in range(w[2]):
                            table[cx * w[0] + ix, cy * w[1] + iy, cz * w[2] + iz] = n
                            n += 1
    assert n == np.prod(ext) and table.min() == 0
    return table


def _expected_keys(p, dx, shape, patch, nd):
    """Keys of the cells the positions lie in, looked up in _key_table."""
    npc = [shape[d] // patch[d] if d < nd else 1 for d in range(3)]
    pid = p["patch"].astype(np.int64)
    pc = [pid // (npc[1] * npc[2]), (pid // npc[2]) % npc[1], pid % npc[2]]
    pos = [p["x"], p["y"], p["z"]]
    ic = []
    for d in range(3):
        if d < nd:
            c = np.floor(pos[d] * (1.0 / dx[d])).astype(np.int64) - pc[d] * patch[d]
            ic.append(np.clip(c, 0, patch[d] - 1))
        else:
            ic.append(np.zeros_like(pid))
    return _key_table(patch, nd)[ic[0], ic[1], ic[2]]


def test_key_table_matches_reference_examples():
    """Hand-checked keys of a 8x8x8 patch (2x2x2 clusters): cell (0,0,1) is
    the second cell of cluster 0, (1,0,0) the 17th, (0,0,4) starts the
    second cluster (z is the fastest cluster axis), (4,0,0) the fifth."""
    t = _key_table((8, 8, 8), 3)
    assert t[0, 0, 1] == 1 and t[0, 1, 0] == 4 and t[1, 0, 0] == 16
    assert t[0, 0, 4] == 64 and t[0, 4, 0] == 128 and t[4, 0, 0] == 256 and t[7, 7, 7] == 511
    t2 = _key_table((8, 12), 2)
    assert t2[0, 1, 0] == 1 and t2[1, 0, 0] == 4 and t2[0, 4, 0] == 16 and t2[4, 0, 0] == 48


@pytest.mark.parametrize("nd,order,shape,patch", [
    (3, 2, (16, 16, 16), (8, 8, 8)),
    (3, 4, (24, 12, 12), (12, 12, 12)),
    (2, 2, (32, 32), (8, 8)),
])
def test_step_invariants(nd, order, shape, patch):
    S, dx, parts = _sim_with_plasma(nd, order, shape, patch, 4, seed=11)
    g0 = S.gauss()
    S.step(12)
    g1 = S.gauss()
    rho = S.rho()
    # Esirkepov + Yee keep div(E) - rho constant
    assert np.abs(g1 - g0).max() < 1e-12 * np.abs(rho).max()
    for s in range(2):
        p = S.get_particles(s)
        assert len(p["x"]) == len(parts[s]["x"])
        np.testing.assert_allclose(np.sort(p["w"]), np.sort(parts[s]["w"]))
        # owner patch is the one containing the position
        pos = [p["x"], p["y"], p["z"]]
        np.testing.assert_array_equal(p["patch"], _patch_of(pos, dx, shape, patch, nd))
        # every key is the enumerated key of the cell holding the position
        np.testing.assert_array_equal(p["key"], _expected_keys(p, dx, shape, patch, nd))
        # keys sorted inside each patch and counts equal to the histogram
        for q in range(S.npatch):
            kq = p["key"][p["patch"] == q]
            assert np.all(np.diff(kq) >= 0)
        np.testing.assert_array_equal(S.patch_counts(s), np.bincount(p["patch"], minlength=S.npatch))
        # positions inside the box
        for d in range(nd):
            assert np.all(pos[d] >= 0) and np.all(pos[d] < shape[d] * dx[d])


@pytest.mark.parametrize("nd,order,shape,patches", [
    (3, 2, (16, 16, 16), [(16, 16, 16), (8, 8, 8), (8, 16, 8)]),
    (2, 2, (32, 16), [(32, 16), (8, 8), (16, 8)]),
])
def test_decomposition_invariance(nd, order, shape, patches):
    """Per-patch ghosted arrays + ghost sums + B exchange give the same
    fields as a single patch: only the J summation order may differ."""
    results = []
    for patch in patches:
        dx = (0.2, 0.2, 0.2)[:nd] if nd == 3 else (0.125, 0.125)
        S = O.OracleSim(nd, order, shape, patch, dx, cfl_dt(dx))
        p = thermal(shape, dx, 4, 0.1, 5)
        # unique weights identify particles across decompositions
        p["w"] = p["w"] * (1 + 1e-9 * np.arange(len(p["w"])))
        sid = S.add_species(1.0, -1.0)
        S.set_particles(sid, p["x"], p["y"], p["z"], p["px"], p["py"], p["pz"], p["w"])
        S.step(10)
        out = {k: S.get_field(k) for k in ("Ex", "Ey", "Ez", "Bx", "By", "Bz", "Jx", "Jy", "Jz")}
        q = S.get_particles(sid)
        o = np.argsort(q["w"])
        out["parts"] = np.array([q[k][o] for k in ("x", "y", "z", "px", "py", "pz")])
        results.append(out)
    for r in results[1:]:
        for k in ("Ex", "Ey", "Ez", "Bx", "By", "Bz", "Jx", "Jy", "Jz"):
            scale = np.abs(results[0][k]).max()
            assert np.abs(r[k] - results[0][k]).max() <= 1e-12 * scale, k
        np.testing.assert_allclose(r["parts"], results[0]["parts"], rtol=1e-12, atol=1e-14)


_DUAL = {"Ex": (1, 0, 0), "Ey": (0, 1, 0), "Ez": (0, 0, 1), "Bx_m": (0, 1, 1), "By_m": (1, 0, 1), "Bz_m": (1, 1, 0)}
_EFIELDS = ("Ex", "Ey", "Ez", "Bx_m", "By_m", "Bz_m")


def _norm2_extended(a, dual, nd):
    """Sum of squares over global indices 0..N (primal) or 0..N+1 (dual)
    per axis, read periodically: the points the reference's norm2 visits
    over all patches (ElectroMagn3D.cpp:185-226)."""
    idx = [np.arange(a.shape[d] + 1 + dual[d]) % a.shape[d] if d < nd else np.arange(a.shape[d])
           for d in range(a.ndim)]
    return float((a[np.ix_(*idx)] ** 2).sum())


@pytest.mark.parametrize("nd,shape,patch", [
    (3, (16, 8, 12), (8, 4, 4)),   # several patches per axis
    (3, (8, 8, 8), (8, 4, 8)),     # one patch along x and z
    (2, (16, 12), (8, 4)),
])
def test_field_energy_closed_form(nd, shape, patch):
    """Uelm_f = 0.5 V_cell sum f^2 over the istart/bufsize ranges: a uniform
    field c_f gives 0.5 V c_f^2 prod_d (N_d + 1 + dual_d); a random periodic
    field the sum over the periodically extended index ranges.  B_m (not B)
    enters; no ghost point beyond the ranges, no interior point twice."""
    dx = (0.3, 0.2, 0.25)[:nd]
    S = O.OracleSim(nd, 2, shape, patch, dx, 0.5 * cfl_dt(dx))
    V = float(np.prod(dx))
    full = tuple(shape) + (1,) * (3 - nd)
    for i, k in enumerate(_EFIELDS):
        S.set_field(k, np.full(full, 1.0 + i))
    for k in ("Bx", "By", "Bz"):
        S.set_field(k, np.full(full, 100.0))  # must not enter
    fe = S.field_energies()
    for i, k in enumerate(_EFIELDS):
        pts = np.prod([shape[d] + 1 + _DUAL[k][d] for d in range(nd)])
        np.testing.assert_allclose(fe[i], 0.5 * V * (1.0 + i) ** 2 * pts, rtol=1e-14, err_msg=k)
    rng = np.random.default_rng(4)
    arrs = {k: rng.standard_normal(full) for k in _EFIELDS}
    for k in _EFIELDS:
        S.set_field(k, arrs[k])
    fe = S.field_energies()
    for i, k in enumerate(_EFIELDS):
        np.testing.assert_allclose(fe[i], 0.5 * V * _norm2_extended(arrs[k], _DUAL[k], nd), rtol=1e-13,
                                   err_msg=k)
    np.testing.assert_allclose(S.energies()[0], fe.sum(), rtol=1e-15)


def test_kinetic_energy_closed_form():
    """Ukin_s = m_s sum w (gamma - 1) for particles of known momenta."""
    shape, patch, dx = (8, 8, 8), (4, 4, 4), (0.2, 0.2, 0.2)
    S = O.OracleSim(3, 2, shape, patch, dx, 0.5 * cfl_dt(dx))
    masses = (1.0, 1836.0)
    u = [np.array([[3.0, 0.0, 4.0], [0.0, 0.0, 0.0], [1.0, 1.0, 1.0]]),
         np.array([[0.0, 1.0, 0.0], [2.0, 2.0, 1.0]])]
    w = [np.array([0.5, 2.0, 1.5]), np.array([0.25, 4.0])]
    # gamma = sqrt(1 + |u|^2): 26**.5, 1, 2; 2**.5, sqrt(10)
    expect = [1.0 * (0.5 * (26 ** 0.5 - 1) + 0 + 1.5 * (2 - 1)),
              1836.0 * (0.25 * (2 ** 0.5 - 1) + 4.0 * (10 ** 0.5 - 1))]
    rng = np.random.default_rng(1)
    for s in range(2):
        sid = S.add_species(masses[s], -1.0 if s == 0 else 1.0)
        n = len(w[s])
        pos = [rng.random(n) * shape[d] * dx[d] for d in range(3)]
        S.set_particles(sid, *pos, u[s][:, 0], u[s][:, 1], u[s][:, 2], w[s])
    e = S.energies()
    assert e[0] == 0.0
    np.testing.assert_allclose(e[1:], expect, rtol=1e-14)


def test_thermal_energy_history_short():
    """Normalised energy check in the style of
    validation/analyses/validate_tst3d_gpu_o2_thermal_plasma_short.py."""
    S, dx, _ = _sim_with_plasma(3, 2, (16, 16, 16), (8, 8, 8), 8, seed=3)
    hist = []
    for _ in range(10):
        S.step(5)
        hist.append(S.energies().sum())
    hist = np.array(hist)
    assert np.abs(hist / hist.mean() - 1).max() < 1e-2


def test_frozen_species_stays_put_until_time_frozen():
    """Species::time_frozen_: a species is pushed by the step whose
    time_dual = (steps done + 1.5) dt exceeds time_frozen.  With
    time_frozen = 4.2 dt the protons stay bit-identical through steps 1-3
    (time_dual 1.5, 2.5, 3.5 dt), the electrons move and the protons' charge
    still counts in rho; from step 4 (4.5 dt) the protons move."""
    shape, patch, dx = (16, 16, 16), (8, 8, 8), (0.2, 0.2, 0.2)
    dt = cfl_dt(dx)
    S = O.OracleSim(3, 2, shape, patch, dx, dt)
    ps = thermal(shape, dx, 2, 0.3, 12)
    pe = thermal(shape, dx, 2, 0.3, 13)
    e = S.add_species(1.0, -1.0)
    p = S.add_species(1.0, 1.0)  # light ions: they would move at once
    S.set_particles(e, pe["x"], pe["y"], pe["z"], pe["px"], pe["py"], pe["pz"], pe["w"])
    S.set_particles(p, ps["x"], ps["y"], ps["z"], ps["px"], ps["py"], ps["pz"], ps["w"])
    S.set_time_frozen(p, 4.2 * dt)
    p0, e0 = S.get_particles(p), S.get_particles(e)
    g0 = S.gauss()
    S.step(3)
    p3, e3 = S.get_particles(p), S.get_particles(e)
    for k in ("x", "y", "z", "px", "py", "pz", "w", "key", "patch"):
        np.testing.assert_array_equal(p3[k], p0[k], err_msg=k)
    assert not np.array_equal(e3["x"], e0["x"])
    # Gauss stays put with the frozen charge included in rho
    assert np.abs(S.gauss() - g0).max() < 1e-12 * np.abs(S.rho()).max()
    S.step(1)
    assert not np.array_equal(S.get_particles(p)["x"], p0["x"])
