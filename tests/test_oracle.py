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
