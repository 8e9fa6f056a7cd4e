"""Seeded synthetic inputs shared by the tests (not collected by pytest).

Thermal plasmas as described by the workloads: ppc particles per cell,
uniform-random positions inside each cell, Gaussian momenta of std sigma
(units of m_species*c), weight n*V_cell/ppc (ParticleCreator.cpp:259).
"""
import numpy as np


def thermal(shape, dx, ppc, sigma, seed, density=1.0, drift=0.0, drift_ux=None, radius=0.0):
    """radius > 0 keeps only the cells whose centre lies within `radius`
    cells of the x axis (the beam cylinder of the beam-plasma workload)."""
    if drift_ux is not None:
        drift = drift_ux
    rng = np.random.default_rng(seed)
    nd = len(shape)
    ncell = int(np.prod(shape))
    act = np.arange(ncell)
    if radius > 0:
        idx = np.unravel_index(act, shape)
        ry = idx[1] + 0.5 - 0.5 * shape[1]
        rz = idx[2] + 0.5 - 0.5 * shape[2] if nd == 3 else 0.0
        act = act[ry * ry + rz * rz <= radius * radius]
    n = len(act) * ppc
    cells = np.repeat(act, ppc)
    idx = np.unravel_index(cells, shape)
    pos = []
    for d in range(3):
        if d < nd:
            v = (idx[d] + rng.random(n)) * dx[d]
            v = np.minimum(v, np.nextafter(shape[d] * dx[d], 0.0))
            pos.append(v)
        else:
            pos.append(np.zeros(n))
    u = rng.normal(0.0, sigma, (3, n))
    u[0] += drift
    vol = float(np.prod(dx[:nd]))
    w = np.full(n, density * vol / ppc)
    return dict(x=pos[0], y=pos[1], z=pos[2], px=u[0], py=u[1], pz=u[2], w=w)


def cfl_dt(dx, factor=0.95):
    return factor / np.sqrt(sum(1.0 / d ** 2 for d in dx))


def load(sim, sid, p):
    sim.set_particles(sid, p["x"], p["y"], p["z"], p["px"], p["py"], p["pz"], p["w"])


def lwfa_profiles(shape=(512, 128, 128), dx=(0.1, 0.5, 0.5)):
    """Laser-wakefield-like setup: electron density 0.01 (critical-density
    units) behind a linear ramp over 100 cells from x = 64 cells (both scaled
    with shape[0] / 512), parabolic channel n0 (1 + 0.5 r^2 / r0^2), r0 = 20,
    around the box axis; one particle in every cell with an even index sum
    (a checkerboard particles-per-cell profile: 4096 particles in every full
    32x16x16 patch, none before the ramp); a y-polarised Gaussian
    laser as initial field, a0 = 2, lambda = 2 pi, waist 16, duration 20
    (FWHM of the field envelope), centred 160 cells (scaled) from x = 0 on
    the axis, B_z = E_y (propagation along +x)."""
    s = shape[0] / 512.0
    L = [shape[d] * dx[d] for d in range(3)]
    x0, ramp = 64 * s * dx[0], 100 * s * dx[0]
    yc, zc = 0.5 * L[1], 0.5 * L[2]
    n0, r0 = 0.01, 20.0

    def density(x, y, z):
        r2 = (y - yc) ** 2 + (z - zc) ** 2
        lon = np.clip((x - x0) / ramp, 0.0, 1.0)
        return n0 * lon * (1.0 + 0.5 * r2 / r0 ** 2)

    a0, waist, dur = 2.0, 16.0, 20.0
    xl = 160 * s * dx[0]

    def laser(x, y, z):
        r2 = (y - yc) ** 2 + (z - zc) ** 2
        env = np.exp(-4.0 * np.log(2.0) * (x - xl) ** 2 / dur ** 2)
        return a0 * np.exp(-r2 / waist ** 2) * env * np.sin(x - xl)

    def ppc(x, y, z):
        i = np.floor(x / dx[0]) + np.floor(y / dx[1]) + np.floor(z / dx[2])
        return np.where(np.fmod(i, 2.0) == 0.0, 1.0, 0.0)

    return density, ppc, {"Ey": laser, "Bz": laser}


def lwfa(sim_cls, precision, shape=(512, 128, 128), patch=(32, 16, 16), dx=(0.1, 0.5, 0.5), seed=7, **kw):
    """Build the LWFA-like simulation (electrons only; the immobile ions are a
    uniform neutralising background that no step touches)."""
    from smilei_amd import profiles as P
    sim = sim_cls("3Dcartesian", 2, shape, patch, dx, cfl_dt(dx), precision=precision, **kw)
    density, ppc, fields = lwfa_profiles(shape, dx)
    sid = sim.add_species("electron", 1.0, -1.0, P.count_particles(sim, ppc, density) + 4096)
    P.create_species(sim, sid, density, ppc, sigma=0.0, seed=seed)
    P.apply_external_fields(sim, fields)
    return sim
