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
