"""Namelist-side helpers: space profiles evaluated on the grid.

Mirrors what the reference does on the host before the time loop:

* species creation from density / particles-per-cell profiles evaluated at
  the cell centres (ParticleCreator::create, src/Particles/ParticleCreator.cpp:
  123-205: positions `cell_position + (i + 0.5) * cell_length`);
* external fields evaluated at each field's staggered positions and added to
  it, with B_m set to B when only B is given (ElectroMagn::applyExternalFields,
  src/ElectroMagn/ElectroMagn.cpp:480-501; ElectroMagn3D::applyExternalField,
  src/ElectroMagn/ElectroMagn3D.cpp:1950-2010, dual axes at (g - 1/2) * dx).

A profile is a Python callable of (x, y, z) (numpy arrays of positions in
the namelist's length units) returning an array or scalar, as the
reference's Python profiles are.
"""
import numpy as np

# dual axes of each field (ElectroMagn3D.cpp:1955-1972)
DUAL = {
    "Ex": (1, 0, 0), "Ey": (0, 1, 0), "Ez": (0, 0, 1),
    "Bx": (0, 1, 1), "By": (1, 0, 1), "Bz": (1, 1, 0),
    "Bx_m": (0, 1, 1), "By_m": (1, 0, 1), "Bz_m": (1, 1, 0),
    "Jx": (1, 0, 0), "Jy": (0, 1, 0), "Jz": (0, 0, 1),
}


def _axes(sim, offsets):
    out = []
    for d in range(3):
        if d < sim.ndim:
            out.append((np.arange(sim.shape[d]) + offsets[d]) * sim.dx[d])
        else:
            out.append(np.zeros(1))
    return np.meshgrid(*out, indexing="ij")


def cell_centres(sim):
    """Positions of the cell centres, arrays of the grid's shape."""
    return _axes(sim, (0.5, 0.5, 0.5))


def field_positions(sim, name):
    """Positions of the points of field `name` (global index g sits at
    g * dx on a primal axis, (g - 1/2) * dx on a dual axis)."""
    return _axes(sim, tuple(-0.5 * f for f in DUAL[name]))


def evaluate(profile, x, y, z, shape):
    """Values of a profile on the points (x, y, z).  Like the reference's
    Profile (src/Profiles/Profile.cpp:199-260), a callable is first tried on
    numpy arrays, and evaluated point by point with floats when it does not
    accept them (e.g. a profile written with math.cos)."""
    if not callable(profile):
        v = profile
    else:
        try:
            v = np.asarray(profile(x, y, z), dtype=np.float64)
            if v.size not in (1, int(np.prod(shape))):
                raise ValueError("profile returned %d values for %d points" % (v.size, int(np.prod(shape))))
        except (TypeError, ValueError):
            xb, yb, zb = np.broadcast_arrays(x, y, z)
            v = np.array([float(profile(float(a), float(b), float(c)))
                          for a, b, c in zip(xb.ravel(), yb.ravel(), zb.ravel())]).reshape(xb.shape)
    return np.ascontiguousarray(np.broadcast_to(np.asarray(v, dtype=np.float64), shape))


def create_species(sim, sid, density, particles_per_cell, sigma=0.0, drift_ux=0.0, seed=0, theta=None):
    """Species creation from density and particles-per-cell profiles (each
    a callable of the cell-centre positions, or a number); momenta Gaussian
    of std sigma, or Maxwell-Juttner of temperature theta when given."""
    x, y, z = cell_centres(sim)
    n = evaluate(density, x, y, z, sim.shape)
    ppc = evaluate(particles_per_cell, x, y, z, sim.shape)
    if theta is not None:
        sim.init_species_profile_mj(sid, n, ppc, theta, drift_ux=drift_ux, seed=seed)
    else:
        sim.init_species_profile(sid, n, ppc, sigma=sigma, drift_ux=drift_ux, seed=seed)
    return n, ppc


def count_particles(sim, particles_per_cell, density):
    """Number of particles create_species makes (same rounding as
    ParticleCreator.cpp:217-246), to size the species capacity."""
    x, y, z = cell_centres(sim)
    n = evaluate(density, x, y, z, sim.shape)
    ppc = evaluate(particles_per_cell, x, y, z, sim.shape)
    base = np.floor(ppc)
    frac = ppc - base
    extra = np.zeros(sim.shape, dtype=bool)
    has = frac > 0
    if np.any(has):
        rem = np.where(has, np.power(np.where(has, frac, 1.0), -1.0 / sim.ndim), 1.0)
        idx = np.indices(sim.shape)
        extra = has.copy()
        for d in range(sim.ndim):
            extra &= np.fmod(idx[d].astype(np.float64), rem) < 1.0
    count = base + extra
    count[(count <= 0) | (np.abs(n) < 1e-200)] = 0
    return int(count.sum())


def apply_external_fields(sim, fields):
    """Add the profiles {field name: callable} to the fields at t = 0; a B
    component given without its B_m gets B_m = B."""
    given = set(fields)
    for name, prof in fields.items():
        x, y, z = field_positions(sim, name)
        sim.set_field(name, sim.get_field(name) + evaluate(prof, x, y, z, sim.shape))
    for c in ("x", "y", "z"):
        if "B" + c in given and "B%s_m" % c not in given:
            sim.set_field("B%s_m" % c, sim.get_field("B" + c))
