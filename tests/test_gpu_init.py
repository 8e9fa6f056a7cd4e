"""Profile-driven particle creation (ParticleCreator's per-cell rounding and
weights) and the beam-cylinder initialiser, on the GPU engine."""
import numpy as np
import pytest

import smilei_amd as sa
from tests.plasma import cfl_dt

pytestmark = pytest.mark.gpu


def make(shape=(16, 16, 16), patch=(8, 8, 8), dx=(0.2, 0.3, 0.4), precision=sa.PLAIN_FP64):
    return sa.Simulation("3Dcartesian", 2, shape, patch, dx, cfl_dt(dx), precision=precision)


def cell_counts(sim, s, dx):
    p = sim.get_particles(s)
    idx = [np.floor(np.array(p[k]) / dx[d]).astype(int) for d, k in enumerate("xyz")]
    c = np.zeros(sim.shape, int)
    np.add.at(c, tuple(idx), 1)
    return c, p


@pytest.mark.parametrize("precision", [sa.PLAIN_FP64, sa.MIXED])
def test_profile_fractional_ppc_lattice(precision):
    """ppc = 1/8 in 3D: the fractional-part rule (remainder (1/8)^(-1/3) = 2)
    gives one particle exactly in the cells whose three indices are even."""
    dx = (0.2, 0.3, 0.4)
    sim = make(dx=dx, precision=precision)
    sid = sim.add_species("e", 1.0, -1.0, 16 ** 3)
    sim.init_species_profile(sid, 2.0, 0.125, sigma=0.0, seed=3)
    c, p = cell_counts(sim, sid, dx)
    i, j, k = np.indices(sim.shape)
    expect = ((i % 2 == 0) & (j % 2 == 0) & (k % 2 == 0)).astype(int)
    np.testing.assert_array_equal(c, expect)
    # weight = density * V_cell / nPart
    np.testing.assert_allclose(p["w"], 2.0 * 0.2 * 0.3 * 0.4, rtol=1e-7)


def test_profile_density_and_counts():
    """Integer ppc profile, zero density gives no particles, weights follow
    density * V / nPart cell by cell, momenta carry the drift."""
    dx = (0.2, 0.3, 0.4)
    sim = make(dx=dx)
    i, j, k = np.indices(sim.shape)
    dens = np.where(i < 4, 0.0, 0.5 + 0.1 * i)
    ppc = 1 + (j % 3)
    sid = sim.add_species("e", 1.0, -1.0, int(ppc.sum()))
    sim.init_species_profile(sid, dens, ppc, sigma=0.0, drift_ux=0.3, seed=5)
    c, p = cell_counts(sim, sid, dx)
    np.testing.assert_array_equal(c, np.where(dens > 0, ppc, 0))
    V = 0.2 * 0.3 * 0.4
    cell = [np.floor(np.array(p[q]) / dx[d]).astype(int) for d, q in enumerate("xyz")]
    w_expect = dens[tuple(cell)] * V / ppc[tuple(cell)]
    np.testing.assert_allclose(p["w"], w_expect, rtol=1e-14)
    np.testing.assert_allclose(p["px"], 0.3, rtol=0, atol=0)
    assert sim.num_particles(sid) == int(np.where(dens > 0, ppc, 0).sum())


def test_beam_cylinder_counts():
    """Cells whose centre lies within the radius of the x axis get ppc
    particles each, all with the beam drift."""
    shape, dx = (16, 16, 16), (0.25, 0.25, 0.25)
    sim = make(shape=shape, dx=dx)
    j = np.arange(16) + 0.5 - 8
    inside = (j[:, None] ** 2 + j[None, :] ** 2 <= 5.0 ** 2)
    sid = sim.add_species("beam", 1.0, -1.0, int(inside.sum()) * 16 * 4)
    sim.init_species(sid, 4, 0.1, 0.0, drift_ux=49.99, radius=5.0, seed=2)
    c, p = cell_counts(sim, sid, dx)
    np.testing.assert_array_equal(c, np.broadcast_to(inside * 4, shape))
    np.testing.assert_allclose(p["px"], 49.99, rtol=0)
    np.testing.assert_allclose(p["w"], 0.1 * 0.25 ** 3 / 4, rtol=1e-15)
