"""Frozen species (Species::time_frozen_, Species.cpp:568, 832-866) on the
GPU engine: not pushed, no current, not re-sorted while the step's
time_dual = (steps done + 1.5) dt <= time_frozen, counted in rho.  Against
the oracle from the same particles across the step at which the species
starts to move, with the step captured in graphs (the captured step changes
when the set of moving species does)."""
import numpy as np
import pytest

import smilei_amd as sa
from oracle import pyoracle as O
from tests.plasma import cfl_dt, load, thermal
from tests.test_gpu_configs import compare

pytestmark = pytest.mark.gpu


def setup(deposit, precision=sa.PLAIN_FP64, use_graph=True):
    shape, patch, dx = (32, 16, 16), (16, 8, 8), (0.2, 0.2, 0.2)
    dt = cfl_dt(dx)
    g = sa.Simulation("3Dcartesian", 2, shape, patch, dx, dt, deposit=deposit, precision=precision,
                      use_graph=use_graph)
    o = O.OracleSim(3, 2, shape, patch, dx, dt)
    parts = [("electron", 1.0, -1.0, thermal(shape, dx, 4, 0.2, 21)),
             ("ion", 4.0, 1.0, thermal(shape, dx, 4, 0.2, 22))]
    for name, m, q, p in parts:
        load(g, g.add_species(name, m, q, len(p["x"])), p)
        o.set_particles(o.add_species(m, q), p["x"], p["y"], p["z"], p["px"], p["py"], p["pz"], p["w"])
    g.set_time_frozen(1, 4.2 * dt)
    o.set_time_frozen(1, 4.2 * dt)
    return g, o


@pytest.mark.parametrize("use_graph", [True, False])
def test_frozen_species_deterministic_bit_exact(use_graph):
    g, o = setup(sa.DEPOSIT_DETERMINISTIC, use_graph=use_graph)
    p0 = g.get_particles(1)
    g.step(3)
    o.step(3)
    compare(g, o, 2, exact=True)
    p3 = g.get_particles(1)
    for k in ("x", "px", "key"):
        np.testing.assert_array_equal(p3[k], p0[k])
    g.step(5)
    o.step(5)
    assert g.steps_done == 8
    compare(g, o, 2, exact=True)
    assert not np.array_equal(g.get_particles(1)["x"], p0["x"])


def test_frozen_species_atomic_within_1e12():
    g, o = setup(sa.DEPOSIT_ATOMIC)
    g.step(8)
    o.step(8)
    compare(g, o, 2, exact=False, tol=1e-12)


def test_frozen_species_mixed_form():
    """Mixed form: the frozen species' particles are untouched (bit for
    bit) while frozen and move afterwards; Gauss stays at its initial value."""
    g, _ = setup(sa.DEPOSIT_ATOMIC, precision=sa.MIXED)
    p0 = g.get_particles(1)
    g0, rho0 = g.gauss(), np.abs(g.rho()).max()
    g.step(3)
    p3 = g.get_particles(1)
    for k in ("x", "y", "z", "px", "py", "pz", "w", "key", "patch"):
        np.testing.assert_array_equal(p3[k], p0[k], err_msg=k)
    g.step(5)
    assert not np.array_equal(g.get_particles(1)["x"], p0["x"])
    assert np.abs(g.gauss() - g0).max() < 1e-6 * rho0
