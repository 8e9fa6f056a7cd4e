"""Long FP64 runs against the oracle from identical particles.

* The 3D thermal plasma of the FP64 configuration (16^3-cell patches,
  electrons and protons at 8 particles per cell each, sigma 0.05 m_e c, the
  protons' sigma in their own mass units) on a 64^3-cell box, 4.2e6
  particles: 100 steps with the atomic tile deposit (keys, owner patches and
  per-patch counts bit-exact; particles and E, B, J within 1e-12 of their
  largest value: the only difference is the order of the FP64 J sums), and
  200 steps with the deterministic deposit (everything bit-exact).
* The beam-plasma (cold gamma = 50 electron beam through an e-/p+
  background) on a 128x64x64 box, 9.2e6 particles, 100 steps of the atomic
  FP64 form: same bounds as above.
"""
import numpy as np
import pytest

import smilei_amd as sa
from oracle import pyoracle as O
from tests.plasma import cfl_dt, load, thermal
from tests.test_gpu_configs import compare

pytestmark = pytest.mark.gpu


def thermal_species(shape, dx):
    return [("electron", 1.0, -1.0, thermal(shape, dx, 8, 0.05, 301)),
            ("proton", 1836.0, 1.0, thermal(shape, dx, 8, 0.05 * np.sqrt(1.0 / 1836.0), 302))]


def beam_species(shape, dx):
    ub = float(np.sqrt(50.0 ** 2 - 1.0))
    return [("electron", 1.0, -1.0, thermal(shape, dx, 8, 0.02, 311)),
            ("proton", 1836.0, 1.0, thermal(shape, dx, 8, 0.02 / np.sqrt(1836.0), 312)),
            ("beam", 1.0, -1.0, thermal(shape, dx, 8, 0.0, 313, density=0.1, drift_ux=ub, radius=16.0))]


def pair(shape, patch, dx, species, deposit):
    dt = cfl_dt(dx)
    g = sa.Simulation("3Dcartesian", 2, shape, patch, dx, dt, deposit=deposit)
    o = O.OracleSim(3, 2, shape, patch, dx, dt)
    for name, m, q, p in species:
        load(g, g.add_species(name, m, q, len(p["x"])), p)
        o.set_particles(o.add_species(m, q), p["x"], p["y"], p["z"], p["px"], p["py"], p["pz"], p["w"])
    return g, o


def test_thermal3d_64cubed_atomic_100_steps_vs_oracle():
    shape, patch, dx = (64, 64, 64), (16, 16, 16), (0.2, 0.2, 0.2)
    g, o = pair(shape, patch, dx, thermal_species(shape, dx), sa.DEPOSIT_ATOMIC)
    g.step(100)
    o.step(100)
    compare(g, o, 2, exact=False, tol=1e-12)
    np.testing.assert_allclose(g.energies(), o.energies(), rtol=1e-12)


def test_thermal3d_64cubed_deterministic_200_steps_bit_exact():
    shape, patch, dx = (64, 64, 64), (16, 16, 16), (0.2, 0.2, 0.2)
    g, o = pair(shape, patch, dx, thermal_species(shape, dx), sa.DEPOSIT_DETERMINISTIC)
    g.step(200)
    o.step(200)
    compare(g, o, 2, exact=True)
    np.testing.assert_allclose(g.energies(), o.energies(), rtol=1e-12)


def test_beam_plasma_fp64_100_steps_vs_oracle():
    shape, patch, dx = (128, 64, 64), (32, 16, 16), (0.25, 0.25, 0.25)
    g, o = pair(shape, patch, dx, beam_species(shape, dx), sa.DEPOSIT_ATOMIC)
    g.step(100)
    o.step(100)
    compare(g, o, 3, exact=False, tol=1e-12)
    np.testing.assert_allclose(g.energies(), o.energies(), rtol=1e-12)
