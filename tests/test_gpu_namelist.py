"""Namelist front-end and the initialisers it relies on, on the GPU."""
import math

import numpy as np
import pytest
from scipy.special import kve

import smilei_amd as sa
from smilei_amd import namelist as N
from tests.plasma import cfl_dt

pytestmark = pytest.mark.gpu


def sim3d(precision=sa.PLAIN_FP64, shape=(32, 32, 32)):
    dx = (0.2, 0.2, 0.2)
    return sa.Simulation("3Dcartesian", 2, shape, (16, 16, 16), dx, cfl_dt(dx), precision=precision)


@pytest.mark.parametrize("theta", [0.02, 0.2])
def test_maxwell_juttner_mean_energy_and_isotropy(theta):
    """<gamma> of the Maxwell-Juttner distribution is 3 theta +
    K1(1/theta) / K2(1/theta) (closed form); components isotropic."""
    sim = sim3d()
    sid = sim.add_species("e", 1.0, -1.0, 32 ** 3 * 16)
    sim.init_species_mj(sid, 16, 1.0, theta, seed=4)
    p = sim.get_particles(sid)
    u = np.stack([p["px"], p["py"], p["pz"]])
    gam = np.sqrt(1.0 + (u ** 2).sum(0))
    x = 1.0 / theta
    expect = 3.0 * theta + kve(1, x) / kve(2, x)   # exponentially scaled Bessel K: ratio unchanged
    err = gam.std() / math.sqrt(gam.size)
    assert abs(gam.mean() - expect) < 5 * err + 1e-6 * expect, (gam.mean(), expect, err)
    m2 = (u ** 2).mean(1)
    np.testing.assert_allclose(m2, m2.mean(), rtol=0.02)
    np.testing.assert_allclose(u.mean(1), 0.0, atol=5 * np.sqrt(m2.mean() / u.shape[1]))


@pytest.mark.parametrize("precision", [sa.PLAIN_FP64, sa.MIXED])
def test_positions_copied_from_another_species(precision):
    sim = sim3d(precision)
    a = sim.add_species("ion", 1836.0, 1.0, 32 ** 3 * 4)
    sim.init_species_mj(a, 4, 1.0, 1e-3 / 1836.0, seed=1)
    b = sim.add_species("e", 1.0, -1.0, 32 ** 3 * 4)
    sim.init_species_like(b, a, weight_scale=2.0, theta=1e-3, seed=2)
    pa, pb = sim.get_particles(a), sim.get_particles(b)
    for k in ("x", "y", "z", "key", "patch"):
        np.testing.assert_array_equal(pa[k], pb[k])
    np.testing.assert_allclose(pb["w"], 2.0 * pa["w"], rtol=1e-7)
    assert not np.allclose(pa["px"] * math.sqrt(1836.0), pb["px"])  # own momenta
    # neutral at t = 0 up to the weight ratio: rho = 1 - 2 everywhere the species sit
    sim.step(3)
    sim.synchronize()


NAMELIST = """
import math as m
T = 10./511.
dx = 0.5*m.sqrt(T)
Main(geometry = "3Dcartesian", interpolation_order = 2, timestep_over_CFL = 0.95,
     number_of_timesteps = 40, cell_length = [dx]*3, number_of_cells = [32, 32, 32],
     number_of_patches = [4, 4, 2], EM_boundary_conditions = [["periodic"]], random_seed = 3)
Vectorization(mode = "off")
Species(name = "proton", position_initialization = "random", momentum_initialization = "mj",
        particles_per_cell = 8, mass = 1836.0, charge = 1.0, charge_density = 1.,
        temperature = [T], boundary_conditions = [["periodic", "periodic"]] * 3)
Species(name = "electron", position_initialization = "proton", momentum_initialization = "mj",
        particles_per_cell = 8, mass = 1.0, charge = -1.0, charge_density = 1.,
        temperature = [T], boundary_conditions = [["periodic", "periodic"]] * 3)
DiagScalar(every = 10)
"""


def test_namelist_thermal_plasma_runs_both_forms(tmp_path):
    out = {}
    for name, prec in (("fp64", sa.PLAIN_FP64), ("mixed", sa.MIXED)):
        sim, rows = N.run(NAMELIST, precision=prec, scalars_path=str(tmp_path / ("scalars_%s.txt" % name)))
        assert rows.shape == (5, 6)  # t = 0, 10, 20, 30, 40 steps; time Utot Ukin Uelm Ukin_p Ukin_e
        assert sim.num_particles(0) == sim.num_particles(1) == 32 ** 3 * 8
        # fields start at zero (electrons sit on the protons: the plasma is neutral)
        assert rows[0, 3] == 0.0
        # total energy conserved to the level of a short thermal run
        assert abs(rows[-1, 1] / rows[0, 1] - 1) < 1e-2
        out[name] = rows
        sim.close()
    # same seeds, same initial state: mixed and FP64 agree
    np.testing.assert_allclose(out["mixed"][:, 1], out["fp64"][:, 1], rtol=1e-5)
    text = (tmp_path / "scalars_fp64.txt").read_text().splitlines()[0]
    assert text.split()[1:] == ["time", "Utot", "Ukin", "Uelm", "Ukin_proton", "Ukin_electron"]
