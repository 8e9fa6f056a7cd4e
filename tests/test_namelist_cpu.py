"""Namelist front-end on the CPU: block parsing, derived parameters, and
refusal of what the hot path does not implement (no device calls)."""
import math

import pytest

from smilei_amd import namelist as N

THERMAL = """
import math as m
T = 10./511.
dx = 0.5*m.sqrt(T)
dt = 0.95 * dx/m.sqrt(3.)
Main(
    geometry = "3Dcartesian",
    interpolation_order = 2,
    timestep = dt,
    simulation_time = 2.*m.pi,
    cell_length = [dx, dx, dx],
    grid_length = [32.*dx, 32.*dx, 32.*dx],
    number_of_patches = [4, 4, 4],
    EM_boundary_conditions = [["periodic"]],
)
LoadBalancing(every = 20)
Vectorization(mode = "off")
Species(name = "proton", position_initialization = "random", momentum_initialization = "mj",
        particles_per_cell = 8, mass = 1836.0, charge = 1.0, charge_density = 1.,
        temperature = [T], boundary_conditions = [["periodic", "periodic"]] * 3)
Species(name = "electron", position_initialization = "proton", momentum_initialization = "mj",
        particles_per_cell = 8, mass = 1.0, charge = -1.0, charge_density = 1.,
        temperature = [T], boundary_conditions = [["periodic", "periodic"]] * 3)
DiagFields(every = 4)
DiagScalar(every = 2)
"""


def test_derived_parameters_follow_the_reference_rules():
    nl = N.load(THERMAL)
    m = nl.main
    T = 10. / 511.
    dx = 0.5 * math.sqrt(T)
    assert m.number_of_cells == [32, 32, 32]
    dt = 0.95 * dx / math.sqrt(3.)
    # n_time = (int)(simulation_time / timestep), Params.cpp:1154
    assert m.number_of_timesteps == int(2. * math.pi / dt)
    assert [s.name for s in nl.species] == ["proton", "electron"]
    assert nl.diag_scalar.every == 2
    assert nl.ignored == ["LoadBalancing", "Vectorization", "DiagFields"]


def test_timestep_from_cfl_and_cells_from_number():
    nl = N.load("""
Main(geometry = "2Dcartesian", number_of_cells = [64, 32], cell_length = [0.1, 0.2],
     timestep_over_CFL = 0.95, number_of_timesteps = 7, number_of_patches = [4, 2])
""")
    assert nl.main.grid_length == pytest.approx([6.4, 6.4])
    assert nl.main.timestep == pytest.approx(0.95 / math.sqrt(1 / 0.01 + 1 / 0.04))
    assert nl.main.number_of_timesteps == 7


@pytest.mark.parametrize("block", ["LaserGaussian3D(a0=1.)", "MovingWindow(time_start=0.)",
                                   "Collisions(species1=['e'])", "CurrentFilter(passes=[2])"])
def test_physics_outside_the_hot_path_is_refused(block):
    with pytest.raises(N.NamelistError, match="not part of this library"):
        N.load(THERMAL + block + "\n")


@pytest.mark.parametrize("text,msg", [
    ("Main(geometry='1Dcartesian', cell_length=[0.1], grid_length=[1.], number_of_patches=[1], timestep=0.05, "
     "simulation_time=1.)\n", "geometry"),
    ("Main(geometry='3Dcartesian', cell_length=[0.1]*3, grid_length=[1.6]*3, number_of_patches=[3]*3, "
     "timestep=0.05, simulation_time=1.)\n", "divide"),
    ("Main(geometry='3Dcartesian', cell_length=[0.1]*3, grid_length=[1.6]*3, number_of_patches=[2]*3, "
     "timestep=0.05, simulation_time=1., EM_boundary_conditions=[['silver-muller']])\n", "periodic"),
])
def test_main_validation(text, msg):
    with pytest.raises(N.NamelistError, match=msg):
        N.load(text)


def test_species_validation():
    base = ("Main(geometry='3Dcartesian', cell_length=[0.1]*3, grid_length=[1.6]*3, number_of_patches=[2]*3, "
            "timestep=0.05, simulation_time=1.)\n")
    with pytest.raises(N.NamelistError, match="momentum_initialization"):
        N.load(base + "Species(name='e', mass=1., charge=-1., particles_per_cell=4, number_density=1., "
                      "momentum_initialization='rectangular', temperature=[0.5])\n")
    with pytest.raises(N.NamelistError, match="isotropic"):
        N.load(base + "Species(name='e', mass=1., charge=-1., particles_per_cell=4, number_density=1., "
                      "momentum_initialization='mj', temperature=[0.5, 0.1, 0.1])\n")
    with pytest.raises(N.NamelistError, match="earlier species"):
        N.load(base + "Species(name='e', mass=1., charge=-1., particles_per_cell=4, number_density=1., "
                      "position_initialization='ion')\n")
    with pytest.raises(N.NamelistError, match="mean_velocity"):
        N.load(base + "Species(name='e', mass=1., charge=-1., particles_per_cell=4, number_density=1., "
                      "mean_velocity=[0., 0.1, 0.])\n")
