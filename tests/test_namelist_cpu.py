"""Namelist front-end on the CPU: block parsing, derived parameters, and
refusal of what the hot path does not implement (no device calls)."""
import math

import numpy as np
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


def happi_read(path, timestep):
    """Read scalars.txt the way happi does (happi/_Diagnostics/Scalar.py:
    64-85): the names are the last comment line, each row is split on
    whitespace, the time column becomes a step number."""
    with open(path) as f:
        lines = f.read().splitlines()
    k = 0
    while lines[k].startswith("#"):
        k += 1
    names = lines[k - 1][1:].split()
    rows = [[float(v) for v in L.split()] for L in lines[k:]]
    steps = [int(round(r[0] / timestep)) for r in rows]
    return names, steps, {n: [r[i] for r in rows] for i, n in enumerate(names)}


def test_scalar_file_layout(tmp_path):
    """The text layout of DiagnosticScalar::write: "# j name" lines, a "#"
    line, the names right-aligned in 2 + max(len, precision + 8) columns
    after a 1 + precision + 9 wide time column, then %.{precision}e values."""
    path = tmp_path / "scalars.txt"
    f = N.ScalarFile(str(path), ["Utot", "Ukin_electron"], precision=10)
    f.write(0.5, [1.25, -3e-5])
    f.write(1.0, [2.0, float("nan")])
    f.close()
    assert path.read_text().splitlines() == [
        "# 1 time",
        "# 2 Utot",
        "# 3 Ukin_electron",
        "#",
        "#               time                Utot       Ukin_electron",
        "    5.0000000000e-01    1.2500000000e+00   -3.0000000000e-05",
        "    1.0000000000e+00    2.0000000000e+00                 nan",
    ]
    names, steps, v = happi_read(str(path), 0.25)
    assert names == ["time", "Utot", "Ukin_electron"] and steps == [2, 4]
    assert v["Utot"] == [1.25, 2.0]
    # a long name widens its column; precision 3 narrows the others
    f = N.ScalarFile(str(tmp_path / "s3.txt"), ["Uelm_Bx_m", "Ubal_norm"], precision=3)
    f.write(0.0, [1.0, 0.0])
    f.close()
    lines = (tmp_path / "s3.txt").read_text().splitlines()
    assert lines[-2] == "#        time    Uelm_Bx_m    Ubal_norm"
    assert lines[-1] == "    0.000e+00    1.000e+00    0.000e+00"


class _FakeSim:
    species = ["ion", "electron"]

    def __init__(self, uelm, ntot, dens, ukin):
        self.v = dict(zip(N._UELM, uelm))
        self.v.update(Ntot=ntot, Dens=dens, Ukin=ukin)

    def scalars(self):
        return self.v


def test_scalar_energy_balance():
    """Utot = Ukin + Uelm, Uexp = Utot at t = 0, Ubal = Utot - Uexp,
    Ubal_norm = Ubal / Utot, Zavg = charge (SmileiMPI.cpp:2366-2424)."""
    names = N.scalar_names(_FakeSim.species)
    assert names[:3] == ["Ubal_norm", "Ubal", "Utot"] and names[15:20] == [
        "Dens_ion", "Ntot_ion", "Zavg_ion", "Ukin_ion", "Urad_ion"]
    assert names[-6:] == ["Uelm_Ex", "Uelm_Ey", "Uelm_Ez", "Uelm_Bx_m", "Uelm_By_m", "Uelm_Bz_m"]
    s0 = _FakeSim([1., 2., 0., 0., 0., 1.], [10, 20], [5., 4.], [30., 6.])
    v, e0 = N.scalar_values(s0, [1.0, -1.0], names, None)
    d = dict(zip(names, v))
    assert d["Uelm"] == 4.0 and d["Ukin"] == 36.0 and d["Utot"] == 40.0 and e0 == 40.0
    assert d["Uexp"] == 40.0 and d["Ubal"] == 0.0 and d["Ubal_norm"] == 0.0
    assert d["Zavg_ion"] == 1.0 and d["Zavg_electron"] == -1.0 and d["Ntot_electron"] == 20
    s1 = _FakeSim([1., 2., 0., 0., 0., 3.], [10, 20], [5., 0.], [30., 4.])
    v, _ = N.scalar_values(s1, [1.0, -1.0], names, e0)
    d = dict(zip(names, v))
    assert d["Utot"] == 40.0 and d["Ubal"] == 0.0
    s2 = _FakeSim([0.] * 6, [10, 0], [5., 0.], [44., 0.])
    v, _ = N.scalar_values(s2, [1.0, -1.0], names, e0)
    d = dict(zip(names, v))
    assert d["Ubal"] == 4.0 and d["Ubal_norm"] == 4.0 / 44.0
    assert math.isnan(d["Zavg_electron"])  # no weight: 0/0, as the reference divides


def test_diag_and_checkpoint_blocks_are_checked():
    with pytest.raises(N.NamelistError, match="Poynting"):
        N.load(THERMAL.replace("DiagScalar(every = 2)", "DiagScalar(every = 2, vars = ['Utot', 'PoyXmin'])"))
    with pytest.raises(N.NamelistError, match="keep_n_dumps"):
        N.load(THERMAL + "Checkpoints(dump_step = 5, keep_n_dumps = 0)\n")
    nl = N.load(THERMAL + "Checkpoints(dump_step = 5, exit_after_dump = False)\n")
    assert nl.checkpoints.dump_step == 5 and nl.checkpoints.exit_after_dump is False
    assert nl.diag_scalar.every == 2


@pytest.mark.parametrize("shape,dx", [((24, 16, 20), (0.3, 0.2, 0.25)), ((32, 12), (0.125, 0.2))])
def test_poisson_e_satisfies_discrete_gauss(shape, dx):
    """The divergence on the Yee grid, (E[i+1] - E[i]) / dx per axis at the
    primal node, of the field returned for a random periodic density is the
    density minus its mean, to rounding; and E has zero mean."""
    rng = np.random.default_rng(5)
    rho = rng.standard_normal(shape)
    E = N.poisson_e(rho, dx)
    div = sum((np.roll(E[d], -1, axis=d) - E[d]) / dx[d] for d in range(len(shape)))
    np.testing.assert_allclose(div, rho - rho.mean(), atol=1e-12 * np.abs(rho).max())
    for e in E:
        assert abs(e.mean()) < 1e-14 * np.abs(e).max()


def test_poisson_e_points_away_from_positive_charge():
    """A single positive cosine mode along x: E_x at the dual points is
    -d(phi)/dx with phi = rho / K^2, i.e. E_x(x) = sin(k (x - dx/2)) *
    (2 sin(k dx / 2) / dx) / K^2 with K = 2 sin(k dx / 2) / dx: the
    discrete counterpart of E = sin(k x) / k, pointing away from the
    density peak at x = 0."""
    N_, dx = 32, 0.25
    x = np.arange(N_) * dx
    k = 2 * np.pi / (N_ * dx)
    rho = np.cos(k * x)[:, None] * np.ones((1, 8))
    E = N.poisson_e(rho, (dx, 0.5))
    K = 2 * np.sin(k * dx / 2) / dx
    expect = np.sin(k * (x - dx / 2)) / K
    np.testing.assert_allclose(E[0][:, 0], expect, atol=1e-13)
    np.testing.assert_allclose(E[1], 0.0, atol=1e-13)
