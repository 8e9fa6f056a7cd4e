"""Namelist front-end and the initialisers it relies on, on the GPU."""
import math

import numpy as np
import pytest
from scipy.special import kve

import smilei_amd as sa
from smilei_amd import namelist as N
from tests.plasma import cfl_dt
from tests.test_namelist_cpu import happi_read

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


def test_drifting_maxwell_juttner_boost():
    """A drifting Maxwell-Juttner species (theta = 0.2, u_d = g V = 0.5
    along x) is the rest-frame draw Lorentz-boosted with Zenitani's
    volume-transform correction.  The boosted (drifting Juttner) sample has,
    per particle, <gamma> = g (<gamma>_0 + theta) - theta / g and <u_x> =
    u_d (<gamma>_0 + theta), from the lab-frame energy and momentum densities
    of a drifting ideal gas (T^00 = (e + p) g^2 - p, T^0x = (e + p) g^2 V, per
    N^0 = n g, p = n theta).  Boosting every particle without the correction
    would give g <gamma>_0 and u_d <gamma>_0 instead, about 200 standard
    errors away at this sample size."""
    theta, ud = 0.2, 0.5
    sim = sim3d()
    sid = sim.add_species("e", 1.0, -1.0, 32 ** 3 * 16)
    sim.init_species_mj(sid, 16, 1.0, theta, drift_ux=ud, seed=6)
    p = sim.get_particles(sid)
    u = np.stack([p["px"], p["py"], p["pz"]])
    gam = np.sqrt(1.0 + (u ** 2).sum(0))
    x = 1.0 / theta
    g0 = 3.0 * theta + kve(1, x) / kve(2, x)
    gb = math.sqrt(1.0 + ud * ud)
    n = gam.size
    assert abs(gam.mean() - (gb * (g0 + theta) - theta / gb)) < 5 * gam.std() / math.sqrt(n)
    assert abs(u[0].mean() - ud * (g0 + theta)) < 5 * u[0].std() / math.sqrt(n)
    # transverse components are left as drawn: isotropic, zero mean
    np.testing.assert_allclose(u[1:].mean(1), 0.0, atol=5 * np.sqrt((u[1:] ** 2).mean() / n))
    np.testing.assert_allclose((u[1] ** 2).mean(), (u[2] ** 2).mean(), rtol=0.02)


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
        # electrons sit on the protons: rho is zero up to the rounding of its
        # sums, so the Poisson solve at t = 0 (Smilei.cpp:329-332, run when
        # rho is not exactly zero) leaves a field energy at rounding level
        assert rows[0, 3] <= 1e-20 * rows[0, 2]
        # total energy conserved to the level of a short thermal run
        assert abs(rows[-1, 1] / rows[0, 1] - 1) < 1e-2
        out[name] = rows
        sim.close()
    # same seeds, same initial state: mixed and FP64 agree
    np.testing.assert_allclose(out["mixed"][:, 1], out["fp64"][:, 1], rtol=1e-5)
    # scalars.txt as happi reads it: same records, the reference's columns
    nl = N.load(NAMELIST)
    names, steps, v = happi_read(str(tmp_path / "scalars_fp64.txt"), nl.main.timestep)
    assert steps == [0, 10, 20, 30, 40]
    assert names == ["time"] + N.scalar_names(["proton", "electron"])
    rows = out["fp64"]
    np.testing.assert_allclose(v["Utot"], rows[:, 1], rtol=1e-10)  # 10 digits printed
    np.testing.assert_allclose(v["Ukin_electron"], rows[:, 5], rtol=1e-10)
    np.testing.assert_allclose(np.sum([v["Uelm_" + f] for f in ("Ex", "Ey", "Ez", "Bx_m", "By_m", "Bz_m")], 0),
                               rows[:, 3], rtol=1e-10, atol=1e-300)
    assert v["Uexp"] == [v["Utot"][0]] * 5
    np.testing.assert_allclose(v["Ubal"], np.array(v["Utot"]) - v["Utot"][0], rtol=0, atol=1e-10 * v["Utot"][0])
    assert v["Ntot_proton"] == [32 ** 3 * 8.0] * 5 and v["Zavg_electron"] == [-1.0] * 5
    # sum of weights = density (1 / |charge|) x volume
    vol = (32 * nl.main.cell_length[0]) ** 3
    np.testing.assert_allclose(v["Dens_proton"], vol, rtol=1e-10)


def test_checkpoint_dump_then_restart_continues_bit_for_bit(tmp_path):
    """Checkpoints(dump_step, exit_after_dump): the run stops after its first
    dump; a second run with restart_dir continues from it to the namelist's
    end.  Deterministic FP64 deposit: the result equals an uninterrupted run
    bit for bit, and scalars.txt holds each record once."""
    base = NAMELIST.replace("DiagScalar(every = 10)", "DiagScalar(every = 5)")
    ref_dir, a_dir, b_dir = tmp_path / "ref", tmp_path / "a", tmp_path / "b"
    for d in (ref_dir, a_dir, b_dir):
        d.mkdir()
    ref, _ = N.run(base, deposit=sa.DEPOSIT_DETERMINISTIC, output_dir=str(ref_dir))
    first, _ = N.run(base + "Checkpoints(dump_step = 15, exit_after_dump = True)\n",
                     deposit=sa.DEPOSIT_DETERMINISTIC, output_dir=str(a_dir))
    assert first.steps_done == 15 and (a_dir / "checkpoints" / "dump-00000.npz").exists()
    first.close()
    # the restarted run writes into the first run's directory (its scalars.txt continues)
    second, _ = N.run(base + "Checkpoints(restart_dir = %r)\n" % str(a_dir),
                      deposit=sa.DEPOSIT_DETERMINISTIC, output_dir=str(a_dir))
    assert second.steps_done == ref.steps_done == 40
    for s in range(2):
        pa, pb = ref.get_particles(s), second.get_particles(s)
        for k in ("x", "y", "z", "px", "py", "pz", "w"):
            np.testing.assert_array_equal(pa[k], pb[k], err_msg=k)
    for f in ("Ex", "By_m", "Jz"):
        np.testing.assert_array_equal(ref.get_field(f), second.get_field(f))
    nl = N.load(base)
    names, steps, v = happi_read(str(a_dir / "scalars.txt"), nl.main.timestep)
    _, steps_ref, v_ref = happi_read(str(ref_dir / "scalars.txt"), nl.main.timestep)
    assert steps == steps_ref == list(range(0, 41, 5))
    assert v == v_ref  # Uexp too: Energy_time_zero travels with the dump
    ref.close()
    second.close()


def test_signals_dump_and_stop_or_go_on(tmp_path):
    """SIGUSR2 dumps and goes on, SIGUSR1 dumps and stops (with
    exit_after_dump = False), as Checkpoint::dump does."""
    import signal

    def raise_at(it):
        if it == 5:
            signal.raise_signal(signal.SIGUSR2)
        if it == 9:
            signal.raise_signal(signal.SIGUSR1)

    text = NAMELIST + "Checkpoints(exit_after_dump = False, keep_n_dumps = 3)\n"
    before = signal.getsignal(signal.SIGUSR1), signal.getsignal(signal.SIGUSR2)
    sim, _ = N.run(text, output_dir=str(tmp_path), on_step=raise_at)
    assert sim.steps_done == 9
    dumps = sorted((tmp_path / "checkpoints").glob("dump-*.npz"))
    assert [d.name for d in dumps] == ["dump-00000.npz", "dump-00001.npz"]
    assert [sa.Simulation.restart(str(d)).steps_done for d in dumps] == [5, 9]
    # the handlers in place before the run are back
    assert (signal.getsignal(signal.SIGUSR1), signal.getsignal(signal.SIGUSR2)) == before


NONNEUTRAL = """
import math as m
dx = 0.2
Main(geometry = "3Dcartesian", interpolation_order = 2, timestep_over_CFL = 0.95,
     number_of_timesteps = 20, cell_length = [dx]*3, number_of_cells = [16, 16, 16],
     number_of_patches = [2, 2, 2], EM_boundary_conditions = [["periodic"]], random_seed = 9%s)
Species(name = "proton", position_initialization = "random", momentum_initialization = "mj",
        particles_per_cell = 4, mass = 1836.0, charge = 1.0, charge_density = 1.,
        temperature = [0.01], boundary_conditions = [["periodic", "periodic"]] * 3)
Species(name = "electron", position_initialization = "random", momentum_initialization = "mj",
        particles_per_cell = 4, mass = 1.0, charge = -1.0, charge_density = 1.,
        temperature = [0.01], boundary_conditions = [["periodic", "periodic"]] * 3)
ExternalField(field = "Bz", profile = lambda x, y, z: 0.05 * m.cos(2 * m.pi * x / 3.2))
"""


def test_namelist_solves_poisson_for_a_nonneutral_start():
    """Electrons and protons at independent random positions: rho != 0
    locally, so (solve_poisson defaults to True, Smilei.cpp:329-332) E starts
    as the electrostatic field of rho and div E - rho vanishes to rounding at
    step 0 (the total charge is zero); with solve_poisson = False E starts at
    zero and the residual is -rho."""
    sim = N.build(N.load(NONNEUTRAL % ""))
    rho = sim.rho()
    assert np.abs(rho).max() > 0.1
    assert np.abs(sim.gauss()).max() < 1e-12 * np.abs(rho).max()
    off = N.build(N.load(NONNEUTRAL % ", solve_poisson = False"))
    # E = 0: the residual is -rho (two evaluations of rho, each summed by
    # atomics in its own order: equal to rounding)
    r = off.rho()
    np.testing.assert_allclose(off.gauss(), -r, rtol=0, atol=1e-13 * np.abs(r).max())


def test_namelist_run_matches_oracle_bit_exact(tmp_path):
    """A namelist-built simulation (parameters derived from the namelist:
    timestep from timestep_over_CFL, patches from number_of_patches, species
    from profiles and temperatures, E from Poisson, an external B profile)
    run in the deterministic FP64 form equals the oracle started from the
    same particles and fields, bit for bit, after 20 steps; its scalars.txt
    record (written and read back as happi reads it) holds the oracle's
    per-field and per-species energies (DiagnosticScalar.cpp:494-522,
    660-690) to the 10 printed digits."""
    from oracle import pyoracle as O
    from tests.test_gpu_configs import compare, copy_to_oracle
    nl = N.load(NONNEUTRAL % "")
    sim = N.build(nl, deposit=sa.DEPOSIT_DETERMINISTIC)
    dx = (0.2, 0.2, 0.2)
    dt = 0.95 / math.sqrt(3 / 0.2 ** 2)
    assert abs(nl.main.timestep - dt) < 1e-15
    o = O.OracleSim(3, 2, (16, 16, 16), (8, 8, 8), dx, dt)
    copy_to_oracle(sim, 2, [1836.0, 1.0], [1.0, -1.0], o)
    n = nl.main.number_of_timesteps
    sim.step(n)
    o.step(n)
    compare(sim, o, 2, exact=True)
    names = N.scalar_names(sim.species)
    vals, _ = N.scalar_values(sim, [1.0, -1.0], names, None)
    out = N.ScalarFile(str(tmp_path / "scalars.txt"), names)
    out.write(n * nl.main.timestep, vals)
    out.close()
    got, steps, v = happi_read(str(tmp_path / "scalars.txt"), nl.main.timestep)
    assert got == ["time"] + names and steps == [n]
    fe, en = o.field_energies(), o.energies()
    for k, f in enumerate(("Ex", "Ey", "Ez", "Bx_m", "By_m", "Bz_m")):
        np.testing.assert_allclose(v["Uelm_" + f][0], fe[k], rtol=1e-9, atol=1e-12 * en.sum(), err_msg=f)
    for k, sp in enumerate(sim.species):
        np.testing.assert_allclose(v["Ukin_" + sp][0], en[1 + k], rtol=1e-9, err_msg=sp)
    np.testing.assert_allclose(v["Uelm"][0], en[0], rtol=1e-9)
    np.testing.assert_allclose(v["Utot"][0], en.sum(), rtol=1e-9)
