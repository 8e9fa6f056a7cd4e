"""The mixed-precision form against the FP64 form at the full sizes of the
workload configurations, each form run on its own from the same seeds (the
particles are drawn in HBM; the mixed form holds them rounded to its FP32
offsets and momenta).

* 3D thermal plasma with 4th-order shapes: 256x128x128 cells, dx = 0.2,
  16^3-cell patches, electrons and protons at 8 particles per cell each
  (6.7e7 particles), sigma 0.05 m_e c (protons: the same temperature), 1000
  steps: field, kinetic and total energy differences <= 1e-5 of the total
  energy, Gauss-law drift of the mixed form <= 1e-6 of max|rho(0)|.
* 3D beam-plasma: 1024x256x256 cells, dx = 0.25, 32x16x16-cell patches,
  e-/p+ background at 8 particles per cell each (sigma 0.02 m_e c), cold
  electron beam gamma = 50, n = 0.1, 8 particles per cell within 16 cells
  of the axis (1.08e9 particles), its configured 300 steps: field, kinetic
  and total energy differences <= 1e-5 of the total energy, Gauss drift
  <= 1e-6.  One simulation is held at a time (the FP64 state of 1.08e9
  particles takes 134 GB).
* Laser-wakefield-like uneven patches: 512x128x128 cells, dx = 0.1, dy = dz
  = 0.5, 1024 patches of 32x16x16 holding 0 to 4096 electrons (ramp +
  channel, checkerboard particles per cell), a0 = 2 laser as initial field,
  500 graph-captured steps; the FP64 form in its autotuned launch shape, the
  mixed form in the persistent work-list launch: field and kinetic energy
  differences <= 1e-5 of the total energy, Gauss drift <= 1e-6 in both.
"""
import math

import numpy as np
import pytest

import smilei_amd as sa
from tests.plasma import lwfa

pytestmark = pytest.mark.gpu


def cfl_dt(dx, factor=0.95):
    return factor / math.sqrt(sum(1.0 / d ** 2 for d in dx))


def run(shape, patch, dx, order, species, precision, steps):
    """Energies after `steps` (and the Gauss drift) of one run from seeds."""
    sim = sa.Simulation("3Dcartesian", order, shape, patch, dx, cfl_dt(dx), precision=precision,
                        deposit=sa.DEPOSIT_ATOMIC, use_graph=True)
    ncell = int(np.prod(shape))
    for k, (name, mass, q, sigma, ppc, density, drift, cells, radius) in enumerate(species):
        sid = sim.add_species(name, mass, q, (cells or ncell) * ppc)
        sim.init_species(sid, ppc, density, sigma, drift_ux=drift, radius=radius, seed=1 + k)
    g0 = sim.gauss()
    rho0 = float(np.abs(sim.rho()).max())
    print("run: %s, %d steps" % ("mixed" if precision == sa.MIXED else "fp64", steps), flush=True)
    sim.step(steps)
    e = sim.energies()
    drift = float(np.abs(sim.gauss() - g0).max()) / rho0
    sim.close()
    return e, drift


def diff(ea, eb):
    tot = eb.sum()
    return dict(field=abs(ea[0] - eb[0]) / tot, kinetic=abs(ea[1:].sum() - eb[1:].sum()) / tot,
                total=abs(ea.sum() - tot) / tot, field_rel=abs(ea[0] - eb[0]) / abs(eb[0]),
                kinetic_rel=abs(ea[1:].sum() - eb[1:].sum()) / abs(eb[1:].sum()))


def test_thermal_order4_full_size_mixed_vs_fp64_1000_steps():
    shape, patch, dx = (256, 128, 128), (16, 16, 16), (0.2, 0.2, 0.2)
    species = [("electron", 1.0, -1.0, 0.05, 8, 1.0, 0.0, 0, 0.0),
               ("proton", 1836.0, 1.0, 0.05 / math.sqrt(1836.0), 8, 1.0, 0.0, 0, 0.0)]
    em, drift_m = run(shape, patch, dx, 4, species, sa.MIXED, 1000)
    ef, drift_f = run(shape, patch, dx, 4, species, sa.PLAIN_FP64, 1000)
    err = diff(em, ef)
    print("order-4 thermal full size, mixed vs fp64 after 1000 steps:", err, "gauss drift", drift_m, drift_f)
    assert err["field"] < 1e-5 and err["kinetic"] < 1e-5 and err["total"] < 1e-5, err
    assert drift_m < 1e-6 and drift_f < 1e-6


def beam_cells(shape, radius):
    j = np.arange(shape[1]) + 0.5 - 0.5 * shape[1]
    k = np.arange(shape[2]) + 0.5 - 0.5 * shape[2]
    return int(np.count_nonzero(j[:, None] ** 2 + k[None, :] ** 2 <= radius * radius)) * shape[0]


def test_beam_plasma_full_size_mixed_vs_fp64_300_steps():
    shape, patch, dx = (1024, 256, 256), (32, 16, 16), (0.25, 0.25, 0.25)
    ub = math.sqrt(50.0 ** 2 - 1.0)
    species = [("electron", 1.0, -1.0, 0.02, 8, 1.0, 0.0, 0, 0.0),
               ("proton", 1836.0, 1.0, 0.02 / math.sqrt(1836.0), 8, 1.0, 0.0, 0, 0.0),
               ("beam", 1.0, -1.0, 0.0, 8, 0.1, ub, beam_cells(shape, 16.0), 16.0)]
    em, drift_m = run(shape, patch, dx, 2, species, sa.MIXED, 300)
    ef, drift_f = run(shape, patch, dx, 2, species, sa.PLAIN_FP64, 300)
    err = diff(em, ef)
    print("beam-plasma full size, mixed vs fp64 after 300 steps:", err, "gauss drift", drift_m, drift_f)
    assert err["field"] < 1e-5 and err["kinetic"] < 1e-5 and err["total"] < 1e-5, err
    assert drift_m < 1e-6 and drift_f < 1e-6


def test_lwfa_full_size_mixed_vs_fp64_500_steps():
    f = lwfa(sa.Simulation, sa.PLAIN_FP64, use_graph=True)
    counts = f.patch_counts(0)
    assert counts.min() == 0 and counts.max() == 4096  # uneven: empty and full patches
    f.autotune(reps=2)
    m = lwfa(sa.Simulation, sa.MIXED, use_graph=True)
    m.set_tuning(sa.TUNE_PERSISTENT, 1)
    rho0 = float(np.abs(f.rho()).max())
    gf0, gm0 = f.gauss(), m.gauss()
    print("run: lwfa, 500 steps each", flush=True)
    f.step(500)
    m.step(500)
    ef, em = f.energies(), m.energies()
    err = diff(em, ef)
    drift_f = float(np.abs(f.gauss() - gf0).max()) / rho0
    drift_m = float(np.abs(m.gauss() - gm0).max()) / rho0
    print("lwfa full size, mixed vs fp64 after 500 steps:", err, "gauss drift", drift_m, drift_f)
    assert err["field"] < 1e-5 and err["kinetic"] < 1e-5 and err["total"] < 1e-5, err
    assert drift_m < 1e-6 and drift_f < 1e-6
    f.close()
    m.close()
