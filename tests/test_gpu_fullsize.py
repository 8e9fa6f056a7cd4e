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
"""
import math

import numpy as np
import pytest

import smilei_amd as sa

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
