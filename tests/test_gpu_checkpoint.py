"""Checkpoint / restart: a run restarted from a checkpoint continues as the
uninterrupted run (bit for bit in the deterministic FP64 form)."""
import numpy as np
import pytest

import smilei_amd as sa
from tests.plasma import cfl_dt, load, thermal

pytestmark = pytest.mark.gpu

FIELDS = ("Ex", "Ey", "Ez", "Bx", "By", "Bz", "Bx_m", "By_m", "Bz_m", "Jx", "Jy", "Jz")


def make(precision, deposit):
    shape, dx = (32, 32, 32), (0.2, 0.2, 0.2)
    sim = sa.Simulation("3Dcartesian", 2, shape, (16, 16, 16), dx, cfl_dt(dx), precision=precision, deposit=deposit)
    for k, (name, m, q, sig) in enumerate([("electron", 1.0, -1.0, 0.05), ("proton", 1836.0, 1.0, 0.05 / 42.85)]):
        p = thermal(shape, dx, 8, sig, 60 + k)
        load(sim, sim.add_species(name, m, q, len(p["x"]) + 100), p)
    return sim


@pytest.mark.parametrize("precision,deposit,tol", [(sa.PLAIN_FP64, sa.DEPOSIT_DETERMINISTIC, 0.0),
                                                   (sa.MIXED, sa.DEPOSIT_ATOMIC, 1e-6)])
def test_restart_continues_the_run(tmp_path, precision, deposit, tol):
    a = make(precision, deposit)
    a.step(5)
    path = str(tmp_path / "ckpt.npz")
    a.checkpoint(path)
    a.step(5)
    b = sa.Simulation.restart(path)
    assert b.steps_done == 5 and b.species == a.species
    b.step(5)
    assert b.steps_done == a.steps_done == 10
    for s in range(2):
        pa, pb = a.get_particles(s), b.get_particles(s)
        np.testing.assert_array_equal(pa["key"], pb["key"])
        for k in ("x", "y", "z", "px", "py", "pz", "w"):
            if tol == 0.0:
                np.testing.assert_array_equal(pa[k], pb[k], err_msg=k)
            else:
                assert np.abs(pa[k] - pb[k]).max() <= tol * np.abs(pa[k]).max(), k
    for f in FIELDS:
        fa, fb = a.get_field(f), b.get_field(f)
        if tol == 0.0:
            np.testing.assert_array_equal(fa, fb, err_msg=f)
        else:
            assert np.abs(fa - fb).max() <= tol * max(np.abs(fa).max(), 1e-300), f


def test_restarted_run_matches_oracle(tmp_path):
    """Deterministic FP64 form: 5 steps, a checkpoint, a new simulation
    restarted from it, 5 more steps; bit-identical (particles, keys, counts,
    every field) to the oracle run 10 steps from the same initial state."""
    from oracle import pyoracle as O
    from tests.test_gpu_configs import compare, copy_to_oracle
    a = make(sa.PLAIN_FP64, sa.DEPOSIT_DETERMINISTIC)
    dx = (0.2, 0.2, 0.2)
    o = O.OracleSim(3, 2, (32, 32, 32), (16, 16, 16), dx, cfl_dt(dx))
    copy_to_oracle(a, 2, [1.0, 1836.0], [-1.0, 1.0], o)
    a.step(5)
    path = str(tmp_path / "ckpt.npz")
    a.checkpoint(path)
    a.close()
    b = sa.Simulation.restart(path)
    b.step(5)
    o.step(10)
    compare(b, o, 2, exact=True)
