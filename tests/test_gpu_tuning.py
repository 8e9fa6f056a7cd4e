"""Launch-shape knobs and the autotuner: they change neither the physics
nor the simulation state."""
import numpy as np
import pytest

import smilei_amd as sa
from tests.plasma import cfl_dt, load, thermal

pytestmark = pytest.mark.gpu

SPECIES = [("electron", 1.0, -1.0, 0.05), ("proton", 1836.0, 1.0, 0.05 / np.sqrt(1836.0))]


def make(precision, deposit=sa.DEPOSIT_ATOMIC, shape=(32, 32, 32), seed=9, order=2):
    dx = (0.2, 0.2, 0.2)
    sim = sa.Simulation("3Dcartesian", order, shape, (16, 16, 16), dx, cfl_dt(dx), precision=precision,
                        deposit=deposit)
    for k, (name, m, q, sig) in enumerate(SPECIES):
        p = thermal(shape, dx, 8, sig, seed + k)
        load(sim, sim.add_species(name, m, q, len(p["x"])), p)
    return sim


@pytest.mark.parametrize("threads,xcd", [(64, 0), (128, 1), (512, 0), (256, 1)])
def test_deterministic_form_independent_of_launch_shape(threads, xcd):
    a = make(sa.PLAIN_FP64, sa.DEPOSIT_DETERMINISTIC)
    b = make(sa.PLAIN_FP64, sa.DEPOSIT_DETERMINISTIC)
    b.set_tuning(sa.TUNE_THREADS, threads)
    b.set_tuning(sa.TUNE_XCD, xcd)
    a.step(5)
    b.step(5)
    for f in ("Ex", "Jy", "Bz_m"):
        np.testing.assert_array_equal(a.get_field(f), b.get_field(f))
    for s in range(2):
        pa, pb = a.get_particles(s), b.get_particles(s)
        np.testing.assert_array_equal(pa["x"], pb["x"])
        np.testing.assert_array_equal(pa["key"], pb["key"])


@pytest.mark.parametrize("precision", [sa.MIXED, sa.PLAIN_FP64])
def test_autotune_keeps_state(precision):
    a = make(precision)
    b = make(precision)
    a.step(2)
    b.step(2)
    knobs = b.autotune(reps=2)
    # layouts: cell-major, lane-interleaved, and (mixed form) pair-interleaved
    assert knobs["threads"] in (128, 256, 512)
    assert knobs["interleave"] in ((0, 1, 2) if precision == sa.MIXED else (0, 1))
    assert knobs["persistent"] in (0, 1)
    assert knobs["xcd"] in (0, 1) and not (knobs["xcd"] and knobs["persistent"])
    assert knobs["tile"] in (0, 1) and knobs["tile_x"] in (1, 2) and knobs["node_acc"] in (0, 1)
    if precision != sa.MIXED:
        assert knobs["tile_x"] == 1 and knobs["node_acc"] == 0
    assert knobs["particles_ms"] > 0
    a.step(5)
    b.step(5)
    for s in range(2):
        pa, pb = a.get_particles(s), b.get_particles(s)
        np.testing.assert_array_equal(pa["key"], pb["key"])
        np.testing.assert_array_equal(a.patch_counts(s), b.patch_counts(s))
        np.testing.assert_allclose(pb["x"], pa["x"], rtol=1e-10)
    for f in ("Ex", "Jx", "By"):
        ref = a.get_field(f)
        assert np.abs(b.get_field(f) - ref).max() <= 1e-10 * np.abs(ref).max(), f


def test_tuning_rejects_bad_values():
    sim = make(sa.MIXED, shape=(16, 16, 16))
    with pytest.raises(RuntimeError, match="SMI_TUNE_THREADS"):
        sim.set_tuning(sa.TUNE_THREADS, 96)
    with pytest.raises(RuntimeError, match="unknown tuning key"):
        sim.set_tuning(99, 1)
    with pytest.raises(RuntimeError, match="SMI_TUNE_TILE"):
        sim.set_tuning(sa.TUNE_TILE, 2)
    with pytest.raises(RuntimeError, match="SMI_TUNE_INTERLEAVE"):
        sim.set_tuning(sa.TUNE_INTERLEAVE, 3)


@pytest.mark.parametrize("precision", [sa.MIXED, sa.PLAIN_FP64])
@pytest.mark.parametrize("uneven", [False, True])
def test_persistent_work_list_matches_grid_launch(uneven, precision):
    """The persistent work-list launch (both forms) visits exactly the
    non-empty clusters: same keys and counts, fields equal up to the order of
    the J flush."""
    from tests.plasma import lwfa
    if uneven:
        shape, patch, dx = (256, 32, 32), (32, 16, 16), (0.1, 0.5, 0.5)
        a = lwfa(sa.Simulation, precision, shape, patch, dx)
        b = lwfa(sa.Simulation, precision, shape, patch, dx)
        ns = 1
    else:
        a, b, ns = make(precision), make(precision), 2
    b.set_tuning(sa.TUNE_PERSISTENT, 1)
    assert b.get_tuning(sa.TUNE_PERSISTENT) == 1
    a.step(6)
    b.step(6)
    for s in range(ns):
        pa, pb = a.get_particles(s), b.get_particles(s)
        np.testing.assert_array_equal(pa["key"], pb["key"])
        np.testing.assert_array_equal(a.patch_counts(s), b.patch_counts(s))
        np.testing.assert_allclose(pb["x"], pa["x"], rtol=1e-10)
    for f in ("Ey", "Jx", "Bz"):
        ref = a.get_field(f)
        assert np.abs(b.get_field(f) - ref).max() <= 1e-10 * np.abs(ref).max(), f


def test_unordered_cells_hold_the_oracle_multisets():
    """SMI_TUNE_CELL_ORDER = 0 (in-cell order left to the atomic scatter, as
    the reference's GPU sort): per cell the same particles as the oracle."""
    from oracle import pyoracle as O
    shape, dx = (32, 32, 32), (0.2, 0.2, 0.2)
    g = make(sa.PLAIN_FP64, shape=shape)
    g.set_tuning(sa.TUNE_CELL_ORDER, 0)
    assert g.get_tuning(sa.TUNE_CELL_ORDER) == 0
    o = O.OracleSim(3, 2, shape, (16, 16, 16), dx, cfl_dt(dx))
    for k, (name, m, q, sig) in enumerate(SPECIES):
        p = thermal(shape, dx, 8, sig, 9 + k)
        o.set_particles(o.add_species(m, q), p["x"], p["y"], p["z"], p["px"], p["py"], p["pz"], p["w"])
    g.step(6)
    o.step(6)
    for s in range(2):
        a, b = g.get_particles(s), o.get_particles(s)
        np.testing.assert_array_equal(g.patch_counts(s), o.patch_counts(s))
        ia = np.lexsort((a["z"], a["y"], a["x"], a["key"], a["patch"]))
        ib = np.lexsort((b["z"], b["y"], b["x"], b["key"], b["patch"]))
        np.testing.assert_array_equal(a["key"][ia], b["key"][ib])
        for k in ("x", "y", "z", "px", "py", "pz", "w"):
            scale = max(np.abs(b[k]).max(), 1e-300)
            assert np.abs(a[k][ia] - b[k][ib]).max() <= 1e-12 * scale, k


def test_xcd_order_mixed_matches_grid_order():
    a, b = make(sa.MIXED), make(sa.MIXED)
    b.set_tuning(sa.TUNE_XCD, 1)
    a.step(6)
    b.step(6)
    for s in range(2):
        np.testing.assert_array_equal(a.get_particles(s)["key"], b.get_particles(s)["key"])
    for f in ("Ex", "Jz"):
        ref = a.get_field(f)
        assert np.abs(b.get_field(f) - ref).max() <= 1e-10 * np.abs(ref).max(), f


def test_persistent_deterministic_fp64_is_bit_identical():
    """The deterministic FP64 form run by the persistent work-list kernel:
    particles, keys and every field bit-identical to the grid launch."""
    from tests.plasma import lwfa
    shape, patch, dx = (256, 32, 32), (32, 16, 16), (0.1, 0.5, 0.5)
    a = lwfa(sa.Simulation, sa.PLAIN_FP64, shape, patch, dx, deposit=sa.DEPOSIT_DETERMINISTIC)
    b = lwfa(sa.Simulation, sa.PLAIN_FP64, shape, patch, dx, deposit=sa.DEPOSIT_DETERMINISTIC)
    b.set_tuning(sa.TUNE_PERSISTENT, 1)
    a.step(5)
    b.step(5)
    pa, pb = a.get_particles(0), b.get_particles(0)
    for k in ("x", "y", "z", "px", "py", "pz", "w", "key", "patch"):
        np.testing.assert_array_equal(pa[k], pb[k], err_msg=k)
    for f in ("Ex", "Ey", "Ez", "Bx_m", "By_m", "Bz_m", "Jx", "Jy", "Jz"):
        np.testing.assert_array_equal(a.get_field(f), b.get_field(f), err_msg=f)


@pytest.mark.parametrize("threads,xcd,pad", [(512, 0, 1), (256, 1, 0), (128, 0, 1)])
def test_two_cluster_patch_tile_agrees(threads, xcd, pad):
    """The 8x4x4-cell patch tile (two clusters per workgroup) against the
    one-cluster tile: the first step reads the same fields, so particles and
    keys are identical after it.  J differs by the rounding of each
    contribution to its tile's fixed-point quantum (2^-46 N cr_max for a
    tile of N particles: about 3e-11 cr_max here, over a few hundred
    contributions per word) and by the FP64 flush order, far below the 1e-8
    bound a misplaced tile would break.  Over six steps keys and counts stay
    equal; the fields then differ at the FP32 level of the mixed form (the
    J differences move FP32 offsets by an ulp here and there), within 1e-5."""
    a, b = make(sa.MIXED), make(sa.MIXED)
    for sim in (a, b):
        sim.set_tuning(sa.TUNE_THREADS, threads)
        sim.set_tuning(sa.TUNE_XCD, xcd)
        sim.set_tuning(sa.TUNE_TILE, pad)
    b.set_tuning(sa.TUNE_TILE_X, 2)
    assert (a.get_tuning(sa.TUNE_TILE_X), b.get_tuning(sa.TUNE_TILE_X)) == (1, 2)
    a.step(1)
    b.step(1)
    for s in range(2):
        pa, pb = a.get_particles(s), b.get_particles(s)
        for k in ("key", "x", "y", "z", "px", "py", "pz", "w"):
            np.testing.assert_array_equal(pa[k], pb[k], err_msg=k)
    for f in ("Jx", "Jy", "Jz"):
        ref = a.get_field(f)
        assert np.abs(b.get_field(f) - ref).max() <= 1e-8 * np.abs(ref).max(), f
    a.step(5)
    b.step(5)
    for s in range(2):
        pa, pb = a.get_particles(s), b.get_particles(s)
        np.testing.assert_array_equal(pa["key"], pb["key"])
        np.testing.assert_array_equal(a.patch_counts(s), b.patch_counts(s))
    for f in ("Ex", "Jz", "By"):
        ref = a.get_field(f)
        assert np.abs(b.get_field(f) - ref).max() <= 1e-5 * np.abs(ref).max(), f


@pytest.mark.parametrize("threads,tile_x,order", [(256, 1, 2), (512, 2, 2), (256, 1, 4)])
def test_pair_interleaved_layout_pre_reduction_is_exact(threads, tile_x, order):
    """Pair-interleaved clusters (SMI_TUNE_INTERLEAVE 2): the mixed kernel
    adds the tile words of two lanes whose particles share a node before one
    atomic.  The fixed-point words are integers, so the J tiles hold exactly
    the same sums as with the lane-interleaved layout: after one step the
    particles are identical and J differs only by the order of the FP64
    flush (1e-13); over six steps keys and counts stay equal."""
    a, b = make(sa.MIXED, order=order), make(sa.MIXED, order=order)
    for sim in (a, b):
        sim.set_tuning(sa.TUNE_THREADS, threads)
        sim.set_tuning(sa.TUNE_TILE_X, tile_x)
    b.set_tuning(sa.TUNE_INTERLEAVE, 2)
    assert (a.get_tuning(sa.TUNE_INTERLEAVE), b.get_tuning(sa.TUNE_INTERLEAVE)) == (1, 2)
    a.step(1)
    b.step(1)
    for s in range(2):
        pa, pb = a.get_particles(s), b.get_particles(s)
        for k in ("key", "x", "y", "z", "px", "py", "pz", "w"):
            np.testing.assert_array_equal(pa[k], pb[k], err_msg=k)
    for f in ("Jx", "Jy", "Jz"):
        ref = a.get_field(f)
        assert np.abs(b.get_field(f) - ref).max() <= 1e-13 * np.abs(ref).max(), f
    a.step(5)
    b.step(5)
    for s in range(2):
        pa, pb = a.get_particles(s), b.get_particles(s)
        np.testing.assert_array_equal(pa["key"], pb["key"])
        np.testing.assert_array_equal(a.patch_counts(s), b.patch_counts(s))
    for f in ("Ex", "Jz", "By"):
        ref = a.get_field(f)
        assert np.abs(b.get_field(f) - ref).max() <= 1e-10 * np.abs(ref).max(), f


@pytest.mark.parametrize("layout,tile", [(1, 1), (0, 1), (1, 0)])
def test_node_accumulated_deposit_agrees(layout, tile):
    """Node-accumulated deposit (SMI_TUNE_NODE_ACC 1): the stayers of each
    256-particle round are staged per node and their tile words summed in
    registers, one atomic per node and word.  Same fixed-point words as the
    per-particle deposit: after one step the particles are identical and J
    differs only by the order of the FP64 flush (1e-13); over six steps keys
    and counts stay equal.  The cell-major layout (0) puts up to 256
    particles of few nodes in one round, past the four staging entries of a
    node: those are deposited at once (the overflow path)."""
    a, b = make(sa.MIXED), make(sa.MIXED)
    for sim in (a, b):
        sim.set_tuning(sa.TUNE_THREADS, 256)
        sim.set_tuning(sa.TUNE_INTERLEAVE, layout)
        sim.set_tuning(sa.TUNE_TILE, tile)
    b.set_tuning(sa.TUNE_NODE_ACC, 1)
    assert (a.get_tuning(sa.TUNE_NODE_ACC), b.get_tuning(sa.TUNE_NODE_ACC)) == (0, 1)
    a.step(1)
    b.step(1)
    for s in range(2):
        pa, pb = a.get_particles(s), b.get_particles(s)
        for k in ("key", "x", "y", "z", "px", "py", "pz", "w"):
            np.testing.assert_array_equal(pa[k], pb[k], err_msg=k)
    for f in ("Jx", "Jy", "Jz"):
        ref = a.get_field(f)
        assert np.abs(b.get_field(f) - ref).max() <= 1e-13 * np.abs(ref).max(), f
    a.step(5)
    b.step(5)
    for s in range(2):
        pa, pb = a.get_particles(s), b.get_particles(s)
        np.testing.assert_array_equal(pa["key"], pb["key"])
        np.testing.assert_array_equal(a.patch_counts(s), b.patch_counts(s))
    for f in ("Ex", "Jz", "By"):
        ref = a.get_field(f)
        assert np.abs(b.get_field(f) - ref).max() <= 1e-10 * np.abs(ref).max(), f
    np.testing.assert_allclose(b.gauss(), a.gauss(), rtol=0, atol=1e-12 * np.abs(a.rho()).max())


def test_node_accumulated_deposit_refused_where_unavailable():
    """The node-accumulated deposit is an order-2 mixed-form mode; other
    launch shapes of the mixed form keep the per-particle deposit."""
    sim = make(sa.PLAIN_FP64)
    with pytest.raises(RuntimeError, match="NODE_ACC"):
        sim.set_tuning(sa.TUNE_NODE_ACC, 1)
    assert sim.get_tuning(sa.TUNE_NODE_ACC) == 0
    a, b = make(sa.MIXED), make(sa.MIXED)
    b.set_tuning(sa.TUNE_NODE_ACC, 1)
    b.set_tuning(sa.TUNE_THREADS, 128)  # not a node-accumulation launch: per particle
    a.set_tuning(sa.TUNE_THREADS, 128)
    a.step(1)
    b.step(1)
    for s in range(2):
        np.testing.assert_array_equal(a.get_particles(s)["x"], b.get_particles(s)["x"])


def test_two_cluster_patch_tile_refused_where_unavailable():
    """The FP64 form has no two-cluster tile: the knob is refused."""
    sim = make(sa.PLAIN_FP64)
    with pytest.raises(RuntimeError, match="TILE_X"):
        sim.set_tuning(sa.TUNE_TILE_X, 2)
    assert sim.get_tuning(sa.TUNE_TILE_X) == 1


def test_tile_shapes_agree():
    """The two J-tile shapes of the mixed kernel: same keys and counts,
    fields equal up to the order of the J flush."""
    a, b = make(sa.MIXED), make(sa.MIXED)
    a.set_tuning(sa.TUNE_TILE, 1)
    b.set_tuning(sa.TUNE_TILE, 0)
    assert (a.get_tuning(sa.TUNE_TILE), b.get_tuning(sa.TUNE_TILE)) == (1, 0)
    a.step(6)
    b.step(6)
    for s in range(2):
        pa, pb = a.get_particles(s), b.get_particles(s)
        np.testing.assert_array_equal(pa["key"], pb["key"])
        np.testing.assert_array_equal(a.patch_counts(s), b.patch_counts(s))
        np.testing.assert_allclose(pb["x"], pa["x"], rtol=1e-10)
    for f in ("Ex", "Jz", "By"):
        ref = a.get_field(f)
        assert np.abs(b.get_field(f) - ref).max() <= 1e-10 * np.abs(ref).max(), f
