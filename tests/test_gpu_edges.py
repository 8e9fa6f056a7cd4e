"""Edge cases of the engine: empty species, the capacity limit, very dense
cells (the workgroup sort of the in-cell fix-up, the mixed kernel's
node-changer queue overflowing into the inline deposit), particles exactly
on cell faces and at the end of the box."""
import numpy as np
import pytest

import smilei_amd as sa
from oracle import pyoracle as O
from tests.plasma import cfl_dt, load, thermal

pytestmark = pytest.mark.gpu

FIELDS = ["Ex", "Ey", "Ez", "Bx", "By", "Bz", "Bx_m", "By_m", "Bz_m", "Jx", "Jy", "Jz"]
SHAPE, PATCH, DX = (16, 16, 16), (8, 8, 8), (0.2, 0.2, 0.2)


def new(precision=sa.PLAIN_FP64, deposit=sa.DEPOSIT_ATOMIC, use_graph=True):
    return sa.Simulation("3Dcartesian", 2, SHAPE, PATCH, DX, cfl_dt(DX), precision=precision, deposit=deposit,
                         use_graph=use_graph)


@pytest.mark.parametrize("precision", [sa.PLAIN_FP64, sa.MIXED])
def test_empty_species_step_and_report(precision):
    """A species without particles (capacity 0, and one emptied by loading
    zero particles) steps with the others, graph-captured, and reports
    nothing."""
    sim = new(precision)
    a = sim.add_species("electron", 1.0, -1.0, 16 ** 3 * 2)
    load(sim, a, thermal(SHAPE, DX, 2, 0.1, 1))
    b = sim.add_species("none", 1.0, -1.0, 0)
    c = sim.add_species("emptied", 1836.0, 1.0, 100)
    z = np.zeros(0)
    sim.set_particles(c, z, z, z, z, z, z, z)
    sim.step(4)
    sim.synchronize()
    for s in (b, c):
        assert sim.num_particles(s) == 0
        assert len(sim.get_particles(s)["x"]) == 0
        assert sim.patch_counts(s).sum() == 0
    e = sim.energies()
    assert np.all(np.isfinite(e)) and e[2] == 0.0 and e[3] == 0.0
    assert sim.num_particles(a) == 16 ** 3 * 2


def test_capacity_is_enforced():
    sim = new()
    s = sim.add_species("e", 1.0, -1.0, 10)
    p = thermal(SHAPE, DX, 1, 0.1, 2)
    with pytest.raises(RuntimeError, match="capacity"):
        load(sim, s, p)
    with pytest.raises(RuntimeError):
        sim.add_species("huge", 1.0, -1.0, 1 << 31)


@pytest.mark.parametrize("precision", [sa.PLAIN_FP64, sa.MIXED])
@pytest.mark.parametrize("key,value", [("x", np.nan), ("x", -0.1), ("px", np.nan), ("pz", np.inf), ("w", np.nan)])
def test_corrupt_particles_are_refused(precision, key, value):
    """Non-finite or out-of-box input is refused by smi_set_particles
    (SMI_ERR_INVALID) before anything reaches the device; the species keeps
    its previous particles."""
    sim = new(precision)
    p = thermal(SHAPE, DX, 1, 0.1, 2)
    s = sim.add_species("e", 1.0, -1.0, len(p["x"]))
    load(sim, s, p)
    bad = {k: v.copy() for k, v in p.items()}
    bad[key][7] = value
    with pytest.raises(RuntimeError, match="outside the simulation box|non-finite"):
        load(sim, s, bad)
    assert sim.num_particles(s) == len(p["x"])
    sim.step(1)


def dense_plasma(seed=5, n_dense=3000, sigma=0.3):
    """A thin background plus n_dense particles in one cell (8, 8, 8), fast
    enough that most of them change node every step."""
    rng = np.random.default_rng(seed)
    p = thermal(SHAPE, DX, 1, 0.05, seed)
    c = np.array([8, 8, 8])
    pos = (c + rng.random((n_dense, 3))) * np.array(DX)
    u = rng.normal(0.0, sigma, (n_dense, 3))
    w = np.full(n_dense, p["w"][0])
    for k, key in enumerate(("x", "y", "z")):
        p[key] = np.concatenate([p[key], pos[:, k]])
    for k, key in enumerate(("px", "py", "pz")):
        p[key] = np.concatenate([p[key], u[:, k]])
    p["w"] = np.concatenate([p["w"], w])
    return p


@pytest.mark.parametrize("n_dense", [3000, 5000])
def test_dense_cell_deterministic_bit_exact(n_dense):
    """3000 (5000) particles in one cell: a bucket far beyond the register
    ranking, ordered by one workgroup with a bitonic sort in LDS (by rank
    counting over LDS-staged chunks beyond 4096), and a long per-patch
    deposit list; the deterministic engine stays bit-identical to the
    oracle."""
    p = dense_plasma(n_dense=n_dense)
    g = new(deposit=sa.DEPOSIT_DETERMINISTIC)
    o = O.OracleSim(3, 2, SHAPE, PATCH, DX, cfl_dt(DX))
    load(g, g.add_species("e", 1.0, -1.0, len(p["x"])), p)
    o.set_particles(o.add_species(1.0, -1.0), p["x"], p["y"], p["z"], p["px"], p["py"], p["pz"], p["w"])
    g.step(3)
    o.step(3)
    a, b = g.get_particles(0), o.get_particles(0)
    for k in ("x", "y", "z", "px", "py", "pz", "w", "key", "patch"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)
    for f in FIELDS:
        np.testing.assert_array_equal(g.get_field(f), o.get_field(f), err_msg=f)


def test_dense_cell_mixed_queue_overflow_conserves_charge():
    """The mixed kernel defers node-changing particles to an LDS queue of a
    few hundred entries; 3000 fast particles in one cluster overflow it and
    the rest deposit inline.  Charge conservation (Gauss law) and the
    energies against the FP64 form hold all the same."""
    p = dense_plasma(sigma=0.5)
    m, f = new(sa.MIXED), new()
    for sim in (m, f):
        load(sim, sim.add_species("e", 1.0, -1.0, len(p["x"])), p)
    g0 = m.gauss()
    rho0 = np.abs(m.rho()).max()
    m.step(10)
    f.step(10)
    drift = np.abs(m.gauss() - g0).max() / rho0
    assert drift < 1e-6, drift
    em, ef = m.energies(), f.energies()
    np.testing.assert_allclose(em.sum(), ef.sum(), rtol=1e-5)


@pytest.mark.parametrize("precision", [sa.PLAIN_FP64, sa.MIXED])
def test_positions_on_cell_faces_and_box_end(precision):
    """Particles exactly on cell faces, on patch faces, at 0 and just below
    the box length: binned in the cell they lie in, as the reference's
    floor(x / dx) (FP64: keys equal the oracle's)."""
    L = np.array(SHAPE) * np.array(DX)
    vals = np.array([0.0, 1 * DX[0], 7 * DX[0], 8 * DX[0], 15 * DX[0], np.nextafter(L[0], 0.0), 0.5 * DX[0]])
    g = np.stack(np.meshgrid(vals, vals, vals, indexing="ij"), -1).reshape(-1, 3)
    n = len(g)
    z = np.zeros(n)
    w = np.full(n, 0.001)
    sim = new(precision)
    s = sim.add_species("e", 1.0, -1.0, n)
    sim.set_particles(s, g[:, 0], g[:, 1], g[:, 2], z, z, z, w)
    got = sim.get_particles(s)
    pos = np.stack([got["x"], got["y"], got["z"]], 1)
    # the reference's rule: floor(x * (1 / dx)) (nvidiaParticles.cu:639-660), clamped
    cell = np.minimum(np.floor(pos * (1.0 / np.array(DX))).astype(int), np.array(SHAPE) - 1)
    pc = cell // np.array(PATCH)
    npc = np.array(SHAPE) // np.array(PATCH)
    np.testing.assert_array_equal(got["patch"], (pc[:, 0] * npc[1] + pc[:, 1]) * npc[2] + pc[:, 2])
    np.testing.assert_array_equal(sim.patch_counts(s), np.bincount(got["patch"], minlength=sim.npatch))
    if precision == sa.PLAIN_FP64:
        o = O.OracleSim(3, 2, SHAPE, PATCH, DX, cfl_dt(DX))
        o.set_particles(o.add_species(1.0, -1.0), g[:, 0], g[:, 1], g[:, 2], z, z, z, w)
        b = o.get_particles(0)
        for k in ("x", "y", "z", "key", "patch"):
            np.testing.assert_array_equal(got[k], b[k], err_msg=k)
    else:
        # the split representation keeps the positions to FP32 offsets
        np.testing.assert_allclose(np.sort(pos.sum(1)), np.sort(g.sum(1)), rtol=0, atol=1e-6)
    # at rest, nothing moves; one step later the bookkeeping is unchanged
    sim.step(1)
    after = sim.get_particles(s)
    np.testing.assert_array_equal(after["patch"], got["patch"])


def cell_multisets_agree(a, b, s, tol):
    """Same particles per (patch, cell) in both simulations, positions and
    momenta to tol relative to their largest magnitude."""
    pa, pb = a.get_particles(s), b.get_particles(s)
    np.testing.assert_array_equal(a.patch_counts(s), b.patch_counts(s))
    ia = np.lexsort((pa["z"], pa["y"], pa["x"], pa["key"], pa["patch"]))
    ib = np.lexsort((pb["z"], pb["y"], pb["x"], pb["key"], pb["patch"]))
    for k in ("patch", "key"):
        np.testing.assert_array_equal(pa[k][ia], pb[k][ib], err_msg=k)
    for k in ("x", "y", "z", "px", "py", "pz", "w"):
        scale = max(np.abs(pb[k]).max(), 1e-300)
        assert np.abs(pa[k][ia] - pb[k][ib]).max() <= tol * scale, k


@pytest.mark.parametrize("precision,interleave", [(sa.PLAIN_FP64, 1), (sa.PLAIN_FP64, 0), (sa.MIXED, 1),
                                                  (sa.MIXED, 0), (sa.MIXED, 2)])
def test_placed_stayers_keep_the_cell_contents(precision, interleave):
    """In-cell order left to the sort (SMI_TUNE_CELL_ORDER 0): the particle
    kernel places the first 31 of its cluster's stayers of each cell (their
    rank in the key's top bits), the others and the particles arriving from
    other clusters take the atomic count-down.  With 3000 slow particles in
    one cell (thousands of stayers, most of them past the placed ranks), a
    switch to the canonical order and back, in the three cluster layouts, every
    cell holds the same particles as in the canonical run."""
    p = dense_plasma(n_dense=3000, sigma=0.05)
    a, b = new(precision), new(precision)
    for sim in (a, b):
        sim.set_tuning(sa.TUNE_INTERLEAVE, interleave)
        load(sim, sim.add_species("e", 1.0, -1.0, len(p["x"])), p)
    b.set_tuning(sa.TUNE_CELL_ORDER, 0)
    a.step(3)
    b.step(3)
    b.set_tuning(sa.TUNE_CELL_ORDER, 1)
    a.step(1)
    b.step(1)
    b.set_tuning(sa.TUNE_CELL_ORDER, 0)
    a.step(3)
    b.step(3)
    tol = 1e-12 if precision == sa.PLAIN_FP64 else 1e-6
    cell_multisets_agree(b, a, 0, tol)
    for f in ("Ex", "Jy", "Bz"):
        ref = a.get_field(f)
        assert np.abs(b.get_field(f) - ref).max() <= 1e-10 * np.abs(ref).max(), f


def test_placed_stayers_2d():
    """The same in 2D (16-cell clusters), FP64 form."""
    shape, patch, dx = (64, 32), (16, 16), (0.2, 0.2)
    sims = []
    for order in (1, 0):
        sim = sa.Simulation("2Dcartesian", 2, shape, patch, dx, cfl_dt(dx))
        sim.set_tuning(sa.TUNE_CELL_ORDER, order)
        p = thermal(shape, dx, 40, 0.05, 11)
        load(sim, sim.add_species("e", 1.0, -1.0, len(p["x"])), p)
        sim.step(5)
        sims.append(sim)
    cell_multisets_agree(sims[1], sims[0], 0, 1e-12)


@pytest.mark.parametrize("shape,patch", [((16, 16, 16), (8, 8, 8)), ((16, 16, 8), (8, 8, 8))])
def test_canonical_order_independent_of_layout(shape, patch):
    """The canonical in-cell order (direction of motion, then the previous
    cell-major position) is the same with the clusters stored
    lane-interleaved or pair-interleaved -- where the in-cell pass computes
    the tie-break and
    the direction of the particles that stayed in their cluster instead of
    reading them -- as cell-major; also in a box of one 8-cell patch along
    z (two clusters), where particles wrap around the periodic box from one
    cluster into the other through their own patch's faces."""
    dx = (0.2, 0.2, 0.2)
    # FP64 form: cell-major against lane-interleaved; mixed form (which also
    # has the pair-interleaved layout): lane- against pair-interleaved
    for prec, lays, tol in ((sa.PLAIN_FP64, (0, 1), 1e-12), (sa.MIXED, (1, 2), 1e-6)):
        sims = []
        for lay in lays:
            sim = sa.Simulation("3Dcartesian", 2, shape, patch, dx, cfl_dt(dx), precision=prec)
            sim.set_tuning(sa.TUNE_INTERLEAVE, lay)
            p = thermal(shape, dx, 12, 0.3, 21)
            load(sim, sim.add_species("e", 1.0, -1.0, len(p["x"])), p)
            sim.step(6)
            sims.append(sim)
        a, b = sims[1].get_particles(0), sims[0].get_particles(0)
        for k in ("patch", "key"):
            np.testing.assert_array_equal(a[k], b[k], err_msg=k)
        for k in ("x", "y", "z", "px", "py", "pz"):
            scale = np.abs(b[k]).max()
            assert np.abs(a[k] - b[k]).max() <= tol * scale, k


@pytest.mark.parametrize("order,persistent", [(4, 0), (2, 1)])
def test_canonical_order_pair_layout_without_pre_reduction(order, persistent):
    """The pair-interleaved layout also reaches the mixed kernels that do not
    pre-reduce pairs (order 4; the persistent work-list launch), which must
    still rank the particles leaving their cluster by the pair layout: the
    canonical in-cell order is the same as with the lane-interleaved
    layout."""
    shape, patch, dx = (32, 32, 32), (16, 16, 16), (0.2, 0.2, 0.2)
    sims = []
    for lay in (1, 2):
        sim = sa.Simulation("3Dcartesian", order, shape, patch, dx, cfl_dt(dx), precision=sa.MIXED)
        sim.set_tuning(sa.TUNE_INTERLEAVE, lay)
        sim.set_tuning(sa.TUNE_PERSISTENT, persistent)
        p = thermal(shape, dx, 4, 0.3, 23)
        load(sim, sim.add_species("e", 1.0, -1.0, len(p["x"])), p)
        sim.step(6)
        sims.append(sim)
    a, b = sims[1].get_particles(0), sims[0].get_particles(0)
    for k in ("patch", "key"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)
    for k in ("x", "y", "z", "px", "py", "pz"):
        assert np.abs(a[k] - b[k]).max() <= 1e-6 * np.abs(b[k]).max(), k
