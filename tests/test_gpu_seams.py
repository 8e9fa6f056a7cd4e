"""Entry points of the reference's remaining seams vs the oracle.

* smi_current_deposition_3d / smi_density_deposition_3d keep the argument
  lists of currentDeposition3DOnDevice / densityDeposition3DOnDevice
  (Projector3D2OrderGPU.cpp:55-120): per-particle int16 charge, cluster
  bins, both loop orders (cell_sorting) and both J layouts (not_spectral).
  Each particle's contributions are the CPU operator's to the bit; the
  tile and global FP64 atomics only reorder the sums, so J and rho agree
  with the oracle to 1e-13 of their largest value.
* smi_density_3d / smi_density_2d (Projector::basic type 0), same bound.
* smi_maxwell_*_2d, smi_sum_j_ghosts, smi_exchange_b: bit-exact.
* smi_sort_exchange: order, keys, counts and bins bit-exact against the
  oracle's exchange and stable cell sort.
"""
import ctypes

import numpy as np
import pytest

from oracle import pyoracle as O
from smilei_amd import _lib
from tests.plasma import cfl_dt

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def I3(v):
    return (ctypes.c_int * 3)(*[int(a) for a in v])


def D3(v):
    return (ctypes.c_double * 3)(*[float(a) for a in v])


def I2(v):
    return (ctypes.c_int * 2)(*[int(a) for a in v])


def D2(v):
    return (ctypes.c_double * 2)(*[float(a) for a in v])


def binned_patch(rng, n, patch=8, move=0.9):
    """One order-2 patch of patch^3 cells starting at global cell (16, 8,
    24): particles sorted by 4^3 cluster (x-major), the interpolator's
    iold / deltaold at the old positions, new positions moved by less than
    `move` cells per axis, charges -1 / +1 / +2, weights in [0.5, 1.5)."""
    os_ = 2
    start = np.array([16, 8, 24])
    dom = start - os_
    dx = np.array([0.3, 0.2, 0.25])
    np3 = [patch + 2 * os_ + 1] * 3
    cell = np.stack([rng.integers(0, patch, n) for _ in range(3)])
    old = (start[:, None] + cell + rng.random((3, n))) * dx[:, None]
    nb = patch // 4
    cl = ((cell[0] // 4) * nb + cell[1] // 4) * nb + cell[2] // 4
    o = np.argsort(cl, kind="stable")
    old, cl = old[:, o], cl[o]
    new = old + (rng.random((3, n)) - 0.5) * 2 * move * dx[:, None]
    xn = old / dx[:, None]
    r = np.floor(xn + 0.5)
    iold = (r - dom[:, None]).astype(np.int32)
    dold = xn - r
    bin_index = np.cumsum(np.bincount(cl, minlength=nb ** 3)).astype(np.int32)
    q = rng.choice(np.array([-1, 1, 2], np.int16), n)
    w = rng.random(n) + 0.5
    return dict(np3=np3, dom=dom, dx=dx, old=old, new=new, iold=iold, dold=dold, bin_index=bin_index, nb=nb, q=q,
                w=w)


def oracle_currents(P, dt):
    np3, dom, dx = P["np3"], P["dom"], P["dx"]
    J = [np.zeros((np3[0] + 1, np3[1], np3[2])), np.zeros((np3[0], np3[1] + 1, np3[2])),
         np.zeros((np3[0], np3[1], np3[2] + 1))]
    icv = 1.0 / np.prod(dx)
    for qv in np.unique(P["q"]):
        m = P["q"] == qv
        O.proj3d(2, J, np3, dom, dx, dt, icv, float(qv), *P["new"][:, m], P["w"][m], P["iold"].reshape(3, -1)[:, m],
                 P["dold"].reshape(3, -1)[:, m])
    return J


def seam_args(P, dt):
    dx = P["dx"]
    n = len(P["w"])
    return [P["nb"]] * 3, n, 1.0 / np.prod(dx), [1 / d for d in dx], [d / dt for d in dx]


@pytest.mark.parametrize("cell_sorting", [False, True])
@pytest.mark.parametrize("not_spectral", [1, 0])
def test_current_deposition_seam(cell_sorting, not_spectral):
    L = _lib.load()
    rng = np.random.default_rng(40 + cell_sorting + 2 * not_spectral)
    P = binned_patch(rng, 20000)
    dt = 0.9 * cfl_dt(P["dx"])
    ref = oracle_currents(P, dt)
    np3 = P["np3"]
    shapes = [ref[0].shape, ref[1].shape, ref[2].shape] if not_spectral else [tuple(np3)] * 3
    J = [torch.zeros(int(np.prod(s)), dtype=torch.float64, device="cuda") for s in shapes]
    nb, n, icv, dinv, dxdt = seam_args(P, dt)
    pos = [dev(P["new"][d]) for d in range(3)]
    q, w, bi = dev(P["q"]), dev(P["w"]), dev(P["bin_index"])
    io, do = dev(P["iold"].ravel()), dev(P["dold"].ravel())
    _lib.check(L.smi_current_deposition_3d(*[ptr(t) for t in J], *[t.numel() for t in J], *[ptr(t) for t in pos],
                                           ptr(q), ptr(w), ptr(bi), *nb, None, ptr(io), ptr(do), n, icv, *dinv,
                                           *dxdt, *P["dom"], np3[1], np3[2], not_spectral, cell_sorting))
    for c in range(3):
        got = host(J[c]).reshape(shapes[c])
        want = ref[c]
        if not not_spectral:  # the tile never reaches the extra dual plane
            want = want[:np3[0], :np3[1], :np3[2]]
        scale = np.abs(want).max()
        assert scale > 0
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-13 * scale, err_msg="J%s" % "xyz"[c])


def test_density_deposition_seam():
    L = _lib.load()
    rng = np.random.default_rng(50)
    P = binned_patch(rng, 20000)
    dt = 0.9 * cfl_dt(P["dx"])
    np3, dom, dx = P["np3"], P["dom"], P["dx"]
    ref = np.zeros(np3)
    icv = 1.0 / np.prod(dx)
    for qv in np.unique(P["q"]):
        m = P["q"] == qv
        O.rho3d(2, ref, np3, dom, [1 / d for d in dx], icv, float(qv), *P["new"][:, m], P["w"][m])
    rho = torch.zeros(int(np.prod(np3)), dtype=torch.float64, device="cuda")
    nb, n, icv, dinv, dxdt = seam_args(P, dt)
    pos = [dev(P["new"][d]) for d in range(3)]
    q, w, bi = dev(P["q"]), dev(P["w"]), dev(P["bin_index"])
    io, do = dev(P["iold"].ravel()), dev(P["dold"].ravel())
    _lib.check(L.smi_density_deposition_3d(ptr(rho), rho.numel(), *[ptr(t) for t in pos], ptr(q), ptr(w), ptr(bi),
                                           *nb, None, ptr(io), ptr(do), n, icv, *dinv, *dxdt, *dom, np3[1], np3[2],
                                           1, False))
    got = host(rho).reshape(np3)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-13 * np.abs(ref).max())


def test_current_deposition_seam_rejects_misbinned_particle():
    """A particle outside the bin that bin_index assigns it to would write
    outside the bin's tile: it is skipped and the call reports it."""
    L = _lib.load()
    rng = np.random.default_rng(51)
    P = binned_patch(rng, 512)
    dt = 0.9 * cfl_dt(P["dx"])
    P["bin_index"][:] = P["bin_index"][-1]  # every particle in bin 0
    np3 = P["np3"]
    J = [torch.zeros((np3[0] + (c == 0)) * (np3[1] + (c == 1)) * (np3[2] + (c == 2)), dtype=torch.float64,
                     device="cuda") for c in range(3)]
    nb, n, icv, dinv, dxdt = seam_args(P, dt)
    pos = [dev(P["new"][d]) for d in range(3)]
    q, w, bi = dev(P["q"]), dev(P["w"]), dev(P["bin_index"])
    io, do = dev(P["iold"].ravel()), dev(P["dold"].ravel())
    rc = L.smi_current_deposition_3d(*[ptr(t) for t in J], *[t.numel() for t in J], *[ptr(t) for t in pos], ptr(q),
                                     ptr(w), ptr(bi), *nb, None, ptr(io), ptr(do), n, icv, *dinv, *dxdt, *P["dom"],
                                     np3[1], np3[2], 1, False)
    assert rc == -1 and b"tile" in L.smi_last_error()


@pytest.mark.parametrize("order", [2, 4])
def test_density_3d_basic(order):
    L = _lib.load()
    rng = np.random.default_rng(60 + order)
    npatch, n = 8, 30000
    np3 = [npatch + 2 * order + 1] * 3
    dom = [32 - order, 16 - order, 48 - order]
    dx = [0.3, 0.2, 0.25]
    pos = [(dom[d] + order + rng.random(n) * npatch) * dx[d] for d in range(3)]
    w = rng.random(n) + 0.5
    icv = 1 / np.prod(dx)
    ref = np.zeros(np3)
    O.rho3d(order, ref, np3, dom, [1 / d for d in dx], icv, -1.0, *pos, w)
    rho = torch.zeros(int(np.prod(np3)), dtype=torch.float64, device="cuda")
    xs = [dev(p) for p in pos] + [dev(w)]  # held: a temporary's memory is reused by the next one
    _lib.check(L.smi_density_3d(order, ptr(rho), I3(np3), I3(dom), D3(dx), -1.0, n, *[ptr(t) for t in xs], None))
    np.testing.assert_allclose(host(rho).reshape(np3), ref, rtol=0, atol=1e-13 * np.abs(ref).max())


def test_density_2d_basic():
    L = _lib.load()
    rng = np.random.default_rng(64)
    npatch, n = 16, 20000
    np2 = [npatch + 5, npatch + 5]
    dom = [30, 14]
    dx = [0.125, 0.2]
    pos = [(dom[d] + 2 + rng.random(n) * npatch) * dx[d] for d in range(2)]
    w = rng.random(n) + 0.5
    ref = np.zeros(np2)
    O.rho2d(2, ref, np2, dom, [1 / d for d in dx], 1 / np.prod(dx), 2.0, *pos, w)
    rho = torch.zeros(int(np.prod(np2)), dtype=torch.float64, device="cuda")
    xs = [dev(p) for p in pos] + [dev(w)]
    _lib.check(L.smi_density_2d(2, ptr(rho), I2(np2), I2(dom), D2(dx), 2.0, n, ptr(xs[0]), ptr(xs[1]), ptr(xs[2]),
                                None))
    np.testing.assert_allclose(host(rho).reshape(np2), ref, rtol=0, atol=1e-13 * np.abs(ref).max())


def test_maxwell_2d_bit_exact():
    L = _lib.load()
    rng = np.random.default_rng(70)
    np2 = [21, 13]
    dx = [0.125, 0.2]
    dt = 0.95 * cfl_dt(dx)
    shp = {"Ex": (np2[0] + 1, np2[1]), "Ey": (np2[0], np2[1] + 1), "Ez": tuple(np2),
           "Bx": (np2[0], np2[1] + 1), "By": (np2[0] + 1, np2[1]), "Bz": (np2[0] + 1, np2[1] + 1)}
    f = {k: rng.standard_normal(s) for k, s in shp.items()}
    Jf = [rng.standard_normal(shp[k]) for k in ("Ex", "Ey", "Ez")]
    E = [f[k].copy() for k in ("Ex", "Ey", "Ez")]
    B = [f[k].copy() for k in ("Bx", "By", "Bz")]
    O.yee2d_ma(E, B, Jf, np2, dt, dx)
    O.yee2d_mf(B, E, np2, dt, dx)
    dE = [dev(f[k]) for k in ("Ex", "Ey", "Ez")]
    dB = [dev(f[k]) for k in ("Bx", "By", "Bz")]
    dJ = [dev(a) for a in Jf]
    _lib.check(L.smi_maxwell_ampere_2d(*[ptr(t) for t in dE + dB + dJ], I2(np2), dt, D2(dx), None))
    _lib.check(L.smi_maxwell_faraday_2d(*[ptr(t) for t in dB + dE], I2(np2), dt, D2(dx), None))
    for c in range(3):
        np.testing.assert_array_equal(host(dE[c]), E[c])
        np.testing.assert_array_equal(host(dB[c]), B[c])


GEOMS = [(3, 2, (16, 24, 16), (8, 8, 8)), (3, 4, (24, 12, 24), (12, 12, 12)), (2, 2, (24, 16), (8, 8))]


def _flat_blocks(rng, nd, order, n_cells, patch, dual):
    npt = int(np.prod([n_cells[d] // patch[d] for d in range(nd)]))
    dims = O.patch_block_dims(nd, order, patch, dual)
    per = sum(int(np.prod(s)) for s in dims)
    return rng.standard_normal(npt * per)


@pytest.mark.parametrize("nd,order,n_cells,patch", GEOMS)
def test_sum_j_ghosts_bit_exact(nd, order, n_cells, patch):
    L = _lib.load()
    rng = np.random.default_rng(80 + order + nd)
    J = _flat_blocks(rng, nd, order, n_cells, patch, [(1, 0, 0), (0, 1, 0), (0, 0, 1)])
    ref = J.copy()
    O.sum_j_ghosts(nd, order, n_cells, patch, ref)
    d = dev(J)
    nc, ps = list(n_cells) + [1] * (3 - nd), list(patch) + [1] * (3 - nd)
    _lib.check(L.smi_sum_j_ghosts(nd, order, I3(nc), I3(ps), ptr(d), None))
    np.testing.assert_array_equal(host(d), ref)


@pytest.mark.parametrize("nd,order,n_cells,patch", GEOMS)
def test_exchange_b_bit_exact(nd, order, n_cells, patch):
    L = _lib.load()
    rng = np.random.default_rng(90 + order + nd)
    B = _flat_blocks(rng, nd, order, n_cells, patch, [(0, 1, 1), (1, 0, 1), (1, 1, 0)])
    ref = B.copy()
    O.exchange_b(nd, order, n_cells, patch, ref)
    d = dev(B)
    nc, ps = list(n_cells) + [1] * (3 - nd), list(patch) + [1] * (3 - nd)
    _lib.check(L.smi_exchange_b(nd, order, I3(nc), I3(ps), ptr(d), None))
    np.testing.assert_array_equal(host(d), ref)


@pytest.mark.parametrize("nd,n_cells,patch,n", [(3, (32, 16, 16), (8, 8, 8), 60000), (2, (64, 32), (16, 8), 40000),
                                               (3, (16, 16, 16), (16, 16, 16), 0)])
def test_sort_exchange_matches_oracle(nd, n_cells, patch, n):
    """Particles in any order, a tenth of them pushed up to one patch length
    out of the periodic box, are wrapped, assigned to their patches and
    sorted by cell key, stably: the oracle's exchange and sort."""
    L = _lib.load()
    rng = np.random.default_rng(100 + nd)
    dx = [0.3, 0.2, 0.25][:nd]
    xmax = [n_cells[d] * dx[d] for d in range(nd)]
    pos = [rng.random(n) * xmax[d] for d in range(nd)]
    out = rng.random(n) < 0.1
    for d in range(nd):
        shift = rng.choice([-1.0, 1.0], n) * rng.random(n) * patch[d] * dx[d]
        pos[d] = np.where(out, pos[d] + shift, pos[d])
    mom = [rng.standard_normal(n) for _ in range(3)]
    w = rng.random(n) + 0.5
    q = rng.choice(np.array([-1, 1], np.int16), n)
    wrapped = [np.where(p < 0, p + xm, np.where(p >= xm, p - xm, p)) for p, xm in zip(pos, xmax)]
    S = O.OracleSim(nd, 2, n_cells, patch, dx, 0.1)
    sid = S.add_species(1.0, -1.0)
    # the charge rides along as a tag in the weight's place for the oracle
    zz = wrapped[2] if nd == 3 else np.zeros(n)
    S.set_particles(sid, wrapped[0], wrapped[1], zz, *mom, np.arange(n, dtype=np.float64))
    ref = S.get_particles(sid)
    order = ref["w"].astype(np.int64)
    dpos = [dev(p) for p in pos] + ([] if nd == 3 else [None])
    dmom = [dev(m) for m in mom]
    dw, dq = dev(w), dev(q)
    key = torch.empty(max(n, 1), dtype=torch.int32, device="cuda")
    npt = S.npatch
    cpc = 64 if nd == 3 else 16
    clpp = int(np.prod([patch[d] for d in range(nd)])) // cpc
    counts = torch.empty(npt, dtype=torch.int32, device="cuda")
    bins = torch.empty(npt * clpp, dtype=torch.int32, device="cuda")
    nc, ps = list(n_cells) + [1] * (3 - nd), list(patch) + [1] * (3 - nd)
    _lib.check(L.smi_sort_exchange(nd, I3(nc), I3(ps), D3(list(dx) + [1.0] * (3 - nd)), n,
                                   *[ptr(t) if t is not None else None for t in dpos], *[ptr(t) for t in dmom],
                                   ptr(dw), ptr(dq), ptr(key), ptr(counts), ptr(bins), None))
    for d, k in enumerate("xyz"[:nd]):
        np.testing.assert_array_equal(host(dpos[d]), ref[k])
    for d, k in enumerate(("px", "py", "pz")):
        np.testing.assert_array_equal(host(dmom[d]), ref[k])
    np.testing.assert_array_equal(host(dw), w[order])
    np.testing.assert_array_equal(host(dq), q[order])
    np.testing.assert_array_equal(host(key)[:n], ref["key"])
    np.testing.assert_array_equal(host(counts), S.patch_counts(sid))
    # bins: per patch, cumulative ends of its clusters (last_index)
    want = []
    for p in range(npt):
        kp = ref["key"][ref["patch"] == p]
        want.append(np.cumsum(np.bincount(kp // cpc, minlength=clpp)))
    np.testing.assert_array_equal(host(bins), np.concatenate(want).astype(np.int32))
