"""Per-operator entry points (reference per-patch layout) vs the oracle.

Inputs are random per-patch arrays and particles; the HIP kernels must
reproduce the CPU operators bit for bit (deterministic deposit) or within
1e-13 of max|J| (atomic deposit, only the summation order differs).
"""
import ctypes

import numpy as np
import pytest

from oracle import pyoracle as O
from smilei_amd import _lib
from tests.plasma import cfl_dt

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def dev(a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a)).to("cuda")
    return t


def ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def I3(v):
    return (ctypes.c_int * 3)(*[int(a) for a in v])


def D3(v):
    return (ctypes.c_double * 3)(*[float(a) for a in v])


STAG = {"Ex": (1, 0, 0), "Ey": (0, 1, 0), "Ez": (0, 0, 1), "Bx_m": (0, 1, 1), "By_m": (1, 0, 1), "Bz_m": (1, 1, 0)}


def setup_patch(order, rng, n, move=False):
    npatch = 12 if order == 4 else 8
    np3 = [npatch + 2 * order + 1] * 3
    dom = [32 - order, 16 - order, 48 - order]
    dx = [0.3, 0.2, 0.25]
    f = {k: rng.standard_normal(tuple(np3[d] + dd[d] for d in range(3))) for k, dd in STAG.items()}
    pos = [(dom[d] + order + rng.random(n) * npatch) * dx[d] for d in range(3)]
    return np3, dom, dx, f, pos


@pytest.mark.parametrize("order", [2, 4])
def test_interpolate_3d_bit_exact(order):
    L = _lib.load()
    rng = np.random.default_rng(100 + order)
    n = 5000
    np3, dom, dx, f, pos = setup_patch(order, rng, n)
    Ep, Bp, iold, dold = O.interp3d(order, f, np3, dom, [1 / d for d in dx], *pos)
    df = [dev(f[k]) for k in ("Ex", "Ey", "Ez", "Bx_m", "By_m", "Bz_m")]
    dp = [dev(p) for p in pos]
    gE = torch.empty(3 * n, dtype=torch.float64, device="cuda")
    gB = torch.empty_like(gE)
    gi = torch.empty(3 * n, dtype=torch.int32, device="cuda")
    gd = torch.empty_like(gE)
    _lib.check(L.smi_interpolate_3d(order, *[ptr(t) for t in df], I3(np3), I3(dom), D3(dx), n,
                                    *[ptr(t) for t in dp], ptr(gE), ptr(gB), ptr(gi), ptr(gd), None))
    np.testing.assert_array_equal(host(gE), Ep)
    np.testing.assert_array_equal(host(gB), Bp)
    np.testing.assert_array_equal(host(gi), iold)
    np.testing.assert_array_equal(host(gd), dold)


def test_interpolate_2d_bit_exact():
    L = _lib.load()
    rng = np.random.default_rng(7)
    n = 4000
    np2 = [8 + 5, 8 + 5]
    dom = [6, 14]
    dx = [0.125, 0.2]
    st2 = {k: v[:2] for k, v in STAG.items()}
    f = {k: rng.standard_normal(tuple(np2[d] + dd[d] for d in range(2))) for k, dd in st2.items()}
    pos = [(dom[d] + 2 + rng.random(n) * 8) * dx[d] for d in range(2)]
    Ep, Bp, iold, dold = O.interp2d(2, f, np2, dom, [1 / d for d in dx], *pos)
    df = [dev(f[k]) for k in ("Ex", "Ey", "Ez", "Bx_m", "By_m", "Bz_m")]
    dp = [dev(p) for p in pos]
    gE = torch.empty(3 * n, dtype=torch.float64, device="cuda")
    gB = torch.empty_like(gE)
    gi = torch.empty(2 * n, dtype=torch.int32, device="cuda")
    gd = torch.empty(2 * n, dtype=torch.float64, device="cuda")
    _lib.check(L.smi_interpolate_2d(2, *[ptr(t) for t in df], (ctypes.c_int * 2)(*np2), (ctypes.c_int * 2)(*dom),
                                    (ctypes.c_double * 2)(*dx), n, *[ptr(t) for t in dp], ptr(gE), ptr(gB), ptr(gi),
                                    ptr(gd), None))
    np.testing.assert_array_equal(host(gE), Ep)
    np.testing.assert_array_equal(host(gB), Bp)
    np.testing.assert_array_equal(host(gi), iold)
    np.testing.assert_array_equal(host(gd), dold)


@pytest.mark.parametrize("ndim", [2, 3])
def test_push_boris_bit_exact(ndim):
    L = _lib.load()
    rng = np.random.default_rng(11)
    n = 20000
    x = [rng.random(n) * 10 for _ in range(3)]
    u = [rng.standard_normal(n) * s for s in (0.1, 1.0, 30.0)]
    Ep = rng.standard_normal(3 * n) * 0.3
    Bp = rng.standard_normal(3 * n) * 2.0
    q, m, dt = -1.0, 1.0, 0.0913
    ref = O.push_boris(ndim, q, m, dt, *x, *u, Ep, Bp)
    dx_ = [dev(a) for a in x]
    du = [dev(a) for a in u]
    gE, gB = dev(Ep), dev(Bp)  # kept alive until the kernel has run
    inv = torch.empty(n, dtype=torch.float64, device="cuda")
    _lib.check(L.smi_push_boris(ndim, q, m, dt, n, *[ptr(t) for t in dx_ + du], ptr(gE), ptr(gB), ptr(inv), None))
    for a, b in zip([host(t) for t in dx_ + du] + [host(inv)], ref):
        np.testing.assert_array_equal(a, b)


def _deposit_inputs(order, rng, n):
    npatch = 12 if order == 4 else 8
    np3 = [npatch + 2 * order + 1] * 3
    dom = [-order, 8 - order, 4 - order]
    dx = [0.3, 0.2, 0.25]
    dt = 0.9 * cfl_dt(dx)
    old = [(dom[d] + order + 1 + rng.random(n) * (npatch - 2)) * dx[d] for d in range(3)]
    new = [old[d] + (rng.random(n) - 0.5) * 1.6 * dx[d] for d in range(3)]
    iold = np.empty(3 * n, np.int32)
    dold = np.empty(3 * n)
    for d in range(3):
        xn = old[d] / dx[d]
        r = np.floor(xn + 0.5)
        iold[d * n:(d + 1) * n] = r.astype(np.int32) - dom[d]
        dold[d * n:(d + 1) * n] = xn - r
    w = rng.random(n) + 0.5
    return np3, dom, dx, dt, new, w, iold, dold


@pytest.mark.parametrize("order", [2, 4])
@pytest.mark.parametrize("deterministic", [1, 0])
def test_currents_3d(order, deterministic):
    L = _lib.load()
    rng = np.random.default_rng(300 + order)
    n = 3000 if deterministic else 20000
    np3, dom, dx, dt, new, w, iold, dold = _deposit_inputs(order, rng, n)
    shapes = [(np3[0] + 1, np3[1], np3[2]), (np3[0], np3[1] + 1, np3[2]), (np3[0], np3[1], np3[2] + 1)]
    J = [np.zeros(s) for s in shapes]
    q = -1.0
    O.proj3d(order, J, np3, dom, dx, dt, 1 / (dx[0] * dx[1] * dx[2]), q, *new, w, iold, dold)
    gJ = [torch.zeros(s, dtype=torch.float64, device="cuda") for s in shapes]
    ins = [dev(a) for a in new + [w, iold, dold]]  # kept alive until the kernel has run
    _lib.check(L.smi_currents_3d(order, *[ptr(t) for t in gJ], I3(np3), I3(dom), D3(dx), dt, q, n,
                                 *[ptr(t) for t in ins], deterministic, None))
    for a, b in zip(gJ, J):
        if deterministic:
            np.testing.assert_array_equal(host(a), b)
        else:
            assert np.abs(host(a) - b).max() <= 1e-13 * np.abs(b).max()


@pytest.mark.parametrize("deterministic", [1, 0])
def test_currents_2d(deterministic):
    L = _lib.load()
    rng = np.random.default_rng(17)
    n = 3000 if deterministic else 20000
    np2 = [8 + 5, 8 + 5]
    dom = [-2, 6]
    dx = [0.125, 0.2]
    dt = 0.9 * cfl_dt(dx)
    old = [(dom[d] + 3 + rng.random(n) * 6) * dx[d] for d in range(2)]
    new = [old[d] + (rng.random(n) - 0.5) * 1.6 * dx[d] for d in range(2)]
    iold = np.empty(2 * n, np.int32)
    dold = np.empty(2 * n)
    for d in range(2):
        xn = old[d] / dx[d]
        r = np.floor(xn + 0.5)
        iold[d * n:(d + 1) * n] = r.astype(np.int32) - dom[d]
        dold[d * n:(d + 1) * n] = xn - r
    pz = rng.standard_normal(n)
    invgf = 1 / np.sqrt(1 + pz ** 2)
    w = rng.random(n) + 0.5
    shapes = [(np2[0] + 1, np2[1]), (np2[0], np2[1] + 1), (np2[0], np2[1])]
    J = [np.zeros(s) for s in shapes]
    O.proj2d(2, J, np2, dom, dx, dt, 1 / (dx[0] * dx[1]), -1.0, *new, pz, w, invgf, iold, dold)
    gJ = [torch.zeros(s, dtype=torch.float64, device="cuda") for s in shapes]
    ins = [dev(a) for a in new + [pz, w, invgf, iold, dold]]  # kept alive until the kernel has run
    _lib.check(L.smi_currents_2d(2, *[ptr(t) for t in gJ], (ctypes.c_int * 2)(*np2), (ctypes.c_int * 2)(*dom),
                                 (ctypes.c_double * 2)(*dx), dt, -1.0, n, *[ptr(t) for t in ins],
                                 deterministic, None))
    for a, b in zip(gJ, J):
        if deterministic:
            np.testing.assert_array_equal(host(a), b)
        else:
            assert np.abs(host(a) - b).max() <= 1e-13 * np.abs(b).max()


def test_maxwell_3d_patch_bit_exact():
    L = _lib.load()
    rng = np.random.default_rng(23)
    np3 = [21, 13, 17]
    dx = [0.2, 0.25, 0.3]
    dt = cfl_dt(dx)
    stag = {"Ex": (1, 0, 0), "Ey": (0, 1, 0), "Ez": (0, 0, 1), "Bx": (0, 1, 1), "By": (1, 0, 1), "Bz": (1, 1, 0)}
    F = {k: rng.standard_normal(tuple(np3[d] + dd[d] for d in range(3))) for k, dd in stag.items()}
    Jc = [rng.standard_normal(F[k].shape) for k in ("Ex", "Ey", "Ez")]
    E = [F["Ex"].copy(), F["Ey"].copy(), F["Ez"].copy()]
    B = [F["Bx"].copy(), F["By"].copy(), F["Bz"].copy()]
    gE = [dev(a) for a in E]
    gB = [dev(a) for a in B]
    gJ = [dev(a) for a in Jc]
    torch.cuda.synchronize()
    O.yee3d_ma(E, B, Jc, np3, dt, dx)
    O.yee3d_mf(B, E, np3, dt, dx)
    _lib.check(L.smi_maxwell_ampere_3d(*[ptr(t) for t in gE + gB + gJ], I3(np3), dt, D3(dx), None))
    _lib.check(L.smi_maxwell_faraday_3d(*[ptr(t) for t in gB + gE], I3(np3), dt, D3(dx), None))
    for a, b in zip(gE + gB, E + B):
        np.testing.assert_array_equal(host(a), b)
