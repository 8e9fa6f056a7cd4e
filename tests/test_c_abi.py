"""The C ABI used from C: tests/c/abi_check.c is built with gcc -std=c99
-pedantic -Werror against include/smilei_amd.h and linked to the library.

The CPU test builds it and runs its `symbols` mode (every declared entry
point is taken by address, so the link fails on a missing one).  The GPU
tests hand it seeded inputs as raw files, let it call every entry point, and
compare what it writes back with the oracle: bit-exact where the GPU sums in
the reference's order, within 1e-13 of the largest value where FP64 atomics
only reorder the sums.
"""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from oracle import pyoracle as O
from tests.plasma import cfl_dt, thermal

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "c", "abi_check.c")
ROCM = "/opt/rocm"


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "smilei_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(smi_[a-z0-9_]+)\s*\(", txt)))


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    if not shutil.which("gcc"):
        pytest.skip("gcc not available")
    out = str(tmp_path_factory.mktemp("cabi") / "abi_check")
    lib = os.path.join(ROOT, "smilei_amd")
    cmd = ["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-pedantic", "-D__HIP_PLATFORM_AMD__",
           "-isystem", os.path.join(ROCM, "include"), "-I", os.path.join(ROOT, "include"), SRC, "-o", out,
           "-L", lib, "-lsmilei_amd", "-L", os.path.join(ROCM, "lib"), "-lamdhip64",
           "-Wl,-rpath," + lib, "-Wl,-rpath," + os.path.join(ROCM, "lib")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return out


def run(harness, mode, d=None, timeout=120):
    args = [harness, mode] + ([str(d)] if d is not None else [])
    r = subprocess.run(args, capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stderr + r.stdout
    return r.stdout


def test_c_symbols_table_covers_the_header(harness):
    table = set(re.findall(r"\(fn_t\)(smi_[a-z0-9_]+)", open(SRC).read()))
    assert table == set(declared_symbols())
    out = run(harness, "symbols")
    assert out.startswith("%d entry points" % len(table))


# ------------------------------------------------------------------ GPU modes

def put(d, name, a):
    np.ascontiguousarray(a).tofile(os.path.join(d, name))


def get(d, name, dtype=np.float64):
    return np.fromfile(os.path.join(d, "out_" + name), dtype=dtype)


@pytest.mark.gpu
def test_c_context_api_bit_exact(harness, tmp_path):
    """Whole-step API from C, deterministic FP64, 3 steps: fields and
    particles bit-identical to the oracle."""
    shape, patch, dx = (16, 16, 16), (8, 8, 8), (0.2, 0.2, 0.2)
    dt = cfl_dt(dx)
    p = thermal(shape, dx, 4, 0.1, 3)
    put(tmp_path, "geom.i32", np.array([3, 2, *shape, *patch, 1, 3, 0], np.int32))
    put(tmp_path, "cell.f64", np.array([*dx, dt]))
    put(tmp_path, "species.f64", np.array([1.0, -1.0]))
    for k in ("x", "y", "z", "px", "py", "pz", "w"):
        put(tmp_path, k + ".f64", p[k])
    run(harness, "context", tmp_path)
    o = O.OracleSim(3, 2, shape, patch, dx, dt)
    o.set_particles(o.add_species(1.0, -1.0), p["x"], p["y"], p["z"], p["px"], p["py"], p["pz"], p["w"])
    o.step(3)
    F = get(tmp_path, "fields.f64").reshape(12, -1)
    for i, name in enumerate(O.FIELDS):
        np.testing.assert_array_equal(F[i], o.get_field(name).ravel(), err_msg=name)
    ref = o.get_particles(0)
    P = get(tmp_path, "parts.f64").reshape(7, -1)
    for i, k in enumerate(("x", "y", "z", "px", "py", "pz", "w")):
        np.testing.assert_array_equal(P[i], ref[k], err_msg=k)
    np.testing.assert_array_equal(get(tmp_path, "key.i32", np.int32), ref["key"])
    np.testing.assert_array_equal(get(tmp_path, "patch.i32", np.int32), ref["patch"])
    np.testing.assert_allclose(get(tmp_path, "energies.f64"), o.energies(), rtol=1e-12)


STAG3 = {"Ex": (1, 0, 0), "Ey": (0, 1, 0), "Ez": (0, 0, 1), "Bxm": (0, 1, 1), "Bym": (1, 0, 1), "Bzm": (1, 1, 0)}


@pytest.mark.gpu
@pytest.mark.parametrize("order", [2, 4])
def test_c_operators_3d(harness, tmp_path, order):
    rng = np.random.default_rng(200 + order)
    npatch, n = 8, 3000
    np3 = [npatch + 2 * order + 1] * 3
    dom = [32 - order, 16 - order, 48 - order]
    dx = [0.3, 0.2, 0.25]
    dt = 0.9 * cfl_dt(dx)
    q, m = -1.0, 1.0
    f = {k: rng.standard_normal(tuple(np3[d] + s[d] for d in range(3))) * 0.1 for k, s in STAG3.items()}
    pos = [(dom[d] + order + 1 + rng.random(n) * (npatch - 2)) * dx[d] for d in range(3)]
    mom = [rng.standard_normal(n) * 0.3 for _ in range(3)]
    w = rng.random(n) + 0.5
    put(tmp_path, "geom.i32", np.array([order, *np3, *dom], np.int32))
    put(tmp_path, "cell.f64", np.array([*dx, dt, q, m]))
    for k in STAG3:
        put(tmp_path, k + ".f64", f[k])
    for k, a in zip(("x", "y", "z", "px", "py", "pz", "w"), pos + mom + [w]):
        put(tmp_path, k + ".f64", a)
    mshape = {"mE0": (1, 0, 0), "mE1": (0, 1, 0), "mE2": (0, 0, 1), "mB0": (0, 1, 1), "mB1": (1, 0, 1),
              "mB2": (1, 1, 0), "mJ0": (1, 0, 0), "mJ1": (0, 1, 0), "mJ2": (0, 0, 1)}
    M = {k: rng.standard_normal(tuple(np3[d] + s[d] for d in range(3))) for k, s in mshape.items()}
    for k, a in M.items():
        put(tmp_path, k + ".f64", a)
    run(harness, "ops3d", tmp_path)
    # oracle: interpolate, push, deposit (in particle order), rho, Maxwell
    fo = {"Ex": f["Ex"], "Ey": f["Ey"], "Ez": f["Ez"], "Bx_m": f["Bxm"], "By_m": f["Bym"], "Bz_m": f["Bzm"]}
    Ep, Bp, iold, dold = O.interp3d(order, fo, np3, dom, [1 / d for d in dx], *pos)
    np.testing.assert_array_equal(get(tmp_path, "Epart.f64"), Ep)
    np.testing.assert_array_equal(get(tmp_path, "Bpart.f64"), Bp)
    np.testing.assert_array_equal(get(tmp_path, "iold.i32", np.int32), iold)
    np.testing.assert_array_equal(get(tmp_path, "deltaold.f64"), dold)
    x, y, z, px, py, pz, invgf = O.push_boris(3, q, m, dt, *pos, *mom, Ep, Bp)
    for k, a in zip(("x", "y", "z", "px", "py", "pz", "invgf"), (x, y, z, px, py, pz, invgf)):
        np.testing.assert_array_equal(get(tmp_path, k + ".f64"), a, err_msg=k)
    J = [np.zeros((np3[0] + 1, np3[1], np3[2])), np.zeros((np3[0], np3[1] + 1, np3[2])),
         np.zeros((np3[0], np3[1], np3[2] + 1))]
    O.proj3d(order, J, np3, dom, dx, dt, 1 / np.prod(dx), q, x, y, z, w, iold, dold)
    for c, k in enumerate(("Jx", "Jy", "Jz")):
        np.testing.assert_array_equal(get(tmp_path, k + ".f64"), J[c].ravel(), err_msg=k)
    rho = np.zeros(np3)
    O.rho3d(order, rho, np3, dom, [1 / d for d in dx], 1 / np.prod(dx), q, x, y, z, w)
    np.testing.assert_allclose(get(tmp_path, "rho.f64"), rho.ravel(), rtol=0, atol=1e-13 * np.abs(rho).max())
    E = [M["mE0"].copy(), M["mE1"].copy(), M["mE2"].copy()]
    B = [M["mB0"].copy(), M["mB1"].copy(), M["mB2"].copy()]
    O.yee3d_ma(E, B, [M["mJ0"], M["mJ1"], M["mJ2"]], np3, dt, dx)
    O.yee3d_mf(B, E, np3, dt, dx)
    for c in range(3):
        np.testing.assert_array_equal(get(tmp_path, "mE%d.f64" % c), E[c].ravel())
        np.testing.assert_array_equal(get(tmp_path, "mB%d.f64" % c), B[c].ravel())


@pytest.mark.gpu
def test_c_operators_2d(harness, tmp_path):
    rng = np.random.default_rng(210)
    npatch, n = 16, 3000
    np2 = [npatch + 5, npatch + 5]
    dom = [30, 14]
    dx = [0.125, 0.2]
    dt = 0.9 * cfl_dt(dx)
    q, m = -1.0, 1.0
    stag = {"Ex": (1, 0), "Ey": (0, 1), "Ez": (0, 0), "Bxm": (0, 1), "Bym": (1, 0), "Bzm": (1, 1)}
    f = {k: rng.standard_normal(tuple(np2[d] + s[d] for d in range(2))) * 0.1 for k, s in stag.items()}
    pos = [(dom[d] + 3 + rng.random(n) * (npatch - 2)) * dx[d] for d in range(2)]
    mom = [rng.standard_normal(n) * 0.3 for _ in range(3)]
    w = rng.random(n) + 0.5
    put(tmp_path, "geom.i32", np.array([*np2, *dom], np.int32))
    put(tmp_path, "cell.f64", np.array([*dx, dt, q, m]))
    for k in stag:
        put(tmp_path, k + ".f64", f[k])
    for k, a in zip(("x", "y", "px", "py", "pz", "w"), pos + mom + [w]):
        put(tmp_path, k + ".f64", a)
    mshape = {"mE0": (1, 0), "mE1": (0, 1), "mE2": (0, 0), "mB0": (0, 1), "mB1": (1, 0), "mB2": (1, 1),
              "mJ0": (1, 0), "mJ1": (0, 1), "mJ2": (0, 0)}
    M = {k: rng.standard_normal(tuple(np2[d] + s[d] for d in range(2))) for k, s in mshape.items()}
    for k, a in M.items():
        put(tmp_path, k + ".f64", a)
    run(harness, "ops2d", tmp_path)
    fo = {"Ex": f["Ex"], "Ey": f["Ey"], "Ez": f["Ez"], "Bx_m": f["Bxm"], "By_m": f["Bym"], "Bz_m": f["Bzm"]}
    Ep, Bp, iold, dold = O.interp2d(2, fo, np2, dom, [1 / d for d in dx], *pos)
    np.testing.assert_array_equal(get(tmp_path, "Epart.f64"), Ep)
    np.testing.assert_array_equal(get(tmp_path, "Bpart.f64"), Bp)
    np.testing.assert_array_equal(get(tmp_path, "iold.i32", np.int32), iold)
    x, y, z, px, py, pz, invgf = O.push_boris(2, q, m, dt, pos[0], pos[1], np.zeros(n), *mom, Ep, Bp)
    for k, a in zip(("x", "y", "px", "py", "pz", "invgf"), (x, y, px, py, pz, invgf)):
        np.testing.assert_array_equal(get(tmp_path, k + ".f64"), a, err_msg=k)
    J = [np.zeros((np2[0] + 1, np2[1])), np.zeros((np2[0], np2[1] + 1)), np.zeros(tuple(np2))]
    O.proj2d(2, J, np2, dom, dx, dt, 1 / np.prod(dx), q, x, y, pz, w, invgf, iold, dold)
    for c, k in enumerate(("Jx", "Jy", "Jz")):
        np.testing.assert_array_equal(get(tmp_path, k + ".f64"), J[c].ravel(), err_msg=k)
    rho = np.zeros(np2)
    O.rho2d(2, rho, np2, dom, [1 / d for d in dx], 1 / np.prod(dx), q, x, y, w)
    np.testing.assert_allclose(get(tmp_path, "rho.f64"), rho.ravel(), rtol=0, atol=1e-13 * np.abs(rho).max())
    E = [M["mE0"].copy(), M["mE1"].copy(), M["mE2"].copy()]
    B = [M["mB0"].copy(), M["mB1"].copy(), M["mB2"].copy()]
    O.yee2d_ma(E, B, [M["mJ0"], M["mJ1"], M["mJ2"]], np2, dt, dx)
    O.yee2d_mf(B, E, np2, dt, dx)
    for c in range(3):
        np.testing.assert_array_equal(get(tmp_path, "mE%d.f64" % c), E[c].ravel())
        np.testing.assert_array_equal(get(tmp_path, "mB%d.f64" % c), B[c].ravel())


@pytest.mark.gpu
def test_c_reference_seams(harness, tmp_path):
    from tests.test_gpu_seams import binned_patch, oracle_currents
    rng = np.random.default_rng(220)
    P = binned_patch(rng, 8000)
    dx, np3, dom = P["dx"], P["np3"], P["dom"]
    dt = 0.9 * cfl_dt(dx)
    ref = oracle_currents(P, dt)
    rho = np.zeros(np3)
    for qv in np.unique(P["q"]):
        msk = P["q"] == qv
        O.rho3d(2, rho, np3, dom, [1 / d for d in dx], 1 / np.prod(dx), float(qv), *P["new"][:, msk], P["w"][msk])
    sizes = [r.size for r in ref] + [rho.size]
    put(tmp_path, "geom.i32", np.array([P["nb"], *np3, *dom, *sizes], np.int32))
    put(tmp_path, "cell.f64", np.array([1 / np.prod(dx), *[1 / d for d in dx], *[d / dt for d in dx]]))
    for d, k in enumerate("xyz"):
        put(tmp_path, k + ".f64", P["new"][d])
    put(tmp_path, "w.f64", P["w"])
    put(tmp_path, "q.i16", P["q"])
    put(tmp_path, "bins.i32", P["bin_index"])
    put(tmp_path, "iold.i32", P["iold"].ravel())
    put(tmp_path, "deltaold.f64", P["dold"].ravel())
    run(harness, "seams", tmp_path)
    for c, k in enumerate(("Jx", "Jy", "Jz")):
        np.testing.assert_allclose(get(tmp_path, k + ".f64"), ref[c].ravel(), rtol=0,
                                   atol=1e-13 * np.abs(ref[c]).max(), err_msg=k)
    np.testing.assert_allclose(get(tmp_path, "rho.f64"), rho.ravel(), rtol=0, atol=1e-13 * np.abs(rho).max())


@pytest.mark.gpu
def test_c_patch_synchronisation(harness, tmp_path):
    rng = np.random.default_rng(230)
    nd, order, n_cells, patch = 3, 2, (16, 24, 16), (8, 8, 8)
    dx = (0.3, 0.2, 0.25)
    npt = int(np.prod([n_cells[d] // patch[d] for d in range(3)]))
    nJ = npt * sum(int(np.prod(s)) for s in O.patch_block_dims(3, 2, patch, [(1, 0, 0), (0, 1, 0), (0, 0, 1)]))
    nB = npt * sum(int(np.prod(s)) for s in O.patch_block_dims(3, 2, patch, [(0, 1, 1), (1, 0, 1), (1, 1, 0)]))
    J, B = rng.standard_normal(nJ), rng.standard_normal(nB)
    n = 20000
    xmax = [n_cells[d] * dx[d] for d in range(3)]
    pos = [rng.random(n) * xmax[d] + (rng.random(n) < 0.1) * (rng.random(n) - 0.5) * 2 * patch[d] * dx[d]
           for d in range(3)]
    mom = [rng.standard_normal(n) for _ in range(3)]
    w = rng.random(n) + 0.5
    q = rng.choice(np.array([-1, 1], np.int16), n)
    put(tmp_path, "geom.i32", np.array([nd, order, *n_cells, *patch], np.int32))
    put(tmp_path, "cell.f64", np.array(dx))
    put(tmp_path, "J.f64", J)
    put(tmp_path, "B.f64", B)
    for k, a in zip(("x", "y", "z", "px", "py", "pz", "w"), pos + mom + [w]):
        put(tmp_path, k + ".f64", a)
    put(tmp_path, "q.i16", q)
    run(harness, "sync", tmp_path)
    Jr, Br = J.copy(), B.copy()
    O.sum_j_ghosts(nd, order, n_cells, patch, Jr)
    O.exchange_b(nd, order, n_cells, patch, Br)
    np.testing.assert_array_equal(get(tmp_path, "J.f64"), Jr)
    np.testing.assert_array_equal(get(tmp_path, "B.f64"), Br)
    wrapped = [np.where(p < 0, p + xm, np.where(p >= xm, p - xm, p)) for p, xm in zip(pos, xmax)]
    S = O.OracleSim(3, 2, n_cells, patch, dx, 0.1)
    sid = S.add_species(1.0, -1.0)
    S.set_particles(sid, *wrapped, *mom, np.arange(n, dtype=np.float64))
    ref = S.get_particles(sid)
    order_ = ref["w"].astype(np.int64)
    for k in ("x", "y", "z", "px", "py", "pz"):
        np.testing.assert_array_equal(get(tmp_path, k + ".f64"), ref[k], err_msg=k)
    np.testing.assert_array_equal(get(tmp_path, "w.f64"), w[order_])
    np.testing.assert_array_equal(get(tmp_path, "q.i16", np.int16), q[order_])
    np.testing.assert_array_equal(get(tmp_path, "key.i32", np.int32), ref["key"])
    np.testing.assert_array_equal(get(tmp_path, "counts.i32", np.int32), S.patch_counts(sid))
