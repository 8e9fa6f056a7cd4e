"""Host-side checks of the C-ABI library that need no GPU."""
import ctypes
import os
import re

import pytest

from smilei_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "smilei_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(smi_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    syms = declared_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(L, s), s
    # and the Python binding describes each of them
    assert set(syms) == set(_lib.SIGNATURES)


def _params(**kw):
    p = _lib.SmiParams()
    p.geometry = kw.get("geometry", b"3Dcartesian")
    p.interpolation_order = kw.get("order", 2)
    p.interpolator = kw.get("interpolator", b"momentum-conserving")
    p.pusher = kw.get("pusher", b"boris")
    p.maxwell_solver = kw.get("maxwell_solver", b"Yee")
    for d in range(3):
        p.n_cells[d] = kw.get("n_cells", (32, 32, 32))[d]
        p.patch_size[d] = kw.get("patch_size", (16, 16, 16))[d]
        p.cell_length[d] = 0.2
    p.timestep = kw.get("timestep", 0.1)
    p.precision = kw.get("precision", 0)
    p.deposit = 0
    return p


@pytest.mark.parametrize("kw,msg", [
    (dict(geometry=b"1Dcartesian"), "geometry"),
    (dict(geometry=b"AMcylindrical"), "geometry"),
    (dict(order=3), "interpolation_order"),
    (dict(geometry=b"2Dcartesian", order=4, n_cells=(32, 32, 1), patch_size=(16, 16, 1)), "interpolation_order"),
    (dict(interpolator=b"wt"), "interpolator"),
    (dict(pusher=b"vay"), "pusher"),
    (dict(maxwell_solver=b"Lehe"), "maxwell_solver"),
    (dict(patch_size=(16, 10, 16)), "multiple of 4"),
    (dict(n_cells=(40, 32, 32)), "divide"),
    (dict(patch_size=(4, 16, 16), n_cells=(32, 32, 32)), "oversize"),
    (dict(timestep=0.0), "timestep"),
    (dict(timestep=0.2), "one cell per step"),
])
def test_factories_reject_unsupported_namelist_values(kw, msg):
    """Like the reference factories, unsupported choices are errors; the
    check happens before any device call, so it runs without a GPU."""
    L = _lib.load()
    h = ctypes.c_void_p()
    rc = L.smi_create(ctypes.byref(_params(**kw)), ctypes.byref(h))
    assert rc == -1
    assert msg in L.smi_last_error().decode()
    assert not h.value


def test_per_operator_entry_points_validate_arguments():
    L = _lib.load()
    z = (ctypes.c_int * 3)(0, 0, 0)
    d = (ctypes.c_double * 3)(1.0, 1.0, 1.0)
    rc = L.smi_interpolate_3d(3, *[None] * 6, z, z, d, 0, None, None, None, None, None, None, None, None)
    assert rc == -1 and "order" in L.smi_last_error().decode()
    rc = L.smi_push_boris(3, -1.0, 0.0, 0.1, 0, *[None] * 6, None, None, None, None)
    assert rc == -1 and "mass" in L.smi_last_error().decode()


def test_seam_and_sync_entry_points_validate_arguments():
    """Geometry and bin checks fail before any device work."""
    L = _lib.load()
    I3 = ctypes.c_int * 3
    D3 = ctypes.c_double * 3
    # patch size not a multiple of 4
    rc = L.smi_sum_j_ghosts(3, 2, I3(32, 32, 32), I3(16, 10, 16), ctypes.c_void_p(1), None)
    assert rc == -1 and "multiple of 4" in L.smi_last_error().decode()
    rc = L.smi_exchange_b(2, 4, I3(32, 32, 1), I3(16, 16, 1), ctypes.c_void_p(1), None)
    assert rc == -1 and "2D order" in L.smi_last_error().decode()
    rc = L.smi_sort_exchange(3, I3(32, 32, 32), I3(16, 16, 16), D3(0.1, 0.1, 0.1), 10, *[None] * 7, None, None,
                             None, None, None)
    assert rc == -1 and "null particle" in L.smi_last_error().decode()
    args = [None] * 3 + [1] * 3 + [None] * 6 + [0, 1, 1] + [None] * 3 + [0] + [1.0] * 7 + [0] * 3 + [13, 13, 1, False]
    rc = L.smi_current_deposition_3d(*args)
    assert rc == -1 and "bin counts" in L.smi_last_error().decode()
    args = [None, 1] + [None] * 6 + [2, 2, 2] + [None] * 3 + [0] + [1.0] * 7 + [0] * 3 + [13, 13, 2, False]
    rc = L.smi_density_deposition_3d(*args)
    assert rc == -1 and "not_spectral" in L.smi_last_error().decode()
