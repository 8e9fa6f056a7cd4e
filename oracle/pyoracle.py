"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke``
and by the ``cpu_baseline`` leg of bench.py, never by the product package.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

FIELDS = ["Ex", "Ey", "Ez", "Bx", "By", "Bz", "Bx_m", "By_m", "Bz_m", "Jx", "Jy", "Jz"]
FID = {n: i for i, n in enumerate(FIELDS)}

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int32)


class OraParams(ctypes.Structure):
    _fields_ = [
        ("ndim", ctypes.c_int),
        ("order", ctypes.c_int),
        ("n_cells", ctypes.c_int * 3),
        ("patch_size", ctypes.c_int * 3),
        ("cell_length", ctypes.c_double * 3),
        ("timestep", ctypes.c_double),
    ]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.ora_create.restype = ctypes.c_void_p
        L.ora_create.argtypes = [ctypes.POINTER(OraParams)]
        L.ora_destroy.argtypes = [ctypes.c_void_p]
        L.ora_set_threads.argtypes = [ctypes.c_int]
        L.ora_add_species.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]
        L.ora_set_time_frozen.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double]
        L.ora_set_particles.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int64] + [_D] * 7
        L.ora_num_particles.restype = ctypes.c_int64
        L.ora_num_particles.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.ora_get_particles.argtypes = [ctypes.c_void_p, ctypes.c_int] + [_D] * 7 + [_I, _I]
        L.ora_patch_counts.argtypes = [ctypes.c_void_p, ctypes.c_int, _I]
        L.ora_set_field.argtypes = [ctypes.c_void_p, ctypes.c_int, _D]
        L.ora_get_field.argtypes = [ctypes.c_void_p, ctypes.c_int, _D]
        L.ora_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.ora_energies.argtypes = [ctypes.c_void_p, _D]
        L.ora_field_energies.argtypes = [ctypes.c_void_p, _D]
        L.ora_rho.argtypes = [ctypes.c_void_p, _D]
        L.ora_gauss.argtypes = [ctypes.c_void_p, _D]
        L.ora_timers.argtypes = [ctypes.c_void_p, _D]
        _I3, _D3, _I2, _D2 = ctypes.c_int * 3, ctypes.c_double * 3, ctypes.c_int * 2, ctypes.c_double * 2
        L.ora_interp3d.argtypes = [ctypes.c_int] + [_D] * 6 + [_I3, _I3, _D3, ctypes.c_int64] + [_D] * 3 + \
            [_D, _D, _I, _D]
        L.ora_interp2d.argtypes = [ctypes.c_int] + [_D] * 6 + [_I2, _I2, _D2, ctypes.c_int64] + [_D] * 2 + \
            [_D, _D, _I, _D]
        L.ora_push_boris.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_int64] + [_D] * 6 + [_D, _D, _D]
        L.ora_proj3d.argtypes = [ctypes.c_int] + [_D] * 3 + [_I3, _I3, _D3, ctypes.c_double, ctypes.c_double,
                                                          ctypes.c_double, ctypes.c_int64] + [_D] * 4 + [_I, _D]
        L.ora_proj2d.argtypes = [ctypes.c_int] + [_D] * 3 + [_I2, _I2, _D2, ctypes.c_double, ctypes.c_double,
                                                          ctypes.c_double, ctypes.c_int64] + [_D] * 5 + [_I, _D]
        L.ora_rho3d.argtypes = [ctypes.c_int, _D, _I3, _I3, _D3, ctypes.c_double, ctypes.c_double,
                                ctypes.c_int64] + [_D] * 4
        L.ora_rho2d.argtypes = [ctypes.c_int, _D, _I2, _I2, _D2, ctypes.c_double, ctypes.c_double,
                                ctypes.c_int64] + [_D] * 3
        L.ora_yee3d_ma.argtypes = [_D] * 9 + [_I3, ctypes.c_double, _D3]
        L.ora_yee3d_mf.argtypes = [_D] * 6 + [_I3, ctypes.c_double, _D3]
        L.ora_yee2d_ma.argtypes = [_D] * 9 + [_I2, ctypes.c_double, _D2]
        L.ora_yee2d_mf.argtypes = [_D] * 6 + [_I2, ctypes.c_double, _D2]
        L.ora_sum_j_ghosts.argtypes = [ctypes.POINTER(OraParams), _D]
        L.ora_exchange_b.argtypes = [ctypes.POINTER(OraParams), _D]
        _lib = L
    return _lib


def dptr(a):
    return a.ctypes.data_as(_D)


def iptr(a):
    return a.ctypes.data_as(_I)


class OracleSim:
    """Whole-simulation oracle: per-patch arrays, reference step order."""

    def __init__(self, ndim, order, n_cells, patch_size, cell_length, timestep):
        L = lib()
        p = OraParams()
        p.ndim, p.order = ndim, order
        nc = list(n_cells) + [1] * (3 - len(n_cells))
        ps = list(patch_size) + [1] * (3 - len(patch_size))
        cl = list(cell_length) + [1.0] * (3 - len(cell_length))
        for d in range(3):
            p.n_cells[d], p.patch_size[d], p.cell_length[d] = nc[d], ps[d], cl[d]
        p.timestep = timestep
        self.h = L.ora_create(ctypes.byref(p))
        if not self.h:
            raise ValueError("oracle rejected the parameters")
        self.ndim = ndim
        self.shape = tuple(nc)
        self.nspecies = 0
        self.npatch = int(np.prod([nc[d] // ps[d] for d in range(3)]))

    def __del__(self):
        if getattr(self, "h", None):
            lib().ora_destroy(self.h)
            self.h = None

    def add_species(self, mass, charge):
        self.nspecies += 1
        return lib().ora_add_species(self.h, float(mass), float(charge))

    def set_time_frozen(self, s, time_frozen):
        if lib().ora_set_time_frozen(self.h, s, float(time_frozen)) != 0:
            raise ValueError("bad species id")

    def set_particles(self, s, x, y, z, px, py, pz, w):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z, px, py, pz, w)]
        rc = lib().ora_set_particles(self.h, s, len(arrs[-1]), *[dptr(a) for a in arrs])
        if rc != 0:
            raise ValueError("oracle set_particles failed: %d" % rc)

    def num_particles(self, s):
        return lib().ora_num_particles(self.h, s)

    def get_particles(self, s):
        n = self.num_particles(s)
        out = {k: np.empty(n) for k in ("x", "y", "z", "px", "py", "pz", "w")}
        out["patch"] = np.empty(n, np.int32)
        out["key"] = np.empty(n, np.int32)
        lib().ora_get_particles(self.h, s, *[dptr(out[k]) for k in ("x", "y", "z", "px", "py", "pz", "w")],
                                iptr(out["patch"]), iptr(out["key"]))
        return out

    def patch_counts(self, s):
        c = np.empty(self.npatch, np.int32)
        lib().ora_patch_counts(self.h, s, iptr(c))
        return c

    def set_field(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(self.shape)
        lib().ora_set_field(self.h, FID[name], dptr(a))

    def get_field(self, name):
        a = np.empty(self.shape)
        lib().ora_get_field(self.h, FID[name], dptr(a))
        return a

    def step(self, n=1):
        lib().ora_step(self.h, n)

    def energies(self):
        out = np.empty(1 + self.nspecies)
        lib().ora_energies(self.h, dptr(out))
        return out

    def field_energies(self):
        out = np.empty(6)
        lib().ora_field_energies(self.h, dptr(out))
        return out

    def rho(self):
        a = np.empty(self.shape)
        lib().ora_rho(self.h, dptr(a))
        return a

    def gauss(self):
        a = np.empty(self.shape)
        lib().ora_gauss(self.h, dptr(a))
        return a

    def timers(self):
        out = np.empty(4)
        lib().ora_timers(self.h, dptr(out))
        return out


def set_threads(n):
    lib().ora_set_threads(int(n))


# ---- single-patch operators (reference per-patch layout) -------------------
def _i(v):
    return (ctypes.c_int * len(v))(*[int(a) for a in v])


def _d(v):
    return (ctypes.c_double * len(v))(*[float(a) for a in v])


def interp3d(order, fields, np3, dom, dinv, x, y, z):
    """fields: dict Ex,Ey,Ez,Bx_m,By_m,Bz_m of per-patch arrays."""
    n = len(x)
    Ep, Bp = np.empty(3 * n), np.empty(3 * n)
    iold, dold = np.empty(3 * n, np.int32), np.empty(3 * n)
    f = [np.ascontiguousarray(fields[k], dtype=np.float64) for k in ("Ex", "Ey", "Ez", "Bx_m", "By_m", "Bz_m")]
    xs = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z)]
    lib().ora_interp3d(order, *[dptr(a) for a in f], _i(np3), _i(dom), _d(dinv), n, *[dptr(a) for a in xs],
                       dptr(Ep), dptr(Bp), iptr(iold), dptr(dold))
    return Ep, Bp, iold, dold


def interp2d(order, fields, np2, dom, dinv, x, y):
    n = len(x)
    Ep, Bp = np.empty(3 * n), np.empty(3 * n)
    iold, dold = np.empty(2 * n, np.int32), np.empty(2 * n)
    f = [np.ascontiguousarray(fields[k], dtype=np.float64) for k in ("Ex", "Ey", "Ez", "Bx_m", "By_m", "Bz_m")]
    xs = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y)]
    lib().ora_interp2d(order, *[dptr(a) for a in f], _i(np2), _i(dom), _d(dinv), n, *[dptr(a) for a in xs],
                       dptr(Ep), dptr(Bp), iptr(iold), dptr(dold))
    return Ep, Bp, iold, dold


def push_boris(ndim, charge, mass, dt, x, y, z, px, py, pz, Ep, Bp):
    """Returns updated copies (x, y, z, px, py, pz, invgf)."""
    arrs = [np.array(a, dtype=np.float64) for a in (x, y, z, px, py, pz)]
    n = len(arrs[0])
    invgf = np.empty(n)
    lib().ora_push_boris(ndim, float(charge), float(mass), float(dt), n, *[dptr(a) for a in arrs],
                         dptr(np.ascontiguousarray(Ep, dtype=np.float64)),
                         dptr(np.ascontiguousarray(Bp, dtype=np.float64)), dptr(invgf))
    return (*arrs, invgf)


def proj3d(order, J, np3, dom, dx, dt, inv_cell_volume, charge, x, y, z, w, iold, dold):
    """J: [Jx, Jy, Jz] per-patch arrays, updated in place."""
    n = len(x)
    xs = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z, w)]
    lib().ora_proj3d(order, *[dptr(a) for a in J], _i(np3), _i(dom), _d(dx), float(dt), float(inv_cell_volume),
                     float(charge), n, *[dptr(a) for a in xs], iptr(np.ascontiguousarray(iold, dtype=np.int32)),
                     dptr(np.ascontiguousarray(dold, dtype=np.float64)))


def proj2d(order, J, np2, dom, dx, dt, inv_cell_volume, charge, x, y, pz, w, invgf, iold, dold):
    n = len(x)
    xs = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, pz, w, invgf)]
    lib().ora_proj2d(order, *[dptr(a) for a in J], _i(np2), _i(dom), _d(dx), float(dt), float(inv_cell_volume),
                     float(charge), n, *[dptr(a) for a in xs], iptr(np.ascontiguousarray(iold, dtype=np.int32)),
                     dptr(np.ascontiguousarray(dold, dtype=np.float64)))


def rho3d(order, rho, np3, dom, dinv, inv_cell_volume, charge, x, y, z, w):
    n = len(x)
    xs = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z, w)]
    lib().ora_rho3d(order, dptr(rho), _i(np3), _i(dom), _d(dinv), float(inv_cell_volume), float(charge), n,
                    *[dptr(a) for a in xs])


def rho2d(order, rho, np2, dom, dinv, inv_cell_volume, charge, x, y, w):
    n = len(x)
    xs = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, w)]
    lib().ora_rho2d(order, dptr(rho), _i(np2), _i(dom), _d(dinv), float(inv_cell_volume), float(charge), n,
                    *[dptr(a) for a in xs])


def yee3d_ma(E, B, J, np3, dt, dx):
    lib().ora_yee3d_ma(*[dptr(a) for a in E], *[dptr(a) for a in B], *[dptr(a) for a in J], _i(np3), float(dt),
                       _d(dx))


def yee3d_mf(B, E, np3, dt, dx):
    lib().ora_yee3d_mf(*[dptr(a) for a in B], *[dptr(a) for a in E], _i(np3), float(dt), _d(dx))


def yee2d_ma(E, B, J, np2, dt, dx):
    lib().ora_yee2d_ma(*[dptr(a) for a in E], *[dptr(a) for a in B], *[dptr(a) for a in J], _i(np2), float(dt),
                       _d(dx))


def yee2d_mf(B, E, np2, dt, dx):
    lib().ora_yee2d_mf(*[dptr(a) for a in B], *[dptr(a) for a in E], _i(np2), float(dt), _d(dx))


# ---- multi-patch synchronisation on patch-major flat arrays --------------
def _params(ndim, order, n_cells, patch_size):
    p = OraParams()
    p.ndim, p.order = ndim, order
    nc = list(n_cells) + [1] * (3 - len(n_cells))
    ps = list(patch_size) + [1] * (3 - len(patch_size))
    for d in range(3):
        p.n_cells[d], p.patch_size[d], p.cell_length[d] = nc[d], ps[d], 1.0
    p.timestep = 0.1
    return p


def patch_block_dims(ndim, order, patch_size, dual):
    """Dims of the three components of one patch block (dual: 3x3 flags)."""
    ps = list(patch_size) + [1] * (3 - len(patch_size))
    return [tuple(ps[d] + 2 * order + 1 + dual[c][d] if d < ndim else 1 for d in range(3)) for c in range(3)]


def sum_j_ghosts(ndim, order, n_cells, patch_size, J):
    """J: flat patch-major float64 array [patch][Jx|Jy|Jz], updated in place."""
    p = _params(ndim, order, n_cells, patch_size)
    if lib().ora_sum_j_ghosts(ctypes.byref(p), dptr(J)) != 0:
        raise ValueError("oracle rejected the geometry")


def exchange_b(ndim, order, n_cells, patch_size, B):
    """B: flat patch-major float64 array [patch][Bx|By|Bz], updated in place."""
    p = _params(ndim, order, n_cells, patch_size)
    if lib().ora_exchange_b(ctypes.byref(p), dptr(B)) != 0:
        raise ValueError("oracle rejected the geometry")
