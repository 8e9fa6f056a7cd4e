"""smilei_amd -- MI355X-native PIC time step of Smilei behind a C ABI.

Python only orchestrates: the particle loop, the sort/exchange and the Yee
solver run in libsmilei_amd.so (HIP, gfx950).  ``Simulation`` wraps one
context of the C ABI (include/smilei_amd.h); ``profiles`` evaluates
namelist-style density / particles-per-cell profiles and external fields on
the grid.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import SmiParams, check

__all__ = ["Simulation", "FIELDS", "PLAIN_FP64", "MIXED", "DEPOSIT_ATOMIC", "DEPOSIT_DETERMINISTIC",
           "TUNE_INTERLEAVE", "TUNE_THREADS", "TUNE_PERSISTENT", "TUNE_CELL_ORDER", "TUNE_XCD", "TUNE_TILE",
           "TUNE_TILE_X", "TUNE_NODE_ACC"]

PLAIN_FP64, MIXED = 0, 1
DEPOSIT_ATOMIC, DEPOSIT_DETERMINISTIC = 0, 1
TUNE_INTERLEAVE, TUNE_THREADS, TUNE_PERSISTENT, TUNE_CELL_ORDER, TUNE_XCD, TUNE_TILE, TUNE_TILE_X = 1, 2, 3, 4, 5, 6, 7
TUNE_NODE_ACC = 8
FIELDS = ["Ex", "Ey", "Ez", "Bx", "By", "Bz", "Bx_m", "By_m", "Bz_m", "Jx", "Jy", "Jz"]
_FID = {n: i for i, n in enumerate(FIELDS)}
_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int32)


def _dp(a):
    return a.ctypes.data_as(_D) if a is not None else None


class Simulation:
    """One simulation (all patches, all species) resident on one GPU.

    Parameters follow Smilei's namelist: ``geometry`` ("2Dcartesian" /
    "3Dcartesian"), ``interpolation_order`` (2, or 4 in 3D), ``n_cells``
    (global cells per axis), ``patch_size`` (cells per patch, multiple of 4),
    ``cell_length``, ``timestep``; ``deposit`` selects the FP64 atomic tile
    deposit or the deterministic per-patch one (reference summation order).
    """

    def __init__(self, geometry="3Dcartesian", interpolation_order=2, n_cells=(32, 32, 32),
                 patch_size=(16, 16, 16), cell_length=(0.2, 0.2, 0.2), timestep=0.1,
                 precision=PLAIN_FP64, deposit=DEPOSIT_ATOMIC, use_graph=True, device=0,
                 interpolator="momentum-conserving", pusher="boris", maxwell_solver="Yee"):
        self._L = _lib.load()
        self._args = dict(geometry=geometry, interpolation_order=int(interpolation_order),
                          n_cells=[int(v) for v in n_cells], patch_size=[int(v) for v in patch_size],
                          cell_length=[float(v) for v in cell_length], timestep=float(timestep),
                          precision=int(precision), deposit=int(deposit), use_graph=bool(use_graph),
                          interpolator=interpolator, pusher=pusher, maxwell_solver=maxwell_solver)
        self._species_args = []
        self.extra = {}
        nd = 2 if geometry == "2Dcartesian" else 3
        p = SmiParams()
        self._keep = [s.encode() for s in (geometry, interpolator, pusher, maxwell_solver)]
        p.geometry, p.interpolator, p.pusher, p.maxwell_solver = self._keep
        p.interpolation_order = int(interpolation_order)
        nc = list(n_cells) + [1] * (3 - len(n_cells))
        ps = list(patch_size) + [1] * (3 - len(patch_size))
        cl = list(cell_length) + [1.0] * (3 - len(cell_length))
        for d in range(3):
            p.n_cells[d], p.patch_size[d], p.cell_length[d] = int(nc[d]), int(ps[d]), float(cl[d])
        p.timestep = float(timestep)
        p.precision, p.deposit, p.use_graph, p.device = int(precision), int(deposit), int(bool(use_graph)), int(device)
        h = ctypes.c_void_p()
        check(self._L.smi_create(ctypes.byref(p), ctypes.byref(h)))
        self.h = h
        self.ndim = nd
        self.dx = tuple(float(cl[d]) for d in range(3))
        self.dt = float(timestep)
        self.shape = tuple(int(nc[d]) if d < nd else 1 for d in range(3))
        self.npatch = int(np.prod([self.shape[d] // int(ps[d]) if d < nd else 1 for d in range(3)]))
        self.species = []

    def close(self):
        if getattr(self, "h", None):
            self._L.smi_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    # ---- species / particles
    def add_species(self, name, mass, charge, capacity):
        sid = ctypes.c_int()
        check(self._L.smi_add_species(self.h, name.encode(), float(mass), float(charge), int(capacity),
                                      ctypes.byref(sid)))
        self.species.append(name)
        self._species_args.append(dict(name=name, mass=float(mass), charge=float(charge), capacity=int(capacity)))
        return sid.value

    def set_time_frozen(self, s, time_frozen):
        """Species::time_frozen_: the species is not pushed, deposits no
        current and is not re-sorted while a step's time_dual (steps done +
        1.5, times the timestep) is at most time_frozen."""
        check(self._L.smi_set_time_frozen(self.h, s, float(time_frozen)))
        self._species_args[s]["time_frozen"] = float(time_frozen)

    @property
    def steps_done(self):
        """Time steps done (the engine's counter)."""
        return int(self._L.smi_step_count(self.h))

    @steps_done.setter
    def steps_done(self, n):
        check(self._L.smi_set_step_count(self.h, int(n)))

    def set_particles(self, s, x, y, z, px, py, pz, w):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z, px, py, pz, w)]
        n = len(arrs[-1])
        check(self._L.smi_set_particles(self.h, s, n, *[_dp(a) for a in arrs]))

    def init_species(self, s, ppc, density, sigma, drift_ux=0.0, radius=0.0, seed=0):
        check(self._L.smi_init_species(self.h, s, int(ppc), float(density), float(sigma), float(drift_ux),
                                       float(radius), int(seed)))

    # ---- launch tuning
    def set_tuning(self, key, value):
        check(self._L.smi_set_tuning(self.h, int(key), int(value)))

    def get_tuning(self, key):
        v = ctypes.c_int()
        check(self._L.smi_get_tuning(self.h, int(key), ctypes.byref(v)))
        return v.value

    def autotune(self, reps=2):
        """Pick the particle-kernel launch shape by timing the candidates on
        the current state; returns the chosen knobs and their time."""
        ms = ctypes.c_float()
        check(self._L.smi_autotune(self.h, int(reps), ctypes.byref(ms)))
        return dict(interleave=self.get_tuning(TUNE_INTERLEAVE), threads=self.get_tuning(TUNE_THREADS),
                    persistent=self.get_tuning(TUNE_PERSISTENT), xcd=self.get_tuning(TUNE_XCD),
                    tile=self.get_tuning(TUNE_TILE), tile_x=self.get_tuning(TUNE_TILE_X),
                    node_acc=self.get_tuning(TUNE_NODE_ACC), particles_ms=ms.value)

    def init_species_mj(self, s, ppc, density, theta, drift_ux=0.0, radius=0.0, seed=0):
        """Maxwell-Juttner momenta of temperature theta = T / (m c^2)."""
        check(self._L.smi_init_species_mj(self.h, s, int(ppc), float(density), float(theta), float(drift_ux),
                                          float(radius), int(seed)))

    def init_species_like(self, s, src, weight_scale=1.0, theta=0.0, drift_ux=0.0, seed=0):
        """Particles at the positions of species `src`, weights scaled,
        fresh Maxwell-Juttner momenta of temperature theta (0: cold)."""
        check(self._L.smi_init_species_like(self.h, s, src, float(weight_scale), float(theta), float(drift_ux),
                                            int(seed)))

    def init_species_profile(self, s, density, ppc, sigma=0.0, drift_ux=0.0, seed=0):
        """Per-cell density and particles-per-cell arrays (shape of the grid,
        values at cell centres), ParticleCreator's rounding and weights."""
        d = np.ascontiguousarray(np.broadcast_to(density, self.shape), dtype=np.float64)
        n = np.ascontiguousarray(np.broadcast_to(ppc, self.shape), dtype=np.float64)
        check(self._L.smi_init_species_profile(self.h, s, _dp(d), _dp(n), float(sigma), float(drift_ux), int(seed)))

    def init_species_profile_mj(self, s, density, ppc, theta, drift_ux=0.0, seed=0):
        """Profiles as init_species_profile, Maxwell-Juttner momenta."""
        d = np.ascontiguousarray(np.broadcast_to(density, self.shape), dtype=np.float64)
        n = np.ascontiguousarray(np.broadcast_to(ppc, self.shape), dtype=np.float64)
        check(self._L.smi_init_species_profile_mj(self.h, s, _dp(d), _dp(n), float(theta), float(drift_ux),
                                                  int(seed)))

    def num_particles(self, s):
        return self._L.smi_num_particles(self.h, s)

    def get_particles(self, s):
        n = self.num_particles(s)
        out = {k: np.empty(n) for k in ("x", "y", "z", "px", "py", "pz", "w")}
        out["patch"] = np.empty(n, np.int32)
        out["key"] = np.empty(n, np.int32)
        check(self._L.smi_get_particles(self.h, s, *[_dp(out[k]) for k in ("x", "y", "z", "px", "py", "pz", "w")],
                                        out["patch"].ctypes.data_as(_I), out["key"].ctypes.data_as(_I)))
        return out

    def patch_counts(self, s):
        c = np.empty(self.npatch, np.int32)
        check(self._L.smi_patch_counts(self.h, s, c.ctypes.data_as(_I)))
        return c

    # ---- fields
    def set_field(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(self.shape)
        check(self._L.smi_set_field(self.h, _FID[name], _dp(a)))

    def get_field(self, name):
        a = np.empty(self.shape)
        check(self._L.smi_get_field(self.h, _FID[name], _dp(a)))
        return a

    # ---- time loop
    def step(self, n=1):
        check(self._L.smi_step(self.h, int(n)))

    # ---- checkpoint / restart (the reference dumps HDF5, Checkpoint.cpp:212-587;
    # here a NumPy archive of the same state: every field array, every
    # species' particles in canonical order, the step count)
    def checkpoint(self, path, extra=None):
        """Write the full simulation state to `path` (.npz); `extra`: a JSON-able
        dict stored with it (returned as `.extra` by `restart`)."""
        import json
        data = {"meta": np.array(json.dumps(dict(args=self._args, species=self._species_args,
                                                 steps_done=self.steps_done, extra=extra or {})))}
        for f in FIELDS:
            data["field_" + f] = self.get_field(f)
        for s in range(len(self.species)):
            p = self.get_particles(s)
            for k in ("x", "y", "z", "px", "py", "pz", "w"):
                data["sp%d_%s" % (s, k)] = p[k]
        np.savez(path, **data)

    @classmethod
    def restart(cls, path, **overrides):
        """A Simulation holding the state written by `checkpoint`
        (constructor arguments may be overridden, e.g. precision)."""
        import json
        with np.load(path, allow_pickle=False) as z:
            meta = json.loads(str(z["meta"]))
            args = dict(meta["args"])
            args.update(overrides)
            sim = cls(**args)
            for s, sp in enumerate(meta["species"]):
                sid = sim.add_species(sp["name"], sp["mass"], sp["charge"], sp["capacity"])
                sim.set_particles(sid, *[z["sp%d_%s" % (s, k)] for k in ("x", "y", "z", "px", "py", "pz", "w")])
                if sp.get("time_frozen", 0.0):
                    sim.set_time_frozen(sid, sp["time_frozen"])
            for f in FIELDS:
                sim.set_field(f, z["field_" + f])
            sim.steps_done = int(meta["steps_done"])
            sim.extra = meta.get("extra", {})
        return sim

    def synchronize(self):
        check(self._L.smi_synchronize(self.h))

    def profile_step(self):
        ms = (ctypes.c_float * 4)()
        check(self._L.smi_profile_step(self.h, ms))
        return {"particles": ms[0], "sort": ms[1], "j_reduce": ms[2], "maxwell": ms[3]}

    # ---- diagnostics
    def energies(self):
        out = np.empty(1 + len(self.species))
        check(self._L.smi_energies(self.h, _dp(out)))
        return out

    def scalars(self):
        """Per-field and per-species sums of DiagnosticScalar (see
        smi_scalars): {"Uelm_Ex": .., ..., "Uelm_Bz_m": .., "Ntot": [..],
        "Dens": [..], "Ukin": [..]} (lists over species)."""
        ns = len(self.species)
        out = np.empty(6 + 3 * ns)
        check(self._L.smi_scalars(self.h, _dp(out)))
        d = {"Uelm_" + f: float(out[k]) for k, f in enumerate(("Ex", "Ey", "Ez", "Bx_m", "By_m", "Bz_m"))}
        for k, key in enumerate(("Ntot", "Dens", "Ukin")):
            d[key] = [float(out[6 + 3 * s + k]) for s in range(ns)]
        return d

    def rho(self):
        a = np.empty(self.shape)
        check(self._L.smi_rho(self.h, _dp(a)))
        return a

    def gauss(self):
        a = np.empty(self.shape)
        check(self._L.smi_gauss(self.h, _dp(a)))
        return a
