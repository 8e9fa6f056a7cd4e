"""Namelist front-end: run a Smilei-style namelist on this library.

A Smilei namelist is a Python script that calls ``Main(...)``,
``Species(...)``, ``ExternalField(...)``, ``DiagScalar(...)`` and other
blocks (src/Python/pyinit.py:156-640).  ``load`` executes such a script with
this module's block classes in its namespace and checks the blocks against
what the hot path implements; ``build`` turns the result into a
``Simulation`` (species created in namelist order, external fields
applied); ``run`` steps it for the namelist's number of time steps and
records the scalar diagnostics.

Supported (the periodic, Yee, Boris, momentum-conserving hot path):

* ``Main``: geometry ("2Dcartesian", "3Dcartesian"), interpolation_order
  (2, 4), two of cell_length / grid_length / number_of_cells,
  number_of_patches, timestep or timestep_over_CFL, simulation_time or
  number_of_timesteps (``n_time = int(simulation_time / timestep)``,
  Params.cpp:1154), periodic EM_boundary_conditions, random_seed,
  solve_poisson (default True: when the species' charge density is not
  zero everywhere, E starts as its periodic electrostatic field, solved
  exactly for the Yee grid's Laplacian; see ``poisson_e``);
* ``Species``: mass, charge, particles_per_cell, number_density or
  charge_density (numbers or profiles of the position), temperature in
  m_e c^2 (one value or three equal values), mean_velocity along x (added
  to u_x as gamma v), momentum_initialization "maxwell-juttner"/"mj"
  (isotropic Maxwell-Juttner of temperature T/m, drawn exactly by the
  library) or "cold", position_initialization "random" or the name of an
  earlier species (copy of its positions), periodic boundary_conditions,
  pusher "boris", time_frozen (immobile until then: no push, no current);
* ``ExternalField`` (initial fields, applied as applyExternalFields does);
* ``DiagScalar``: ``scalars.txt`` in the reference's text layout (every,
  vars, precision; the energy balance, per-species and per-field scalars;
  no field min/max or Poynting columns);
* ``Checkpoints``: dump_step, dump_minutes, exit_after_dump, keep_n_dumps,
  restart_dir, and the signals SIGUSR1 (dump and stop) / SIGUSR2 (dump and
  go on) of the reference (src/Checkpoint/Checkpoint.cpp:172-210); dumps
  are NumPy archives (``checkpoints/dump-NNNNN.npz``), not HDF5.

Blocks that change nothing on this path (load balancing, vectorisation and
the other diagnostics) are accepted and recorded; blocks
that would change the physics (lasers, moving window, collisions,
injectors, filters, antennas, walls, radiation, ionisation) raise ``NamelistError``
rather than run a different simulation.
"""
import glob
import math
import os
import signal
import time

import numpy as np

from . import MIXED, PLAIN_FP64, DEPOSIT_ATOMIC, Simulation
from . import profiles as P


class NamelistError(Exception):
    pass


class _Block:
    """A namelist block: its keyword arguments as attributes."""
    defaults = {}

    def __init__(self, **kw):
        for k, v in self.defaults.items():
            setattr(self, k, v)
        for k, v in kw.items():
            setattr(self, k, v)
        self._kw = kw


class _Main(_Block):
    defaults = dict(geometry=None, interpolation_order=2, interpolator="momentum-conserving", cell_length=[],
                    grid_length=[], number_of_cells=[], number_of_patches=None, timestep=None,
                    timestep_over_CFL=None, simulation_time=None, number_of_timesteps=None,
                    EM_boundary_conditions=[["periodic"]], maxwell_solver="Yee", random_seed=None,
                    print_every=None, solve_poisson=True, gpu_computing=False)


class _Species(_Block):
    defaults = dict(name=None, mass=None, charge=None, particles_per_cell=None, number_density=None,
                    charge_density=None, temperature=[1e-10], mean_velocity=[0.0, 0.0, 0.0],
                    momentum_initialization="cold", position_initialization="random", pusher="boris",
                    boundary_conditions=[["periodic"]], time_frozen=0.0, c_part_max=1.0, is_test=False,
                    ionization_model="none", radiation_model="none")


class _ExternalField(_Block):
    defaults = dict(field=None, profile=None)


class _DiagScalar(_Block):
    defaults = dict(every=1, vars=[], precision=10)


class _Checkpoints(_Block):
    defaults = dict(restart_dir=None, restart_number=None, dump_step=0, dump_minutes=0.0, keep_n_dumps=2,
                    exit_after_dump=True, dump_deflate=0, file_grouping=None)


_IGNORED = ("LoadBalancing", "Vectorization", "DiagFields", "DiagProbe", "DiagParticleBinning",
            "DiagScreen", "DiagTrackParticles", "DiagPerformances", "DiagRadiationSpectrum", "DiagNewParticles",
            "MultipleDecomposition")
_REFUSED = ("Laser", "LaserEnvelope", "LaserGaussian2D", "LaserGaussian3D", "LaserPlanar1D", "LaserOffset",
            "LaserGaussianAM", "MovingWindow", "Collisions", "ParticleInjector", "PrescribedField", "Antenna",
            "PartWall", "RadiationReaction", "MultiphotonBreitWheeler", "CurrentFilter", "FieldFilter",
            "SmileiSingleton")


class Namelist:
    """Blocks collected from one namelist."""

    def __init__(self):
        self.main = None
        self.species = []
        self.external_fields = []
        self.diag_scalar = None
        self.checkpoints = None
        self.ignored = []

    # ---- derived parameters (after validation)
    @property
    def ndim(self):
        return 2 if self.main.geometry == "2Dcartesian" else 3

    def _namespace(self):
        nl = self

        def Main(**kw):
            if nl.main is not None:
                raise NamelistError("Main defined twice")
            nl.main = _Main(**kw)

        def Species(**kw):
            nl.species.append(_Species(**kw))

        def ExternalField(**kw):
            nl.external_fields.append(_ExternalField(**kw))

        def DiagScalar(**kw):
            nl.diag_scalar = _DiagScalar(**kw)

        def Checkpoints(**kw):
            nl.checkpoints = _Checkpoints(**kw)

        ns = {"Main": Main, "Species": Species, "ExternalField": ExternalField, "DiagScalar": DiagScalar,
              "Checkpoints": Checkpoints, "math": math, "__name__": "__namelist__"}
        for name in _IGNORED:
            ns[name] = (lambda name: lambda *a, **kw: nl.ignored.append(name))(name)
        for name in _REFUSED:
            def refuse(*a, _name=name, **kw):
                raise NamelistError("%s is not part of this library's hot path (periodic, Yee, Boris, "
                                    "Esirkepov); the namelist cannot run here" % _name)
            ns[name] = refuse
        return ns

    def validate(self):
        m = self.main
        if m is None:
            raise NamelistError("the namelist defines no Main block")
        if m.geometry not in ("2Dcartesian", "3Dcartesian"):
            raise NamelistError("Main.geometry %r: only 2Dcartesian and 3Dcartesian are supported" % m.geometry)
        nd = self.ndim
        given = [len(m.cell_length) > 0, len(m.grid_length) > 0, len(m.number_of_cells) > 0]
        if sum(given) != 2:
            raise NamelistError("Main: define exactly two of grid_length, number_of_cells and cell_length")
        if not given[2]:
            m.number_of_cells = [int(round(L / d)) for L, d in zip(m.grid_length, m.cell_length)]
        elif not given[0]:
            m.cell_length = [L / n for L, n in zip(m.grid_length, m.number_of_cells)]
        m.grid_length = [n * d for n, d in zip(m.number_of_cells, m.cell_length)]
        if len(m.cell_length) != nd or len(m.number_of_cells) != nd:
            raise NamelistError("Main: cell_length / number_of_cells need %d values" % nd)
        if m.number_of_patches is None or len(m.number_of_patches) != nd:
            raise NamelistError("Main.number_of_patches needs %d values" % nd)
        for n, p in zip(m.number_of_cells, m.number_of_patches):
            if p <= 0 or n % p:
                raise NamelistError("Main: number_of_patches must divide the number of cells")
        if m.timestep is None:
            if m.timestep_over_CFL is None:
                raise NamelistError("Main: timestep and timestep_over_CFL not defined")
            m.timestep = m.timestep_over_CFL / math.sqrt(sum(1.0 / d ** 2 for d in m.cell_length))
        if m.number_of_timesteps is None:
            if m.simulation_time is None:
                raise NamelistError("Main: simulation_time and number_of_timesteps not defined")
            m.number_of_timesteps = int(m.simulation_time / m.timestep)
        for bc in m.EM_boundary_conditions:
            if any(b != "periodic" for b in bc):
                raise NamelistError("Main.EM_boundary_conditions: only periodic boundaries are supported")
        names = []
        for s in self.species:
            if s.name is None or s.mass is None or s.charge is None:
                raise NamelistError("Species needs name, mass and charge")
            if s.particles_per_cell is None:
                raise NamelistError("Species %s: particles_per_cell not defined" % s.name)
            if (s.number_density is None) == (s.charge_density is None):
                raise NamelistError("Species %s: define one of number_density and charge_density" % s.name)
            for bc in s.boundary_conditions:
                if any(b != "periodic" for b in bc):
                    raise NamelistError("Species %s: only periodic boundary conditions are supported" % s.name)
            if s.pusher != "boris":
                raise NamelistError("Species %s: pusher %r is not supported (boris only)" % (s.name, s.pusher))
            if callable(s.time_frozen) or not float(s.time_frozen) >= 0:
                raise NamelistError("Species %s: time_frozen must be a number >= 0" % s.name)
            if s.ionization_model != "none" or s.radiation_model != "none":
                raise NamelistError("Species %s: ionisation / radiation are not supported" % s.name)
            mv = list(s.mean_velocity) + [0.0] * (3 - len(s.mean_velocity))
            if callable(mv[0]) or any(callable(v) for v in mv) or mv[1] != 0 or mv[2] != 0:
                raise NamelistError("Species %s: mean_velocity must be a constant along x" % s.name)
            if not abs(mv[0]) < 1:
                raise NamelistError("Species %s: |mean_velocity| must be below 1 (c)" % s.name)
            T = list(s.temperature)
            if any(callable(t) for t in T) or len(set(T)) != 1:
                raise NamelistError("Species %s: temperature must be one constant (isotropic)" % s.name)
            mi = s.momentum_initialization
            if mi not in ("maxwell-juttner", "mj", "cold"):
                raise NamelistError("Species %s: momentum_initialization %r not supported" % (s.name, mi))
            pi = s.position_initialization
            if pi != "random" and pi not in names:
                raise NamelistError("Species %s: position_initialization %r (random or an earlier species name)"
                                    % (s.name, pi))
            names.append(s.name)
        for e in self.external_fields:
            if e.field not in P.DUAL or e.field.startswith("J"):
                raise NamelistError("ExternalField: field %r is not an E, B or B_m component" % e.field)
        if self.diag_scalar is not None:
            d = self.diag_scalar
            if int(d.every) < 1:
                raise NamelistError("DiagScalar.every must be a positive number of steps")
            known = set(scalar_names([s.name for s in self.species]))
            bad = [v for v in d.vars if v not in known]
            if bad:
                raise NamelistError("DiagScalar.vars %s: not computed here (field min/max and Poynting "
                                    "scalars are not part of this library)" % bad)
        c = self.checkpoints
        if c is not None:
            if int(c.keep_n_dumps) < 1:
                raise NamelistError("Checkpoints.keep_n_dumps must be >= 1")
            if c.restart_number is not None:
                raise NamelistError("Checkpoints.restart_number is not supported (the latest dump is used)")
        return self


# DiagnosticScalar's scalars, in the order of DiagnosticScalar::init
# (src/Diagnostic/DiagnosticScalar.cpp:249-286); the field min/max and
# Poynting scalars that follow there are not computed here.
_GENERAL = ["Ubal_norm", "Ubal", "Utot", "Uexp", "Ukin", "Urad", "UmBWpairs", "Uelm", "Ukin_bnd", "Ukin_out_mvw",
            "Ukin_inj_mvw", "Ukin_new", "Uelm_bnd", "Uelm_out_mvw", "Uelm_inj_mvw"]
_UELM = ["Uelm_Ex", "Uelm_Ey", "Uelm_Ez", "Uelm_Bx_m", "Uelm_By_m", "Uelm_Bz_m"]


def scalar_names(species):
    names = list(_GENERAL)
    for s in species:
        names += ["Dens_" + s, "Ntot_" + s, "Zavg_" + s, "Ukin_" + s, "Urad_" + s]
    return names + _UELM


def scalar_values(sim, charges, names, e_time_zero):
    """The scalars of `names` for the current state, with the energy
    balance of SmileiMPI::computeGlobalDiags (SmileiMPI.cpp:2366-2424):
    Utot = Ukin + Uelm + Urad, Uexp = Utot at t = 0 plus the boundary and
    injection terms (zero with periodic boundaries), Ubal = Utot - Uexp,
    Ubal_norm = Ubal / Utot; Zavg = sum(w q) / sum(w)."""
    d = sim.scalars()
    v = dict.fromkeys(_GENERAL, 0.0)
    for f in _UELM:
        v[f] = d[f]
    v["Uelm"] = sum(d[f] for f in _UELM)
    v["Ukin"] = sum(d["Ukin"])
    v["Utot"] = v["Ukin"] + v["Uelm"] + v["Urad"]
    e0 = v["Utot"] if e_time_zero is None else e_time_zero
    v["Uexp"] = e0
    v["Ubal"] = v["Utot"] - v["Uexp"]
    v["Ubal_norm"] = v["Ubal"] / v["Utot"] if v["Utot"] > 0 else 0.0
    for k, s in enumerate(sim.species):
        dens = d["Dens"][k]
        v["Dens_" + s] = dens
        v["Ntot_" + s] = d["Ntot"][k]
        v["Zavg_" + s] = charges[k] * dens / dens if dens else float("nan")
        v["Ukin_" + s] = d["Ukin"][k]
        v["Urad_" + s] = 0.0
    return [v[n] for n in names], e0


class ScalarFile:
    """``scalars.txt`` in the layout of DiagnosticScalar::write
    (src/Diagnostic/DiagnosticScalar.cpp:350-397): a "# j name" line per
    column, a "#" line, the names in one line, then one row per recorded
    step, numbers in scientific notation with `precision` digits, each
    column right-aligned in 2 + max(len(name), precision + 8) characters."""

    def __init__(self, path, names, precision=10, append=False):
        self.names, self.p = names, int(precision)
        self.f = open(path, "a" if append else "w")
        if self.f.tell() == 0:
            self.f.write("# 1 time\n")
            for j, n in enumerate(names):
                self.f.write("# %d %s\n" % (j + 2, n))
            self.f.write("#\n#" + "time".rjust(self.p + 9)
                         + "".join(n.rjust(self.width(n)) for n in names) + "\n")

    def width(self, name):
        return 2 + max(len(name), self.p + 8)

    def write(self, t, values):
        fmt = "%%.%de" % self.p
        self.f.write((fmt % t).rjust(self.p + 10)
                     + "".join((fmt % v).rjust(self.width(n)) for n, v in zip(self.names, values)) + "\n")
        self.f.flush()

    def close(self):
        self.f.close()


def load(source):
    """Execute a namelist (a file path, or the script text when it contains
    a newline) and return its validated blocks."""
    if "\n" not in source:
        with open(source) as f:
            source = f.read()
    nl = Namelist()
    exec(compile(source, "<namelist>", "exec"), nl._namespace())
    return nl.validate()


def _density(s):
    if s.number_density is not None:
        return s.number_density
    q = abs(float(s.charge))
    cd = s.charge_density
    return (lambda x, y, z: np.abs(np.asarray(cd(x, y, z), dtype=float)) / q) if callable(cd) else abs(cd) / q


def poisson_e(rho, dx):
    """Electrostatic field of a periodic charge density on the Yee grid.

    Solves -lap(phi) = rho - <rho> with the discrete 5/7-point Laplacian of
    the primal grid (the operator the reference's conjugate-gradient solver
    inverts, ElectroMagn3D::compute_Ap, ElectroMagn3D.cpp:312-330) exactly,
    by its eigen-decomposition: the Laplacian is diagonal in the discrete
    Fourier basis with eigenvalues -sum_d (2 sin(pi m_d / N_d) / dx_d)^2.
    E_d = -(phi[i] - phi[i-1]) / dx_d sits at the dual index i ((i - 1/2) dx_d)
    as in ElectroMagn3D::initE (ElectroMagn3D.cpp:529-), so the discrete
    div E equals rho - <rho> to rounding.  The mean of rho (the k = 0 mode,
    which no periodic field can carry) is left out, which is what the
    reference's centring of E (VectorPatch::solvePoisson, VectorPatch.cpp:
    1645-1670) amounts to.  rho has the simulation's dimensions; returns one
    array per dimension."""
    rho = np.asarray(rho, dtype=np.float64)
    nd = rho.ndim
    rhat = np.fft.fftn(rho)
    k2 = np.zeros(rho.shape)
    for d in range(nd):
        m = np.arange(rho.shape[d])
        kd = (2.0 * np.sin(np.pi * m / rho.shape[d]) / dx[d]) ** 2
        shp = [1] * nd
        shp[d] = rho.shape[d]
        k2 = k2 + kd.reshape(shp)
    k2.flat[0] = 1.0
    phihat = rhat / k2
    phihat.flat[0] = 0.0
    phi = np.fft.ifftn(phihat).real
    return [-(phi - np.roll(phi, 1, axis=d)) / dx[d] for d in range(nd)]


def solve_poisson(sim):
    """Main.solve_poisson at t = 0 (Smilei.cpp:329-332): when the charge
    density of the species is not zero everywhere, E becomes the periodic
    electrostatic field of that density (poisson_e).  Returns whether it
    was solved."""
    nd = sim.ndim
    rho = sim.rho().reshape(sim.shape[:nd])
    if not np.any(rho):
        return False
    E = poisson_e(rho, sim.dx[:nd])
    for d, name in enumerate(("Ex", "Ey", "Ez")[:nd]):
        sim.set_field(name, E[d].reshape(sim.shape))
    return True


def build(nl, precision=PLAIN_FP64, deposit=DEPOSIT_ATOMIC, use_graph=True):
    """Create the Simulation of a loaded namelist: species in namelist order,
    the electrostatic field of their charge when Main.solve_poisson (the
    reference default), then the external fields (added onto it, as
    ElectroMagn::applyExternalFields adds its profiles)."""
    m = nl.main
    patch = [n // p for n, p in zip(m.number_of_cells, m.number_of_patches)]
    sim = Simulation(m.geometry, m.interpolation_order, m.number_of_cells, patch, m.cell_length, m.timestep,
                     precision=precision, deposit=deposit, use_graph=use_graph, interpolator=m.interpolator,
                     maxwell_solver=m.maxwell_solver)
    seed = int(m.random_seed or 0)
    ids = {}
    ncell = int(np.prod(m.number_of_cells))
    for k, s in enumerate(nl.species):
        dens = _density(s)
        theta = float(s.temperature[0]) / float(s.mass) if s.momentum_initialization != "cold" else 0.0
        v = float(s.mean_velocity[0])
        drift = v / math.sqrt(1.0 - v * v)
        ppc = s.particles_per_cell
        if s.position_initialization != "random":
            src = ids[s.position_initialization]
            cap = sim.num_particles(src)
            sid = sim.add_species(s.name, s.mass, s.charge, cap)
            src_dens = _density(next(b for b in nl.species if b.name == s.position_initialization))
            if callable(dens) or callable(src_dens):
                raise NamelistError("Species %s: copied positions need constant densities" % s.name)
            sim.init_species_like(sid, src, weight_scale=dens / src_dens if src_dens else 0.0, theta=theta,
                                  drift_ux=drift, seed=seed + 1 + k)
        elif not callable(dens) and not callable(ppc) and float(ppc) == int(ppc) and dens > 0:
            sid = sim.add_species(s.name, s.mass, s.charge, ncell * int(ppc))
            sim.init_species_mj(sid, int(ppc), dens, theta, drift_ux=drift, seed=seed + 1 + k)
        else:
            cap = P.count_particles(sim, ppc, dens)
            sid = sim.add_species(s.name, s.mass, s.charge, max(cap, 1))
            P.create_species(sim, sid, dens, ppc, theta=theta, drift_ux=drift, seed=seed + 1 + k)
        ids[s.name] = sid
        if float(s.time_frozen) > 0:
            sim.set_time_frozen(sid, float(s.time_frozen))
    if m.solve_poisson:
        solve_poisson(sim)
    if nl.external_fields:
        P.apply_external_fields(sim, {e.field: e.profile for e in nl.external_fields})
    return sim


def _latest_dump(restart_dir):
    """The dump of restart_dir/checkpoints with the most steps done."""
    import json
    best, best_step = None, -1
    for path in glob.glob(os.path.join(restart_dir, "checkpoints", "dump-*.npz")):
        with np.load(path, allow_pickle=False) as z:
            step = int(json.loads(str(z["meta"]))["steps_done"])
        if step > best_step:
            best, best_step = path, step
    if best is None:
        raise NamelistError("Checkpoints.restart_dir %r holds no dump" % restart_dir)
    return best


def run(source, precision=PLAIN_FP64, deposit=DEPOSIT_ATOMIC, steps=None, scalars_path=None, output_dir=".",
        on_step=None):
    """Load, build and run a namelist for its number of time steps (or
    `steps`, counted from the start of this run).  Returns (simulation,
    rows): rows = (time, Utot, Ukin, Uelm, Ukin_<species>...) at step 0 and
    every DiagScalar.every steps (every step without a DiagScalar block).
    With a DiagScalar block its scalars go to `scalars_path` (default
    output_dir/scalars.txt) in the reference's layout; a Checkpoints block
    dumps to output_dir/checkpoints and restarts from restart_dir.
    SIGUSR1 during the run dumps and stops, SIGUSR2 dumps and goes on;
    on_step(it), when given, is called after every step (progress, or
    raising a signal at a chosen step)."""
    nl = load(source)
    ck = nl.checkpoints
    extra = {}
    if ck is not None and ck.restart_dir:
        sim = Simulation.restart(_latest_dump(ck.restart_dir), precision=precision, deposit=deposit)
        extra = sim.extra
    else:
        sim = build(nl, precision=precision, deposit=deposit)
    start = sim.steps_done
    n_end = nl.main.number_of_timesteps if steps is None else start + int(steps)
    dt = nl.main.timestep
    diag = nl.diag_scalar
    every = max(int(diag.every), 1) if diag is not None else 1
    charges = [float(s.charge) for s in nl.species]
    names = scalar_names(sim.species)
    if diag is not None and diag.vars:
        names = [n for n in names if n in diag.vars]
    sfile = None
    if diag is not None:
        sfile = ScalarFile(scalars_path or os.path.join(output_dir, "scalars.txt"), names, diag.precision,
                           append=start > 0)
    e0 = extra.get("Energy_time_zero")
    rows = []

    def record(it):
        nonlocal e0
        e = sim.energies()
        rows.append([it * dt, float(e.sum()), float(e[1:].sum()), float(e[0])] + list(e[1:]))
        if sfile is not None:
            vals, e0 = scalar_values(sim, charges, names, e0)
            sfile.write(it * dt, vals)

    got = []
    handlers = {}
    for sig in (signal.SIGUSR1, signal.SIGUSR2):
        try:
            handlers[sig] = signal.signal(sig, lambda signum, frame: got.append(signum))
        except ValueError:  # not the main thread: no signal triggers
            pass
    dump_number = int(extra.get("dump_number", 0))
    t_ref = time.time()
    try:
        if start == 0:
            record(0)  # a restarted run continues the file of the dumped one
        it = start
        while it < n_end:
            sim.step(1)
            it += 1
            if on_step is not None:
                on_step(it)
            if it % every == 0:
                record(it)
            if ck is None and not got:
                continue
            due = got or (ck is not None and ((int(ck.dump_step) > 0 and (it - start) % int(ck.dump_step) == 0)
                                              or (float(ck.dump_minutes) > 0
                                                  and time.time() - t_ref > 60.0 * float(ck.dump_minutes))))
            if due:
                keep = int(ck.keep_n_dumps) if ck is not None else 2
                os.makedirs(os.path.join(output_dir, "checkpoints"), exist_ok=True)
                sim.checkpoint(os.path.join(output_dir, "checkpoints", "dump-%05d.npz" % (dump_number % keep)),
                               extra={"Energy_time_zero": e0, "dump_number": dump_number + 1})
                dump_number += 1
                t_ref = time.time()
                stop = (ck is None or bool(ck.exit_after_dump)) or (got and got[-1] != signal.SIGUSR2)
                del got[:]
                if stop:
                    break
    finally:
        for sig, h in handlers.items():
            signal.signal(sig, h)
        if sfile is not None:
            sfile.close()
    return sim, np.array(rows)
