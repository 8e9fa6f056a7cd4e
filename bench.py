#!/usr/bin/env python3
"""Particle-pushes/s of the PIC time step on one MI355X.

Default workload (3D beam-plasma, 1.08e9 particles): 1024x256x256 cells,
dx = 0.25, dt = 0.95*CFL, 32x16x16-cell patches (8192 patches); electron
and proton background at 8 particles/cell each, thermal sigma = 0.02 m_e c
(protons: same temperature, 0.02/sqrt(1836) in m_p c); a cold electron beam,
density 0.1, gamma = 50 along x (u_x = sqrt(50^2 - 1)), 8 particles/cell in
the cells whose centre lies within 16 cells of the axis.  2nd-order shapes,
Yee, periodic.  `--workload thermal` selects the 3D thermal plasma
(256x128x128 cells, dx = 0.2, 16^3 patches, e-/p+ 8 ppc, sigma 0.05 m_e c,
67,108,864 particles); `--workload lwfa` the laser-wakefield-like setup of
uneven patches (512x128x128 cells, 1024 patches holding 0 to 4096
electrons, laser as initial field; no CPU baseline).

The particles are initialised in HBM; one "step" is the full PIC step
(gather, Boris push, Esirkepov deposit, cell sort + patch exchange, Yee
update) over every species.  After the warm-up steps the library's
autotuner picks the particle kernels' launch shape (cluster storage layout,
workgroup size, persistent work list) on the live state, outside the timed
region.
The order of particles inside a cell is left to the atomic scatter, as in
the reference's GPU sort (`--cell-order canonical` adds the in-cell sort).  Both precision forms run in turn (mixed first:
it is the headline); each is timed over the same number of steps.

Prints one JSON line.  `value` = particles * steps / wall time of the timed
steps.  `roofline` rates the dominant kernel (the fused gather-push-deposit
kernel) against HBM: algorithmic bytes per push (mixed: read cell key,
3 FP32 offsets, 3 FP32 momenta, FP32 weight = 32 B, write key, offsets and
momenta = 28 B: 60 B; FP64: read x,y,z,ux,uy,uz,w = 56 B, write x,y,z,
ux,uy,uz = 48 B, write the key 4 B: 108 B) divided by that kernel's device
time, measured with HIP events on the library's stream over the timed
steps.  `cpu_baseline` times the CPU oracle (OpenMP over patches, the
reference's CPU structure) on a bounded sub-box of the same plasma.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# algorithmic particle bytes per push (read + write of the particle state)
BYTES_PER_PUSH = {"fp64": 108.0, "mixed": 60.0}
# HBM bytes per push of the particle kernels on the beam-plasma workload,
# measured with rocprofv3 FETCH_SIZE / WRITE_SIZE (scaled by a 4 GiB device
# copy): profiles/k_particles_beam1B_traffic.md.  Re-measured when the kernels
# change; None for workloads without a measurement.
TRAFFIC_PER_PUSH = {"beam": {"mixed": 135.0, "fp64": 157.7}}
# VALU instructions per push of the mixed particle kernel on the beam-plasma
# workload (SQ_INSTS_VALU per launch / particles, rocprofv3 --pmc:
# profiles/k_particles_mixed_beam1B_sq.md), and the chip's VALU issue peak:
# 256 CUs x 4 SIMDs x 2.4 GHz, one wave64 instruction per 4 cycles per SIMD.
VALU_INSTR_PER_PUSH = {"beam": {"mixed": 969.0 / 64.0}}
VALU_PEAK_WAVE_INSTR = 256 * 4 * 2.4e9 / 4
# configured run length of each workload (steps)
RUN_STEPS = {"beam": 300, "thermal": 200, "lwfa": 0}


def cfl_dt(dx, factor=0.95):
    return factor / np.sqrt(sum(1.0 / d ** 2 for d in dx))


def build_lwfa(sim):
    """Laser-wakefield-like setup (profiles and laser as initial field,
    electrons only; the immobile ions are a uniform background no step
    touches): per-patch counts from 0 to 4096."""
    from smilei_amd import profiles as P
    from tests.plasma import lwfa_profiles
    density, ppc, fields = lwfa_profiles(sim.shape, sim.dx)
    sid = sim.add_species("electron", 1.0, -1.0, P.count_particles(sim, ppc, density))
    P.create_species(sim, sid, density, ppc, sigma=0.0, seed=7)
    P.apply_external_fields(sim, fields)
    return [sim.num_particles(sid)]


def workload(name="beam", ppc=8):
    """Species: (name, mass, charge, sigma, ppc, density, drift_ux, radius)."""
    if name == "lwfa":
        dx = (0.1, 0.5, 0.5)
        return dict(name=name, shape=(512, 128, 128), patch=(32, 16, 16), dx=dx, dt=cfl_dt(dx), species=[],
                    build=build_lwfa,
                    desc="3D laser-wakefield-like setup 512x128x128 cells, dx=0.1, dy=dz=0.5, 1024 patches of "
                         "32x16x16, n=0.01 ramp over 100 cells from x=64 + parabolic channel r0=20, checkerboard "
                         "1 ppc (0..4096 per patch), a0=2 Gaussian laser as initial field, electrons")
    if name == "thermal":
        shape = (256, 128, 128)
        dx = (0.2, 0.2, 0.2)
        return dict(name=name, shape=shape, patch=(16, 16, 16), dx=dx, dt=cfl_dt(dx),
                    species=[("electron", 1.0, -1.0, 0.05, 8, 1.0, 0.0, 0.0),
                             ("proton", 1836.0, 1.0, 0.05 / np.sqrt(1836.0), 8, 1.0, 0.0, 0.0)],
                    desc="3D thermal plasma 256x128x128 cells, dx=0.2, 16^3-cell patches, e-/p+ 8 ppc each, "
                         "sigma 0.05 m_e c")
    shape = (1024, 256, 256)
    dx = (0.25, 0.25, 0.25)
    ub = float(np.sqrt(50.0 ** 2 - 1.0))
    return dict(name=name, shape=shape, patch=(32, 16, 16), dx=dx, dt=cfl_dt(dx),
                species=[("electron", 1.0, -1.0, 0.02, ppc, 1.0, 0.0, 0.0),
                         ("proton", 1836.0, 1.0, 0.02 / np.sqrt(1836.0), ppc, 1.0, 0.0, 0.0),
                         ("beam", 1.0, -1.0, 0.0, 8, 0.1, ub, 16.0)],
                desc="3D beam-plasma 1024x256x256 cells, dx=0.25, 32x16x16-cell patches, e-/p+ background "
                     "%d ppc each sigma 0.02 m_e c, cold e- beam gamma=50 n=0.1 8 ppc r<=16 cells" % ppc)


def cylinder_cells(shape, radius):
    """Number of cells whose centre lies within `radius` cells of the x axis."""
    j = np.arange(shape[1]) + 0.5 - 0.5 * shape[1]
    k = np.arange(shape[2]) + 0.5 - 0.5 * shape[2]
    return int(np.count_nonzero(j[:, None] ** 2 + k[None, :] ** 2 <= radius * radius)) * shape[0]


def make_sim(args, W, form):
    """The workload's simulation in one form, particles created in HBM from
    fixed seeds (both forms draw the same particles; the mixed form stores
    them rounded to its FP32 offsets and momenta)."""
    import smilei_amd as sa
    prec = sa.MIXED if form == "mixed" else sa.PLAIN_FP64
    sim = sa.Simulation("3Dcartesian", args.order, W["shape"], W["patch"], W["dx"], W["dt"],
                        precision=prec, deposit=sa.DEPOSIT_ATOMIC, use_graph=True)
    sim.set_tuning(sa.TUNE_CELL_ORDER, 1 if args.cell_order == "canonical" else 0)
    if args.threads:
        sim.set_tuning(sa.TUNE_THREADS, args.threads)
    if args.tile_x and form == "mixed":
        sim.set_tuning(sa.TUNE_TILE_X, args.tile_x)
    if args.layout >= 0:
        sim.set_tuning(sa.TUNE_INTERLEAVE, args.layout)
    if args.tile >= 0 and form == "mixed":
        sim.set_tuning(sa.TUNE_TILE, args.tile)
    if args.node_acc >= 0 and form == "mixed":
        sim.set_tuning(sa.TUNE_NODE_ACC, args.node_acc)
    ncell = int(np.prod(W["shape"]))
    per_species = W["build"](sim) if "build" in W else []
    for k, (name, m, q, sig, ppc, dens, drift, radius) in enumerate(W["species"]):
        cells = cylinder_cells(W["shape"], radius) if radius > 0 else ncell
        sid = sim.add_species(name, m, q, cells * ppc)
        sim.init_species(sid, ppc, dens, sig, drift_ux=drift, radius=radius, seed=1 + k)
        per_species.append(sim.num_particles(sid))
    return sim, per_species


def profile(sim, steps):
    """Device time per kernel class over `steps` eager steps (HIP events on
    the library's stream), ms summed over the steps."""
    acc = {"particles": 0.0, "sort": 0.0, "j_reduce": 0.0, "maxwell": 0.0}
    for _ in range(steps):
        ms = sim.profile_step()
        for k in acc:
            acc[k] += ms[k]
    return acc


def run_gpu(args, W, form):
    """One form along the workload's trajectory: W warm-up steps, the
    autotuner, the K timed graph-replayed steps (the headline: wall clock
    between two device synchronisations), P eager steps with per-kernel-class
    events; then, with --run-steps R > 0, the trajectory continues to step
    R in blocks of 50 graph-replayed steps timed with device
    synchronisations, and P more eager steps are profiled at its end (the
    kernel times late in the run, once the beam has bunched)."""
    sim, per_species = make_sim(args, W, form)
    npart = sum(per_species)
    e0 = sim.energies()
    sim.step(args.warmup)
    sim.synchronize()
    tuned = sim.autotune(reps=2) if not args.no_autotune else None
    done = args.warmup
    # timed region: K graph-replayed steps (the production path)
    sim.synchronize()
    t0 = time.perf_counter()
    sim.step(args.steps)
    sim.synchronize()
    t1 = time.perf_counter()
    done += args.steps
    P = max(1, min(args.steps, 10))
    acc = profile(sim, P)
    done += P
    out = dict(npart=npart, per_species=per_species, wall=t1 - t0, kernel_ms=acc, profiled_steps=P, tuned=tuned)
    if args.run_steps > done:
        blocks = []
        while done < args.run_steps:
            k = min(50, args.run_steps - done)
            sim.synchronize()
            b0 = time.perf_counter()
            sim.step(k)
            sim.synchronize()
            blocks.append((done, k, time.perf_counter() - b0))
            done += k
        print("%s: trajectory to step %d done" % (form, done), file=sys.stderr, flush=True)
        out["kernel_ms_late"] = profile(sim, P)
        tot = sum(b[2] for b in blocks) + (t1 - t0)
        nst = sum(b[1] for b in blocks) + args.steps
        last, nlast = 0.0, 0
        for first, k, t in reversed(blocks):
            if nlast >= 50:
                break
            last += t
            nlast += k
        out["run"] = dict(first_step=args.warmup, last_step=done, steps=nst, ms_per_step=1e3 * tot / nst,
                          ms_per_step_last=1e3 * last / nlast, last_steps=nlast,
                          pushes_per_s=npart * nst / tot, pushes_per_s_last=npart * nlast / last)
    out["energies"] = sim.energies().tolist()
    out["energies0"] = e0.tolist()
    sim.close()
    return out


def energy_errors(ea, eb):
    """Energy differences of run a against run b over the sampled steps
    (rows: samples; columns: Uelm, Ukin of each species): the largest
    |dUelm| / Uelm(b), |dUkin| / Ukin(b) (each energy against itself, at the
    same step), and the same differences over the initial total energy."""
    ua, ub = ea[:, 0], eb[:, 0]
    ka, kb = ea[:, 1:].sum(1), eb[:, 1:].sum(1)
    tot0 = abs(eb[0].sum())
    live = np.abs(ub) > 0  # Uelm is zero at step 0 when the fields start at zero
    return {
        "field_vs_field": float((np.abs(ua - ub)[live] / np.abs(ub[live])).max()) if live.any() else 0.0,
        "kinetic_vs_kinetic": float((np.abs(ka - kb) / np.abs(kb)).max()),
        "field_vs_total": float(np.abs(ua - ub).max() / tot0),
        "kinetic_vs_total": float(np.abs(ka - kb).max() / tot0),
        "total_vs_total": float(np.abs(ea.sum(1) - eb.sum(1)).max() / tot0),
    }


def run_accuracy(args, W):
    """Three runs of the workload from the same seeds for args.accuracy
    steps, energies every 100 steps: the mixed form, the FP64 form, and the
    FP64 form again with another launch shape of the particle kernel (128
    threads per cluster, XCD-aware order: the same arithmetic summed in
    another order).  The third run measures the floor: how far two FP64 runs
    that differ only by the order of the atomic J sums drift apart.  Reports
    each energy's error against itself and against the initial total, and
    each run's Gauss-law drift max|(div E - rho)(t) - (div E - rho)(0)| /
    max|rho(0)|."""
    import smilei_amd as sa
    runs = [("mixed", "mixed", None), ("fp64", "fp64", None), ("fp64_alt", "fp64", {sa.TUNE_THREADS: 128,
                                                                                   sa.TUNE_XCD: 1})]
    hist = {}
    for label, form, tuning in runs:
        sim, per_species = make_sim(args, W, form)
        for k, v in (tuning or {}).items():
            sim.set_tuning(k, v)
        g0 = sim.gauss()
        rho0 = float(np.abs(sim.rho()).max())
        e = [sim.energies().tolist()]
        done = 0
        while done < args.accuracy:
            k = min(100, args.accuracy - done)
            sim.step(k)
            done += k
            e.append(sim.energies().tolist())
            print("accuracy %s: step %d, energies %s" % (label, done, e[-1]), file=sys.stderr, flush=True)
        drift = float(np.abs(sim.gauss() - g0).max() / rho0)
        hist[label] = dict(energies=np.array(e), gauss_drift=drift, npart=sum(per_species))
        sim.close()
    em, ef, ea = hist["mixed"]["energies"], hist["fp64"]["energies"], hist["fp64_alt"]["energies"]
    out = {
        "accuracy_steps": args.accuracy,
        "workload": W["desc"],
        "order": args.order,
        "particles": hist["mixed"]["npart"],
        "mixed_vs_fp64": energy_errors(em, ef),
        "fp64_floor": energy_errors(ea, ef),
        "gauss_drift_mixed": hist["mixed"]["gauss_drift"],
        "gauss_drift_fp64": hist["fp64"]["gauss_drift"],
        "gauss_drift_fp64_alt": hist["fp64_alt"]["gauss_drift"],
        "energy_change_mixed": float(em[-1].sum() / em[0].sum() - 1),
        "energy_change_fp64": float(ef[-1].sum() / ef[0].sum() - 1),
        "energies": {k: v["energies"].tolist() for k, v in hist.items()},
    }
    print(json.dumps(out))


def copy_probe(nbytes=4 << 30, reps=10):
    """Device-to-device copy of `nbytes` (hipMemcpyAsync through torch):
    the HBM bandwidth a plain stream reaches on this box, read + write bytes
    per second.  Also the known-size kernel that calibrates the FETCH_SIZE /
    WRITE_SIZE counters."""
    import torch
    src = torch.empty(nbytes // 4, dtype=torch.int32, device="cuda").fill_(1)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(reps):
        dst.copy_(src)
    t1.record()
    torch.cuda.synchronize()
    ms = t0.elapsed_time(t1) / reps
    del src, dst
    torch.cuda.empty_cache()
    return 2.0 * nbytes / (ms / 1e3) / 1e9


def run_cpu_baseline(args, W):
    """The oracle on a sub-box of the same plasma (same dx, dt, patch size,
    ppc and species; for the beam-plasma a 128x64x64-cell box with the beam
    on its axis), a few steps, all host cores."""
    from oracle import pyoracle as O
    from tests.plasma import thermal
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    O.set_threads(threads)
    shape = (64, 64, 64) if W["name"] == "thermal" else (128, 64, 64)
    S = O.OracleSim(3, args.order, shape, W["patch"], W["dx"], W["dt"])
    n = 0
    for k, (name, m, q, sig, ppc, dens, drift, radius) in enumerate(W["species"]):
        p = thermal(shape, W["dx"], ppc, sig, 100 + k, density=dens, drift_ux=drift, radius=radius)
        sid = S.add_species(m, q)
        S.set_particles(sid, p["x"], p["y"], p["z"], p["px"], p["py"], p["pz"], p["w"])
        n += len(p["x"])
    S.step(1)  # warm-up
    steps = 0
    t0 = time.perf_counter()
    while True:
        S.step(1)
        steps += 1
        if time.perf_counter() - t0 > args.cpu_seconds or steps >= 50:
            break
    dt = time.perf_counter() - t0
    rate = n * steps / dt
    return dict(value=rate, unit="particle-pushes/s", cores=threads, kind="port",
                sample="CPU oracle, %dx%dx%d-cell sub-box of the same plasma (%s patches, same species and ppc, "
                       "%d particles), %d steps, OpenMP over patches; %.1f ns*core/push"
                       % (shape + ("x".join(map(str, W["patch"])), n, steps, 1e9 * dt * threads / (n * steps))))


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="beam", choices=["beam", "thermal", "lwfa"])
    ap.add_argument("--ppc", type=int, default=8, help="background particles/cell of the beam-plasma workload")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--forms", default="mixed,fp64", help="comma list of mixed,fp64; the first is the headline")
    ap.add_argument("--copy-only", action="store_true", help="run the device copy probe only")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shape")
    ap.add_argument("--threads", type=int, default=0, help="particle-kernel workgroup size (with --no-autotune)")
    ap.add_argument("--layout", type=int, default=-1,
                    help="cluster layout of the sort: 0 cell-major, 1 lane-, 2 pair-interleaved (with --no-autotune)")
    ap.add_argument("--tile-x", type=int, default=0,
                    help="mixed form: clusters per workgroup along x, 1 or 2 (with --no-autotune)")
    ap.add_argument("--tile", type=int, default=-1,
                    help="mixed form: 1 = J-tile rows padded, 0 = unpadded (with --no-autotune)")
    ap.add_argument("--node-acc", type=int, default=-1,
                    help="mixed form, order 2: 1 = stayers deposited node by node, 0 = by particle "
                         "(with --no-autotune)")
    ap.add_argument("--order", type=int, default=2, choices=[2, 4], help="interpolation order")
    ap.add_argument("--accuracy", type=int, default=0,
                    help="instead of timing: run both forms N steps from the same seeds and print the mixed "
                         "form's energy errors and both forms' Gauss-law drift")
    ap.add_argument("--run-steps", type=int, default=-1,
                    help="follow the trajectory to this step after the timed steps and report its mean step time "
                         "and that of its last 50 steps (-1: the workload's configured length, 0: off)")
    ap.add_argument("--cell-order", default="any", choices=["any", "canonical"],
                    help="order of particles inside a cell after the sort: 'any' (left to the atomic scatter, "
                         "as the reference's GPU sort) or the canonical order")
    args = ap.parse_args()
    if args.copy_only:
        print(json.dumps({"copy_GBs": copy_probe(), "bytes_per_copy": 4 << 30}))
        return
    if args.gpus != 1:
        print("bench.py: one GPU only (the library does not shard the PIC step across GPUs); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    W = workload(args.workload, args.ppc)
    if args.run_steps < 0:
        args.run_steps = RUN_STEPS[W["name"]]
    if args.accuracy > 0:
        run_accuracy(args, W)
        return
    forms = [f for f in args.forms.split(",") if f]
    copy_gbs = copy_probe()
    res = {}
    for form in forms:
        g = run_gpu(args, W, form)
        npart, K = g["npart"], args.steps
        t_part = g["kernel_ms"]["particles"] / 1e3 / g["profiled_steps"]  # s per step, all species
        achieved = BYTES_PER_PUSH[form] * npart / t_part / 1e9
        e0, e1 = np.array(g["energies0"]), np.array(g["energies"])
        P = g["profiled_steps"]
        res[form] = dict(value=npart * K / g["wall"], ms_per_step=1e3 * g["wall"] / K,
                         kernel_ms_per_step={k: v / P for k, v in g["kernel_ms"].items()},
                         achieved=achieved, npart=npart, particles_per_species=g["per_species"],
                         autotuned=g["tuned"],
                         energy_change=float(e1.sum() / e0.sum() - 1), energies=g["energies"])
        if "run" in g:
            res[form]["run"] = g["run"]
            res[form]["kernel_ms_per_step_late"] = {k: v / P for k, v in g["kernel_ms_late"].items()}
    head = forms[0]
    h = res[head]
    # HBM bytes per particle-kernel launch (one launch per step runs every
    # species of the workload)
    tpp = TRAFFIC_PER_PUSH.get(W["name"], {}).get(head)
    traffic = tpp * h["npart"] if tpp else None
    # pushes per second of the particle kernel alone (its device time)
    kernel_rate = h["npart"] / (h["kernel_ms_per_step"]["particles"] / 1e3)
    vpp = VALU_INSTR_PER_PUSH.get(W["name"], {}).get(head)
    out = {
        "metric": "particle-pushes/s (gather+Boris+Esirkepov) 1xMI355X, mixed vs FP64; %HBM peak",
        "value": h["value"],
        "unit": "particle-pushes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": h["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if head == "mixed" else "f64",
        "data": "synthetic (seeded plasma initialised in HBM)",
        "config": {
            "workload": "%s (%d particles), order %d, Yee, periodic, atomic tile deposit, in-cell order %s"
                        % (W["desc"], h["npart"], args.order, args.cell_order),
            "form": head,
            "particles": h["npart"],
            "forms": {f: {k: v for k, v in r.items() if k != "energies"} for f, r in res.items()},
        },
        "roofline": {
            "bound": "hbm",
            "achieved": h["achieved"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": h["achieved"] / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_bytes_per_push": tpp,
            "achieved_traffic_GBs": tpp * kernel_rate / 1e9 if tpp else None,
            "measured_copy_GBs": copy_gbs,
            "kernel": "%s particle kernel (fused gather+Boris+key+Esirkepov), %d B/push algorithmic"
                      % (head, BYTES_PER_PUSH[head]),
            "frac_note": "frac counts algorithmic particle bytes only; the kernel is bound by VALU issue and LDS "
                         "(atomics, bank conflicts), not by HBM",
            "valu_instr_per_push": vpp,
            "valu_frac": vpp * kernel_rate / VALU_PEAK_WAVE_INSTR if vpp else None,
        },
    }
    if "mixed" in res and "fp64" in res:
        out["config"]["speedup_mixed_vs_fp64"] = res["mixed"]["value"] / res["fp64"]["value"]
        em, ef = np.array(res["mixed"]["energies"]), np.array(res["fp64"]["energies"])
        out["config"]["rel_energy_diff_mixed_vs_fp64"] = float(abs(em.sum() - ef.sum()) / abs(ef.sum()))
    if not args.no_cpu_baseline and W["species"]:
        out["cpu_baseline"] = run_cpu_baseline(args, W)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
