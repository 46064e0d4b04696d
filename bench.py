"""Benchmark: FDTD time steps of the Fullwave 2.5 engine on one GPU.

Default workload: the 3D heterogeneous inclusion configuration, 768 x 768 x 640 interior
(904 x 904 x 776 cells with the 24-cell absorbing layer, 36-cell transition layer and 8 ghost
cells per side), 20 seeded ellipsoidal inclusions, spatially varying gamma, 32 x 32 matrix-array
plane-wave transmit at 2.5 MHz, a 6000-step job whose full-aperture receive plane and three
orthogonal mid-planes (2.06 M sensors) are recorded every step. A "step" is one time step
(velocity update, pressure update, source/air/sensor pass) over the whole grid with every input
resident in HBM; ``--steps`` of the job are timed. ``--config focused_2d`` runs the 4096 x 2048
2D heterogeneous focused-transmit workload instead.

Prints one JSON line. ``value`` is GLUPS over the updated cells (the ghost border is never
updated). ``roofline`` is for the dominant kernel, timed with HIP events around each launch on
the engine's stream during a second pass over the same steps; its bytes are the algorithmic
bytes per updated cell (each array element the kernel must read or write, once).
``cpu_baseline`` times the FP32 build of the oracle (OpenMP) on a bounded sample of the same
grid (an x-slab of it in 3D). Other ``--config`` values print the batching, grouping, serving
and gradient figures.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
# bytes per lattice update of a whole isotropic time step in the step-level accounting the
# throughput targets are stated in (2D: 140, 3D: 180); GLUPS x this / peak = step roofline
STEP_BYTES_PER_LUP = {2: 140, 3: 180}
METRIC = "GLUPS + % HBM roofline (1 GPU); served jobs/hr with p50/p99 latency"


def algorithmic_bytes_per_cell(ndim: int, isotropic: bool, *, uniform: bool = False,
                               coded_fraction: float | None = None) -> dict:
    """Bytes each kernel must move per updated cell.

    Wave state and medium, float32 words: velocity reads p, rho, psi (2 per axis), v and writes
    v, psi; pressure reads v, K, p, phi (2 per axis) and writes p, phi. Relaxation coefficients:
    1/kappa, a1, a2, b1, b2 per map set (one set when isotropic) and the int32 dcmap (not read
    when every cell uses one coefficient set). With coded coefficients (``coded_fraction`` =
    share of cells carrying a code) every cell reads a 2-byte code instead, and the cells
    without a code also read the maps."""
    nd = ndim
    nax = 1 if isotropic else nd
    coef = 4 * (5 * nax + (0 if uniform else 1))
    if coded_fraction is not None:
        coef = 2 + (1.0 - coded_fraction) * coef
    vel = 4 * (2 + 6 * nd) + coef
    pre = 4 * (3 + 5 * nd) + coef
    return {"velocity": round(vel, 2), "pressure": round(pre, 2)}


def profiled_kernels(workload: str) -> str | None:
    """The kernel configuration (FULLWAVE_HIP_KERNELS syntax) that the committed counter summary
    of this workload measured, if there is one."""
    path = ROOT / "profiles" / f"pmc_traffic_{workload}.json"
    if not path.exists():
        return None
    from fullwave25_amd import engine

    return engine.kernel_override_string(json.loads(path.read_text())["kernel_config"])


def pin_kernels(args, workload: str, default_size: bool) -> str:
    """Kernel selection for a bench run: ``--kernels SHAPES`` pins them, ``--kernels tune``
    lets the engine's autotuner choose; by default the full-size workloads run the configuration
    their committed counter summary measured (the autotuner's choices among shapes within a few
    per cent of each other vary from run to run, and the traffic figure belongs to one
    configuration), and anything else is autotuned. Returns how the kernels were chosen."""
    if args.kernels and args.kernels != "tune":
        return "pinned"
    if not args.kernels and default_size:
        k = profiled_kernels(workload)
        if k:
            os.environ["FULLWAVE_HIP_KERNELS"] = k
            return "profiled"
    os.environ.pop("FULLWAVE_HIP_KERNELS", None)
    return "autotuned"


def coefficient_layout(sim, inp) -> dict:
    """How the handle reads relaxation coefficients, for the byte accounting."""
    n, frac = sim.relaxation_coding()
    ndmap = int(np.asarray(inp.dmap).shape[-1])
    return {"uniform": ndmap == 1, "coded_fraction": frac if n else None, "codes": n}


def pmc_summary(fetch_dir: Path, write_dir: Path, copy_fetch_dir: Path, copy_write_dir: Path,
                copy_bytes: int, config_file: Path) -> dict:
    """Per-kernel HBM traffic from two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE; KiB per
    dispatch), each scaled by what the same counter reads for the plain copy kernel moving
    ``copy_bytes`` each way. The kernels are the velocity and pressure kernels of the kernel
    configuration in ``config_file`` (written by ``--profile-steps``; the pmc runs pin it)."""
    import csv

    def per_kernel(d: Path) -> dict:
        rows = {}
        for r in csv.DictReader(open(d / "run_counter_collection.csv")):
            rows.setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]))
        return rows

    def copy_kib(d: Path) -> float:
        vals = [v for k, v in per_kernel(d).items() if "k_copy" in k]
        return float(np.mean(vals[0]))

    f_scale = copy_bytes / 1024.0 / copy_kib(copy_fetch_dir)
    w_scale = copy_bytes / 1024.0 / copy_kib(copy_write_dir)
    fetch, write = per_kernel(fetch_dir), per_kernel(write_dir)
    cfg = json.loads(Path(config_file).read_text())
    out = {"kernel_config": cfg["kernel_config"], "codes": cfg.get("codes", 0), "fetch_scale": round(f_scale, 4),
           "write_scale": round(w_scale, 4), "kernels": {}}
    for kind, tag in (("velocity", "k_velocity"), ("pressure", "k_pressure")):
        # the pinned kernel is the one dispatched every step; others ran only once or twice
        name = max((k for k in fetch if tag in k), key=lambda k: len(fetch[k]))
        f_kib, w_kib = float(np.mean(fetch[name])), float(np.mean(write[name]))
        out["kernels"][kind] = {
            "symbol": name, "dispatches": len(fetch[name]), "fetch_kib": round(f_kib, 1), "write_kib": round(w_kib, 1),
            "bytes_per_launch": int(round((f_kib * f_scale + w_kib * w_scale) * 1024)),
        }
    return out


def measured_traffic(workload: str, kind: str, kernel_config: dict, codes: int):
    """HBM bytes per launch of ``kind`` from the committed counter summary for this workload, if
    it was measured with the same kernel configuration and coefficient coding; else None."""
    path = ROOT / "profiles" / f"pmc_traffic_{workload}.json"
    if not path.exists():
        return None
    d = json.loads(path.read_text())
    if d.get("kernel_config", {}).get(kind) != kernel_config.get(kind) or bool(d.get("codes", 0)) != bool(codes):
        return None
    return d["kernels"][kind]["bytes_per_launch"]


def bench_batched(args) -> dict:
    """Many focused transmits on one 2048 x 1536 medium: every transmit as its own plain engine
    run (create, upload, tune, run, download), one after the other, against all of them as the
    shots of one batched handle. Also timed (on the first 8 transmits): the reference's loop,
    one engine run per transmit on a ``.dat`` directory whose maps are written once
    (``fw_run_dir``, the drop-in for its launcher)."""
    import dataclasses

    from fullwave25_amd import engine, workloads

    t0 = time.time()
    wl, shots = workloads.steered_transmits_2d(nx=args.nx, ny=args.ny, nt=args.steps, n_transmits=args.transmits)
    inp = wl.inputs
    print(f"# {wl.name} {inp.shape} nt={inp.nt}, {args.transmits} transmits, built in {time.time() - t0:.1f}s",
          file=sys.stderr, flush=True)
    n = shots.shape[0]
    with engine.HipSimulation(inp) as warm:  # first-use costs (library load, code objects)
        warm.run(min(10, inp.nt))
    t1 = time.perf_counter()
    seq = []
    for b in range(n):
        with engine.HipSimulation(dataclasses.replace(inp, icmat=shots[b])) as sim:
            sim.run(inp.nt)
            seq.append(sim.sensors())
        print(f"# sequential transmit {b + 1}/{n}", file=sys.stderr, flush=True)
    t_seq = time.perf_counter() - t1
    t2 = time.perf_counter()
    with engine.HipSimulation(inp, shots=shots) as sim:
        sim.run(inp.nt)
        bat = sim.sensors()
        kcfg = sim.kernel_config()
    t_bat = time.perf_counter() - t2
    # the reference's own loop over transmits: the .dat directory written once (static maps),
    # then per transmit a new icmat.dat and one run of the engine on the directory
    import tempfile

    n_file = min(n, 8)
    with tempfile.TemporaryDirectory() as tmp:
        sim_dir = Path(tmp) / "sim"
        inp.write_dir(sim_dir)
        t3 = time.perf_counter()
        for b in range(n_file):
            shots[b].astype(np.float32).tofile(sim_dir / "icmat.dat")
            engine.run_dir(sim_dir)
            np.fromfile(sim_dir / "genout.dat", dtype=np.float32)
        t_file = (time.perf_counter() - t3) / n_file * n
    worst = max(float(np.linalg.norm(bat[b] - seq[b]) / max(np.linalg.norm(seq[b]), 1e-30)) for b in range(n))
    cells = inp.ncells * inp.nt * n
    return {
        "metric": "batched transmits: jobs/hour vs sequential plain launches (1 GPU)",
        "value": round(n / t_bat * 3600, 1), "unit": "jobs/hour", "n_gpus": 1, "steps": inp.nt,
        "higher_is_better": True, "dtype": "f32", "data": "synthetic",
        "sequential_jobs_per_hour": round(n / t_seq * 3600, 1), "speedup_vs_sequential": round(t_seq / t_bat, 3),
        "file_contract_jobs_per_hour": round(n / t_file * 3600, 1),
        "speedup_vs_file_contract": round(t_file / t_bat, 3),
        "glups_batched": round(cells / t_bat / 1e9, 3), "glups_sequential": round(cells / t_seq / 1e9, 3),
        "max_rel_l2_batched_vs_sequential": worst, "kernel_config": kcfg,
        "config": {"workload": f"{wl.name}: {args.nx}x{args.ny} interior ({inp.nx}x{inp.ny} extended), "
                               f"{n} steered focused transmits, {inp.nt} steps each",
                   "transmits": n, "cells": inp.ncells},
    }


def bench_grouped(args) -> dict:
    """Jobs on uneven media in shared launches: ``--jobs`` 2D jobs on 1024^2, 1536x1024, 2048^2
    and 3072x2048 grids (one seeded heterogeneous medium per size) with 3000..6000 steps, run
    one after the other as plain engine runs (create, upload, tune, run, download) against all
    of them in one group of handles (fw_group_run: one velocity, one pressure and one
    source/sensor launch per step for every job still running; jobs leave the group as they
    finish)."""
    import dataclasses

    from fullwave25_amd import engine, workloads

    t0 = time.time()
    rng = np.random.default_rng(2028)
    small = args.mix == "small"
    sizes = [(256, 256), (384, 256), (512, 512), (768, 512)] if small else \
        [(1024, 1024), (1536, 1024), (2048, 2048), (3072, 2048)]
    unit = 250 if small else 1000  # jobs run 3..6 units of steps
    media = {sz: workloads.focused_linear_array_2d(nx=sz[0], ny=sz[1], nt=6 * unit,
                                                   n_elements=sz[1] // 12 if small else 128).inputs for sz in sizes}
    jobs = []
    for _ in range(args.jobs):
        sz = sizes[int(rng.integers(len(sizes)))]
        jobs.append(dataclasses.replace(media[sz], nt=int(rng.integers(3, 7)) * unit))
    print(f"# grouped: {len(jobs)} jobs built in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    with engine.HipSimulation(jobs[0]) as warm:  # first-use costs (library load, code objects)
        warm.run(10)
    t1 = time.perf_counter()
    seq = []
    for inp in jobs:
        with engine.HipSimulation(inp) as sim:
            sim.run(inp.nt)
            seq.append(sim.sensors())
    t_seq = time.perf_counter() - t1
    print(f"# sequential: {t_seq:.1f}s", file=sys.stderr, flush=True)
    t2 = time.perf_counter()
    sims = [engine.HipSimulation(inp) for inp in jobs]
    done = 0
    got = [None] * len(jobs)
    live = list(range(len(jobs)))
    # kernel family and shape: timed on the first (fresh) group where left open, then kept
    variant, width, chunk = args.group_variant, args.group_width, args.group_chunk
    while live:
        end = min(jobs[i].nt for i in live)
        with engine.HipGroup([sims[i] for i in live], width=width, chunk=chunk, variant=variant) as grp:
            shape = grp.kernel_config()["velocity"]
            variant, width, chunk = shape["variant"], shape["width"], shape["chunk"]
            if args.graph_steps:
                grp.set_graph_steps(args.graph_steps)
            grp.run(end - done)
        done = end
        for i in [i for i in live if jobs[i].nt == end]:
            got[i] = sims[i].sensors()
            sims[i].close()
        live = [i for i in live if jobs[i].nt > end]
    t_grp = time.perf_counter() - t2
    worst = max(float(np.linalg.norm(g - q) / max(np.linalg.norm(q), 1e-30)) for g, q in zip(got, seq))
    cells = sum(inp.ncells * inp.nt for inp in jobs)
    return {
        "metric": "grouped uneven jobs: jobs/hour vs sequential plain launches (1 GPU)",
        "value": round(len(jobs) / t_grp * 3600, 1), "unit": "jobs/hour", "n_gpus": 1,
        "higher_is_better": True, "dtype": "f32", "data": "synthetic",
        "sequential_jobs_per_hour": round(len(jobs) / t_seq * 3600, 1),
        "speedup_vs_sequential": round(t_seq / t_grp, 3),
        "glups_grouped": round(cells / t_grp / 1e9, 3), "glups_sequential": round(cells / t_seq / 1e9, 3),
        "max_rel_l2_grouped_vs_sequential": worst,
        "config": {"workload": f"{len(jobs)} focused-transmit jobs on "
                               + ", ".join(f"{a}x{b}" for a, b in sizes)
                               + f" interiors, {3 * unit}-{6 * unit} steps; group kernel "
                               + (f"tile:{chunk}" if variant == "tile" else f"stream:{width}:{chunk}")
                               + (f", steps replayed from {args.graph_steps}-step graphs" if args.graph_steps else ""),
                   "mix": args.mix, "jobs": len(jobs), "graph_steps": args.graph_steps},
    }


def bench_batched_mix(args) -> dict:
    """The batched-transmit configuration as one workload: 128 steered focused transmits on one
    2048 x 1536 medium (3000 steps each) and 32 jobs on uneven grids (1024^2, 1536x1024, 2048^2,
    3072x2048; 3000..6000 steps), all submitted at once to the simulation server (the transmits
    as batched jobs of ``--shots-per-job`` shots on one handle each, the uneven jobs as jobs of
    their own, ``--slots`` running at once), against every transmit and job run one after the
    other as a plain engine run (create, upload, tune, run, download). Jobs/hour counts each
    transmit and each uneven job as one job; the latency percentiles are over the submitted jobs."""
    import dataclasses

    from fullwave25_amd import engine, workloads
    from fullwave25_amd.serving import SimulationServer

    t0 = time.time()
    wl, shots = workloads.steered_transmits_2d(nt=3000, n_transmits=args.transmits)
    inp = wl.inputs
    rng = np.random.default_rng(2031)
    sizes = [(1024, 1024), (1536, 1024), (2048, 2048), (3072, 2048)]
    media = {sz: workloads.focused_linear_array_2d(nx=sz[0], ny=sz[1], nt=6000, seed=2032 + i).inputs
             for i, sz in enumerate(sizes)}
    uneven = [dataclasses.replace(media[sizes[int(rng.integers(len(sizes)))]], nt=int(rng.integers(3, 7)) * 1000)
              for _ in range(args.jobs)]
    n_jobs = shots.shape[0] + len(uneven)
    print(f"# batched mix: {shots.shape[0]} transmits on {inp.shape} + {len(uneven)} uneven jobs, built in "
          f"{time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    with engine.HipSimulation(inp) as warm:  # first-use costs (library load, code objects)
        warm.run(10)
    t1 = time.perf_counter()
    for b in range(shots.shape[0]):
        with engine.HipSimulation(dataclasses.replace(inp, icmat=shots[b])) as sim:
            sim.run(inp.nt)
            sim.sensors()
        if (b + 1) % 16 == 0:
            print(f"# sequential transmit {b + 1}/{shots.shape[0]}", file=sys.stderr, flush=True)
    t_tx = time.perf_counter() - t1
    for q, job in enumerate(uneven):
        with engine.HipSimulation(job) as sim:
            sim.run(job.nt)
            sim.sensors()
        if (q + 1) % 8 == 0:
            print(f"# sequential uneven job {q + 1}/{len(uneven)}", file=sys.stderr, flush=True)
    t_seq = time.perf_counter() - t1
    print(f"# sequential: {t_seq:.1f}s ({t_tx:.1f}s of transmits)", file=sys.stderr, flush=True)
    k = max(1, int(args.shots_per_job))
    t2 = time.perf_counter()
    with SimulationServer(max_slots=args.slots) as srv:
        futs = [srv.submit(inp, shots=shots[b:b + k]) for b in range(0, shots.shape[0], k)]
        futs += [srv.submit(job) for job in uneven]
        for f in futs:
            f.result()
    t_srv = time.perf_counter() - t2
    st = srv.stats()
    return {
        "metric": "batched transmits + uneven jobs: jobs/hour vs sequential plain launches (1 GPU)",
        "value": round(n_jobs / t_srv * 3600, 1), "unit": "jobs/hour", "n_gpus": 1, "higher_is_better": True,
        "dtype": "f32", "data": "synthetic",
        "sequential_jobs_per_hour": round(n_jobs / t_seq * 3600, 1), "speedup_vs_sequential": round(t_seq / t_srv, 3),
        "wall_s": round(t_srv, 2), "sequential_wall_s": round(t_seq, 2),
        "latency_p50_s": round(st.latency_p50_s, 3), "latency_p99_s": round(st.latency_p99_s, 3),
        "submitted_jobs": len(futs), "peak_concurrent_jobs": srv.peak_running,
        "config": {"workload": f"{shots.shape[0]} steered focused transmits on {inp.nx}x{inp.ny} cells (2048x1536 "
                               f"interior, 3000 steps, {k} shots per batched job) + {len(uneven)} focused-transmit "
                               f"jobs on 1024^2..3072x2048 interiors with 3000-6000 steps, submitted together",
                   "transmits": int(shots.shape[0]), "uneven_jobs": len(uneven), "slots": args.slots},
    }


def bench_adjoint(args) -> dict:
    """The misfit-gradient job of the served mix on its own: 1024^2 interior, 2000 steps, the L2
    misfit of a smoothed initial model against the traces of a seeded heterogeneous true medium;
    the time of the forward run (which writes the tape and the checkpoints) and of the reverse
    sweep, against a plain forward run of the same job, each from a handle that is set up (the
    handles' creation, uploads and kernel tuning are outside all three times). ``--checkpoint-every 0`` (default) lets
    the engine choose the segment length (the whole run taped when it fits in half of the free
    device memory, so nothing is recomputed); the same job with 200-step segments (each segment
    but the last recomputed from its checkpoint) is timed beside it."""
    from fullwave25_amd import adjoint, engine, workloads

    initial, true = workloads.misfit_media_2d(nx=args.nx, ny=args.ny, nt=args.steps)
    inp = initial.inputs
    with engine.HipSimulation(true.inputs) as sim:  # the observed data; also first-use costs
        sim.run(inp.nt)
        observed = sim.sensors().astype(np.float64)
    with engine.HipSimulation(inp) as sim:
        sim.run(10)
    # the plain run timed as the gradient's legs are: from a handle whose set-up (uploads, codes,
    # kernel tuning) is done, to the downloaded traces; with the set-up beside it
    t0 = time.perf_counter()
    with engine.HipSimulation(inp) as sim:
        sim.kernel_config()
        t1 = time.perf_counter()
        sim.run(inp.nt)
        sim.sensors()
        t_plain = time.perf_counter() - t1
    t_plain_setup = time.perf_counter() - t0

    def gradient(every: int):
        with adjoint.HipAdjoint(inp, checkpoint_every=every) as adj:
            t1 = time.perf_counter()
            tr = adj.forward()
            t2 = time.perf_counter()
            gk, gr = adj.backward((tr - observed).astype(np.float32))
            t3 = time.perf_counter()
            return adj.checkpoint_every, t2 - t1, t3 - t2, gk, gr, adj.sim.kernel_config()

    every, t_fwd, t_rev, gk, gr, kcfg = gradient(args.checkpoint_every)
    seg = min(200, inp.nt)
    alt = gradient(seg) if every != seg else None
    tape_gb = every * (inp.ndim + 1) * inp.ncells * 4 / 1e9
    cells = inp.ncells * inp.nt
    out = {
        "metric": "misfit gradient (dJ/dK, dJ/drho): time vs a plain forward run (1 GPU)",
        "value": round((t_fwd + t_rev) / t_plain, 3), "unit": "forward-run equivalents", "n_gpus": 1,
        "higher_is_better": False, "dtype": "f32", "data": "synthetic",
        "plain_forward_s": round(t_plain, 4), "plain_forward_with_setup_s": round(t_plain_setup, 4),
        "forward_with_tape_s": round(t_fwd, 4),
        "reverse_sweep_s": round(t_rev, 4), "gradient_jobs_per_hour": round(3600 / (t_fwd + t_rev), 1),
        "checkpoint_every": every, "tape_gb": round(tape_gb, 2), "kernel_config": kcfg,
        "forward_glups": round(cells / t_plain / 1e9, 3), "finite": bool(np.isfinite(gk).all() and np.isfinite(gr).all()),
        "config": {"workload": f"misfit_media_2d: {args.nx}x{args.ny} interior ({inp.nx}x{inp.ny} extended), "
                               f"smoothed initial model of a seeded layered medium ({inp.ndmap} sound speeds), "
                               f"{inp.nt} steps, segments of {every} steps", "cells": inp.ncells},
    }
    if alt is not None:
        out[f"segments_of_{seg}"] = {
            "value": round((alt[1] + alt[2]) / t_plain, 3), "forward_with_tape_s": round(alt[1], 4),
            "reverse_sweep_s": round(alt[2], 4),
            "max_rel_diff_grad_K": float(np.abs(alt[3] - gk).max() / max(np.abs(gk).max(), 1e-30)),
        }
    return out


def _served_jobs(args):
    """The served mix, seeded. ``--mix standard``: 2D forward jobs on 512^2..4096^2 grids with
    1000..8000 steps (70 %), misfit-gradient jobs on 1024^2 (20 %), small 3D forward jobs
    (10 %). ``--mix small``: the same shares on 256^2..1024^2 grids with 500..4000 steps, 512^2
    misfit gradients and 96^3 3D jobs: requests none of which fills the GPU alone."""
    from fullwave25_amd import workloads

    import dataclasses

    rng = np.random.default_rng(7)
    small = args.mix == "small"
    sizes = (256, 512, 1024) if small else (512, 1024, 2048, 4096)
    adj_n, adj_steps = (512, 1000) if small else (1024, 2000)
    media = {}

    def medium(n, steps):
        # one medium per grid size, built for the longest run; a job runs its own step count
        if n not in media:
            media[n] = workloads.plane_wave_2d(nx=n, ny=n, nt=8000, sensor_depth=n // 2).inputs
        return dataclasses.replace(media[n], nt=steps)

    jobs = []
    adj_inp = None
    for _ in range(args.jobs):
        u = rng.random()
        steps = int(rng.integers(2, 17)) * (250 if small else 500)
        if u < 0.7:
            jobs.append(("forward", medium(int(rng.choice(sizes)), steps)))
        elif u < 0.9:
            adj_inp = adj_inp if adj_inp is not None else medium(adj_n, adj_steps)
            jobs.append(("adjoint", adj_inp))
        else:
            jobs.append(("forward3d", None))
    return jobs, media


def bench_served(args) -> dict:
    """Poisson arrivals of the served mix through the simulation server, against the same jobs
    run one at a time on plain engine handles."""
    from fullwave25_amd import adjoint, engine, workloads
    from fullwave25_amd.serving import SimulationServer, adjoint_segment

    t0 = time.time()
    jobs, _ = _served_jobs(args)
    # gradient segments as the server picks them for its default budget (90 % of the free memory)
    share = int(0.9 * engine.device_memory(0)[0]) // 4
    n3 = 96 if args.mix == "small" else args.nz
    w3 = workloads.plane_wave_3d(n3, n3, n3, nt=1000) if any(k == "forward3d" for k, _ in jobs) else None
    jobs = [(k, w3.inputs if k == "forward3d" else inp) for k, inp in jobs]
    observed = {}
    for k, inp in jobs:
        if k == "adjoint" and id(inp) not in observed:
            # "observed" data: the same setup in a 2 % faster medium
            true = adjoint.with_medium(inp, 1.02 * np.sqrt(inp.maps["K"] / inp.maps["rho"]), inp.maps["rho"])
            with engine.HipSimulation(true) as sim:
                sim.run(true.nt)
                observed[id(inp)] = sim.sensors().astype(np.float64)
    print(f"# served mix: {len(jobs)} jobs built in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)

    def run_plain(kind, inp):
        if kind == "adjoint":  # the server's segment length (SimulationServer.submit_adjoint)
            with adjoint.HipAdjoint(inp, checkpoint_every=adjoint_segment(inp, share)) as adj:
                tr = adj.forward()
                adj.backward((tr - observed[id(inp)]).astype(np.float32))
            return
        with engine.HipSimulation(inp) as sim:
            sim.run(inp.nt)
            sim.sensors()

    t1 = time.perf_counter()
    for k, inp in jobs:
        run_plain(k, inp)
    t_seq = time.perf_counter() - t1
    print(f"# sequential: {t_seq:.1f}s", file=sys.stderr, flush=True)

    rng = np.random.default_rng(8)
    rate = args.arrival_factor * len(jobs) / t_seq  # arrivals per second
    gaps = rng.exponential(1.0 / rate, len(jobs))
    with SimulationServer(max_slots=args.slots) as srv:
        futs = []
        t_next = time.perf_counter()
        for (k, inp), gap in zip(jobs, gaps):
            t_next += gap
            while (wait := t_next - time.perf_counter()) > 0:
                time.sleep(min(0.002, wait))
            if k == "adjoint":
                futs.append(srv.submit_adjoint(inp, observed[id(inp)]))
            else:
                futs.append(srv.submit(inp))
        for f in futs:
            f.result()
    st = srv.stats()
    return {
        "metric": "served jobs/hour with p50/p99 latency (1 GPU)", "value": round(st.jobs_per_hour, 1),
        "unit": "jobs/hour", "n_gpus": 1, "higher_is_better": True, "dtype": "f32", "data": "synthetic",
        "latency_p50_s": round(st.latency_p50_s, 3), "latency_p99_s": round(st.latency_p99_s, 3),
        "queue_p50_s": round(st.queue_p50_s, 3), "sequential_jobs_per_hour": round(len(jobs) / t_seq * 3600, 1),
        "speedup_vs_sequential": round(st.jobs_per_hour / (len(jobs) / t_seq * 3600), 3),
        "served_glups": round(st.glups, 3), "peak_concurrent_jobs": srv.peak_running,
        "config": {"workload": (f"{len(jobs)} jobs: 2D forward 256^2-1024^2 x 500-4000 steps, 512^2 misfit "
                                f"gradients (1000 steps), 96^3 3D forward (1000 steps)" if args.mix == "small" else
                                f"{len(jobs)} jobs: 2D forward 512^2-4096^2 x 1000-8000 steps, 1024^2 misfit "
                                f"gradients (2000 steps, whole run taped when it fits a quarter of the budget), "
                                f"{args.nz}^3 3D forward (1000 steps)")
                               + f"; Poisson arrivals at {args.arrival_factor} x the sequential service rate",
                   "mix": args.mix, "slots": args.slots, "jobs": len(jobs)},
    }




def x_slab(inp, n_planes: int, x0: int | None = None):
    """Planes ``[x0, x0 + n_planes)`` of ``inp`` as inputs of their own: every map sliced along x
    (the slowest axis, so each slice is contiguous), the source, sensor and air points inside the
    slab kept and shifted. By default the slab is centred on the first source plane, so it holds
    the transmit and the receive plane. Used to time the CPU oracle on a bounded sample of a grid
    too large for it to step in seconds."""
    import dataclasses

    n = min(int(n_planes), inp.nx)
    if x0 is None:
        first = int(inp.icc[:, 0].min()) if inp.ncoords else inp.nx // 2
        x0 = first - n // 2
    x0 = max(0, min(int(x0), inp.nx - n))
    sl = slice(x0, x0 + n)

    def keep(coords):
        c = np.asarray(coords).reshape(-1, inp.ndim)
        m = (c[:, 0] >= x0) & (c[:, 0] < x0 + n)
        out = c[m].astype(np.int32)
        out[:, 0] -= x0
        return out, m

    icc, m = keep(inp.icc)
    outc, _ = keep(inp.outc)
    zero = keep(inp.icczero)[0] if inp.ncoordszero else np.zeros((0, inp.ndim), np.int32)
    return dataclasses.replace(
        inp, nx=n, maps={k: np.ascontiguousarray(v[sl]) for k, v in inp.maps.items()},
        dcmap=np.ascontiguousarray(inp.dcmap[sl]), icc=icc, icmat=np.ascontiguousarray(inp.icmat[m]),
        outc=outc, icczero=zero)


def updated_cells(inp) -> int:
    """Cells the engine updates: every coordinate in [8, n - 8) (the 8-cell ghost border stays 0)."""
    return int(np.prod([s - 16 for s in inp.shape]))


def cpu_baseline(inp, steps: int, *, slab_planes: int = 0) -> dict:
    """The FP32 oracle (C, OpenMP) timed on this host's cores over ``steps`` steps after 5
    warm-up steps, on ``inp`` or, with ``slab_planes``, on an x-slab of it. The steps are clamped
    to what the inputs' ``nt`` leaves after the warm-up. GLUPS over updated cells."""
    import oracle

    oracle.build()
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    model = ""
    try:
        model = next(l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("model name"))
    except (OSError, StopIteration):
        pass
    sample = inp if not slab_planes else x_slab(inp, slab_planes)
    warm = min(5, max(sample.nt - 1, 0))
    steps = max(1, min(int(steps), sample.nt - warm))
    sim = oracle.OracleSimulation(sample, precision=32)
    try:
        sim.run(warm)  # first touch of the state arrays
        t0 = time.perf_counter()
        sim.run(steps)
        dt = time.perf_counter() - t0
    finally:
        sim.close()
    where = (f"an x-slab of {sample.nx} planes ({sample.nx}x{sample.ny}" + (f"x{sample.nz}" if sample.ndim == 3 else "")
             + f" cells, {sample.ncoords} sources, {sample.ncoordsout} sensors) of the same grid") if slab_planes \
        else "the same workload"
    return {
        "value": round(updated_cells(sample) * steps / dt / 1e9, 4),
        "unit": "GLUPS",
        "cores": threads,
        "kind": "port",
        "cpu_model": model,
        "seconds": round(dt, 2),
        "sample": f"FP32 oracle (C, OpenMP, {threads} threads) on {where}: steps {warm + 1}..{warm + steps} "
                  f"after {warm} warm-up steps; GLUPS over updated cells",
    }


# Headline single-grid workloads and their full sizes and step counts.
SINGLE_CONFIGS = {
    "hetero_3d": {"workload": "inclusions_3d", "nx": 768, "ny": 768, "nz": 640, "nt": 6000,
                  "desc": "3D heterogeneous, ellipsoidal inclusions, 32x32 matrix-array plane wave, 2.5 MHz"},
    "focused_2d": {"workload": "focused_linear_array_2d", "nx": 4096, "ny": 2048, "nz": 1, "nt": 8000,
                   "desc": "2D heterogeneous six-layer medium with scatterers, 128-element array focused at 40 mm, 3 MHz"},
    "plane_wave_2d": {"workload": "plane_wave_2d", "nx": 512, "ny": 512, "nz": 1, "nt": 2000,
                      "desc": "2D homogeneous plane wave, 3 MHz"},
    "random_3d": {"workload": "random_medium_3d", "nx": 256, "ny": 256, "nz": 256, "nt": 1000,
                  "desc": "3D seeded random smooth medium"},
}
# the CPU sample for grids the oracle cannot step in seconds: an x-slab of this many planes
CPU_SLAB_PLANES = {"hetero_3d": 40}
# CPU baseline steps (about 10-30 s of oracle time on the GPU host's cores)
CPU_STEPS = {"hetero_3d": 30, "random_3d": 30}


def build_single(args):
    """The workload of a single-grid config at ``args``' sizes (the config's own by default) with
    the config's full step count (at least ``--warmup`` and ``--steps``)."""
    from fullwave25_amd import workloads

    cfg = SINGLE_CONFIGS[args.config]
    nt = max(cfg["nt"], args.steps, args.warmup)
    if args.config == "hetero_3d":
        return workloads.inclusions_3d(args.nx, args.ny, args.nz, nt=nt)
    if args.config == "focused_2d":
        return workloads.focused_linear_array_2d(nx=args.nx, ny=args.ny, nt=nt)
    if args.config == "plane_wave_2d":
        return workloads.plane_wave_2d(nx=args.nx, ny=args.ny, nt=nt)
    return workloads.random_medium_3d(args.nx, args.ny, args.nz, nt, seed=5)


def bench_single(args) -> dict:
    """One grid, one transmit: ``--warmup`` untimed steps, reset, then exactly ``--steps`` timed
    steps (fw_sim_run synchronizes the engine's stream before it returns, so the wall clock
    brackets the steps); a second pass over the same steps with HIP events around every launch
    gives the dominant kernel's roofline; a third with the relaxation coefficients read from
    the per-cell maps instead of codes gives the uncoded rate."""
    from fullwave25_amd import engine

    cfg = SINGLE_CONFIGS[args.config]
    t0 = time.time()
    wl = build_single(args)
    inp = wl.inputs
    print(f"# workload {wl.name} {inp.shape} nt={inp.nt}, {inp.ncoordsout} sensors, built in {time.time() - t0:.1f}s",
          file=sys.stderr, flush=True)
    full_size = (args.nx, args.ny, args.nz if inp.ndim == 3 else 1) == (cfg["nx"], cfg["ny"], cfg["nz"])
    source = pin_kernels(args, wl.name, full_size)
    upd = updated_cells(inp)
    if inp.ndim == 3:
        # The first large device allocation of a process maps memory that the 3D kernels
        # stream more slowly than memory the process has allocated and released before
        # (904x904x776 grid: pressure kernel 10.9 against 10.1 ms, velocity 11.3 against
        # 11.1 ms; an untouched allocate-and-release of 100 GB before the first handle gives
        # the same gain). A server is in the second state from its second job on. The set-up
        # therefore creates and releases one handle of the workload (its allocations, no
        # steps) before the measured one.
        engine.HipSimulation(inp, device=0).close()
    with engine.HipSimulation(inp, device=0) as sim:
        kcfg = sim.kernel_config()
        layout = coefficient_layout(sim, inp)
        print(f"# handle ready: kernels {engine.kernel_override_string(kcfg)}, {layout['codes']} codes", file=sys.stderr,
              flush=True)
        if args.profile_steps:
            sim.run(args.profile_steps)
            return {"kernel_config": kcfg, "kernels": engine.kernel_override_string(kcfg), "codes": layout["codes"]}
        sim.run(args.warmup)
        sim.reset()
        t1 = time.perf_counter()
        sim.run(args.steps)
        elapsed = time.perf_counter() - t1
        finite = bool(np.isfinite(sim.sensors()).all())
        print(f"# timed {args.steps} steps: {elapsed * 1e3 / args.steps:.3f} ms/step", file=sys.stderr, flush=True)
        sim.reset()
        kms = sim.run_timed(args.steps)
        uncoded = None
        if layout["codes"]:
            # the same kernels reading the ten relaxation maps and dcmap per cell
            sim.set_relaxation_codes(False)
            sim.reset()
            sim.run(min(args.warmup, 5))
            sim.reset()
            t2 = time.perf_counter()
            sim.run(args.steps)
            t_unc = time.perf_counter() - t2
            uncoded = {"glups": round(upd * args.steps / t_unc / 1e9, 3), "ms_per_step": round(t_unc / args.steps * 1e3, 4),
                       "step_frac": round(upd * args.steps / t_unc * STEP_BYTES_PER_LUP[inp.ndim] / 1e9 / HBM_PEAK_GBS, 4)}
            print(f"# uncoded: {t_unc * 1e3 / args.steps:.3f} ms/step", file=sys.stderr, flush=True)
    copy_gbs = engine.copy_bandwidth(1 << 30, repeats=10)
    glups = upd * args.steps / elapsed / 1e9
    bpc = algorithmic_bytes_per_cell(inp.ndim, inp.isotropic, uniform=layout["uniform"],
                                     coded_fraction=layout["coded_fraction"])
    bpc_unc = algorithmic_bytes_per_cell(inp.ndim, inp.isotropic, uniform=layout["uniform"])
    dominant = max(("velocity", "pressure"), key=lambda k: kms[k])
    avg_s = kms[dominant] / 1e3 / args.steps
    achieved = bpc[dominant] * upd / avg_s / 1e9
    # the dominant kernel's per-cell streams without its stencils, on arrays of the grid's size:
    # velocity reads p, rho, the code, 2 ndim memory variables and ndim velocities and writes the
    # last two; pressure reads ndim velocities, K, the code, 2 ndim memory variables and p and
    # writes p and the memory variables
    nd = inp.ndim
    streams = (3 + 3 * nd, 3 * nd) if dominant == "velocity" else (3 + 3 * nd, 1 + 2 * nd)
    stream_gbs = engine.stream_bandwidth(inp.ncells, *streams, repeats=5)
    dims = f"{args.nx}x{args.ny}" + (f"x{args.nz}" if inp.ndim == 3 else "")
    ext = "x".join(str(s) for s in inp.shape)
    return {
        "metric": METRIC,
        "value": round(glups, 3),
        "unit": "GLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{wl.name}: {cfg['desc']}; {dims} interior ({ext} extended), {args.steps} timed steps "
                        f"of the {inp.nt}-step job, isotropic relaxation, {inp.ncoords} source points, "
                        f"{inp.ncoordsout} sensors recorded every {inp.modt} step(s); GLUPS over the {upd} updated "
                        f"cells (the 8-cell ghost border is not updated)",
            "cells_extended": inp.ncells,
            "cells_updated": upd,
            "sources": inp.ncoords,
            "sensors": inp.ncoordsout,
            "record_every": inp.modt,
            "nt": inp.nt,
            "finite_output": finite,
        },
        "glups_extended_grid": round(inp.ncells * args.steps / elapsed / 1e9, 3),
        "uncoded": uncoded,
        "roofline": {
            "bound": "hbm",
            "kernel": dominant,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(wl.name, dominant, kcfg, layout["codes"]),
            "bytes_per_cell": bpc[dominant],
            "bytes_per_cell_uncoded": bpc_unc[dominant],
            "updated_cells_per_launch": upd,
            "avg_launch_us": round(avg_s * 1e6, 2),
            "kernel_ms_per_step": {k: round(v / args.steps, 5) for k, v in kms.items()},
            "copy_kernel_gbs": round(copy_gbs, 1),
            "frac_of_copy": round(achieved / copy_gbs, 4),
            "stream_probe": {"reads": streams[0], "writes": streams[1], "cells": inp.ncells, "gbs": round(stream_gbs, 1)},
            "frac_of_stream_probe": round(achieved / stream_gbs, 4),
            "kernel_config": kcfg,
            "kernels": engine.kernel_override_string(kcfg),
            "kernel_source": source,
            "relaxation_codes": layout["codes"],
            "coded_fraction": layout["coded_fraction"],
            "step_bytes_per_lup": STEP_BYTES_PER_LUP[inp.ndim],
            "step_frac": round(glups * STEP_BYTES_PER_LUP[inp.ndim] / HBM_PEAK_GBS, 4),
        },
        "_inputs": inp,
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", default="hetero_3d",
                    choices=(*SINGLE_CONFIGS, "batched_2d", "batched_mix", "served_mix", "grouped_2d", "adjoint_2d",
                             "inclusions_3d"))
    ap.add_argument("--nx", type=int, default=None)
    ap.add_argument("--ny", type=int, default=None)
    ap.add_argument("--nz", type=int, default=None)
    ap.add_argument("--cpu-steps", type=int, default=0,
                    help="steps of the CPU baseline (default: 30 on the 3D slab, 150 on 2D grids)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=0,
                    help="run only this many steps and exit (for rocprofv3)")
    ap.add_argument("--sweep", type=int, default=0,
                    help="time every kernel family/shape over this many steps and exit")
    ap.add_argument("--sweep-shapes", default="", help="comma-separated kernel shapes for --sweep")
    ap.add_argument("--stream-probe", default="",
                    help="R,W,CELLS: only run the streaming probe (R read and W write arrays of CELLS floats)")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run the plain copy kernel over this many bytes (counter calibration)")
    ap.add_argument("--kernels", default="",
                    help="pin the kernel configuration (FULLWAVE_HIP_KERNELS syntax), or 'tune' to autotune; "
                         "default: the profiled configuration for the full-size workloads, else autotune")
    ap.add_argument("--pmc-summary", nargs=5, metavar=("FETCH_DIR", "WRITE_DIR", "COPY_FETCH_DIR", "COPY_WRITE_DIR",
                                                       "CONFIG_JSON"),
                    help="summarize rocprofv3 --pmc runs into per-launch HBM traffic (JSON) and exit")
    ap.add_argument("--copy-bytes", type=int, default=268435456)
    ap.add_argument("--transmits", type=int, default=32,
                    help="batched_2d, batched_mix: transmits on the shared medium (batched_mix: 128)")
    ap.add_argument("--shots-per-job", type=int, default=32, help="batched_mix: transmits per batched job")
    ap.add_argument("--jobs", type=int, default=20, help="served_mix, grouped_2d: number of jobs")
    ap.add_argument("--checkpoint-every", type=int, default=0,
                    help="adjoint_2d: segment length (0: the engine's choice)")
    ap.add_argument("--group-width", type=int, default=0, help="grouped_2d: lanes per workgroup (0: timed)")
    ap.add_argument("--group-chunk", type=int, default=0, help="grouped_2d: rows per workgroup (0: timed)")
    ap.add_argument("--group-variant", choices=("auto", "stream", "tile"), default="auto",
                    help="grouped_2d: kernel family of the group (auto: timed)")
    ap.add_argument("--graph-steps", type=int, default=0,
                    help="grouped_2d: replay the group's steps from a captured graph of this many steps")
    ap.add_argument("--slots", type=int, default=4, help="served_mix: concurrent jobs")
    ap.add_argument("--mix", choices=("standard", "small"), default="standard",
                    help="served_mix, grouped_2d: job sizes (small: none fills the GPU alone)")
    ap.add_argument("--arrival-factor", type=float, default=2.0,
                    help="served_mix: arrival rate as a multiple of the sequential service rate")
    args = ap.parse_args()
    if args.config == "inclusions_3d":  # the workload's own name
        args.config = "hetero_3d"
    # grid sizes: the config's own unless given
    own = {"batched_2d": (2048, 1536, 1), "batched_mix": (2048, 1536, 1), "grouped_2d": (2048, 2048, 1),
           "adjoint_2d": (1024, 1024, 1), "served_mix": (1024, 1024, 256)}
    dflt = own.get(args.config) or tuple(SINGLE_CONFIGS[args.config][k] for k in ("nx", "ny", "nz"))
    args.nx = dflt[0] if args.nx is None else args.nx
    args.ny = dflt[1] if args.ny is None else args.ny
    args.nz = dflt[2] if args.nz is None else args.nz
    if args.kernels and args.kernels != "tune":
        os.environ["FULLWAVE_HIP_KERNELS"] = args.kernels
    if args.pmc_summary:
        dirs = [Path(p) for p in args.pmc_summary]
        print(json.dumps(pmc_summary(*dirs[:4], args.copy_bytes, dirs[4]), indent=1))
        return 0
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU (one simulation or batch per GPU); "
                                   "run with --gpus 1"}), flush=True)
        return 2
    if args.copy_probe:
        from fullwave25_amd import engine

        gbs = engine.copy_bandwidth(args.copy_probe, repeats=5)
        print(json.dumps({"copy_bytes": args.copy_probe, "copy_gbs": round(gbs, 1)}), flush=True)
        return 0
    if args.stream_probe:
        from fullwave25_amd import engine

        r, w, cells = (int(x) for x in args.stream_probe.split(","))
        gbs = engine.stream_bandwidth(cells, r, w, repeats=5)
        print(json.dumps({"reads": r, "writes": w, "cells": cells, "stream_gbs": round(gbs, 1)}), flush=True)
        return 0
    if args.config in ("batched_2d", "batched_mix", "served_mix", "grouped_2d", "adjoint_2d"):
        if args.config in ("batched_2d", "adjoint_2d") and "--steps" not in sys.argv:
            args.steps = 3000 if args.config == "batched_2d" else 2000
        if args.config == "batched_mix":
            if "--transmits" not in sys.argv:
                args.transmits = 128
            if "--jobs" not in sys.argv:
                args.jobs = 32
        fn = {"batched_2d": bench_batched, "batched_mix": bench_batched_mix, "served_mix": bench_served,
              "grouped_2d": bench_grouped, "adjoint_2d": bench_adjoint}[args.config]
        print(json.dumps(fn(args)), flush=True)
        return 0
    if args.sweep:
        return sweep(args)
    line = bench_single(args)
    inp = line.pop("_inputs", None)
    if "metric" not in line:  # --profile-steps
        print(json.dumps(line), flush=True)
        return 0
    if not args.no_cpu_baseline:
        # the GPU line is printed whatever happens here
        try:
            steps = args.cpu_steps or CPU_STEPS.get(args.config, 150)
            line["cpu_baseline"] = cpu_baseline(inp, steps, slab_planes=CPU_SLAB_PLANES.get(args.config, 0))
        except Exception as exc:  # noqa: BLE001
            line["cpu_baseline"] = {"value": None, "unit": "GLUPS", "error": f"{type(exc).__name__}: {exc}"}
    print(json.dumps(line), flush=True)
    return 0


def sweep(args) -> int:
    """Time every kernel family and shape over ``--sweep`` steps on the config's workload."""
    from fullwave25_amd import engine

    wl = build_single(args)
    inp = wl.inputs
    shapes = ["percell"] + [f"stream:{w}:{c}" for w in (128, 256) for c in (8, 16, 32)]
    shapes += [f"tile:{c}" for c in (8, 16, 32)]
    if inp.ndim == 3:
        shapes = ["percell"] + [f"{v}:{w}:{c}" for v in ("stream", "pipe", "pipelite") for w in (512, 1024)
                                for c in (8, 16, 32)]
    if args.sweep_shapes:
        shapes = args.sweep_shapes.split(",")
    upd = updated_cells(inp)
    for shape in shapes:
        os.environ["FULLWAVE_HIP_KERNELS"] = shape
        with engine.HipSimulation(inp, device=0) as sim:
            layout = coefficient_layout(sim, inp)
            bpc = algorithmic_bytes_per_cell(inp.ndim, inp.isotropic, uniform=layout["uniform"],
                                             coded_fraction=layout["coded_fraction"])
            sim.run_timed(min(20, args.sweep))
            sim.reset()
            kms = sim.run_timed(args.sweep)
        gbs = {k: bpc[k] * upd / (kms[k] / 1e3 / args.sweep) / 1e9 for k in ("velocity", "pressure")}
        print(json.dumps({"shape": shape, "us_per_step": {k: round(v / args.sweep * 1e3, 2) for k, v in kms.items()},
                          "GBs": {k: round(v, 1) for k, v in gbs.items()}}), flush=True)
    os.environ.pop("FULLWAVE_HIP_KERNELS", None)
    return 0


if __name__ == "__main__":
    sys.exit(main())
