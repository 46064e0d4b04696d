"""Benchmark: FDTD time steps of the 2D heterogeneous focused-transmit workload on one GPU.

Workload: 4096 x 2048 interior (4256 x 2208 with the 36-cell PML, 36-cell transition layer and
8 ghost cells per side), seeded six-layer medium with sub-resolution scatterers, 128-element
linear array focused at 40 mm, 3 MHz, per-element receive, 8000 steps. A "step" is one time step
(velocity update, pressure update, source/sensor pass) over the whole grid; the default timed
region is the full 8000-step transmit with every input resident in HBM.

Prints one JSON line. ``value`` is GLUPS (extended-grid cells x steps / s). ``roofline`` is for
the dominant kernel, timed with HIP events around each launch on the engine's stream during a
second pass over the same steps; its bytes are the algorithmic bytes per updated cell (each
array element the kernel must read or write, once). ``cpu_baseline`` times the FP32 build of the
oracle (OpenMP) on the first steps of the same workload.
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
METRIC = "GLUPS + % HBM roofline (1 GPU); served jobs/hr with p50/p99 latency"


def algorithmic_bytes_per_cell(ndim: int, isotropic: bool) -> dict:
    """Bytes each kernel must move per updated cell (float32 words and the int32 dcmap)."""
    nd = ndim
    nax = 1 if isotropic else nd
    # velocity: reads p, rho, dcmap, 1/kappa (nax), a/b (4 nax), psi (2 nd), v (nd); writes v, psi
    vel = 4 * ((1 + 1 + 1 + nax + 4 * nax + 2 * nd + nd) + (nd + 2 * nd))
    # pressure: reads v (nd), K, dcmap, 1/kappa (nax), a/b (4 nax), phi (2 nd), p; writes p, phi
    pre = 4 * ((nd + 1 + 1 + nax + 4 * nax + 2 * nd + 1) + (1 + 2 * nd))
    return {"velocity": vel, "pressure": pre}


def build_workload(args):
    from fullwave25_amd import workloads

    if args.config == "focused_2d":
        return workloads.focused_linear_array_2d(nx=args.nx, ny=args.ny, nt=args.steps)
    if args.config == "plane_wave_2d":
        return workloads.plane_wave_2d(nx=args.nx, ny=args.ny, nt=args.steps)
    raise SystemExit(f"unknown config {args.config}")


def cpu_baseline(inp, steps: int) -> dict:
    import oracle

    oracle.build()
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    sim = oracle.OracleSimulation(inp, precision=32)
    sim.run(1)  # first-touch of the state arrays
    t0 = time.perf_counter()
    sim.run(steps)
    dt = time.perf_counter() - t0
    sim.close()
    return {
        "value": inp.ncells * steps / dt / 1e9,
        "unit": "GLUPS",
        "cores": threads,
        "kind": "port",
        "sample": f"FP32 oracle (C, OpenMP, {threads} threads), steps 1..{steps + 1} of the same workload",
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--config", default="focused_2d")
    ap.add_argument("--nx", type=int, default=4096)
    ap.add_argument("--ny", type=int, default=2048)
    ap.add_argument("--cpu-steps", type=int, default=150)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=0,
                    help="run only this many steps and exit (for rocprofv3)")
    ap.add_argument("--sweep", type=int, default=0,
                    help="time every kernel family/shape over this many steps and exit")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run the plain copy kernel over this many bytes (counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        from fullwave25_amd import engine

        gbs = engine.copy_bandwidth(args.copy_probe, repeats=5)
        print(json.dumps({"copy_bytes": args.copy_probe, "copy_gbs": round(gbs, 1)}), flush=True)
        return 0
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs one simulation per GPU and this build measures one GPU; "
                          "run with --gpus 1"}))
        return 2

    from fullwave25_amd import engine

    t0 = time.time()
    wl = build_workload(args)
    inp = wl.inputs
    print(f"# workload {wl.name} {inp.shape} nt={inp.nt} built in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    if args.sweep:
        shapes = ["percell"] + [f"stream:{w}:{c}" for w in (128, 256) for c in (8, 16, 32)]
        shapes += [f"tile:{c}" for c in (8, 16, 32)]
        if inp.ndim == 3:
            shapes = ["percell"] + [f"stream:64:{c}" for c in (8, 16, 32, 64)]
        bpc = algorithmic_bytes_per_cell(inp.ndim, inp.isotropic)
        updated = int(np.prod([s - 16 for s in inp.shape]))
        for shape in shapes:
            os.environ["FULLWAVE_HIP_KERNELS"] = shape
            with engine.HipSimulation(inp, device=0) as sim:
                sim.run_timed(min(20, args.sweep))
                sim.reset()
                kms = sim.run_timed(args.sweep)
            gbs = {k: bpc[k] * updated / (kms[k] / 1e3 / args.sweep) / 1e9 for k in ("velocity", "pressure")}
            print(json.dumps({"shape": shape, "us_per_step": {k: round(v / args.sweep * 1e3, 2) for k, v in kms.items()},
                              "GBs": {k: round(v, 1) for k, v in gbs.items()}}), flush=True)
        os.environ.pop("FULLWAVE_HIP_KERNELS", None)
        return 0
    sim = engine.HipSimulation(inp, device=0)
    kcfg = sim.kernel_config()
    if args.profile_steps:
        sim.run(args.profile_steps)
        sim.close()
        return 0
    sim.run(min(args.warmup, inp.nt))
    sim.reset()
    # fw_sim_run returns after synchronizing the engine's stream, so the wall clock around it
    # brackets exactly the K steps (one GPU: no barrier to take)
    t1 = time.perf_counter()
    sim.run(args.steps)
    elapsed = time.perf_counter() - t1
    traces = sim.sensors()
    finite = bool(np.isfinite(traces).all())
    # second pass over the same steps with HIP events around every launch
    sim.reset()
    kms = sim.run_timed(args.steps)
    sim.close()
    copy_gbs = engine.copy_bandwidth(1 << 30, repeats=10)

    cells = inp.ncells
    updated = int(np.prod([s - 16 for s in inp.shape]))
    glups = cells * args.steps / elapsed / 1e9
    bpc = algorithmic_bytes_per_cell(inp.ndim, inp.isotropic)
    dominant = max(("velocity", "pressure"), key=lambda k: kms[k])
    avg_s = kms[dominant] / 1e3 / args.steps
    achieved = bpc[dominant] * updated / avg_s / 1e9
    line = {
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
            "workload": f"{wl.name}: 2D heterogeneous {args.nx}x{args.ny} interior "
                        f"({inp.nx}x{inp.ny} extended), {args.steps} steps, isotropic relaxation",
            "cells": cells,
            "sources": inp.ncoords,
            "sensors": inp.ncoordsout,
            "finite_output": finite,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dominant,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "bytes_per_cell": bpc[dominant],
            "avg_launch_us": round(avg_s * 1e6, 2),
            "kernel_ms_per_step": {k: round(v / args.steps, 5) for k, v in kms.items()},
            "copy_kernel_gbs": round(copy_gbs, 1),
            "kernel_config": kcfg,
        },
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(inp, args.cpu_steps)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
