"""The batched-transmit and served-mix configurations at their own sizes, on the GPU.

Batched transmits: focused transmits, each steered to its own lateral focus (4 of the 128 of the
configuration: both ends of the sweep and two inside), on one seeded 2048 x 1536 heterogeneous
medium (2208 x 1696 cells with the absorbing and transition layers), 3000 steps each, mixed with
jobs on the uneven grids {1024^2, 1536 x 1024, 2048^2, 3072 x 2048}. Served mix: 2D forward
jobs of 512^2 and 2048^2, a misfit gradient on 1024^2 (a seeded true medium against its smoothed
initial model) with checkpoints every 200 steps, and a 256^3 3D forward job, submitted together
to the simulation server.

The reference runs one transmit per engine call (fullwave/solver/solver.py:620-745); batching,
grouping and serving are this engine's own, so each is held to two bars: every job's result
equals the same job run alone (bit for bit: same kernels, same arithmetic), and the FP64 oracle
on the full grids over truncated horizons (rel-L2 <= 1e-4 for traces; <= 1e-3 for the FP32
adjoint's gradients against the FP64 host adjoint). Full-length runs are held to
superposition.
"""

import dataclasses
import functools

import numpy as np
import pytest

import oracle
from fullwave25_amd import engine, workloads
from fullwave25_amd.adjoint import HipAdjoint
from fullwave25_amd.serving import SimulationServer
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-4
PICKS = (0, 42, 85, 127)
UNEVEN = [(1024, 1024), (1536, 1024), (2048, 2048), (3072, 2048)]


@functools.lru_cache(maxsize=None)
def _steered():
    wl, shots = workloads.steered_transmits_2d(transmits=PICKS)
    return wl.inputs, shots


def _kernels_env(shape: dict) -> str:
    """FULLWAVE_HIP_KERNELS for a single handle with a group's kernel family and shape."""
    if shape["variant"] == "tile":
        return f"tile:{shape['chunk']}"
    return f"stream:{shape['width']}:{shape['chunk']}"


def test_steered_batch_full_grid_against_oracle_and_single_runs():
    inp, shots = _steered()
    n = 300
    print(f"steered batch: grid {inp.shape}, {len(PICKS)} transmits, {n} of {inp.nt} steps", flush=True)
    with engine.HipSimulation(inp, shots=shots) as sim:
        sim.run(n)
        batch = sim.sensors().copy()
        cfg = sim.kernel_config()
    assert batch.shape == (len(PICKS), n, inp.ncoordsout)
    for b in range(len(PICKS)):
        with engine.HipSimulation(dataclasses.replace(inp, icmat=shots[b])) as one:
            one.set_kernel_config(cfg)
            one.run(n)
            np.testing.assert_array_equal(batch[b], one.sensors())
    print("  each shot equals its single run", flush=True)
    for b in (0, len(PICKS) - 1):
        o = oracle.OracleSimulation(dataclasses.replace(inp, icmat=shots[b], nt=n), 64)
        o.run(n)
        err = rel_l2(batch[b], o.sensors())
        print(f"  transmit {PICKS[b]}: rel-L2 vs FP64 oracle on the full grid {err:.3e}", flush=True)
        assert np.abs(o.sensors()).max() > 0
        assert err <= TOL


def test_steered_batch_full_length_superposition():
    """Two transmits and their sum as three shots of one handle, over the whole 3000 steps: the
    engine is linear in the source, so the sum's traces equal the sum of the traces up to FP32
    accumulation (1e-7 to 1e-6 relative); an error in the update would show far above 1e-5."""
    inp, shots = _steered()
    three = np.stack([shots[0], shots[-1], shots[0] + shots[-1]]).astype(np.float32)
    with engine.HipSimulation(inp, shots=three) as sim:
        sim.run(inp.nt)
        rec = sim.sensors().astype(np.float64)
    assert np.isfinite(rec).all() and np.abs(rec[0]).max() > 0
    err = rel_l2(rec[2], rec[0] + rec[1])
    print(f"steered transmits, {inp.nt} steps: superposition rel-L2 {err:.2e}", flush=True)
    assert err <= 1e-5


def test_uneven_group_with_steered_batch(monkeypatch):
    """One group of handles: the steered batch (4 shots) and one job on each uneven grid size,
    advanced by shared launches; each member's traces equal its run alone with the group's
    kernel shape, and the smallest member meets the FP64 oracle."""
    inp, shots = _steered()
    members = [workloads.focused_linear_array_2d(nx=nx, ny=ny, nt=3000, seed=2030 + i).inputs
               for i, (nx, ny) in enumerate(UNEVEN)]
    n = 200
    sims = [engine.HipSimulation(inp, shots=shots)] + [engine.HipSimulation(m) for m in members]
    try:
        with engine.HipGroup(sims) as grp:  # family and shape timed on the fresh members
            shape = grp.kernel_config()["velocity"]
            print(f"group of {len(sims)}: {shape}", flush=True)
            grp.run(n)
        got = [s.sensors().copy() for s in sims]
    finally:
        for s in sims:
            s.close()
    monkeypatch.setenv("FULLWAVE_HIP_KERNELS", _kernels_env(shape))
    for (m, sh), g in zip([(inp, shots)] + [(m, None) for m in members], got):
        with engine.HipSimulation(m, shots=sh) as alone:
            alone.run(n)
            np.testing.assert_array_equal(g, alone.sensors())
        print(f"  member {m.shape}{'' if sh is None else f' x {sh.shape[0]} shots'} equals its run alone", flush=True)
    small = members[0]
    o = oracle.OracleSimulation(dataclasses.replace(small, nt=n), 64)
    o.run(n)
    err = rel_l2(got[1], o.sensors())
    print(f"  {small.shape} member: rel-L2 vs FP64 oracle {err:.3e}", flush=True)
    assert err <= TOL


@functools.lru_cache(maxsize=None)
def _served():
    """The served mix through one server; returns the jobs, their results and the kernel
    configurations the server used."""
    fwd_small = workloads.plane_wave_2d(nx=512, ny=512, nt=1000, sensor_depth=256).inputs
    fwd_big = workloads.plane_wave_2d(nx=2048, ny=2048, nt=1000, sensor_depth=1024).inputs
    initial, true = workloads.misfit_media_2d(nt=240)
    with engine.HipSimulation(true.inputs) as sim:
        sim.run(true.inputs.nt)
        observed = sim.sensors().astype(np.float64)
    vol = workloads.plane_wave_3d(256, 256, 256, nt=200).inputs
    jobs = [("forward", fwd_small), ("forward", fwd_big), ("adjoint", initial.inputs), ("forward", vol),
            ("forward", fwd_small)]
    print("served mix: " + ", ".join(f"{k} {i.shape} x {i.nt}" for k, i in jobs), flush=True)
    with SimulationServer(max_slots=4) as srv:
        futs = [srv.submit_adjoint(i, observed) if k == "adjoint" else srv.submit(i)
                for k, i in jobs]
        results = [f.result(timeout=600) for f in futs]
        configs = [srv.kernel_config_for(i, adjoint=k == "adjoint") for k, i in jobs]
        st = srv.stats()
    print(f"  served {st.jobs} jobs: p50 {st.latency_p50_s:.2f} s, p99 {st.latency_p99_s:.2f} s, "
          f"peak running {srv.peak_running}", flush=True)
    assert st.failed == 0
    return jobs, results, configs, observed


def test_served_mix_equals_alone():
    jobs, results, configs, observed = _served()
    for (kind, inp), got, cfg in zip(jobs, results, configs):
        if kind == "adjoint":
            assert got["checkpoint_every"] == inp.nt  # the whole run's tape fits the server's budget
            with HipAdjoint(inp, checkpoint_every=got["checkpoint_every"], kernel_config=cfg) as adj:
                traces = adj.forward()
                resid = traces.astype(np.float64) - observed
                gk, gr = adj.backward(resid.astype(np.float32))
            np.testing.assert_array_equal(got["traces"], traces)
            np.testing.assert_array_equal(got["grad_K"], gk)
            np.testing.assert_array_equal(got["grad_rho"], gr)
            assert np.abs(gk).max() > 0
        else:
            with engine.HipSimulation(inp) as alone:
                alone.set_kernel_config(cfg)
                alone.run(inp.nt)
                np.testing.assert_array_equal(got, alone.sensors())
        print(f"  {kind} {inp.shape} equals its run alone", flush=True)
    o = oracle.OracleSimulation(jobs[0][1], 64)
    o.run(jobs[0][1].nt)
    assert rel_l2(results[0], o.sensors()) <= TOL


def test_served_gradient_against_fp64_adjoint():
    """The 1024^2 misfit gradient (240 steps, the whole run taped as the server chose) against
    the FP64 host adjoint fed the same residual."""
    jobs, results, _, observed = _served()
    inp, got = jobs[2][1], results[2]
    resid = (got["traces"].astype(np.float64) - observed).astype(np.float32).astype(np.float64)
    want_k, want_r = oracle.OracleSimulation(inp, 64).adjoint(resid)
    ek, er = rel_l2(got["grad_K"], want_k), rel_l2(got["grad_rho"], want_r)
    print(f"1024^2 misfit gradient: rel-L2 dK {ek:.2e}, drho {er:.2e}", flush=True)
    assert np.abs(want_k).max() > 0
    assert ek <= 1e-3 and er <= 1e-3
