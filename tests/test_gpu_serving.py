"""The simulation server on the GPU: a mix of 2D, 3D, exponential-model and batched jobs runs
concurrently under a memory budget, and every result equals the same job run alone (the server
reuses one kernel configuration per grid shape, so the kernels and their arithmetic are the
same) and matches the FP64 oracle."""

import dataclasses

import numpy as np
import pytest

import oracle
from fullwave25_amd import engine, workloads
from fullwave25_amd.serving import SimulationServer
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu


def _jobs():
    a = workloads.random_medium_2d(60, 52, 200, seed=50, air=3).inputs
    b = workloads.random_medium_2d(96, 140, 260, seed=51).inputs
    c = workloads.random_medium_3d(20, 18, 22, 80, seed=52).inputs
    d = workloads.random_medium_2d(50, 46, 180, seed=53, exponential=True).inputs
    rng = np.random.default_rng(54)
    shots = np.stack([a.icmat * rng.uniform(0.5, 2.0) for _ in range(3)]).astype(np.float32)
    e = workloads.random_medium_2d(40, 131, 150, seed=56).inputs  # ny % 4 != 0: the marching group
    # each job twice, interleaved, plus a batch of transmits on a's medium
    return [(a, None), (b, None), (c, None), (d, None), (a, shots), (b, None), (c, None), (d, None), (a, None),
            (e, None)]


@pytest.mark.parametrize("batching", ["group", "mixed", "off"])
def test_served_mix_equals_alone_and_oracle(batching):
    """With batching, the 2D isotropic jobs (a, b and the batch on a's medium in the tiled
    group, e in the marching group) run as group members while the 3D and exponential jobs run
    on their own handles; "mixed" caps the groups at a's size, so that b, e and the batch run
    alone beside them."""
    jobs = _jobs()
    a, b = jobs[0][0], jobs[1][0]
    limit = a.ncells if batching == "mixed" else 1 << 21
    assert b.ncells > a.ncells  # b (and the batch of three) exceed the "mixed" cap
    budget = 3 * max(engine.device_bytes(inp, 1 if s is None else s.shape[0]) for inp, s in jobs)
    with SimulationServer(hbm_budget=budget, max_slots=4, chunk_steps=64, batching=batching != "off",
                          quantum_steps=48, group_max_cells=limit) as srv:
        futs = [srv.submit(inp, shots=s) for inp, s in jobs]
        results = [f.result(timeout=300) for f in futs]
        configs = [srv.kernel_config_for(inp, 1 if s is None else s.shape[0]) for inp, s in jobs]
        grouped = [srv._grouped(inp, 1 if s is None else s.shape[0]) for inp, s in jobs]
    want = {"group": [True, True, False, False, True, True, False, False, True, True],
            "mixed": [True, False, False, False, False, False, False, False, True, False],
            "off": [False] * 10}[batching]
    assert grouped == want
    if batching == "group":  # tiled group for aligned rows, marching group for the rest
        assert configs[0]["velocity"]["variant"] == "tile" and configs[9]["velocity"]["variant"] == "stream"
    st = srv.stats()
    print(f"served {st.jobs} jobs: p50 {st.latency_p50_s * 1e3:.1f} ms, p99 {st.latency_p99_s * 1e3:.1f} ms, "
          f"peak running {srv.peak_running}, peak bytes {srv.peak_in_use}", flush=True)
    assert st.failed == 0 and srv.peak_in_use <= budget and srv.peak_running >= 2
    for (inp, shots), got, cfg in zip(jobs, results, configs):
        with engine.HipSimulation(inp, shots=shots) as alone:
            alone.set_kernel_config(cfg)
            alone.run(inp.nt)
            np.testing.assert_array_equal(got, alone.sensors())
    for inp, shots in jobs[:4]:
        o = oracle.OracleSimulation(inp, 64)
        o.run(inp.nt)
        got = results[jobs.index((inp, shots))]
        assert rel_l2(got, o.sensors()) <= 1e-4


def test_failed_job_does_not_stop_the_server():
    good = workloads.random_medium_2d(40, 40, 60, seed=55).inputs
    bad = dataclasses.replace(good, outc=np.array([[10**6, 3]], np.int32))
    with SimulationServer(max_slots=2) as srv:
        f1, f2, f3 = srv.submit(good), srv.submit(bad), srv.submit(good)
        r1, r3 = f1.result(timeout=120), f3.result(timeout=120)
        with pytest.raises(engine.EngineError, match="sensor coordinate"):
            f2.result(timeout=120)
    np.testing.assert_array_equal(r1, r3)
