"""Scheduling logic of the simulation server, with a host-only runner (no GPU): memory-budget
admission, the slot limit, arrival order with bounded bypassing, failure isolation and the
latency statistics."""

import threading
import time

import numpy as np
import pytest

from fullwave25_amd import engine
from fullwave25_amd.serving import SimulationServer, adjoint_device_bytes, adjoint_segment
from tests.helpers import raw_inputs


def _inp(n, nt=10):
    return raw_inputs((n, n), nt=nt, icc=[[20, 20]], icmat=np.ones((1, nt)), outc=[[21, 21]])


class _Recorder:
    def __init__(self, delay=0.05):
        self.delay = delay
        self.lock = threading.Lock()
        self.order = []

    def __call__(self, server, inp, shots):
        with self.lock:
            self.order.append(inp.nx)
        time.sleep(self.delay)
        if inp.nt == 13:
            raise RuntimeError("planted failure")
        return np.full((inp.n_records, inp.ncoordsout), float(inp.nx), np.float32)


def test_budget_and_slots_respected():
    jobs = [_inp(n) for n in (40, 60, 80, 40, 100, 60, 40, 80)]
    need = [engine.device_bytes(j) for j in jobs]
    budget = need[4] + need[0]  # the largest job fits only next to the smallest
    rec = _Recorder()
    with SimulationServer(hbm_budget=budget, max_slots=3, runner=rec) as srv:
        futs = [srv.submit(j) for j in jobs]
        outs = [f.result(timeout=30) for f in futs]
    for j, o in zip(jobs, outs):
        assert o[0, 0] == j.nx
    assert srv.peak_in_use <= budget
    assert srv.peak_running <= 3
    st = srv.stats()
    assert st.jobs == len(jobs) and st.failed == 0
    assert st.latency_p99_s >= st.latency_p50_s > 0


def test_too_large_job_rejected_at_submit():
    with SimulationServer(hbm_budget=1000, runner=_Recorder()) as srv:
        with pytest.raises(ValueError, match="budget"):
            srv.submit(_inp(40))


def test_bypass_is_bounded():
    """A big job at the head lets small ones pass only max_bypass times."""
    small, big = _inp(40), _inp(120)
    budget = engine.device_bytes(big)
    rec = _Recorder(delay=0.02)
    with SimulationServer(hbm_budget=budget, max_slots=4, max_bypass=2, runner=rec) as srv:
        first = srv.submit(small)  # occupies part of the budget, so `big` cannot start yet
        time.sleep(0.005)
        fb = srv.submit(big)
        rest = [srv.submit(small) for _ in range(6)]
        for f in [first, fb, *rest]:
            f.result(timeout=30)
    pos_big = rec.order.index(120)
    assert pos_big <= 1 + 2  # after the first job and at most two bypassing ones


def test_failure_is_isolated():
    rec = _Recorder(delay=0.01)
    with SimulationServer(hbm_budget=10**9, max_slots=2, runner=rec) as srv:
        ok = srv.submit(_inp(40))
        bad = srv.submit(_inp(40, nt=13))
        ok2 = srv.submit(_inp(60))
        assert ok.result(timeout=30)[0, 0] == 40
        with pytest.raises(RuntimeError, match="planted"):
            bad.result(timeout=30)
        assert ok2.result(timeout=30)[0, 0] == 60
    st = srv.stats()
    assert st.jobs == 2 and st.failed == 1


def test_group_membership_and_shapes():
    """Which forward jobs join the batcher's groups, and with which fixed kernel shape: 2D
    isotropic jobs up to the cell cap (all shots together) join, the tiled shape for rows of a
    multiple of 4 cells and the marching shape otherwise; larger, 3D and exponential jobs run
    alone on the shape tuned for them (none before their first handle)."""
    aligned, unaligned, larger = _inp(40), _inp(41), _inp(48)
    cube = raw_inputs((20, 20, 20), nt=10, icc=[[10, 10, 10]], icmat=np.ones((1, 10)), outc=[[11, 11, 11]])
    expo = raw_inputs((40, 40), nt=10, icc=[[20, 20]], icmat=np.ones((1, 10)), outc=[[21, 21]], a_exp=0.999)
    with SimulationServer(hbm_budget=10**9, batching=True, group_max_cells=45 * 45) as srv:
        assert srv._grouped(aligned, 1) and srv._grouped(unaligned, 1)
        assert not srv._grouped(aligned, 2)  # two shots: 3200 cells
        assert not srv._grouped(larger, 1) and not srv._grouped(cube, 1) and not srv._grouped(expo, 1)
        assert srv.kernel_config_for(aligned)["velocity"] == {"variant": "tile", "width": 256, "chunk": 8}
        assert srv.kernel_config_for(unaligned)["pressure"] == {"variant": "stream", "width": 256, "chunk": 16}
        assert srv.kernel_config_for(larger) is None
        many = _inp(40)
        many.dmap = np.zeros((9, 2, 800), np.float32)  # more sound speeds than the tiled kernels hold
        assert srv.kernel_config_for(many)["velocity"]["variant"] == "stream"
    with SimulationServer(hbm_budget=10**9, batching=False) as srv:
        assert not srv._grouped(aligned, 1) and srv.kernel_config_for(aligned) is None


def test_adjoint_segment_policy():
    """A gradient job tapes its whole run when that fits its share of the budget; otherwise its
    segment length minimises checkpoints (ceil(nT / C) states) plus tape (C steps of ndim + 1
    grids), checked here against every C by brute force."""
    import math

    inp = raw_inputs((64, 48), nt=400, icc=[[20, 20]], icmat=np.ones((1, 400)), outc=[[21, 21]])
    full = adjoint_device_bytes(inp, inp.nt)
    assert adjoint_segment(inp, full) == inp.nt
    c = adjoint_segment(inp, full - 1)
    assert 1 <= c < inp.nt
    state = 1 + 2 + 4 * 2

    def floats(e):
        return math.ceil(inp.nt / e) * state + e * 3

    best = min(floats(e) for e in range(1, inp.nt + 1))
    assert floats(c) <= 1.05 * best
