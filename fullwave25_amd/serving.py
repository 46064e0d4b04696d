"""Serving: many simulation jobs on one GPU, admitted by device-memory budget.

The reference runs one simulation per engine call and has no server; this module has no
counterpart there. A :class:`SimulationServer` accepts jobs (engine inputs, optionally a batch
of source matrices on one medium) from any thread and returns a future per job. A scheduler
thread admits waiting jobs in arrival order whenever the device memory they need fits the
budget and a slot is free (a job that does not fit lets smaller later jobs pass it, up to
``max_bypass`` times, then it waits for memory to free up); each admitted job runs on its own
engine handle, which owns its own HIP stream, from its own worker thread, so the kernels of
concurrent jobs overlap on the device. Inputs stay resident in HBM from admission to
completion; the host only moves the inputs in and the traces out.

Continuous batching (``batching=True``, the default): 2D isotropic forward jobs of at most
``group_max_cells`` cells (all shots together; larger ones fill the GPU alone and keep their
own tuned kernels) do not run their own launch loop. Once a job's handle holds its inputs, it joins a shared group
(:class:`~fullwave25_amd.engine.HipGroup`), and a batcher thread advances the whole group
``quantum_steps`` at a time with one velocity, one pressure and one source/sensor launch per
step for all members; jobs that arrive meanwhile join at the next quantum, at their own step 0,
and finished jobs leave. A grouped job's traces are bit-identical to its own run with the
group's kernel shape.

Kernel shapes: the first handle of each grid shape times the candidates (the engine's
autotuner); later jobs of that shape reuse the choice, so a job's numbers do not depend on
what else is running. Grouped jobs use a fixed group shape chosen by the job's own grid: the
tiled kernels ``tile:8`` where the rows are 16-byte aligned (ny % 4 == 0) and the grid has at
most 768 sound speeds, else the marching kernels ``stream:256:16``; the batcher keeps one group
per shape.

Latency of a job is measured from ``submit`` to its result being available.
"""

from __future__ import annotations

import math
import threading
import time
from collections import deque
from concurrent.futures import Future
from dataclasses import dataclass, field

import numpy as np

from . import engine
from .solver.inputs import SimulationInputs


@dataclass
class JobRecord:
    """Timing of one job (seconds, ``time.perf_counter`` clock)."""

    job_id: int
    cells: int
    steps: int
    n_shots: int
    device_bytes: int
    submitted: float
    started: float = 0.0
    finished: float = 0.0
    error: str | None = None

    @property
    def latency(self) -> float:
        return self.finished - self.submitted

    @property
    def queued(self) -> float:
        return self.started - self.submitted


@dataclass
class _Job:
    record: JobRecord
    inputs: SimulationInputs
    shots: np.ndarray | None
    future: Future
    bypassed: int = 0
    observed: np.ndarray | None = None  # adjoint jobs: the traces to fit
    checkpoint_every: int = 0


def adjoint_device_bytes(inp: SimulationInputs, checkpoint_every: int) -> int:
    """Device memory of an adjoint job with segments of ``checkpoint_every`` steps (> 0): the
    forward handle, its checkpoints, the tape of one segment (ndim + 1 grids a step), the
    adjoint state and scratch."""
    nd = inp.ndim
    cells = inp.nx * inp.ny * (-(-inp.nz // 32) * 32) if nd == 3 else inp.ncells
    state = (1 + nd + (0 if inp.exponential else 4 * nd)) * cells
    every = max(1, min(checkpoint_every, inp.nt))
    n_ckpt = -(-inp.nt // every)
    extra = (n_ckpt + 1) * state + every * (nd + 1) * cells + (3 * nd + 2) * cells + inp.n_records * inp.ncoordsout
    return engine.device_bytes(inp) + 4 * extra


def adjoint_segment(inp: SimulationInputs, share_bytes: int) -> int:
    """Segment length of an adjoint job allowed ``share_bytes`` of device memory: the whole run
    (taped in its forward run, nothing recomputed) when that fits, else the length that needs
    the least memory (ceil(nT / C) checkpoints of the state against C steps of tape)."""
    if adjoint_device_bytes(inp, inp.nt) <= share_bytes:
        return int(inp.nt)
    nd = inp.ndim
    state = 1 + nd + (0 if inp.exponential else 4 * nd)
    return int(min(inp.nt, max(1, math.ceil(math.sqrt(inp.nt * state / (nd + 1))))))


@dataclass
class ServerStats:
    jobs: int
    failed: int
    wall_s: float
    jobs_per_hour: float
    latency_p50_s: float
    latency_p99_s: float
    queue_p50_s: float
    glups: float
    records: list = field(repr=False, default_factory=list)


GROUP_WIDTH, GROUP_CHUNK = 256, 16  # marching kernel shape of grouped jobs
GROUP_TILE_ROWS = 8  # tiled kernel shape of grouped jobs with ny % 4 == 0
TILE_MAX_NDMAP = 768  # sound speeds the tiled kernels take (their coefficient sets sit in LDS)


def _group_shape(inp: SimulationInputs) -> dict:
    """The kernel shape a grouped job runs with (depends on the job's own grid only)."""
    if inp.ny % 4 == 0 and inp.ndmap <= TILE_MAX_NDMAP:
        return {"variant": "tile", "width": 256, "chunk": GROUP_TILE_ROWS}
    return {"variant": "stream", "width": GROUP_WIDTH, "chunk": GROUP_CHUNK}
# Jobs of more cells (all shots together) than this fill the GPU on their own, where their own
# tuned kernels (the float4 tiles) outrun the grouped marching kernels: they run alone.
GROUP_MAX_CELLS = 1 << 21


def _shape_key(inp: SimulationInputs, n_shots: int) -> tuple:
    return (inp.ndim, inp.shape, bool(inp.isotropic), bool(inp.exponential), inp.ndmap == 1, n_shots)


class SimulationServer:
    """Concurrent simulation jobs on one device under a memory budget.

    ``hbm_budget``: bytes of device memory jobs may hold at once (default: 90 % of the free
    memory when the server starts). ``max_slots``: jobs running at once. ``chunk_steps``: steps
    per engine call of a worker (a job's progress granularity).
    """

    def __init__(self, device: int = 0, *, hbm_budget: int | None = None, max_slots: int = 8,
                 chunk_steps: int = 500, max_bypass: int = 8, runner=None, batching: bool = True,
                 quantum_steps: int = 250, group_max_cells: int = GROUP_MAX_CELLS) -> None:
        self.device = device
        self.group_max_cells = int(group_max_cells)
        self.batching = bool(batching) and runner is None
        self.quantum_steps = int(quantum_steps)
        self._batch_lock = threading.Condition()
        self._batch_in: list = []
        self._batch_closed = False
        self._batcher = None
        # runner(server, inputs, shots) -> traces; the engine by default (tests of the
        # scheduling logic substitute a host-only one)
        self._runner = runner if runner is not None else SimulationServer._run_engine
        if hbm_budget is None:
            free, _ = engine.device_memory(device)
            hbm_budget = int(0.9 * free)
        self.budget = int(hbm_budget)
        self.max_slots = int(max_slots)
        self.chunk_steps = int(chunk_steps)
        self.max_bypass = int(max_bypass)
        self._lock = threading.Condition()
        self._pending: deque[_Job] = deque()
        self._in_use = 0
        self.peak_in_use = 0
        self.peak_running = 0
        self._running = 0
        self._next_id = 0
        self._closed = False
        self._records: list[JobRecord] = []
        self._configs: dict = {}
        self._config_locks: dict = {}
        self._workers: list[threading.Thread] = []
        self._t0 = time.perf_counter()
        self._scheduler = threading.Thread(target=self._schedule, name="fullwave-scheduler", daemon=True)
        self._scheduler.start()
        if self.batching:
            self._batcher = threading.Thread(target=self._batch_loop, name="fullwave-batcher", daemon=True)
            self._batcher.start()

    # ---- public ---------------------------------------------------------------------------

    def submit(self, inputs: SimulationInputs, *, shots: np.ndarray | None = None) -> Future:
        """Queue a job; the future's result is the traces (``[n_records, n_sensors]``, or
        ``[n_shots, n_records, n_sensors]`` with ``shots``) and ``future.record`` its timing."""
        n_shots = 1 if shots is None else int(np.asarray(shots).shape[0])
        need = engine.device_bytes(inputs, n_shots)
        if need > self.budget:
            raise ValueError(f"job needs {need} bytes of device memory, the budget is {self.budget}")
        fut: Future = Future()
        with self._lock:
            if self._closed:
                raise RuntimeError("server is closed")
            rec = JobRecord(self._next_id, inputs.ncells, inputs.nt, n_shots, need, time.perf_counter())
            self._next_id += 1
            fut.record = rec
            self._pending.append(_Job(rec, inputs, shots, fut))
            self._lock.notify_all()
        return fut

    def submit_adjoint(self, inputs: SimulationInputs, observed: np.ndarray, *,
                       checkpoint_every: int = 0) -> Future:
        """Queue a misfit-gradient job: the result is ``{"loss", "grad_K", "grad_rho", "traces",
        "checkpoint_every"}`` for ``0.5 * ||traces - observed||^2`` (gradients with respect to the K and
        rho maps of ``inputs``; see :mod:`fullwave25_amd.adjoint`). ``checkpoint_every=0``: the
        whole run taped when that needs at most a quarter of the memory budget, else the segment
        length that needs the least memory (:func:`adjoint_segment`)."""
        if checkpoint_every <= 0:
            checkpoint_every = adjoint_segment(inputs, self.budget // 4)
        need = adjoint_device_bytes(inputs, checkpoint_every)
        if need > self.budget:
            raise ValueError(f"job needs {need} bytes of device memory, the budget is {self.budget}")
        fut: Future = Future()
        with self._lock:
            if self._closed:
                raise RuntimeError("server is closed")
            rec = JobRecord(self._next_id, inputs.ncells, inputs.nt, 1, need, time.perf_counter())
            self._next_id += 1
            fut.record = rec
            self._pending.append(_Job(rec, inputs, None, fut, observed=np.asarray(observed, np.float64),
                                      checkpoint_every=int(checkpoint_every)))
            self._lock.notify_all()
        return fut

    def drain(self) -> None:
        """Block until every submitted job has finished."""
        with self._lock:
            while self._pending or self._running:
                self._lock.wait()

    def close(self) -> None:
        self.drain()
        with self._lock:
            self._closed = True
            self._lock.notify_all()
        self._scheduler.join()
        for w in self._workers:
            w.join()
        if self._batcher is not None:
            with self._batch_lock:
                self._batch_closed = True
                self._batch_lock.notify_all()
            self._batcher.join()

    def __enter__(self) -> SimulationServer:
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def stats(self) -> ServerStats:
        with self._lock:
            recs = [r for r in self._records]
        done = [r for r in recs if r.error is None]
        if not done:
            return ServerStats(0, len(recs), 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, recs)
        t_first = min(r.submitted for r in recs)
        t_last = max(r.finished for r in recs)
        wall = max(t_last - t_first, 1e-9)
        lat = np.array([r.latency for r in done])
        queued = np.array([r.queued for r in done])
        cell_steps = sum(r.cells * r.steps * r.n_shots for r in done)
        return ServerStats(
            jobs=len(done), failed=len(recs) - len(done), wall_s=wall, jobs_per_hour=len(done) / wall * 3600.0,
            latency_p50_s=float(np.percentile(lat, 50)), latency_p99_s=float(np.percentile(lat, 99)),
            queue_p50_s=float(np.percentile(queued, 50)), glups=cell_steps / wall / 1e9, records=recs,
        )

    # ---- scheduler and workers ------------------------------------------------------------

    def _pick(self) -> _Job | None:
        """First waiting job that fits, in arrival order; a job passed over too often blocks
        the ones behind it until it fits."""
        free = self.budget - self._in_use
        for idx, job in enumerate(self._pending):
            if job.record.device_bytes <= free:
                for earlier in list(self._pending)[:idx]:
                    earlier.bypassed += 1
                del self._pending[idx]
                return job
            if job.bypassed >= self.max_bypass:
                return None
        return None

    def _schedule(self) -> None:
        while True:
            with self._lock:
                while True:
                    if self._closed and not self._pending:
                        return
                    job = self._pick() if self._running < self.max_slots else None
                    if job is not None:
                        break
                    self._lock.wait()
                self._in_use += job.record.device_bytes
                self._running += 1
                self.peak_in_use = max(self.peak_in_use, self._in_use)
                self.peak_running = max(self.peak_running, self._running)
            w = threading.Thread(target=self._work, args=(job,), name=f"fullwave-job-{job.record.job_id}", daemon=True)
            self._workers.append(w)
            w.start()

    def _kernel_config(self, key, sim: engine.HipSimulation):
        """Reuse the kernel shapes tuned for this grid shape; the first handle tunes them."""
        with self._lock:
            lk = self._config_locks.setdefault(key, threading.Lock())
        with lk:
            cfg = self._configs.get(key)
            if cfg is not None:
                sim.set_kernel_config(cfg)
            else:
                self._configs[key] = sim.kernel_config()  # tunes now, on this handle's arrays

    def kernel_config_for(self, inp: SimulationInputs, n_shots: int = 1, *, adjoint: bool = False) -> dict | None:
        """The kernel configuration this server runs (or ran) jobs of this shape with: the group
        shape for grouped forward jobs, else the one tuned on the first handle of the shape (None
        if no such job has run yet). Misfit-gradient jobs (``adjoint``) never join a group and
        keep their own configurations (their handles run kernels that write the adjoint tape)."""
        if adjoint:
            with self._lock:
                return self._configs.get(_shape_key(inp, 1) + ("adjoint",))
        if self._grouped(inp, n_shots):
            shape = _group_shape(inp)
            return {"velocity": shape, "pressure": shape}
        with self._lock:
            return self._configs.get(_shape_key(inp, n_shots))

    @staticmethod
    def _groupable(inp: SimulationInputs) -> bool:
        """Jobs the grouped kernels take: 2D isotropic relaxation, coefficient sets in LDS."""
        return (inp.ndim == 2 and bool(inp.isotropic) and not bool(inp.exponential) and inp.ndmap <= 1024
                and inp.ncells < (1 << 30))

    def _grouped(self, inp: SimulationInputs, n_shots: int) -> bool:
        """Whether a forward job joins the shared group: batching on, a groupable model, and
        small enough that its own launches would leave the GPU partly idle."""
        return self.batching and self._groupable(inp) and inp.ncells * n_shots <= self.group_max_cells

    def _run_engine(self, inp: SimulationInputs, shots: np.ndarray | None) -> np.ndarray:
        n_shots = 1 if shots is None else int(shots.shape[0])
        if self._grouped(inp, n_shots):
            with engine.HipSimulation(inp, self.device, shots=shots) as sim:
                shape = _group_shape(inp)
                sim.set_kernel_config({"velocity": shape, "pressure": shape})
                sim.relaxation_codes()  # first-run set-up here, on this worker's thread
                done = threading.Event()
                failure: list = []
                with self._batch_lock:
                    self._batch_in.append((sim, inp.nt, done, failure, shape["variant"]))
                    self._batch_lock.notify_all()
                done.wait()
                if failure:
                    raise failure[0]
                return sim.sensors()
        with engine.HipSimulation(inp, self.device, shots=shots) as sim:
            self._kernel_config(_shape_key(inp, n_shots), sim)
            left = inp.nt
            while left > 0:
                n = min(self.chunk_steps, left)
                sim.run(n)
                left -= n
            return sim.sensors()

    def _batch_loop(self) -> None:
        """Advance the grouped jobs: a quantum of steps (or fewer, up to the next job to finish)
        per round; arrivals join and finished jobs leave between rounds. Jobs of each kernel
        shape form their own group; a round advances every group by the same steps."""
        members: list = []  # [sim, steps left, done event, failure list, shape]
        groups: dict = {}   # shape variant -> HipGroup of the members with that shape
        while True:
            with self._batch_lock:
                while not self._batch_in and not members:
                    if self._batch_closed:
                        return
                    self._batch_lock.wait()
                new, self._batch_in = self._batch_in, []
            changed = set()
            for sim, steps, done, failure, variant in new:
                members.append([sim, steps, done, failure, variant])
                changed.add(variant)
            try:
                for v in changed:
                    if v in groups:
                        groups.pop(v).close()
                for v in {m[4] for m in members}:
                    if v not in groups:
                        mine = [m[0] for m in members if m[4] == v]
                        groups[v] = (engine.HipGroup(mine, variant="tile", chunk=GROUP_TILE_ROWS) if v == "tile" else
                                     engine.HipGroup(mine, width=GROUP_WIDTH, chunk=GROUP_CHUNK))
                q = min(self.quantum_steps, min(m[1] for m in members))
                for grp in groups.values():
                    grp.run(q)
            except Exception as exc:  # every member of the failed round fails; the server goes on
                for grp in groups.values():
                    grp.close()
                groups = {}
                for m in members:
                    m[3].append(exc)
                    m[2].set()
                members = []
                continue
            for m in members:
                m[1] -= q
            finished = [m for m in members if m[1] == 0]
            if finished:
                for v in {m[4] for m in finished}:
                    groups.pop(v).close()
                members = [m for m in members if m[1] > 0]
                for m in finished:
                    m[2].set()

    def _run_adjoint(self, job: _Job) -> dict:
        from .adjoint import HipAdjoint

        inp = job.inputs
        key = _shape_key(inp, 1) + ("adjoint",)
        with self._lock:
            cfg = self._configs.get(key)
        # the adjoint handle fixes its kernels when it is created: pass a known configuration in
        with HipAdjoint(inp, checkpoint_every=job.checkpoint_every, device=self.device, kernel_config=cfg) as adj:
            if cfg is None:
                with self._lock:
                    self._configs.setdefault(key, adj.sim.kernel_config())
            traces = adj.forward()
            resid = traces.astype(np.float64) - job.observed
            gk, gr = adj.backward(resid.astype(np.float32))
        return {"loss": 0.5 * float(np.sum(resid**2)), "grad_K": gk, "grad_rho": gr, "traces": traces,
                "checkpoint_every": job.checkpoint_every}

    def _work(self, job: _Job) -> None:
        rec = job.record
        rec.started = time.perf_counter()
        try:
            if job.observed is not None:
                out = self._run_adjoint(job)
            else:
                out = self._runner(self, job.inputs, job.shots)
            rec.finished = time.perf_counter()
            job.future.set_result(out)
        except Exception as exc:  # the job fails, the server goes on
            rec.finished = time.perf_counter()
            rec.error = repr(exc)
            job.future.set_exception(exc)
        finally:
            with self._lock:
                self._in_use -= rec.device_bytes
                self._running -= 1
                self._records.append(rec)
                self._lock.notify_all()


def run_sequential(jobs: list[tuple[SimulationInputs, np.ndarray | None]], device: int = 0) -> tuple[list, float]:
    """The plain path for comparison: one engine handle per job, one job at a time (create,
    upload, tune, run, download). Returns ``(results, wall seconds)``."""
    t0 = time.perf_counter()
    out = []
    for inp, shots in jobs:
        with engine.HipSimulation(inp, device, shots=shots) as sim:
            sim.run(inp.nt)
            out.append(sim.sensors())
    return out, time.perf_counter() - t0


__all__ = ["JobRecord", "ServerStats", "SimulationServer", "run_sequential"]
