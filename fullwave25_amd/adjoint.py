"""Gradients of sensor-trace functionals with respect to the sound-speed and density maps.

The reference simulates forward only; this module has no counterpart there. The gradient comes
from the engine's hand-written adjoint kernels (the transposed time step, run backwards over a
tape of the forward run, with checkpointed recomputation where the whole tape does not fit;
``fw_adj_*`` in include/fullwave_hip.h), exposed as

* :class:`HipAdjoint` — one forward run from rest and any number of reverse sweeps on its tape
  and checkpoints: ``backward(w)`` returns ``dJ/dK`` and ``dJ/drho`` for ``J = sum(w * traces)``;
* :class:`TraceFunction` — a ``torch.autograd.Function``: ``traces = TraceFunction.apply(c, rho,
  template)`` with ``c`` and ``rho`` maps on the padded grid, so any torch loss of the traces
  back-propagates into ``c`` and ``rho``;
* :func:`misfit_gradient` — the L2 sensor misfit ``0.5 * ||traces - observed||^2`` and its
  gradient in one call.

Model: ``K = rho c^2`` (float32), ``rho`` enters the velocity update through the two-point
average on each staggered face. The coefficient sets (``dcmap``/``dmap``, a function of the
rounded sound speed) and the attenuation maps are held fixed at the template's values, so the
gradient is that of the simulation with those sets.
"""

from __future__ import annotations

import ctypes
import dataclasses

import numpy as np

from . import engine
from .solver.inputs import SimulationInputs


class HipAdjoint:
    """A forward run of ``inputs`` kept for reverse sweeps. The sweep reads a tape of the run
    (``ndim + 1`` grids a step); the run is cut into segments of ``checkpoint_every`` steps whose
    tapes are recomputed from checkpoints, except the last. ``checkpoint_every=0`` (default) tapes
    the whole run when that fits in half of the device's free memory (no recomputation), else
    picks the segment length that needs the least memory; ``self.checkpoint_every`` says which."""

    def __init__(self, inputs: SimulationInputs, *, checkpoint_every: int = 0, device: int = 0,
                 kernel_config: dict | None = None) -> None:
        self.inp = inputs
        self.sim = engine.HipSimulation(inputs, device)
        if kernel_config:
            self.sim.set_kernel_config(kernel_config)
        self.lib = self.sim.lib
        self._a = ctypes.c_void_p()
        try:
            engine._check(self.lib, self.lib.fw_adj_create(self.sim._h, int(checkpoint_every), ctypes.byref(self._a)))
            every = ctypes.c_int64()
            engine._check(self.lib, self.lib.fw_adj_checkpoint_every(self._a, ctypes.byref(every)))
            self.checkpoint_every = int(every.value)
        except Exception:
            self.close()
            raise

    def forward(self) -> np.ndarray:
        """Run ``nT`` steps from rest; returns the traces ``[n_records, n_sensors]``."""
        engine._check(self.lib, self.lib.fw_adj_forward(self._a))
        return self.sim.sensors()

    def backward(self, weights: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
        """``(dJ/dK, dJ/drho)`` on the grid for ``J = sum(weights * traces)``."""
        w = np.ascontiguousarray(weights, dtype=np.float32)
        want = (self.inp.n_records, self.inp.ncoordsout)
        if w.shape != want:
            raise ValueError(f"weights must be {want}, got {w.shape}")
        gk = np.empty(self.inp.shape, np.float32)
        gr = np.empty(self.inp.shape, np.float32)
        engine._check(self.lib, self.lib.fw_adj_backward(self._a, w.ctypes.data_as(ctypes.c_void_p), w.nbytes,
                                                         gk.ctypes.data_as(ctypes.c_void_p),
                                                         gr.ctypes.data_as(ctypes.c_void_p), gk.nbytes))
        return gk, gr

    def close(self) -> None:
        if self._a:
            self.lib.fw_adj_destroy(self._a)
            self._a = ctypes.c_void_p()
        self.sim.close()

    def __enter__(self) -> HipAdjoint:
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def __del__(self) -> None:
        try:
            self.close()
        except Exception:
            pass


def with_medium(template: SimulationInputs, c: np.ndarray, rho: np.ndarray) -> SimulationInputs:
    """``template`` with ``K = rho c^2`` and ``rho`` replaced (float32), everything else kept."""
    c64 = np.asarray(c, dtype=np.float64)
    r64 = np.asarray(rho, dtype=np.float64)
    if c64.shape != tuple(template.shape) or r64.shape != tuple(template.shape):
        raise ValueError(f"maps must have the grid's shape {template.shape}")
    maps = dict(template.maps)
    maps["K"] = (r64 * c64**2).astype(np.float32)
    maps["rho"] = r64.astype(np.float32)
    return dataclasses.replace(template, maps=maps)


def chain_to_c_rho(c, rho, gk, grho):
    """``(dJ/dc, dJ/drho)`` from ``(dJ/dK, dJ/drho at fixed K)`` with ``K = rho c^2``."""
    c = np.asarray(c, dtype=np.float64)
    rho = np.asarray(rho, dtype=np.float64)
    gk = np.asarray(gk, dtype=np.float64)
    return gk * 2.0 * rho * c, np.asarray(grho, dtype=np.float64) + gk * c**2


def misfit_gradient(template: SimulationInputs, c, rho, observed, *, checkpoint_every: int = 0,
                    device: int = 0) -> tuple[float, np.ndarray, np.ndarray, np.ndarray]:
    """L2 sensor misfit ``0.5 * sum((traces - observed)^2)`` of the medium ``(c, rho)`` and its
    gradient: returns ``(loss, dL/dc, dL/drho, traces)``."""
    inp = with_medium(template, c, rho)
    with HipAdjoint(inp, checkpoint_every=checkpoint_every, device=device) as adj:
        traces = adj.forward()
        resid = traces.astype(np.float64) - np.asarray(observed, dtype=np.float64)
        gk, gr = adj.backward(resid.astype(np.float32))
    gc, grho = chain_to_c_rho(c, rho, gk, gr)
    return 0.5 * float(np.sum(resid**2)), gc, grho, traces


def _torch():
    import torch

    return torch


class TraceFunction:
    """``traces = TraceFunction.apply(c, rho, template, checkpoint_every=0)``: sensor traces of
    the template simulation with the sound-speed and density maps ``c``, ``rho`` (torch tensors
    on the padded grid), differentiable with respect to both."""

    _fn = None

    @classmethod
    def _build(cls):
        torch = _torch()

        class _Fn(torch.autograd.Function):
            @staticmethod
            def forward(ctx, c, rho, template, checkpoint_every):
                c_np = c.detach().cpu().numpy()
                r_np = rho.detach().cpu().numpy()
                adj = HipAdjoint(with_medium(template, c_np, r_np), checkpoint_every=int(checkpoint_every))
                traces = adj.forward()
                ctx.adj = adj
                ctx.c_np, ctx.r_np = c_np, r_np
                ctx.dev, ctx.dtype = c.device, c.dtype
                return torch.from_numpy(traces).to(device=c.device)

            @staticmethod
            def backward(ctx, grad_out):
                w = grad_out.detach().cpu().numpy().astype(np.float32)
                gk, gr = ctx.adj.backward(w)
                ctx.adj.close()
                gc, grho = chain_to_c_rho(ctx.c_np, ctx.r_np, gk, gr)
                to = dict(device=ctx.dev, dtype=ctx.dtype)
                return torch.from_numpy(gc).to(**to), torch.from_numpy(grho).to(**to), None, None

        return _Fn

    @classmethod
    def apply(cls, c, rho, template: SimulationInputs, checkpoint_every: int = 0):
        if cls._fn is None:
            cls._fn = cls._build()
        return cls._fn.apply(c, rho, template, checkpoint_every)


__all__ = ["HipAdjoint", "TraceFunction", "chain_to_c_rho", "misfit_gradient", "with_medium"]
