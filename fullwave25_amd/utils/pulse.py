"""Source waveforms.

``gaussian_modulated_sinusoidal_signal`` reproduces the waveform of
fullwave/utils/pulse.py:7-72: a sine at ``f0`` under a super-Gaussian envelope
``exp(-(1.05 * t * w0 / (ncycles * pi)) ** (2 * drop_off))``, on the time axis
``t = n / nt * duration - ncycles / f0 - delay``, optionally shifted by one cell transit time
(``dt / cfl``) per source layer so that a thick layered source emits a one-way wave.
"""

from __future__ import annotations

import numpy as np
from numpy.typing import NDArray


def gaussian_modulated_sinusoidal_signal(
    nt: int,
    duration: float,
    ncycles: int,
    drop_off: int,
    f0: float,
    p0: float,
    delay_sec: float = 0.0,
    i_layer: int | None = None,
    dt_for_layer_delay: float | None = None,
    cfl_for_layer_delay: float | None = None,
) -> NDArray[np.float64]:
    """Gaussian-windowed tone burst sampled on ``nt`` points over ``duration`` seconds."""
    # same operation order as the reference so the samples agree bit for bit
    t = np.arange(nt, dtype=np.float64) / nt * duration - ncycles / f0
    t = t - delay_sec
    if i_layer:
        if not dt_for_layer_delay:
            raise AssertionError("dt must be provided if i_layer is provided")
        if not cfl_for_layer_delay:
            raise AssertionError("cfl must be provided if i_layer is provided")
        t = t - i_layer * (dt_for_layer_delay / cfl_for_layer_delay)
    w0 = 2.0 * np.pi * f0
    envelope = np.exp(-((1.05 * t * w0 / (ncycles * np.pi)) ** (2 * drop_off)))
    return envelope * np.sin(w0 * t) * p0
