"""Plane-wave attenuation measurement for the attenuation-accuracy check.

A line source on the top row of a homogeneous medium (c = 1540 m/s, rho = 1000 kg/m^3, power-law
attenuation built by ``Medium.build``) launches a plane wave. Pressure is recorded on two rows at
depths z1 < z2 over the central columns, the direct arrival is gated at each depth with the same
Tukey window, and the attenuation is the log spectral ratio over the travel distance:

    alpha(f) = ln(|P(z1, f)| / |P(z2, f)|) / (z2 - z1)    [Np/m]

The grid is designed for 10 MHz (12 points per wavelength, CFL 0.4) so 1-20 MHz is resolved;
the source pulse is a Gaussian derivative whose spectrum covers that band; the source row spans
the full width, so edge diffraction reaches the centre long after the gate closes.
"""

from __future__ import annotations

import json
import tempfile
from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden" / "attenuation_modeling_reference.json"
NP_PER_DB_CM = 100.0 / (20.0 * np.log10(np.e))


def reference_panels() -> list[dict]:
    return json.loads(GOLDEN.read_text())["panels"]


def build_inputs(alpha0: float, gamma: float, *, f_design=10e6, nx=1100, ny=1024, z1=300, z2=900,
                 half_width=16, nt=2500):
    from fullwave25_amd.grid import Grid
    from fullwave25_amd.medium import Medium
    from fullwave25_amd.sensor import Sensor
    from fullwave25_amd.solver.solver import Solver
    from fullwave25_amd.source import Source

    c0 = 1540.0
    dx = c0 / f_design / 12
    dt = 0.4 * dx / c0
    grid = Grid(domain_size=(nx * dx, ny * dx), f0=f_design, duration=nt * dt, c0=c0, ppw=12, cfl=0.4)
    shape = (grid.nx, grid.ny)
    medium = Medium(grid, np.full(shape, c0), np.full(shape, 1000.0), np.full(shape, alpha0),
                    np.full(shape, gamma), np.zeros(shape))
    sigma = 1.0 / (2 * np.pi * 5e6)  # spectrum peaks near 5 MHz
    t = np.arange(grid.nt) * dt - 6 * sigma
    pulse = -(t / sigma) * np.exp(-0.5 * (t / sigma) ** 2)
    mask = np.zeros(shape, bool)
    mask[0, :] = True
    source = Source(np.tile(pulse, (shape[1], 1)), mask)
    smask = np.zeros(shape, bool)
    cols = slice(shape[1] // 2 - half_width, shape[1] // 2 + half_width + 1)
    smask[z1, cols] = True
    smask[z2, cols] = True
    sensor = Sensor(smask)
    with tempfile.TemporaryDirectory() as tmp:
        solver = Solver(Path(tmp), grid, medium, source, sensor)
        inputs, _ = solver.build_inputs()
    meta = {"dx": grid.dx, "dt": grid.dt, "z1": z1, "z2": z2, "t0": 6 * sigma, "n_per_row": 2 * half_width + 1}
    return inputs, meta


def attenuation_from_traces(traces: np.ndarray, meta: dict, freqs_mhz) -> np.ndarray:
    """Attenuation in Np/m at ``freqs_mhz`` from ``[n_steps, n_sensors]`` traces."""
    n = meta["n_per_row"]
    s1 = traces[:, :n].mean(axis=1)
    s2 = traces[:, n:].mean(axis=1)
    dt, dx = meta["dt"], meta["dx"]
    half = int(round(0.45e-6 / dt))
    spec = []
    for s, z in ((s1, meta["z1"]), (s2, meta["z2"])):
        k0 = int(round((meta["t0"] + z * dx / 1540.0) / dt))
        seg = np.zeros(2 * half)
        lo = max(k0 - half, 0)
        piece = s[lo : k0 + half]
        seg[: piece.size] = piece
        # Tukey window (25 % cosine tapers)
        w = np.ones(seg.size)
        taper = seg.size // 8
        ramp = 0.5 * (1 - np.cos(np.pi * np.arange(taper) / taper))
        w[:taper] = ramp
        w[-taper:] = ramp[::-1]
        nfft = 1 << 16
        spec.append(np.fft.rfft(seg * w, nfft))
    f = np.fft.rfftfreq(1 << 16, dt) / 1e6
    ratio = np.abs(spec[0]) / np.abs(spec[1])
    alpha = np.log(ratio) / ((meta["z2"] - meta["z1"]) * dx)
    return np.interp(np.asarray(freqs_mhz), f, alpha)


def target(alpha0, gamma, freqs_mhz):
    return alpha0 * np.asarray(freqs_mhz) ** gamma * NP_PER_DB_CM
