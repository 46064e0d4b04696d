"""Plane-wave attenuation measurement for the attenuation-accuracy check.

A line source on the top row of a homogeneous medium (c = 1540 m/s, rho = 1000 kg/m^3, power-law
attenuation built by ``Medium.build``) launches a plane wave. Pressure is recorded on two rows at
depths z1 < z2 over the central columns, the direct arrival is gated at each depth with the same
Tukey window, and the attenuation is the log spectral ratio over the travel distance:

    alpha(f) = ln(|P(z1, f)| / |P(z2, f)|) / (z2 - z1)    [Np/m]

The grid is designed for 10 MHz (12 points per wavelength, CFL 0.4) so 1-20 MHz is resolved;
the source pulse is a Gaussian derivative with its spectral peak at 8 MHz, which keeps 1 MHz and
20 MHz within 15 dB of the peak. The separation z2 - z1 is chosen per (alpha0, gamma) so that the
attenuation at 20 MHz over it is about 30 dB (between 150 and 900 cells): long enough to measure
weak attenuation, short enough that strongly attenuated high frequencies stay far above the
window's spectral leakage. The gate is 4 us long (1 MHz is resolved), the domain is wide enough
that the diffraction from the ends of the line source reaches the sensors only after the gate
closes, and deep enough that nothing from the bottom layer returns within it.
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


GATE_HALF_S = 2.0e-6


def geometry(alpha0: float, gamma: float, dx: float) -> dict:
    """Depths, domain size and step count for one (alpha0, gamma)."""
    db_per_cell_20mhz = alpha0 * 20.0**gamma * dx * 100.0
    d = int(np.clip(round(30.0 / db_per_cell_20mhz), 150, 900))
    z1 = 120
    z2 = z1 + d
    gate = int(np.ceil(1540.0 * GATE_HALF_S / dx))  # cells travelled during half a gate
    half_width = int(np.ceil(1.1 * np.sqrt((z2 + gate) ** 2 - z2**2)))
    return {"z1": z1, "z2": z2, "nx": z2 + gate + 100, "ny": 2 * half_width, "gate_cells": gate}


def build_inputs(alpha0: float, gamma: float, *, f_design=10e6, half_width=16):
    from fullwave25_amd.grid import Grid
    from fullwave25_amd.medium import Medium
    from fullwave25_amd.sensor import Sensor
    from fullwave25_amd.solver.solver import Solver
    from fullwave25_amd.source import Source

    c0 = 1540.0
    dx = c0 / f_design / 12
    dt = 0.4 * dx / c0
    geo = geometry(alpha0, gamma, dx)
    nx, ny, z1, z2 = geo["nx"], geo["ny"], geo["z1"], geo["z2"]
    sigma = 1.0 / (2 * np.pi * 8e6)  # spectrum peaks at 8 MHz
    t0 = 6 * sigma
    nt = int(np.ceil((t0 + z2 * dx / c0 + GATE_HALF_S) / dt)) + 20
    grid = Grid(domain_size=(nx * dx, ny * dx), f0=f_design, duration=nt * dt, c0=c0, ppw=12, cfl=0.4)
    shape = (grid.nx, grid.ny)
    medium = Medium(grid, np.full(shape, c0), np.full(shape, 1000.0), np.full(shape, alpha0),
                    np.full(shape, gamma), np.zeros(shape))
    t = np.arange(grid.nt) * dt - t0
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
    meta = {"dx": grid.dx, "dt": grid.dt, "z1": z1, "z2": z2, "t0": t0, "n_per_row": 2 * half_width + 1}
    return inputs, meta


def attenuation_from_traces(traces: np.ndarray, meta: dict, freqs_mhz) -> np.ndarray:
    """Attenuation in Np/m at ``freqs_mhz`` from ``[n_steps, n_sensors]`` traces."""
    n = meta["n_per_row"]
    s1 = traces[:, :n].mean(axis=1)
    s2 = traces[:, n:].mean(axis=1)
    dt, dx = meta["dt"], meta["dx"]
    half = int(round(GATE_HALF_S / dt))
    nfft = 1 << 16
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
        spec.append(np.fft.rfft(seg * w, nfft))
    f = np.fft.rfftfreq(nfft, dt) / 1e6
    ratio = np.abs(spec[0]) / np.abs(spec[1])
    alpha = np.log(ratio) / ((meta["z2"] - meta["z1"]) * dx)
    return np.interp(np.asarray(freqs_mhz), f, alpha)


def target(alpha0, gamma, freqs_mhz):
    return alpha0 * np.asarray(freqs_mhz) ** gamma * NP_PER_DB_CM
