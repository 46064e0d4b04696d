"""Synthetic workloads built through the public front end (Grid, Medium, Source, Sensor,
Solver), at the sizes of the benchmark configurations and at small sizes for checks.

Every builder returns a :class:`~fullwave25_amd.solver.inputs.SimulationInputs` (plus, where a
transmit is involved, what is needed to post-process the traces). Random media use
``numpy.random.default_rng(seed)``.
"""

from __future__ import annotations

import tempfile
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from .grid import Grid
from .medium import Medium
from .sensor import Sensor
from .solver.solver import Solver
from .source import Source
from .utils.pulse import gaussian_modulated_sinusoidal_signal


@dataclass
class Workload:
    name: str
    inputs: object  # SimulationInputs
    element_of_sensor: np.ndarray | None = None  # per-sensor element index (-1 = none)
    n_elements: int = 0

    @property
    def cells(self) -> int:
        return self.inputs.ncells


def _grid(nx: int, ny: int, *, f0=3e6, c0=1540.0, ppw=12, cfl=0.4, nt=2000, nz: int | None = None) -> Grid:
    dx = c0 / f0 / ppw
    dt = cfl * dx / c0
    size = (nx * dx, ny * dx) if nz is None else (nx * dx, ny * dx, nz * dx)
    return Grid(domain_size=size, f0=f0, duration=nt * dt, c0=c0, ppw=ppw, cfl=cfl)


def _inputs(grid, medium, source, sensor, *, n_pml: int, n_transition: int | None = None,
            isotropic: bool = True, exponential: bool = False):
    n_tr = grid.ppw * 3 if n_transition is None else n_transition
    if not exponential and n_pml > 0 and n_tr > 0:
        # the plane-wise builder gives the Solver path's inputs byte for byte, faster and in
        # far less memory on large grids
        from .solver.fast_inputs import build_inputs_planewise

        medium.use_isotropic_relaxation = isotropic
        return build_inputs_planewise(grid, medium, source, sensor, n_pml=n_pml, n_transition=n_tr,
                                      isotropic=isotropic)
    with tempfile.TemporaryDirectory() as tmp:
        solver = Solver(Path(tmp), grid, medium, source, sensor, pml_layer_thickness_px=n_pml,
                        n_transition_layer=n_transition, use_isotropic_relaxation=isotropic,
                        use_exponential_attenuation=exponential)
        inputs, _ = solver.build_inputs()
    return inputs


def plane_wave_2d(nx: int = 512, ny: int = 512, nt: int = 2000, *, n_pml: int = 20,
                  n_transition: int | None = None, sensor_depth: int | None = 400,
                  alpha0: float = 0.5, gamma: float = 1.1, ncycles: int = 3,
                  isotropic: bool = True, f0: float = 3e6) -> Workload:
    """Homogeneous 2D medium (c = 1540 m/s, rho = 1000 kg/m^3, alpha0, gamma), a 3-cycle
    Gaussian tone burst on the top row as a line source, a line of receivers at
    ``sensor_depth`` (the 512 x 512 / 2000-step plane-wave configuration at default sizes)."""
    grid = _grid(nx, ny, nt=nt, f0=f0)
    shape = (grid.nx, grid.ny)
    medium = Medium(grid, np.full(shape, 1540.0), np.full(shape, 1000.0), np.full(shape, alpha0),
                    np.full(shape, gamma), np.zeros(shape), use_isotropic_relaxation=isotropic)
    mask = np.zeros(shape, dtype=bool)
    mask[0, :] = True
    sig = gaussian_modulated_sinusoidal_signal(grid.nt, grid.duration, ncycles, 2, f0, 1e5)
    source = Source(np.tile(sig, (int(mask.sum()), 1)), mask)
    smask = np.zeros(shape, dtype=bool)
    smask[min(sensor_depth if sensor_depth is not None else nx // 2, nx - 1), :] = True
    sensor = Sensor(smask)
    return Workload("plane_wave_2d", _inputs(grid, medium, source, sensor, n_pml=n_pml,
                                             n_transition=n_transition, isotropic=isotropic))


def plane_wave_3d(nx: int = 256, ny: int = 256, nz: int = 256, nt: int = 1000, *, n_pml: int = 12,
                  n_transition: int = 8, alpha0: float = 0.5, gamma: float = 1.1, f0: float = 3e6) -> Workload:
    """Homogeneous 3D medium with a plane-wave transmit from the x = 0 face (a 3-cycle tone
    burst on every point of it) and a receive plane at mid-depth."""
    grid = _grid(nx, ny, nt=nt, f0=f0, nz=nz)
    shape = (grid.nx, grid.ny, grid.nz)
    medium = Medium(grid, np.full(shape, 1540.0), np.full(shape, 1000.0), np.full(shape, alpha0),
                    np.full(shape, gamma), np.zeros(shape))
    mask = np.zeros(shape, dtype=bool)
    mask[0] = True
    sig = gaussian_modulated_sinusoidal_signal(grid.nt, grid.duration, 3, 2, f0, 1e5)
    source = Source(np.tile(sig, (int(mask.sum()), 1)), mask)
    smask = np.zeros(shape, dtype=bool)
    smask[nx // 2, ::4, ::4] = True
    return Workload("plane_wave_3d", _inputs(grid, medium, source, Sensor(smask), n_pml=n_pml,
                                             n_transition=n_transition))


def inclusions_3d(nx: int = 768, ny: int = 768, nz: int = 640, nt: int = 6000, *, n_inclusions: int = 20,
                  seed: int = 2028, n_pml: int = 24, n_transition: int = 36, f0: float = 2.5e6,
                  n_el: int = 32, ntic: int = 400, record_every: int = 1, radii=(20, 120)) -> Workload:
    """3D heterogeneous medium with ellipsoidal inclusions and a matrix-array plane-wave transmit
    (the 768 x 768 x 640 / 6000-step configuration at default sizes; 904 x 904 x 776 cells with
    the 24-cell absorbing layer, 36-cell transition and ghost border).

    Background c = 1540 m/s, rho = 1000 kg/m^3, alpha0 = 0.5 dB/(MHz cm); gamma varies with depth
    from 0.7 to 1.5 (steps of 0.1); ``n_inclusions`` seeded ellipsoids (radii ``radii`` cells)
    with c ~ U[1450, 1620], alpha0 ~ U[0.3, 1.5] (steps of 0.1) and rho within +-3 %. The
    transmit: a ``n_el`` x ``n_el`` array (pitch one wavelength, one-cell kerf) on the first
    interior plane, every element firing the same 3-cycle pulse (``ntic`` samples). Receivers:
    the full-aperture receive plane one plane deeper (every y-z point of the interior) and the
    three orthogonal mid-planes x = nx/2, y = ny/2, z = nz/2 (every point; 2,060,802 distinct
    sensors at the default size), recorded every ``record_every`` steps. Built plane by plane
    (:func:`~fullwave25_amd.solver.fast_inputs.build_inputs_planewise`).
    """
    from .solver.fast_inputs import build_inputs_planewise

    grid = _grid(nx, ny, nt=nt, f0=f0, nz=nz)
    shape = (grid.nx, grid.ny, grid.nz)
    rng = np.random.default_rng(seed)
    c = np.full(shape, 1540.0)
    rho = np.full(shape, 1000.0)
    a0 = np.full(shape, 0.5)
    gm = np.empty(shape)
    gm[:] = np.round(0.7 + 0.8 * np.arange(nx) / max(nx - 1, 1), 1)[:, None, None]
    for _ in range(n_inclusions):
        ctr = [rng.uniform(0.1, 0.9) * n for n in shape]
        rad = [rng.uniform(*radii) for _ in range(3)]
        lo = [max(0, int(ctr[a] - rad[a])) for a in range(3)]
        hi = [min(shape[a], int(ctr[a] + rad[a]) + 1) for a in range(3)]
        if any(h <= l for l, h in zip(lo, hi)):
            continue
        ax = [((np.arange(lo[a], hi[a]) - ctr[a]) / rad[a]) ** 2 for a in range(3)]
        inside = ax[0][:, None, None] + ax[1][None, :, None] + ax[2][None, None, :] <= 1.0
        box = tuple(slice(lo[a], hi[a]) for a in range(3))
        c[box] = np.where(inside, rng.uniform(1450, 1620), c[box])
        rho[box] = np.where(inside, 1000.0 * (1 + rng.uniform(-0.03, 0.03)), rho[box])
        a0[box] = np.where(inside, np.round(rng.uniform(0.3, 1.5), 1), a0[box])
    medium = Medium(grid, c, rho, a0, gm, np.zeros(shape))
    pitch = int(round(grid.wavelength / grid.dx))
    width = pitch - 1
    el = np.zeros(ny, bool)
    el_z = np.zeros(nz, bool)
    y0, z0 = (ny - n_el * pitch) // 2, (nz - n_el * pitch) // 2
    for e in range(n_el):
        el[max(0, y0 + e * pitch): max(0, y0 + e * pitch + width)] = True
        el_z[max(0, z0 + e * pitch): max(0, z0 + e * pitch + width)] = True
    face = el[:, None] & el_z[None, :]
    mask = np.zeros(shape, bool)
    mask[0] = face
    sig = gaussian_modulated_sinusoidal_signal(ntic, ntic * grid.dt, 3, 2, f0, 1e5)
    source = Source(np.broadcast_to(sig, (int(face.sum()), ntic)), mask)
    smask = np.zeros(shape, bool)
    smask[1] = True
    smask[nx // 2] = True
    smask[:, ny // 2] = True
    smask[:, :, nz // 2] = True
    sensor = Sensor(smask, sampling_modulus_time=record_every)
    inputs = build_inputs_planewise(grid, medium, source, sensor, n_pml=n_pml, n_transition=n_transition)
    return Workload("inclusions_3d", inputs)


def layered_medium_2d(shape, rng, *, n_layers=6, amp_cells=10.0, dx=None, scatter_per_cell=15,
                      wavelength_cells=12):
    """Seeded layered medium: ``n_layers`` horizontal layers whose interfaces carry a random
    sinusoidal perturbation of ``amp_cells``; per-layer c ~ U[1450, 1620], rho ~ U[900, 1100],
    alpha0 ~ U[0.3, 1.5], gamma ~ U[0.8, 1.6]; sub-resolution scatterers at
    ``scatter_per_cell`` per resolution cell (wavelength^2) with a +-2 % Gaussian density
    fluctuation."""
    nx, ny = shape
    c = np.empty(shape)
    rho = np.empty(shape)
    a0 = np.empty(shape)
    gm = np.empty(shape)
    props = np.stack([rng.uniform(1450, 1620, n_layers), rng.uniform(900, 1100, n_layers),
                      rng.uniform(0.3, 1.5, n_layers), rng.uniform(0.8, 1.6, n_layers)], 1)
    depth = np.arange(nx)[:, None].astype(np.float64)
    lat = np.arange(ny)[None, :].astype(np.float64)
    layer = np.zeros(shape, dtype=np.int64)
    bounds = np.linspace(0, nx, n_layers + 1)[1:-1]
    for b in bounds:
        k = rng.uniform(1, 4) * 2 * np.pi / ny
        ph = rng.uniform(0, 2 * np.pi)
        iface = b + amp_cells * np.sin(k * lat + ph)
        layer += (depth >= iface).astype(np.int64)
    for i in range(n_layers):
        sel = layer == i
        c[sel], rho[sel], a0[sel], gm[sel] = props[i]
    density = scatter_per_cell / float(wavelength_cells**2)
    scat = rng.random(shape) < density
    rho = rho * np.where(scat, 1.0 + 0.02 * rng.standard_normal(shape), 1.0)
    return c, rho, a0, gm


def _array_geometry(ny: int, n_elements: int, pitch_cells: int):
    width = pitch_cells - 1
    aperture = n_elements * pitch_cells
    y0 = (ny - aperture) // 2
    elem_of_col = np.full(ny, -1)
    for e in range(n_elements):
        elem_of_col[y0 + e * pitch_cells : y0 + e * pitch_cells + width] = e
    centres = np.array([y0 + e * pitch_cells + (width - 1) / 2 for e in range(n_elements)])
    return elem_of_col, centres


def focused_signals(grid, coords, elem_of_col, centres, *, focus_m: float, lateral_m: float = 0.0,
                    f0: float = 3e6, ncycles: int = 3) -> np.ndarray:
    """Source signals ``[n_points, nt]`` of a transmit focused at depth ``focus_m`` and lateral
    offset ``lateral_m`` from the array centre, for source points ``coords`` (rows of a thick
    source get the one-cell transit delay of their row)."""
    yc = (grid.ny - 1) / 2
    dist = np.sqrt(focus_m**2 + ((centres - yc) * grid.dx - lateral_m) ** 2)
    delay = (dist.max() - dist) / grid.c0
    p0 = np.zeros((coords.shape[0], grid.nt))
    cache = {}
    for r, (ix, iy) in enumerate(coords):
        e = elem_of_col[iy]
        key = (e, ix)
        if key not in cache:
            cache[key] = gaussian_modulated_sinusoidal_signal(
                grid.nt, grid.duration, ncycles, 2, f0, 1e5, delay_sec=delay[e],
                i_layer=int(ix), dt_for_layer_delay=grid.dt, cfl_for_layer_delay=grid.cfl)
        p0[r] = cache[key]
    return p0


def focused_linear_array_2d(nx: int = 4096, ny: int = 2048, nt: int = 8000, *, n_elements: int = 128,
                            pitch_cells: int = 12, focus_m: float = 40e-3, seed: int = 2026,
                            n_pml: int = 36, layers: int = 3, isotropic: bool = True,
                            f0: float = 3e6, ncycles: int = 3) -> Workload:
    """Heterogeneous 2D medium with a focused transmit from a linear array
    (the 4096 x 2048 / 8000-step configuration at default sizes).

    The array sits on the top ``layers`` rows (a thick source with one-cell transit delays per
    row, as the reference's examples build it), element width ``pitch_cells - 1`` (one-cell
    kerf), focused at ``focus_m`` on the centre line. Receivers are the element faces one row
    below the source layers; traces are averaged per element afterwards.
    """
    grid = _grid(nx, ny, nt=nt, f0=f0)
    shape = (grid.nx, grid.ny)
    rng = np.random.default_rng(seed)
    c, rho, a0, gm = layered_medium_2d(shape, rng, wavelength_cells=int(round(grid.wavelength / grid.dx)))
    medium = Medium(grid, c, rho, a0, gm, np.zeros(shape), use_isotropic_relaxation=isotropic)
    elem_of_col, centres = _array_geometry(ny, n_elements, pitch_cells)
    cols = np.nonzero(elem_of_col >= 0)[0]
    mask = np.zeros(shape, dtype=bool)
    mask[:layers, cols] = True
    coords = np.argwhere(mask)
    p0 = focused_signals(grid, coords, elem_of_col, centres, focus_m=focus_m, f0=f0, ncycles=ncycles)
    source = Source(p0, mask)
    smask = np.zeros(shape, dtype=bool)
    smask[layers, cols] = True
    sensor = Sensor(smask)
    inputs = _inputs(grid, medium, source, sensor, n_pml=n_pml, isotropic=isotropic)
    elem = elem_of_col[np.argwhere(smask)[:, 1]]
    return Workload("focused_linear_array_2d", inputs, element_of_sensor=elem, n_elements=n_elements)


def steered_transmits_2d(nx: int = 2048, ny: int = 1536, nt: int = 3000, *, n_transmits: int = 128,
                         n_elements: int = 128, pitch_cells: int = 12, focus_m: float = 30e-3,
                         seed: int = 2027, n_pml: int = 36, layers: int = 3, f0: float = 3e6,
                         transmits=None):
    """Many focused transmits on one seeded heterogeneous medium (the batched-transmit
    configuration: 2048 x 1536, 128 transmits, each focused at ``focus_m`` and steered to its own
    lateral position, spread evenly across the aperture). Returns ``(workload, shots)``:
    ``workload.inputs`` carries the first transmit built, ``shots`` the source matrices
    ``[n, n_points, nt]`` (float32) in the order of the inputs' source points, of every transmit
    or, with ``transmits`` (indices into the ``n_transmits``), of those only."""
    grid = _grid(nx, ny, nt=nt, f0=f0)
    shape = (grid.nx, grid.ny)
    rng = np.random.default_rng(seed)
    c, rho, a0, gm = layered_medium_2d(shape, rng, wavelength_cells=int(round(grid.wavelength / grid.dx)))
    medium = Medium(grid, c, rho, a0, gm, np.zeros(shape))
    elem_of_col, centres = _array_geometry(ny, n_elements, pitch_cells)
    cols = np.nonzero(elem_of_col >= 0)[0]
    mask = np.zeros(shape, dtype=bool)
    mask[:layers, cols] = True
    coords = np.argwhere(mask)
    half = 0.5 * n_elements * pitch_cells * grid.dx
    laterals = np.linspace(-0.8 * half, 0.8 * half, n_transmits)
    if transmits is not None:
        laterals = laterals[np.asarray(transmits, dtype=np.int64)]
    shots = np.stack([focused_signals(grid, coords, elem_of_col, centres, focus_m=focus_m, lateral_m=x, f0=f0)
                      for x in laterals]).astype(np.float32)
    smask = np.zeros(shape, dtype=bool)
    smask[layers, cols] = True
    inputs = _inputs(grid, medium, Source(shots[0].astype(np.float64), mask), Sensor(smask), n_pml=n_pml)
    elem = elem_of_col[np.argwhere(smask)[:, 1]]
    return Workload("steered_transmits_2d", inputs, element_of_sensor=elem, n_elements=n_elements), shots


def misfit_media_2d(nx: int = 1024, ny: int = 1024, nt: int = 2000, *, seed: int = 2029,
                    smooth_cells: float = 12.0, n_pml: int = 20, f0: float = 3e6):
    """The misfit-gradient job of the served mix: a seeded "true" heterogeneous medium
    (:func:`layered_medium_2d`) and a smoothed initial model of it (c and rho filtered with a
    Gaussian of ``smooth_cells``; the attenuation maps are the true ones), both with a 3-cycle
    tone burst on the top row and a line of receivers two rows below it (reflection geometry, as
    with an array on the skin). Returns
    ``(initial, true)`` workloads; the true medium's traces are the observed data the initial
    model is fitted to."""
    from scipy.ndimage import gaussian_filter

    grid = _grid(nx, ny, nt=nt, f0=f0)
    shape = (grid.nx, grid.ny)
    rng = np.random.default_rng(seed)
    c, rho, a0, gm = layered_medium_2d(shape, rng, wavelength_cells=int(round(grid.wavelength / grid.dx)))
    mask = np.zeros(shape, dtype=bool)
    mask[0, :] = True
    sig = gaussian_modulated_sinusoidal_signal(grid.nt, grid.duration, 3, 2, f0, 1e5)
    smask = np.zeros(shape, dtype=bool)
    smask[2, :] = True
    out = []
    for cc, rr in ((gaussian_filter(c, smooth_cells), gaussian_filter(rho, smooth_cells)), (c, rho)):
        medium = Medium(grid, cc, rr, a0, gm, np.zeros(shape))
        source = Source(np.tile(sig, (int(mask.sum()), 1)), mask)
        out.append(Workload("misfit_media_2d", _inputs(grid, medium, source, Sensor(smask), n_pml=n_pml)))
    return out[0], out[1]


def element_traces(traces: np.ndarray, workload: Workload) -> np.ndarray:
    """Average ``[n_records, n_sensors]`` traces per array element -> ``[n_elements, n_records]``."""
    e = workload.element_of_sensor
    out = np.zeros((workload.n_elements, traces.shape[0]))
    for k in range(workload.n_elements):
        out[k] = traces[:, e == k].mean(axis=1)
    return out


def random_medium_2d(nx: int, ny: int, nt: int, *, seed: int = 0, n_pml: int = 6, n_transition: int = 4,
                     isotropic: bool = True, n_src: int = 5, n_sensors: int = 17, air: int = 0,
                     exponential: bool = False) -> Workload:
    """Small seeded heterogeneous 2D problem for parity checks: random smooth c/rho/alpha/gamma,
    a few point sources with distinct pulses, scattered sensors and optional air points
    (``exponential``: the exponential-attenuation model instead of relaxation)."""
    grid = _grid(nx, ny, nt=nt)
    shape = (grid.nx, grid.ny)
    rng = np.random.default_rng(seed)
    from scipy.ndimage import gaussian_filter

    def field(lo, hi):
        f = gaussian_filter(rng.random(shape), 2.0)
        f = (f - f.min()) / (np.ptp(f) + 1e-12)
        return lo + (hi - lo) * f

    c = field(1450, 1620)
    rho = field(900, 1100)
    # a few attenuation levels keep the number of distinct relaxation fits small
    a0 = np.round(field(0.3, 1.5), 1)
    gm = np.round(field(0.8, 1.6), 1)
    air_map = np.zeros(shape, dtype=bool)
    if air:
        idx = rng.choice(nx * ny, size=air, replace=False)
        air_map.reshape(-1)[idx] = True
    medium = Medium(grid, c, rho, a0, gm, np.zeros(shape), air_map=air_map, use_isotropic_relaxation=isotropic)
    mask = np.zeros(shape, dtype=bool)
    mask.reshape(-1)[rng.choice(nx * ny, size=n_src, replace=False)] = True
    p0 = np.stack([gaussian_modulated_sinusoidal_signal(grid.nt, grid.duration, 2, 2, 3e6, 1e5,
                                                         delay_sec=1e-7 * i) for i in range(n_src)])
    source = Source(p0, mask)
    smask = np.zeros(shape, dtype=bool)
    smask.reshape(-1)[rng.choice(nx * ny, size=n_sensors, replace=False)] = True
    sensor = Sensor(smask, sampling_modulus_time=1)
    return Workload("random_medium_2d", _inputs(grid, medium, source, sensor, n_pml=n_pml,
                                                n_transition=n_transition, isotropic=isotropic,
                                                exponential=exponential))


def random_medium_3d(nx: int, ny: int, nz: int, nt: int, *, seed: int = 0, n_pml: int = 4,
                     n_transition: int = 3, isotropic: bool = True, n_src: int = 4,
                     n_sensors: int = 13, air: int = 0, exponential: bool = False) -> Workload:
    """Small seeded heterogeneous 3D problem for parity checks."""
    grid = _grid(nx, ny, nt=nt, nz=nz)
    shape = (grid.nx, grid.ny, grid.nz)
    rng = np.random.default_rng(seed)
    from scipy.ndimage import gaussian_filter

    def field(lo, hi):
        f = gaussian_filter(rng.random(shape), 1.5)
        f = (f - f.min()) / (np.ptp(f) + 1e-12)
        return lo + (hi - lo) * f

    air_map = np.zeros(shape, dtype=bool)
    if air:
        air_map.reshape(-1)[rng.choice(air_map.size, size=air, replace=False)] = True
    medium = Medium(grid, field(1450, 1620), field(900, 1100), np.round(field(0.3, 1.5), 1),
                    np.round(field(0.7, 1.5), 1),
                    np.zeros(shape), air_map=air_map, use_isotropic_relaxation=isotropic)
    n = int(np.prod(shape))
    mask = np.zeros(shape, dtype=bool)
    mask.reshape(-1)[rng.choice(n, size=n_src, replace=False)] = True
    p0 = np.stack([gaussian_modulated_sinusoidal_signal(grid.nt, grid.duration, 2, 2, 3e6, 1e5,
                                                         delay_sec=1e-7 * i) for i in range(n_src)])
    source = Source(p0, mask)
    smask = np.zeros(shape, dtype=bool)
    smask.reshape(-1)[rng.choice(n, size=n_sensors, replace=False)] = True
    sensor = Sensor(smask)
    return Workload("random_medium_3d", _inputs(grid, medium, source, sensor, n_pml=n_pml,
                                                n_transition=n_transition, isotropic=isotropic,
                                                exponential=exponential))
