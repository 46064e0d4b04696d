"""The front end's engine inputs against the reference's own output (golden vectors written by
tests/golden/make_frontend_golden.py from the reference front end).

Every ``.dat`` file the reference's InputFileWriter writes for a 2D and a 3D case, isotropic and
per-axis relaxation and the exponential-attenuation model, is compared with what this package's
Solver path produces: padding, absorbing-layer ramps, relaxation maps, the exponential model's
decay map with its radial absorbing mask, coefficient tables, coordinate lists, scalars and the
source matrix.
Every file must match byte for byte.
"""

import tempfile
from pathlib import Path

import numpy as np
import pytest

from fullwave25_amd.grid import Grid
from fullwave25_amd.medium import Medium, MediumRelaxationMaps
from fullwave25_amd.sensor import Sensor
from fullwave25_amd.solver.inputs import build_simulation_inputs
from fullwave25_amd.solver.pml_builder import PMLBuilder, PMLBuilderExponentialAttenuation
from fullwave25_amd.source import Source
from fullwave25_amd.utils.pulse import gaussian_modulated_sinusoidal_signal

GOLD = Path(__file__).resolve().parent / "golden" / "frontend_golden.npz"
INT_FILES = {"icc", "outc", "icczero", "nX", "nY", "nZ", "nT", "ncoords", "ncoordsout", "ncoordszero",
             "nTic", "modT", "ndmap", "dcmap"}


@pytest.fixture(scope="module")
def gold():
    return dict(np.load(GOLD, allow_pickle=False))


def _grid(shape, nt):
    f0, c0, ppw, cfl = 3e6, 1540.0, 12, 0.4
    dx = c0 / f0 / ppw
    dt = cfl * dx / c0
    return Grid(domain_size=tuple(s * dx for s in shape), f0=f0, duration=nt * dt, c0=c0, ppw=ppw, cfl=cfl)


def _our_files(gold, name, iso):
    g = {k[len(name) + 4:]: v for k, v in gold.items() if k.startswith(f"{name}/in/")}
    shape = tuple(int(v) for v in g["shape"])
    relax = {k[len("relax/"):]: v for k, v in g.items() if k.startswith("relax/")}
    grid = _grid(shape, int(g["nt"]))
    medium = MediumRelaxationMaps(grid, g["c"], g["rho"], g["beta"], relax, air_map=g["air"],
                                  use_isotropic_relaxation=iso)
    pml = PMLBuilder(grid, medium, Source(g["p0"], g["src"]), Sensor(g["sen"], int(g["modT"])),
                     m_spatial_order=8, n_pml_layer=int(g["n_pml"]), n_transition_layer=int(g["n_transition"]),
                     use_isotropic_relaxation=iso)
    ext = pml.run(use_pml=True)
    inp = build_simulation_inputs(pml.extended_grid, ext, pml.extended_source, pml.extended_sensor, isotropic=iso)
    with tempfile.TemporaryDirectory() as tmp:
        inp.write_dir(Path(tmp))
        return {f.stem: np.fromfile(f, dtype=np.uint8) for f in Path(tmp).glob("*.dat")}


def _our_files_exp(gold, name):
    g = {k[len(name) + 4:]: v for k, v in gold.items() if k.startswith(f"{name}/in/")}
    shape = tuple(int(v) for v in g["shape"])
    grid = _grid(shape, int(g["nt"]))
    medium = Medium(grid, g["c"], g["rho"], g["alpha_coeff"], g["alpha_power"], g["beta"], air_map=g["air"])
    pml = PMLBuilderExponentialAttenuation(grid, medium, Source(g["p0"], g["src"]), Sensor(g["sen"], int(g["modT"])),
                                           m_spatial_order=8, n_pml_layer=int(g["n_pml"]))
    ext = pml.run(use_pml=True)
    inp = build_simulation_inputs(pml.extended_grid, ext, pml.extended_source, pml.extended_sensor)
    assert inp.exponential
    with tempfile.TemporaryDirectory() as tmp:
        inp.write_dir(Path(tmp))
        return {f.stem: np.fromfile(f, dtype=np.uint8) for f in Path(tmp).glob("*.dat")}


@pytest.mark.parametrize("case", ["2d_iso", "2d_aniso", "3d_iso", "3d_aniso", "2d_exp", "3d_exp"])
def test_engine_inputs_match_reference(gold, case):
    name, kind = case.split("_")
    ours = _our_files_exp(gold, name) if kind == "exp" else _our_files(gold, name, kind == "iso")
    theirs = {k[len(case) + 1:]: v for k, v in gold.items() if k.startswith(case + "/")}
    assert theirs, "no golden files"
    missing = sorted(set(theirs) - set(ours))
    assert not missing, f"files the reference writes but this package does not: {missing}"
    for stem, ref_bytes in sorted(theirs.items()):
        got = ours[stem]
        dtype = np.int32 if stem in INT_FILES else np.float32
        np.testing.assert_array_equal(got.view(dtype), ref_bytes.view(dtype), err_msg=stem)


def test_pulse_matches_reference(gold):
    for i, (nt, dur, nc, do, f0, p0, dl, il) in enumerate([
        (500, 5e-6, 3, 2, 3e6, 1e5, 0.0, None),
        (777, 7.3e-6, 2, 1, 2.5e6, 2.0, 3e-7, 2),
    ]):
        kw = dict(delay_sec=dl)
        if il:
            kw.update(i_layer=il, dt_for_layer_delay=1.1e-8, cfl_for_layer_delay=0.4)
        np.testing.assert_array_equal(gaussian_modulated_sinusoidal_signal(nt, dur, nc, do, f0, p0, **kw), gold[f"pulse{i}"])
