"""The drop-in on the reference's own files: simulation directories holding the exact bytes the
reference's InputFileWriter wrote (tests/golden/frontend_golden.npz: 2D and 3D, isotropic and
per-axis relaxation, exponential attenuation) run through HipLauncher.run -> fw_run_dir, and the
static-map layout (fullwave/solver/input_file_writer.py:156-170, 647-714; the maps written once
to the work directory and linked from every txrx_N directory) through Solver.run.

Bar: genout.dat against the FP64 oracle run on the same directory's inputs, rel-L2 <= 1e-4.
"""

from pathlib import Path

import numpy as np
import pytest

import oracle
from fullwave25_amd import workloads
from fullwave25_amd.solver.inputs import SimulationInputs
from fullwave25_amd.solver.launcher import HipLauncher
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-4
GOLD = Path(__file__).resolve().parent / "golden" / "frontend_golden.npz"
CASES = ["2d_iso", "2d_aniso", "3d_iso", "3d_aniso", "2d_exp", "3d_exp"]


@pytest.mark.parametrize("case", CASES)
def test_reference_written_directory_runs(tmp_path, case):
    gold = np.load(GOLD, allow_pickle=False)
    sim_dir = tmp_path / case
    sim_dir.mkdir()
    for key in gold.files:
        if key.startswith(case + "/"):
            gold[key].tofile(sim_dir / f"{key[len(case) + 1:]}.dat")  # the reference's bytes, as written
    inp = SimulationInputs.read_dir(sim_dir)
    print(f"{case}: reference-written directory, grid {inp.shape}, {inp.nt} steps, {inp.ncoords} sources, "
          f"{inp.ncoordsout} sensors", flush=True)
    flat = HipLauncher(is_3d=inp.ndim == 3).run(sim_dir)
    got = np.asarray(flat, dtype=np.float64).reshape(-1, inp.ncoordsout)
    o = oracle.OracleSimulation(inp, 64)
    o.run(inp.nt)
    want = o.sensors()
    assert got.shape == want.shape
    assert np.abs(want).max() > 0
    err = rel_l2(got, want)
    print(f"{case}: genout.dat rel-L2 vs FP64 oracle {err:.3e}", flush=True)
    assert err <= TOL


def test_static_map_transmits_through_solver(tmp_path, monkeypatch):
    """Two transmits on one medium the reference's way: the first writes the maps to the work
    directory (is_static_map=True) and links them from txrx_0, the second only writes its source
    matrix into txrx_1 and links the same maps (recalculate_pml=False). Each equals the same
    transmit run in memory (kernel shapes pinned, so bit for bit) and meets the FP64 oracle."""
    import fullwave25_amd as fw

    monkeypatch.setenv("FULLWAVE_HIP_KERNELS", "stream:256:8")
    from fullwave25_amd.utils.pulse import gaussian_modulated_sinusoidal_signal

    grid = workloads._grid(64, 56, nt=180)
    shape = (grid.nx, grid.ny)
    rng = np.random.default_rng(81)
    c = 1500 + 80 * rng.random(shape)
    medium = fw.Medium(grid, c, np.full(shape, 1000.0), np.full(shape, 0.6), np.full(shape, 1.1), np.zeros(shape))
    mask = np.zeros(shape, bool)
    mask[0, 10:46] = True
    smask = np.zeros(shape, bool)
    smask[2, ::3] = True
    sensor = fw.Sensor(smask)
    outs, sources = [], []
    for tx, delay in enumerate((0.0, 2e-7)):
        sig = gaussian_modulated_sinusoidal_signal(grid.nt, grid.duration, 2, 2, 3e6, 1e5, delay_sec=delay)
        ramp = np.linspace(0.5, 1.5, int(mask.sum()))[:, None]  # a different apodization per transmit
        src = fw.Source((sig[None, :] * (ramp if tx == 0 else ramp[::-1])), mask)
        sources.append(src)
        solver = fw.Solver(tmp_path, grid, medium, src, sensor, pml_layer_thickness_px=10, n_transition_layer=8,
                           write_input_files=True)
        outs.append(solver.run(f"txrx_{tx}", is_static_map=True, recalculate_pml=tx == 0))
    for tx in (0, 1):
        d = tmp_path / f"txrx_{tx}"
        assert (d / "K.dat").is_symlink() and (d / "K.dat").resolve() == (tmp_path / "K.dat").resolve()
        assert (d / "dcmap.dat").is_symlink() and not (d / "icmat.dat").is_symlink()
    assert not np.array_equal(np.fromfile(tmp_path / "txrx_0" / "icmat.dat", np.float32),
                              np.fromfile(tmp_path / "txrx_1" / "icmat.dat", np.float32))
    for tx in (0, 1):
        # the same transmit in memory (no files), and the FP64 oracle on its inputs
        solver = fw.Solver(tmp_path / f"mem_{tx}", grid, medium, sources[tx], sensor, pml_layer_thickness_px=10,
                           n_transition_layer=8)
        np.testing.assert_array_equal(outs[tx], solver.run())
        inp, _ = solver.build_inputs()
        o = oracle.OracleSimulation(inp, 64)
        o.run(inp.nt)
        err = rel_l2(outs[tx].T, o.sensors())
        print(f"static-map transmit {tx}: rel-L2 vs FP64 oracle {err:.3e}", flush=True)
        assert err <= TOL
    assert rel_l2(outs[0], outs[1]) > 1e-2  # the transmits differ

