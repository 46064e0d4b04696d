"""fw_run_dir_noalloc, the directory contract run without any device allocation, against
fw_run_dir on the same directories (bit for bit with the same kernel shapes) and against the
FP64 oracle (rel-L2 <= 1e-4), twice in a row on one GPU with the second run smaller than the
first, so anything the first run left in the shared device array would show."""

import numpy as np
import pytest

import oracle
from fullwave25_amd import engine, workloads
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-4
SHAPES = {"2d": "stream:256:8", "3d": "pipelite:512:8"}


def _cases(dim):
    if dim == "2d":
        return [workloads.random_medium_2d(90, 84, 160, seed=91, air=4).inputs,
                workloads.random_medium_2d(40, 36, 120, seed=92, n_src=3).inputs]
    return [workloads.random_medium_3d(30, 28, 26, 80, seed=93, air=3).inputs,
            workloads.random_medium_3d(20, 18, 22, 60, seed=94).inputs]


def _genout(d, inp):
    return np.fromfile(d / "genout.dat", dtype=np.float32).reshape(-1, inp.ncoordsout)


@pytest.mark.parametrize("dim", ["2d", "3d"])
def test_noalloc_matches_run_dir_on_two_calls(tmp_path, monkeypatch, dim):
    monkeypatch.setenv("FULLWAVE_HIP_KERNELS", SHAPES[dim])
    for n, inp in enumerate(_cases(dim)):
        a, b = tmp_path / f"alloc_{n}", tmp_path / f"noalloc_{n}"
        inp.write_dir(a)
        inp.write_dir(b)
        print(f"{dim} call {n}: grid {inp.shape}, {inp.nt} steps, kernels {SHAPES[dim]}", flush=True)
        engine.run_dir(a)
        engine.run_dir_noalloc(b)
        want, got = _genout(a, inp), _genout(b, inp)
        assert got.shape == (inp.n_records, inp.ncoordsout)
        assert np.abs(want).max() > 0
        np.testing.assert_array_equal(got, want)
        ref = oracle.OracleSimulation(inp, precision=64)
        ref.run(inp.nt)
        err = rel_l2(got, ref.sensors())
        print(f"{dim} call {n}: rel-L2 vs FP64 oracle {err:.3e}", flush=True)
        assert err <= TOL


def test_noalloc_default_kernels(tmp_path, monkeypatch):
    """Without FULLWAVE_HIP_KERNELS the allocation-free run takes the default shapes (not timed)
    and still meets the oracle."""
    monkeypatch.delenv("FULLWAVE_HIP_KERNELS", raising=False)
    inp = workloads.random_medium_2d(50, 44, 100, seed=95, air=2).inputs
    inp.write_dir(tmp_path / "run")
    engine.run_dir_noalloc(tmp_path / "run")
    ref = oracle.OracleSimulation(inp, precision=64)
    ref.run(inp.nt)
    assert rel_l2(_genout(tmp_path / "run", inp), ref.sensors()) <= TOL
