"""bench.py's host-side legs, on the CPU: the CPU baseline (clamped to the workload's nt), the
x-slab sample it times for large grids, the argument handling of the headline configs, and
that a line cannot be lost to a failing CPU leg."""

import json
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

from fullwave25_amd import workloads

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
import bench  # noqa: E402


def test_cpu_baseline_clamps_steps_to_nt(oracle_lib):
    inp = workloads.random_medium_2d(24, 20, 12, seed=3).inputs
    assert inp.nt == 12
    out = bench.cpu_baseline(inp, 150)
    assert out["value"] > 0 and out["kind"] == "port" and out["cores"] >= 1
    assert "steps 6..12" in out["sample"]


def test_cpu_baseline_on_a_slab(oracle_lib):
    inp = workloads.random_medium_3d(30, 20, 18, 9, seed=2).inputs
    out = bench.cpu_baseline(inp, 100, slab_planes=20)
    assert out["value"] > 0
    assert "x-slab of 20 planes" in out["sample"]


def test_x_slab_keeps_and_shifts_points():
    inp = workloads.random_medium_3d(30, 20, 18, 9, seed=2, n_src=6, n_sensors=40).inputs
    x0 = 9
    s = bench.x_slab(inp, 20, x0=x0)
    assert s.shape == (20, inp.ny, inp.nz)
    for k, v in inp.maps.items():
        np.testing.assert_array_equal(s.maps[k], v[x0:x0 + 20])
    np.testing.assert_array_equal(s.dcmap, inp.dcmap[x0:x0 + 20])
    inside = (inp.icc[:, 0] >= x0) & (inp.icc[:, 0] < x0 + 20)
    np.testing.assert_array_equal(s.icc, inp.icc[inside] - np.array([x0, 0, 0]))
    np.testing.assert_array_equal(s.icmat, inp.icmat[inside])
    sin = (inp.outc[:, 0] >= x0) & (inp.outc[:, 0] < x0 + 20)
    np.testing.assert_array_equal(s.outc, inp.outc[sin] - np.array([x0, 0, 0]))
    s.validate()
    # default placement: centred on the first source plane, inside the grid
    d = bench.x_slab(inp, 20)
    assert d.nx == 20 and d.ncoords >= 1


def test_updated_cells():
    inp = workloads.random_medium_2d(24, 20, 5).inputs
    assert bench.updated_cells(inp) == (inp.nx - 16) * (inp.ny - 16)


def test_more_than_one_gpu_is_refused():
    res = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                         timeout=120)
    assert res.returncode == 2
    assert "one GPU" in json.loads(res.stdout.strip().splitlines()[-1])["error"]


@pytest.mark.parametrize("config,shape", [("hetero_3d", (768, 768, 640)), ("focused_2d", (4096, 2048, 1))])
def test_headline_configs_have_their_full_sizes(config, shape):
    c = bench.SINGLE_CONFIGS[config]
    assert (c["nx"], c["ny"], c["nz"]) == shape
    assert c["nt"] >= 6000


def test_workload_nt_does_not_follow_steps(monkeypatch):
    """The driver's short run (--steps 20) must still build the config's full-length job."""
    seen = {}

    def fake(nx, ny, nz, nt):
        seen["nt"] = nt
        return None

    monkeypatch.setattr(workloads, "inclusions_3d", fake)

    class A:
        config, nx, ny, nz, steps, warmup = "hetero_3d", 8, 8, 8, 20, 5

    bench.build_single(A())
    assert seen["nt"] == 6000
