"""The plane-wise builder for large grids gives the same engine inputs as the Solver path,
byte for byte (2D and 3D, isotropic and per-axis, with air points, sources and sensors)."""

import tempfile
from pathlib import Path

import numpy as np
import pytest

from fullwave25_amd import workloads
from fullwave25_amd.medium import Medium
from fullwave25_amd.sensor import Sensor
from fullwave25_amd.solver.fast_inputs import build_inputs_planewise, ramp_ops
from fullwave25_amd.solver.pml_builder import _LayerShape, apply_transition_and_pml
from fullwave25_amd.solver.solver import Solver
from fullwave25_amd.source import Source


def _case(shape, seed, iso):
    rng = np.random.default_rng(seed)
    g = workloads._grid(*shape, nt=30) if len(shape) == 2 else workloads._grid(shape[0], shape[1], nt=30, nz=shape[2])
    sh = (g.nx, g.ny, g.nz) if g.is_3d else (g.nx, g.ny)
    c = np.round(1450 + 170 * rng.random(sh), 1)
    rho = 900 + 200 * rng.random(sh)
    a0 = np.round(0.3 + 1.2 * rng.random(sh), 1)
    gm = np.round(0.8 + 0.8 * rng.random(sh), 1)
    med = Medium(g, c, rho, a0, gm, np.zeros(sh), air_map=rng.random(sh) < 0.01, use_isotropic_relaxation=iso)
    mask = np.zeros(sh, bool)
    mask.reshape(-1)[rng.choice(mask.size, 5, replace=False)] = True
    smask = np.zeros(sh, bool)
    smask.reshape(-1)[rng.choice(smask.size, 9, replace=False)] = True
    return g, med, Source(rng.standard_normal((5, g.nt)), mask), Sensor(smask, sampling_modulus_time=2)


@pytest.mark.parametrize("shape,iso", [((30, 26), True), ((30, 26), False), ((12, 10, 14), True), ((12, 10, 14), False)])
def test_planewise_equals_solver_path(shape, iso):
    g, med, src, sen = _case(shape, 7 + len(shape) + iso, iso)
    with tempfile.TemporaryDirectory() as t:
        ref, _ = Solver(Path(t), g, med, src, sen, pml_layer_thickness_px=6 if len(shape) == 2 else 4,
                        n_transition_layer=4 if len(shape) == 2 else 3, use_isotropic_relaxation=iso).build_inputs()
    fast = build_inputs_planewise(g, med, src, sen, n_pml=6 if len(shape) == 2 else 4,
                                  n_transition=4 if len(shape) == 2 else 3, isotropic=iso)
    assert set(fast.maps) == set(ref.maps) - {"c", "beta"}
    for k, v in fast.maps.items():
        assert v.dtype == ref.maps[k].dtype
        np.testing.assert_array_equal(v.view(np.uint8), ref.maps[k].view(np.uint8), err_msg=k)
    for k in ("dcmap", "dmap", "d", "icc", "icmat", "outc", "icczero"):
        np.testing.assert_array_equal(getattr(fast, k), getattr(ref, k), err_msg=k)
    for k in ("nx", "ny", "nz", "nt", "dx", "dy", "dz", "dt", "c0", "modt", "isotropic"):
        assert getattr(fast, k) == getattr(ref, k), k


@pytest.mark.parametrize("kind,region", [("polynomial", "both"), ("linear", "both"), ("cosine", "transition")])
def test_ramp_ops_reproduce_the_ramp(kind, region):
    rng = np.random.default_rng(3)
    values = rng.random((40, 3))
    want = apply_transition_and_pml(values.copy(), 7.5, 0, _LayerShape(8, 6, 5), kind=kind, region=region)
    mode, src, f = ramp_ops(40, 8, 6, 5, kind=kind, region=region)
    base = values[src]
    got = np.where((mode == 1)[:, None], base - f[:, None] * (base - 7.5), base)
    got = np.where((mode == 2)[:, None], 7.5, got)
    np.testing.assert_array_equal(got, want)
