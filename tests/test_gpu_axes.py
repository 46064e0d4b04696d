"""Per-axis relaxation maps on the GPU: every kernel family against the FP64 oracle on inputs
whose axes carry different maps (anisotropic model; fullwave/medium.py:374-406,
fullwave/solver/input_file_writer.py:592-620).

The reference's front end writes the same values for every axis, so inputs built through it
cannot show a kernel reading one axis's maps for another; here axis 1 and axis 2 get axis 0's
maps mirrored (tests.helpers.distinct_axis_maps). The oracle's own axis mapping is pinned by
tests/test_oracle_axes.py. Tolerance: rel-L2 of the traces and of the final pressure field
<= 1e-4.
"""

import functools

import numpy as np
import pytest

import oracle
from fullwave25_amd import engine, workloads
from tests.helpers import distinct_axis_maps, rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-4

FAMILIES = {
    2: ["percell", "stream:128:8", "stream:256:32", "tile:8", "tile:16", "tile:32"],
    3: ["percell", "stream:512:8", "stream:1024:16", "pipe:512:8", "pipe:1024:16", "pipelite:512:8",
        "pipelite:1024:32"],
}


@functools.lru_cache(maxsize=None)
def _case(ndim):
    if ndim == 2:
        base = workloads.random_medium_2d(52, 72, 180, seed=71, isotropic=False, air=3).inputs
    else:
        base = workloads.random_medium_3d(22, 26, 24, 100, seed=72, isotropic=False, air=2).inputs
    return base, distinct_axis_maps(base)


@functools.lru_cache(maxsize=None)
def _oracle(ndim, distinct=True):
    inp = _case(ndim)[1 if distinct else 0]
    o = oracle.OracleSimulation(inp, 64)
    o.run(inp.nt)
    return o.sensors(), o.field("p")


@pytest.mark.parametrize("ndim", [2, 3])
def test_mirrored_maps_change_the_answer(ndim):
    """The test inputs have teeth: with the axes' maps made distinct the traces move far more
    than the tolerance, so a kernel reading axis 0's maps on every axis cannot pass."""
    base, inp = _case(ndim)
    for stem in ("kappaw", "apmly1", "bpmlw2"):
        assert not np.array_equal(inp.maps[stem], base.maps[stem])
    a, _ = _oracle(ndim, True)
    b, _ = _oracle(ndim, False)
    print(f"{ndim}D: rel-L2 distinct vs identical per-axis maps = {rel_l2(a, b):.3e}", flush=True)
    assert rel_l2(a, b) > 100 * TOL


@pytest.mark.parametrize("ndim,shape", [(d, s) for d in (2, 3) for s in FAMILIES[d]])
def test_family_reads_each_axis_maps(monkeypatch, ndim, shape):
    _, inp = _case(ndim)
    if shape.startswith("tile"):
        assert inp.ny % 4 == 0
    monkeypatch.setenv("FULLWAVE_HIP_KERNELS", shape)
    print(f"{ndim}D per-axis maps, kernels {shape}: grid {inp.shape}, {inp.nt} steps", flush=True)
    with engine.HipSimulation(inp) as sim:
        sim.run(inp.nt)
        got, p_got = sim.sensors().astype(np.float64), sim.field("p").astype(np.float64)
        cfg = sim.kernel_config()
    want, p_want = _oracle(ndim)
    assert np.abs(want).max() > 0
    err, ferr = rel_l2(got, want), rel_l2(p_got, p_want)
    print(f"{ndim}D {shape} ({cfg['velocity']['variant']}): rel-L2 traces {err:.3e}, field {ferr:.3e}", flush=True)
    assert err <= TOL and ferr <= TOL
