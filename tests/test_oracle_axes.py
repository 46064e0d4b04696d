"""Per-axis relaxation maps reach the axis they name, in the oracle.

In the anisotropic relaxation model every axis has its own maps: u (velocity side) and x
(pressure side) for axis 0, w and y for axis 1, v and z for axis 2
(fullwave/medium.py:374-406, fullwave/solver/input_file_writer.py:592-620). The reference's front
end writes the same values for every axis, so its files cannot show which axis reads which map;
physics can. A plane wave travelling along one axis has no gradient along the others, so only
its own axis's memory variables act on it: with axis 0 lossless and axis 1 absorbing, a wave
along x keeps its amplitude while a wave along y loses a good part of it (3D: axis 2 absorbs
less than axis 1). The negative controls plant the two plausible mistakes in the oracle's inputs -- every
axis reading axis 0's maps, and the axes' maps swapped -- and the same check fails for both.
"""

import numpy as np
import pytest

import oracle
from tests.helpers import raw_inputs

LOSSLESS = {"kappa": 1.0, "d": (0.0, 0.0), "alpha": (1e6, 2e6)}
LOSSY = {"kappa": 1.0, "d": (3e6, 1.5e6), "alpha": (2e7, 3e7)}
MILD = {"kappa": 1.0, "d": (1e6, 5e5), "alpha": (2e7, 3e7)}
STEPS = {2: 100, 3: 60}
SHAPE = {2: (200, 200), 3: (96, 96, 96)}


def _pulse(shape, axis):
    """Unit Gaussian pulse along `axis`, constant across the others, zero on the ghost border."""
    n = shape[axis]
    g = np.exp(-(((np.arange(n) - n // 2) / 4.0) ** 2))
    p = np.moveaxis(np.broadcast_to(g, shape[:axis] + shape[axis + 1:] + (n,)), -1, axis).copy()
    inner = tuple(slice(8, m - 8) for m in shape)
    out = np.zeros(shape)
    out[inner] = p[inner]
    return out


def _peak(inp, axis):
    """Peak |p| of the pulse started along `axis` after the test's steps, in a band around the
    centre of the other axes (the grid edges' disturbance does not reach it in that time)."""
    sim = oracle.OracleSimulation(inp, precision=64)
    sim.set_field("p", _pulse(inp.shape, axis))
    sim.run(STEPS[inp.ndim])
    p = sim.field("p")
    band = tuple(slice(None) if a == axis else slice(m // 2 - 6, m // 2 + 6) for a, m in enumerate(inp.shape))
    return float(np.abs(p[band]).max())


def _inputs(ndim, per_axis):
    return raw_inputs(SHAPE[ndim], nt=STEPS[ndim], isotropic=False, relax={"v": per_axis, "p": per_axis})


def _axes_ok(peaks) -> bool:
    """Axis 0 keeps the split pulse's amplitude (0.5), axis 1 loses a good part of it, axis 2
    (3D) lies clearly between the two."""
    ok = peaks[0] > 0.45 and peaks[1] < 0.8 * peaks[0]
    if len(peaks) == 3:
        ok = ok and peaks[1] + 0.03 < peaks[2] < peaks[0] - 0.03
    return ok


@pytest.mark.parametrize("ndim", [2, 3])
def test_each_axis_reads_its_own_maps(ndim):
    per_axis = [LOSSLESS, LOSSY, MILD][:ndim]
    peaks = [_peak(_inputs(ndim, per_axis), a) for a in range(ndim)]
    print(f"{ndim}D peaks per axis: {peaks}")
    assert _axes_ok(peaks), peaks
    # the lossless axis: the pulse splits into two halves of (almost) unit shape
    assert peaks[0] < 0.55


@pytest.mark.parametrize("ndim", [2, 3])
def test_planted_axis_mistakes_fail(ndim):
    maps = [LOSSLESS, LOSSY, MILD][:ndim]
    mistakes = [[maps[0]] * ndim,                # every axis reading axis 0's maps
                [maps[1], maps[0]] + maps[2:]]   # axes 0 and 1 swapped
    if ndim == 3:
        mistakes += [[maps[0], maps[2], maps[1]],  # axes 1 and 2 swapped
                     [maps[0], maps[1], maps[1]]]  # axis 2 reading axis 1's maps
    for wrong in mistakes:
        assert not _axes_ok([_peak(_inputs(ndim, wrong), a) for a in range(ndim)])


def test_raw_inputs_write_each_axis():
    inp = raw_inputs((40, 40), isotropic=False, relax={"v": [LOSSLESS, LOSSY], "p": [LOSSY, LOSSLESS]})
    assert np.all(inp.maps["apmlu1"] == 0) and np.all(inp.maps["apmlw1"] < 0)
    assert np.all(inp.maps["apmlx1"] < 0) and np.all(inp.maps["apmly1"] == 0)
    # b = exp(-(d / kappa + alpha) dt): the absorbing axis's is the smaller
    assert np.all(inp.maps["bpmlw1"] < inp.maps["bpmlu1"]) and np.all(inp.maps["bpmlx1"] < inp.maps["bpmly1"])
