"""The C oracle against an independent numpy statement of the same update, on tiny grids,
and its internal consistency (FP32 vs FP64, isotropic vs per-axis maps)."""

import numpy as np
import pytest

import oracle
from fullwave25_amd import workloads
from tests.helpers import raw_inputs, rel_l2


@pytest.fixture(scope="module")
def small_2d():
    return workloads.random_medium_2d(30, 26, 90, seed=3, air=4).inputs


@pytest.fixture(scope="module")
def small_3d():
    return workloads.random_medium_3d(12, 11, 10, 40, seed=1, air=3).inputs


def _oracle(inp, n, precision=64):
    o = oracle.OracleSimulation(inp, precision)
    o.run(n)
    return o


def test_c_oracle_matches_numpy_2d(oracle_lib, small_2d):
    o = _oracle(small_2d, small_2d.nt)
    rec, p = oracle.brute_force_run(small_2d, small_2d.nt)
    assert np.abs(rec).max() > 0
    assert rel_l2(o.sensors(), rec) < 1e-13
    assert rel_l2(o.field("p"), p) < 1e-13


def test_c_oracle_matches_numpy_3d(oracle_lib, small_3d):
    o = _oracle(small_3d, small_3d.nt)
    rec, p = oracle.brute_force_run(small_3d, small_3d.nt)
    assert np.abs(rec).max() > 0
    assert rel_l2(o.sensors(), rec) < 1e-13
    assert rel_l2(o.field("p"), p) < 1e-13


def test_fp32_oracle_close_to_fp64(oracle_lib, small_2d):
    a = _oracle(small_2d, small_2d.nt, 64).sensors()
    b = _oracle(small_2d, small_2d.nt, 32).sensors()
    assert rel_l2(b, a) < 1e-5


def test_anisotropic_maps_equal_isotropic(oracle_lib):
    """The reference ramps every per-axis copy identically, so the per-axis (anisotropic) maps
    hold the isotropic values and must give the same result."""
    iso = workloads.random_medium_2d(24, 22, 50, seed=5, isotropic=True).inputs
    ani = workloads.random_medium_2d(24, 22, 50, seed=5, isotropic=False).inputs
    assert set(ani.maps) > set(iso.maps)
    for ax in ("w",):
        np.testing.assert_array_equal(ani.maps[f"apml{ax}1"], iso.maps["apmlu1"])
    a = _oracle(iso, iso.nt).sensors()
    b = _oracle(ani, ani.nt).sensors()
    np.testing.assert_array_equal(a, b)


def test_source_air_and_sensor_order(oracle_lib):
    """Injection, then pressure release, then recording: a sensor on a source records the
    injected sample, a sensor on an air point records 0, a repeated source keeps its last row."""
    shape = (24, 24)
    icc = [[12, 12], [12, 12], [14, 10], [15, 15]]
    sig = np.stack([np.full(20, 1.0), np.full(20, 2.0), np.full(20, 3.0), np.full(20, 4.0)])
    inp = raw_inputs(shape, nt=20, icc=icc, icmat=sig, outc=[[12, 12], [14, 10], [15, 15], [10, 10]],
                     icczero=[[15, 15]])
    o = _oracle(inp, 20)
    rec = o.sensors()
    np.testing.assert_allclose(rec[:, 0], 2.0)
    np.testing.assert_allclose(rec[:, 1], 3.0)
    np.testing.assert_allclose(rec[:, 2], 0.0)
    assert np.abs(rec[5:, 3]).max() > 0
    rb, _ = oracle.brute_force_run(inp, 20)
    np.testing.assert_allclose(rec, rb, rtol=1e-13, atol=1e-15)


def test_modt_records(oracle_lib, small_2d):
    import dataclasses

    inp = dataclasses.replace(small_2d, modt=7)
    o = _oracle(inp, inp.nt)
    full = _oracle(small_2d, small_2d.nt).sensors()
    rec = o.sensors()
    assert rec.shape[0] == -(-inp.nt // 7)
    np.testing.assert_array_equal(rec, full[::7])


def test_ghost_border_stays_zero(oracle_lib, small_2d):
    p = _oracle(small_2d, small_2d.nt).field("p")
    assert np.all(p[:8] == 0) and np.all(p[-8:] == 0)
    assert np.all(p[:, :8] == 0) and np.all(p[:, -8:] == 0)


@pytest.mark.parametrize("is_3d", [False, True])
def test_exponential_model_c_oracle_matches_numpy(oracle_lib, is_3d):
    """The exponential-attenuation model (decay map with the radial absorbing mask, air points,
    sources): C oracle against the numpy statement."""
    from fullwave25_amd import workloads

    if is_3d:
        inp = workloads.random_medium_3d(12, 10, 11, 30, seed=4, air=2, exponential=True).inputs
    else:
        inp = workloads.random_medium_2d(30, 26, 60, seed=3, air=3, exponential=True).inputs
    assert inp.exponential and "a_exp" in inp.maps and "kappax" not in inp.maps
    o = oracle.OracleSimulation(inp, 64)
    o.run(inp.nt)
    rec, pf = oracle.brute_force_run(inp, inp.nt)
    assert np.abs(rec).max() > 0
    assert rel_l2(o.sensors(), rec) < 1e-12
    assert rel_l2(o.field("p"), pf) < 1e-12
