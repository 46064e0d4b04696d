"""Coded relaxation coefficients (16-bit code per cell + tuple table) against the per-cell maps.

The coded kernels read the maps' own values from the table, so every kernel family must give
bit-identical traces and fields with codes on and off; the oracle comparisons in
test_gpu_parity.py run with codes on (the default).
"""

import dataclasses

import numpy as np
import pytest

import oracle
from fullwave25_amd import engine, workloads
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu


def _run(inp, codes: bool, n=None, graph_steps=0):
    with engine.HipSimulation(inp) as sim:
        sim.set_relaxation_codes(codes)
        if graph_steps:
            sim.set_graph_steps(graph_steps)
        ncodes = sim.relaxation_codes()
        sim.run(inp.nt if n is None else n)
        return ncodes, sim.sensors().copy(), sim.field("p").copy()


def test_codes_in_use_only_for_isotropic_relaxation():
    iso = workloads.random_medium_2d(60, 52, 10, seed=41, air=4).inputs
    aniso = workloads.random_medium_2d(44, 70, 10, seed=42, isotropic=False).inputs
    expo = workloads.random_medium_2d(50, 46, 10, seed=43, exponential=True).inputs
    with engine.HipSimulation(iso) as sim:
        n = sim.relaxation_codes()
        print(f"isotropic 2D: {n} codes", flush=True)
        assert 1 <= n <= 65535
    for inp in (aniso, expo):
        with engine.HipSimulation(inp) as sim:
            assert sim.relaxation_codes() == 0


def _tuple_counts(inp):
    """How many cells hold each distinct 11-word tuple the encoder sees (the ten isotropic
    relaxation maps, with 1/kappa as the device holds it, and dcmap), host cells only."""
    names = ["kappau", "apmlu1", "apmlu2", "bpmlu1", "bpmlu2", "kappax", "apmlx1", "apmlx2", "bpmlx1", "bpmlx2"]
    cols = []
    for k in names:
        m = np.ascontiguousarray(inp.maps[k], np.float32).reshape(-1)
        if k.startswith("kappa"):  # the device holds 1/kappa (IEEE float32 division)
            m = np.where(m != 0, np.float32(1) / np.where(m != 0, m, np.float32(1)), np.float32(0)).astype(np.float32)
        cols.append(m.view(np.uint32))
    cols.append(np.ascontiguousarray(inp.dcmap, np.int32).reshape(-1).view(np.uint32))
    st = np.ascontiguousarray(np.stack(cols, 1))
    rows = st.view(np.dtype((np.void, st.shape[1] * 4))).reshape(-1)
    return np.unique(rows, return_counts=True)[1]


def test_codes_count_matches_distinct_tuples():
    inp = workloads.random_medium_3d(30, 26, 41, 10, seed=44).inputs
    want = len(_tuple_counts(inp))
    with engine.HipSimulation(inp) as sim:
        got = sim.relaxation_codes()
    # 3D rows are padded on the device to 32-float multiples; padding cells add the all-zero tuple
    print(f"distinct tuples {want}, codes {got}", flush=True)
    assert got in (want, want + 1)


@pytest.mark.parametrize("shape", ["stream:256:8", "stream:64:3", "tile:8", "tile:32", "tile:8/stream:256:8"])
def test_coded_2d_bitwise_equal_maps(monkeypatch, shape):
    for inp in (workloads.random_medium_2d(70, 300, 120, seed=31, air=5).inputs,
                workloads.random_medium_2d(53, 532, 60, seed=37, n_pml=7, n_transition=5).inputs):
        if shape.startswith("tile") and inp.ny % 4:
            continue
        monkeypatch.setenv("FULLWAVE_HIP_KERNELS", shape)
        print(f"{shape}: {inp.shape}", flush=True)
        n, a, pa = _run(inp, True)
        n0, b, pb = _run(inp, False)
        assert n > 0 and n0 == 0
        assert np.abs(a).max() > 0
        np.testing.assert_array_equal(a, b)
        np.testing.assert_array_equal(pa, pb)


@pytest.mark.parametrize("shape", ["pipe:512:7", "pipe:1024:16", "pipelite:512:2", "pipelite:1024:9", "pipelite:512:32"])
def test_coded_3d_bitwise_equal_maps(monkeypatch, shape):
    inp = workloads.random_medium_3d(40, 30, 150, 60, seed=33, air=4).inputs
    monkeypatch.setenv("FULLWAVE_HIP_KERNELS", shape)
    print(f"{shape}: {inp.shape}", flush=True)
    n, a, pa = _run(inp, True)
    n0, b, pb = _run(inp, False)
    assert n > 0 and n0 == 0
    assert np.abs(a).max() > 0
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(pa, pb)


def test_codes_switched_off_by_environment(monkeypatch):
    inp = workloads.random_medium_2d(40, 40, 10, seed=45).inputs
    monkeypatch.setenv("FULLWAVE_HIP_CODES", "0")
    with engine.HipSimulation(inp) as sim:
        assert sim.relaxation_codes() == 0
        sim.set_relaxation_codes(True)  # the call overrides the environment, also after set-up
        assert sim.relaxation_codes() > 0


def _perturbed(inp, seed, region=None):
    """Copy of ``inp`` whose apmlu1 map gets a distinct value in (nearly) every cell of
    ``region`` (a boolean mask; all cells if None)."""
    rng = np.random.default_rng(seed)
    maps = dict(inp.maps)
    f = 1 + 0.05 * rng.random(inp.shape)
    if region is not None:
        f = np.where(region, f, 1.0)
    maps["apmlu1"] = (inp.maps["apmlu1"] * f).astype(np.float32)
    return dataclasses.replace(inp, maps=maps)


def _oracle_traces(inp):
    # long enough runs: the sensors must see the wave itself, not only the stencil's precursor
    o = oracle.OracleSimulation(inp, 64)
    o.run(inp.nt)
    return o.sensors()


def test_code_coverage_matches_exact_tuple_counts():
    """The encoder counts a wavefront whose lanes hold one tuple in registers across its
    iterations and the others with atomics; on a 6.4 M-cell grid (more cells than the encoder
    has lanes) with more distinct tuples than codes, the cells its 65535 codes cover are exactly
    those of the 65535 most frequent tuples counted on the host."""
    base = workloads.random_medium_2d(2600, 2400, 10, seed=50).inputs
    region = np.zeros(base.shape, bool)
    region[:40] = True
    inp = _perturbed(base, 50, region)
    cnt = _tuple_counts(inp)
    assert len(cnt) > 65535
    want = np.sort(cnt)[::-1][:65535].sum() / cnt.sum()
    with engine.HipSimulation(inp) as sim:
        n, frac = sim.relaxation_coding()
    print(f"{len(cnt)} tuples over {cnt.sum()} cells: coverage {frac:.9f}, host counts {want:.9f}", flush=True)
    assert n == 65535
    assert abs(frac - want) < 1e-12


@pytest.mark.parametrize("shape", ["stream:256:8", "tile:8", "tile:16/stream:64:4"])
def test_cells_without_codes_read_the_maps_2d(monkeypatch, shape):
    """More distinct tuples than codes: the most frequent tuples are coded, the rest of the
    cells read the per-cell maps, with the same results."""
    inp = _perturbed(workloads.random_medium_2d(300, 280, 320, seed=46).inputs, 46)
    monkeypatch.setenv("FULLWAVE_HIP_KERNELS", shape)
    with engine.HipSimulation(inp) as sim:
        n, frac = sim.relaxation_coding()
    print(f"{shape}: {n} codes cover {frac:.3f} of the cells", flush=True)
    assert n == 65535 and 0.5 <= frac < 1.0
    _, a, pa = _run(inp, True)
    _, b, pb = _run(inp, False)
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(pa, pb)
    assert rel_l2(a.astype(np.float64), _oracle_traces(inp)) <= 1e-4


@pytest.mark.parametrize("shape", ["pipe:512:7", "pipelite:1024:9"])
def test_cells_without_codes_read_the_maps_3d(monkeypatch, shape):
    base = workloads.random_medium_3d(60, 50, 80, 50, seed=48).inputs
    region = np.zeros(base.shape, bool)
    region[:25] = True
    inp = _perturbed(base, 48, region)
    monkeypatch.setenv("FULLWAVE_HIP_KERNELS", shape)
    with engine.HipSimulation(inp) as sim:
        n, frac = sim.relaxation_coding()
    print(f"{shape}: {n} codes cover {frac:.3f} of the cells", flush=True)
    assert n == 65535 and 0.5 <= frac < 1.0
    _, a, pa = _run(inp, True)
    _, b, pb = _run(inp, False)
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(pa, pb)


def test_low_coverage_keeps_the_maps():
    """Codes that would cover less than half of the cells are not used."""
    inp = _perturbed(workloads.random_medium_2d(380, 360, 320, seed=49).inputs, 49)
    with engine.HipSimulation(inp) as sim:
        assert sim.relaxation_coding() == (0, 0.0)
        sim.run(inp.nt)
        got = sim.sensors().astype(np.float64)
    assert rel_l2(got, _oracle_traces(inp)) <= 1e-4


@pytest.mark.parametrize("graph_steps", [0, 8])
def test_map_upload_after_run_recodes(monkeypatch, graph_steps):
    """A relaxation map uploaded between runs takes effect in the coded kernels (and in a
    captured graph) exactly as in a fresh handle built with that map."""
    monkeypatch.setenv("FULLWAVE_HIP_KERNELS", "tile:8/stream:256:8")
    inp = workloads.random_medium_2d(64, 256, 80, seed=47).inputs
    maps = dict(inp.maps)
    maps["bpmlx2"] = (inp.maps["bpmlx2"] * 0.97).astype(np.float32)
    changed = dataclasses.replace(inp, maps=maps)
    with engine.HipSimulation(inp) as sim:
        if graph_steps:
            sim.set_graph_steps(graph_steps)
        n_before = sim.relaxation_codes()
        sim.run(40)
        sim.reset()
        sim.upload("bpmlx2", maps["bpmlx2"])
        assert n_before > 0 and sim.relaxation_codes() > 0
        sim.run(inp.nt)
        got = sim.sensors().copy()
    _, want, _ = _run(changed, True, graph_steps=graph_steps)
    _, want_maps, _ = _run(changed, False)
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(got, want_maps)
