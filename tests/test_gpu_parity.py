"""HIP engine vs the FP64 oracle, through the C ABI (needs an MI355X).

Tolerance: relative L2 error of the sensor traces <= 1e-4 (FP32 engine vs FP64 oracle on the
same float32 inputs).
"""

import dataclasses

import numpy as np
import pytest

import oracle
from fullwave25_amd import engine, workloads
from tests.helpers import raw_inputs, rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-4


def _gpu(inp, n=None):
    with engine.HipSimulation(inp) as sim:
        sim.run(inp.nt if n is None else n)
        return sim.sensors().astype(np.float64), sim.field("p").astype(np.float64)


def _ref(inp, n=None):
    o = oracle.OracleSimulation(inp, 64)
    o.run(inp.nt if n is None else n)
    return o.sensors(), o.field("p")


CASES = {
    "2d_iso": lambda: workloads.random_medium_2d(60, 52, 300, seed=11, air=6).inputs,
    "2d_aniso": lambda: workloads.random_medium_2d(44, 70, 200, seed=12, isotropic=False).inputs,
    "2d_ragged": lambda: workloads.random_medium_2d(37, 131, 150, seed=13, n_src=9, n_sensors=40).inputs,
    "3d_iso": lambda: workloads.random_medium_3d(20, 18, 22, 120, seed=14, air=4).inputs,
    "3d_aniso": lambda: workloads.random_medium_3d(16, 21, 19, 90, seed=15, isotropic=False).inputs,
    "2d_exp": lambda: workloads.random_medium_2d(50, 46, 260, seed=16, air=5, exponential=True).inputs,
    "3d_exp": lambda: workloads.random_medium_3d(18, 20, 17, 100, seed=17, air=3, exponential=True).inputs,
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_engine_matches_fp64_oracle(case):
    inp = CASES[case]()
    print(f"case {case}: shape {inp.shape}, nt {inp.nt}", flush=True)
    got, p_got = _gpu(inp)
    want, p_want = _ref(inp)
    assert np.abs(want).max() > 0
    assert got.shape == want.shape
    err = rel_l2(got, want)
    print(f"case {case}: rel-L2 traces {err:.3e}, field {rel_l2(p_got, p_want):.3e}", flush=True)
    assert err <= TOL
    assert rel_l2(p_got, p_want) <= TOL


def test_plane_wave_512_full_length():
    """The 512 x 512, 2000-step homogeneous plane-wave configuration at full size."""
    inp = workloads.plane_wave_2d().inputs
    print(f"plane wave: shape {inp.shape}, nt {inp.nt}", flush=True)
    got, _ = _gpu(inp)
    want, _ = _ref(inp)
    err = rel_l2(got, want)
    print(f"plane wave 512: rel-L2 {err:.3e}", flush=True)
    assert err <= TOL


def test_source_air_sensor_semantics():
    shape = (24, 24)
    icc = [[12, 12], [12, 12], [14, 10], [15, 15]]
    sig = np.stack([np.full(20, 1.0), np.full(20, 2.0), np.full(20, 3.0), np.full(20, 4.0)])
    inp = raw_inputs(shape, nt=20, icc=icc, icmat=sig, outc=[[12, 12], [14, 10], [15, 15], [10, 10]],
                     icczero=[[15, 15]])
    got, _ = _gpu(inp)
    want, _ = _ref(inp)
    np.testing.assert_allclose(got[:, :3], want[:, :3], rtol=0, atol=0)
    assert rel_l2(got, want) <= TOL


def test_deterministic():
    inp = workloads.random_medium_2d(40, 40, 100, seed=21).inputs
    a, _ = _gpu(inp)
    b, _ = _gpu(inp)
    np.testing.assert_array_equal(a, b)


def test_run_in_pieces_and_reset():
    inp = workloads.random_medium_2d(40, 36, 120, seed=22).inputs
    full, _ = _gpu(inp)
    with engine.HipSimulation(inp) as sim:
        sim.run(50)
        sim.run(70)
        pieces = sim.sensors().astype(np.float64)
        sim.reset()
        assert sim.step == 0
        sim.run(120)
        again = sim.sensors().astype(np.float64)
    np.testing.assert_array_equal(pieces, full)
    np.testing.assert_array_equal(again, full)


def test_modt_and_record_count():
    inp = dataclasses.replace(workloads.random_medium_2d(30, 30, 95, seed=23).inputs, modt=4)
    got, _ = _gpu(inp)
    assert got.shape[0] == -(-95 // 4)
    want, _ = _ref(inp)
    assert rel_l2(got, want) <= TOL


def test_reciprocity_on_gpu():
    """Initial-pressure reciprocity p_{A->B} / K_B = p_{B->A} / K_A (uniform c, varying rho,
    no absorbing layer): holds when the velocity and pressure stencils are adjoint."""
    shape = (40, 44)
    rng = np.random.default_rng(5)
    rho = 1000 + 200 * rng.random(shape)
    inp = raw_inputs(shape, rho=rho, nt=60)
    K = inp.maps["K"].astype(np.float64)
    A, B = (15, 17), (24, 29)
    out = []
    for src, dst in ((A, B), (B, A)):
        p0 = np.zeros(shape, np.float32)
        p0[src] = 1.0
        with engine.HipSimulation(inp) as sim:
            sim.set_field("p", p0)
            sim.run(60)
            out.append(sim.field("p").astype(np.float64)[dst] / K[dst])
    assert abs(out[0]) > 0
    assert abs(out[0] - out[1]) <= 1e-4 * abs(out[0])


def test_run_dir_contract(tmp_path):
    inp = workloads.random_medium_2d(32, 30, 80, seed=24, air=3).inputs
    inp.write_dir(tmp_path / "txrx_0")
    engine.run_dir(tmp_path / "txrx_0")
    genout = np.fromfile(tmp_path / "txrx_0" / "genout.dat", dtype=np.float32).reshape(-1, inp.ncoordsout)
    want, _ = _ref(inp)
    assert rel_l2(genout, want) <= TOL


def test_invalid_coordinates_raise():
    inp = workloads.random_medium_2d(30, 30, 20, seed=25).inputs
    bad = dataclasses.replace(inp, outc=np.array([[0, 10_000]], np.int32))
    with engine.HipSimulation(bad) as sim, pytest.raises(engine.EngineError, match="outside"):
        sim.run(1)


def test_run_past_nt_raises():
    inp = workloads.random_medium_2d(30, 30, 10, seed=26).inputs
    with engine.HipSimulation(inp) as sim:
        sim.run(10)
        with pytest.raises(engine.EngineError, match="exceeds"):
            sim.run(1)


STREAM_SHAPES_2D = ["stream:64:2", "stream:128:7", "stream:256:64", "stream:128:33", "tile:8", "tile:16", "tile:32"]


@pytest.mark.parametrize("shape", STREAM_SHAPES_2D)
def test_streaming_kernels_match_per_cell_kernels(monkeypatch, shape):
    """The two GPU implementations (x-marching at several tile shapes, and per-cell) on the
    same inputs: 2D isotropic with air, 2D per-axis maps, 2D uniform sound speed."""
    for inp in (workloads.random_medium_2d(70, 300, 120, seed=31, air=5).inputs,
                workloads.random_medium_2d(150, 92, 100, seed=32, isotropic=False).inputs,
                workloads.plane_wave_2d(60, 280, 150, n_pml=10, sensor_depth=40).inputs,
                workloads.random_medium_2d(53, 531, 60, seed=37, n_pml=7, n_transition=5).inputs):
        if shape.startswith("tile") and inp.ny % 4:
            continue
        print(f"{shape}: {inp.shape}", flush=True)
        monkeypatch.setenv("FULLWAVE_HIP_KERNELS", shape)
        a, pa = _gpu(inp)
        monkeypatch.setenv("FULLWAVE_HIP_KERNELS", "percell")
        b, pb = _gpu(inp)
        monkeypatch.delenv("FULLWAVE_HIP_KERNELS")
        assert rel_l2(a, b) <= 1e-6
        assert rel_l2(pa, pb) <= 1e-6


@pytest.mark.parametrize("shape", ["tile:8", "tile:16", "tile:32", "stream:256:8"])
def test_unequal_spacing(monkeypatch, shape):
    """dy != dx: the x and y stencil weights differ, so the tiled kernels stage both axes'
    coefficient sets (with equal spacings they stage one set per sound speed). Against the
    per-cell kernels and the FP64 oracle."""
    base = workloads.random_medium_2d(70, 300, 120, seed=38).inputs
    inp = dataclasses.replace(base, dy=base.dy * 1.25)
    monkeypatch.setenv("FULLWAVE_HIP_KERNELS", shape)
    a, pa = _gpu(inp)
    monkeypatch.setenv("FULLWAVE_HIP_KERNELS", "percell")
    b, pb = _gpu(inp)
    monkeypatch.delenv("FULLWAVE_HIP_KERNELS")
    assert rel_l2(a, b) <= 1e-6 and rel_l2(pa, pb) <= 1e-6
    ref, _ = _ref(inp)
    err = rel_l2(a, ref)
    print(f"{shape}, dy = 1.25 dx: rel-L2 {err:.2e} against the oracle", flush=True)
    assert err <= 1e-4
    assert rel_l2(a, _gpu(base)[0]) > 1e-3  # the spacing reaches the stencil


def _many_speeds_2d(shape, nt, seed, n_speeds=767):
    """2D inputs whose sound speed spans n_speeds integer values (the tiled kernels' limit is
    768 coefficient sets), smooth in space, with dy = 1.25 dx so both axes' sets are staged."""
    from scipy.ndimage import gaussian_filter

    rng = np.random.default_rng(seed)
    f = gaussian_filter(rng.random(shape), 6.0)
    c = 1000.0 + (n_speeds - 1) * (f - f.min()) / np.ptp(f)
    src = [[shape[0] // 2, shape[1] // 2], [20, 40]]
    sig = np.sin(np.arange(nt) * 0.3)[None, :] * np.array([[1.0], [0.7]])
    inp = raw_inputs(shape, c=c, rho=1000 + 50 * rng.random(shape), nt=nt, icc=src, icmat=sig,
                     outc=[[30, 30], [50, 150], [12, 180]])
    return dataclasses.replace(inp, dy=inp.dy * 1.25)


@pytest.mark.parametrize("shape", ["tile:8", "tile:32"])
def test_tiled_kernels_at_the_sound_speed_limit(monkeypatch, shape):
    """767 sound speeds with unequal spacings: the most coefficient sets the tiled kernels (and
    the tiled group kernels) stage in LDS, next to their largest tile; against the per-cell
    kernels and the FP64 oracle."""
    inp = _many_speeds_2d((72, 200), 120, seed=41)
    assert inp.ndmap == 767 and inp.ny % 4 == 0
    monkeypatch.setenv("FULLWAVE_HIP_KERNELS", shape)
    a, pa = _gpu(inp)
    monkeypatch.setenv("FULLWAVE_HIP_KERNELS", "percell")
    b, pb = _gpu(inp)
    monkeypatch.delenv("FULLWAVE_HIP_KERNELS")
    assert rel_l2(a, b) <= 1e-6 and rel_l2(pa, pb) <= 1e-6
    ref, _ = _ref(inp)
    err = rel_l2(a, ref)
    print(f"{shape}, {inp.ndmap} sound speeds, dy = 1.25 dx: rel-L2 {err:.2e} against the oracle", flush=True)
    assert err <= 1e-4
    rows = int(shape.split(":")[1])
    with engine.HipSimulation(inp) as sim:
        with engine.HipGroup([sim], variant="tile", chunk=rows) as grp:
            assert grp.kernel_config()["velocity"]["variant"] == "tile"
            grp.run(inp.nt)
        assert rel_l2(sim.sensors().astype(np.float64), ref) <= 1e-4


@pytest.mark.parametrize("shape", ["pipelite:512:8", "pipe:1024:7", "stream:512:9"])
def test_unequal_spacing_3d(monkeypatch, shape):
    """dz != dx = dy in 3D: the three axes' stencil weights differ. The 3D kernels read them
    from LDS for the three-speed medium (61 sound speeds) and from the global table for 171."""
    for base in (_three_speed_3d((44, 50, 90), 60, seed=36), workloads.random_medium_3d(30, 26, 40, 50, seed=39).inputs):
        inp = dataclasses.replace(base, dz=base.dz * 1.2)
        monkeypatch.setenv("FULLWAVE_HIP_KERNELS", shape)
        a, pa = _gpu(inp)
        monkeypatch.setenv("FULLWAVE_HIP_KERNELS", "percell")
        b, pb = _gpu(inp)
        monkeypatch.delenv("FULLWAVE_HIP_KERNELS")
        assert rel_l2(a, b) <= 1e-6 and rel_l2(pa, pb) <= 1e-6
        ref, _ = _ref(inp)
        err = rel_l2(a, ref)
        print(f"{shape}, {inp.shape}, {inp.ndmap} sound speeds, dz = 1.2 dx: rel-L2 {err:.2e} against the oracle",
              flush=True)
        assert err <= 1e-4
        assert rel_l2(a, _gpu(base)[0]) > 1e-3


def _three_speed_3d(shape, nt, seed):
    """3D inputs whose sound speed takes three values (61 coefficient sets, LDS-resident)."""
    rng = np.random.default_rng(seed)
    c = rng.choice([1500.0, 1540.0, 1560.0], size=shape)
    rho = 950 + 100 * rng.random(shape)
    relax = {"v": {"kappa": 1.01, "d": (2e5, 4e5), "alpha": (3e6, 2e7)},
             "p": {"kappa": 0.99, "d": (1e5, 3e5), "alpha": (5e6, 4e7)}}
    src = [[shape[0] // 2, shape[1] // 2, shape[2] // 2], [12, 14, 20]]
    sig = np.sin(np.arange(nt) * 0.3)[None, :] * np.array([[1.0], [0.5]])
    return raw_inputs(shape, c=c, rho=rho, nt=nt, relax=relax, icc=src, icmat=sig,
                      outc=[[20, 20, 20], [15, 30, 40], [30, 22, 11]])


@pytest.mark.parametrize("shape", ["stream:512:2", "stream:1024:16", "stream:512:9", "stream:1024:5", "pipe:512:7",
                                   "pipe:1024:16", "pipelite:512:2", "pipelite:1024:9", "pipelite:512:64"])
def test_streaming_3d_matches_per_cell(monkeypatch, shape):
    for inp in (workloads.random_medium_3d(40, 30, 150, 60, seed=33, air=4).inputs,
                workloads.random_medium_3d(26, 70, 36, 50, seed=34, isotropic=False).inputs,
                _three_speed_3d((44, 50, 90), 60, seed=36)):
        print(f"{shape}: {inp.shape}", flush=True)
        monkeypatch.setenv("FULLWAVE_HIP_KERNELS", shape)
        a, pa = _gpu(inp)
        monkeypatch.setenv("FULLWAVE_HIP_KERNELS", "percell")
        b, pb = _gpu(inp)
        monkeypatch.delenv("FULLWAVE_HIP_KERNELS")
        assert rel_l2(a, b) <= 1e-6
        assert rel_l2(pa, pb) <= 1e-6


def test_autotuned_config_is_reported():
    # long enough for the sensors to see the wave itself, not only the stencil's precursor
    inp = workloads.random_medium_2d(80, 200, 240, seed=35).inputs
    with engine.HipSimulation(inp) as sim:
        cfg = sim.kernel_config()
        assert cfg["velocity"]["variant"] in ("percell", "stream", "tile")
        assert cfg["pressure"]["variant"] in ("percell", "stream", "tile")
        sim.run(inp.nt)
        got = sim.sensors().astype(np.float64)
    want, _ = _ref(inp)
    assert rel_l2(got, want) <= TOL


def test_solver_end_to_end(tmp_path):
    """Front end -> Solver.run -> engine, in memory and through the directory contract."""
    import fullwave25_amd as fw
    from fullwave25_amd.utils.pulse import gaussian_modulated_sinusoidal_signal

    grid = workloads._grid(60, 50, nt=160)
    shape = (grid.nx, grid.ny)
    medium = fw.Medium(grid, np.full(shape, 1540.0), np.full(shape, 1000.0), np.full(shape, 0.5),
                       np.full(shape, 1.0), np.zeros(shape))
    mask = np.zeros(shape, bool)
    mask[0, :] = True
    sig = gaussian_modulated_sinusoidal_signal(grid.nt, grid.duration, 2, 2, 3e6, 1e5)
    source = fw.Source(np.tile(sig, (shape[1], 1)), mask)
    smask = np.zeros(shape, bool)
    smask[40, ::5] = True
    sensor = fw.Sensor(smask, sampling_modulus_time=2)
    solver = fw.Solver(tmp_path, grid, medium, source, sensor, pml_layer_thickness_px=10, n_transition_layer=8)
    out = solver.run()
    assert out.shape == (sensor.n_sensors, -(-grid.nt // 2))
    inputs, _ = solver.build_inputs()
    want, _ = _ref(inputs)
    assert rel_l2(out.T, want) <= TOL
    path = solver.run("txrx_files", load_results=False)
    assert path.name == "genout.dat" and (path.parent / "fw2_execution.log").exists()
    files = np.fromfile(path, np.float32).reshape(-1, sensor.n_sensors)
    np.testing.assert_array_equal(files, out.T)


def _shot_stack(inp, n_shots, seed):
    """Distinct source signals on the same source points: scaled, delayed copies."""
    rng = np.random.default_rng(seed)
    out = []
    for b in range(n_shots):
        sh = np.roll(inp.icmat, int(rng.integers(0, 12)), axis=1) * float(rng.uniform(0.5, 2.0))
        out.append(sh.astype(np.float32))
    return np.stack(out)


@pytest.mark.parametrize("case", ["2d_iso", "3d_aniso", "2d_exp"])
def test_batched_shots_match_oracle(case):
    """Several transmits on one medium in one handle: every shot against its own FP64 oracle run."""
    inp = CASES[case]()
    shots = _shot_stack(inp, 3, seed=40)
    print(f"batch {case}: shape {inp.shape}, 3 shots", flush=True)
    with engine.HipSimulation(inp, shots=shots) as sim:
        sim.run(inp.nt)
        got = sim.sensors().astype(np.float64)
        p_all = sim.field("p", all_shots=True)
    assert got.shape[0] == 3
    for b in range(3):
        one = dataclasses.replace(inp, icmat=shots[b])
        want, p_want = _ref(one)
        err = rel_l2(got[b], want)
        print(f"  shot {b}: rel-L2 {err:.3e}", flush=True)
        assert err <= TOL
        assert rel_l2(p_all[b], p_want) <= TOL


@pytest.mark.parametrize("shape", ["percell", "stream:128:16", "tile:16", "tile:8/stream:256:8"])
def test_batched_shots_bitwise_equal_single_runs(monkeypatch, shape):
    """With the kernel shape pinned, each shot of a batch is bit-identical to its own run."""
    inp = workloads.random_medium_2d(52, 64, 160, seed=41, air=3).inputs
    shots = _shot_stack(inp, 4, seed=42)
    monkeypatch.setenv("FULLWAVE_HIP_KERNELS", shape)
    with engine.HipSimulation(inp, shots=shots) as sim:
        sim.run(inp.nt)
        got = sim.sensors()
    for b in (0, 3):
        with engine.HipSimulation(dataclasses.replace(inp, icmat=shots[b])) as one:
            one.run(inp.nt)
            np.testing.assert_array_equal(got[b], one.sensors())


@pytest.mark.parametrize("shape", ["stream:1024:16", "pipelite:512:8"])
def test_batched_shots_3d_bitwise_and_reuse(monkeypatch, shape):
    """3D batch (x-marching kernels) equals single runs; new source matrices on the same
    handle after a reset give the same result as a fresh handle."""
    inp = workloads.random_medium_3d(20, 22, 70, 60, seed=43).inputs
    a, b = _shot_stack(inp, 2, seed=44), _shot_stack(inp, 2, seed=45)
    monkeypatch.setenv("FULLWAVE_HIP_KERNELS", shape)
    with engine.HipSimulation(inp, shots=a) as sim:
        sim.run(inp.nt)
        got_a = sim.sensors()
        sim.set_shots(b)
        sim.reset()
        sim.run(inp.nt)
        got_b = sim.sensors()
    with engine.HipSimulation(dataclasses.replace(inp, icmat=a[1])) as one:
        one.run(inp.nt)
        np.testing.assert_array_equal(got_a[1], one.sensors())
    with engine.HipSimulation(inp, shots=b) as fresh:
        fresh.run(inp.nt)
        np.testing.assert_array_equal(got_b, fresh.sensors())


@pytest.mark.parametrize("pinned", ["percell", "tile:8/stream:256:8"])
def test_graph_replay_matches_plain_launches(monkeypatch, pinned):
    """Steady steps replayed from a captured graph (device step counter) give the same bits as
    plain launches, across several replays with plain steps between them, after a reset, and
    for a batch of shots."""
    inp = workloads.random_medium_2d(60, 64, 300, seed=70, air=4).inputs
    monkeypatch.setenv("FULLWAVE_HIP_KERNELS", pinned)
    with engine.HipSimulation(inp) as sim:
        sim.run(inp.nt)
        want = sim.sensors()
        p_want = sim.field("p")
    with engine.HipSimulation(inp) as sim:
        sim.set_graph_steps(16)
        for n in (50, 37, 100, 113):
            sim.run(n)
        np.testing.assert_array_equal(sim.sensors(), want)
        np.testing.assert_array_equal(sim.field("p"), p_want)
        sim.reset()
        sim.run(inp.nt)
        np.testing.assert_array_equal(sim.sensors(), want)
    shots = _shot_stack(inp, 3, seed=71)
    with engine.HipSimulation(inp, shots=shots) as plain, engine.HipSimulation(inp, shots=shots) as graph:
        plain.run(inp.nt)
        graph.set_graph_steps(32)
        graph.run(inp.nt)
        np.testing.assert_array_equal(graph.sensors(), plain.sensors())


def test_empty_lists_and_degenerate_counts():
    """No source points (the wave starts from an initial pressure field instead), no sensors,
    a recording interval longer than the run, and a run of zero steps: against the oracle
    started from the same field."""
    shape = (48, 52)
    rng = np.random.default_rng(42)
    p0 = np.zeros(shape)
    p0[20:28, 22:30] = rng.random((8, 8))
    cases = {
        "no sources": raw_inputs(shape, nt=40, outc=[[10, 10], [30, 40]]),
        "no sensors": raw_inputs(shape, nt=40, icc=[[24, 26]], icmat=np.ones((1, 5))),
        "modT > nT": raw_inputs(shape, nt=30, modt=45, outc=[[24, 26]]),
    }
    for name, inp in cases.items():
        with engine.HipSimulation(inp) as sim:
            sim.set_field("p", p0.astype(np.float32))
            sim.run(0)
            assert sim.step == 0 and sim.sensors().shape == (0, inp.ncoordsout)
            sim.run(inp.nt)
            got, p_got = sim.sensors().astype(np.float64), sim.field("p").astype(np.float64)
        o = oracle.OracleSimulation(inp, 64)
        o.set_field("p", p0)
        o.run(inp.nt)
        assert got.shape == o.sensors().shape == (-(-inp.nt // inp.modt), inp.ncoordsout), name
        if inp.ncoordsout:
            assert np.abs(got).max() > 0, name
            assert rel_l2(got, o.sensors()) <= TOL, name
        assert rel_l2(p_got, o.field("p")) <= TOL, name
        print(f"{name}: records {got.shape}, field rel-L2 {rel_l2(p_got, o.field('p')):.2e}", flush=True)


@pytest.mark.parametrize("shape", ["tile:8", "tile:32", "stream:256:16", "stream:64:8"])
def test_fused_boundary_matches_boundary_launch(monkeypatch, shape):
    """The 2D pressure kernels do the step's sources, pressure release and sensors for the points
    their tiles own; the records equal those of the same kernels followed by the boundary pass
    launched on its own, bit for bit: points in the ghost border, on tile edges, repeated and on
    each other, three shots, a recording interval of 3, sources that stop before the run ends,
    and steps replayed from a graph."""
    from tests.test_gpu_group import _edge_point_media

    inp = _edge_point_media()[0]
    rng = np.random.default_rng(72)
    shots = (inp.icmat[None] * (1 + rng.random((3, 1, 1)))).astype(np.float32)
    monkeypatch.setenv("FULLWAVE_HIP_KERNELS", shape)
    out = {}
    for fused in ("1", "0"):
        monkeypatch.setenv("FULLWAVE_HIP_FUSED_BOUNDARY", fused)
        with engine.HipSimulation(inp, shots=shots) as sim:
            sim.run(25)
            sim.set_graph_steps(10)
            sim.run(inp.nt - 25)
            out[fused] = sim.sensors().copy()
    print(f"{shape}: records {out['1'].shape}, max |rec| {np.abs(out['1']).max():.3e}", flush=True)
    assert np.abs(out["1"]).max() > 0
    np.testing.assert_array_equal(out["1"], out["0"])
