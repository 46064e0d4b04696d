"""Grouped launches: several 2D simulations of different sizes and media advanced by shared
kernel launches (fw_group_*), against each simulation run on its own with the same kernel shape
(bit for bit), and the state checks of the group API."""

import numpy as np
import pytest

from fullwave25_amd import engine, workloads
from tests.helpers import raw_inputs, rel_l2

pytestmark = pytest.mark.gpu
SHAPE = "stream:256:16"


def _media():
    return [
        workloads.random_medium_2d(60, 52, 200, seed=61, air=4).inputs,
        workloads.plane_wave_2d(60, 280, 150, n_pml=10, sensor_depth=40).inputs,
        workloads.random_medium_2d(150, 92, 260, seed=62, n_pml=7, n_transition=5).inputs,
        workloads.random_medium_2d(37, 131, 180, seed=63, n_src=9, n_sensors=40).inputs,
    ]


def _alone(inp, n, monkeypatch, codes=True, shape=SHAPE):
    monkeypatch.setenv("FULLWAVE_HIP_KERNELS", shape)
    with engine.HipSimulation(inp) as sim:
        sim.set_relaxation_codes(codes)
        sim.run(n)
        out = sim.sensors().copy()
    monkeypatch.delenv("FULLWAVE_HIP_KERNELS")
    return out


def test_group_matches_members_run_alone(monkeypatch):
    media = _media()
    k = min(m.nt for m in media)
    sims = [engine.HipSimulation(m) for m in media]
    sims[2].set_relaxation_codes(False)  # one member on the per-cell maps, the rest coded
    try:
        with engine.HipGroup(sims, width=256, chunk=16) as grp:
            grp.run(k // 2)
            grp.run(k - k // 2)
        got = [s.sensors().copy() for s in sims]
        for i, (m, g) in enumerate(zip(media, got)):
            want = _alone(m, k, monkeypatch, codes=(i != 2))
            print(f"member {i}: shape {m.shape}, {k} steps, max |rec| {np.abs(g).max():.3e}", flush=True)
            assert np.abs(want).max() > 0
            np.testing.assert_array_equal(g, want)
        # after the group, each member finishes its own run alone
        for i, (m, s) in enumerate(zip(media, sims)):
            s.run(m.nt - k)
            np.testing.assert_array_equal(s.sensors(), _alone(m, m.nt, monkeypatch, codes=(i != 2)))
    finally:
        for s in sims:
            s.close()


def test_group_with_batched_member(monkeypatch):
    a = workloads.random_medium_2d(60, 52, 120, seed=64).inputs
    b = workloads.random_medium_2d(80, 60, 120, seed=65).inputs
    rng = np.random.default_rng(66)
    shots = (a.icmat[None] * (1 + rng.random((3, 1, 1)))).astype(np.float32)
    sa = engine.HipSimulation(a, shots=shots)
    sb = engine.HipSimulation(b)
    try:
        with engine.HipGroup([sa, sb]) as grp:  # shape timed on the fresh members
            cfg = grp.kernel_config()["velocity"]
            grp.run(a.nt)
        shape = f"stream:{cfg['width']}:{cfg['chunk']}"
        print(f"timed group shape {shape}", flush=True)
        monkeypatch.setenv("FULLWAVE_HIP_KERNELS", shape)
        with engine.HipSimulation(a, shots=shots) as alone:
            alone.run(a.nt)
            np.testing.assert_array_equal(sa.sensors(), alone.sensors())
        monkeypatch.delenv("FULLWAVE_HIP_KERNELS")
        np.testing.assert_array_equal(sb.sensors(), _alone(b, b.nt, monkeypatch, shape=shape))
    finally:
        sa.close()
        sb.close()


def test_group_state_checks():
    m2 = workloads.random_medium_2d(40, 40, 50, seed=67).inputs
    m3 = workloads.random_medium_3d(20, 18, 22, 20, seed=68).inputs
    with engine.HipSimulation(m2) as a, engine.HipSimulation(m2) as b:
        with pytest.raises(engine.EngineError, match="twice"):
            engine.HipGroup([a, a])
        with engine.HipSimulation(m3) as c, pytest.raises(engine.EngineError, match="2D isotropic"):
            engine.HipGroup([a, c])
        b.run(3)
        grp = engine.HipGroup([a, b])  # b has run: no timing, the default shape
        assert grp.kernel_config()["velocity"] == {"variant": "stream", "width": 256, "chunk": 16}
        grp.run(5)
        assert (a.step, b.step) == (5, 8)
        with pytest.raises(engine.EngineError, match="steps left"):
            grp.run(m2.nt - 7)
        a.set_relaxation_codes(False)  # new coefficient layout: the group's view is stale
        with pytest.raises(engine.EngineError, match="changed"):
            grp.run(1)
        grp.close()


def test_members_at_different_steps(monkeypatch):
    """Members join at different steps (a job that arrives while others run): each member's
    records equal its own run alone, step for step."""
    ma = workloads.random_medium_2d(60, 52, 200, seed=69).inputs
    mb = workloads.random_medium_2d(44, 90, 160, seed=70).inputs
    a, b = engine.HipSimulation(ma), engine.HipSimulation(mb)
    try:
        with engine.HipGroup([a, b], width=256, chunk=16) as grp:
            a.run(37)  # a runs ahead on its own, then both advance together
            grp.run(60)
            b.run(11)
            grp.run(50)
        assert (a.step, b.step) == (147, 121)
        np.testing.assert_array_equal(a.sensors(), _alone_partial(ma, 147, monkeypatch))
        np.testing.assert_array_equal(b.sensors(), _alone_partial(mb, 121, monkeypatch))
    finally:
        a.close()
        b.close()


def _alone_partial(inp, n, monkeypatch):
    return _alone(inp, n, monkeypatch)


def test_group_against_oracle():
    """Two members of different sizes, grouped for the first 150 steps; the longer one then
    finishes alone. Both full records against the FP64 oracle."""
    import oracle

    media = [workloads.random_medium_2d(60, 52, 300, seed=11, air=6).inputs,
             workloads.random_medium_2d(37, 131, 150, seed=13, n_src=9, n_sensors=40).inputs]
    sims = [engine.HipSimulation(m) for m in media]
    try:
        with engine.HipGroup(sims) as grp:
            grp.run(150)
        sims[0].run(150)
        for m, s in zip(media, sims):
            o = oracle.OracleSimulation(m, 64)
            o.run(m.nt)
            err = rel_l2(s.sensors().astype(np.float64), o.sensors())
            print(f"{m.shape}: rel-L2 {err:.3e}", flush=True)
            assert err <= 1e-4
    finally:
        for s in sims:
            s.close()


def _tile_media():
    """Four media with ny % 4 == 0 (the tiled kernels' row alignment)."""
    return _media()[:3] + [workloads.random_medium_2d(37, 128, 180, seed=63, n_src=9, n_sensors=40).inputs]


@pytest.mark.parametrize("rows,uncoded", [(8, 1), (16, None), (32, 1), (8, None)])
def test_tiled_group_matches_members_run_alone(monkeypatch, rows, uncoded):
    """The tiled group kernels: each member's records equal its own run with tile:<rows>, bit
    for bit; every member coded (the kernels then hold the coded path only), or one member on
    its per-cell maps."""
    media = _tile_media()
    assert all(m.ny % 4 == 0 for m in media)
    k = min(m.nt for m in media)
    sims = [engine.HipSimulation(m) for m in media]
    if uncoded is not None:
        sims[uncoded].set_relaxation_codes(False)
    try:
        with engine.HipGroup(sims, variant="tile", chunk=rows) as grp:
            assert grp.kernel_config()["pressure"] == {"variant": "tile", "width": 256, "chunk": rows}
            grp.run(k // 3)
            grp.run(k - k // 3)
        for i, (m, s) in enumerate(zip(media, sims)):
            want = _alone(m, k, monkeypatch, codes=(i != uncoded), shape=f"tile:{rows}")
            print(f"member {i}: shape {m.shape}, {k} steps, max |rec| {np.abs(want).max():.3e}", flush=True)
            assert np.abs(want).max() > 0
            np.testing.assert_array_equal(s.sensors(), want)
    finally:
        for s in sims:
            s.close()


def test_timed_family_with_batched_member(monkeypatch):
    """variant="auto" times both families on fresh members (one of them a batch of three
    shots); every member equals its own run with the chosen family and shape."""
    a = workloads.random_medium_2d(60, 52, 120, seed=64).inputs
    b = workloads.random_medium_2d(80, 60, 120, seed=65).inputs
    rng = np.random.default_rng(66)
    shots = (a.icmat[None] * (1 + rng.random((3, 1, 1)))).astype(np.float32)
    sa = engine.HipSimulation(a, shots=shots)
    sb = engine.HipSimulation(b)
    try:
        with engine.HipGroup([sa, sb], variant="auto") as grp:
            cfg = grp.kernel_config()["velocity"]
            grp.run(a.nt)
        shape = f"tile:{cfg['chunk']}" if cfg["variant"] == "tile" else f"stream:{cfg['width']}:{cfg['chunk']}"
        print(f"timed group shape {shape}", flush=True)
        monkeypatch.setenv("FULLWAVE_HIP_KERNELS", shape)
        with engine.HipSimulation(a, shots=shots) as alone:
            alone.run(a.nt)
            np.testing.assert_array_equal(sa.sensors(), alone.sensors())
        monkeypatch.delenv("FULLWAVE_HIP_KERNELS")
        np.testing.assert_array_equal(sb.sensors(), _alone(b, b.nt, monkeypatch, shape=shape))
    finally:
        sa.close()
        sb.close()


def test_tiled_group_rejects_unaligned_member():
    ok = workloads.random_medium_2d(60, 52, 50, seed=61).inputs
    bad = workloads.random_medium_2d(37, 131, 50, seed=63).inputs  # ny % 4 == 3
    with engine.HipSimulation(ok) as a, engine.HipSimulation(bad) as b:
        with pytest.raises(engine.EngineError, match="ny % 4"):
            engine.HipGroup([a, b], variant="tile")
        with engine.HipGroup([a, b], variant="auto") as grp:  # times the marching family only
            assert grp.kernel_config()["velocity"]["variant"] == "stream"
            grp.run(5)


def test_group_graph_replay_matches_plain_launches():
    """Group steps replayed from a captured graph -- several replays in one call, a remainder
    launched plainly, members at different steps, traces read and other work done between the
    calls -- give every member the plain launches' records bit for bit."""
    media = _media()
    k = min(m.nt for m in media)
    plain = [engine.HipSimulation(m) for m in media]
    graph = [engine.HipSimulation(m) for m in media]
    try:
        plain[1].run(5)  # one member ahead of the others, in both groups
        graph[1].run(5)
        with engine.HipGroup(plain, width=256, chunk=16) as gp, engine.HipGroup(graph, width=256, chunk=16) as gg:
            gg.set_graph_steps(16)
            for n in (40, 16, 37, k - 5 - 93):  # 2 replays + 8 steps; 1 replay; 2 + 5; the rest
                gp.run(n)
                gg.run(n)
                for i, (a, b) in enumerate(zip(plain, graph)):
                    np.testing.assert_array_equal(b.sensors(), a.sensors(), err_msg=f"member {i} after {n}")
        assert np.abs(plain[0].sensors()).max() > 0
    finally:
        for s in plain + graph:
            s.close()


def _edge_point_media():
    """Two media whose source, pressure-release and sensor points sit where the group's pressure
    launch hands them to different workgroups: in the ghost border, on both sides of tile
    boundaries (rows 8 + 8k, columns 64k/128k/256k), repeated, and on each other; the sources
    stop before the run ends and every third step is recorded."""
    rng = np.random.default_rng(71)
    out = []
    for shape, nt in (((75, 300), 60), ((44, 132), 60)):
        nx, ny = shape
        c = rng.uniform(1450.0, 1600.0, shape)
        rho = rng.uniform(900.0, 1100.0, shape)
        icc = [[3, 5], [8, 255], [8, 256], [15, 63], [16, 64], [nx - 1, ny - 1], [30, 30], [30, 30], [20, 40],
               [20, 9], [nx - 9, 127], [24, 128]]
        zero = [[20, 40], [1, 1], [nx - 12, ny - 3], [23, 23], [24, 128]]
        outc = [[3, 5], [30, 30], [20, 40], [1, 1], [8, 256], [15, 63], [16, 64], [nx - 1, ny - 1], [40, 100],
                [40, 100], [nx - 12, ny - 3], [0, 0], [24, 128], [nx - 9, 127], [31, 255], [32, 256]]
        icc = [[min(i, nx - 1), min(j, ny - 1)] for i, j in icc]
        zero = [[min(i, nx - 1), min(j, ny - 1)] for i, j in zero]
        outc = [[min(i, nx - 1), min(j, ny - 1)] for i, j in outc]
        sig = rng.standard_normal((len(icc), 40)).astype(np.float32)
        out.append(raw_inputs(shape, c=c, rho=rho, nt=nt, icc=icc, icmat=sig, outc=outc, icczero=zero, modt=3))
    return out


@pytest.mark.parametrize("variant,width,chunk", [("stream", 256, 16), ("stream", 64, 8), ("tile", 256, 8),
                                                 ("tile", 256, 32)])
def test_group_boundary_points_at_tile_edges(monkeypatch, variant, width, chunk):
    """The pressure launch's per-tile boundary work: every member equals its own run (source,
    release and sensor passes of the plain engine) bit for bit, and the first meets the FP64
    oracle."""
    import oracle

    media = _edge_point_media()
    shape = f"tile:{chunk}" if variant == "tile" else f"stream:{width}:{chunk}"
    sims = [engine.HipSimulation(m) for m in media]
    try:
        with engine.HipGroup(sims, variant=variant, width=width, chunk=chunk) as grp:
            grp.run(25)
            grp.run(35)
        for i, (m, s) in enumerate(zip(media, sims)):
            want = _alone(m, m.nt, monkeypatch, shape=shape)
            print(f"{shape} member {i}: shape {m.shape}, max |rec| {np.abs(want).max():.3e}", flush=True)
            assert np.abs(want).max() > 0
            np.testing.assert_array_equal(s.sensors(), want)
        ref = oracle.OracleSimulation(media[0], 64)
        ref.run(media[0].nt)
        assert rel_l2(sims[0].sensors().astype(np.float64), ref.sensors()) <= 1e-4
    finally:
        for s in sims:
            s.close()
