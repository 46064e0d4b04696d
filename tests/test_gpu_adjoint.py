"""The engine's adjoint (hand-written transposed-step kernels over a tape of the forward run,
with checkpointed recomputation) on the GPU, against the FP64 host adjoint of the oracle (itself
checked against finite differences in tests/test_oracle_adjoint.py), and through torch autograd
against finite differences of the FP64 oracle.

Tolerance: relative L2 error of the gradient maps <= 1e-3 (FP32 adjoint vs FP64 adjoint)."""

import dataclasses

import numpy as np
import pytest

import oracle
from fullwave25_amd import engine, workloads
from fullwave25_amd.adjoint import HipAdjoint, TraceFunction, chain_to_c_rho
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-3

CASES = {
    "2d_iso": lambda: workloads.random_medium_2d(40, 36, 150, seed=60, air=3).inputs,
    "2d_aniso": lambda: workloads.random_medium_2d(34, 40, 120, seed=61, isotropic=False).inputs,
    "2d_exp": lambda: workloads.random_medium_2d(40, 36, 150, seed=62, exponential=True).inputs,
    "3d_iso": lambda: workloads.random_medium_3d(14, 12, 16, 60, seed=63, air=2).inputs,
    "2d_one_speed": lambda: workloads.plane_wave_2d(48, 40, 150, n_pml=8, sensor_depth=20).inputs,
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_gpu_adjoint_matches_fp64_adjoint(case):
    inp = CASES[case]()
    rng = np.random.default_rng(64)
    w = rng.standard_normal((inp.n_records, inp.ncoordsout)).astype(np.float32)
    print(f"adjoint {case}: shape {inp.shape}, nt {inp.nt}", flush=True)
    with HipAdjoint(inp, checkpoint_every=17) as adj:
        traces = adj.forward()
        gk, gr = adj.backward(w)
        gk2, gr2 = adj.backward(w)  # the checkpoints serve any number of sweeps
    o = oracle.OracleSimulation(inp, 64)
    want_k, want_r = o.adjoint(w.astype(np.float64))
    o2 = oracle.OracleSimulation(inp, 64)
    o2.run(inp.nt)
    assert rel_l2(traces, o2.sensors()) <= 1e-4
    ek, er = rel_l2(gk, want_k), rel_l2(gr, want_r)
    print(f"  rel-L2 dK {ek:.2e}, drho {er:.2e}", flush=True)
    assert ek <= TOL and er <= TOL
    np.testing.assert_array_equal(gk, gk2)
    np.testing.assert_array_equal(gr, gr2)


PERCELL = {"velocity": {"variant": "percell", "width": 0, "chunk": 0},
           "pressure": {"variant": "percell", "width": 0, "chunk": 0}}
TILE8 = {"velocity": {"variant": "tile", "width": 256, "chunk": 8},
         "pressure": {"variant": "tile", "width": 256, "chunk": 8}}


def test_checkpoint_spacing_does_not_change_the_gradient():
    """Segments of 1, 23 and nT steps, and the automatic choice (which tapes the whole run of a
    grid this small): the same tape, so the same gradient bit for bit."""
    inp = CASES["2d_iso"]()
    w = np.random.default_rng(65).standard_normal((inp.n_records, inp.ncoordsout)).astype(np.float32)
    out = []
    for every in (1, 23, inp.nt, 0):
        with HipAdjoint(inp, checkpoint_every=every, kernel_config=PERCELL) as adj:
            assert adj.checkpoint_every == (every or inp.nt)
            adj.forward()
            out.append(adj.backward(w))
    for gk, gr in out[1:]:
        np.testing.assert_array_equal(gk, out[0][0])
        np.testing.assert_array_equal(gr, out[0][1])


def _overwritten_sensors_case():
    """2d_iso with sensors added on source points and on pressure-release points, a repeated
    sensor, and sources that stop injecting half-way through the run: the reverse step drops the
    records of a point while the same step overwrites it."""
    inp = CASES["2d_iso"]()
    assert inp.icczero is not None and len(inp.icczero) >= 2
    outc = np.concatenate([inp.outc, inp.icc[:3], inp.icczero[:2], inp.outc[:1]]).astype(np.int32)
    return dataclasses.replace(inp, outc=outc, icmat=np.ascontiguousarray(inp.icmat[:, : inp.nt // 2]))


def test_sensors_on_overwritten_points_match_fp64_adjoint():
    inp = _overwritten_sensors_case()
    assert inp.ntic < inp.nt
    w = np.random.default_rng(68).standard_normal((inp.n_records, inp.ncoordsout)).astype(np.float32)
    with HipAdjoint(inp, checkpoint_every=19) as adj:
        traces = adj.forward()
        gk, gr = adj.backward(w)
    o = oracle.OracleSimulation(inp, 64)
    want_k, want_r = o.adjoint(w.astype(np.float64))
    o2 = oracle.OracleSimulation(inp, 64)
    o2.run(inp.nt)
    assert rel_l2(traces, o2.sensors()) <= 1e-4
    ek, er = rel_l2(gk, want_k), rel_l2(gr, want_r)
    print(f"sensors on sources and air: rel-L2 dK {ek:.2e}, drho {er:.2e}", flush=True)
    assert ek <= TOL and er <= TOL


@pytest.mark.parametrize("pin", ["percell", "tile"])
def test_taped_forward_run_equals_plain_run(pin):
    """Writing the tape leaves the forward run as it is: the traces of the adjoint's forward run
    equal those of a plain run with the same kernels, bit for bit."""
    inp = CASES["2d_iso"]()
    cfg = PERCELL if pin == "percell" else TILE8
    with HipAdjoint(inp, checkpoint_every=inp.nt, kernel_config=cfg) as adj:
        assert adj.sim.kernel_config() == cfg
        taped = adj.forward()
    with engine.HipSimulation(inp) as sim:
        sim.set_kernel_config(cfg)
        sim.run(inp.nt)
        plain = sim.sensors()
    np.testing.assert_array_equal(taped, plain)


def test_tiled_tape_equals_per_cell_tape():
    """The tape written by the tiled 2D forward kernels against the per-cell kernels' tape: the
    same arithmetic, so gradients equal to FP32 rounding (bit for bit where the kernels agree)."""
    inp = CASES["2d_iso"]()
    w = np.random.default_rng(69).standard_normal((inp.n_records, inp.ncoordsout)).astype(np.float32)
    out = []
    for cfg in (PERCELL, TILE8):
        with HipAdjoint(inp, checkpoint_every=inp.nt, kernel_config=cfg) as adj:
            adj.forward()
            out.append(adj.backward(w))
    ek, er = rel_l2(out[1][0], out[0][0]), rel_l2(out[1][1], out[0][1])
    print(f"tiled vs per-cell tape: rel-L2 dK {ek:.2e}, drho {er:.2e}", flush=True)
    assert np.abs(out[0][0]).max() > 0
    assert ek <= 1e-6 and er <= 1e-6


def test_handle_on_other_kernels_tapes_with_the_tiled_ones():
    """A handle pinned to the x-marching kernels (which write no tape) runs the tiled kernels
    once it is bound to an adjoint; the gradient still matches the FP64 adjoint."""
    inp = CASES["2d_iso"]()
    stream = {"velocity": {"variant": "stream", "width": 128, "chunk": 8},
              "pressure": {"variant": "stream", "width": 128, "chunk": 8}}
    w = np.random.default_rng(70).standard_normal((inp.n_records, inp.ncoordsout)).astype(np.float32)
    with HipAdjoint(inp, checkpoint_every=40, kernel_config=stream) as adj:
        assert adj.sim.kernel_config() == TILE8
        adj.forward()
        gk, gr = adj.backward(w)
    want_k, want_r = oracle.OracleSimulation(inp, 64).adjoint(w.astype(np.float64))
    assert rel_l2(gk, want_k) <= TOL and rel_l2(gr, want_r) <= TOL


def _oracle_loss(inp, K, rho, observed):
    maps = dict(inp.maps)
    maps["K"], maps["rho"] = K, rho  # float64: the oracle keeps them exact
    o = oracle.OracleSimulation(dataclasses.replace(inp, maps=maps), 64)
    o.run(inp.nt)
    return 0.5 * float(np.sum((o.sensors() - observed) ** 2))


def test_torch_misfit_gradient_matches_finite_differences():
    """L2 misfit against traces of a perturbed ("true") medium: the gradient from
    TraceFunction's backward, dotted with a smooth perturbation of c (and of rho), equals the
    central difference of the FP64 oracle's loss along it."""
    import torch

    inp = workloads.random_medium_2d(40, 36, 160, seed=66).inputs
    c0 = np.sqrt(inp.maps["K"].astype(np.float64) / inp.maps["rho"].astype(np.float64))
    r0 = inp.maps["rho"].astype(np.float64)
    idx = np.indices(inp.shape)
    bump = np.exp(-(((idx[0] - inp.nx / 2) ** 2 + (idx[1] - inp.ny / 2) ** 2) / 40.0))
    true = dataclasses.replace(inp, maps={**inp.maps, "K": (r0 * (c0 * (1 + 0.02 * bump)) ** 2).astype(np.float32)})
    o = oracle.OracleSimulation(true, 64)
    o.run(inp.nt)
    observed = o.sensors()

    c = torch.tensor(c0, dtype=torch.float64, requires_grad=True)
    rho = torch.tensor(r0, dtype=torch.float64, requires_grad=True)
    traces = TraceFunction.apply(c, rho, inp, 40)
    loss = 0.5 * torch.sum((traces.double() - torch.from_numpy(observed)) ** 2)
    loss.backward()
    gc, grho = c.grad.numpy(), rho.grad.numpy()

    d = bump * c0 * 1e-3
    eps = 1e-3
    fd_c = (_oracle_loss(inp, r0 * (c0 + eps * d) ** 2, r0, observed)
            - _oracle_loss(inp, r0 * (c0 - eps * d) ** 2, r0, observed)) / (2 * eps)
    dr = bump * r0 * 1e-3
    fd_r = (_oracle_loss(inp, (r0 + eps * dr) * c0**2, r0 + eps * dr, observed)
            - _oracle_loss(inp, (r0 - eps * dr) * c0**2, r0 - eps * dr, observed)) / (2 * eps)
    ad_c, ad_r = float(np.sum(gc * d)), float(np.sum(grho * dr))
    print(f"misfit: dc fd {fd_c:.6e} adj {ad_c:.6e}; drho fd {fd_r:.6e} adj {ad_r:.6e}", flush=True)
    assert abs(ad_c - fd_c) <= 1e-3 * abs(fd_c)
    assert abs(ad_r - fd_r) <= 1e-3 * abs(fd_r)


def test_one_speed_kernels_equal_general_kernels():
    """A medium of one sound speed takes the adjoint kernels that skip the per-neighbour
    coefficient lookup; the same medium with an unused second coefficient set (every dcmap
    value still 0) takes the general kernels. Same weights, same arithmetic: the gradients
    agree to FP32 rounding of the forward kernels' shapes (bit for bit when those match)."""
    inp = CASES["2d_one_speed"]()
    assert inp.ndmap == 1
    two = dataclasses.replace(inp, dmap=np.concatenate([inp.dmap, inp.dmap], axis=2))
    assert two.ndmap == 2
    rng = np.random.default_rng(65)
    w = rng.standard_normal((inp.n_records, inp.ncoordsout)).astype(np.float32)
    out = []
    for case in (inp, two):
        with HipAdjoint(case, checkpoint_every=25) as adj:
            adj.forward()
            out.append(adj.backward(w))
    (gk1, gr1), (gk2, gr2) = out
    ek, er = rel_l2(gk1, gk2), rel_l2(gr1, gr2)
    print(f"one-speed vs general adjoint kernels: rel-L2 dK {ek:.2e}, drho {er:.2e}", flush=True)
    assert np.abs(gk1).max() > 0
    assert ek <= 1e-5 and er <= 1e-5


@pytest.mark.parametrize("case", ["2d_iso", "2d_aniso", "2d_exp", "2d_one_speed"])
def test_tiled_gathers_equal_per_cell_gathers(monkeypatch, case):
    """The transposed gathers on LDS tiles (2D; the first fused with the velocity transposed)
    against the per-cell kernels (FULLWAVE_HIP_ADJ_GATHERS=percell): the same arithmetic in the
    same order, so the same bits; the forward kernels pinned so that both sweeps read the same
    tape."""
    inp = CASES[case]()
    w = np.random.default_rng(67).standard_normal((inp.n_records, inp.ncoordsout)).astype(np.float32)
    pin = PERCELL
    out = {}
    for mode in ("tiles", "percell"):
        if mode == "percell":
            monkeypatch.setenv("FULLWAVE_HIP_ADJ_GATHERS", "percell")
        else:
            monkeypatch.delenv("FULLWAVE_HIP_ADJ_GATHERS", raising=False)
        with HipAdjoint(inp, checkpoint_every=23, kernel_config=pin) as adj:
            adj.forward()
            out[mode] = adj.backward(w)
    assert np.abs(out["tiles"][0]).max() > 0
    np.testing.assert_array_equal(out["tiles"][0], out["percell"][0])
    np.testing.assert_array_equal(out["tiles"][1], out["percell"][1])
