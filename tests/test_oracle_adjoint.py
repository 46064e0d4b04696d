"""The oracle's reverse-mode (adjoint) sweep against central finite differences of the forward
oracle, FP64: the directional derivative of sum(grad * traces) along random perturbations of
the K and rho maps, for isotropic and per-axis relaxation, the exponential model and 3D."""

import dataclasses

import numpy as np
import pytest

import oracle
from fullwave25_amd import workloads


def _functional(inp, grad, K=None, rho=None):
    i2 = dataclasses.replace(inp, maps=dict(inp.maps))
    if K is not None:
        i2.maps["K"] = K
    if rho is not None:
        i2.maps["rho"] = rho
    o = oracle.OracleSimulation(i2, 64)
    o.run(inp.nt)
    return float(np.sum(o.sensors() * grad))


CASES = {
    "2d_iso": lambda: workloads.random_medium_2d(24, 20, 50, seed=3, air=3).inputs,
    "2d_aniso": lambda: workloads.random_medium_2d(22, 18, 40, seed=4, isotropic=False).inputs,
    "2d_exp": lambda: workloads.random_medium_2d(24, 20, 50, seed=5, exponential=True).inputs,
    "3d_iso": lambda: workloads.random_medium_3d(8, 7, 9, 25, seed=6).inputs,
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_adjoint_matches_finite_differences(oracle_lib, case):
    inp = CASES[case]()
    rng = np.random.default_rng(1)
    grad = rng.standard_normal((inp.n_records, inp.ncoordsout))
    gK, grho = oracle.OracleSimulation(inp, 64).adjoint(grad)
    K0 = inp.maps["K"].astype(np.float64)
    r0 = inp.maps["rho"].astype(np.float64)
    dK = rng.standard_normal(K0.shape) * K0
    dr = rng.standard_normal(r0.shape) * r0
    eps = 1e-6
    fd_k = (_functional(inp, grad, K=K0 + eps * dK) - _functional(inp, grad, K=K0 - eps * dK)) / (2 * eps)
    fd_r = (_functional(inp, grad, rho=r0 + eps * dr) - _functional(inp, grad, rho=r0 - eps * dr)) / (2 * eps)
    assert abs(np.sum(gK * dK) - fd_k) <= 1e-6 * abs(fd_k)
    assert abs(np.sum(grho * dr) - fd_r) <= 1e-6 * abs(fd_r)
    # a perturbation confined to the ghost border changes nothing
    ghost = np.ones(inp.shape, bool)
    ghost[tuple(slice(8, -8) for _ in inp.shape)] = False
    assert np.all(gK[ghost] == 0)


def test_adjoint_is_linear_in_the_weights(oracle_lib):
    inp = CASES["2d_iso"]()
    rng = np.random.default_rng(2)
    g1 = rng.standard_normal((inp.n_records, inp.ncoordsout))
    g2 = rng.standard_normal((inp.n_records, inp.ncoordsout))
    a1, b1 = oracle.OracleSimulation(inp, 64).adjoint(g1)
    a2, b2 = oracle.OracleSimulation(inp, 64).adjoint(g2)
    a3, b3 = oracle.OracleSimulation(inp, 64).adjoint(2 * g1 - g2)
    np.testing.assert_allclose(a3, 2 * a1 - a2, rtol=1e-9, atol=1e-9 * np.abs(a1).max())
    np.testing.assert_allclose(b3, 2 * b1 - b2, rtol=1e-9, atol=1e-9 * np.abs(b1).max())
