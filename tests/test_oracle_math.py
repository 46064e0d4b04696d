"""What the mathematics fixes about the update the oracle (and the engine) implement.

The reference ships no engine, so these pin the oracle without it:

* the coefficient tables satisfy the consistency condition of a first derivative, and with the
  transverse term in the form the oracle uses they give a leapfrog scheme whose phase error is
  below 6e-4 at every angle for >= 4 points per wavelength, while the alternatives are >= 30x
  worse (this is what identifies the stencil form);
* one time step of the oracle applied to a plane wave equals the analytic symbol of that
  stencil (both half steps, 2D and 3D);
* velocity and pressure operators are adjoint: initial-pressure reciprocity and exact
  conservation of the leapfrog's discrete energy in a lossless box;
* the absorbing layer removes outgoing waves (reflection below -40 dB).
"""

import numpy as np
import pytest

import oracle
from fullwave25_amd.solver import coefficients
from tests.helpers import raw_inputs, rel_l2


def _sym(w, th, others, transverse="oracle"):
    """Symbol S (D f = i S f / h) of the staggered operator along one axis."""
    ax = sum(w[k - 1] * 2 * np.sin((2 * k - 1) * th / 2) for k in range(1, 9))
    if transverse == "none":
        return ax
    cos_sum = sum(2 * np.cos(t) for t in others)
    if transverse == "laplacian":
        cos_sum = sum(2 * np.cos(t) - 2 for t in others)
    return ax + w[8] * 2 * np.sin(th / 2) * cos_sum


def _max_phase_error(is_3d, r, transverse):
    w = coefficients.stencil_weights(r, is_3d=is_3d)
    worst = 0.0
    for ppw in (4, 6, 8, 12, 20):
        kdx = 2 * np.pi / ppw
        for az in np.linspace(0, np.pi / 4, 9):
            for el in (np.linspace(0, np.arctan(1 / np.sqrt(2)), 5) if is_3d else [0.0]):
                k = np.array([np.cos(el) * np.cos(az), np.cos(el) * np.sin(az), np.sin(el)]) * kdx
                k = k if is_3d else k[:2]
                s2 = sum(_sym(w, k[a], [k[b] for b in range(len(k)) if b != a], transverse) ** 2
                         for a in range(len(k)))
                wdt = 2 * np.arcsin(min(r / 2 * np.sqrt(s2), 1.0))
                worst = max(worst, abs(wdt / (r * kdx) - 1))
    return worst


@pytest.mark.parametrize("is_3d", [False, True])
def test_consistency_condition(is_3d):
    nt = coefficients.n_transverse_lines(is_3d)
    for r in np.linspace(0, 0.5, 11):
        w = coefficients.stencil_weights(r, is_3d=is_3d)
        s = sum((2 * k - 1) * w[k - 1] for k in range(1, 9)) + nt * w[8]
        assert abs(s - 1) < 6e-5
    assert abs(coefficients.stencil_weights(0.0, is_3d=is_3d)[8]) < 1e-6


@pytest.mark.parametrize("is_3d", [False, True])
def test_dispersion_identifies_stencil_form(is_3d):
    for r in (0.2, 0.4):
        good = _max_phase_error(is_3d, r, "oracle")
        assert good < 6e-4
        assert _max_phase_error(is_3d, r, "none") > 30 * good
        assert _max_phase_error(is_3d, r, "laplacian") > 30 * good


def test_dmap_at_reference_speed_is_d():
    c = np.full((5, 5), 1540.0)
    dx = 1540 / 3e6 / 12
    dt = 0.4 * dx / 1540
    dmap = coefficients.d_map(c, dt, dx, is_3d=False)
    np.testing.assert_allclose(dmap[:, :, 0], coefficients.d_matrix(0.4, is_3d=False), rtol=1e-12, atol=1e-15)
    c[0, 0] = 1600.4
    assert coefficients.ndmap_value(c) == 61
    assert coefficients.dc_map(c)[0, 0] == 60 and coefficients.dc_map(c)[1, 1] == 0


@pytest.mark.parametrize("shape", [(48, 52), (30, 32, 34)])
def test_one_step_equals_analytic_symbol(oracle_lib, shape):
    nd = len(shape)
    rho, c = 1000.0, 1540.0
    inp = raw_inputs(shape, nt=1, c=c, rho=rho)
    h = inp.dx
    w = np.asarray(inp.dmap, np.float64)[1:9, 0, 0].tolist() + [float(inp.dmap[1, 1, 0])]
    k = np.array([0.31, 0.17, 0.23][:nd]) * 2 * np.pi / 12 * 12 / 6  # ~6 points per wavelength
    idx = np.indices(shape).astype(np.float64)
    phase = sum(k[a] * idx[a] for a in range(nd))
    sim = oracle.OracleSimulation(inp, 64)
    sim.set_field("p", np.sin(phase))
    sim.run(1)
    inner = tuple(slice(17, s - 17) for s in shape)
    S = [_sym(w, k[a], [k[b] for b in range(nd) if b != a]) / h for a in range(nd)]
    for a, name in enumerate(["vx", "vy", "vz"][:nd]):
        stag = phase + k[a] / 2
        want = -inp.dt / rho * S[a] * np.cos(stag)
        got = sim.field(name)
        assert rel_l2(got[inner], want[inner]) < 1e-9
    K = float(inp.maps["K"][0, 0] if nd == 2 else inp.maps["K"][0, 0, 0])  # float32-rounded rho c^2
    lam = inp.dt**2 * K / rho * sum(s * s for s in S)
    assert rel_l2(sim.field("p")[inner], (1 - lam) * np.sin(phase)[inner]) < 1e-9


def test_reciprocity(oracle_lib):
    shape = (40, 44)
    rng = np.random.default_rng(5)
    inp = raw_inputs(shape, rho=1000 + 200 * rng.random(shape), nt=80)
    K = inp.maps["K"].astype(np.float64)
    A, B = (15, 17), (24, 29)
    vals = []
    for src, dst in ((A, B), (B, A)):
        sim = oracle.OracleSimulation(inp, 64)
        p0 = np.zeros(shape)
        p0[src] = 1.0
        sim.set_field("p", p0)
        sim.run(80)
        vals.append(sim.field("p")[dst] / K[dst])
    assert abs(vals[0]) > 0
    assert abs(vals[0] - vals[1]) < 1e-12 * abs(vals[0])


def test_reciprocity_detects_a_shifted_operator(oracle_lib):
    """The same check fails when the pressure-side stencil is misplaced by one cell (a
    deliberate mistake planted in a copy of the numpy statement), so it does discriminate."""
    shape = (30, 30)
    inp = raw_inputs(shape, nt=30)
    K = inp.maps["K"].astype(np.float64)

    def run(src, shift):
        p = np.zeros(shape)
        p[src] = 1.0
        v = [np.zeros(shape), np.zeros(shape)]
        w = oracle.coefficient_sets(inp.dmap)[0]
        inner = (slice(8, -8), slice(8, -8))
        for _ in range(30):
            for a in range(2):
                d = sum(w[k - 1] * (np.roll(p, -k, a) - np.roll(p, k - 1, a)) for k in range(1, 9))
                b = 1 - a
                d = d + w[8] * sum(np.roll(np.roll(p, -1, a), s, b) - np.roll(p, s, b) for s in (1, -1))
                v[a][inner] -= inp.dt / 1000.0 * d[inner] / inp.dx
            div = 0
            for a in range(2):
                sh = shift if a == 0 else 0
                d = sum(w[k - 1] * (np.roll(v[a], -(k - 1) + sh, a) - np.roll(v[a], k + sh, a)) for k in range(1, 9))
                b = 1 - a
                d = d + w[8] * sum(np.roll(v[a], s, b) - np.roll(np.roll(v[a], 1, a), s, b) for s in (1, -1))
                div = div + d / inp.dx
            p[inner] -= inp.dt * K[inner] * div[inner]
        return p

    A, B = (12, 13), (17, 18)
    good = abs(run(A, 0)[B] / K[B] - run(B, 0)[A] / K[A]) / abs(run(A, 0)[B] / K[B])
    bad = abs(run(A, 1)[B] / K[B] - run(B, 1)[A] / K[A]) / abs(run(A, 1)[B] / K[B])
    assert good < 1e-10
    assert bad > 1e-3


def test_discrete_energy_conserved(oracle_lib):
    shape = (36, 40)
    rng = np.random.default_rng(9)
    rho = 1000 + 150 * rng.random(shape)
    inp = raw_inputs(shape, rho=rho, nt=200)
    K = inp.maps["K"].astype(np.float64)
    rhof = inp.maps["rho"].astype(np.float64)
    rbar = [0.5 * (rhof + np.roll(rhof, -1, 0)), 0.5 * (rhof + np.roll(rhof, -1, 1))]
    idx = np.indices(shape)
    p0 = np.exp(-((idx[0] - 18.0) ** 2 + (idx[1] - 20.0) ** 2) / 6.0)
    ghost = np.ones(shape, bool)
    ghost[8:-8, 8:-8] = False
    p0[ghost] = 0.0  # the ghost border is never updated, so it must start at zero
    sim = oracle.OracleSimulation(inp, 64)
    sim.set_field("p", p0)
    energies = []
    v_prev = [np.zeros(shape), np.zeros(shape)]
    p_n = p0.copy()
    for _ in range(150):
        sim.run(1)
        v_now = [sim.field("vx"), sim.field("vy")]
        e = 0.5 * np.sum(p_n**2 / K) + 0.5 * sum(np.sum(rbar[a] * v_prev[a] * v_now[a]) for a in range(2))
        energies.append(e)
        v_prev = v_now
        p_n = sim.field("p")
    energies = np.array(energies)
    assert np.max(np.abs(energies / energies[0] - 1)) < 1e-10


def test_exponential_model_energy_decay(oracle_lib):
    """Exponential-attenuation model, uniform decay factor A: rescaling p_n by A^-n and
    v_{n+1/2} by A^-(n+1) turns the update into the lossless leapfrog with K replaced by A K, whose
    discrete energy is conserved; so the lossless energy form of the actual fields,
    E_n = sum p_n^2 / 2K + sum rho_bar v_{n-1/2} v_{n+1/2} / 2, satisfies E_n = A^(2n) E_0 exactly.
    A model that scaled only one of the two fields would decay like A^n instead (determinant A
    per step), 10 % away after 120 steps at A = 0.999."""
    shape = (34, 38)
    rng = np.random.default_rng(21)
    rho = 1000 + 150 * rng.random(shape)
    inp = raw_inputs(shape, rho=rho, nt=200, a_exp=0.999)
    K = inp.maps["K"].astype(np.float64)
    A = float(inp.maps["a_exp"][0, 0])
    rhof = inp.maps["rho"].astype(np.float64)
    rbar = [0.5 * (rhof + np.roll(rhof, -1, 0)), 0.5 * (rhof + np.roll(rhof, -1, 1))]
    idx = np.indices(shape)
    p0 = np.exp(-((idx[0] - 17.0) ** 2 + (idx[1] - 19.0) ** 2) / 6.0)
    ghost = np.ones(shape, bool)
    ghost[8:-8, 8:-8] = False
    p0[ghost] = 0.0
    sim = oracle.OracleSimulation(inp, 64)
    sim.set_field("p", p0)
    v_prev = [np.zeros(shape), np.zeros(shape)]
    p_n = p0.copy()
    energies = []
    for _ in range(120):
        sim.run(1)
        v_now = [sim.field("vx"), sim.field("vy")]
        energies.append(0.5 * np.sum(p_n**2 / K) + 0.5 * sum(np.sum(rbar[a] * v_prev[a] * v_now[a]) for a in range(2)))
        v_prev = v_now
        p_n = sim.field("p")
    e = np.array(energies) / energies[0]
    n = np.arange(e.size)
    assert np.max(np.abs(e / A ** (2 * n) - 1)) < 1e-10
    assert abs(e[-1] / A ** n[-1] - 1) > 0.05


def test_absorbing_layer_reflection(oracle_lib):
    """A pulse from a point source leaves through the absorbing layer: the trace next to the
    source in a small domain matches the trace in a domain large enough that nothing comes
    back within the window, to better than -40 dB of the direct arrival."""
    from fullwave25_amd import workloads
    from fullwave25_amd.grid import Grid
    from fullwave25_amd.medium import Medium
    from fullwave25_amd.sensor import Sensor
    from fullwave25_amd.source import Source
    from fullwave25_amd.utils.pulse import gaussian_modulated_sinusoidal_signal

    def trace(n_inner, n_pml):
        g = workloads._grid(n_inner, n_inner, nt=420)
        shape = (g.nx, g.ny)
        med = Medium(g, np.full(shape, 1540.0), np.full(shape, 1000.0), np.zeros(shape),
                     np.ones(shape), np.zeros(shape))
        m = np.zeros(shape, bool)
        m[n_inner // 2, n_inner // 2] = True
        sig = gaussian_modulated_sinusoidal_signal(g.nt, g.duration, 2, 2, 3e6, 1.0)
        sm = np.zeros(shape, bool)
        sm[n_inner // 2 + 6, n_inner // 2] = True
        inp = workloads._inputs(g, med, Source(sig[None, :], m), Sensor(sm), n_pml=n_pml)
        o = oracle.OracleSimulation(inp, 64)
        o.run(inp.nt)
        return o.sensors()[:, 0]

    small = trace(60, 20)
    big = trace(260, 20)
    direct = np.max(np.abs(big))
    assert direct > 0
    assert np.max(np.abs(small - big)) < 1e-2 * direct
