"""Host-side front end, mirroring the reference's own unit tests where they exist
(tests/test_grid.py, tests/test_medium.py:590-716, tests/solver/test_launcher.py:119-297,
tests/utils/test_relaxation_parameters.py:120-191 of the reference)."""

import numpy as np
import pytest

import fullwave25_amd as fw
from fullwave25_amd.medium import MediumRelaxationMaps
from fullwave25_amd.relaxation import RelaxationTable, generate_relaxation_params
from fullwave25_amd.solver.launcher import HipLauncher, parse_device_id
from fullwave25_amd.solver.utils import relaxation_param_keys


def test_grid_2d_properties():
    grid = fw.Grid((0.12, 0.12), 1e6, 1e-5, c0=1540, ppw=12, cfl=0.2)
    wavelength = 1540 / 1e6
    dx = wavelength / 12
    assert np.isclose(grid.omega, 2 * np.pi * 1e6)
    assert grid.nx == int(0.12 / wavelength * 12) and grid.ny == grid.nx and grid.nz == 0
    assert np.isclose(grid.dx, dx) and np.isclose(grid.dy, dx) and grid.dz == 0
    assert np.isclose(grid.dt, 0.2 * dx / 1540)
    assert grid.nt == int(1e-5 / (0.2 * dx / 1540))


def test_grid_3d_and_checks():
    grid = fw.Grid((0.12, 0.12, 0.12), 1e6, 1e-5, c0=1540, ppw=12, cfl=0.2)
    assert grid.is_3d and grid.nz == grid.nx
    for bad in (dict(cfl=0.6), dict(cfl=0.001), dict(ppw=4)):
        with pytest.raises(AssertionError):
            fw.Grid((0.1, 0.1), 1e6, 1e-5, **bad)
    with pytest.raises(AssertionError):
        fw.Grid((0.1,), 1e6, 1e-5)


def _dummy_relax(shape, value=0.1):
    return {k: np.full(shape, value) for k in relaxation_param_keys(2)}


def _grid_small(shape=(2, 2)):
    dx = 1540 / 1e6 / 12
    return fw.Grid((shape[0] * dx, shape[1] * dx), 1e6, 1e-5, c0=1540, ppw=12, cfl=0.2)


def test_relaxation_coefficients_closed_form():
    """The a/b values of the reference's test_calc_relaxation_param_dict_for_fw2_values."""
    grid = _grid_small()
    dt = grid.dt
    shape = (2, 2)
    rd = {
        "kappa_x1": np.full(shape, 2.0), "kappa_x2": np.full(shape, 3.0),
        "d_x1_nu1": np.full(shape, 0.1), "alpha_x1_nu1": np.full(shape, 0.2),
        "d_x1_nu2": np.full(shape, 0.11), "alpha_x1_nu2": np.full(shape, 0.22),
        "d_x2_nu1": np.full(shape, 0.15), "alpha_x2_nu1": np.full(shape, 0.25),
        "d_x2_nu2": np.full(shape, 0.165), "alpha_x2_nu2": np.full(shape, 0.275),
    }
    m = MediumRelaxationMaps(grid, np.full(shape, 1500.0), np.full(shape, 1000.0), np.zeros(shape), rd)
    fw2 = m.relaxation_param_dict_for_fw2
    b = np.exp(-(0.1 / 2.0 + 0.2) * dt)
    a = 0.1 / (2.0 * (0.1 + 2.0 * 0.2) + 1e-10) * (b - 1)
    np.testing.assert_allclose(fw2["b_pml_u1"], b, rtol=1e-10)
    np.testing.assert_allclose(fw2["a_pml_u1"], a, rtol=1e-10)


def test_kappa_mapping_iso_and_aniso():
    grid = _grid_small()
    shape = (2, 2)
    rd = _dummy_relax(shape)
    rd["kappa_x1"] = np.full(shape, 5.0)
    rd["kappa_x2"] = np.full(shape, 7.0)
    iso = MediumRelaxationMaps(grid, np.full(shape, 1500.0), np.full(shape, 1000.0), np.zeros(shape), rd)
    np.testing.assert_allclose(iso.relaxation_param_dict_for_fw2["kappa_u"], 5.0)
    np.testing.assert_allclose(iso.relaxation_param_dict_for_fw2["kappa_x"], 7.0)
    ani = MediumRelaxationMaps(grid, np.full(shape, 1500.0), np.full(shape, 1000.0), np.zeros(shape), rd,
                               use_isotropic_relaxation=False)
    f = ani.relaxation_param_dict_for_fw2
    for k, v in (("kappa_u", 5.0), ("kappa_w", 5.0), ("kappa_x", 7.0), ("kappa_y", 7.0)):
        np.testing.assert_allclose(f[k], v)
    assert {"a_pml_w1", "b_pml_y2"} <= set(f)


def test_mechanisms_sorted_by_time_constant():
    grid = _grid_small()
    shape = (2, 2)
    rd = _dummy_relax(shape)
    rd["kappa_x1"] = np.ones(shape)
    rd["d_x1_nu1"], rd["alpha_x1_nu1"] = np.full(shape, 5.0), np.full(shape, 5.0)  # tau 10
    rd["d_x1_nu2"], rd["alpha_x1_nu2"] = np.full(shape, 1.0), np.full(shape, 1.0)  # tau 2
    m = MediumRelaxationMaps(grid, np.full(shape, 1500.0), np.full(shape, 1000.0), np.zeros(shape), rd)
    np.testing.assert_allclose(m.relaxation_param_dict["d_x1_nu1"], 1.0)
    np.testing.assert_allclose(m.relaxation_param_dict["d_x1_nu2"], 5.0)


def test_relaxation_keys_checked():
    grid = _grid_small()
    shape = (2, 2)
    rd = _dummy_relax(shape)
    rd["bogus"] = np.zeros(shape)
    with pytest.raises(ValueError, match="Unknown relaxation parameter keys"):
        MediumRelaxationMaps(grid, np.full(shape, 1500.0), np.full(shape, 1000.0), np.zeros(shape), rd)


@pytest.mark.parametrize(("given", "want"), [(None, 0), (0, 0), ("3", 3), ([2], 2), ("1", 1)])
def test_device_id_parsing(given, want):
    assert parse_device_id(given) == want


@pytest.mark.parametrize("bad", [-1, "x", "-2", [1, -1], 1.5])
def test_device_id_rejected(bad):
    with pytest.raises(ValueError):
        parse_device_id(bad)


def test_multi_device_list_is_not_supported():
    with pytest.raises(NotImplementedError):
        parse_device_id([0, 1])


def test_launcher_cpu_not_implemented(tmp_path):
    with pytest.raises(NotImplementedError):
        HipLauncher(use_gpu=False).run(tmp_path)


def test_lookup_binning_and_clipping():
    """searchsorted onto the table grid, then clip (relaxation_parameters.py:47-54)."""
    t = RelaxationTable(alpha_list=[0.1, 0.2, 0.3], power_list=[1.0, 1.5])
    ia, ip = t.indices(np.array([0.05, 0.1, 0.15, 0.3, 0.9]), np.array([0.5, 1.0, 1.2, 1.5, 3.0]))
    np.testing.assert_array_equal(ia, [0, 0, 1, 2, 2])
    np.testing.assert_array_equal(ip, [0, 0, 1, 1, 1])


def test_table_save_load_roundtrip(tmp_path):
    t = RelaxationTable(alpha_list=[0.25, 0.5], power_list=[1.0, 1.2])
    t.fill()
    for suffix in (".npz", ".mat"):
        t.save(tmp_path / f"db{suffix}")
        u = RelaxationTable.load(tmp_path / f"db{suffix}")
        np.testing.assert_allclose(u.database, t.database)
        out = generate_relaxation_params(np.full((2, 2), 0.5), np.full((2, 2), 1.2),
                                         path_database=tmp_path / f"db{suffix}")
        np.testing.assert_allclose(out["kappa_x1"], t.database[1, 1, 0])


def test_fit_meets_power_law_in_model():
    """The FP64 fit reproduces alpha0 f^gamma in the engine's discrete-time model inside the
    band (the fit range without its margins): 2 % for mid-range parameters, 6 % at gamma 1.6
    with a strong alpha0 (kappa >= 1 limits how steep the model can rise)."""
    from fullwave25_amd import relaxation as R

    dt = 0.4 * (1540 / 3e6 / 12) / 1540
    band = R.fit_band_for_grid(1540 / 3e6 / 12)
    f = np.geomspace(band[0] * R.FIT_MARGIN, band[1] / R.FIT_MARGIN, 60)
    w = 2 * np.pi * f
    for a0, g, tol in ((0.5, 1.1, 0.02), (0.3, 0.8, 0.02), (1.5, 1.0, 0.02), (1.5, 1.6, 0.06)):
        p, _ = R.fit_relaxation_parameters(a0, g, dt=dt, band=band)
        s1 = R.stretch_inverse_discrete(w, p[0], p[2::4], p[3::4], dt)
        s2 = R.stretch_inverse_discrete(w, p[1], p[4::4], p[5::4], dt)
        err = np.abs(np.imag((w / R.C_REF) / np.sqrt(s1 * s2)) / R.target_attenuation(f, a0, g) - 1).max()
        assert err < tol, (a0, g, err)
        assert p[0] >= 1.0 and p[1] >= 1.0


def test_medium_build_uses_grid_time_step():
    grid = _grid_small((4, 4))
    shape = (4, 4)
    med = fw.Medium(grid, np.full(shape, 1540.0), np.full(shape, 1000.0), np.full(shape, 0.5),
                    np.full(shape, 1.1), np.zeros(shape))
    built = med.build()
    assert isinstance(built, MediumRelaxationMaps)
    assert np.all(built.relaxation_param_dict["kappa_x1"] >= 1.0)
    assert np.all(np.isfinite(built.relaxation_param_dict_for_fw2["a_pml_u1"]))


def test_saved_table_refused_for_another_time_step(tmp_path):
    """A table is a fit of the engine's discrete-time recursion at one time step: it records the
    step, and a grid with another step refuses it; a table without a step (the reference's
    continuous-time format) is taken with a warning."""
    dt = 0.4 * (1540 / 3e6 / 12) / 1540
    t = RelaxationTable(alpha_list=[0.5], power_list=[1.1], dt=dt)
    t.fill()
    t.save(tmp_path / "db.mat")
    a, g = np.full((2, 2), 0.5), np.full((2, 2), 1.1)
    out = generate_relaxation_params(a, g, path_database=tmp_path / "db.mat", dt=dt)
    np.testing.assert_allclose(out["kappa_x1"], t.database[0, 0, 0])
    with pytest.raises(ValueError, match="fitted for dt"):
        generate_relaxation_params(a, g, path_database=tmp_path / "db.mat", dt=1.5 * dt)
    u = RelaxationTable(alpha_list=[0.5], power_list=[1.1])
    u.fill()
    u.save(tmp_path / "ct.npz")
    assert RelaxationTable.load(tmp_path / "ct.npz").dt is None
    generate_relaxation_params(a, g, path_database=tmp_path / "ct.npz", dt=dt)  # warns, does not raise
