"""Acoustic media (mirror of fullwave/medium.py).

* :class:`MediumRelaxationMaps` — per-voxel relaxation parameters given directly
  (fullwave/medium.py:20-557). It sorts the relaxation mechanisms of every voxel by time
  constant ``d / kappa + alpha`` (fullwave/medium.py:121-222) and derives the engine maps:
  kappa and the recursive-convolution coefficients
  ``b = exp(-(d / kappa + alpha) dt)``, ``a = d / (kappa (d + kappa alpha) + 1e-10) (b - 1)``
  (fullwave/medium.py:273-289, 299-406).
* :class:`Medium` — sound speed, density, power-law attenuation ``alpha0 f^gamma`` and
  nonlinearity; ``build()`` turns (alpha0, gamma) into relaxation maps with this package's FP64
  relaxation fit (fullwave/medium.py:754-1066 does the same through a look-up table that the
  reference does not ship).
* :class:`MediumExponentialAttenuation` — single-rate attenuation variant
  (fullwave/medium.py:560-751, 1068-1111).
"""

from __future__ import annotations

import logging
from pathlib import Path

import numpy as np
from numpy.typing import NDArray

from .grid import Grid
from .solver.utils import initialize_relaxation_param_dict, relaxation_param_keys

logger = logging.getLogger("__main__." + __name__)

_SUPPORTED_N_RELAX = (2,)


def calc_a_and_b(d, kappa, alpha, dt):
    """Recursive-convolution coefficients of one relaxation mechanism.

    ``b = exp(-(d / kappa + alpha) dt)`` and ``a = d / (kappa (d + kappa alpha) + eps) (b - 1)``
    with ``eps = 1e-10`` (fullwave/medium.py:273-289, Komatitsch & Martin 2007 CPML form).
    """
    d = np.asarray(d, dtype=np.float64)
    kappa = np.asarray(kappa, dtype=np.float64)
    alpha = np.asarray(alpha, dtype=np.float64)
    b = np.exp(-(d / kappa + alpha) * dt)
    a = d / (kappa * (d + kappa * alpha) + 1e-10) * (b - 1)
    return a, b


def _check_n_relax(n: int) -> None:
    if n not in _SUPPORTED_N_RELAX:
        raise ValueError("Only n_relaxation_mechanisms=2 are supported currently.")


def fw2_map_names(*, is_3d: bool, isotropic: bool, n_relax: int = 2) -> list[str]:
    """Names of the per-voxel relaxation maps the engine consumes, in upload order.

    ``u, w, v`` are the velocity-update side (x, y, z derivative of p) and ``x, y, z`` the
    pressure-update side (divergence of v); the isotropic model shares one ``u`` and one ``x``
    set over all axes (fullwave/medium.py:332-406, fullwave/solver/input_file_writer.py:581-620).
    """
    if isotropic:
        v_axes, p_axes = ["u"], ["x"]
    else:
        v_axes = ["u", "w", "v"] if is_3d else ["u", "w"]
        p_axes = ["x", "y", "z"] if is_3d else ["x", "y"]
    names = [f"kappa_{a}" for a in v_axes + p_axes]
    for nu in range(1, n_relax + 1):
        for a in v_axes + p_axes:
            names += [f"a_pml_{a}{nu}", f"b_pml_{a}{nu}"]
    return names


def sort_mechanisms(params: dict, n: int) -> dict:
    """Per voxel and side (x1 = velocity side, x2 = pressure side), the mechanisms ordered by
    ``d / kappa + alpha``; ties keep their input order (fullwave/medium.py:121-222)."""
    out = {"kappa_x1": np.asarray(params["kappa_x1"]), "kappa_x2": np.asarray(params["kappa_x2"])}
    for side in ("x1", "x2"):
        kappa = np.asarray(params[f"kappa_{side}"], dtype=np.float64)
        d = np.stack([np.asarray(params[f"d_{side}_nu{nu}"], dtype=np.float64) for nu in range(1, n + 1)], -1)
        al = np.stack([np.asarray(params[f"alpha_{side}_nu{nu}"], dtype=np.float64) for nu in range(1, n + 1)], -1)
        tau = d / kappa[..., None] + al
        order = np.argsort(tau, axis=-1, kind="stable")
        d = np.take_along_axis(d, order, -1)
        al = np.take_along_axis(al, order, -1)
        for nu in range(1, n + 1):
            out[f"d_{side}_nu{nu}"] = d[..., nu - 1]
            out[f"alpha_{side}_nu{nu}"] = al[..., nu - 1]
    return {k: out[k] for k in relaxation_param_keys(n)}


class MediumRelaxationMaps:
    """Medium described by per-voxel relaxation parameters (the engine's native input)."""

    def __init__(
        self,
        grid: Grid,
        sound_speed: NDArray[np.float64],
        density: NDArray[np.float64],
        beta: NDArray[np.float64],
        relaxation_param_dict: dict,
        *,
        air_map: NDArray | None = None,
        n_relaxation_mechanisms: int = 2,
        use_isotropic_relaxation: bool = True,
    ) -> None:
        _check_n_relax(n_relaxation_mechanisms)
        self.n_relaxation_mechanisms = n_relaxation_mechanisms
        self.grid = grid
        self.is_3d = grid.is_3d
        self.sound_speed = np.atleast_2d(sound_speed)
        self.density = np.atleast_2d(density)
        self.beta = np.atleast_2d(beta)
        self.air_map = (
            np.zeros_like(self.sound_speed, dtype=bool) if air_map is None else air_map
        )
        self.relaxation_param_dict = initialize_relaxation_param_dict(
            n_relaxation_mechanisms, value=np.zeros_like(self.sound_speed)
        )
        self._update_relaxation_param_dict(relaxation_param_dict)
        self.use_isotropic_relaxation = use_isotropic_relaxation
        self.relaxation_param_dict_for_fw2 = self._calc_relaxation_param_dict_for_fw2(
            use_isotropic_relaxation=use_isotropic_relaxation,
        )
        self.check_fields()

    # --- relaxation parameters -------------------------------------------------------------

    def check_relaxation_param_dict(self, relaxation_param_dict, contents_shape, n_relaxation_mechanisms=2):
        want = set(relaxation_param_keys(n_relaxation_mechanisms))
        got = set(relaxation_param_dict.keys())
        if got != want:
            raise ValueError(f"Unknown relaxation parameter keys: {got - want}")
        for value in relaxation_param_dict.values():
            if np.shape(value) != tuple(contents_shape):
                raise ValueError(
                    "Relaxation parameter map shape error: "
                    f"{np.shape(value)} != {self.sound_speed.shape} (domain shape)",
                )

    def _update_relaxation_param_dict(self, updates: dict) -> None:
        """Store the maps with the mechanisms of each voxel sorted by time constant.

        Sorting is per voxel and per side (x1 = velocity side, x2 = pressure side) by
        ``d / kappa + alpha``; ties keep their input order (fullwave/medium.py:121-222).
        """
        self.check_relaxation_param_dict(updates, self.sound_speed.shape, self.n_relaxation_mechanisms)
        out = sort_mechanisms(updates, self.n_relaxation_mechanisms)
        self.relaxation_param_dict = {k: np.atleast_2d(v) for k, v in out.items()}

    def _calculate_relaxation_coefficients(self) -> dict:
        r = self.relaxation_param_dict
        out = {"kappa_x1": r["kappa_x1"], "kappa_x2": r["kappa_x2"]}
        for nu in range(1, self.n_relaxation_mechanisms + 1):
            for side in ("x1", "x2"):
                a, b = calc_a_and_b(r[f"d_{side}_nu{nu}"], r[f"kappa_{side}"], r[f"alpha_{side}_nu{nu}"], self.grid.dt)
                out[f"a_pml_{side}_nu{nu}"] = a
                out[f"b_pml_{side}_nu{nu}"] = b
        return out

    def _calc_relaxation_param_dict_for_fw2(self, *, use_isotropic_relaxation: bool = True) -> dict:
        """Engine-named maps: every velocity-side axis takes the x1 set, every pressure-side
        axis the x2 set (fullwave/medium.py:332-406)."""
        coef = self._calculate_relaxation_coefficients()
        side_of = {"u": "x1", "w": "x1", "v": "x1", "x": "x2", "y": "x2", "z": "x2"}
        out = {}
        for name in fw2_map_names(is_3d=self.is_3d, isotropic=use_isotropic_relaxation,
                                  n_relax=self.n_relaxation_mechanisms):
            if name.startswith("kappa_"):
                out[name] = coef[f"kappa_{side_of[name[-1]]}"].copy()
            else:
                kind, axis_nu = name[:5], name[6:]  # "a_pml" / "b_pml", "<axis><nu>"
                axis, nu = axis_nu[0], axis_nu[1:]
                out[name] = coef[f"{kind}_{side_of[axis]}_nu{nu}"].copy()
        return out

    # --- properties -------------------------------------------------------------------------

    @property
    def bulk_modulus(self) -> NDArray[np.float64]:
        return np.multiply(self.sound_speed**2, self.density)

    @property
    def n_air(self) -> int:
        return int(np.asarray(self.air_map).sum())

    @property
    def n_coords_zero(self) -> int:
        return self.n_air

    def check_fields(self) -> None:
        shape = tuple(self.grid.shape)
        for name in ("sound_speed", "density", "beta"):
            f = getattr(self, name)
            assert f.shape == shape, f"map shape error: {f.shape} != {shape}"
        for value in self.relaxation_param_dict.values():
            assert value.shape == shape, f"map shape error: {value.shape} != {shape}"

    def build(self) -> MediumRelaxationMaps:
        """Already built; returned as is so a Solver accepts either medium type."""
        return self

    def print_info(self) -> None:
        print(str(self))

    def summary(self) -> None:
        self.print_info()

    def __str__(self) -> str:
        lines = [
            "Relaxation Medium:",
            f"  Grid: {self.grid}",
            f"  Sound speed: min {np.min(self.sound_speed):.2f} m/s, max {np.max(self.sound_speed):.2f} m/s",
            f"  Density: min {np.min(self.density):.2f} kg/m^3, max {np.max(self.density):.2f} kg/m^3",
            f"  Number of air coordinates: {self.n_air}",
            f"  Number of relaxation mechanisms: {self.n_relaxation_mechanisms}",
        ]
        lines += [f"    {k}: min {np.min(v):.4e}, max {np.max(v):.4e}" for k, v in self.relaxation_param_dict.items()]
        return "\n".join(lines) + "\n"

    __repr__ = __str__


class MediumExponentialAttenuation:
    """Medium with a per-voxel single-rate amplitude decay ``alpha_exp`` (per time step)."""

    def __init__(self, grid, sound_speed, density, alpha_exp, beta, *, air_map=None) -> None:
        if not isinstance(grid, Grid):
            raise TypeError("grid must be a Grid")
        self.grid = grid
        self.is_3d = grid.is_3d
        self.sound_speed = np.atleast_2d(sound_speed)
        self.density = np.atleast_2d(density)
        self.alpha_exp = np.atleast_2d(alpha_exp)
        self.beta = np.atleast_2d(beta)
        self.air_map = np.zeros_like(self.sound_speed, dtype=bool) if air_map is None else air_map
        self.check_fields()

    @property
    def bulk_modulus(self):
        return np.multiply(self.sound_speed**2, self.density)

    @property
    def n_air(self) -> int:
        return int(np.asarray(self.air_map).sum())

    @property
    def n_coords_zero(self) -> int:
        return self.n_air

    def check_fields(self) -> None:
        shape = tuple(self.grid.shape)
        for name in ("sound_speed", "density", "alpha_exp", "beta"):
            f = getattr(self, name)
            assert f.shape == shape, f"map shape error: {f.shape} != {shape}"

    def build(self):
        return self


class Medium:
    """Medium given by sound speed, density, power-law attenuation and nonlinearity.

    ``alpha_coeff`` is alpha0 in dB/(MHz^gamma cm) and ``alpha_power`` is gamma. ``build()``
    maps every voxel's (alpha0, gamma) to relaxation parameters. With
    ``path_relaxation_parameters_database=None`` (the default) the table is this package's FP64
    fit, filled lazily on the reference's binning (``searchsorted`` onto the table grid);
    a path to a saved table (``.npz`` or ``.mat`` with ``database``, ``alpha_0_list``,
    ``power_list``, ``invalid_matrix``) uses that table instead.
    """

    def __init__(
        self,
        grid: Grid,
        sound_speed,
        density,
        alpha_coeff,
        alpha_power,
        beta,
        *,
        air_map=None,
        path_relaxation_parameters_database: Path | None = None,
        n_relaxation_mechanisms: int = 2,
        attenuation_builder: str = "lookup",
        use_isotropic_relaxation: bool = True,
    ) -> None:
        _check_n_relax(n_relaxation_mechanisms)
        if not isinstance(grid, Grid):
            raise TypeError("grid must be a Grid")
        if path_relaxation_parameters_database is not None and not Path(path_relaxation_parameters_database).exists():
            raise FileNotFoundError(f"{path_relaxation_parameters_database} does not exist")
        self.grid = grid
        self.is_3d = grid.is_3d
        self.sound_speed = np.atleast_2d(sound_speed)
        self.density = np.atleast_2d(density)
        self.alpha_coeff = np.atleast_2d(alpha_coeff)
        self.alpha_power = np.atleast_2d(alpha_power)
        self.beta = np.atleast_2d(beta)
        self.air_map = np.zeros_like(self.sound_speed, dtype=bool) if air_map is None else air_map
        self.path_relaxation_parameters_database = path_relaxation_parameters_database
        self.n_relaxation_mechanisms = n_relaxation_mechanisms
        self.use_isotropic_relaxation = use_isotropic_relaxation
        self.attenuation_builder = attenuation_builder
        self.check_fields()

    def check_fields(self) -> None:
        shape = tuple(self.grid.shape)
        for name in ("sound_speed", "density", "alpha_coeff", "alpha_power", "beta"):
            f = getattr(self, name)
            assert f.shape == shape, f"map shape error: {f.shape} != {shape}"

    @property
    def bulk_modulus(self):
        return np.multiply(self.sound_speed**2, self.density)

    @property
    def n_air(self) -> int:
        return int(np.asarray(self.air_map).sum())

    @property
    def n_coords_zero(self) -> int:
        return self.n_air

    def build(self) -> MediumRelaxationMaps:
        """Relaxation maps for every voxel's (alpha0, gamma) (fullwave/medium.py:1029-1066)."""
        if self.attenuation_builder != "lookup":
            raise ValueError(
                f"Unknown attenuation_builder: {self.attenuation_builder}. "
                'Only "lookup" is supported currently.',
            )
        from .relaxation import fit_band_for_grid, generate_relaxation_params

        params = generate_relaxation_params(
            alpha_coeff=self.alpha_coeff,
            alpha_power=self.alpha_power,
            n_relaxation_mechanisms=self.n_relaxation_mechanisms,
            path_database=self.path_relaxation_parameters_database,
            dt=self.grid.dt,
            band=fit_band_for_grid(self.grid.dx, self.grid.c0),
        )
        return MediumRelaxationMaps(
            grid=self.grid,
            sound_speed=self.sound_speed,
            density=self.density,
            beta=self.beta,
            relaxation_param_dict=params,
            air_map=self.air_map,
            n_relaxation_mechanisms=self.n_relaxation_mechanisms,
            use_isotropic_relaxation=self.use_isotropic_relaxation,
        )

    def _db_mhz_cm_to_a_exp(self, alpha_coeff):
        """Per-step decay factor ``exp(-dt * alpha0 / 2 * f0 * c0 / (1e-2 * Np))`` with
        ``Np = 10 log10(e)`` (fullwave/medium.py:1068-1090)."""
        np_factor = -10 * np.log10(np.exp(-1))
        f0 = self.grid.omega / (2.0 * np.pi * 1e6)
        texp = alpha_coeff / 2.0 * f0 * self.grid.c0 / (1e-2 * np_factor)
        return np.exp(-self.grid.dt * texp)

    def build_exponential(self) -> MediumExponentialAttenuation:
        return MediumExponentialAttenuation(
            grid=self.grid,
            sound_speed=self.sound_speed,
            density=self.density,
            alpha_exp=self._db_mhz_cm_to_a_exp(self.alpha_coeff),
            beta=self.beta,
            air_map=self.air_map,
        )
