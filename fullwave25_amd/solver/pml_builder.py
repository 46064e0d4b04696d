"""Padding of the medium and the stretched-coordinate (CPML) absorbing layer.

Behaviour follows fullwave/solver/pml_builder.py:

* every map is padded by ``nb = n_transition + n_pml + m_spatial_order`` cells per side; maps
  are padded by edge replication, masks (source, sensor, air) by zeros
  (``_extend_map_for_pml``, :321-704 — its face/edge/corner fills are edge replication);
* the relaxation parameters are ramped into the absorbing layer (``_apply_pml``, :842-1025;
  ``_apply_pml_3d``, :1027-1254; ``_apply_transition_and_pml``, :1256-1498), for every map and
  along every axis in turn:

  - mechanism 1, ``d``: polynomial (n = 2) ramp to
    ``d_target = -(n + 1) c0 ln(R) / (2 (L_pml + L_transition))``, ``R = 1e-30``, over the PML
    plus transition layers;
  - mechanism 1, ``alpha``: linear ramp to 0 over the same layers;
  - mechanisms >= 2, ``d`` and ``alpha``: cosine ramp to 0 inside the transition layer, 0 in
    the PML;
  - ``kappa`` unchanged;

  then ``a, b`` are recomputed from the ramped maps.

Source and sensor coordinates shift by ``nb`` on every axis; their C order is unchanged, so the
padded coordinate lists are computed directly instead of through padded masks.
"""

from __future__ import annotations

import logging
from functools import cached_property

import numpy as np
from numpy.typing import NDArray

from ..grid import Grid
from ..medium import Medium, MediumExponentialAttenuation, MediumRelaxationMaps, calc_a_and_b, fw2_map_names
from ..sensor import Sensor
from ..source import Source

logger = logging.getLogger("__main__." + __name__)


def _ramp(kind: str, n: int, order: float = 2) -> NDArray[np.float64]:
    """Transition profile sampled at ``linspace(0, 1, n + 1)``
    (fullwave/solver/pml_builder.py:20-38)."""
    x = np.linspace(0, 1, n + 1)
    if kind == "linear":
        return x
    if kind == "polynomial":
        return x**order
    if kind == "cosine":
        return 0.5 * (1 - np.cos(np.pi * x))
    if kind == "smooth":
        def part(y):
            return np.where(y > 0, np.exp(-1 / (y + 1e-20)), 0)

        return part(x) / (part(x) + part(1 - x))
    raise ValueError(
        f"Invalid transition type: {kind}. Choose from 'smooth', 'linear', or 'polynomial'.",
    )


def pad_map(a: NDArray, nb: int, *, fill_edge: bool = True) -> NDArray[np.float64]:
    """Pad ``nb`` cells on every side: edge replication for maps, zeros for masks."""
    a = np.asarray(a, dtype=np.float64)
    return np.pad(a, nb, mode="edge") if fill_edge else np.pad(a, nb, mode="constant")


class _LayerShape:
    def __init__(self, m_spatial_order: int, n_pml: int, n_transition: int) -> None:
        self.m = m_spatial_order
        self.n_pml = n_pml
        self.n_tr = n_transition


def apply_transition_and_pml(
    values: NDArray[np.float64],
    target: float,
    axis: int,
    layers: _LayerShape,
    *,
    kind: str,
    order: float = 2,
    region: str = "both",
) -> NDArray[np.float64]:
    """Ramp ``values`` toward ``target`` near both ends of ``axis`` (in place, returned).

    ``region``: ``"both"`` ramps over PML + transition from the ghost border inward,
    ``"transition"`` ramps over the transition layer only and holds ``target`` in the PML,
    ``"pml"`` ramps over the PML only. The low side's ramp reaches the interior value one cell
    before the untouched interior; the high side's starts at the last untouched cell; beyond the
    ramps the value is ``target`` (fullwave/solver/pml_builder.py:1342-1398).
    """
    if region == "transition":
        if layers.n_tr == 0:
            raise ValueError(
                "Transition layer is not defined. "
                "Set transit_within_transition_layer to False or define n_transition_layer.",
            )
        thick, off = layers.n_tr, layers.n_pml
    elif region == "pml":
        thick, off = layers.n_pml, 0
    else:
        thick, off = layers.n_pml + layers.n_tr, 0
    f = _ramp(kind, thick, order)
    v = np.moveaxis(values, axis, 0)  # view
    n = v.shape[0]
    m = layers.m
    v[: m + off + thick] = target
    v[n - m - thick - off :] = target
    lo = slice(m + off - 1, m + off + thick)
    hi_start = n - m - thick - off - 1
    hi = slice(hi_start, n - m - off)
    lo_val = v[m + off + thick][None].copy()
    hi_val = v[hi_start][None].copy()
    shape = (-1,) + (1,) * (v.ndim - 1)
    v[lo] = lo_val - f[::-1].reshape(shape) * (lo_val - target)
    v[hi] = hi_val - f.reshape(shape) * (hi_val - target)
    return values


class PMLBuilder:
    """Pads grid, medium, source and sensor and builds the absorbing layer maps."""

    def __init__(
        self,
        grid: Grid,
        medium: Medium | MediumRelaxationMaps,
        source: Source,
        sensor: Sensor,
        *,
        m_spatial_order: int = 8,
        n_pml_layer: int = 40,
        n_transition_layer: int = 40,
        use_isotropic_relaxation: bool = False,
    ) -> None:
        if not isinstance(grid, Grid):
            raise TypeError("grid must be a Grid")
        if not isinstance(medium, (Medium, MediumRelaxationMaps)):
            raise TypeError("medium must be a Medium or MediumRelaxationMaps")
        if not isinstance(source, Source) or not isinstance(sensor, Sensor):
            raise TypeError("source/sensor have the wrong type")
        self.grid_org = grid
        self.medium_org = medium
        self.source_org = source
        self.sensor_org = sensor
        self.is_3d = grid.is_3d
        self.use_isotropic_relaxation = use_isotropic_relaxation
        self.m_spatial_order = m_spatial_order
        self.n_pml_layer = n_pml_layer
        self.n_transition_layer = n_transition_layer

        nb = self.num_boundary_points
        spacing = (grid.dx, grid.dy, grid.dz)
        domain_size = tuple(
            (medium.sound_speed.shape[a] + 2 * nb) * spacing[a] for a in range(3 if self.is_3d else 2)
        )
        self.extended_grid = Grid(
            domain_size=domain_size, f0=grid.f0, duration=grid.duration,
            c0=grid.c0, ppw=grid.ppw, cfl=grid.cfl,
        )
        self.extended_medium = self._extend_medium(medium)
        self.extended_source = _shifted_source(source, nb)
        self.extended_sensor = _shifted_sensor(sensor, nb)
        self.pml_layer_m = self.extended_grid.dx * self.n_pml_layer
        self.transition_layer_m = self.extended_grid.dx * self.n_transition_layer
        self.n_polynomial = 2
        self.theoritical_reflection_coefficient = 10 ** (-30)
        if self.n_pml_layer == 0:
            self.n_transition_layer = 0

    @cached_property
    def num_boundary_points(self) -> int:
        return self.n_transition_layer + self.n_pml_layer + self.m_spatial_order

    @property
    def nx(self) -> int:
        return self.extended_grid.nx

    @property
    def ny(self) -> int:
        return self.extended_grid.ny

    @property
    def nz(self) -> int:
        return self.extended_grid.nz

    @property
    def nt(self) -> int:
        return self.extended_grid.nt

    @property
    def n_sources(self) -> int:
        return self.extended_source.n_sources

    @property
    def n_sensors(self) -> int:
        return self.extended_sensor.n_sensors

    @property
    def n_air(self) -> int:
        return self.extended_medium.n_air

    def _extend_map_for_pml(self, a, *, fill_edge: bool = True):
        return pad_map(a, self.num_boundary_points, fill_edge=fill_edge)

    def _extend_medium(self, medium):
        ext = self._extend_map_for_pml
        if isinstance(medium, MediumRelaxationMaps):
            return MediumRelaxationMaps(
                grid=self.extended_grid,
                sound_speed=ext(medium.sound_speed),
                density=ext(medium.density),
                beta=ext(medium.beta),
                relaxation_param_dict={k: ext(v) for k, v in medium.relaxation_param_dict.items()},
                air_map=ext(medium.air_map, fill_edge=False),
                n_relaxation_mechanisms=medium.n_relaxation_mechanisms,
            )
        return Medium(
            grid=self.extended_grid,
            sound_speed=ext(medium.sound_speed),
            density=ext(medium.density),
            beta=ext(medium.beta),
            alpha_coeff=ext(medium.alpha_coeff),
            alpha_power=ext(medium.alpha_power),
            air_map=ext(medium.air_map, fill_edge=False),
            n_relaxation_mechanisms=medium.n_relaxation_mechanisms,
            path_relaxation_parameters_database=medium.path_relaxation_parameters_database,
            attenuation_builder=medium.attenuation_builder,
        )

    def _localize_pml_region(self):
        """Per-axis PML depth masks in [0, 1) (1 in the ghost border), as
        fullwave/solver/pml_builder.py:715-792; used for inspection only."""
        shape = self.extended_medium.sound_speed.shape
        m, n = self.m_spatial_order, self.n_pml_layer
        out = []
        for axis in range(len(shape)):
            prof = np.zeros(shape[axis])
            for i in range(n):
                prof[i + shape[axis] - m - n] = i / n
                prof[m + n - i - 1] = i / n
            prof[:m] = 1
            prof[shape[axis] - m:] = 1
            view = [1] * len(shape)
            view[axis] = -1
            out.append(np.broadcast_to(prof.reshape(view), shape).copy())
        return tuple(out)

    def run(self, *, use_pml: bool = True) -> MediumRelaxationMaps:
        """Relaxation maps on the padded grid with the absorbing layer applied."""
        built = self.extended_medium.build()
        if use_pml:
            return self._apply_pml(built)
        return built

    def _apply_pml(self, built: MediumRelaxationMaps) -> MediumRelaxationMaps:
        layers = _LayerShape(self.m_spatial_order, self.n_pml_layer, self.n_transition_layer)
        n_axes = 3 if self.is_3d else 2
        d_target = (
            -(self.n_polynomial + 1)
            * self.extended_grid.c0
            * np.log(self.theoritical_reflection_coefficient)
            / (2 * (self.pml_layer_m + self.transition_layer_m))
        )
        rp = built.relaxation_param_dict
        n_relax = built.n_relaxation_mechanisms
        # one ramped set per side; per-axis engine maps share it (the reference ramps every
        # per-axis copy identically, along every axis)
        ramped = {}
        for side in ("x1", "x2"):
            ramped[f"kappa_{side}"] = rp[f"kappa_{side}"].copy()
            for nu in range(1, n_relax + 1):
                for var in ("d", "alpha"):
                    arr = rp[f"{var}_{side}_nu{nu}"].copy()
                    for axis in range(n_axes):
                        if nu >= 2:
                            apply_transition_and_pml(arr, 0.0, axis, layers, kind="cosine", region="transition")
                        elif var == "d":
                            apply_transition_and_pml(arr, d_target, axis, layers, kind="polynomial",
                                                     order=self.n_polynomial, region="both")
                        else:
                            apply_transition_and_pml(arr, 0.0, axis, layers, kind="linear", region="both")
                    ramped[f"{var}_{side}_nu{nu}"] = arr
        ab = {}
        for side in ("x1", "x2"):
            for nu in range(1, n_relax + 1):
                ab[(side, nu)] = calc_a_and_b(ramped[f"d_{side}_nu{nu}"], ramped[f"kappa_{side}"],
                                              ramped[f"alpha_{side}_nu{nu}"], built.grid.dt)
        side_of = {"u": "x1", "w": "x1", "v": "x1", "x": "x2", "y": "x2", "z": "x2"}
        out = {}
        for name in fw2_map_names(is_3d=self.is_3d, isotropic=self.use_isotropic_relaxation, n_relax=n_relax):
            if name.startswith("kappa_"):
                out[name] = ramped[f"kappa_{side_of[name[-1]]}"].copy()
            else:
                axis, nu = name[6], int(name[7:])
                a, b = ab[(side_of[axis], nu)]
                out[name] = (a if name[0] == "a" else b).copy()
                out[f"d_{axis}_nu{nu}"] = ramped[f"d_{side_of[axis]}_nu{nu}"]
                out[f"alpha_{axis}_nu{nu}"] = ramped[f"alpha_{side_of[axis]}_nu{nu}"]
        built.relaxation_param_dict_for_fw2.update(out)
        return built


def _shifted_source(source: Source, nb: int) -> Source:
    new = Source.__new__(Source)
    new.p0 = source.p0
    new.grid_shape = tuple(s + 2 * nb for s in source.grid_shape)
    new.is_3d = source.is_3d
    new.incoords = source.incoords + nb
    return new


def _shifted_sensor(sensor: Sensor, nb: int) -> Sensor:
    new = Sensor.__new__(Sensor)
    new.grid_shape = tuple(s + 2 * nb for s in sensor.grid_shape)
    new.sampling_modulus_time = sensor.sampling_modulus_time
    new.is_3d = sensor.is_3d
    new.outcoords = sensor.outcoords + nb
    return new


class PMLBuilderExponentialAttenuation(PMLBuilder):
    """Padding and absorbing layer for the exponential-attenuation model
    (fullwave/solver/pml_builder.py:1652-1935).

    The border is ``n_pml + m_spatial_order`` cells (no transition layer); maps, and here also
    the air map, are padded by edge replication. ``run`` turns alpha0 into the per-step decay
    factor ``exp(-dt alpha0/2 f0 c0 / (1e-2 * 10 log10 e))`` (``Medium.build_exponential``) and
    scales it by ``1 - m``, where ``m`` is the Euclidean distance into the border, counted from
    1 at the first border cell, normalized to 1 at the grid corner.
    """

    def __init__(self, grid, medium, source, sensor, *, m_spatial_order: int = 8, n_pml_layer: int = 40):
        if not isinstance(medium, Medium):
            raise TypeError("medium must be a Medium")
        super().__init__(grid, medium, source, sensor, m_spatial_order=m_spatial_order,
                         n_pml_layer=n_pml_layer, n_transition_layer=0)
        # the reference pads the air map of this model like a medium map
        self.extended_medium.air_map = self._extend_map_for_pml(medium.air_map)

    @staticmethod
    def _mask_body(shape, n_body: int) -> NDArray[np.float64]:
        dist2 = 0.0
        for axis, n in enumerate(shape):
            x = np.arange(1, n + 1, dtype=np.float64)
            r = np.where(x <= n_body, n_body - x + 1, np.where(x > n - n_body, x - (n - n_body), 0))
            view = [1] * len(shape)
            view[axis] = -1
            dist2 = dist2 + (r**2).reshape(view)
        mask = np.sqrt(np.broadcast_to(dist2, shape))
        if mask.max() > 0:
            mask = mask / mask.max()
        return mask

    def run(self, *, use_pml: bool = True) -> MediumExponentialAttenuation:
        built = self.extended_medium.build_exponential()
        if use_pml:
            built.alpha_exp = built.alpha_exp * (1 - self._mask_body(built.alpha_exp.shape, self.num_boundary_points))
        return built
