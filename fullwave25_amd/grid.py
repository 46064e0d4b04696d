"""Computational grid (mirror of fullwave/grid.py:11-227).

dx = dy = dz = c0 / f0 / ppw, dt = cfl * dx / c0, nt = round(duration / dt).
"""

from __future__ import annotations

import logging
from dataclasses import dataclass

import numpy as np

logger = logging.getLogger("__main__." + __name__)


@dataclass
class Grid:
    """Grid size, spacing and time step of a 2D or 3D simulation.

    Same constructor and properties as the reference ``fullwave.Grid``
    (fullwave/grid.py:11-137), including its input checks (fullwave/grid.py:60-70).
    """

    domain_size: tuple
    f0: float
    duration: float
    c0: float = 1540
    ppw: int = 12
    cfl: float = 0.2

    def __post_init__(self) -> None:
        self._check_input()
        self.is_3d = len(self.domain_size) == 3

    def _check_input(self) -> None:
        assert len(self.domain_size) in [2, 3], "Domain size must be 2D or 3D"
        assert self.f0 > 0, "Center frequency must be greater than 0"
        assert self.f0 < 30e6, "Center frequency must be less than 30 MHz"
        assert self.duration > 0, "Duration must be greater than 0"
        assert self.c0 > 0, "Speed of sound must be greater than 0"
        assert self.ppw > 5, "Points per wavelength must be greater than 5"
        assert self.cfl > 0.01, "CFL number must be greater than 0.01"
        assert self.cfl < 0.5, "CFL number must be less than 0.5"

    @property
    def omega(self) -> float:
        return 2 * np.pi * self.f0

    @property
    def wavelength(self) -> float:
        return self.c0 / self.f0

    @property
    def dx(self) -> float:
        return self.wavelength / self.ppw

    @property
    def dy(self) -> float:
        return self.wavelength / self.ppw

    @property
    def dz(self) -> float:
        return self.wavelength / self.ppw if len(self.domain_size) == 3 else 0

    @property
    def nx(self) -> int:
        return int(np.round(self.domain_size[0] / self.dx).item())

    @property
    def ny(self) -> int:
        return int(np.round(self.domain_size[1] / self.dy).item())

    @property
    def nz(self) -> int:
        if len(self.domain_size) != 3:
            return 0
        return int(np.round(self.domain_size[2] / self.dz).item())

    @property
    def n_axial(self) -> int:
        return self.nx

    @property
    def n_lateral(self) -> int:
        return self.ny

    @property
    def n_elevational(self) -> int:
        return self.nz

    @property
    def dt(self) -> float:
        return self.cfl * self.dx / self.c0

    @property
    def nt(self) -> int:
        return round(self.duration / self.dt)

    @property
    def t(self) -> np.ndarray:
        return np.arange(0, self.duration, self.dt)

    @property
    def shape(self) -> tuple:
        return (self.nx, self.ny, self.nz) if self.is_3d else (self.nx, self.ny)

    def print_info(self) -> None:
        print(str(self))

    def summary(self) -> None:
        self.print_info()

    def __str__(self) -> str:
        dims = f"({self.nx}, {self.ny}, {self.nz})" if self.is_3d else f"({self.nx}, {self.ny})"
        return (
            "Grid Information:\n"
            f"  Domain size (grid.domain_size): {tuple(self.domain_size)} m\n"
            f"  Center frequency (grid.f0): {self.f0 / 1e6} MHz\n"
            f"  Duration (grid.duration): {self.duration:.2e} s\n"
            f"  Speed of sound (grid.c0): {self.c0} m/s\n"
            f"  Points per wavelength (grid.ppw): {self.ppw}\n"
            f"  CFL number (grid.cfl): {self.cfl}\n"
            f"  Grid spacing (grid.dx): {self.dx:.4e} m\n"
            f"  Number of grid points: {dims}\n"
            f"  Time step (grid.dt): {self.dt:.2e} s\n"
            f"  Number of time steps (grid.nt): {self.nt}\n"
            f"  is 3D simulation (grid.is_3d): {self.is_3d}"
        )

    __repr__ = __str__
