"""Pressure source (mirror of fullwave/source.py:18-176).

A source is a mask plus one time signal per masked voxel: ``p0[n_sources, nt]``. Voxel order is
the C order of the mask (``map_to_coords``), which is the row order of ``icmat``.
"""

from __future__ import annotations

import logging

import numpy as np
from numpy.typing import NDArray

from .utils.coordinates import coords_to_map, map_to_coords

logger = logging.getLogger("__main__." + __name__)


class Source:
    """Time-varying pressure imposed on the voxels of ``mask``."""

    def __init__(self, p0: NDArray[np.float64], mask: NDArray[np.bool_]) -> None:
        self.p0 = np.atleast_2d(p0)
        mask = np.atleast_2d(mask)
        self.grid_shape = mask.shape
        self.is_3d = len(self.grid_shape) == 3
        self.incoords = map_to_coords(mask)
        if self.p0.shape[0] != self.incoords.shape[0]:
            raise ValueError("Input signal has the wrong number of elements")

    def validate(self, grid_shape) -> None:
        grid_shape = tuple(grid_shape)
        assert self.mask.shape == grid_shape, f"{self.mask.shape} != {grid_shape}"
        assert np.any(self.mask), "No active source found."

    @property
    def icmat(self) -> NDArray[np.float64]:
        """``[n_sources, nt]`` source signals, the engine's ``icmat``."""
        return self.p0

    @property
    def mask(self) -> NDArray[np.int64]:
        return coords_to_map(self.incoords, grid_shape=self.grid_shape, is_3d=self.is_3d)

    @property
    def n_sources(self) -> int:
        return int(self.incoords.shape[0])

    @property
    def ncoords(self) -> int:
        return self.n_sources

    def print_info(self) -> None:
        print(str(self))

    def summary(self) -> None:
        self.print_info()

    def __str__(self) -> str:
        return (
            "Source: \n"
            f"  Number of sources: {self.n_sources}\n"
            f"  Grid shape: {self.grid_shape}\n"
            f"  Is 3D: {self.is_3d}\n"
            f"  p0 shape: {self.p0.shape}\n"
        )

    __repr__ = __str__
