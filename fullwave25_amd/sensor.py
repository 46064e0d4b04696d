"""Pressure sensors (mirror of fullwave/sensor.py:18-164).

Sensors are the non-zero voxels of a mask, in C order. Pressure is recorded every
``sampling_modulus_time`` steps; the engine writes ``[n_out, n_sensors]`` time-major and the
solver returns ``[n_sensors, n_out]``.
"""

from __future__ import annotations

import logging

import numpy as np
from numpy.typing import NDArray

from .utils.coordinates import coords_to_index_map, coords_to_map, map_to_coords

logger = logging.getLogger("__main__." + __name__)


class Sensor:
    """Pressure recording positions and the recording period in time steps."""

    def __init__(self, mask: NDArray[np.bool_], sampling_modulus_time: int = 1) -> None:
        mask = np.atleast_2d(mask)
        self.grid_shape = mask.shape
        self.sampling_modulus_time = int(sampling_modulus_time)
        self.is_3d = len(self.grid_shape) == 3
        self.outcoords = map_to_coords(mask)

    def validate(self, grid_shape) -> None:
        grid_shape = tuple(grid_shape)
        assert self.mask.shape == grid_shape, f"{self.mask.shape} != {grid_shape}"
        assert np.any(self.mask), "No active sensor found."

    @property
    def mask(self) -> NDArray[np.int64]:
        return coords_to_map(self.outcoords, grid_shape=self.grid_shape, is_3d=self.is_3d)

    @property
    def indexed_mask(self) -> NDArray[np.int64]:
        return coords_to_index_map(self.outcoords, grid_shape=self.grid_shape, is_3d=self.is_3d)

    @property
    def n_sensors(self) -> int:
        return int(self.outcoords.shape[0])

    @property
    def ncoordsout(self) -> int:
        return self.n_sensors

    def print_info(self) -> None:
        print(str(self))

    def summary(self) -> None:
        self.print_info()

    def __str__(self) -> str:
        return (
            "Sensor: \n"
            f"  Number of sensors: {self.n_sensors}\n"
            f"  Grid shape: {self.grid_shape}\n"
            f"  Is 3D: {self.is_3d}\n"
        )

    __repr__ = __str__
