"""Mask <-> coordinate-list conversion.

The order of coordinates is the order of ``np.nonzero`` on a C-order array (row-major, last
axis fastest), which is what the engine contract uses for ``icc``/``outc``/``icczero``
(fullwave/utils/coordinates.py:28-53, 56-84).
"""

from __future__ import annotations

import numpy as np
from numpy.typing import NDArray


def map_to_coords(map_data: NDArray, *, export_as_xyz: bool = False) -> NDArray[np.int64]:
    """Return an ``[n, ndim]`` int64 array of the non-zero positions of ``map_data``.

    Rows are in C order. ``export_as_xyz`` reverses the column order
    (fullwave/utils/coordinates.py:28-53).
    """
    arr = np.asarray(map_data)
    coords = np.argwhere(arr != 0).astype(np.int64)
    if export_as_xyz:
        coords = coords[:, ::-1].copy()
    return coords


def coords_to_map(coords: NDArray, grid_shape, *, is_3d: bool = False) -> NDArray[np.int64]:
    """Inverse of :func:`map_to_coords`: a 0/1 int map with ones at ``coords``."""
    shape = tuple(int(s) for s in grid_shape[: 3 if is_3d else 2])
    mask = np.zeros(shape, dtype=int)
    c = np.asarray(coords).astype(int)
    if c.size:
        mask[tuple(c[:, a] for a in range(len(shape)))] = 1
    return mask


def coords_to_index_map(coords: NDArray, grid_shape, *, is_3d: bool = False) -> NDArray[np.int64]:
    """Map of 1-based coordinate indices (0 where no coordinate), as in
    fullwave/utils/coordinates.py:87-102."""
    shape = tuple(int(s) for s in grid_shape[: 3 if is_3d else 2])
    mask = np.zeros(shape, dtype=int)
    c = np.asarray(coords).astype(int)
    if c.size:
        mask[tuple(c[:, a] for a in range(len(shape)))] = np.arange(c.shape[0]) + 1
    return mask
