"""Engine output loading and relaxation-dictionary helpers.

``load_dat_data`` mirrors fullwave/solver/utils.py:12-45 (float32 read, NaN/Inf *warning*
only); ``initialize_relaxation_param_dict`` mirrors fullwave/solver/utils.py:68-86, whose key
order is the column order of the relaxation look-up table.
"""

from __future__ import annotations

import logging
from pathlib import Path

import numpy as np
from numpy.typing import DTypeLike, NDArray

logger = logging.getLogger("__main__." + __name__)


def warn_non_finite(data: NDArray) -> None:
    """Log (not raise) when the engine output holds NaN or Inf, like the reference loader."""
    if np.isnan(data).any():
        logger.warning(
            "The simulation contains NaN values. Check the simulation domains or PML settings.",
        )
    if np.isinf(data).any():
        logger.warning(
            "The simulation contains Inf values. Check the simulation domains or PML settings.",
        )


def load_dat_data(dat_file_path: Path, dtype: DTypeLike = np.float32) -> NDArray:
    """Read a raw little-endian ``.dat`` file as a flat array."""
    dat_file_path = Path(dat_file_path)
    if not dat_file_path.exists():
        msg = f"dat_file_path {dat_file_path} does not exist"
        logger.error(msg)
        raise ValueError(msg)
    data = np.fromfile(dat_file_path, dtype=dtype)
    warn_non_finite(data)
    return data


def load_dat_and_reshape(dat_file_path: Path, n_sensors: int, dtype: DTypeLike = np.float32):
    """Read ``genout.dat`` and return it as ``[n_sensors, n_out]``."""
    return load_dat_data(dat_file_path, dtype=dtype).reshape(-1, n_sensors).T


def relaxation_param_keys(n_relaxation_mechanisms: int = 2) -> list[str]:
    """Keys of a relaxation dictionary in look-up-table column order."""
    keys = ["kappa_x1", "kappa_x2"]
    for nu in range(1, n_relaxation_mechanisms + 1):
        keys += [f"d_x1_nu{nu}", f"alpha_x1_nu{nu}", f"d_x2_nu{nu}", f"alpha_x2_nu{nu}"]
    return keys


def initialize_relaxation_param_dict(
    n_relaxation_mechanisms: int = 2,
    value: NDArray | None = None,
) -> dict:
    """Dictionary with every relaxation key, each a copy of ``value`` (or ``None``)."""
    return {
        k: (value.copy() if value is not None else None)
        for k in relaxation_param_keys(n_relaxation_mechanisms)
    }
