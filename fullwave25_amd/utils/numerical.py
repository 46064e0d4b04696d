"""Small numerical helpers shared by the host-side front end.

``matlab_round`` follows fullwave/utils/numerical.py:20-22 (round half away from zero by
adding 1e-9 before numpy's round, then casting to int).
"""

from __future__ import annotations

import numpy as np


def matlab_round(value):
    """Round like MATLAB's ``round`` for non-negative inputs (fullwave/utils/numerical.py:20-22)."""
    return np.round(np.asarray(value) + 1e-9).astype(int)
