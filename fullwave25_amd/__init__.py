"""fullwave25_amd — an MI355X-native time-stepping engine for Fullwave 2.5.

The front-end classes keep the reference's API (``fullwave/__init__.py:31-45``): build a
:class:`Grid`, a :class:`Medium` (or :class:`MediumRelaxationMaps`), a :class:`Source` and a
:class:`Sensor`, then ``Solver(...).run()``. The time stepping runs in hand-written HIP kernels
for gfx950 behind the C ABI of ``include/fullwave_hip.h``.
"""

import logging

from . import utils
from .grid import Grid
from .medium import Medium, MediumExponentialAttenuation, MediumRelaxationMaps
from .sensor import Sensor
from .solver.launcher import HipLauncher, SimulationError
from .solver.solver import Solver
from .source import Source

__version__ = "0.1.0"
VERSION = __version__

logger = logging.getLogger("__main__." + __name__)

__all__ = [
    "Grid",
    "HipLauncher",
    "Medium",
    "MediumExponentialAttenuation",
    "MediumRelaxationMaps",
    "Sensor",
    "SimulationError",
    "Solver",
    "Source",
    "utils",
]
