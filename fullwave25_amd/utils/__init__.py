"""Host-side helpers (coordinates, pulses, rounding)."""

from . import coordinates, numerical, pulse
from .coordinates import coords_to_index_map, coords_to_map, map_to_coords
from .numerical import matlab_round
from .pulse import gaussian_modulated_sinusoidal_signal

__all__ = [
    "coordinates",
    "coords_to_index_map",
    "coords_to_map",
    "gaussian_modulated_sinusoidal_signal",
    "map_to_coords",
    "matlab_round",
    "numerical",
    "pulse",
]
