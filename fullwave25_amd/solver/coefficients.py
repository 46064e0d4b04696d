"""Finite-difference coefficient tables of the staggered 16-point stencil.

The engine's spatial operator along one axis, at a staggered point, is

    D f = sum_{k=1..8} d_k [f(+k) - f(-k+1)] + e * sum_{transverse lines} [f(+1) - f(0)]

where ``d_1..d_8`` are the axial weights and ``e`` (``d[1, 1]`` in the reference's 9x2 table)
weights the nearest-neighbour difference on each of the 2 (2D) or 4 (3D) transverse lines. The
weights are degree-7 polynomials of the local Courant number ``r = c dt / dx``; the polynomial
coefficients below are the reference's values (fullwave/solver/input_file_writer.py:183-367 for
``d`` at ``r = cfl``, :369-556 for ``dmap`` at ``r = (c_min + i) dt / dx``). They are data, so
they must be kept identical; they are stored as a table and evaluated with Horner's rule.

``dcmap`` is the per-voxel index ``round(c) - round(c_min)`` into ``dmap``'s last axis
(fullwave/solver/input_file_writer.py:558-559, 816-821), with MATLAB-style rounding.
"""

from __future__ import annotations

import numpy as np
from numpy.typing import NDArray

from ..utils.numerical import matlab_round

# rows: d1..d8, e ; columns: coefficient of r^7, r^6, ..., r^1, r^0
_POLY_2D = np.array(
    [
        [-0.000874634088067635, -0.00180530560296097, -0.000440512972481673, 0.00474018847663366,
         -1.93097802254349e-05, -0.292328221171893, -6.58101498708345e-08, 1.25420636437969],
        [0.000793317828964018, 0.00161433256585486, 0.000397244786277123, 0.00546057645976549,
         1.73781972873916e-05, 0.0588754971188371, 5.91706982879834e-08, -0.123406473759703],
        [-0.000650217700538851, -0.00116449260340413, -0.000324403734066325, -0.00911483710059994,
         -1.417399823126e-05, 0.0233184077551615, -4.82326094707544e-08, 0.0346342451534453],
        [0.000467529510541428, 0.000732736676632388, 0.000232444388955328, 0.00846419766685254,
         1.01438593426278e-05, -0.0317586249260511, 3.44988852042879e-08, -0.0119674942518101],
        [-0.000298416281187033, -0.000399380750669364, -0.000148203388388213, -0.00601788793192501,
         -6.46543538517443e-06, 0.0241912754935119, -2.19855171569984e-08, 0.00415554391204146],
        [0.000167882669698981, 0.000188195874702691, 8.3057921860396e-05, 0.00348461963201376,
         3.61873162287129e-06, -0.0149875789940005, 1.22979142197165e-08, -0.00129213888778954],
        [-6.22209937489143e-05, -6.44890425871692e-05, -3.02936928954918e-05, -0.00133386143898282,
         -1.31215186728213e-06, 0.00670228205200379, -4.44653967516776e-09, 0.000315659916047599],
        [6.8474088109024e-06, 1.14082245705934e-05, 3.0372759370575e-06, 0.000236122782444105,
         1.26768491232397e-07, -0.00153347270556276, 4.21617557752767e-10, -4.51948990428065e-05],
        [2.13188763071246e-06, -7.41025068776257e-05, 2.31652037371554e-06, -0.00259495924602038,
         1.20637183170338e-07, 0.0521123771632193, 4.42258843694177e-10, -4.20967682664542e-07],
    ],
    dtype=np.float64,
)

_POLY_3D = np.array(
    [
        [3.26627215252963e-3, -7.91679373564790e-4, 1.08663532410570e-3, 2.54974226454794e-2,
         3.23083288193913e-5, -3.97704676886853e-1, 7.95584310128586e-8, 1.25425295688331],
        [-2.83291379048757e-3, 8.52796449228369e-4, -9.45353822586534e-4, -8.82015372858580e-3,
         -2.81364895458027e-5, 6.73021045987599e-2, -6.93180036837075e-8, -1.23448809066664e-1],
        [2.32775473203342e-3, -5.56793042789852e-4, 7.77649035879584e-4, 2.45547234243566e-3,
         2.31537892801923e-5, 1.61900960524164e-2, 5.70523152308121e-8, 3.46683979649506e-2],
        [-1.68883462553539e-3, 3.03535823592644e-4, -5.64777117315819e-4, 2.44582905523866e-4,
         -1.68215579314751e-5, -2.62344345204941e-2, -4.14559953526389e-8, -1.19918511290930e-2],
        [1.08994931098070e-3, -1.41445142143525e-4, 3.64794490139160e-4, -8.86057426195227e-4,
         1.08681882832738e-5, 2.07238558666603e-2, 2.67876079477806e-8, 4.17058420250698e-3],
        [-6.39950124405340e-4, 6.06079815415080e-5, -2.14633466007892e-4, 6.84580412267934e-4,
         -6.39907927898092e-6, -1.29825288653404e-2, -1.57775422151124e-8, -1.29998325971518e-3],
        [2.92716539609611e-4, -1.87446062803024e-5, 9.85389372183761e-5, -2.40360290348543e-4,
         2.94166215515130e-6, 5.57066438452790e-3, 7.25741366376659e-9, 3.18698432679400e-4],
        [-6.42183857909518e-5, 3.38552867751042e-6, -2.17377151411164e-5, 4.98269067389945e-5,
         -6.50197868987757e-7, -1.19096089679178e-3, -1.60559948991172e-9, -4.57795411807702e-5],
        [-4.47723278782936e-5, -7.69502473399932e-5, -1.41765498250133e-5, -2.54672045901272e-3,
         -4.14343385915353e-7, 5.00210047924752e-2, -1.01220354410507e-9, -8.07139347787336e-8],
    ],
    dtype=np.float64,
)

N_AXIAL = 8  # half-width of the axial stencil; also the ghost-cell border of the grid


def stencil_weights(r, *, is_3d: bool) -> NDArray[np.float64]:
    """Return ``[..., 9]``: the 8 axial weights ``d_1..d_8`` then the transverse weight ``e``
    for Courant number(s) ``r``."""
    poly = _POLY_3D if is_3d else _POLY_2D
    r = np.asarray(r, dtype=np.float64)[..., None]
    acc = np.zeros(r.shape[:-1] + (9,), dtype=np.float64)
    for col in range(poly.shape[1]):
        acc = acc * r + poly[:, col]
    return acc


def n_transverse_lines(is_3d: bool) -> int:
    """Number of transverse nearest-neighbour lines the ``e`` term sums over."""
    return 4 if is_3d else 2


def d_matrix(cfl: float, *, is_3d: bool) -> NDArray[np.float64]:
    """The reference's ``d`` table ``[9, 2]``: rows 1..8 of column 0 are ``d_1..d_8`` at
    ``r = cfl`` and ``[1, 1]`` is ``e``; everything else is zero
    (fullwave/solver/input_file_writer.py:183-367)."""
    w = stencil_weights(cfl, is_3d=is_3d)
    d = np.zeros((9, 2))
    d[1:9, 0] = w[:8]
    d[1, 1] = w[8]
    return d


def dmap_dimension(sound_speed: NDArray) -> int:
    """``round(c_max) - round(c_min)`` (fullwave/solver/input_file_writer.py:94-97)."""
    c = np.asarray(sound_speed)
    return int(matlab_round(c.max()) - matlab_round(c.min()))


def d_map(sound_speed: NDArray, dt: float, dx: float, *, is_3d: bool) -> NDArray[np.float64]:
    """``dmap [9, 2, dim + 1]``: the ``d`` table at ``r = (c_min + i) dt / dx`` for
    ``i = 0..dim`` (fullwave/solver/input_file_writer.py:369-556)."""
    dim = dmap_dimension(sound_speed)
    cmin = float(np.asarray(sound_speed).min())
    r = (np.arange(dim + 1) + cmin) * dt / dx
    w = stencil_weights(r, is_3d=is_3d)  # [dim+1, 9]
    out = np.zeros((9, 2, dim + 1))
    out[1:9, 0, :] = w[:, :8].T
    out[1, 1, :] = w[:, 8]
    return out


def dc_map(sound_speed: NDArray) -> NDArray[np.int64]:
    """0-based per-voxel index into ``dmap``: ``round(c) - round(c_min)``
    (fullwave/solver/input_file_writer.py:558-559 stores it 1-based and writes ``-1``)."""
    c = np.asarray(sound_speed)
    return matlab_round(c) - matlab_round(c.min())


def ndmap_value(sound_speed: NDArray) -> int:
    """Number of coefficient sets the engine reads (``ndmap.dat``): 1 when ``c`` is uniform
    after rounding, else ``dim + 1`` (fullwave/solver/input_file_writer.py:809-813)."""
    dim = dmap_dimension(sound_speed)
    return 1 if dim == 0 else dim + 1


def pack_coefficient_sets(dmap: NDArray[np.float64]) -> NDArray[np.float64]:
    """``[ndmap, 9]`` rows of ``(d_1..d_8, e)`` from a ``[9, 2, ndmap]`` ``dmap`` table."""
    dmap = np.asarray(dmap)
    out = np.empty((dmap.shape[2], 9))
    out[:, :8] = dmap[1:9, 0, :].T
    out[:, 8] = dmap[1, 1, :]
    return out
