"""Relaxation parameters for power-law attenuation, fitted on the host in FP64.

The reference maps every voxel's (alpha0, gamma) to 10 relaxation parameters by a
``searchsorted`` lookup into a precomputed table (fullwave/utils/relaxation_parameters.py:18-242)
that it does not ship (``.MISSING_LARGE_BLOBS``). This module produces that table itself.

Model (README.md:36-44 of the reference; Komatitsch & Martin 2007). Each side n of the first-order
system (n = 1: gradient in the velocity update, n = 2: divergence in the pressure update) uses a
stretched derivative with, for time dependence exp(-i w t),

    S_n(w) = 1 / s_n(w) = 1 / kappa_n - sum_nu (d_n,nu / kappa_n^2) / (d_n,nu / kappa_n + alpha_n,nu - i w)

so a plane wave has wavenumber k(w) = (w / c) / sqrt(S_1 S_2) and attenuation Im k. The fit
chooses ``kappa_1, kappa_2`` and ``d, alpha`` of two mechanisms per side so that Im k matches
alpha0 f^gamma (dB/(MHz^gamma cm) -> Np/m) over 1-20 MHz in relative terms, while the mean
phase velocity over the band stays c. The reference frequency for the conversion is
c = 1540 m/s, as the table has no sound-speed axis.

Column order of a table row (the reference's, fullwave/solver/utils.py:68-86):
``kappa_x1, kappa_x2, d_x1_nu1, alpha_x1_nu1, d_x2_nu1, alpha_x2_nu1, d_x1_nu2, ...``.
"""

from __future__ import annotations

import logging
import threading
from pathlib import Path

import numpy as np
from numpy.typing import NDArray
from scipy.optimize import least_squares

from .solver.utils import initialize_relaxation_param_dict, relaxation_param_keys

logger = logging.getLogger("__main__." + __name__)

NP_PER_DB = 1.0 / (20.0 * np.log10(np.e))  # 1 dB = 0.1151 Np
C_REF = 1540.0
# the band over which the attenuation must follow alpha0 f^gamma, and how far past both of its
# ends the fit reaches: a least-squares fit is worst at the ends of its frequency range, so
# fitting over [f_lo / 1.25, 1.25 f_hi] moves that edge error out of the band (at gamma = 1.6
# the worst error inside 1-20 MHz drops from about 9 % to about 3.5 %)
FIT_BAND_HZ = (1.0e6, 20.0e6)
FIT_MARGIN = 1.25
DEFAULT_FIT_RANGE_HZ = (FIT_BAND_HZ[0] / FIT_MARGIN, FIT_BAND_HZ[1] * FIT_MARGIN)
N_FIT_FREQ = 48

# table grid (the binning the reference's searchsorted lookup implies)
ALPHA_0_LIST = np.round(np.arange(0, 201) * 0.01, 10)  # 0.00 .. 2.00 dB/(MHz^gamma cm)
POWER_LIST = np.round(np.arange(10, 201) * 0.01, 10)  # 0.10 .. 2.00


def target_attenuation(f_hz, alpha0: float, gamma: float) -> NDArray[np.float64]:
    """alpha0 (f / 1 MHz)^gamma in Np/m."""
    return alpha0 * (np.asarray(f_hz) / 1e6) ** gamma * NP_PER_DB * 100.0


def stretch_inverse(w, kappa, d, alpha) -> NDArray[np.complex128]:
    """S(w) = 1/kappa - sum_nu (d_nu / kappa^2) / (d_nu / kappa + alpha_nu - i w)."""
    w = np.asarray(w, dtype=np.float64)
    out = np.full(w.shape, 1.0 / kappa, dtype=np.complex128)
    for dn, an in zip(d, alpha):
        out -= (dn / kappa**2) / (dn / kappa + an - 1j * w)
    return out


def model_wavenumber(f_hz, params: NDArray[np.float64], c: float = C_REF, n_relax: int = 2):
    """Complex wavenumber of the relaxation model for one table row ``params``."""
    w = 2 * np.pi * np.asarray(f_hz, dtype=np.float64)
    k1, k2 = params[0], params[1]
    d1 = [params[2 + 4 * i] for i in range(n_relax)]
    a1 = [params[3 + 4 * i] for i in range(n_relax)]
    d2 = [params[4 + 4 * i] for i in range(n_relax)]
    a2 = [params[5 + 4 * i] for i in range(n_relax)]
    s1 = stretch_inverse(w, k1, d1, a1)
    s2 = stretch_inverse(w, k2, d2, a2)
    return (w / c) / np.sqrt(s1 * s2)


def model_attenuation(f_hz, params, c: float = C_REF) -> NDArray[np.float64]:
    """Im k in Np/m."""
    return np.imag(model_wavenumber(f_hz, params, c))


def model_phase_velocity(f_hz, params, c: float = C_REF) -> NDArray[np.float64]:
    w = 2 * np.pi * np.asarray(f_hz, dtype=np.float64)
    return w / np.real(model_wavenumber(f_hz, params, c))


def _unpack(theta):
    kap = np.exp(theta[:2])
    rates = np.exp(theta[2:])
    p = np.empty(10)
    p[0], p[1] = kap
    # rates: d_x1_nu1, alpha_x1_nu1, d_x2_nu1, alpha_x2_nu1, d_x1_nu2, ...
    p[2:] = rates
    return p


def stretch_inverse_discrete(w, kappa, d, alpha, dt) -> NDArray[np.complex128]:
    """The engine's S(w): the recursive convolution psi_n = b psi_(n-1) + a D_n with
    a, b of (d, kappa, alpha, dt), i.e. 1/kappa + sum_nu a_nu / (1 - b_nu exp(i w dt))."""
    w = np.asarray(w, dtype=np.float64)
    out = np.full(w.shape, 1.0 / kappa, dtype=np.complex128)
    for dn, an in zip(d, alpha):
        b = np.exp(-(dn / kappa + an) * dt)
        a = dn / (kappa * (dn + kappa * an) + 1e-10) * (b - 1)
        out += a / (1 - b * np.exp(1j * w * dt))
    return out


def fit_relaxation_parameters(alpha0: float, gamma: float, *, band=DEFAULT_FIT_RANGE_HZ, n_freq: int = N_FIT_FREQ,
                              dt: float | None = None,
                              warm_start: NDArray[np.float64] | None = None, warm_alpha0: float | None = None):
    """Fit one (alpha0, gamma) pair; returns ``(params[10], max relative error over band)``.

    With ``dt`` the model is the engine's discrete-time recursive convolution at that time
    step (what the simulation actually attenuates by); without it, the continuous-time model.

    ``warm_start`` (a fitted row, for ``warm_alpha0``) seeds a single local solve with its
    ``d`` rates scaled by ``alpha0 / warm_alpha0`` (loss is nearly linear in ``d``); the
    multi-start search runs only if that solve misses by more than 2 %.
    """
    if alpha0 <= 0:
        p = np.zeros(10)
        p[0] = p[1] = 1.0
        p[3] = p[5] = p[7] = p[9] = 2 * np.pi * 1e6  # alpha > 0, d = 0: a = 0, b < 1
        return p, 0.0
    f = np.geomspace(band[0], band[1], n_freq)
    tgt = target_attenuation(f, alpha0, gamma)
    w = 2 * np.pi * f

    def sides(p):
        if dt is None:
            return (stretch_inverse(w, p[0], p[2::4], p[3::4]), stretch_inverse(w, p[1], p[4::4], p[5::4]))
        return (stretch_inverse_discrete(w, p[0], p[2::4], p[3::4], dt),
                stretch_inverse_discrete(w, p[1], p[4::4], p[5::4], dt))

    def resid(theta):
        p = _unpack(theta)
        s1, s2 = sides(p)
        k = (w / C_REF) / np.sqrt(s1 * s2)
        att = np.imag(k)
        r_att = np.log(np.maximum(att, 1e-30) / tgt)
        vel = np.log(np.real(k) * C_REF / w)  # log(c / c_phase)
        # keep the two sides balanced: the wave impedance carries sqrt(S2 / S1), and an
        # imbalance of 1e-3 costs as much as a 1 % attenuation error
        imb = 10.0 * np.log(s1 / s2)
        return np.concatenate([r_att, [10.0 * vel.mean()], imb.real, imb.imag])

    # kappa >= 1 on both sides: the high-frequency wave speed c / sqrt(kappa_1 kappa_2) must not
    # exceed c, or the time step would leave the stable range
    lo = np.array([0.0, 0.0] + [np.log(1e2)] * 8)
    hi = np.array([0.5, 0.5] + [np.log(1e10)] * 8)

    def max_err(p):
        s1, s2 = sides(p)
        att = np.imag((w / C_REF) / np.sqrt(s1 * s2))
        return float(np.max(np.abs(att / tgt - 1)))

    if warm_start is not None and warm_alpha0 and warm_alpha0 > 0:
        ws = np.array(warm_start, dtype=np.float64)
        ws[2::2] *= alpha0 / warm_alpha0  # d entries sit at 2, 4, 6, 8
        theta0 = np.clip(np.concatenate([np.log(ws[:2]), np.log(np.maximum(ws[2:], 1e-300))]), lo + 1e-6, hi - 1e-6)
        sol = least_squares(resid, theta0, method="trf", x_scale="jac", bounds=(lo, hi), max_nfev=400)
        p = _unpack(sol.x)
        err = max_err(p)
        if err < 0.02:
            return p, err

    best = None
    # relaxation rates spread over the band; d sized so the loss per mechanism is right
    centres = [(0.5, 4.0, 1.5, 12.0), (1.0, 8.0, 2.5, 25.0), (0.3, 2.0, 5.0, 40.0)]
    for cs in centres:
        a_nominal = target_attenuation(2e6 * np.array(cs), alpha0, gamma)  # Np/m
        theta0 = [0.0, 0.0]
        for i, cmhz in enumerate(cs):
            beta = 2 * np.pi * cmhz * 1e6
            # loss of one Debye term near its peak ~ (w/c) * (d / beta) / 2 ... size d from it
            d = max(2.0 * a_nominal[i] * C_REF / (2 * np.pi * cmhz * 1e6) * beta, 1.0)
            theta0 += [np.log(d), np.log(beta)]
        # reorder to (d_x1_nu1, a_x1_nu1, d_x2_nu1, a_x2_nu1, d_x1_nu2, a_x1_nu2, ...)
        r = theta0[2:]
        theta0 = np.array(theta0[:2] + [r[0], r[1], r[2], r[3], r[4], r[5], r[6], r[7]])
        theta0 = np.clip(theta0, lo + 1e-6, hi - 1e-6)
        sol = least_squares(resid, theta0, method="trf", x_scale="jac", bounds=(lo, hi), max_nfev=4000)
        if best is None or sol.cost < best.cost:
            best = sol
    p = _unpack(best.x)
    return p, max_err(p)


class RelaxationTable:
    """The (alpha0, gamma) -> 10-parameter table, on the reference's binning.

    Rows are fitted on first use and kept; ``fill()`` fits every row (to save a complete table
    with :meth:`save`). A table saved by :meth:`save` (``.npz``) or a MATLAB ``.mat`` with the
    reference's variable names loads with :meth:`load`.
    """

    def __init__(self, alpha_list=ALPHA_0_LIST, power_list=POWER_LIST, *, dt: float | None = None,
                 band=DEFAULT_FIT_RANGE_HZ) -> None:
        self.dt = dt
        self.band = tuple(band)
        self.alpha_list = np.asarray(alpha_list, dtype=np.float64).reshape(-1)
        self.power_list = np.asarray(power_list, dtype=np.float64).reshape(-1)
        na, npw = self.alpha_list.size, self.power_list.size
        self.database = np.full((na, npw, 10), np.nan)
        self.invalid_matrix = np.zeros((na, npw), dtype=bool)
        self.fit_error = np.full((na, npw), np.nan)
        self._lock = threading.Lock()

    @property
    def alpha_min(self):
        return self.alpha_list.min()

    @property
    def alpha_max(self):
        return self.alpha_list.max()

    @property
    def power_min(self):
        return self.power_list.min()

    @property
    def power_max(self):
        return round(float(self.power_list.max()), 4)

    def _ensure(self, ia: NDArray, ip: NDArray) -> None:
        with self._lock:
            need = np.isnan(self.database[ia, ip, 0])
            todo = sorted(set(zip(ip[need].tolist(), ia[need].tolist())))  # by gamma, then alpha0
            for p, a in todo:
                # warm start from the nearest fitted row of the same gamma, else any gamma
                done_a = np.nonzero(~np.isnan(self.database[:, p, 0]) & (self.alpha_list > 0))[0]
                warm = warm_a0 = None
                if done_a.size:
                    j = done_a[np.argmin(np.abs(done_a - a))]
                    warm, warm_a0 = self.database[j, p], self.alpha_list[j]
                else:
                    filled = np.argwhere(~np.isnan(self.database[:, :, 0]) & (self.alpha_list[:, None] > 0))
                    if filled.size:
                        k = filled[np.argmin(np.abs(filled[:, 1] - p) * 1000 + np.abs(filled[:, 0] - a))]
                        warm, warm_a0 = self.database[k[0], k[1]], self.alpha_list[k[0]]
                params, err = fit_relaxation_parameters(self.alpha_list[a], self.power_list[p],
                                                        band=self.band, dt=self.dt,
                                                        warm_start=warm, warm_alpha0=warm_a0)
                self.database[a, p] = params
                self.fit_error[a, p] = err

    def fill(self) -> None:
        ia, ip = np.meshgrid(np.arange(self.alpha_list.size), np.arange(self.power_list.size), indexing="ij")
        self._ensure(ia.reshape(-1), ip.reshape(-1))

    def indices(self, alpha_coeff, alpha_power):
        """``searchsorted`` onto the table grid, clipped (relaxation_parameters.py:47-54)."""
        ia = np.searchsorted(self.alpha_list.round(10), alpha_coeff)
        ip = np.searchsorted(self.power_list.round(10), alpha_power)
        ia = np.clip(ia, 0, self.alpha_list.size - 1)
        ip = np.clip(ip, 0, self.power_list.size - 1)
        return ia, ip

    def lookup(self, alpha_coeff, alpha_power) -> NDArray[np.float64]:
        a = np.asarray(alpha_coeff, dtype=np.float64)
        g = np.asarray(alpha_power, dtype=np.float64)
        ia, ip = self.indices(a, g)
        uniq = np.unique(np.stack([ia.reshape(-1), ip.reshape(-1)], 1), axis=0)
        self._ensure(uniq[:, 0], uniq[:, 1])
        bad = self.invalid_matrix[ia, ip]
        if np.any(bad):
            logger.warning("Some attenuation values correspond to invalid relaxation parameters "
                           "(%d points).", int(bad.sum()))
        return self.database[ia, ip, :]

    def save(self, path: Path) -> None:
        path = Path(path)
        data = {
            "dt": np.array(np.nan if self.dt is None else self.dt),
            "band_hz": np.array(self.band),
            "database": self.database, "alpha_0_list": self.alpha_list[None, :],
            "power_list": self.power_list[None, :], "invalid_matrix": self.invalid_matrix,
        }
        if path.suffix == ".mat":
            from scipy.io import savemat

            savemat(path, data)
        else:
            np.savez(path, **data)

    @classmethod
    def load(cls, path: Path) -> RelaxationTable:
        path = Path(path)
        if path.suffix == ".mat":
            from scipy.io import loadmat

            data = loadmat(path)
        else:
            data = dict(np.load(path, allow_pickle=False))
        dt = float(np.asarray(data["dt"]).reshape(-1)[0]) if "dt" in data else float("nan")
        band = tuple(np.asarray(data["band_hz"]).reshape(-1)) if "band_hz" in data else DEFAULT_FIT_RANGE_HZ
        t = cls(data["alpha_0_list"], data["power_list"], dt=None if np.isnan(dt) else dt, band=band)
        t.database = np.asarray(data["database"], dtype=np.float64)
        t.invalid_matrix = np.asarray(data["invalid_matrix"]).astype(bool)
        if t.database.ndim != 3 or t.database.shape[2] != 10:
            raise ValueError("look_up_table must have 3 dimensions and 10 columns.")
        if np.isnan(t.database).any():
            raise ValueError("look_up_table must not contain NaN values.")
        return t


_tables: dict = {}


def default_table(dt: float | None = None, band=DEFAULT_FIT_RANGE_HZ) -> RelaxationTable:
    """The process-wide table for one time step and band (filled lazily)."""
    key = (None if dt is None else float(np.float64(dt)), tuple(float(b) for b in band))
    if key not in _tables:
        _tables[key] = RelaxationTable(dt=dt, band=band)
    return _tables[key]


def fit_band_for_grid(dx: float, c0: float = C_REF) -> tuple[float, float]:
    """Fit range for a grid: the band 1 MHz up to 20 MHz or to 4 points per wavelength,
    whichever is lower (its low end at most a tenth of its high end), widened by ``FIT_MARGIN``
    at both ends."""
    f_hi = min(FIT_BAND_HZ[1], c0 / (4.0 * dx))
    f_lo = min(FIT_BAND_HZ[0], f_hi / 10.0)
    return (f_lo / FIT_MARGIN, f_hi * FIT_MARGIN)


def generate_relaxation_params(
    alpha_coeff,
    alpha_power,
    n_relaxation_mechanisms: int = 2,
    path_database: Path | None = None,
    *,
    dt: float | None = None,
    band=DEFAULT_FIT_RANGE_HZ,
) -> dict:
    """Per-voxel relaxation dictionary for ``alpha_coeff`` / ``alpha_power`` maps
    (same signature and clipping as fullwave/utils/relaxation_parameters.py:78-112).

    ``dt`` fits the engine's discrete-time model for that time step (``Medium.build`` passes
    its grid's); ``path_database`` uses a saved table instead of fitting. A saved table fitted
    for another time step is refused (its discrete-time fit would give another attenuation on
    this grid); one saved without a time step (the reference's format, a continuous-time fit)
    is used with a warning."""
    if n_relaxation_mechanisms != 2:
        raise ValueError("Only n_relaxation_mechanisms=2 are supported currently.")
    table = RelaxationTable.load(path_database) if path_database is not None else default_table(dt, band)
    if path_database is not None and dt is not None:
        if table.dt is None:
            logger.warning("relaxation table %s has no time step: it is a continuous-time fit, not the "
                           "discrete-time fit for dt = %.6g s", path_database, dt)
        elif abs(table.dt - dt) > 1e-9 * abs(dt):
            raise ValueError(f"relaxation table {path_database} was fitted for dt = {table.dt:.9g} s, this grid's "
                             f"time step is {dt:.9g} s; fit a table for it (RelaxationTable(dt=...).fill())")
    a = np.asarray(alpha_coeff, dtype=np.float64)
    g = np.asarray(alpha_power, dtype=np.float64)
    if np.any(a < table.alpha_min) or np.any(g < table.power_min):
        logger.warning("attenuation is out of range. the out-of-range values will be clipped to the min value.")
    if np.any(a > table.alpha_max) or np.any(g > table.power_max):
        logger.warning("attenuation is out of range. the out-of-range values will be clipped to the max value.")
    a = np.clip(a, table.alpha_min, table.alpha_max)
    g = np.clip(g, table.power_min, table.power_max)
    rows = table.lookup(a, g)
    out = initialize_relaxation_param_dict(n_relaxation_mechanisms)
    for i, key in enumerate(relaxation_param_keys(n_relaxation_mechanisms)):
        out[key] = rows[..., i]
    return out
