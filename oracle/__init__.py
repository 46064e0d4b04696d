"""TEST INFRASTRUCTURE ONLY — the oracle.

A CPU restatement of the Fullwave 2.5 time-stepping engine (the update is written out at the
top of oracle/fw_oracle.c), used by tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of bench.py to check and time the HIP engine. The product never imports
this package.

* :class:`OracleSimulation` — the C oracle (``libfw_oracle.so``, FP64 or FP32, OpenMP) on the
  same :class:`~fullwave25_amd.solver.inputs.SimulationInputs` the engine takes.
* :func:`brute_force_run` — an independent pure-numpy statement of the same update for tiny
  grids (whole-array slicing instead of loops), to check the C oracle's indexing.

Parity status of each function is recorded in the DESIGN document of the repository; the
update itself has no reference implementation (the reference ships no engine) and is pinned by
the mathematical checks in tests/test_oracle_math.py.
"""

from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libfw_oracle.so"

_P = ctypes.POINTER(ctypes.c_double)


class _Problem(ctypes.Structure):
    _fields_ = [
        ("ndim", ctypes.c_int32), ("n_relax", ctypes.c_int32),
        ("nx", ctypes.c_int64), ("ny", ctypes.c_int64), ("nz", ctypes.c_int64),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double), ("dt", ctypes.c_double),
        ("K", _P), ("rho", _P),
        ("kappa_v", _P * 3), ("kappa_p", _P * 3),
        ("a_v", (_P * 2) * 3), ("b_v", (_P * 2) * 3), ("a_p", (_P * 2) * 3), ("b_p", (_P * 2) * 3),
        ("dcmap", ctypes.POINTER(ctypes.c_int32)), ("coef", _P), ("ndmap", ctypes.c_int64),
        ("ncoords", ctypes.c_int64), ("icc", ctypes.POINTER(ctypes.c_int64)), ("icmat", _P), ("ntic", ctypes.c_int64),
        ("nzero", ctypes.c_int64), ("zero", ctypes.POINTER(ctypes.c_int64)),
        ("nout", ctypes.c_int64), ("outc", ctypes.POINTER(ctypes.c_int64)), ("modt", ctypes.c_int64),
        ("aexp", _P),
    ]


def build(force: bool = False) -> Path:
    """Compile ``libfw_oracle.so`` with the Makefile next to this file."""
    if force or not LIB_PATH.exists():
        subprocess.run(["make", "-C", str(HERE), "-B" if force else "all"], check=True,
                       stdout=subprocess.DEVNULL)
    return LIB_PATH


_lib = None


def load() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        lib = ctypes.CDLL(str(LIB_PATH))
        for sfx in ("f64", "f32"):
            getattr(lib, f"fwo_create_{sfx}").restype = ctypes.c_void_p
            getattr(lib, f"fwo_create_{sfx}").argtypes = [ctypes.POINTER(_Problem)]
            getattr(lib, f"fwo_destroy_{sfx}").argtypes = [ctypes.c_void_p]
            getattr(lib, f"fwo_run_{sfx}").argtypes = [ctypes.c_void_p, ctypes.c_int64, _P]
            getattr(lib, f"fwo_get_field_{sfx}").argtypes = [ctypes.c_void_p, ctypes.c_int, _P]
            getattr(lib, f"fwo_set_field_{sfx}").argtypes = [ctypes.c_void_p, ctypes.c_int, _P]
            getattr(lib, f"fwo_adjoint_{sfx}").argtypes = [ctypes.c_void_p, ctypes.c_int64, _P, _P, _P]
        _lib = lib
    return _lib


def coefficient_sets(dmap: np.ndarray) -> np.ndarray:
    """``[ndmap, 9]`` rows ``(w_1..w_8, e)`` from the contract's ``dmap [9, 2, ndmap]``."""
    out = np.empty((dmap.shape[2], 9))
    out[:, :8] = np.asarray(dmap, dtype=np.float64)[1:9, 0, :].T
    out[:, 8] = dmap[1, 1, :]
    return out


def linear_indices(coords: np.ndarray, shape: tuple) -> np.ndarray:
    c = np.asarray(coords, dtype=np.int64).reshape(-1, len(shape))
    if c.shape[0] == 0:
        return np.zeros(0, dtype=np.int64)
    return np.ravel_multi_index(tuple(c[:, a] for a in range(len(shape))), shape).astype(np.int64)


_V_AXES = ("u", "w", "v")
_P_AXES = ("x", "y", "z")


class OracleSimulation:
    """The C oracle on a SimulationInputs, FP64 (``precision=64``) or FP32 (``32``)."""

    def __init__(self, inp, precision: int = 64) -> None:
        self.lib = load()
        self.inp = inp
        self.sfx = {64: "f64", 32: "f32"}[precision]
        keep = []

        def dbl(a):
            arr = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
            keep.append(arr)
            return arr.ctypes.data_as(_P)

        def i64(a):
            arr = np.ascontiguousarray(np.asarray(a, dtype=np.int64))
            keep.append(arr)
            return arr.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))

        P = _Problem()
        P.ndim, P.n_relax = inp.ndim, 2
        P.nx, P.ny, P.nz = inp.nx, inp.ny, inp.nz if inp.ndim == 3 else 1
        P.dx, P.dy, P.dz, P.dt = inp.dx, inp.dy, inp.dz, inp.dt
        P.K, P.rho = dbl(inp.maps["K"]), dbl(inp.maps["rho"])
        if getattr(inp, "exponential", False):
            P.aexp = dbl(inp.maps["a_exp"])
        for a in range(inp.ndim if not getattr(inp, "exponential", False) else 0):
            va = "u" if inp.isotropic else _V_AXES[a]
            pa = "x" if inp.isotropic else _P_AXES[a]
            P.kappa_v[a] = dbl(inp.maps[f"kappa{va}"])
            P.kappa_p[a] = dbl(inp.maps[f"kappa{pa}"])
            for nu in range(2):
                P.a_v[a][nu] = dbl(inp.maps[f"apml{va}{nu + 1}"])
                P.b_v[a][nu] = dbl(inp.maps[f"bpml{va}{nu + 1}"])
                P.a_p[a][nu] = dbl(inp.maps[f"apml{pa}{nu + 1}"])
                P.b_p[a][nu] = dbl(inp.maps[f"bpml{pa}{nu + 1}"])
        dc = np.ascontiguousarray(inp.dcmap, dtype=np.int32)
        keep.append(dc)
        P.dcmap = dc.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        P.coef = dbl(coefficient_sets(inp.dmap))
        P.ndmap = inp.ndmap
        P.ncoords = inp.ncoords
        P.icc = i64(linear_indices(inp.icc, inp.shape))
        P.icmat = dbl(inp.icmat)
        P.ntic = inp.ntic
        zero = linear_indices(inp.icczero, inp.shape) if inp.ncoordszero else np.zeros(0, np.int64)
        P.nzero = zero.shape[0]
        P.zero = i64(zero)
        P.nout = inp.ncoordsout
        P.outc = i64(linear_indices(inp.outc, inp.shape))
        P.modt = inp.modt
        self._keep = keep
        self._P = P
        self._h = getattr(self.lib, f"fwo_create_{self.sfx}")(ctypes.byref(P))
        if not self._h:
            raise MemoryError("oracle allocation failed")
        self._rec = np.zeros((inp.n_records, inp.ncoordsout), dtype=np.float64)
        self.step = 0

    def run(self, n_steps: int) -> None:
        if self.step + n_steps > self.inp.nt:
            raise ValueError("run past nt")
        getattr(self.lib, f"fwo_run_{self.sfx}")(self._h, int(n_steps), self._rec.ctypes.data_as(_P))
        self.step += n_steps

    def sensors(self) -> np.ndarray:
        n = -(-self.step // self.inp.modt)
        return self._rec[:n].copy()

    def field(self, name: str) -> np.ndarray:
        which = {"p": 0, "vx": 1, "vy": 2, "vz": 3}[name]
        out = np.empty(self.inp.shape, dtype=np.float64)
        getattr(self.lib, f"fwo_get_field_{self.sfx}")(self._h, which, out.ctypes.data_as(_P))
        return out

    def set_field(self, name: str, values: np.ndarray) -> None:
        which = {"p": 0, "vx": 1, "vy": 2, "vz": 3}[name]
        arr = np.ascontiguousarray(values, dtype=np.float64)
        getattr(self.lib, f"fwo_set_field_{self.sfx}")(self._h, which, arr.ctypes.data_as(_P))

    def adjoint(self, grad_rec: np.ndarray, n_steps: int | None = None) -> tuple[np.ndarray, np.ndarray]:
        """Reverse-mode product of ``grad_rec`` ([n_records, n_sensors]) with the Jacobian of
        the traces of a full run from rest: returns ``(dK, drho)`` on the grid, the coefficient
        sets held fixed. Runs the forward pass itself (the state is left at its end)."""
        n = self.inp.nt if n_steps is None else int(n_steps)
        g = np.ascontiguousarray(grad_rec, dtype=np.float64)
        if g.shape != (-(-n // self.inp.modt), self.inp.ncoordsout):
            raise ValueError(f"grad_rec must be [{-(-n // self.inp.modt)}, {self.inp.ncoordsout}]")
        gk = np.zeros(self.inp.shape)
        gr = np.zeros(self.inp.shape)
        rc = getattr(self.lib, f"fwo_adjoint_{self.sfx}")(self._h, n, g.ctypes.data_as(_P), gk.ctypes.data_as(_P),
                                                          gr.ctypes.data_as(_P))
        if rc:
            raise MemoryError("oracle adjoint: allocation failed")
        self.step = n
        return gk, gr

    def close(self) -> None:
        if self._h:
            getattr(self.lib, f"fwo_destroy_{self.sfx}")(self._h)
            self._h = None

    def __del__(self) -> None:
        try:
            self.close()
        except Exception:
            pass


def brute_force_run(inp, n_steps: int) -> tuple[np.ndarray, np.ndarray]:
    """Independent numpy statement of the update for tiny grids (FP64).

    Whole-array shifted slices instead of the C oracle's per-cell loops; returns
    ``(records [n_rec, n_out], p_final)``.
    """
    nd = inp.ndim
    shape = inp.shape
    H = 8
    inner = tuple(slice(H, s - H) for s in shape)
    coef = coefficient_sets(inp.dmap)
    w = coef[np.asarray(inp.dcmap)[inner]]  # [*inner, 9]
    h = (inp.dx, inp.dy, inp.dz)

    def sh(f, off):
        """f shifted: value at (inner cell + off)."""
        return f[tuple(slice(H + off[a], shape[a] - H + off[a]) for a in range(nd))]

    def unit(a, s=1):
        o = [0] * nd
        o[a] = s
        return o

    def add(*offs):
        return [sum(o[a] for o in offs) for a in range(nd)]

    m = {k: np.asarray(v, dtype=np.float64) for k, v in inp.maps.items()}
    expo = getattr(inp, "exponential", False)
    p = np.zeros(shape)
    v = [np.zeros(shape) for _ in range(nd)]
    psi = [[np.zeros(shape) for _ in range(2)] for _ in range(nd)]
    phi = [[np.zeros(shape) for _ in range(2)] for _ in range(nd)]
    rec = np.zeros((inp.n_records, inp.ncoordsout))
    icc = linear_indices(inp.icc, shape)
    zero = linear_indices(inp.icczero, shape) if inp.ncoordszero else np.zeros(0, np.int64)
    outc = linear_indices(inp.outc, shape)
    for n in range(n_steps):
        for a in range(nd):
            va = "u" if inp.isotropic else _V_AXES[a]
            ea = unit(a)
            acc = sum(w[..., k - 1] * (sh(p, unit(a, k)) - sh(p, unit(a, -(k - 1)))) for k in range(1, 9))
            tr = 0.0
            for b in range(nd):
                if b == a:
                    continue
                for s in (1, -1):
                    eb = unit(b, s)
                    tr = tr + (sh(p, add(ea, eb)) - sh(p, eb))
            D = (acc + w[..., 8] * tr) / h[a]
            rho_bar = 0.5 * (m["rho"][inner] + sh(m["rho"], ea))
            if expo:
                v[a][inner] = m["a_exp"][inner] * (v[a][inner] - inp.dt / rho_bar * D)
                continue
            mem = 0.0
            for nu in range(2):
                q = m[f"bpml{va}{nu + 1}"][inner] * psi[a][nu][inner] + m[f"apml{va}{nu + 1}"][inner] * D
                psi[a][nu][inner] = q
                mem = mem + q
            v[a][inner] -= inp.dt / rho_bar * (D / m[f"kappa{va}"][inner] + mem)
        div = 0.0
        for a in range(nd):
            pa = "x" if inp.isotropic else _P_AXES[a]
            ea = unit(a)
            acc = sum(w[..., k - 1] * (sh(v[a], unit(a, k - 1)) - sh(v[a], unit(a, -k))) for k in range(1, 9))
            tr = 0.0
            for b in range(nd):
                if b == a:
                    continue
                for s in (1, -1):
                    eb = unit(b, s)
                    tr = tr + (sh(v[a], eb) - sh(v[a], add(unit(a, -1), eb)))
            D = (acc + w[..., 8] * tr) / h[a]
            if expo:
                div = div + D
                continue
            mem = 0.0
            for nu in range(2):
                q = m[f"bpml{pa}{nu + 1}"][inner] * phi[a][nu][inner] + m[f"apml{pa}{nu + 1}"][inner] * D
                phi[a][nu][inner] = q
                mem = mem + q
            div = div + D / m[f"kappa{pa}"][inner] + mem
        if expo:
            p[inner] = m["a_exp"][inner] * (p[inner] - inp.dt * m["K"][inner] * div)
        else:
            p[inner] -= inp.dt * m["K"][inner] * div
        pf = p.reshape(-1)
        if n < inp.ntic:
            pf[icc] = np.asarray(inp.icmat, dtype=np.float64)[:, n]
        pf[zero] = 0.0
        if n % inp.modt == 0:
            rec[n // inp.modt] = pf[outc]
    return rec[: -(-n_steps // inp.modt)], p
