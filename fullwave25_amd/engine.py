"""ctypes binding of ``libfullwave_hip.so`` (declared in include/fullwave_hip.h).

This module loads the in-tree HIP library and fails loudly when it is missing: there is no
CPU path. :class:`HipSimulation` owns one engine handle; :func:`run_dir` is the directory
contract used by :class:`~fullwave25_amd.solver.launcher.HipLauncher`.
"""

from __future__ import annotations

import ctypes
import threading
from pathlib import Path

import numpy as np

from .solver.inputs import SimulationInputs

LIB_NAME = "libfullwave_hip.so"
LIB_PATH = Path(__file__).resolve().parent / LIB_NAME

FW_OK = 0
ABI_VERSION = 4
STATUS_NAMES = {
    0: "FW_OK", 1: "FW_ERR_INVALID", 2: "FW_ERR_DEVICE", 3: "FW_ERR_NOMEM",
    4: "FW_ERR_IO", 5: "FW_ERR_STATE", 6: "FW_ERR_UNSUPPORTED",
}
KERNEL_KINDS = ("velocity", "pressure", "boundary")

# every symbol include/fullwave_hip.h declares
EXPORTED_SYMBOLS = (
    "fw_abi_version", "fw_version_string", "fw_device_count", "fw_last_error", "fw_run_dir", "fw_run_dir_noalloc",
    "fw_device_copy_bandwidth", "fw_device_stream_bandwidth", "fw_device_memory",
    "fw_sim_create", "fw_sim_upload", "fw_sim_run", "fw_sim_reset", "fw_sim_step_count",
    "fw_sim_record_count", "fw_sim_download_sensors", "fw_sim_download_field",
    "fw_sim_set_field", "fw_sim_kernel_config", "fw_sim_set_kernel_config", "fw_sim_set_graph_steps",
    "fw_sim_set_relaxation_codes", "fw_sim_relaxation_codes", "fw_sim_run_timed",
    "fw_sim_destroy", "fw_group_create", "fw_group_config", "fw_group_create_shaped", "fw_group_shape",
    "fw_group_run", "fw_group_set_graph_steps", "fw_group_destroy",
    "fw_adj_create", "fw_adj_checkpoint_every", "fw_adj_forward", "fw_adj_backward", "fw_adj_destroy",
)
VARIANTS = {"percell": 0, "stream": 1, "tile": 2, "pipe": 3, "pipelite": 4}


class EngineError(RuntimeError):
    """A non-zero status from the engine, with its message."""

    def __init__(self, status: int, message: str) -> None:
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class FwDesc(ctypes.Structure):
    _fields_ = [
        ("abi_version", ctypes.c_int32), ("ndim", ctypes.c_int32), ("isotropic", ctypes.c_int32),
        ("n_relax", ctypes.c_int32), ("nx", ctypes.c_int64), ("ny", ctypes.c_int64),
        ("nz", ctypes.c_int64), ("nt", ctypes.c_int64), ("dx", ctypes.c_double),
        ("dy", ctypes.c_double), ("dz", ctypes.c_double), ("dt", ctypes.c_double),
        ("c0", ctypes.c_double), ("ncoords", ctypes.c_int64), ("ntic", ctypes.c_int64),
        ("ncoordsout", ctypes.c_int64), ("modt", ctypes.c_int64), ("ncoordszero", ctypes.c_int64),
        ("ndmap", ctypes.c_int64), ("attenuation_model", ctypes.c_int32), ("n_shots", ctypes.c_int32),
    ]


_lib = None
_lib_lock = threading.Lock()


def load_library(path: Path | None = None) -> ctypes.CDLL:
    """Load (once) and return the engine library; raise if it has not been built."""
    global _lib
    with _lib_lock:
        if _lib is not None and path is None:
            return _lib
        p = Path(path) if path is not None else LIB_PATH
        if not p.exists():
            raise RuntimeError(
                f"{p} is missing: build the HIP engine first (python -c 'import __graft_entry__ as g; g.build()')",
            )
        lib = ctypes.CDLL(str(p))
        vp, i32, i64, sz = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_size_t
        lib.fw_abi_version.restype = ctypes.c_int
        lib.fw_version_string.restype = ctypes.c_char_p
        lib.fw_last_error.restype = ctypes.c_char_p
        lib.fw_device_count.argtypes = [ctypes.POINTER(ctypes.c_int)]
        lib.fw_run_dir.argtypes = [ctypes.c_char_p, ctypes.c_int]
        lib.fw_run_dir_noalloc.argtypes = [ctypes.c_char_p, ctypes.c_int]
        lib.fw_device_copy_bandwidth.argtypes = [ctypes.c_int, sz, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        lib.fw_device_stream_bandwidth.argtypes = [ctypes.c_int, i64, i32, i32, i32, ctypes.POINTER(ctypes.c_double)]
        lib.fw_device_memory.argtypes = [ctypes.c_int, ctypes.POINTER(sz), ctypes.POINTER(sz)]
        lib.fw_sim_create.argtypes = [ctypes.POINTER(FwDesc), ctypes.c_int, ctypes.POINTER(vp)]
        lib.fw_sim_upload.argtypes = [vp, ctypes.c_char_p, vp, sz]
        lib.fw_sim_run.argtypes = [vp, i64]
        lib.fw_sim_reset.argtypes = [vp]
        lib.fw_sim_step_count.argtypes = [vp, ctypes.POINTER(i64)]
        lib.fw_sim_record_count.argtypes = [vp, ctypes.POINTER(i64)]
        lib.fw_sim_download_sensors.argtypes = [vp, vp, sz]
        lib.fw_sim_download_field.argtypes = [vp, ctypes.c_char_p, vp, sz]
        lib.fw_sim_set_field.argtypes = [vp, ctypes.c_char_p, vp, sz]
        lib.fw_sim_kernel_config.argtypes = [vp, i32, ctypes.POINTER(i32), ctypes.POINTER(i32), ctypes.POINTER(i32)]
        lib.fw_sim_set_kernel_config.argtypes = [vp, i32, i32, i32, i32]
        lib.fw_sim_set_graph_steps.argtypes = [vp, i64]
        lib.fw_sim_set_relaxation_codes.argtypes = [vp, i32]
        lib.fw_sim_relaxation_codes.argtypes = [vp, ctypes.POINTER(i32), ctypes.POINTER(ctypes.c_double)]
        lib.fw_sim_run_timed.argtypes = [vp, i64, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(i32)]
        lib.fw_sim_destroy.argtypes = [vp]
        lib.fw_group_create.argtypes = [ctypes.POINTER(vp), i32, i32, i32, ctypes.POINTER(vp)]
        lib.fw_group_run.argtypes = [vp, i64]
        lib.fw_group_config.argtypes = [vp, ctypes.POINTER(i32), ctypes.POINTER(i32)]
        lib.fw_group_create_shaped.argtypes = [ctypes.POINTER(vp), i32, i32, i32, i32, ctypes.POINTER(vp)]
        lib.fw_group_shape.argtypes = [vp, ctypes.POINTER(i32), ctypes.POINTER(i32), ctypes.POINTER(i32)]
        lib.fw_group_destroy.argtypes = [vp]
        lib.fw_group_set_graph_steps.argtypes = [vp, i64]
        lib.fw_adj_create.argtypes = [vp, i64, ctypes.POINTER(vp)]
        lib.fw_adj_checkpoint_every.argtypes = [vp, ctypes.POINTER(i64)]
        lib.fw_adj_forward.argtypes = [vp]
        lib.fw_adj_backward.argtypes = [vp, vp, sz, vp, vp, sz]
        lib.fw_adj_destroy.argtypes = [vp]
        if lib.fw_abi_version() != ABI_VERSION:
            raise RuntimeError(f"{p}: ABI version {lib.fw_abi_version()}, expected {ABI_VERSION}")
        if path is None:
            _lib = lib
        return lib


def _check(lib, status: int) -> None:
    if status != FW_OK:
        raise EngineError(status, lib.fw_last_error().decode(errors="replace"))


def device_count() -> int:
    lib = load_library()
    n = ctypes.c_int(0)
    _check(lib, lib.fw_device_count(ctypes.byref(n)))
    return n.value


def copy_bandwidth(nbytes: int = 1 << 30, repeats: int = 10, device: int = 0) -> float:
    """GB/s (read + write) of a plain device-to-device copy kernel."""
    lib = load_library()
    g = ctypes.c_double(0)
    _check(lib, lib.fw_device_copy_bandwidth(int(device), int(nbytes), int(repeats), ctypes.byref(g)))
    return g.value


def stream_bandwidth(cells: int, nread: int, nwrite: int, repeats: int = 5, device: int = 0) -> float:
    """GB/s (read + write) of a kernel that reads ``nread`` and writes ``nwrite`` arrays of
    ``cells`` floats, one float per lane (the stencil kernels' per-cell operand pattern)."""
    lib = load_library()
    g = ctypes.c_double(0)
    _check(lib, lib.fw_device_stream_bandwidth(int(device), int(cells), int(nread), int(nwrite), int(repeats),
                                               ctypes.byref(g)))
    return g.value


def device_memory(device: int = 0) -> tuple[int, int]:
    """``(free, total)`` device memory in bytes."""
    lib = load_library()
    f, t = ctypes.c_size_t(0), ctypes.c_size_t(0)
    _check(lib, lib.fw_device_memory(int(device), ctypes.byref(f), ctypes.byref(t)))
    return f.value, t.value


def device_bytes(inp: SimulationInputs, n_shots: int = 1) -> int:
    """Device memory one handle for ``inp`` allocates (maps once, wave state per shot)."""
    nd = inp.ndim
    # 3D rows are padded to a multiple of 32 floats on the device
    cells = inp.nx * inp.ny * (-(-inp.nz // 32) * 32) if nd == 3 else inp.ncells
    n_maps = 3 + len(inp.attenuation_map_stems())  # K, rho, dcmap + attenuation maps
    state = (1 + nd + (0 if inp.exponential else 4 * nd)) * cells
    per_shot = state + inp.ncoords * inp.ntic + inp.n_records * inp.ncoordsout
    fixed = n_maps * cells + 2 * (inp.ncoords + inp.ncoordsout + inp.ncoordszero) + inp.ndmap * nd * 9
    return 4 * (fixed + n_shots * per_shot)


def run_dir(sim_dir: Path, device: int = 0) -> None:
    """Run the engine on a simulation directory; writes ``genout.dat`` there."""
    lib = load_library()
    _check(lib, lib.fw_run_dir(str(Path(sim_dir).absolute()).encode(), int(device)))


# device bytes fw_run_dir_noalloc works in (FW_NOALLOC_SCRATCH_BYTES)
NOALLOC_SCRATCH_BYTES = 256 << 20


def run_dir_noalloc(sim_dir: Path, device: int = 0) -> None:
    """:func:`run_dir` without device allocation: every device array of the run sits in the
    library's static per-GPU array of :data:`NOALLOC_SCRATCH_BYTES` (runs that need more fail
    with FW_ERR_UNSUPPORTED); kernel shapes are the defaults or FULLWAVE_HIP_KERNELS's."""
    lib = load_library()
    _check(lib, lib.fw_run_dir_noalloc(str(Path(sim_dir).absolute()).encode(), int(device)))


def make_desc(inp: SimulationInputs, n_shots: int = 1) -> FwDesc:
    return FwDesc(
        abi_version=ABI_VERSION, ndim=inp.ndim, isotropic=1 if inp.isotropic else 0, n_relax=inp.n_relax,
        nx=inp.nx, ny=inp.ny, nz=inp.nz if inp.ndim == 3 else 1, nt=inp.nt,
        dx=inp.dx, dy=inp.dy, dz=inp.dz, dt=inp.dt, c0=inp.c0,
        ncoords=inp.ncoords, ntic=inp.ntic, ncoordsout=inp.ncoordsout, modt=inp.modt,
        ncoordszero=inp.ncoordszero, ndmap=inp.ndmap,
        attenuation_model=1 if inp.exponential else 0, n_shots=int(n_shots),
    )


def kernel_override_string(cfg: dict) -> str:
    """``{"velocity": {...}, "pressure": {...}}`` (as ``HipSimulation.kernel_config`` returns)
    -> ``"<velocity>/<pressure>"`` for ``FULLWAVE_HIP_KERNELS``."""

    def one(c):
        if c["variant"] == "percell":
            return "percell"
        if c["variant"] == "tile":
            return f"tile:{c['chunk']}"
        return f"{c['variant']}:{c['width']}:{c['chunk']}"

    return f"{one(cfg['velocity'])}/{one(cfg['pressure'])}"


class HipSimulation:
    """One engine handle with its inputs resident in HBM.

    ``shots``: source matrices ``[n_shots, ncoords, ntic]`` of several transmits on this one
    medium (same source points, different signals), advanced together with the maps stored
    once; then :meth:`sensors` returns ``[n_shots, n_records, n_sensors]``. Without it the
    handle runs ``inp.icmat`` alone.
    """

    def __init__(self, inp: SimulationInputs, device: int = 0, *, shots: np.ndarray | None = None) -> None:
        self.lib = load_library()
        self.inp = inp
        self.batched = shots is not None
        if shots is not None:
            shots = np.ascontiguousarray(shots, dtype=np.float32)
            if shots.ndim != 3 or shots.shape[1:] != (inp.ncoords, inp.ntic):
                raise ValueError(f"shots must be [n_shots, {inp.ncoords}, {inp.ntic}], got {shots.shape}")
        self.n_shots = 1 if shots is None else int(shots.shape[0])
        self._shots = shots
        self._h = ctypes.c_void_p()
        desc = make_desc(inp, self.n_shots)
        _check(self.lib, self.lib.fw_sim_create(ctypes.byref(desc), int(device), ctypes.byref(self._h)))
        try:
            self._upload_all()
        except Exception:
            self.close()
            raise

    def upload(self, name: str, arr: np.ndarray) -> None:
        a = np.ascontiguousarray(arr)
        _check(self.lib, self.lib.fw_sim_upload(self._h, name.encode(), a.ctypes.data_as(ctypes.c_void_p), a.nbytes))

    def _upload_all(self) -> None:
        inp = self.inp
        for name in ["K", "rho", *inp.attenuation_map_stems()]:
            self.upload(name, inp.maps[name].astype(np.float32, copy=False))
        self.upload("dcmap", inp.dcmap.astype(np.int32, copy=False))
        self.upload("dmap", inp.dmap.astype(np.float32, copy=False))
        self.upload("icc", inp.icc.astype(np.int32, copy=False))
        self.upload("icmat", inp.icmat.astype(np.float32, copy=False) if self._shots is None else self._shots)
        self.upload("outc", inp.outc.astype(np.int32, copy=False))
        if inp.ncoordszero:
            self.upload("icczero", inp.icczero.astype(np.int32, copy=False))

    def set_shots(self, shots: np.ndarray) -> None:
        """New source matrices (``[n_shots, ncoords, ntic]``, or ``[ncoords, ntic]`` for a
        single handle) on the same medium; call :meth:`reset` before running them."""
        a = np.ascontiguousarray(shots, dtype=np.float32)
        want = (self.n_shots, self.inp.ncoords, self.inp.ntic) if self.batched else (self.inp.ncoords, self.inp.ntic)
        if a.shape != want:
            raise ValueError(f"source matrices must be {want}, got {a.shape}")
        self.upload("icmat", a)

    def run(self, n_steps: int) -> None:
        _check(self.lib, self.lib.fw_sim_run(self._h, int(n_steps)))

    def run_timed(self, n_steps: int) -> dict:
        """Run with HIP events around every launch; returns summed ms per kernel kind."""
        ms = (ctypes.c_double * 8)()
        nk = ctypes.c_int32(0)
        _check(self.lib, self.lib.fw_sim_run_timed(self._h, int(n_steps), ms, ctypes.byref(nk)))
        return {KERNEL_KINDS[i]: ms[i] for i in range(nk.value)}

    def kernel_config(self) -> dict:
        """Kernel family/shape of the velocity and pressure kernels (chosen on first use)."""
        out = {}
        for kind, label in ((0, "velocity"), (1, "pressure")):
            v, w, c = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
            _check(self.lib, self.lib.fw_sim_kernel_config(self._h, kind, ctypes.byref(v), ctypes.byref(w), ctypes.byref(c)))
            name = {0: "percell", 1: "stream", 2: "tile", 3: "pipe", 4: "pipelite"}.get(v.value, str(v.value))
            out[label] = {"variant": name, "width": w.value, "chunk": c.value}
        return out

    def set_kernel_config(self, cfg: dict) -> None:
        """Pin the kernels (a ``kernel_config()`` dict, either key optional) before the first run."""
        for kind, label in ((0, "velocity"), (1, "pressure")):
            if label in cfg:
                c = cfg[label]
                _check(self.lib, self.lib.fw_sim_set_kernel_config(self._h, kind, VARIANTS[c["variant"]],
                                                                     int(c["width"]), int(c["chunk"])))

    def set_graph_steps(self, steps: int) -> None:
        """Replay steady steps from a captured HIP graph of ``steps`` steps (0 turns it off)."""
        _check(self.lib, self.lib.fw_sim_set_graph_steps(self._h, int(steps)))

    def set_relaxation_codes(self, enable: bool) -> None:
        """Let the marching kernels read coded relaxation coefficients (the default) or the
        ten per-cell maps; results are the same either way."""
        _check(self.lib, self.lib.fw_sim_set_relaxation_codes(self._h, 1 if enable else 0))

    def relaxation_codes(self) -> int:
        """Relaxation tuples coded for this medium (0: per-cell maps in use)."""
        return self.relaxation_coding()[0]

    def relaxation_coding(self) -> tuple[int, float]:
        """(codes in use, fraction of device cells carrying a code)."""
        n = ctypes.c_int32(0)
        f = ctypes.c_double(0.0)
        _check(self.lib, self.lib.fw_sim_relaxation_codes(self._h, ctypes.byref(n), ctypes.byref(f)))
        return n.value, f.value

    def kernel_override(self) -> str:
        """The kernel configuration as a ``FULLWAVE_HIP_KERNELS`` value that pins it."""
        return kernel_override_string(self.kernel_config())

    def reset(self) -> None:
        _check(self.lib, self.lib.fw_sim_reset(self._h))

    @property
    def step(self) -> int:
        v = ctypes.c_int64(0)
        _check(self.lib, self.lib.fw_sim_step_count(self._h, ctypes.byref(v)))
        return v.value

    def sensors(self) -> np.ndarray:
        """Recorded traces ``[n_records, n_sensors]`` (genout.dat layout), or
        ``[n_shots, n_records, n_sensors]`` for a batch."""
        n = ctypes.c_int64(0)
        _check(self.lib, self.lib.fw_sim_record_count(self._h, ctypes.byref(n)))
        out = np.empty((self.n_shots, n.value, self.inp.ncoordsout), dtype=np.float32)
        _check(self.lib, self.lib.fw_sim_download_sensors(self._h, out.ctypes.data_as(ctypes.c_void_p), out.nbytes))
        return out if self.batched else out[0]

    def field(self, name: str, *, all_shots: bool = False) -> np.ndarray:
        """Wave field ``name`` of the first shot, or ``[n_shots, ...]`` with ``all_shots``."""
        out = np.empty(((self.n_shots,) if all_shots else ()) + tuple(self.inp.shape), dtype=np.float32)
        _check(self.lib, self.lib.fw_sim_download_field(self._h, name.encode(), out.ctypes.data_as(ctypes.c_void_p), out.nbytes))
        return out

    def set_field(self, name: str, values: np.ndarray) -> None:
        a = np.ascontiguousarray(values, dtype=np.float32)
        if a.shape != tuple(self.inp.shape):
            raise ValueError(f"field shape {a.shape} != grid {self.inp.shape}")
        _check(self.lib, self.lib.fw_sim_set_field(self._h, name.encode(), a.ctypes.data_as(ctypes.c_void_p), a.nbytes))

    def close(self) -> None:
        if self._h:
            self.lib.fw_sim_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self) -> HipSimulation:
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def __del__(self) -> None:
        try:
            self.close()
        except Exception:
            pass


class HipGroup:
    """Several 2D isotropic simulations on one device (any grid sizes, media and shot counts)
    advanced together: each time step is one velocity and one pressure launch over all of them
    (``fw_group_*``), the pressure launch also injecting each member's sources, zeroing its
    pressure-release points and recording its sensors. Members keep their own records
    (``member.sensors()``), bit-identical to ``member.run`` with the group's kernel shape
    (``kernel_config()``): ``variant="stream"``, the x-marching kernels ``stream:<width>:<chunk>``;
    ``variant="tile"``, the tiled kernels ``tile:<chunk>`` (members with ny % 4 == 0);
    ``variant="auto"``, whichever family and shape runs fastest. ``width``/``chunk`` 0: timed on
    the group when every member is fresh. Members not yet set up skip their own autotuning.
    Members may be at different steps (a job that joins a running group) and may run alone
    between group runs; they must stay open while the group exists."""

    _FAMILIES = {"auto": 0, "stream": 1, "tile": 2}

    def __init__(self, sims: list[HipSimulation], *, width: int = 0, chunk: int = 0, variant: str = "stream") -> None:
        self.lib = load_library()
        self.sims = list(sims)
        self._h = ctypes.c_void_p()
        if variant not in self._FAMILIES:
            raise ValueError(f"variant must be one of {sorted(self._FAMILIES)}, got {variant!r}")
        # members not yet set up skip their own autotuning (the group runs its own shape)
        if variant == "tile":
            shape = {"variant": "tile", "width": 256, "chunk": int(chunk) or 16}
        else:
            shape = {"variant": "stream", "width": int(width) or 256, "chunk": int(chunk) or 16}
        for sim in self.sims:
            try:
                sim.set_kernel_config({"velocity": shape, "pressure": shape})
            except EngineError as e:
                if e.status != 5:  # FW_ERR_STATE: already set up, keeps its own kernels
                    raise
        arr = (ctypes.c_void_p * len(self.sims))(*[sim._h.value for sim in self.sims])
        _check(self.lib, self.lib.fw_group_create_shaped(arr, len(self.sims), self._FAMILIES[variant], int(width),
                                                         int(chunk), ctypes.byref(self._h)))

    def kernel_config(self) -> dict:
        """The group's kernel shape, as ``HipSimulation.kernel_config`` reports one."""
        v, w, c = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        _check(self.lib, self.lib.fw_group_shape(self._h, ctypes.byref(v), ctypes.byref(w), ctypes.byref(c)))
        shape = {"variant": "tile" if v.value == 2 else "stream", "width": w.value, "chunk": c.value}
        return {"velocity": shape, "pressure": shape}

    def set_graph_steps(self, steps: int) -> None:
        """Replay steady steps from a captured HIP graph of ``steps`` group steps (0: plain
        launches); records are the same either way."""
        _check(self.lib, self.lib.fw_group_set_graph_steps(self._h, int(steps)))

    def run(self, n_steps: int) -> None:
        _check(self.lib, self.lib.fw_group_run(self._h, int(n_steps)))

    def close(self) -> None:
        if self._h:
            self.lib.fw_group_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self) -> HipGroup:
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def __del__(self) -> None:
        try:
            self.close()
        except Exception:
            pass
