"""The C-ABI library: it loads, exports every declared symbol, and validates without a GPU."""

import ctypes
import re
from pathlib import Path

import pytest

from fullwave25_amd import engine

HEADER = Path(__file__).resolve().parents[1] / "include" / "fullwave_hip.h"


def declared_functions() -> set[str]:
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(fw_[a-z_]+)\s*\(", text))


def test_library_is_built():
    assert engine.LIB_PATH.exists(), "run __graft_entry__.build() first"


def test_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(engine.LIB_PATH))
    declared = declared_functions()
    assert declared, "no declarations parsed"
    missing = [name for name in sorted(declared) if not hasattr(lib, name)]
    assert not missing, missing
    assert declared == set(engine.EXPORTED_SYMBOLS)


def test_abi_version_and_string():
    lib = engine.load_library()
    assert lib.fw_abi_version() == engine.ABI_VERSION
    assert b"gfx950" in lib.fw_version_string()


def test_invalid_descriptor_rejected_before_device():
    lib = engine.load_library()
    desc = engine.FwDesc(abi_version=engine.ABI_VERSION, ndim=4, isotropic=1, n_relax=2, nx=32, ny=32, nz=1, nt=1,
                         dx=1e-4, dy=1e-4, dz=0, dt=1e-8, c0=1540, modt=1, ndmap=1)
    h = ctypes.c_void_p()
    rc = lib.fw_sim_create(ctypes.byref(desc), 0, ctypes.byref(h))
    assert rc == 1  # FW_ERR_INVALID
    assert b"ndim" in lib.fw_last_error()
    desc.ndim, desc.n_relax = 2, 3
    assert lib.fw_sim_create(ctypes.byref(desc), 0, ctypes.byref(h)) == 6  # FW_ERR_UNSUPPORTED
    desc.n_relax, desc.nx = 2, 10
    assert lib.fw_sim_create(ctypes.byref(desc), 0, ctypes.byref(h)) == 1
    assert b"ghost" in lib.fw_last_error()


def test_too_many_shots_rejected_before_device():
    lib = engine.load_library()
    desc = engine.FwDesc(abi_version=engine.ABI_VERSION, ndim=3, isotropic=1, n_relax=2, nx=1000, ny=32, nz=32,
                         nt=1, dx=1e-4, dy=1e-4, dz=1e-4, dt=1e-8, c0=1540, modt=1, ndmap=1, n_shots=70)
    h = ctypes.c_void_p()
    assert lib.fw_sim_create(ctypes.byref(desc), 0, ctypes.byref(h)) == 6  # FW_ERR_UNSUPPORTED
    assert b"split the batch" in lib.fw_last_error()


def test_null_arguments_rejected():
    lib = engine.load_library()
    assert lib.fw_sim_run(None, 1) == 1
    assert lib.fw_device_count(None) == 1
    assert lib.fw_run_dir(None, 0) == 1
    assert lib.fw_sim_destroy(None) == 0


def test_group_arguments_rejected_before_device():
    """fw_group_create_shaped checks the kernel family, width and chunk before it reads any
    handle (the pointer below is never dereferenced)."""
    lib = engine.load_library()
    sims = (ctypes.c_void_p * 1)(ctypes.c_void_p(1))
    g = ctypes.c_void_p()
    for (variant, width, chunk), msg in [((3, 0, 0), b"variant"), ((2, 128, 0), b"256 columns"),
                                         ((2, 0, 12), b"8, 16 or 32"), ((0, 256, 0), b"timed kernel family"),
                                         ((1, 100, 16), b"64, 128 or 256"), ((1, 256, 1), b"chunk")]:
        assert lib.fw_group_create_shaped(sims, 1, variant, width, chunk, ctypes.byref(g)) == 1, (variant, width, chunk)
        assert msg in lib.fw_last_error(), lib.fw_last_error()
        assert not g.value
    assert lib.fw_group_create_shaped(None, 1, 1, 0, 0, ctypes.byref(g)) == 1
    v, w, c = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    assert lib.fw_group_shape(None, ctypes.byref(v), ctypes.byref(w), ctypes.byref(c)) == 1
    assert lib.fw_group_destroy(None) == 0


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(RuntimeError, match="missing"):
        engine.load_library(tmp_path / "libfullwave_hip.so")


def test_noalloc_rejects_oversized_run_before_device(tmp_path):
    """fw_run_dir_noalloc checks the run's device footprint against its fixed array before any
    device call or map read: a directory of scalar files alone is enough to be refused."""
    import numpy as np

    ints = {"nX": 4000, "nY": 4000, "nT": 100, "ncoords": 1, "ncoordsout": 1, "nTic": 100, "modT": 1, "ndmap": 1}
    floats = {"dX": 1e-4, "dY": 1e-4, "dT": 1e-8, "c0": 1540.0}
    for stem, v in ints.items():
        np.array([v], np.int32).tofile(tmp_path / f"{stem}.dat")
    for stem, v in floats.items():
        np.array([v], np.float32).tofile(tmp_path / f"{stem}.dat")
    lib = engine.load_library()
    assert lib.fw_run_dir_noalloc(str(tmp_path).encode(), 0) == 6  # FW_ERR_UNSUPPORTED
    msg = lib.fw_last_error()
    assert b"allocation-free" in msg and str(engine.NOALLOC_SCRATCH_BYTES).encode() in msg, msg
    assert lib.fw_run_dir_noalloc(None, 0) == 1
    (tmp_path / "nX.dat").unlink()
    assert lib.fw_run_dir_noalloc(str(tmp_path).encode(), 0) == 4  # FW_ERR_IO
