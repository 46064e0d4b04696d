"""The host code under AddressSanitizer and UndefinedBehaviorSanitizer, without a GPU:

* the oracle (`make -C oracle sanitize`: oracle/sanitize_check.c runs small 2D and 3D forward and
  adjoint problems, isotropic, per-axis and exponential, in both precisions);
* the engine library's host side (tests/host_api_check.cpp, built with every library source and
  the sanitizers on the host part only): descriptor validation, the directory contract's scalar
  files and the allocation-free run's size check, the resolution of source, air and sensor
  coordinates to device indices, the coefficient table, and the grouping of those points by the
  tile of a grouped launch that owns them.

Any finding aborts the program, so the test checks the exit status and the final line.
"""

import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
BUILD = ROOT / "tests" / "_build"
DRIVER = BUILD / "host_api_check"
CSRC = ROOT / "fullwave25_amd" / "csrc"


def _hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    pytest.skip("hipcc not found")


def test_oracle_under_sanitizers():
    res = subprocess.run(["make", "-C", str(ROOT / "oracle"), "sanitize"], capture_output=True, text=True, timeout=600)
    print(res.stdout[-2000:], res.stderr[-2000:])
    assert res.returncode == 0
    assert "sanitize check: ok" in res.stdout


def _build_driver() -> Path:
    sources = sorted(CSRC.glob("*.cpp")) + sorted(CSRC.glob("*.hip"))
    inputs = sources + [ROOT / "tests" / "host_api_check.cpp", ROOT / "include" / "fullwave_hip.h"] + \
        sorted(CSRC.glob("*.h"))
    if DRIVER.exists() and DRIVER.stat().st_mtime > max(p.stat().st_mtime for p in inputs):
        return DRIVER
    BUILD.mkdir(parents=True, exist_ok=True)
    cmd = [_hipcc(), "-std=c++17", "-O1", "-g", "--offload-arch=gfx950",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
           "-Xarch_host", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer",
           *map(str, sources), str(ROOT / "tests" / "host_api_check.cpp"), "-o", str(DRIVER)]
    res = subprocess.run(cmd, capture_output=True, text=True, timeout=900)
    assert res.returncode == 0, res.stderr[-4000:]
    return DRIVER


def test_host_api_under_sanitizers(tmp_path):
    driver = _build_driver()
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    res = subprocess.run([str(driver), str(tmp_path)], capture_output=True, text=True, timeout=300, env=env)
    print(res.stdout[-2000:], res.stderr[-4000:])
    assert res.returncode == 0
    assert "host api check: ok" in res.stdout
