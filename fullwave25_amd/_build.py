"""Compile the HIP engine (``libfullwave_hip.so``) in-tree for gfx950.

Used by ``__graft_entry__.build()``; also runnable as ``python -m fullwave25_amd._build``.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OUT = PKG / "libfullwave_hip.so"
SOURCES = ["fw_api.cpp", "fw_adjoint.cpp", "fw_group.cpp", "fw_kernels_simple.hip", "fw_kernels_stream.hip", "fw_kernels_stream3d.hip",
           "fw_kernels_tile2d.hip", "fw_kernels_pipe3d.hip", "fw_kernels_adjoint.hip", "fw_kernels_codes.hip",
           "fw_noalloc.hip"]

HIPCC_FLAGS = [
    "-O3",
    "--offload-arch=gfx950",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # packed FP32 instructions are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
]


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> Path:
    """Compile every engine source into ``fullwave25_amd/libfullwave_hip.so`` (always rebuilds)."""
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [hipcc(), *HIPCC_FLAGS, *[str(CSRC / s) for s in SOURCES], "-o", str(tmp)]
    env = dict(os.environ)
    res = subprocess.run(cmd, capture_output=True, text=True, env=env)
    noise = "packed-fp32-ops"  # the host pass does not know the GPU feature flag
    msgs = "\n".join(l for l in (res.stdout + res.stderr).splitlines() if noise not in l)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{' '.join(cmd)}\n{msgs}")
    if verbose and msgs.strip():
        print(msgs, file=sys.stderr)
    tmp.replace(OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
