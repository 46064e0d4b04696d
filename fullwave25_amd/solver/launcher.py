"""Engine launcher: the drop-in for the reference's process launcher.

``HipLauncher.run(simulation_dir, load_results=...)`` keeps the contract of
fullwave/solver/launcher.py:160-254 — same arguments, ``genout.dat`` contents or path as the
return value, ``SimulationError`` naming ``fw2_execution.log`` on failure, ``NotImplementedError``
when ``use_gpu=False`` — but runs the HIP engine in process through the C ABI instead of a
CUDA executable. It does not ``chdir`` (the reference's ``os.chdir`` made it process-global) and
needs neither ``nvidia-smi`` nor ``libcuda``.
"""

from __future__ import annotations

import logging
import time
from pathlib import Path

import numpy as np

from .. import engine
from .utils import load_dat_data, warn_non_finite

logger = logging.getLogger("__main__." + __name__)


class SimulationError(Exception):
    """The engine reported a failure."""


def parse_device_id(device_id) -> int:
    """Accept the reference's ``cuda_device_id`` forms (fullwave/solver/launcher.py:63-105) and
    return the one device this engine runs on."""
    if device_id is None:
        return 0
    if isinstance(device_id, bool):
        raise ValueError("CUDA device ID must be an integer, string, list, or None.")
    if isinstance(device_id, int):
        if device_id < 0:
            raise ValueError("CUDA device ID must be a non-negative integer.")
        return device_id
    if isinstance(device_id, str):
        if "," in device_id:
            ids = device_id.split(",")
            if not all(i.isdigit() for i in ids):
                raise ValueError("CUDA device ID string must represent a non-negative integer.")
            return parse_device_id([int(i) for i in ids])
        if not device_id.isdigit():
            raise ValueError("CUDA device ID string must represent a non-negative integer.")
        return int(device_id)
    if isinstance(device_id, list):
        if not all(isinstance(i, int) and not isinstance(i, bool) and i >= 0 for i in device_id):
            raise ValueError("All CUDA device IDs in the list must be non-negative integers.")
        if len(device_id) != 1:
            raise NotImplementedError(
                "this engine runs one simulation on one GPU; pass a single device id "
                "(run independent simulations on other GPUs in other processes)",
            )
        return device_id[0]
    raise ValueError("CUDA device ID must be an integer, string, list, or None.")


class HipLauncher:
    """Runs the HIP engine on a simulation directory or on in-memory inputs."""

    def __init__(
        self,
        path_fullwave_simulation_bin: Path | None = None,
        *,
        is_3d: bool = False,
        use_gpu: bool = True,
        cuda_device_id=None,
    ) -> None:
        # the engine is a library; a binary path is accepted for call compatibility only
        self._path_fullwave_simulation_bin = path_fullwave_simulation_bin
        self.is_3d = is_3d
        self.use_gpu = use_gpu
        self.device = parse_device_id(cuda_device_id)
        self.cuda_device_id = str(self.device)

    def _verify_device(self) -> None:
        n = engine.device_count()
        if self.device >= n:
            raise ValueError(f"CUDA device ID {self.device} does not exist.")

    def run(self, simulation_dir: Path, *, load_results: bool = True):
        """Run the engine on ``simulation_dir`` and return ``genout.dat`` (float32, flat) or
        its path when ``load_results`` is False."""
        simulation_dir = Path(simulation_dir).absolute()
        if not self.use_gpu:
            msg = "Currently, only GPU version is supported."
            logger.error(msg)
            raise NotImplementedError(msg)
        log_path = simulation_dir / "fw2_execution.log"
        logger.info("Running simulation...")
        t0 = time.time()
        with log_path.open("w", encoding="utf-8") as log:
            log.write(f"{engine.load_library().fw_version_string().decode()}\n")
            log.write(f"device {self.device}, directory {simulation_dir}\n")
            try:
                engine.run_dir(simulation_dir, self.device)
            except Exception as e:  # engine status or missing library
                log.write(f"error: {e}\n")
                log.flush()
                logger.exception("Simulation failed")
                msg = (
                    "Simulation failed. please check the simulation log file for more information.\n"
                    "The log file is located at:\n"
                    f"{log_path}"
                )
                raise SimulationError(msg) from e
            log.write(f"completed in {time.time() - t0:.3e} s\n")
        logger.info("Simulation completed in %.2e seconds.", time.time() - t0)
        if load_results:
            return load_dat_data(simulation_dir / "genout.dat")
        return simulation_dir / "genout.dat"

    def run_inputs(self, inputs) -> np.ndarray:
        """Run in-memory inputs; returns the traces as ``genout.dat`` would hold them
        (``[n_records, n_sensors]`` float32)."""
        if not self.use_gpu:
            raise NotImplementedError("Currently, only GPU version is supported.")
        t0 = time.time()
        try:
            with engine.HipSimulation(inputs, device=self.device) as sim:
                sim.run(inputs.nt)
                out = sim.sensors()
        except engine.EngineError as e:
            raise SimulationError(f"Simulation failed: {e}") from e
        logger.info("Simulation completed in %.2e seconds.", time.time() - t0)
        warn_non_finite(out)
        return out


# the reference's name for the launcher class
Launcher = HipLauncher
