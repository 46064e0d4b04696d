"""Solver: the public entry point (mirror of fullwave/solver/solver.py:276-778).

Same constructor keywords and ``run()`` keywords as the reference. ``run()`` builds the padded
medium and the absorbing layer, assembles the engine inputs and runs the HIP engine, then
returns the sensor traces as ``[n_sensors, n_out]`` (fullwave/solver/solver.py:600-618).

Two routes reach the engine:

* in memory (default): inputs go straight to HBM through the C ABI, no files;
* the directory contract: with ``write_input_files=True`` or ``load_results=False`` the
  reference's ``.dat`` files are written under ``work_dir`` (with static-map symlinks when
  ``is_static_map``) and the engine reads them like the reference's executable did.

The reference picks a CUDA binary by name (fullwave/solver/solver.py:177-273); here the variant
(2D/3D, isotropic or not) is a runtime field of the engine descriptor and no binary or CUDA
probing is involved.
"""

from __future__ import annotations

import logging
import tempfile
import time
from pathlib import Path

import numpy as np

from ..grid import Grid
from ..medium import Medium, MediumRelaxationMaps
from ..sensor import Sensor
from ..source import Source
from .inputs import SimulationInputs, build_simulation_inputs
from .launcher import HipLauncher
from .pml_builder import PMLBuilder, PMLBuilderExponentialAttenuation

logger = logging.getLogger("__main__." + __name__)


class Solver:
    """Sets up and runs one Fullwave simulation on the HIP engine."""

    def __init__(
        self,
        work_dir: Path,
        grid: Grid,
        medium: Medium | MediumRelaxationMaps,
        source: Source | None = None,
        sensor: Sensor | None = None,
        *,
        transducer=None,
        path_fullwave_simulation_bin: Path | None = None,
        use_pml: bool = True,
        m_spatial_order: int = 8,
        pml_layer_thickness_px: int | None = None,
        n_transition_layer: int | None = None,
        run_on_memory: bool = False,
        use_gpu: bool = True,
        use_exponential_attenuation: bool = False,
        use_isotropic_relaxation: bool = True,
        cuda_device_id=None,
        write_input_files: bool = False,
    ) -> None:
        if m_spatial_order != 8:
            raise ValueError("the engine's stencil half-width is 8 (m_spatial_order=8)")
        if run_on_memory:
            self._tempdir = tempfile.TemporaryDirectory()
            self.work_dir = Path(self._tempdir.name)
        else:
            self.work_dir = Path(work_dir)
            self.work_dir.mkdir(exist_ok=True, parents=True)
        self.grid = grid
        self.is_3d = grid.is_3d
        self.use_gpu = use_gpu
        self.use_exponential_attenuation = use_exponential_attenuation
        self.use_isotropic_relaxation = use_isotropic_relaxation
        self.write_input_files = write_input_files
        self.n_relax_mechanisms = getattr(medium, "n_relaxation_mechanisms", 2)
        self._check_input(grid, medium, source, sensor, transducer)

        self.medium = medium
        if use_exponential_attenuation and not isinstance(medium, Medium):
            raise TypeError("the exponential-attenuation model needs a Medium (alpha_coeff)")
        if not use_exponential_attenuation and use_isotropic_relaxation != medium.use_isotropic_relaxation:
            logger.warning(
                "Solver is set to use %s relaxation, but the provided medium is using %s "
                "relaxation. Overriding the medium.",
                "isotropic" if use_isotropic_relaxation else "anisotropic",
                "isotropic" if medium.use_isotropic_relaxation else "anisotropic",
            )
        self.medium.use_isotropic_relaxation = use_isotropic_relaxation

        self.use_pml = use_pml
        if not use_pml:
            pml_layer_thickness_px = 0
            n_transition_layer = 0
        if pml_layer_thickness_px is None:
            pml_layer_thickness_px = self.grid.ppw * 3
        if n_transition_layer is None:
            n_transition_layer = self.grid.ppw * 3

        self.source = source if source is not None else transducer.source
        self.sensor = sensor if sensor is not None else transducer.sensor
        self.transducer = transducer
        self.path_fullwave_simulation_bin = path_fullwave_simulation_bin
        self.cuda_device_id = cuda_device_id
        self.fullwave_launcher = HipLauncher(is_3d=self.is_3d, use_gpu=use_gpu, cuda_device_id=cuda_device_id)
        if use_exponential_attenuation:
            self.pml_builder = PMLBuilderExponentialAttenuation(
                grid=self.grid, medium=self.medium, source=self.source, sensor=self.sensor,
                m_spatial_order=m_spatial_order, n_pml_layer=pml_layer_thickness_px,
            )
        else:
            self.pml_builder = PMLBuilder(
                grid=self.grid, medium=self.medium, source=self.source, sensor=self.sensor,
                m_spatial_order=m_spatial_order, n_pml_layer=pml_layer_thickness_px,
                n_transition_layer=n_transition_layer, use_isotropic_relaxation=use_isotropic_relaxation,
            )
        if self.medium.beta is not None and np.any(np.asarray(self.medium.beta) != 0):
            logger.warning("beta (nonlinearity) is non-zero; this engine solves the linear equations and ignores it")

    @staticmethod
    def _check_input(grid, medium, source, sensor, transducer) -> None:
        """Input checks of fullwave/solver/solver.py:539-597."""
        if source is None and transducer is None:
            raise ValueError("source or transducer must be defined")
        if sensor is None and transducer is None:
            raise ValueError("sensor or transducer must be defined")
        if transducer is not None and source is not None:
            raise ValueError("source and transducer cannot be defined at the same time")
        if transducer is not None and sensor is not None:
            logger.warning("sensor and transducer are defined at the same time. It uses sensor instead of transducer.sensor.")
        if not isinstance(grid, Grid):
            raise TypeError("grid must be a Grid")
        if not isinstance(medium, (Medium, MediumRelaxationMaps)):
            raise TypeError("medium must be a Medium or MediumRelaxationMaps")
        shape = (grid.nx, grid.ny, grid.nz) if grid.is_3d else (grid.nx, grid.ny)
        if source is not None:
            if not isinstance(source, Source):
                raise TypeError("source must be a Source")
            source.validate(grid_shape=shape)
        if sensor is not None:
            if not isinstance(sensor, Sensor):
                raise TypeError("sensor must be a Sensor")
            sensor.validate(grid_shape=shape)

    @staticmethod
    def _reshape_sensor_data(raw, sensor) -> np.ndarray:
        """``[n_out * n_sensors]`` or ``[n_out, n_sensors]`` -> ``[n_sensors, n_out]``."""
        return np.asarray(raw).reshape(-1, sensor.n_sensors).T

    def build_inputs(self, *, record_whole_domain: bool = False,
                     sampling_modulus_time_whole_domain: int = 1) -> tuple[SimulationInputs, Sensor]:
        """Padded medium + absorbing layer -> the engine inputs (no engine call)."""
        t0 = time.time()
        extended_medium = self.pml_builder.run(use_pml=self.use_pml)
        logger.debug("PML building completed in %.2e seconds.", time.time() - t0)
        if sampling_modulus_time_whole_domain != 1 and not record_whole_domain:
            logger.warning(
                "sampling_modulus_time_whole_domain %s is ignored when record_whole_domain is False.",
                sampling_modulus_time_whole_domain,
            )
        if record_whole_domain:
            g = self.pml_builder.extended_grid
            shape = (g.nx, g.ny, g.nz) if self.is_3d else (g.nx, g.ny)
            sensor = Sensor(mask=np.ones(shape, dtype=bool), sampling_modulus_time=sampling_modulus_time_whole_domain)
        else:
            sensor = self.pml_builder.extended_sensor
        inputs = build_simulation_inputs(
            self.pml_builder.extended_grid, extended_medium, self.pml_builder.extended_source, sensor,
            isotropic=self.use_isotropic_relaxation, n_relax=self.n_relax_mechanisms,
        )
        return inputs, sensor

    def run(
        self,
        simulation_dir_name: str | Path = "txrx_0",
        *,
        is_static_map: bool = False,
        recalculate_pml: bool = True,
        record_whole_domain: bool = False,
        sampling_modulus_time_whole_domain: int = 1,
        load_results: bool = True,
    ):
        """Run the simulation; returns ``[n_sensors, n_out]`` traces (or the ``genout.dat``
        path when ``load_results`` is False)."""
        use_files = self.write_input_files or not load_results
        sim_dir = self.work_dir / simulation_dir_name
        if use_files and not recalculate_pml and is_static_map and (self.work_dir / "K.dat").exists():
            # static maps already on disk (fullwave/solver/input_file_writer.py:156-170): only
            # the source matrix is new; everything else is linked from work_dir
            if record_whole_domain:
                raise ValueError("record_whole_domain needs recalculate_pml=True")
            sensor = self.pml_builder.extended_sensor
            sim_dir.mkdir(parents=True, exist_ok=True)
            np.asarray(self.pml_builder.extended_source.icmat).astype(np.float32).tofile(sim_dir / "icmat.dat")
            for f in self.work_dir.glob("*.dat"):
                if f.name in ("icmat.dat", "genout.dat"):
                    continue
                link = sim_dir / f.name
                if link.exists() or link.is_symlink():
                    link.unlink()
                link.symlink_to(f.absolute())
        else:
            inputs, sensor = self.build_inputs(record_whole_domain=record_whole_domain,
                                               sampling_modulus_time_whole_domain=sampling_modulus_time_whole_domain)
            if not use_files:
                traces = self.fullwave_launcher.run_inputs(inputs)
                return self._reshape_sensor_data(traces, sensor)
            inputs.write_dir(sim_dir, maps_dir=self.work_dir if is_static_map else None)
        result = self.fullwave_launcher.run(sim_dir, load_results=load_results)
        if load_results:
            return self._reshape_sensor_data(result, sensor)
        return result

    def print_info(self) -> None:
        print(str(self))

    def summary(self) -> None:
        self.print_info()

    def __str__(self) -> str:
        return (
            "\nSolver(\n"
            f"  work_dir={self.work_dir}\n"
            f"  use_pml={self.use_pml}\n"
            f"  pml_thickness_px={self.pml_builder.n_pml_layer}\n"
            f"  n_transition_layer={self.pml_builder.n_transition_layer}\n"
            f"  is_3d={self.is_3d}\n"
            f"  use_gpu={self.use_gpu}\n"
            f"  use_isotropic_relaxation={self.use_isotropic_relaxation}\n"
            ")"
        )

    __repr__ = __str__


__all__ = ["Solver"]
