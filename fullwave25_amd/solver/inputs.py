"""The engine's input contract held in memory.

:class:`SimulationInputs` carries exactly what the reference's ``InputFileWriter`` writes into
a simulation directory (fullwave/solver/input_file_writer.py:563-881): per-voxel maps, the
coefficient tables ``d``/``dmap``/``dcmap``, coordinate lists, the source matrix and the
scalars. Arrays use the file dtypes (float32 maps, int32 coordinates and indices); scalars are
rounded through float32 the way the ``.dat`` files store them, so the in-memory path and the
directory path hand the engine identical numbers.

``write_dir`` writes the contract (the ``.dat`` files; ``genout.dat`` is the engine's output)
and ``read_dir`` reads one back.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from pathlib import Path

import numpy as np
from numpy.typing import NDArray

from ..medium import fw2_map_names
from ..utils.coordinates import map_to_coords
from . import coefficients

_F32 = np.float32


def _f32(x: float) -> float:
    return float(np.float32(x))


def engine_file_stem(map_name: str) -> str:
    """``kappa_u`` -> ``kappau``, ``a_pml_u1`` -> ``apmlu1`` (file names of the contract,
    fullwave/solver/input_file_writer.py:581-620)."""
    return map_name.replace("_", "")


@dataclass
class SimulationInputs:
    """Everything one engine run reads."""

    ndim: int
    nx: int
    ny: int
    nz: int
    nt: int
    dx: float
    dy: float
    dz: float
    dt: float
    c0: float
    isotropic: bool
    maps: dict = field(default_factory=dict)  # file stem -> float32 [nx, ny(, nz)]
    dcmap: NDArray[np.int32] | None = None
    d: NDArray[np.float32] | None = None  # [9, 2]
    dmap: NDArray[np.float32] | None = None  # [9, 2, ndmap]
    icc: NDArray[np.int32] | None = None  # [ncoords, ndim]
    icmat: NDArray[np.float32] | None = None  # [ncoords, ntic]
    outc: NDArray[np.int32] | None = None  # [nout, ndim]
    icczero: NDArray[np.int32] | None = None  # [nzero, ndim]
    modt: int = 1
    n_relax: int = 2
    exponential: bool = False  # exponential-attenuation model: one "a_exp" map, no relaxation maps

    @property
    def shape(self) -> tuple:
        return (self.nx, self.ny, self.nz) if self.ndim == 3 else (self.nx, self.ny)

    @property
    def ncells(self) -> int:
        return int(np.prod(self.shape))

    @property
    def ncoords(self) -> int:
        return int(self.icc.shape[0])

    @property
    def ntic(self) -> int:
        return int(self.icmat.shape[1])

    @property
    def ncoordsout(self) -> int:
        return int(self.outc.shape[0])

    @property
    def ncoordszero(self) -> int:
        return 0 if self.icczero is None else int(self.icczero.shape[0])

    @property
    def ndmap(self) -> int:
        return int(self.dmap.shape[2])

    @property
    def n_records(self) -> int:
        return -(-self.nt // self.modt)

    def relaxation_map_stems(self) -> list[str]:
        return [engine_file_stem(n) for n in fw2_map_names(is_3d=self.ndim == 3, isotropic=self.isotropic, n_relax=self.n_relax)]

    def attenuation_map_stems(self) -> list[str]:
        """The attenuation maps the engine reads for this model."""
        return ["a_exp"] if self.exponential else self.relaxation_map_stems()

    def validate(self) -> None:
        shape = self.shape
        for name in ["K", "rho", *self.attenuation_map_stems()]:
            if name not in self.maps:
                raise ValueError(f"missing map '{name}'")
            if self.maps[name].shape != shape:
                raise ValueError(f"map '{name}' has shape {self.maps[name].shape}, grid is {shape}")
        if self.dcmap.shape != shape:
            raise ValueError("dcmap shape mismatch")
        if self.icmat.shape[0] != self.icc.shape[0]:
            raise ValueError("icmat rows must match the number of source coordinates")

    # --- the directory contract -------------------------------------------------------------

    def write_dir(self, sim_dir: Path, *, maps_dir: Path | None = None) -> Path:
        """Write the ``.dat`` files of the contract into ``sim_dir``.

        With ``maps_dir`` the medium-dependent files go there and ``sim_dir`` gets symbolic
        links to them (the reference's static-map layout,
        fullwave/solver/input_file_writer.py:156-170, 647-714); ``icmat.dat`` always goes to
        ``sim_dir``.
        """
        sim_dir = Path(sim_dir)
        sim_dir.mkdir(parents=True, exist_ok=True)
        target = Path(maps_dir) if maps_dir is not None else sim_dir
        target.mkdir(parents=True, exist_ok=True)
        self.icmat.astype(_F32).tofile(sim_dir / "icmat.dat")
        files = {}
        for name in ("c", "K", "rho", "beta"):
            if name in self.maps:
                files[name] = self.maps[name].astype(_F32)
        files["icc"] = self.icc.astype(np.int32)
        files["outc"] = self.outc.astype(np.int32)
        files["icczero"] = (self.icczero if self.icczero is not None else np.zeros((0, self.ndim))).astype(np.int32)
        scal_f = {"dX": self.dx, "dY": self.dy, "dT": self.dt, "c0": self.c0}
        scal_i = {
            "nX": self.nx, "nY": self.ny, "nT": self.nt, "ncoords": self.ncoords,
            "ncoordsout": self.ncoordsout, "ncoordszero": self.ncoordszero, "nTic": self.ntic,
            "modT": self.modt, "ndmap": self.ndmap,
        }
        if self.ndim == 3:
            scal_f["dZ"] = self.dz
            scal_i["nZ"] = self.nz
        for k, v in scal_f.items():
            files[k] = np.array(v, dtype=_F32)
        for k, v in scal_i.items():
            files[k] = np.array(v, dtype=np.int32)
        files["d"] = self.d.astype(_F32)
        files["dmap"] = self.dmap.astype(_F32)
        files["dcmap"] = self.dcmap.astype(np.int32)
        for stem in self.attenuation_map_stems():
            files[stem] = self.maps[stem].astype(_F32)
        for k, arr in files.items():
            arr.tofile(target / f"{k}.dat")
            if maps_dir is not None:
                link = sim_dir / f"{k}.dat"
                if link.exists() or link.is_symlink():
                    link.unlink()
                link.symlink_to((target / f"{k}.dat").absolute())
        return sim_dir

    @classmethod
    def read_dir(cls, sim_dir: Path) -> SimulationInputs:
        """Read a simulation directory written by the reference's ``InputFileWriter`` or by
        :meth:`write_dir`."""
        sim_dir = Path(sim_dir)

        def rd(name, dtype, count=-1):
            return np.fromfile(sim_dir / f"{name}.dat", dtype=dtype, count=count)

        ndim = 3 if (sim_dir / "nZ.dat").exists() else 2
        iso = not (sim_dir / "kappaw.dat").exists()
        expo = (sim_dir / "a_exp.dat").exists()
        nx, ny, nt = int(rd("nX", np.int32)[0]), int(rd("nY", np.int32)[0]), int(rd("nT", np.int32)[0])
        nz = int(rd("nZ", np.int32)[0]) if ndim == 3 else 1
        shape = (nx, ny, nz) if ndim == 3 else (nx, ny)
        ndmap = int(rd("ndmap", np.int32)[0])
        ntic = int(rd("nTic", np.int32)[0])
        ncoords = int(rd("ncoords", np.int32)[0])
        nzero = int(rd("ncoordszero", np.int32)[0]) if (sim_dir / "ncoordszero.dat").exists() else 0
        inp = cls(
            ndim=ndim, nx=nx, ny=ny, nz=nz, nt=nt,
            dx=float(rd("dX", _F32)[0]), dy=float(rd("dY", _F32)[0]),
            dz=float(rd("dZ", _F32)[0]) if ndim == 3 else 0.0,
            dt=float(rd("dT", _F32)[0]), c0=float(rd("c0", _F32)[0]), isotropic=iso,
            modt=int(rd("modT", np.int32)[0]), exponential=expo,
        )
        for name in ("K", "rho", "c", "beta"):
            if (sim_dir / f"{name}.dat").exists():
                inp.maps[name] = rd(name, _F32).reshape(shape)
        for stem in inp.attenuation_map_stems():
            inp.maps[stem] = rd(stem, _F32).reshape(shape)
        inp.dcmap = rd("dcmap", np.int32).reshape(shape)
        inp.d = rd("d", _F32).reshape(9, 2)
        inp.dmap = rd("dmap", _F32).reshape(9, 2, -1)[:, :, :ndmap]
        inp.icc = rd("icc", np.int32).reshape(-1, ndim)
        inp.outc = rd("outc", np.int32).reshape(-1, ndim)
        inp.icczero = rd("icczero", np.int32).reshape(-1, ndim)[:nzero] if nzero else np.zeros((0, ndim), np.int32)
        inp.icmat = rd("icmat", _F32).reshape(ncoords, ntic)
        return inp


def build_simulation_inputs(grid, medium, source, sensor, *, isotropic: bool = True,
                            n_relax: int = 2) -> SimulationInputs:
    """Assemble the contract from padded front-end objects.

    ``grid``, ``medium`` (a :class:`~fullwave25_amd.medium.MediumRelaxationMaps` with the PML
    applied, or a :class:`~fullwave25_amd.medium.MediumExponentialAttenuation`), ``source`` and
    ``sensor`` are the padded objects of a ``PMLBuilder``; this does
    what ``InputFileWriter.__init__`` and ``_save_variables_into_dat_file`` do
    (fullwave/solver/input_file_writer.py:26-103, 563-627) without touching the disk.
    """
    is_3d = grid.is_3d
    c = np.asarray(medium.sound_speed, dtype=np.float64)
    exponential = hasattr(medium, "alpha_exp")
    inp = SimulationInputs(
        ndim=3 if is_3d else 2,
        nx=grid.nx, ny=grid.ny, nz=grid.nz if is_3d else 1, nt=grid.nt,
        dx=_f32(grid.dx), dy=_f32(grid.dy), dz=_f32(grid.dz) if is_3d else 0.0,
        dt=_f32(grid.dt), c0=_f32(grid.c0), isotropic=isotropic,
        modt=int(sensor.sampling_modulus_time), n_relax=n_relax, exponential=exponential,
    )
    inp.maps["c"] = c.astype(_F32)
    inp.maps["K"] = np.asarray(medium.bulk_modulus).astype(_F32)
    inp.maps["rho"] = np.asarray(medium.density).astype(_F32)
    inp.maps["beta"] = np.asarray(medium.beta).astype(_F32)
    if exponential:
        inp.maps["a_exp"] = np.asarray(medium.alpha_exp).astype(_F32)
    else:
        fw2 = medium.relaxation_param_dict_for_fw2
        for name in fw2_map_names(is_3d=is_3d, isotropic=isotropic, n_relax=n_relax):
            inp.maps[engine_file_stem(name)] = np.asarray(fw2[name]).astype(_F32)
    inp.d = coefficients.d_matrix(grid.cfl, is_3d=is_3d).astype(_F32)
    inp.dmap = coefficients.d_map(c, grid.dt, grid.dx, is_3d=is_3d).astype(_F32)
    if coefficients.ndmap_value(c) == 1:
        inp.dmap = inp.dmap[:, :, :1]
    inp.dcmap = coefficients.dc_map(c).astype(np.int32)
    inp.icc = np.asarray(source.incoords).astype(np.int32)
    inp.icmat = np.asarray(source.icmat).astype(_F32)
    inp.outc = np.asarray(sensor.outcoords).astype(np.int32)
    inp.icczero = map_to_coords(medium.air_map).astype(np.int32)
    inp.validate()
    return inp
