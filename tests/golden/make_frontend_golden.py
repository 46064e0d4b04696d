"""Write golden vectors of the reference front end's engine inputs.

The padding (``PMLBuilder._extend_map_for_pml``), the absorbing-layer ramps
(``PMLBuilder._apply_transition_and_pml``: exact slice boundaries and the per-axis sequence),
and the byte layout of every ``.dat`` file (``InputFileWriter``) are not fixed by any invariant
or closed form, so they are pinned by running the reference itself on small cases, in this
container only. The reference is imported from /root/reference with a stand-in for ``cv2``
(OpenCV is absent; the front end uses it only for plotting). Its engine is not involved.

Output: tests/golden/frontend_golden.npz holding, per case, the inputs given to the front end and
every file the reference's InputFileWriter writes (``<case>/<file stem>``) for isotropic and
per-axis relaxation and for the exponential-attenuation model, plus pulse samples.

Usage: python tests/golden/make_frontend_golden.py
"""

from __future__ import annotations

import sys
import tempfile
import types
from pathlib import Path

import numpy as np

REF = Path("/root/reference")
OUT = Path(__file__).resolve().parent / "frontend_golden.npz"


def cases():
    """Small 2D and 3D setups (arrays only; no package objects)."""
    rng = np.random.default_rng(20251202)
    out = {}
    for name, shape, npml, ntr in (("2d", (24, 20), 4, 3), ("3d", (6, 5, 4), 2, 2)):
        c = 1500.0 + 80.0 * rng.random(shape)
        c[: shape[0] // 2] = 1540.0
        rho = 950.0 + 100.0 * rng.random(shape)
        relax = {
            "kappa_x1": 1.0 + 0.01 * rng.random(shape),
            "kappa_x2": 1.0 + 0.01 * rng.random(shape),
        }
        for nu in (1, 2):
            for side in ("x1", "x2"):
                relax[f"d_{side}_nu{nu}"] = 1e5 * (1 + rng.random(shape)) * nu
                relax[f"alpha_{side}_nu{nu}"] = 1e6 * (1 + 3 * rng.random(shape)) * nu
        air = np.zeros(shape, dtype=bool)
        air.reshape(-1)[rng.choice(air.size, 3, replace=False)] = True
        src = np.zeros(shape, dtype=bool)
        if len(shape) == 2:
            src[0, 3:11] = True
        else:
            src[0, 1:4, 1:3] = True
        sen = np.zeros(shape, dtype=bool)
        sen.reshape(-1)[rng.choice(sen.size, 7, replace=False)] = True
        nt = 37
        p0 = rng.standard_normal((int(src.sum()), nt))
        out[name] = {
            "shape": np.array(shape), "c": c, "rho": rho, "beta": np.zeros(shape), "air": air,
            "src": src, "sen": sen, "p0": p0, "nt": np.array(nt), "modT": np.array(2),
            "n_pml": np.array(npml), "n_transition": np.array(ntr), **{f"relax/{k}": v for k, v in relax.items()},
        }
    # power-law maps for the exponential-attenuation model (a separate stream, so the draws
    # above stay as they were)
    rng_exp = np.random.default_rng(20251203)
    for cs in out.values():
        shape = tuple(int(v) for v in cs["shape"])
        cs["alpha_coeff"] = 0.2 + 1.3 * rng_exp.random(shape)
        cs["alpha_power"] = 0.8 + 0.8 * rng_exp.random(shape)
    return out


def grid_args(shape, nt):
    f0, c0, ppw, cfl = 3e6, 1540.0, 12, 0.4
    dx = c0 / f0 / ppw
    dt = cfl * dx / c0
    return dict(domain_size=tuple(s * dx for s in shape), f0=f0, duration=nt * dt, c0=c0, ppw=ppw, cfl=cfl)


def main() -> None:
    sys.modules.setdefault("cv2", types.ModuleType("cv2"))
    sys.path.insert(0, str(REF))
    import fullwave  # noqa: E402  (the reference)
    from fullwave.solver.input_file_writer import InputFileWriter
    from fullwave.solver.pml_builder import PMLBuilder, PMLBuilderExponentialAttenuation

    data = {}
    for name, cs in cases().items():
        shape = tuple(int(v) for v in cs["shape"])
        nt = int(cs["nt"])
        for k, v in cs.items():
            data[f"{name}/in/{k}"] = np.asarray(v)
        relax = {k[len("relax/"):]: v for k, v in cs.items() if k.startswith("relax/")}
        for iso in (True, False):
            tag = f"{name}_{'iso' if iso else 'aniso'}"
            grid = fullwave.Grid(**grid_args(shape, nt))
            medium = fullwave.MediumRelaxationMaps(grid, cs["c"], cs["rho"], cs["beta"], relax,
                                                   air_map=cs["air"], use_isotropic_relaxation=iso)
            source = fullwave.Source(cs["p0"], cs["src"])
            sensor = fullwave.Sensor(cs["sen"], sampling_modulus_time=int(cs["modT"]))
            pml = PMLBuilder(grid, medium, source, sensor, m_spatial_order=8, n_pml_layer=int(cs["n_pml"]),
                             n_transition_layer=int(cs["n_transition"]), use_isotropic_relaxation=iso)
            ext = pml.run(use_pml=True)
            with tempfile.TemporaryDirectory() as tmp:
                tmp = Path(tmp)
                fake_bin = tmp / "engine_stand_in"
                fake_bin.write_bytes(b"")
                writer = InputFileWriter(tmp, pml.extended_grid, ext, pml.extended_source, pml.extended_sensor,
                                         path_fullwave_simulation_bin=fake_bin,
                                         use_isotropic_relaxation=iso)
                sim = writer.run("txrx_0")
                for f in sorted(sim.glob("*.dat")):
                    data[f"{tag}/{f.stem}"] = np.fromfile(f, dtype=np.uint8)
        # exponential-attenuation model: its own padding (no transition layer), the air map
        # padded like a medium map, and the a_exp map with the radial absorbing mask
        tag = f"{name}_exp"
        with tempfile.TemporaryDirectory() as tmp:
            tmp = Path(tmp)
            # Medium checks that a relaxation database path exists; the exponential model never
            # reads it, so an empty file satisfies the check
            no_db = tmp / "unused_database.mat"
            no_db.write_bytes(b"")
            grid = fullwave.Grid(**grid_args(shape, nt))
            medium = fullwave.Medium(grid, cs["c"], cs["rho"], cs["alpha_coeff"], cs["alpha_power"], cs["beta"],
                                     air_map=cs["air"], path_relaxation_parameters_database=no_db)
            source = fullwave.Source(cs["p0"], cs["src"])
            sensor = fullwave.Sensor(cs["sen"], sampling_modulus_time=int(cs["modT"]))
            pml = PMLBuilderExponentialAttenuation(grid, medium, source, sensor, m_spatial_order=8,
                                                   n_pml_layer=int(cs["n_pml"]))
            ext = pml.run(use_pml=True)
            fake_bin = tmp / "engine_stand_in"
            fake_bin.write_bytes(b"")
            writer = InputFileWriter(tmp, pml.extended_grid, ext, pml.extended_source, pml.extended_sensor,
                                     path_fullwave_simulation_bin=fake_bin, use_exponential_attenuation=True)
            sim = writer.run("txrx_0")
            for f in sorted(sim.glob("*.dat")):
                data[f"{tag}/{f.stem}"] = np.fromfile(f, dtype=np.uint8)
    from fullwave.utils.pulse import gaussian_modulated_sinusoidal_signal

    for i, (nt, dur, nc, do, f0, p0, dl, il) in enumerate([
        (500, 5e-6, 3, 2, 3e6, 1e5, 0.0, None),
        (777, 7.3e-6, 2, 1, 2.5e6, 2.0, 3e-7, 2),
    ]):
        kw = dict(delay_sec=dl)
        if il:
            kw.update(i_layer=il, dt_for_layer_delay=1.1e-8, cfl_for_layer_delay=0.4)
        data[f"pulse{i}"] = gaussian_modulated_sinusoidal_signal(nt, dur, nc, do, f0, p0, **kw)
    np.savez_compressed(OUT, **data)
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes, {len(data)} arrays)")


if __name__ == "__main__":
    main()
