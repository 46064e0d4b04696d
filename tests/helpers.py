"""Hand-made engine inputs for tests (no front end, no absorbing layer)."""

from __future__ import annotations

import numpy as np

from fullwave25_amd.medium import calc_a_and_b
from fullwave25_amd.solver import coefficients
from fullwave25_amd.solver.inputs import SimulationInputs


def raw_inputs(shape, *, c=1540.0, rho=1000.0, dx=1540.0 / 3e6 / 12, cfl=0.4, nt=100,
               relax=None, icc=None, icmat=None, outc=None, icczero=None, modt=1,
               isotropic=True, rng=None, a_exp=None) -> SimulationInputs:
    """Engine inputs on ``shape`` with maps given as scalars or arrays.

    ``relax=None`` is lossless (kappa = 1, d = 0); otherwise a dict with per-side ("v": the
    velocity side, maps u/w/v; "p": the pressure side, maps x/y/z) ``kappa``, ``d`` (2 values)
    and ``alpha`` (2 values) applied uniformly. With ``isotropic=False`` a side may instead hold a
    list of such dicts, one per axis, so every axis gets its own maps. ``a_exp`` (scalar or
    array) selects the exponential-attenuation model with that per-step decay map instead.
    """
    ndim = len(shape)
    c = np.broadcast_to(np.asarray(c, dtype=np.float64), shape).copy()
    rho = np.broadcast_to(np.asarray(rho, dtype=np.float64), shape).copy()
    c0 = 1540.0
    dt = float(np.float32(cfl * dx / c0))
    dx = float(np.float32(dx))
    inp = SimulationInputs(ndim=ndim, nx=shape[0], ny=shape[1], nz=shape[2] if ndim == 3 else 1, nt=nt,
                           dx=dx, dy=dx, dz=dx if ndim == 3 else 0.0, dt=dt, c0=float(np.float32(c0)),
                           isotropic=isotropic, modt=modt)
    inp.maps["K"] = (rho * c**2).astype(np.float32)
    inp.maps["rho"] = rho.astype(np.float32)
    vax = ["u"] if isotropic else ["u", "w", "v"][:ndim]
    pax = ["x"] if isotropic else ["x", "y", "z"][:ndim]
    if a_exp is not None:
        inp.exponential = True
        inp.maps["a_exp"] = np.broadcast_to(np.asarray(a_exp, np.float32), shape).copy()
        vax, pax = [], []
    for side, axes in (("v", vax), ("p", pax)):
        rs = (relax or {}).get(side, {"kappa": 1.0, "d": (0.0, 0.0), "alpha": (1e6, 2e6)})
        for i, ax in enumerate(axes):
            r = rs[i] if isinstance(rs, (list, tuple)) else rs
            inp.maps[f"kappa{ax}"] = np.full(shape, r["kappa"], np.float32)
            for nu in range(2):
                a, b = calc_a_and_b(r["d"][nu], r["kappa"], r["alpha"][nu], dt)
                inp.maps[f"apml{ax}{nu + 1}"] = np.full(shape, a, np.float32)
                inp.maps[f"bpml{ax}{nu + 1}"] = np.full(shape, b, np.float32)
    inp.d = coefficients.d_matrix(cfl, is_3d=ndim == 3).astype(np.float32)
    inp.dmap = coefficients.d_map(c, dt, dx, is_3d=ndim == 3).astype(np.float32)
    inp.dcmap = coefficients.dc_map(c).astype(np.int32)
    inp.icc = np.zeros((0, ndim), np.int32) if icc is None else np.asarray(icc, np.int32).reshape(-1, ndim)
    ntic = 0 if icmat is None else np.asarray(icmat).shape[1]
    inp.icmat = np.zeros((inp.icc.shape[0], ntic), np.float32) if icmat is None else np.asarray(icmat, np.float32)
    inp.outc = np.zeros((0, ndim), np.int32) if outc is None else np.asarray(outc, np.int32).reshape(-1, ndim)
    inp.icczero = np.zeros((0, ndim), np.int32) if icczero is None else np.asarray(icczero, np.int32).reshape(-1, ndim)
    inp.validate()
    return inp


def rel_l2(a, b) -> float:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def distinct_axis_maps(inp):
    """A copy of anisotropic inputs whose relaxation maps differ from axis to axis: axis 1 (maps
    w, y) gets axis 0's maps mirrored along x, axis 2 (v, z) axis 0's maps mirrored along y. The
    mirrored maps are valid relaxation maps (kappa >= 1, 0 < b <= 1) that differ wherever the
    medium is not symmetric, so a kernel reading one axis's maps for another changes its
    result."""
    import dataclasses

    assert not inp.isotropic and not getattr(inp, "exponential", False)
    maps = dict(inp.maps)
    for a, (va, pa) in enumerate((("w", "y"), ("v", "z"))[: inp.ndim - 1], start=1):
        flip = 0 if a == 1 else 1
        for stem in ["kappa"] + [f"{ab}pml" for ab in "ab"]:
            for side0, side in (("u", va), ("x", pa)):
                for nu in ([""] if stem == "kappa" else ["1", "2"]):
                    maps[f"{stem}{side}{nu}"] = np.ascontiguousarray(np.flip(inp.maps[f"{stem}{side0}{nu}"], flip))
    return dataclasses.replace(inp, maps=maps)
