"""Engine inputs for large grids, built plane by plane.

:func:`build_inputs_planewise` returns exactly what ``Solver(...).build_inputs()`` returns for a
:class:`~fullwave25_amd.medium.Medium` with the relaxation model and the absorbing layer on
(every float32 map byte for byte; ``tests/test_fast_inputs.py``), without ever holding a
float64 copy of a padded map. The padded-grid pipeline of the reference
(fullwave/solver/pml_builder.py: pad, look up the relaxation parameters, ramp every map along
x, then y, then z, recompute a and b) is a per-voxel function of

* the (alpha0, gamma) pair of the voxel the padding replicates, and
* for each axis, where the voxel sits in the absorbing layer,

because each axis' ramp reads only the plane where the ramp starts. So a voxel's value is: the
looked-up parameter of the nearest interior voxel, then each axis' ramp operation in turn
(identity, ``v - f (v - target)`` with the value of that axis' start plane, or the target).
This module applies the same operations in the same order and precision on one x plane at a
time and writes float32 directly: memory is the float32 output plus a few planes.

The reference has no counterpart; it builds every padded map whole.
"""

from __future__ import annotations

import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from ..grid import Grid
from ..medium import Medium, calc_a_and_b, fw2_map_names, sort_mechanisms
from ..relaxation import fit_band_for_grid, generate_relaxation_params
from ..utils.coordinates import map_to_coords
from ..utils.numerical import matlab_round
from . import coefficients
from .inputs import SimulationInputs, engine_file_stem
from .pml_builder import _ramp

_F32 = np.float32

# ramp modes of one coordinate along one axis
_ID, _RAMP, _TARGET = 0, 1, 2


def ramp_ops(n: int, m: int, n_pml: int, n_tr: int, *, kind: str, order: float = 2, region: str = "both"):
    """Per-coordinate ``(mode, src, f)`` of ``apply_transition_and_pml`` along an axis of length
    ``n``: identity, ``v[src] - f (v[src] - target)``, or the target."""
    if region == "transition":
        thick, off = n_tr, n_pml
    elif region == "pml":
        thick, off = n_pml, 0
    else:
        thick, off = n_pml + n_tr, 0
    f = _ramp(kind, thick, order)
    mode = np.full(n, _ID, np.int8)
    src = np.arange(n)
    fac = np.zeros(n)
    mode[: m + off + thick] = _TARGET
    mode[n - m - thick - off:] = _TARGET
    lo = np.arange(m + off - 1, m + off + thick)
    hi_start = n - m - thick - off - 1
    hi = np.arange(hi_start, n - m - off)
    if lo.size and lo[-1] >= hi[0]:
        raise ValueError("grid too small for its absorbing layers")
    mode[lo], src[lo], fac[lo] = _RAMP, m + off + thick, f[::-1]
    mode[hi], src[hi], fac[hi] = _RAMP, hi_start, f
    return mode, src, fac


def _apply(values, mode, src, fac, target, axis):
    """One axis' ramp on an array whose ``axis`` is the ramped one (float64, new array)."""
    base = np.take(values, src, axis=axis)
    shape = [1] * values.ndim
    shape[axis] = -1
    f = fac.reshape(shape)
    out = np.where((mode == _RAMP).reshape(shape), base - f * (base - target), base)
    return np.where((mode == _TARGET).reshape(shape), target, out)


def build_inputs_planewise(grid: Grid, medium: Medium, source, sensor, *, n_pml: int, n_transition: int,
                           m_spatial_order: int = 8, isotropic: bool = True) -> SimulationInputs:
    """Same result as ``Solver(work_dir, grid, medium, source, sensor,
    pml_layer_thickness_px=n_pml, n_transition_layer=n_transition,
    use_isotropic_relaxation=isotropic).build_inputs()[0]``, built one x plane at a time."""
    if not isinstance(medium, Medium):
        raise TypeError("the plane-wise builder takes a Medium (relaxation model)")
    if n_pml == 0:
        n_transition = 0
    m = m_spatial_order
    nb = n_transition + n_pml + m
    is_3d = grid.is_3d
    shape = tuple(np.asarray(medium.sound_speed).shape)
    spacing = (grid.dx, grid.dy, grid.dz)
    egrid = Grid(domain_size=tuple((shape[a] + 2 * nb) * spacing[a] for a in range(len(shape))), f0=grid.f0,
                 duration=grid.duration, c0=grid.c0, ppw=grid.ppw, cfl=grid.cfl)
    eshape = tuple(s + 2 * nb for s in shape)
    n_relax = medium.n_relaxation_mechanisms
    if n_relax != 2:
        raise ValueError("only two relaxation mechanisms are supported")

    c = np.asarray(medium.sound_speed, dtype=np.float64)
    rho = np.asarray(medium.density, dtype=np.float64)
    a0 = np.asarray(medium.alpha_coeff, dtype=np.float64)
    gm = np.asarray(medium.alpha_power, dtype=np.float64)

    band = fit_band_for_grid(egrid.dx, egrid.c0)
    db = medium.path_relaxation_parameters_database

    d_target = -(2 + 1) * egrid.c0 * np.log(10 ** (-30)) / (2 * (egrid.dx * n_pml + egrid.dx * n_transition))
    kinds = {}
    for side in ("x1", "x2"):
        kinds[f"kappa_{side}"] = None
        for nu in (1, 2):
            if nu >= 2:
                kinds[f"d_{side}_nu{nu}"] = ("cosine", 2, "transition", 0.0)
                kinds[f"alpha_{side}_nu{nu}"] = ("cosine", 2, "transition", 0.0)
            else:
                kinds[f"d_{side}_nu{nu}"] = ("polynomial", 2, "both", d_target)
                kinds[f"alpha_{side}_nu{nu}"] = ("linear", 2, "both", 0.0)
    ops = {}
    for key, kd in kinds.items():
        if kd is None or n_pml == 0:
            continue
        kind, order, region, _ = kd
        ops[key] = [ramp_ops(eshape[a], m, n_pml, n_transition, kind=kind, order=order, region=region)
                    for a in range(len(shape))]

    names = fw2_map_names(is_3d=is_3d, isotropic=isotropic, n_relax=n_relax)
    side_of = {"u": "x1", "w": "x1", "v": "x1", "x": "x2", "y": "x2", "z": "x2"}
    out = {"K": np.empty(eshape, _F32), "rho": np.empty(eshape, _F32)}
    for nm in names:
        out[engine_file_stem(nm)] = np.empty(eshape, _F32)
    dcmap = np.empty(eshape, np.int32)
    c_min_round = matlab_round(np.asarray(c.min()))
    nd = len(shape)

    # interior index each padded coordinate replicates, per axis
    clip = [np.clip(np.arange(eshape[a]) - nb, 0, shape[a] - 1) for a in range(nd)]
    tail = tuple(np.ix_(*clip[1:]))
    keys = list(kinds)
    workers = max(1, min(16, os.cpu_count() or 1))

    # every distinct (alpha0, gamma) pair, its sorted parameters and its a/b (one look-up each)
    def plane_pairs(k):
        return np.unique((a0[k] + 1j * gm[k]).reshape(-1))

    with ThreadPoolExecutor(workers) as ex:
        pairs = np.unique(np.concatenate(list(ex.map(plane_pairs, range(shape[0])))))
    params = generate_relaxation_params(pairs.real, pairs.imag, n_relax, db, dt=egrid.dt, band=band)
    table = sort_mechanisms(params, n_relax)
    ab_table = {(side, nu): calc_a_and_b(table[f"d_{side}_nu{nu}"], table[f"kappa_{side}"],
                                         table[f"alpha_{side}_nu{nu}"], egrid.dt)
                for side in ("x1", "x2") for nu in (1, 2)}

    def labels(k):
        return np.searchsorted(pairs, (a0[k] + 1j * gm[k])[tail])

    # rows/columns of a plane that some ramp touches (the rest of the plane is interior)
    edge = []
    for a in range(1, nd):
        touched = np.zeros(eshape[a], bool)
        for key in keys:
            if key in ops:
                touched |= ops[key][a][0] != _ID
        edge.append(np.nonzero(touched)[0])
    x_touched = np.zeros(eshape[0], bool)
    for key in keys:
        if key in ops:
            x_touched |= ops[key][0][0] != _ID

    def ramp_axis(v, key, axis):
        """One in-plane axis' ramp (axis 1 = y, 2 = z of the grid), on the touched lines only."""
        mode, src, fac = ops[key][axis]
        target = kinds[key][3]
        idx = edge[axis - 1]
        base = np.take(v, src[idx], axis=axis - 1)
        shp = [1] * v.ndim
        shp[axis - 1] = -1
        f = fac[idx].reshape(shp)
        m_ = mode[idx].reshape(shp)
        new = np.where(m_ == _RAMP, base - f * (base - target), base)
        new = np.where(m_ == _TARGET, target, new)
        out_v = v.copy()
        if axis == 1:
            out_v[idx] = new
        else:
            out_v[:, idx] = new
        return out_v

    def do_plane(i):
        ii = clip[0][i]
        lab_i = labels(ii)
        if x_touched[i] and n_pml:
            # every key: x ramp of the layer's start plane, then y and z
            start = {}
            ramped = {}
            for key in keys:
                kd = kinds[key]
                if kd is None:
                    ramped[key] = table[key][lab_i]
                    continue
                mx, sx_, fx = ops[key][0]
                if mx[i] == _TARGET:
                    v = np.full(eshape[1:], kd[3])
                else:
                    k = clip[0][sx_[i]]
                    if k not in start:
                        start[k] = labels(k)
                    v = table[key][start[k]]
                    if mx[i] == _RAMP:
                        v = v - fx[i] * (v - kd[3])
                for a in range(1, nd):
                    v = ramp_axis(v, key, a)
                ramped[key] = v
            ab = {sn: calc_a_and_b(ramped[f"d_{sn[0]}_nu{sn[1]}"], ramped[f"kappa_{sn[0]}"],
                                   ramped[f"alpha_{sn[0]}_nu{sn[1]}"], egrid.dt) for sn in ab_table}
        else:
            # interior plane: per-pair values, ramps and a/b only on the touched rows/columns
            ramped = {key: table[key][lab_i] for key in keys}
            if n_pml:
                for key in keys:
                    if kinds[key] is not None:
                        for a in range(1, nd):
                            ramped[key] = ramp_axis(ramped[key], key, a)
            ab = {}
            for sn, (ta, tb) in ab_table.items():
                a_ = ta[lab_i]
                b_ = tb[lab_i]
                if n_pml:
                    for a in range(1, nd):
                        idx = edge[a - 1]
                        sl = (idx,) if a == 1 else (slice(None), idx)
                        ea, eb = calc_a_and_b(ramped[f"d_{sn[0]}_nu{sn[1]}"][sl], ramped[f"kappa_{sn[0]}"][sl],
                                              ramped[f"alpha_{sn[0]}_nu{sn[1]}"][sl], egrid.dt)
                        a_[sl] = ea
                        b_[sl] = eb
                ab[sn] = (a_, b_)
        for nm in names:
            stem = engine_file_stem(nm)
            if nm.startswith("kappa_"):
                out[stem][i] = ramped[f"kappa_{side_of[nm[-1]]}"]
            else:
                axis, nu = nm[6], int(nm[7:])
                a_, b_ = ab[(side_of[axis], nu)]
                out[stem][i] = a_ if nm[0] == "a" else b_
        cp = c[ii][tail]
        out["K"][i] = np.multiply(cp**2, rho[ii][tail])
        out["rho"][i] = rho[ii][tail]
        dcmap[i] = matlab_round(cp) - c_min_round

    with ThreadPoolExecutor(workers) as ex:
        for _ in ex.map(do_plane, range(eshape[0])):
            pass

    inp = SimulationInputs(
        ndim=nd, nx=eshape[0], ny=eshape[1], nz=eshape[2] if is_3d else 1, nt=egrid.nt,
        dx=float(_F32(egrid.dx)), dy=float(_F32(egrid.dy)), dz=float(_F32(egrid.dz)) if is_3d else 0.0,
        dt=float(_F32(egrid.dt)), c0=float(_F32(egrid.c0)), isotropic=isotropic,
        modt=int(sensor.sampling_modulus_time), n_relax=n_relax,
    )
    inp.maps.update(out)
    inp.d = coefficients.d_matrix(egrid.cfl, is_3d=is_3d).astype(_F32)
    inp.dmap = coefficients.d_map(c, egrid.dt, egrid.dx, is_3d=is_3d).astype(_F32)
    if coefficients.ndmap_value(c) == 1:
        inp.dmap = inp.dmap[:, :, :1]
    inp.dcmap = dcmap
    inp.icc = (np.asarray(source.incoords) + nb).astype(np.int32)
    inp.icmat = np.asarray(source.icmat).astype(_F32)
    inp.outc = (np.asarray(sensor.outcoords) + nb).astype(np.int32)
    inp.icczero = (map_to_coords(medium.air_map) + nb).astype(np.int32).reshape(-1, nd)
    inp.validate()
    return inp


__all__ = ["build_inputs_planewise", "ramp_ops"]
