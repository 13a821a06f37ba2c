"""Regional lateral boundary: inflow rates on the device, boundary-value files.

An open lateral edge lets air leave through outward faces but, in the
transport operator, brings nothing in through inward ones.  The missing inflow
is a source that does not depend on the solution, so it rides on the orbit
right-hand side: s_bc = r_in * c_bc per cell and bin, where r_in is the
upwind rate through the cell's inward boundary faces (the device kernel
``orbit_lateral_inflow``; reference orbit/core/lateral_bc.py:58-92) and c_bc
the concentration outside.  The same c_bc is the far-upwind value the FCT
source reads beyond the edge (reference orbit/cli.py:1202-1241).

Data format (one file per month, the reference's NPZ, lateral_bc.py:33-36):
``lat`` (ny,), ``lon`` (nx,), ``nz`` (scalar), ``species`` (n_s,) names and
``c_bc`` (n_s, n_bins, nz, ny, nx) in each species' tracer units.  Species
without an entry get no inflow.

Here a boundary source is kept as one (n_bins, N) array per species, the
rates of all bins are computed once per file, and ``batch_inputs`` turns a
source into the batched device inputs of ``drivers.forward_fct``.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from ._lib import check, lib

ENV_VAR = "ORBIT_LATERAL_BC"  # reference lateral_bc.py:38 (the switch keeps its name)


class BoundaryFileError(ValueError):
    """A boundary-value file that cannot be used with the run's grid."""


def compute_lateral_boundary_inflow(grid, ctx=None) -> np.ndarray:
    """r_in (1/s) of every cell of one bin, k-major (N,), zero off the edges;
    computed on the device from the bin's grid fields."""
    from .assembly import grid_fields
    from .core import default_context
    ctx = ctx or default_context()
    keep: List = []
    fields = grid_fields(grid, keep)
    rate = np.empty(fields.nz * fields.ny * fields.nx)
    check(lib().orbit_lateral_inflow(ctx.handle, C.byref(fields), rate.ctypes.data), allow_warn=False)
    return rate


def inflow_rates(grids: Sequence, ctx=None) -> np.ndarray:
    """(n_bins, N) inflow rates of every bin of a month."""
    return np.stack([compute_lateral_boundary_inflow(g, ctx) for g in grids])


class LateralBCSource(dict):
    """species index -> (n_bins, N) boundary source r_in * c_bc (row t is bin t),
    with ``c_bc``: species index -> (n_bins, N) boundary values (the FCT ghost)."""

    def __init__(self, *a, **k):
        super().__init__(*a, **k)
        self.c_bc: Dict[int, np.ndarray] = {}


def _read(path) -> Tuple[np.ndarray, np.ndarray, int, List[str], np.ndarray]:
    with np.load(path, allow_pickle=False) as f:
        missing = [k for k in ("lat", "lon", "nz", "species", "c_bc") if k not in f.files]
        if missing:
            raise BoundaryFileError(f"{path}: no {', '.join(missing)} in the boundary file")
        return (np.asarray(f["lat"], dtype=np.float64), np.asarray(f["lon"], dtype=np.float64),
                int(f["nz"]), [str(s) for s in f["species"]], np.asarray(f["c_bc"]))


def _same_axis(file_axis: np.ndarray, run_axis) -> bool:
    run_axis = np.asarray(run_axis, dtype=np.float64)
    return file_axis.shape == run_axis.shape and bool(np.all(np.abs(file_axis - run_axis) <= 1e-6))


def load_lateral_bc(path, grids: Sequence, species_names: Sequence[str], n_bins: int = 8,
                    ctx=None) -> LateralBCSource:
    """Boundary source of one month from a boundary-value file.

    Every check runs before any device work; a failed one raises
    BoundaryFileError (a ValueError): a species name the run does not carry,
    levels or horizontal coordinates (to 1e-6 degrees) other than the run
    grid's, a c_bc array of another layout, or a value that is negative or
    not finite."""
    lat, lon, nz, names, c_bc = _read(path)
    g0 = grids[0]
    run_names = list(species_names)
    strangers = [n for n in names if n not in run_names]
    if strangers:
        raise BoundaryFileError(f"{path}: species {strangers} are not among the run's {run_names}")
    if nz != g0.nz or not _same_axis(lat, g0.lat) or not _same_axis(lon, g0.lon):
        raise BoundaryFileError(
            f"{path}: boundary values are on another grid ({nz} levels x {lat.size} latitudes "
            f"[{lat.min():.3f}, {lat.max():.3f}] x {lon.size} longitudes [{lon.min():.3f}, {lon.max():.3f}]) "
            f"than the run ({g0.nz} x {g0.ny} x {g0.nx})")
    want = (len(names), n_bins, g0.nz, g0.ny, g0.nx)
    if c_bc.shape != want:
        raise BoundaryFileError(f"{path}: c_bc has layout {c_bc.shape}, the run needs {want} "
                                "(species, bins, levels, latitudes, longitudes)")
    c_bc = c_bc.astype(np.float64)
    bad = ~np.isfinite(c_bc) | (c_bc < 0)
    if bad.any():
        where = np.argwhere(bad)[0]
        raise BoundaryFileError(f"{path}: c_bc must be finite and >= 0; {int(bad.sum())} entries are not, "
                                f"the first at (species {where[0]}, bin {where[1]})")
    r_in = inflow_rates(grids[:n_bins], ctx)
    src = LateralBCSource()
    for q, name in enumerate(names):
        s = run_names.index(name)
        values = c_bc[q].reshape(n_bins, -1)
        src[s] = r_in * values
        src.c_bc[s] = values
    return src


def from_env(month, grids, species_names, n_bins: int = 8, verbose: bool = True,
             ctx=None) -> Optional[LateralBCSource]:
    """The month's boundary source when ORBIT_LATERAL_BC names a file pattern
    (``{month:02d}`` in it is the month), else None."""
    pattern = os.environ.get(ENV_VAR, "").strip()
    if not pattern:
        return None
    path = pattern.format(month=month)
    src = load_lateral_bc(path, grids, species_names, n_bins=n_bins, ctx=ctx)
    if verbose:
        parts = [f"{species_names[s]} {float(v.sum(axis=1).mean()):.3e}" for s, v in sorted(src.items())]
        print(f"boundary inflow from {path}, bin-mean total per species (tracer units/s): " + "; ".join(parts))
    return src


def merge_into(extra_rhs, src, n_bins: int = 8):
    """extra_rhs (species -> per-bin sources, or None) plus src; a new dict when
    extra_rhs is None, else extra_rhs updated in place.  Rows index bins."""
    if src is None:
        return extra_rhs
    merged = {} if extra_rhs is None else extra_rhs
    for s, rows in src.items():
        rows = np.asarray(rows)[:n_bins]
        merged[s] = rows.copy() if s not in merged else np.asarray(merged[s])[:n_bins] + rows
    return merged


def batch_inputs(src: Optional[LateralBCSource], species: Sequence[int], e):
    """(e + s_bc, ghosts) for a batch of species: ``e`` (nt, n_bins, N) CUDA
    float64 in the order of ``species``; ``ghosts`` has one (n_bins, N) device
    tensor per species, None where the species has no boundary values (its
    FCT edge faces stay closed); (e, None) when there is no source."""
    import torch
    if src is None:
        return e, None
    forced = e.clone()
    ghosts: List = []
    for q, s in enumerate(species):
        if s not in src:
            ghosts.append(None)
            continue
        forced[q] += torch.as_tensor(np.asarray(src[s]), device=e.device)
        ghosts.append(torch.as_tensor(np.ascontiguousarray(src.c_bc[s]), device=e.device))
    return forced, ghosts
