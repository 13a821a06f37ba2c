"""Lateral boundary inflow on regional domains, with the rate on the device.

Mirrors the reference's orbit/core/lateral_bc.py: the open boundary lets air
leave through outward faces but brings nothing in through inward ones, so a
boundary concentration c_bc enters as a solution-independent source
s_bc = r_in * c_bc on the orbit right-hand side (the channel of the FCT
correction), and c_bc is the FCT ghost beyond the edge (orbit/cli.py:1202-1241).

* ``compute_lateral_boundary_inflow(grid)`` -> r_in (N,), computed by
  ``orbit_lateral_inflow`` (reference lateral_bc.py:58-92).
* ``load_lateral_bc(path, grids, species_names)`` reads the monthly NPZ
  (lat, lon, nz, species, c_bc (n_s, 8, nz, ny, nx)) with the reference's
  checks (lateral_bc.py:95-137); ``from_env`` honours ORBIT_LATERAL_BC;
  ``merge_into`` adds the source to an extra-RHS dict.
* ``batch_inputs(src, species, e)`` gives the batched forward driver its
  inputs: e + s_bc per tracer and the per-tracer ghosts
  (``drivers.forward_fct(..., ghost=...)``).
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Dict, List, Optional, Sequence

import numpy as np

from ._lib import check, lib

ENV_VAR = "ORBIT_LATERAL_BC"


def compute_lateral_boundary_inflow(grid, ctx=None) -> np.ndarray:
    """Inflow rate r_in >= 0 (1/s) per cell, shape (N,); zero off the boundary."""
    from .assembly import grid_fields
    from .core import default_context
    ctx = ctx or default_context()
    keep: List = []
    f = grid_fields(grid, keep)
    out = np.empty(f.nz * f.ny * f.nx)
    check(lib().orbit_lateral_inflow(ctx.handle, C.byref(f), out.ctypes.data), allow_warn=False)
    return out


class LateralBCSource(dict):
    """{species index: [r_in * c_bc per bin]} (the RHS source), with
    ``c_bc`` = {species index: [c_bc per bin, flat (N,)]} kept for the
    horizontal FCT's ghost beyond the edge."""

    def __init__(self, *a, **k):
        super().__init__(*a, **k)
        self.c_bc: Dict[int, List[np.ndarray]] = {}


def load_lateral_bc(path, grids: Sequence, species_names: Sequence[str], n_bins: int = 8,
                    ctx=None) -> LateralBCSource:
    """Read a boundary file and return {species_idx: [s_bc per bin (N,)]}.
    The file's lat / lon / nz must match the run grid, c_bc must be finite and
    non-negative and name known species (ValueError otherwise)."""
    g0 = grids[0]
    with np.load(path, allow_pickle=False) as d:
        lat, lon, nz = d["lat"], d["lon"], int(d["nz"])
        c_bc = d["c_bc"]
        names = [str(s) for s in d["species"]]
    glat, glon = np.asarray(g0.lat), np.asarray(g0.lon)
    if (nz != g0.nz or lat.shape != (g0.ny,) or lon.shape != (g0.nx,)
            or not np.allclose(lat, glat, atol=1e-6) or not np.allclose(lon, glon, atol=1e-6)):
        raise ValueError(
            f"lateral BC grid mismatch in {path}: file nz={nz} "
            f"lat {lat[0]:.3f}..{lat[-1]:.3f} ({lat.size}) lon {lon[0]:.3f}..{lon[-1]:.3f} "
            f"({lon.size}); run nz={g0.nz} lat {glat[0]:.3f}..{glat[-1]:.3f} "
            f"({g0.ny}) lon {glon[0]:.3f}..{glon[-1]:.3f} ({g0.nx})")
    if c_bc.shape != (len(names), n_bins, g0.nz, g0.ny, g0.nx):
        raise ValueError(f"c_bc shape {c_bc.shape} != ({len(names)}, {n_bins}, "
                         f"{g0.nz}, {g0.ny}, {g0.nx})")
    if not np.all(np.isfinite(c_bc)) or np.any(c_bc < 0):
        raise ValueError(f"c_bc in {path} has non-finite or negative values")
    unknown = sorted(set(names) - set(species_names))
    if unknown:
        raise ValueError(f"unknown species in {path}: {unknown}")
    rates = [compute_lateral_boundary_inflow(grids[t], ctx) for t in range(n_bins)]
    out = LateralBCSource()
    for i, name in enumerate(names):
        s = list(species_names).index(name)
        cb = [c_bc[i, t].astype(np.float64).ravel() for t in range(n_bins)]
        out[s] = [rates[t] * cb[t] for t in range(n_bins)]
        out.c_bc[s] = cb
    return out


def from_env(month, grids, species_names, n_bins: int = 8, verbose: bool = True,
             ctx=None) -> Optional[LateralBCSource]:
    """Boundary source for ``month`` if ORBIT_LATERAL_BC is set (a path with
    ``{month:02d}``), else None."""
    tmpl = os.environ.get(ENV_VAR, "").strip()
    if not tmpl:
        return None
    path = tmpl.format(month=month)
    src = load_lateral_bc(path, grids, species_names, n_bins=n_bins, ctx=ctx)
    if verbose:
        tot = {species_names[s]: float(sum(v.sum() for v in vecs) / n_bins) for s, vecs in src.items()}
        print(f"  [lateral BC] {path}: inflow source (bin-mean sum over cells, tracer units/s) "
              + ", ".join(f"{k}={v:.3e}" for k, v in tot.items()))
    return src


def merge_into(extra_rhs, src, n_bins: int = 8):
    """Add ``src`` into an extra-RHS dict (creating it if None); returns the dict."""
    if src is None:
        return extra_rhs
    if extra_rhs is None:
        extra_rhs = {}
    for s, vecs in src.items():
        if s in extra_rhs:
            extra_rhs[s] = [extra_rhs[s][t] + vecs[t] for t in range(n_bins)]
        else:
            extra_rhs[s] = list(vecs)
    return extra_rhs


def batch_inputs(src: Optional[LateralBCSource], species: Sequence[int], e):
    """(e + s_bc, ghosts) for a batch of species: e is (nt, n_bins, N) (CUDA
    float64 tensor) in the order of ``species``; ghosts is a list of nt
    (n_bins, N) tensors, None for a species without boundary values (its FCT
    faces keep the closed edge)."""
    import torch
    if src is None:
        return e, None
    e2 = e.clone()
    ghosts: List = []
    for q, s in enumerate(species):
        if s in src:
            e2[q] += torch.from_numpy(np.stack(src[s])).to(e.device)
            ghosts.append(torch.from_numpy(np.stack(src.c_bc[s])).to(e.device))
        else:
            ghosts.append(None)
    return e2, ghosts
