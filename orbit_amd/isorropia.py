"""ISORROPIA partitioning update on the device: the step of the outer closure
that turns the orbit state into the gas/particle fractions the species
operators read.

Replaces partitioning_per_bin and update_grid_partitioning (reference
orbit/core/isorropia_closure.py:75-478).  For every bin and cell, from the
end-of-bin states of pSO4, TotalNH and TotalNO3 (clipped at 0) and the bin's
temperature, relative humidity, fine dust and sea salt, one kernel
(orbit_lut_partitioning_dev) queries the resident 7-D table 19 times around the
state and forms

  f_nh_eq, f_no3_eq       the equilibrium fractions (mass extraction, and the
                          deposition of full solves);
  f_nh_marg, f_no3_marg   d(f c)/dc, the mean of the +/-1 % one-sided
                          differences clipped to [0, 1] (the deposition of the
                          linearised marginal and adjoint solves);
  f_nh_dno3, f_no3_dnh,   the cross partials of the clipped marginal fractions
  f_nh_dso4, f_no3_dso4   (the iso-coupled marginal mode's K blocks);

optionally blended with the previous iteration, alpha * new + (1 - alpha) * prev.

``MonthPartitioning`` keeps a month's met fields on the device and computes the
fields straight from a device orbit tensor; ``partitioning_per_bin`` and
``update_grid_partitioning`` keep the reference's call surface (GridData-like
grids, orbits {species: [c_0, ..., c_8]}, per-bin dicts of (nz, ny, nx) arrays,
the grids' NHPartitioning / NHPartitioningEq / NO3Partitioning /
NO3PartitioningEq set by the update).
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, Optional, Sequence

import numpy as np

from ._lib import check, dev_f64, lib

IDX_TOTAL_NH, IDX_PSO4, IDX_TOTAL_NO3 = 2, 5, 6  # reference orbit/core/deposition.py:60-64
FIELDS = ("f_nh_eq", "f_nh_marg", "f_no3_eq", "f_no3_marg", "f_nh_dno3", "f_no3_dnh", "f_nh_dso4", "f_no3_dso4")


def _field_index(lut, name):
    try:
        return lut._fields.index(name)
    except ValueError:
        raise ValueError(f"the partitioning update needs the table's {name} field loaded") from None


class MonthPartitioning:
    """One month's met fields resident on the device (8 bins): T (K), RH
    (percent), fine dust and sea salt (ug/m3; a grid whose array is empty has
    none, as in the reference)."""

    def __init__(self, grids: Sequence, lut, ctx=None):
        import torch
        from .core import default_context
        self.ctx = ctx or default_context()
        self.lut = lut
        self.n_bins = len(grids)
        g0 = grids[0]
        self.shape = (int(g0.nz), int(g0.ny), int(g0.nx))
        self.N = int(np.prod(self.shape))
        dev = torch.device("cuda", self.ctx.device)

        def stack(name, required):
            arrs = [np.asarray(getattr(g, name, np.array([])), dtype=np.float64) for g in grids]
            if all(a.size == self.N for a in arrs):
                return torch.from_numpy(np.stack([a.ravel() for a in arrs])).to(dev)
            if required or any(a.size == self.N for a in arrs):
                if required:
                    raise ValueError(f"every grid needs {name} of {self.N} cells")
                # some bins with the field, some without: zeros where absent
                return torch.from_numpy(np.stack([a.ravel() if a.size == self.N else np.zeros(self.N)
                                                  for a in arrs])).to(dev)
            return None

        self.T = stack("Temperature", True)
        self.RH = stack("RH", True)
        self.dust = stack("dust_fine", False)
        self.seasalt = stack("sea_salt_fine", False)

    def compute(self, so4, nh, no3, cross: bool = True, alpha: float = 1.0, prev=None):
        """Fields (n_fields, n_bins, N) on the device from end-of-bin states
        (n_bins, N) each (rows may be strided, e.g. orbit[s, 1:]); prev: a
        (n_fields, n_bins, N) tensor of the last iteration for the blend."""
        import torch
        if self.lut.handle is None:
            self.lut._load()
        nf = 8 if cross else 4
        dev = self.T.device
        states = []
        for name, t in (("pSO4", so4), ("TotalNH", nh), ("TotalNO3", no3)):
            dev_f64(t, self.ctx.device, name, last=self.N, rows=self.n_bins)
            if t.stride(-1) != 1:
                raise ValueError(f"{name} rows must be unit stride")
            states.append(t)
        if len({t.stride(0) for t in states}) != 1:
            states = [t.contiguous() for t in states]
        out = torch.empty((nf, self.n_bins, self.N), dtype=torch.float64, device=dev)
        pp = None
        if prev is not None:
            dev_f64(prev, self.ctx.device, "prev", last=self.N)
            if tuple(prev.shape) != tuple(out.shape):
                raise ValueError(f"prev must have shape {tuple(out.shape)}")
            prev = prev.contiguous()
            pp = (C.c_void_p * nf)(*[prev[f].data_ptr() for f in range(nf)])
        sp = (C.c_void_p * 3)(*[t.data_ptr() for t in states])
        mp = (C.c_void_p * 4)(self.T.data_ptr(), self.RH.data_ptr(),
                              None if self.dust is None else self.dust.data_ptr(),
                              None if self.seasalt is None else self.seasalt.data_ptr())
        op = (C.c_void_p * nf)(*[out[f].data_ptr() for f in range(nf)])
        torch.cuda.current_stream(dev).synchronize()
        check(lib().orbit_lut_partitioning_dev(self.ctx.handle, self.lut.handle, _field_index(self.lut, "f_nh4"),
                                               _field_index(self.lut, "f_no3"), self.n_bins, self.N, sp,
                                               states[0].stride(0), mp, self.N, int(bool(cross)), float(alpha),
                                               pp, op))
        self.ctx.sync()
        return out

    def compute_from_orbits(self, orbits, cross: bool = True, alpha: float = 1.0, prev=None):
        """orbits: {species: (n_bins + 1, N) device tensor or [c_0, ..., c_8]}."""
        import torch
        dev = self.T.device

        def states(s):  # end-of-bin states c_1 .. c_n_bins
            o = orbits[s]
            if torch.is_tensor(o):
                return o[1:self.n_bins + 1]
            return torch.from_numpy(np.stack([np.asarray(o[t + 1], dtype=np.float64).ravel()
                                              for t in range(self.n_bins)])).to(dev)

        return self.compute(states(IDX_PSO4), states(IDX_TOTAL_NH), states(IDX_TOTAL_NO3), cross, alpha, prev)

    def as_per_bin(self, fields) -> Dict[int, Dict[str, np.ndarray]]:
        """{bin: {field: (nz, ny, nx) numpy}} as the reference returns it."""
        host = fields.cpu().numpy()
        return {b: {FIELDS[f]: host[f, b].reshape(self.shape) for f in range(host.shape[0])}
                for b in range(self.n_bins)}


def partitioning_per_bin(grids: Sequence, orbits, lut, include_cross_partials: bool = True,
                         ctx=None) -> Dict[int, Dict[str, np.ndarray]]:
    """{bin: {field: (nz, ny, nx)}} (reference isorropia_closure.py:75-417)."""
    mp = MonthPartitioning(grids, lut, ctx)
    return mp.as_per_bin(mp.compute_from_orbits(orbits, include_cross_partials))


def update_grid_partitioning(grids: Sequence, orbits, lut, alpha: float = 0.5, prev_per_bin=None,
                             include_cross_partials: bool = True, ctx=None) -> Dict[int, Dict[str, np.ndarray]]:
    """The update with the under-relaxation blend; sets each grid's
    NHPartitioning (marginal), NHPartitioningEq, NO3Partitioning (marginal) and
    NO3PartitioningEq and returns the blended per-bin dict (reference
    isorropia_closure.py:420-478)."""
    import torch
    mp = MonthPartitioning(grids, lut, ctx)
    nf = 8 if include_cross_partials else 4
    prev = None
    if prev_per_bin and all(b in prev_per_bin for b in range(mp.n_bins)) and 0.0 < alpha < 1.0:
        prev = torch.from_numpy(np.stack([np.stack([np.asarray(prev_per_bin[b][FIELDS[f]], dtype=np.float64).ravel()
                                                    for b in range(mp.n_bins)]) for f in range(nf)])).to(mp.T.device)
    per_bin = mp.as_per_bin(mp.compute_from_orbits(orbits, include_cross_partials, alpha, prev))
    for b, g in enumerate(grids):
        d = per_bin[b]
        g.NHPartitioning = d["f_nh_marg"]
        g.NHPartitioningEq = d["f_nh_eq"]
        g.NO3Partitioning = d["f_no3_marg"]
        g.NO3PartitioningEq = d["f_no3_eq"]
    return per_bin
