"""The 7-D ISORROPIA partitioning table on the device.

Replaces IsorropiaLUT (reference orbit/core/isorropia_lut.py:12-147), which
the partitioning closure queries at every cell and bin of the orbit state.
The table keeps its FP32 on-disk type in HBM (a production table is ~1.7 GB
per field, far below one MI355X's 288 GB, so it stays resident instead of
being released between queries); a query is one kernel
(orbit_lut_query_dev): multilinear interpolation with linear extrapolation,
as scipy's RegularGridInterpolator(method="linear", bounds_error=False,
fill_value=None) does for the reference, then the reference's clips.

Axes and fields use the reference archive's key names, so an archive loads
with np.load (allow_pickle=False) unchanged.
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, Optional, Sequence

import numpy as np

from . import _lib
from ._lib import check, lib

AXES = ("so4_axis", "nh_axis", "no3_axis", "ca_axis", "na_axis", "t_axis", "rh_axis")
ALL_FIELDS = ("f_nh4", "f_no3", "aerosol_water", "ph")
# per-field clamps of IsorropiaLUT.query (isorropia_lut.py:136-146)
CLAMP = {"f_nh4": (0.0, 1.0), "f_no3": (0.0, 1.0), "aerosol_water": (0.0, np.inf), "ph": (-np.inf, np.inf)}


class IsorropiaLUT:
    """Query interface for the 7-D ISORROPIA lookup table on the device.

    lut_path: an .npz archive with the reference's keys; or arrays: a dict with
    the same keys.  fields: the output fields to load (the reference's default
    loads f_nh4 and f_no3 only; query() returns NaN for the others).  The table
    is read on the first query."""

    def __init__(self, lut_path: Optional[str] = None, fields: Sequence[str] = ("f_nh4", "f_no3"),
                 arrays: Optional[Dict[str, np.ndarray]] = None, ctx=None):
        if (lut_path is None) == (arrays is None):
            raise ValueError("give exactly one of lut_path and arrays")
        for f in fields:
            if f not in ALL_FIELDS:
                raise ValueError(f"unknown field {f!r}")
        self._path = lut_path
        self._arrays = arrays
        self._fields = tuple(fields)
        self._ctx = ctx
        self.handle = None
        self.axes = None
        self.n_loads = 0

    @property
    def loaded(self) -> bool:
        return self.handle is not None

    def _load(self):
        from .core import default_context
        self._ctx = self._ctx or default_context()
        src = self._arrays if self._arrays is not None else np.load(self._path, allow_pickle=False)
        axes = [np.ascontiguousarray(src[a], dtype=np.float64) for a in AXES]
        tabs = [np.ascontiguousarray(src[f], dtype=np.float32) for f in self._fields]
        shape = tuple(len(a) for a in axes)
        for f, t in zip(self._fields, tabs):
            if t.shape != shape:
                raise ValueError(f"field {f} has shape {t.shape}, the axes give {shape}")
        lens = (C.c_int * 7)(*shape)
        axp = (C.c_void_p * 7)(*[a.ctypes.data for a in axes])
        fp = (C.c_void_p * len(tabs))(*[t.ctypes.data for t in tabs])
        h = C.c_void_p()
        check(lib().orbit_lut_create(self._ctx.handle, 7, lens, axp, len(tabs), fp, C.byref(h)), allow_warn=False)
        self.handle = h
        self.axes = tuple(axes)
        self.n_loads += 1

    def release(self):
        """Free the device table; the next query loads it again."""
        if self.handle:
            lib().orbit_lut_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass

    def query(self, total_so4, total_nh, total_no3, crustal_ca, sea_salt_na, T, RH):
        """Concentrations in ug/m3 (element mass), T in K, RH in 0-1; scalars or
        arrays/tensors broadcastable to one shape.  Returns (f_nh4, f_no3,
        aerosol_water, ph), flat: CUDA float64 tensors when the inputs are
        CUDA tensors, else numpy arrays."""
        import torch
        if self.handle is None:
            self._load()
        on_dev = any(isinstance(v, torch.Tensor) and v.is_cuda for v in
                     (total_so4, total_nh, total_no3, crustal_ca, sea_salt_na, T, RH))
        dev = torch.device("cuda", self._ctx.device)
        xs = [torch.as_tensor(v, dtype=torch.float64, device=dev) for v in
              (total_so4, total_nh, total_no3, crustal_ca, sea_salt_na, T, RH)]
        xs = [x.reshape(-1).contiguous() for x in torch.broadcast_tensors(*[x.reshape(-1) for x in xs])]
        n = xs[0].numel()
        outs = [torch.empty(n, dtype=torch.float64, device=dev) for _ in self._fields]
        lo = (C.c_double * len(self._fields))(*[CLAMP[f][0] for f in self._fields])
        hi = (C.c_double * len(self._fields))(*[CLAMP[f][1] for f in self._fields])
        cp = (C.c_void_p * 7)(*[x.data_ptr() for x in xs])
        op = (C.c_void_p * len(outs))(*[o.data_ptr() for o in outs])
        torch.cuda.current_stream(dev).synchronize()
        check(lib().orbit_lut_query_dev(self._ctx.handle, self.handle, n, cp, lo, hi, op))
        self._ctx.sync()
        got = dict(zip(self._fields, outs))
        res = [got.get(f, torch.full((n,), float("nan"), dtype=torch.float64, device=dev)) for f in ALL_FIELDS]
        return tuple(res) if on_dev else tuple(r.cpu().numpy() for r in res)
