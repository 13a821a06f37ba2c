"""ISORROPIA cross-coupling blocks of the iso-coupled marginal mode, on the device.

Replaces orbit/modes/iso_coupling.py.  The marginal solve linearises the
deposition of TotalNH and TotalNO3 around the baseline partitioning; with the
ISORROPIA cross partials (``isorropia.partitioning_per_bin``, include_cross_partials)
the linearisation gains four diagonal blocks per bin,

  K[TotalNH,  TotalNO3] = dk_NH  * f_nh_dno3      K[TotalNH,  pSO4] = dk_NH  * f_nh_dso4
  K[TotalNO3, TotalNH ] = dk_NO3 * f_no3_dnh      K[TotalNO3, pSO4] = dk_NO3 * f_no3_dso4

with dk the particle-minus-gas deposition rate of the species: the wet
contrast particle_wet_dep - other_gas_wet_dep on every level, plus on the
surface level the dry contrast over the layer thickness, (particle_dry_dep -
NH3_dry_dep) / Dz for TotalNH and (particle_dry_dep - HNO3_dry_dep) / Dz for
TotalNO3 (HNO3 at 3 cm/s when the grid has no such field).  One kernel
(``orbit_iso_cross_blocks_dev``) forms all four for every cell and bin.

The blocks enter the species layer's K_sources_per_bin like the chemistry
couplings (the source term e_target -= K c_source at the source's end-of-bin
state), summed into a block already present (``add_iso_blocks``, the merge of
orbit/modes/marginal.py:487-511).  The environment variable
ORBIT_ISO_DISABLE_BLOCKS (comma-separated NH_NO3, NH_SO4, NO3_NH, NO3_SO4)
zeroes blocks, as in the reference's diagnostic.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import scipy.sparse as sp

from ._lib import check, dev_f64, lib

IDX_TOTAL_NH, IDX_PSO4, IDX_TOTAL_NO3 = 2, 5, 6  # reference orbit/core/deposition.py:60-64
# the kernel's block order: (target, source, diagnostic name)
BLOCKS = ((IDX_TOTAL_NH, IDX_TOTAL_NO3, "NH_NO3"), (IDX_TOTAL_NH, IDX_PSO4, "NH_SO4"),
          (IDX_TOTAL_NO3, IDX_TOTAL_NH, "NO3_NH"), (IDX_TOTAL_NO3, IDX_PSO4, "NO3_SO4"))
PARTIALS = ("f_nh_dno3", "f_no3_dnh", "f_nh_dso4", "f_no3_dso4")

# the marginal mode's source DAG gains these edges with iso coupling on
# (reference iso_coupling.py:149-172)
ISO_SOURCES_FOR_TARGET: Dict[int, list] = {IDX_TOTAL_NH: [IDX_TOTAL_NO3], IDX_TOTAL_NO3: [IDX_TOTAL_NH]}


def disabled_blocks() -> set:
    """Block names switched off by ORBIT_ISO_DISABLE_BLOCKS."""
    raw = os.environ.get("ORBIT_ISO_DISABLE_BLOCKS", "").strip()
    return {t.strip() for t in raw.split(",") if t.strip()} if raw else set()


def block_mask(disabled: Optional[set] = None) -> int:
    off = disabled_blocks() if disabled is None else set(disabled)
    return sum(1 << q for q, (_, _, name) in enumerate(BLOCKS) if name not in off)


def merged_sources_for_target(base: Dict[int, list], iso: bool) -> Dict[int, list]:
    """The chemistry DAG plus, with iso on, NO3 -> NH, NH -> NO3 and pSO4 ->
    both (reference iso_coupling.py:175-198)."""
    out = {k: list(v) for k, v in base.items()}
    if not iso:
        return out
    for target, sources in ISO_SOURCES_FOR_TARGET.items():
        cur = out.setdefault(target, [])
        for s in list(sources) + [IDX_PSO4]:
            if s not in cur:
                cur.append(s)
    return out


def _field(g, name, optional=False):
    a = np.asarray(getattr(g, name, np.array([])), dtype=np.float64)
    if a.size == 0:
        if optional:
            return None
        raise ValueError(f"the iso cross blocks need the grid's {name}")
    return a


def iso_cross_diagonals_dev(fields: dict, partials, shape, ctx=None, disabled: Optional[set] = None):
    """The four diagonals on the device: fields {Dz, particle_dry_dep,
    NH3_dry_dep, HNO3_dry_dep (or None), particle_wet_dep, other_gas_wet_dep},
    each a (n_bins, N) float64 device tensor, and the partials f_nh_dno3,
    f_no3_dnh, f_nh_dso4, f_no3_dso4 (a dict or a sequence in that order, each
    (n_bins, N)); shape (nz, ny, nx).  Returns (4, n_bins, N) in BLOCKS order."""
    import torch
    from .core import default_context
    ctx = ctx or default_context()
    nz, ny, nx = (int(v) for v in shape)
    N = nz * ny * nx
    plist = [partials[k] for k in PARTIALS] if isinstance(partials, dict) else list(partials)
    if len(plist) != 4:
        raise ValueError("four cross partials: f_nh_dno3, f_no3_dnh, f_nh_dso4, f_no3_dso4")
    names = ("Dz", "particle_dry_dep", "NH3_dry_dep", "HNO3_dry_dep", "particle_wet_dep", "other_gas_wet_dep")
    arrs = {k: fields.get(k) for k in names}
    ref = arrs["Dz"]
    if ref is None:
        raise ValueError("fields must include Dz")
    n_bins = int(ref.shape[0])
    ts = [t for t in list(arrs.values()) + plist if t is not None]
    for t in ts:
        dev_f64(t, ctx.device, "iso field", last=N, rows=n_bins)
    ts = [t.contiguous() for t in ts]
    if len({t.stride(0) for t in ts}) != 1:
        raise ValueError("every field and partial needs the same bin stride")
    stride = ts[0].stride(0)
    out = torch.empty((4, n_bins, N), dtype=torch.float64, device=ref.device)
    cont = {k: (None if v is None else v.contiguous()) for k, v in arrs.items()}
    pl = [p.contiguous() for p in plist]
    P = (C.c_void_p * 4)(*[p.data_ptr() for p in pl])
    O = (C.c_void_p * 4)(*[out[q].data_ptr() for q in range(4)])
    dp = lambda k: None if cont[k] is None else cont[k].data_ptr()  # noqa: E731
    torch.cuda.current_stream(ref.device).synchronize()
    check(lib().orbit_iso_cross_blocks_dev(ctx.handle, n_bins, nz, ny, nx, stride, dp("Dz"), dp("particle_dry_dep"),
                                          dp("NH3_dry_dep"), dp("HNO3_dry_dep"), dp("particle_wet_dep"),
                                          dp("other_gas_wet_dep"), P, block_mask(disabled), O))
    ctx.sync()
    return out


def month_fields_dev(grids: Sequence, device) -> dict:
    """The deposition fields the blocks read, (n_bins, N) device tensors."""
    import torch

    def stack(name, optional=False):
        arrs = [_field(g, name, optional) for g in grids]
        if any(a is None for a in arrs):
            return None
        return torch.from_numpy(np.stack([a.ravel() for a in arrs])).to(device)

    return {"Dz": stack("Dz"), "particle_dry_dep": stack("particle_dry_dep"), "NH3_dry_dep": stack("NH3_dry_dep"),
            "HNO3_dry_dep": stack("HNO3_dry_dep", optional=True), "particle_wet_dep": stack("particle_wet_dep"),
            "other_gas_wet_dep": stack("other_gas_wet_dep")}


def assemble_iso_cross_blocks(grid, indexer=None, f_nh_dno3=None, f_no3_dnh=None, f_nh_dso4=None, f_no3_dso4=None,
                              ctx=None) -> Dict[Tuple[int, int], sp.csc_matrix]:
    """{(target, source): diagonal CSC} for one bin (reference
    iso_coupling.py:101-146); the grid's fields and the partials are (nz, ny, nx)."""
    import torch
    from .core import default_context
    ctx = ctx or default_context()
    dev = torch.device("cuda", ctx.device)
    shape = (int(grid.nz), int(grid.ny), int(grid.nx))
    fields = month_fields_dev([grid], dev)
    parts = [torch.from_numpy(np.asarray(p, dtype=np.float64).reshape(1, -1)).to(dev)
             for p in (f_nh_dno3, f_no3_dnh, f_nh_dso4, f_no3_dso4)]
    d = iso_cross_diagonals_dev(fields, parts, shape, ctx).cpu().numpy()
    N = int(np.prod(shape))
    return {(t, s): sp.diags(d[q, 0], 0, shape=(N, N), format="csc") for q, (t, s, _) in enumerate(BLOCKS)}


def add_iso_blocks(K_sources_per_bin: List[dict], grids: Sequence, iso_cross_3d: dict, endbin_keys: Optional[set] = None,
                   ctx=None) -> List[dict]:
    """Add the iso blocks of every bin to K_sources_per_bin (summed into a
    block already present) and their keys to endbin_keys: the merge of the
    reference's iso-coupled marginal mode (orbit/modes/marginal.py:487-511).
    iso_cross_3d: {partial: (n_bins, nz, ny, nx) or per-bin list}."""
    import torch
    from .core import default_context
    ctx = ctx or default_context()
    dev = torch.device("cuda", ctx.device)
    g0 = grids[0]
    shape = (int(g0.nz), int(g0.ny), int(g0.nx))
    N = int(np.prod(shape))
    parts = [torch.from_numpy(np.stack([np.asarray(iso_cross_3d[k][t], dtype=np.float64).ravel()
                                        for t in range(len(grids))])).to(dev) for k in PARTIALS]
    d = iso_cross_diagonals_dev(month_fields_dev(grids, dev), parts, shape, ctx).cpu().numpy()
    for tau in range(len(grids)):
        for q, (t, s, _) in enumerate(BLOCKS):
            K = sp.diags(d[q, tau], 0, shape=(N, N), format="csc")
            key = (t, s)
            K_sources_per_bin[tau][key] = K_sources_per_bin[tau][key] + K if key in K_sources_per_bin[tau] else K
            if endbin_keys is not None:
                endbin_keys.add(key)
    return K_sources_per_bin
