"""CPU restatement of the ISORROPIA cross-coupling blocks (TEST ORACLE).

Test infrastructure only: the product never imports this module.

  delta_k_nh                orbit/modes/iso_coupling.py:58-74 (wet contrast
                            particle_wet_dep - other_gas_wet_dep everywhere, plus
                            (particle_dry_dep - NH3_dry_dep) / Dz on level 0 where
                            Dz > 0)
  delta_k_no3               orbit/modes/iso_coupling.py:77-98 (same wet contrast,
                            dry contrast against HNO3_dry_dep; a constant 0.03 m/s
                            on level 0 when the grid has no such field,
                            orbit/core/deposition.py:226-245)
  assemble_iso_cross_blocks orbit/modes/iso_coupling.py:101-146 (four diagonal CSC
                            blocks; ORBIT_ISO_DISABLE_BLOCKS zeroes named blocks,
                            :50-55)
  merged_sources_for_target orbit/modes/iso_coupling.py:149-198
  add_iso_blocks            orbit/modes/marginal.py:487-511 (summed into a block
                            already present; every iso key is an end-of-bin key)

Pinned by tests/test_oracle_iso_coupling.py: the reference's own closed forms
(tests/test_iso_coupling_no3_wet.py: the wet contrast aloft, the surface dry
contrast over Dz, the shared wet part; tests/test_simulation_modes.py:913-997:
keys, diagonal-only CSC blocks, magnitudes, the merged DAG) restated, plus
distinct partials per block (a swapped pairing fails), the HNO3 fallback, Dz = 0
and the disabled-block switch.
"""
from __future__ import annotations

import os

import numpy as np
import scipy.sparse as sp

# orbit/core/deposition.py:57-77
IDX_TOTAL_NH, IDX_PSO4, IDX_TOTAL_NO3 = 2, 5, 6
BLOCK_NAMES = {(IDX_TOTAL_NH, IDX_TOTAL_NO3): "NH_NO3", (IDX_TOTAL_NH, IDX_PSO4): "NH_SO4",
               (IDX_TOTAL_NO3, IDX_TOTAL_NH): "NO3_NH", (IDX_TOTAL_NO3, IDX_PSO4): "NO3_SO4"}


def disabled_blocks():
    raw = os.environ.get("ORBIT_ISO_DISABLE_BLOCKS", "").strip()
    return {t.strip() for t in raw.split(",") if t.strip()} if raw else set()


def _surface_dry(grid, vd_gas):
    nz, ny, nx = grid.nz, grid.ny, grid.nx
    dz = np.asarray(grid.Dz, dtype=np.float64)[0]
    contrast = np.asarray(grid.particle_dry_dep, dtype=np.float64)[0] - vd_gas
    out = np.zeros((nz, ny, nx))
    out[0] = np.where(dz > 0, contrast / np.where(dz > 0, dz, 1.0), 0.0)
    return out


def _wet(grid):
    return np.asarray(grid.particle_wet_dep, dtype=np.float64) - np.asarray(grid.other_gas_wet_dep, dtype=np.float64)


def hno3_vd(grid):
    a = np.asarray(getattr(grid, "HNO3_dry_dep", np.array([])), dtype=np.float64)
    if a.size > 0:
        return a
    vd = np.zeros((grid.nz, grid.ny, grid.nx))
    vd[0] = 0.03
    return vd


def delta_k_nh(grid):
    return _wet(grid) + _surface_dry(grid, np.asarray(grid.NH3_dry_dep, dtype=np.float64)[0])


def delta_k_no3(grid):
    return _wet(grid) + _surface_dry(grid, hno3_vd(grid)[0])


def assemble_iso_cross_blocks(grid, f_nh_dno3, f_no3_dnh, f_nh_dso4, f_no3_dso4, disabled=None):
    off = disabled_blocks() if disabled is None else set(disabled)
    dnh = delta_k_nh(grid).ravel()
    dno3 = delta_k_no3(grid).ravel()
    diags = {(IDX_TOTAL_NH, IDX_TOTAL_NO3): dnh * np.ravel(f_nh_dno3),
             (IDX_TOTAL_NH, IDX_PSO4): dnh * np.ravel(f_nh_dso4),
             (IDX_TOTAL_NO3, IDX_TOTAL_NH): dno3 * np.ravel(f_no3_dnh),
             (IDX_TOTAL_NO3, IDX_PSO4): dno3 * np.ravel(f_no3_dso4)}
    N = dnh.size
    idx = np.arange(N)
    out = {}
    for key, d in diags.items():
        if BLOCK_NAMES[key] in off:
            d = np.zeros_like(d)
        out[key] = sp.csc_matrix((d, (idx, idx)), shape=(N, N))
    return out


def merged_sources_for_target(base, iso):
    if not iso:
        return dict(base)
    out = {k: list(v) for k, v in base.items()}
    for target, src in ((IDX_TOTAL_NH, IDX_TOTAL_NO3), (IDX_TOTAL_NO3, IDX_TOTAL_NH)):
        cur = out.setdefault(target, [])
        if src not in cur:
            cur.append(src)
        if IDX_PSO4 not in cur:
            cur.append(IDX_PSO4)
    return out


def add_iso_blocks(K_sources_per_bin, grids, iso_cross_3d, endbin_keys=None, disabled=None):
    for tau, g in enumerate(grids):
        blocks = assemble_iso_cross_blocks(g, iso_cross_3d["f_nh_dno3"][tau], iso_cross_3d["f_no3_dnh"][tau],
                                           iso_cross_3d["f_nh_dso4"][tau], iso_cross_3d["f_no3_dso4"][tau],
                                           disabled=disabled)
        for key, K in blocks.items():
            K_sources_per_bin[tau][key] = K_sources_per_bin[tau][key] + K if key in K_sources_per_bin[tau] else K
            if endbin_keys is not None:
                endbin_keys.add(key)
    return K_sources_per_bin
