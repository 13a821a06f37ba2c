"""The per-species operators of a month assembled on the device.

Replaces assemble_species_operators (reference orbit/core/operator.py:371-595)
with assemble_deposition (orbit/core/deposition.py:101-362): the shared
transport block T (orbit_transport_assemble), every species' diagonal
(deposition, SO2/NOx/O3/CO chemistry losses, VBS aging) and the diagonal
coupling blocks (orbit_species_diagonals), so a month's operator set is built
from grid fields with no host CSR (OperatorSet.from_grids).

Species indices and names are the reference's (deposition.py:55-73).  The
partitioning choice and its fallbacks (deposition.py:180-223) and the
environment switches of the reference (ORBIT_VBS_A_PHOTO, ORBIT_VBS_K_AGE,
ORBIT_VBS_AGING_GAIN; ORBIT_VBS_FRAG refused) are read here.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import scipy.sparse as sp

from ._lib import check, lib

IDX_VBS_C100, IDX_PM25, IDX_TOTAL_NH, IDX_SO2, IDX_NOX, IDX_PSO4, IDX_TOTAL_NO3 = 0, 1, 2, 3, 4, 5, 6
IDX_O3, IDX_CO, IDX_VBS_C10, IDX_VBS_C1, IDX_VBS_C01, IDX_VBS_C1000, IDX_POA = 7, 8, 9, 10, 11, 12, 13
N_SPECIES = 14
# order of the coupling diagonals returned by the device: (target, source)
COUPLINGS = ((IDX_PSO4, IDX_SO2), (IDX_TOTAL_NO3, IDX_NOX), (IDX_VBS_C100, IDX_VBS_C1000),
             (IDX_VBS_C10, IDX_VBS_C100), (IDX_VBS_C1, IDX_VBS_C10), (IDX_VBS_C01, IDX_VBS_C1))

_PTRS = ("particle_dry_dep", "SO2_dry_dep", "NOx_dry_dep", "NH3_dry_dep", "VOC_dry_dep", "HNO3_dry_dep",
         "O3_dry_dep", "is_land", "particle_wet_dep", "SO2_wet_dep", "other_gas_wet_dep", "nh_part", "no3_part",
         "F_p_vbs", "F_p_vbs_surface", "j_no2_soa")
_RATES = ("so2_ox", "nox_to_no3", "o3_loss", "co_loss", "archive_OH")


class SpeciesFields(C.Structure):
    _fields_ = ([(a, C.c_void_p) for a in _PTRS] + [("a_photo", C.c_double)]
                + [(a, C.c_void_p) for a in _RATES] + [("k_age", C.c_double), ("aging_gain", C.c_double)])


def _arr(g, name):
    v = getattr(g, name, None)
    return None if v is None or np.size(v) == 0 else v


def _partitions(g, partitioning: str):
    """TotalNH and TotalNO3 particle fractions (deposition.py:180-223)."""
    if partitioning not in ("marginal", "equilibrium"):
        raise ValueError(f"partitioning must be one of ('marginal', 'equilibrium'), got {partitioning!r}")
    nh_eq = _arr(g, "NHPartitioningEq")
    nh = nh_eq if partitioning == "equilibrium" and nh_eq is not None else _arr(g, "NHPartitioning")
    marg, eq = _arr(g, "NO3Partitioning"), _arr(g, "NO3PartitioningEq")
    no3 = next((f for f in ((marg, eq) if partitioning == "marginal" else (eq, marg)) if f is not None), None)
    return nh, no3


def species_diagonals(g, partitioning: str = "marginal", so2_ox_rate=None, nox_to_no3_rate=None,
                      o3_loss_rate=None, co_loss_rate=None, a_photo: Optional[float] = None,
                      k_age: Optional[float] = None, aging_gain: Optional[float] = None,
                      ctx=None) -> Tuple[np.ndarray, Dict[Tuple[int, int], np.ndarray]]:
    """(d_species (14, N), {(target, source): (N,) diagonal of K}) for one bin,
    computed on the device."""
    from .core import default_context
    ctx = ctx or default_context()
    nz, ny, nx = int(g.nz), int(g.ny), int(g.nx)
    N = nz * ny * nx
    nh, no3 = _partitions(g, partitioning)
    if nh is None:
        raise ValueError("the grid has no NHPartitioning field")
    oh = _arr(g, "archive_OH")
    if nox_to_no3_rate is None and oh is not None:
        raise ValueError("pass nox_to_no3_rate: the grid carries archive oxidants, from which the reference "
                         "derives it (orbit/core/nox_to_no3_rate.py), chemistry outside the transport path")
    if oh is not None and "ORBIT_VBS_FRAG" in os.environ:
        raise ValueError("ORBIT_VBS_FRAG was removed in 1.0.3; set ORBIT_VBS_AGING_GAIN (default 1.075) instead.")
    a_photo = float(os.environ.get("ORBIT_VBS_A_PHOTO", 0.0)) if a_photo is None else float(a_photo)
    k_age = float(os.environ.get("ORBIT_VBS_K_AGE", "4.0e-11")) if k_age is None else float(k_age)
    aging_gain = float(os.environ.get("ORBIT_VBS_AGING_GAIN", "1.075")) if aging_gain is None else float(aging_gain)
    so2 = so2_ox_rate if so2_ox_rate is not None else _arr(g, "SO2oxidation")
    if so2 is None:
        raise ValueError("no SO2 oxidation rate: pass so2_ox_rate or give the grid SO2oxidation")
    keep: List[np.ndarray] = []

    def p(a):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype=np.float64)
        keep.append(a)
        return a.ctypes.data

    f = SpeciesFields()
    Fv, Fs = _arr(g, "F_p_vbs"), _arr(g, "F_p_vbs_surface")
    vals = {n: _arr(g, n) for n in ("particle_dry_dep", "SO2_dry_dep", "NOx_dry_dep", "NH3_dry_dep",
                                    "VOC_dry_dep", "HNO3_dry_dep", "O3_dry_dep", "is_land", "particle_wet_dep",
                                    "SO2_wet_dep", "other_gas_wet_dep", "j_no2_soa")}
    vals.update(nh_part=nh, no3_part=no3,
                F_p_vbs=Fv if Fv is not None and np.shape(Fv)[0] == 5 else None,
                F_p_vbs_surface=Fs if Fs is not None and np.shape(Fs)[0] == 5 else None)
    for name in _PTRS:
        setattr(f, name, p(vals[name]))
    f.a_photo = a_photo
    f.so2_ox, f.nox_to_no3, f.o3_loss, f.co_loss, f.archive_OH = (
        p(so2), p(nox_to_no3_rate), p(o3_loss_rate), p(co_loss_rate), p(oh))
    f.k_age, f.aging_gain = k_age, aging_gain
    Dz = np.ascontiguousarray(g.Dz, dtype=np.float64)
    d = np.empty((N_SPECIES, N))
    K6 = np.empty((len(COUPLINGS), N))
    check(lib().orbit_species_diagonals(ctx.handle, nz, ny, nx, Dz.ctypes.data, C.byref(f), d.ctypes.data,
                                        K6.ctypes.data), allow_warn=False)
    K = {COUPLINGS[0]: K6[0]}
    if nox_to_no3_rate is not None:
        K[COUPLINGS[1]] = K6[1]
    if oh is not None:
        for q in range(2, 6):
            K[COUPLINGS[q]] = K6[q]
    return d, K


def assemble_species_operators(g, indexer=None, verbose: bool = False, scheme: str = "exp",
                               nox_to_no3_rate=None, so2_ox_rate=None, o3_loss_rate=None, co_loss_rate=None,
                               return_D_per_species: bool = False, partitioning: str = "marginal", ctx=None):
    """The reference's assemble_species_operators (orbit/core/operator.py:371-595)
    computed on the device: returns (L_species [14 CSC], K_sources {(t, s): CSC},
    T CSC, d_species (14, N)); with return_D_per_species the deposition part is
    not separated here and D is returned as the diagonals themselves."""
    from .assembly import assemble_transport_block
    if scheme != "exp":
        raise ValueError(f"Only scheme='exp' is supported (got {scheme!r})")
    T = assemble_transport_block(g, ctx)
    d, Kd = species_diagonals(g, partitioning, so2_ox_rate, nox_to_no3_rate, o3_loss_rate, co_loss_rate,
                              ctx=ctx)
    L = [(T + sp.diags(d[s])).tocsc() for s in range(N_SPECIES)]
    K = {key: sp.diags(v).tocsc() for key, v in Kd.items()}
    if verbose:
        print(f"{N_SPECIES} species operators on the device, couplings = {len(K)}")
    if return_D_per_species:
        return L, K, T, d, [sp.diags(v).tocsc() for v in d]
    return L, K, T, d


def month_diagonals(grids: Sequence, species: Optional[Sequence[int]] = None, rates: Optional[Sequence[dict]] = None,
                    partitioning: str = "marginal", ctx=None):
    """D (n_bins, len(species), N) and the per-bin coupling diagonals for a month,
    for OperatorSet.from_grids(grids, D).  rates[bin]: the keyword rates of
    species_diagonals for that bin."""
    species = list(range(N_SPECIES)) if species is None else list(species)
    Ds, Ks = [], []
    for b, g in enumerate(grids):
        d, K = species_diagonals(g, partitioning, ctx=ctx, **((rates[b] if rates else None) or {}))
        Ds.append(d[species])
        Ks.append(K)
    return np.stack(Ds), Ks
