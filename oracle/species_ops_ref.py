"""CPU oracle of the per-species operator diagonals -- TEST INFRASTRUCTURE ONLY.

Restates the deposition diagonal of every species (reference
orbit/core/deposition.py:101-362: wet rate on every level, dry velocity over
the surface layer thickness, partition-weighted gas/particle blends for
TotalNH, TotalNO3 and the VBS bins, the SOA photolytic sink) and the chemistry
parts of assemble_species_operators (orbit/core/operator.py:371-595: SO2
oxidation, NOx -> TotalNO3, O3 and CO losses, VBS aging with its gain) as
d_species (14, N) and the diagonal coupling blocks K[(target, source)] (N,).

Grid objects carry the reference's GridData names; the partitioning choice
("marginal" / "equilibrium") and its fallbacks follow deposition.py:180-223.

Pinned by tests/test_oracle_species_ops.py: golden vectors of the reference's
own assemble_species_operators (tests/golden/species_ops_golden.npz, written by
tests/golden/make_species_ops_golden.py), the blends' end points (p = 0 gives
the gas rates, p = 1 the particle rates) and the coupling conservation
(SO2's oxidation loss equals pSO4's gain; aging moves gain x the loss).
"""
from __future__ import annotations

import numpy as np

IDX_VBS_C100, IDX_PM25, IDX_TOTAL_NH, IDX_SO2, IDX_NOX, IDX_PSO4, IDX_TOTAL_NO3 = 0, 1, 2, 3, 4, 5, 6
IDX_O3, IDX_CO, IDX_VBS_C10, IDX_VBS_C1, IDX_VBS_C01, IDX_VBS_C1000, IDX_POA = 7, 8, 9, 10, 11, 12, 13
N_SPECIES = 14
IDX_VBS_BINS = (IDX_VBS_C01, IDX_VBS_C1, IDX_VBS_C10, IDX_VBS_C100, IDX_VBS_C1000)
VBS_AGING_PAIRS = ((IDX_VBS_C1000, IDX_VBS_C100), (IDX_VBS_C100, IDX_VBS_C10),
                   (IDX_VBS_C10, IDX_VBS_C1), (IDX_VBS_C1, IDX_VBS_C01))


def _a(g, name):
    v = getattr(g, name, None)
    return None if v is None or np.size(v) == 0 else np.asarray(v, dtype=np.float64)


def nh_partitioning(g, partitioning="marginal"):
    """deposition.py:180-199"""
    eq = _a(g, "NHPartitioningEq")
    if partitioning == "equilibrium" and eq is not None:
        return eq
    return _a(g, "NHPartitioning")


def no3_partitioning(g, partitioning="marginal"):
    """deposition.py:202-223"""
    marg, eq = _a(g, "NO3Partitioning"), _a(g, "NO3PartitioningEq")
    for f in ((marg, eq) if partitioning == "marginal" else (eq, marg)):
        if f is not None:
            return f
    return np.zeros((g.nz, g.ny, g.nx))


def _hno3_vd(g):
    v = _a(g, "HNO3_dry_dep")
    if v is not None:
        return v
    vd = np.zeros((g.nz, g.ny, g.nx))
    vd[0] = 0.03
    return vd


def _o3_vd(g):
    v = _a(g, "O3_dry_dep")
    if v is not None:
        return v
    vd = np.zeros((g.nz, g.ny, g.nx))
    land = _a(g, "is_land")
    land = np.ones((g.ny, g.nx)) if land is None else land
    vd[0] = np.where(land > 0, 0.004, 0.0005)
    return vd


def dry_velocity(g, s, partitioning):
    """Surface dry-deposition velocity (ny, nx), deposition.py:263-308."""
    pdd = _a(g, "particle_dry_dep")[0]
    if s in IDX_VBS_BINS:
        Fs = _a(g, "F_p_vbs_surface")
        if Fs is None or Fs.shape[0] != 5:
            return pdd
        F = Fs[IDX_VBS_BINS.index(s)]
        voc = _a(g, "VOC_dry_dep")
        voc = np.full_like(pdd, 1.0e-3) if voc is None else voc[0]
        return F * pdd + (1.0 - F) * voc
    if s in (IDX_PM25, IDX_POA, IDX_PSO4):
        return pdd
    if s == IDX_TOTAL_NH:
        p = nh_partitioning(g, partitioning)[0]
        return (1.0 - p) * _a(g, "NH3_dry_dep")[0] + p * pdd
    if s == IDX_SO2:
        return _a(g, "SO2_dry_dep")[0]
    if s == IDX_NOX:
        return _a(g, "NOx_dry_dep")[0]
    if s == IDX_TOTAL_NO3:
        p = no3_partitioning(g, partitioning)[0]
        return (1.0 - p) * _hno3_vd(g)[0] + p * pdd
    if s == IDX_O3:
        return _o3_vd(g)[0]
    return np.zeros((g.ny, g.nx))


def wet_rate(g, s, partitioning):
    """Wet-deposition rate (nz, ny, nx), deposition.py:311-362."""
    pwd = _a(g, "particle_wet_dep")
    if s in IDX_VBS_BINS:
        F = _a(g, "F_p_vbs")
        if F is None or F.shape[0] != 5:
            return pwd.copy()
        return F[IDX_VBS_BINS.index(s)] * pwd
    if s in (IDX_PM25, IDX_POA, IDX_PSO4):
        return pwd.copy()
    if s == IDX_TOTAL_NH:
        p = nh_partitioning(g, partitioning)
        return (1.0 - p) * _a(g, "other_gas_wet_dep") + p * pwd
    if s == IDX_SO2:
        return _a(g, "SO2_wet_dep").copy()
    if s == IDX_TOTAL_NO3:
        p = no3_partitioning(g, partitioning)
        return (1.0 - p) * _a(g, "other_gas_wet_dep") + p * pwd
    return np.zeros((g.nz, g.ny, g.nx))


def photolysis_rate(g, s, a_photo):
    """A_PHOTO j(NO2) F_p for VBS bins (deposition.py:149-170, soa_photolysis.py:64-76)."""
    j = _a(g, "j_no2_soa")
    if s not in IDX_VBS_BINS or j is None or a_photo <= 0.0:
        return 0.0
    F = _a(g, "F_p_vbs")
    Fi = 1.0 if F is None or F.shape[0] != 5 else F[IDX_VBS_BINS.index(s)]
    return a_photo * j * Fi


def deposition_diag(g, s, partitioning="marginal", a_photo=0.0):
    """deposition.py:101-146, as a flat (N,) diagonal."""
    d = wet_rate(g, s, partitioning).copy()
    dz0 = np.asarray(g.Dz, dtype=np.float64)[0]
    vd = dry_velocity(g, s, partitioning)
    d[0] += np.where(dz0 > 0, vd / np.where(dz0 > 0, dz0, 1.0), 0.0)
    d = d + photolysis_rate(g, s, a_photo)
    return d.ravel()


def species_diagonals(g, partitioning="marginal", so2_ox_rate=None, nox_to_no3_rate=None, o3_loss_rate=None,
                      co_loss_rate=None, a_photo=0.0, k_age=4.0e-11, aging_gain=1.075):
    """(d_species (14, N), K {(target, source): (N,) diagonal}) as
    assemble_species_operators builds them (operator.py:430-583)."""
    D = [deposition_diag(g, s, partitioning, a_photo) for s in range(N_SPECIES)]
    so2 = (np.asarray(so2_ox_rate, dtype=np.float64) if so2_ox_rate is not None
           else _a(g, "SO2oxidation")).ravel()
    K = {(IDX_PSO4, IDX_SO2): -so2}
    extra = {IDX_SO2: so2}
    if nox_to_no3_rate is not None:
        r = np.asarray(nox_to_no3_rate, dtype=np.float64).ravel()
        extra[IDX_NOX] = r
        K[(IDX_TOTAL_NO3, IDX_NOX)] = -r
    if o3_loss_rate is not None:
        extra[IDX_O3] = np.asarray(o3_loss_rate, dtype=np.float64).ravel()
    if co_loss_rate is not None:
        extra[IDX_CO] = np.asarray(co_loss_rate, dtype=np.float64).ravel()
    oh = _a(g, "archive_OH")
    age = {}
    if oh is not None:
        rate_oh = (k_age * oh).ravel()
        F = _a(g, "F_p_vbs")
        for src, dst in VBS_AGING_PAIRS:
            if F is not None and F.shape[0] == 5:
                r = rate_oh * (1.0 - np.clip(F[IDX_VBS_BINS.index(src)], 0.0, 1.0).ravel())
            else:
                r = rate_oh
            age[src] = r
            K[(dst, src)] = -aging_gain * r
    d = np.empty((N_SPECIES, D[0].size))
    for s in range(N_SPECIES):
        d[s] = D[s]
        if s in extra:
            d[s] = d[s] + extra[s]
        if s in age:
            d[s] = d[s] + age[s]
    return d, K
