"""Write tests/golden/species_ops_golden.npz from the reference's own
assemble_species_operators (orbit/core/operator.py:371-595) and its deposition
module (orbit/core/deposition.py:101-362).

The per-species diagonals combine many branches (partition-weighted blends,
fallback fields, the SOA photolytic sink, VBS aging) whose exact values are
fixed only by the reference code, so this script imports those modules
(numpy/scipy only) from the reference checkout in this container on two small
grids -- one with every optional field, one with none -- and records inputs and
outputs.  Run from the repo root:

    python tests/golden/make_species_ops_golden.py /root/reference

The output is data only; no reference source is stored.
"""
import os
import sys
import warnings

import numpy as np

FIELDS = ("particle_dry_dep", "SO2_dry_dep", "NOx_dry_dep", "NH3_dry_dep", "VOC_dry_dep", "HNO3_dry_dep",
          "O3_dry_dep", "is_land", "particle_wet_dep", "SO2_wet_dep", "other_gas_wet_dep", "NHPartitioning",
          "NHPartitioningEq", "NO3Partitioning", "NO3PartitioningEq", "F_p_vbs", "F_p_vbs_surface", "j_no2_soa",
          "archive_OH", "SO2oxidation", "Dz", "dP")
RATES = ("so2_ox_rate", "nox_to_no3_rate", "o3_loss_rate", "co_loss_rate")


def build(GridData, seed, full):
    rng = np.random.default_rng(seed)
    nz, ny, nx = 3, 4, 5
    g = GridData()
    g.nz, g.ny, g.nx = nz, ny, nx
    sh = (nz, ny, nx)
    g.Dz = np.array([120.0, 500.0, 1500.0])[:, None, None] * (1 + 0.1 * rng.random(sh))
    g.Dz[0, 1, 2] = 0.0                       # degenerate surface cell: no dry deposition there
    g.dP = 1.2 * 9.80665 * g.Dz
    for f in ("particle_dry_dep", "SO2_dry_dep", "NOx_dry_dep", "NH3_dry_dep"):
        setattr(g, f, rng.uniform(1e-4, 1e-2, sh))
    for f in ("particle_wet_dep", "SO2_wet_dep", "other_gas_wet_dep"):
        setattr(g, f, rng.uniform(1e-7, 1e-4, sh))
    g.NHPartitioning = rng.uniform(0, 1, sh)
    g.SO2oxidation = rng.uniform(1e-7, 1e-5, sh)
    g.is_land = (rng.random((ny, nx)) < 0.6).astype(np.uint8)
    rates = {}
    if full:
        g.VOC_dry_dep = rng.uniform(1e-4, 5e-3, sh)
        g.HNO3_dry_dep = rng.uniform(5e-3, 5e-2, sh)
        g.O3_dry_dep = rng.uniform(1e-4, 8e-3, sh)
        g.NHPartitioningEq = rng.uniform(0, 1, sh)
        g.NO3Partitioning = rng.uniform(0, 1, sh)
        g.NO3PartitioningEq = rng.uniform(0, 1, sh)
        g.F_p_vbs = rng.uniform(-0.1, 1.1, (5,) + sh)          # clipped only in the aging rate
        g.F_p_vbs_surface = rng.uniform(0, 1, (5, ny, nx))
        g.j_no2_soa = rng.uniform(0, 1e-2, sh)
        g.archive_OH = rng.uniform(1e5, 1e7, sh)
        rates = {r: rng.uniform(1e-7, 1e-5, sh) for r in RATES}
    return g, rates


CASES = {"full_marginal": (1, True, "marginal", 4.0e-4), "full_equilibrium": (2, True, "equilibrium", 4.0e-4),
         "bare_equilibrium": (3, False, "equilibrium", 0.0), "bare_marginal": (4, False, "marginal", 0.0)}


def main(ref_root):
    sys.path.insert(0, ref_root)
    from orbit.core.grid_data import GridData
    from orbit.core.indexing import CellIndexer
    from orbit.core import operator as opmod
    out = {}
    for name, (seed, full, part, a_photo) in CASES.items():
        os.environ["ORBIT_VBS_A_PHOTO"] = repr(a_photo)
        g, rates = build(GridData, seed, full)
        idx = CellIndexer(g.nz, g.ny, g.nx)
        # the transport block is covered by the assembly fixture: zero winds keep it cheap here
        g.UAvg = np.zeros((g.nz, g.ny, g.nx)); g.VAvg = np.zeros_like(g.UAvg)
        g.Kxxyy = np.full_like(g.UAvg, 100.0); g.K_meander_u = np.zeros_like(g.UAvg)
        g.K_meander_v = np.zeros_like(g.UAvg); g.Kzz = np.full_like(g.UAvg, 10.0)
        g.dx = np.full(g.ny, 5e4); g.dy = 5e4; g.lat = np.linspace(10, 12, g.ny); g.dlat = 0.5; g.dlon = 0.625
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            L, K, T, d = opmod.assemble_species_operators(g, idx, verbose=False, partitioning=part, **rates)
        p = f"{name}/"
        out[p + "meta"] = np.array([g.nz, g.ny, g.nx, int(full)])
        out[p + "partitioning"] = np.array(part)
        out[p + "a_photo"] = np.array(a_photo)
        for f in FIELDS:
            v = getattr(g, f, None)
            if v is not None and np.size(v) > 0:
                out[p + "in/" + f] = np.asarray(v, dtype=np.float64)
        for r, v in rates.items():
            out[p + "rate/" + r] = v
        out[p + "d_species"] = d
        for (t, s), M in K.items():
            out[p + f"K/{t}_{s}"] = np.asarray(M.diagonal()).ravel()
        print(name, "couplings", sorted(K))
    os.environ.pop("ORBIT_VBS_A_PHOTO", None)
    np.savez_compressed("tests/golden/species_ops_golden.npz", **out)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
