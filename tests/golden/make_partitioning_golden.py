"""Write tests/golden/partitioning_golden.npz: the reference's own ISORROPIA
partitioning update on synthetic inputs, the golden vectors that pin
oracle/isorropia_ref.py and the device kernel.

Imports the reference's orbit/core/isorropia_closure.py (partitioning_per_bin,
update_grid_partitioning) and orbit/core/isorropia_lut.py (IsorropiaLUT, on a
seeded synthetic 7-D table written to a temporary archive) -- numpy/scipy only
-- in the build container, never on the GPU box.  Inputs: two bins of a 3x4x5
grid whose end-of-bin pSO4 / TotalNH / TotalNO3 states cover the table, its
outside (linear extrapolation), exact zeros (one-sided differences) and
negative values (clipped to 0); the first bin has dust and sea salt, the
second neither (the zero fallbacks).  Outputs: every field with cross
partials, and the blend with alpha = 0.5 against a previous iteration.

    python tests/golden/make_partitioning_golden.py
"""
import os
import sys
import tempfile
import types

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
sys.path.insert(0, ROOT)
sys.path.insert(0, REF)

from oracle import lut_ref  # noqa: E402  (the seeded table only)
from orbit.core import isorropia_closure as ic  # noqa: E402
from orbit.core.isorropia_lut import IsorropiaLUT  # noqa: E402

NZ, NY, NX = 3, 4, 5
N = NZ * NY * NX


def inputs(seed=5):
    rng = np.random.default_rng(seed)
    grids, orbits = [], {2: [None] * 9, 5: [None] * 9, 6: [None] * 9}
    for tau in range(2):
        g = types.SimpleNamespace(nz=NZ, ny=NY, nx=NX)
        g.Temperature = rng.uniform(255.0, 315.0, (NZ, NY, NX))
        g.RH = rng.uniform(5.0, 105.0, (NZ, NY, NX))
        if tau == 0:
            g.dust_fine = rng.uniform(0.0, 40.0, (NZ, NY, NX))
            g.sea_salt_fine = rng.uniform(0.0, 5.0, (NZ, NY, NX))
        else:
            g.dust_fine = np.array([])
            g.sea_salt_fine = np.array([])
        grids.append(g)
        for s, hi in ((2, 35.0), (5, 45.0), (6, 35.0)):
            c = rng.uniform(-2.0, hi, N)
            c[rng.random(N) < 0.1] = 0.0
            orbits[s][tau + 1] = c
    for s in orbits:
        orbits[s][0] = orbits[s][2]
        for t in range(3, 9):
            orbits[s][t] = np.zeros(N)
    return grids, orbits


def main():
    table = lut_ref.synthetic_table(17)
    grids, orbits = inputs()
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "lut.npz")
        np.savez(path, **table)
        lut = IsorropiaLUT(path)
        new = ic.partitioning_per_bin(grids, orbits, lut, include_cross_partials=True)
        prev = {tau: {k: np.full_like(v, 0.25 + 0.1 * tau) for k, v in d.items()} for tau, d in new.items()}
        blended = ic.update_grid_partitioning(grids, orbits, lut, alpha=0.5, prev_per_bin=prev)
    out = {f"table_{k}": v for k, v in table.items() if k not in ("aerosol_water", "ph")}
    for tau, g in enumerate(grids):
        out[f"T_{tau}"] = g.Temperature
        out[f"RH_{tau}"] = g.RH
        out[f"dust_{tau}"] = g.dust_fine
        out[f"seasalt_{tau}"] = g.sea_salt_fine
        for s in (2, 5, 6):
            out[f"state_{s}_{tau}"] = orbits[s][tau + 1]
        for k, v in new[tau].items():
            out[f"new_{k}_{tau}"] = v
            out[f"blend_{k}_{tau}"] = blended[tau][k]
        # the grid attributes the update sets
        out[f"grid_NHPartitioning_{tau}"] = g.NHPartitioning
        out[f"grid_NO3PartitioningEq_{tau}"] = g.NO3PartitioningEq
    path = os.path.join(HERE, "partitioning_golden.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path) / 1e3:.1f} kB)")


if __name__ == "__main__":
    main()
