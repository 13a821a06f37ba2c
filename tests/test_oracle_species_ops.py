"""Pins of the per-species diagonal oracle (oracle/species_ops_ref.py): the
reference's own assemble_species_operators on grids with every optional field
and with none (golden vectors), the blends' end points, and the couplings'
mass bookkeeping."""
import os
import types

import numpy as np
import pytest

from oracle import species_ops_ref as so

GOLD = os.path.join(os.path.dirname(__file__), "golden", "species_ops_golden.npz")


def golden_cases():
    z = np.load(GOLD)
    return z, sorted({k.split("/")[0] for k in z.files})


def golden_grid(z, name):
    nz, ny, nx, full = (int(v) for v in z[f"{name}/meta"])
    g = types.SimpleNamespace(nz=nz, ny=ny, nx=nx)
    pre = f"{name}/in/"
    for k in z.files:
        if k.startswith(pre):
            setattr(g, k[len(pre):], z[k])
    rates = {k.split("/")[-1]: z[k] for k in z.files if k.startswith(f"{name}/rate/")}
    return g, rates, str(z[f"{name}/partitioning"]), float(z[f"{name}/a_photo"])


def golden_K(z, name):
    pre = f"{name}/K/"
    return {tuple(int(x) for x in k[len(pre):].split("_")): z[k] for k in z.files if k.startswith(pre)}


@pytest.mark.parametrize("name", golden_cases()[1])
def test_oracle_matches_reference_golden(name):
    z, _ = golden_cases()
    g, rates, part, a_photo = golden_grid(z, name)
    d, K = so.species_diagonals(g, part, a_photo=a_photo, **rates)
    ref = z[f"{name}/d_species"]
    np.testing.assert_allclose(d, ref, rtol=1e-14, atol=0)
    KG = golden_K(z, name)
    assert set(K) == set(KG)
    for key in K:
        np.testing.assert_allclose(K[key], KG[key], rtol=1e-14, atol=0)


def test_blend_end_points_and_coupling_bookkeeping():
    z, _ = golden_cases()
    g, rates, part, a = golden_grid(z, "full_marginal")
    for p, expect_nh in ((0.0, "NH3_dry_dep"), (1.0, "particle_dry_dep")):
        g2 = types.SimpleNamespace(**vars(g))
        g2.NHPartitioning = np.full_like(g.NHPartitioning, p)
        v = so.dry_velocity(g2, so.IDX_TOTAL_NH, "marginal")
        np.testing.assert_array_equal(v, getattr(g, expect_nh)[0])
        w = so.wet_rate(g2, so.IDX_TOTAL_NH, "marginal")
        np.testing.assert_array_equal(w, g.other_gas_wet_dep if p == 0 else g.particle_wet_dep)
    d, K = so.species_diagonals(g, part, a_photo=a, **rates)
    d0, K0 = so.species_diagonals(g, part, a_photo=a)           # no rates: the grid's SO2oxidation
    # SO2 loses what pSO4 gains; aging moves gain x the source's loss
    dep_so2 = so.deposition_diag(g, so.IDX_SO2, part, a)
    np.testing.assert_allclose(d[so.IDX_SO2] - dep_so2, rates["so2_ox_rate"].ravel(), rtol=1e-9)
    np.testing.assert_allclose(d0[so.IDX_SO2] - dep_so2, g.SO2oxidation.ravel(), rtol=1e-9)
    np.testing.assert_array_equal(K0[(so.IDX_PSO4, so.IDX_SO2)], -g.SO2oxidation.ravel())
    np.testing.assert_array_equal(K[(so.IDX_PSO4, so.IDX_SO2)], -rates["so2_ox_rate"].ravel())
    loss = d[so.IDX_VBS_C1000] - so.deposition_diag(g, so.IDX_VBS_C1000, part, a)
    np.testing.assert_allclose(K[(so.IDX_VBS_C100, so.IDX_VBS_C1000)], -1.075 * loss, rtol=1e-12)
    # the degenerate surface cell (Dz = 0) gets no dry deposition
    flat0 = 1 * g.nx + 2
    assert so.deposition_diag(g, so.IDX_PM25, part)[flat0] == g.particle_wet_dep[0, 1, 2]
