"""GPU parity of the per-species diagonals and couplings against the
reference's golden vectors (tests/golden/species_ops_golden.npz) and the
oracle, and a month's species operator set built from grid fields end to end
(device transport + device diagonals) against the oracle's solve."""
import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu

import test_oracle_species_ops as tso
from oracle import assembly_ref as asm, orbit_ref as orb, species_ops_ref as so
from orbit_amd import species_ops, synth
from orbit_amd.core import OperatorSet


@pytest.mark.parametrize("name", tso.golden_cases()[1])
def test_device_diagonals_match_reference_golden(name):
    z, _ = tso.golden_cases()
    g, rates, part, a_photo = tso.golden_grid(z, name)
    kw = {"so2_ox_rate": rates.get("so2_ox_rate"), "nox_to_no3_rate": rates.get("nox_to_no3_rate"),
          "o3_loss_rate": rates.get("o3_loss_rate"), "co_loss_rate": rates.get("co_loss_rate")}
    d, K = species_ops.species_diagonals(g, part, a_photo=a_photo, **kw)
    np.testing.assert_allclose(d, z[f"{name}/d_species"], rtol=1e-14, atol=0)
    KG = tso.golden_K(z, name)
    assert set(K) == set(KG)
    for key in K:
        np.testing.assert_allclose(K[key], KG[key], rtol=1e-14, atol=0)


def test_species_operator_set_from_grids_matches_the_oracle():
    grids = synth.make_reference_grids(10, 8, 5, seed=41)
    rng = np.random.default_rng(41)
    sh = (5, 8, 10)
    for g in grids:      # deposition fields of the month (same draw for every bin: a month's climatology)
        g.particle_dry_dep = np.full(sh, 2e-3); g.SO2_dry_dep = np.full(sh, 6e-3)
        g.NOx_dry_dep = np.full(sh, 1e-3); g.NH3_dry_dep = np.full(sh, 8e-3)
        g.particle_wet_dep = np.full(sh, 2e-6); g.SO2_wet_dep = np.full(sh, 3e-6)
        g.other_gas_wet_dep = np.full(sh, 4e-6); g.SO2oxidation = np.full(sh, 2e-6)
        g.NHPartitioning = rng.uniform(0, 1, sh)
    species = [species_ops.IDX_SO2, species_ops.IDX_PSO4, species_ops.IDX_TOTAL_NH]
    D, Ks = species_ops.month_diagonals(grids, species)
    op = OperatorSet.from_grids(grids, D)
    e = np.zeros((3, 8, op.N))
    e[:, :, : 80] = rng.lognormal(0, 1, (3, 8, 80)) * 1e-5
    o, info = op.solve_forward(e, tol=1e-11, precision="fp64")
    To = [asm.assemble_transport_block(g) for g in grids]
    for k, s in enumerate(species):
        Do = np.stack([so.species_diagonals(g)[0][s] for g in grids])
        np.testing.assert_allclose(D[:, k], Do, rtol=1e-13, atol=0)
        lus = [orb.factor_one_bin((To[b] + sp.diags(Do[b])).tocsr()) for b in range(8)]
        ref, _ = orb.solve_orbit_one_species(lus, list(e[k]), tol=1e-12, maxiter=500)
        assert np.abs(o[k, 0] - ref[0]).max() <= 1e-8 * np.abs(ref[0]).max(), s
    np.testing.assert_array_equal(Ks[0][(species_ops.IDX_PSO4, species_ops.IDX_SO2)], -grids[0].SO2oxidation.ravel())
