"""GPU parity of the ISORROPIA partitioning update (orbit_lut_partitioning_dev,
reference orbit/core/isorropia_closure.py:75-478): against the reference's
own golden vectors (tests/golden/partitioning_golden.npz), against the oracle
at a 24x20x15 month, and one closure iteration -- orbits -> partitioning
(blended with the last iteration) -> species diagonals -> update_diag ->
orbits -- against the oracle on the same inputs at each stage."""
import copy

import numpy as np
import pytest
import scipy.sparse as sp
import torch

pytestmark = pytest.mark.gpu

import test_oracle_isorropia as toi
from oracle import assembly_ref as asm, isorropia_ref as iso_ref, lut_ref, orbit_ref as orb
from oracle import species_ops_ref as so
from orbit_amd import isorropia, species_ops, synth
from orbit_amd.core import OperatorSet
from orbit_amd.lut import IsorropiaLUT


def _close(got, ref, scale_tol, what):
    scale = max(1.0, float(np.abs(ref).max()))
    err = float(np.abs(got - ref).max()) / scale
    assert err <= scale_tol, (what, err)
    return err


# a cross partial c_self * (m(+) - m(-)) / (other+ - other-) divides round-off
# in the clipped marginal fractions m by its step, 2 % of the other state plus
# 1e-12: where that state is zero the step is 1e-12 and the reference's own
# value carries its round-off times 1e12.  The bar per cell is set by that
# step: 1e-13 * |c_self| / step on top of 1e-10 of the field's scale.
CROSS = {"f_nh_dno3": (2, 6), "f_no3_dnh": (6, 2), "f_nh_dso4": (2, 5), "f_no3_dso4": (6, 5)}


def _close_field(k, got, ref, states, tol=1e-10):
    if k not in CROSS:
        return _close(got, ref, tol, k)
    c_self = np.maximum(states[CROSS[k][0]], 0.0).reshape(ref.shape)
    c_other = np.maximum(states[CROSS[k][1]], 0.0).reshape(ref.shape)
    step = (c_other * 1.01 + 1e-12) - np.maximum(c_other * 0.99, 0.0)
    bar = tol * max(1.0, float(np.abs(ref).max())) + 1e-13 * c_self / step
    bad = np.abs(got - ref) > bar
    assert not bad.any(), (k, float(np.abs(got - ref)[bad].max()), int(bad.sum()))
    well = c_other > 0
    return float(np.abs(got - ref)[well].max()) / max(1.0, float(np.abs(ref).max())) if well.any() else 0.0


def test_device_partitioning_matches_the_reference_golden_vectors():
    z, table, grids, orbits = toi.golden_case()
    lut = IsorropiaLUT(arrays=table)
    new = isorropia.partitioning_per_bin(grids, orbits, lut)
    worst = 0.0
    for tau in range(2):
        st = {s: orbits[s][tau + 1] for s in (2, 5, 6)}
        for k in isorropia.FIELDS:
            worst = max(worst, _close_field(k, new[tau][k], z[f"new_{k}_{tau}"], st))
    print(f"golden: worst scaled difference where no state is zero {worst:.2e}", flush=True)
    prev = {tau: {k: np.full_like(v, 0.25 + 0.1 * tau) for k, v in d.items()} for tau, d in new.items()}
    blended = isorropia.update_grid_partitioning(grids, orbits, lut, alpha=0.5, prev_per_bin=prev)
    for tau, g in enumerate(grids):
        st = {s: orbits[s][tau + 1] for s in (2, 5, 6)}
        for k in isorropia.FIELDS:
            _close_field(k, blended[tau][k], z[f"blend_{k}_{tau}"], st)
        assert g.NHPartitioning is blended[tau]["f_nh_marg"] and g.NO3PartitioningEq is blended[tau]["f_no3_eq"]
    # without the cross partials: the four diagonal fields, the same values
    four = isorropia.partitioning_per_bin(grids, orbits, lut, include_cross_partials=False)
    assert sorted(four[0]) == sorted(isorropia.FIELDS[:4])
    for k in isorropia.FIELDS[:4]:
        np.testing.assert_array_equal(four[1][k], new[1][k])


def test_month_partitioning_matches_the_oracle():
    rng = np.random.default_rng(12)
    nz, ny, nx = 15, 20, 24
    N = nz * ny * nx
    table = lut_ref.synthetic_table(23)
    grids = []
    for b in range(8):
        grids.append(type("G", (), dict(nz=nz, ny=ny, nx=nx, Temperature=rng.uniform(260, 310, (nz, ny, nx)),
                                        RH=rng.uniform(10, 99, (nz, ny, nx)),
                                        dust_fine=rng.uniform(0, 30, (nz, ny, nx)),
                                        sea_salt_fine=rng.uniform(0, 4, (nz, ny, nx))))())
    orbit = rng.uniform(-1.0, 40.0, (3, 9, N))
    orbit[rng.random(orbit.shape) < 0.05] = 0.0
    species = (isorropia.IDX_PSO4, isorropia.IDX_TOTAL_NH, isorropia.IDX_TOTAL_NO3)
    orbits = {s: list(orbit[q]) for q, s in enumerate(species)}
    lut = IsorropiaLUT(arrays=table)
    mp = isorropia.MonthPartitioning(grids, lut)
    dev = torch.from_numpy(orbit).cuda()
    got = mp.compute(dev[0, 1:], dev[1, 1:], dev[2, 1:]).cpu().numpy()  # strided rows of one tensor
    ref = iso_ref.partitioning_per_bin(grids, orbits, table)
    worst = 0.0
    for f, k in enumerate(isorropia.FIELDS):
        r = np.stack([ref[b][k].ravel() for b in range(8)])
        st = {s: orbit[q, 1:] for q, s in enumerate(species)}
        worst = max(worst, _close_field(k, got[f], r, st))
    print(f"24x20x15 x 8 bins: worst scaled difference {worst:.2e}", flush=True)
    # blend in place on the device: out = alpha new + (1 - alpha) prev
    prev = torch.full_like(torch.from_numpy(got).cuda(), 0.3)
    bl = mp.compute(dev[0, 1:], dev[1, 1:], dev[2, 1:], alpha=0.25, prev=prev).cpu().numpy()
    np.testing.assert_allclose(bl, 0.25 * got + 0.75 * 0.3, rtol=1e-14, atol=1e-15)


def test_one_closure_iteration_matches_the_oracle():
    """Orbits of TotalNH, pSO4 and TotalNO3 on a month built from grid fields
    -> the partitioning (marginal fractions, blended with the starting ones at
    alpha = 0.5) on the device -> the species diagonals that read them ->
    update_diag -> the new orbits; each stage against the oracle given the
    same inputs (partitioning 1e-9 scaled, diagonals 1e-13, orbits 1e-8)."""
    grids = synth.make_reference_grids(10, 8, 5, seed=43)
    rng = np.random.default_rng(43)
    sh = (5, 8, 10)
    for g in grids:
        g.particle_dry_dep = np.full(sh, 2e-3); g.SO2_dry_dep = np.full(sh, 6e-3)
        g.NOx_dry_dep = np.full(sh, 1e-3); g.NH3_dry_dep = np.full(sh, 8e-3)
        g.particle_wet_dep = np.full(sh, 2e-6); g.SO2_wet_dep = np.full(sh, 3e-6)
        g.other_gas_wet_dep = np.full(sh, 4e-6); g.SO2oxidation = np.full(sh, 2e-6)
        g.NHPartitioning = np.full(sh, 0.5); g.NO3Partitioning = np.full(sh, 0.5)
        g.Temperature = rng.uniform(265, 305, sh); g.RH = rng.uniform(20, 95, sh)
        g.dust_fine = rng.uniform(0, 20, sh); g.sea_salt_fine = np.array([])
    ref_grids = copy.deepcopy(grids)
    species = [species_ops.IDX_TOTAL_NH, species_ops.IDX_PSO4, species_ops.IDX_TOTAL_NO3]
    D0, _ = species_ops.month_diagonals(grids, species)
    op = OperatorSet.from_grids(grids, D0)
    e = np.zeros((3, 8, op.N))
    e[:, :, :80] = rng.lognormal(0, 1, (3, 8, 80))
    o0, _ = op.solve_forward(e, tol=1e-11, precision="fp64")
    o0 *= 25.0 / np.abs(o0).max()  # concentrations across the table's range
    orbits = {s: list(o0[q]) for q, s in enumerate(species)}
    table = lut_ref.synthetic_table(29)
    lut = IsorropiaLUT(arrays=table)
    start = {b: {k: np.full(sh, 0.5) for k in isorropia.FIELDS} for b in range(8)}
    blended = isorropia.update_grid_partitioning(grids, orbits, lut, alpha=0.5, prev_per_bin=start)
    ref_new = iso_ref.blend(iso_ref.partitioning_per_bin(ref_grids, orbits, table), start, 0.5)
    for b in range(8):
        for k in ("f_nh_marg", "f_no3_marg"):
            _close(blended[b][k], ref_new[b][k], 1e-9, (k, b))
    D1, _ = species_ops.month_diagonals(grids, species)
    op.update_diag(D1)
    o1, info = op.solve_forward(e, tol=1e-11, precision="fp64")
    To = [asm.assemble_transport_block(g) for g in ref_grids]
    for b, g in enumerate(ref_grids):   # the oracle's grids take the device's partition fields
        g.NHPartitioning = blended[b]["f_nh_marg"]
        g.NO3Partitioning = blended[b]["f_no3_marg"]
    for q, s in enumerate(species):
        Do = np.stack([so.species_diagonals(g)[0][s] for g in ref_grids])
        np.testing.assert_allclose(D1[:, q], Do, rtol=1e-13, atol=0)
        lus = [orb.factor_one_bin((To[b] + sp.diags(Do[b])).tocsr()) for b in range(8)]
        ref, _ = orb.solve_orbit_one_species(lus, list(e[q]), tol=1e-12, maxiter=500)
        assert info[q]["converged"]
        assert np.abs(o1[q, 0] - ref[0]).max() <= 1e-8 * np.abs(ref[0]).max(), s
    # the partitioning moved the TotalNH loss: the orbit changed
    assert np.abs(D1[:, 0] - D0[:, 0]).max() > 0
