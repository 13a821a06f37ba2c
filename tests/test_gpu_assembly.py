"""GPU parity of the device-side transport assembly and of operator sets built
from grid fields.

* The device stencil reproduces the reference's own assemble_transport_block
  (golden vectors of tests/golden/assembly_golden.npz) to 1e-12 relative, on
  every case, the two periodic-longitude grids (wrap face) included.
* An operator set built from grids (no host CSR) holds exactly the arrays the
  CSR path builds from the same T: their solves agree bit for bit; its
  periodic solve matches the oracle (splu + GMRES on T_oracle + D) to 1e-8.
* The air-mass pre-flight on the device-assembled blocks matches the oracle's
  statistics and applies the reference's abort threshold.
"""
import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu

import test_oracle_assembly as toa
from oracle import assembly_ref as asm, mass_ref, orbit_ref as orb
from orbit_amd import _lib, assembly, synth
from orbit_amd.core import OperatorSet


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def test_device_stencil_matches_reference_golden_vectors():
    z, names = toa._golden_cases()
    checked = 0
    for name in names:
        g = toa.golden_grid(z, name)
        N = g.nz * g.ny * g.nx
        A = assembly.assemble_transport_block(g)
        B = toa.golden_matrix(z, name, "T", N)
        assert abs(A - B).max() <= 1e-12 * abs(B).max(), name
        # and the stencil entry by entry against the oracle's
        st = assembly.transport_stencil(g)
        ref = asm.transport_stencil(g)
        for k in ref:
            assert np.abs(st[k] - ref[k]).max() <= 1e-12 * max(np.abs(ref[k]).max(), 1e-300), (name, k)
        checked += 1
    assert checked == len(names) >= 6


@pytest.fixture(scope="module")
def gridset():
    grids, D, e = synth.grid_workload(12, 10, 6, 3, seed=21)
    op_g = OperatorSet.from_grids(grids, D)
    return grids, D, e, op_g


def test_grid_opset_equals_the_csr_opset_and_the_oracle(gridset):
    grids, D, e, op_g = gridset
    g0 = grids[0]
    Ts = [assembly.assemble_transport_block(g) for g in grids]
    op_c = OperatorSet.from_shared(Ts, D, g0.nz, g0.ny, g0.nx)
    for prec in ("fp64", "mixed"):
        og, ig = op_g.solve_forward(e, tol=1e-10, precision=prec)
        oc, ic = op_c.solve_forward(e, tol=1e-10, precision=prec)
        assert np.array_equal(og, oc), prec
        assert all(i["converged"] for i in ig)
    lg, _ = op_g.solve_adjoint(e, tol=1e-10, precision="fp64")
    lc, _ = op_c.solve_adjoint(e, tol=1e-10, precision="fp64")
    assert np.array_equal(lg, lc)
    # against the oracle: splu per bin of A = I + dtau (T_oracle + D)
    To = [asm.assemble_transport_block(g) for g in grids]
    og, _ = op_g.solve_forward(e, tol=1e-11, precision="fp64")
    for s in range(e.shape[0]):
        lus = [orb.factor_one_bin((To[t] + sp.diags(D[t, s])).tocsr()) for t in range(8)]
        ref, _ = orb.solve_orbit_one_species(lus, list(e[s]), tol=1e-12, maxiter=500)
        assert _rel(og[s, 0], ref[0]) < 1e-8, s
        assert _rel(og[s, 5], ref[5]) < 1e-8, s


def test_air_mass_preflight_on_device_assembled_blocks(gridset):
    grids, D, e, op_g = gridset
    mb = assembly.air_mass_preflight(op_g, grids)
    g0 = grids[0]
    rho = asm.density_weights(g0)
    ref = mass_ref.mass_balance_diagnostics([asm.assemble_transport_block(g) for g in grids], rho,
                                            g0.nz, g0.ny, g0.nx)
    assert mb["interior_p90"] < 1e-8
    np.testing.assert_allclose(mb["top_p90"], ref["top_p90"], rtol=1e-9)
    np.testing.assert_allclose(mb["layer_p90"][-1], ref["layer_p90"][-1], rtol=1e-9)
    # a grid whose omega does not close the air budget is refused
    bad, Db, _ = synth.grid_workload(12, 10, 6, 1, seed=22, interface_omega=False)
    op_b = OperatorSet.from_grids(bad, Db)
    with pytest.raises(assembly.AirMassImbalance):
        assembly.air_mass_preflight(op_b, bad)
    mbb = assembly.air_mass_preflight(op_b, bad, allow_imbalance=True)
    refb = mass_ref.mass_balance_diagnostics([asm.assemble_transport_block(g) for g in bad],
                                             asm.density_weights(bad[0]), 6, 10, 12)
    for k in ("interior_median", "interior_p90", "interior_p99", "top_median"):
        np.testing.assert_allclose(mbb[k], refb[k], rtol=1e-9)
    assert not mass_ref.mass_balance_gate(refb)


def test_grid_opset_rejects_bad_input(gridset):
    grids, D, _, _ = gridset
    with pytest.raises(ValueError):
        OperatorSet.from_grids(grids, D[:, :, :-1])
    import types
    g = types.SimpleNamespace(**vars(grids[0]))
    g.has_split_fluxes = True          # split fluxes without UAvg_plus etc.
    with pytest.raises(_lib.OrbitError):
        assembly.transport_stencil(g)


def test_periodic_month_from_grids():
    """A periodic longitude: the wrap face in the device rows, in the opset's
    edge stencils, in the solves and in the air-mass check."""
    grids, D, e = synth.grid_workload(12, 10, 6, 2, seed=31, periodic=True)
    g0 = grids[0]
    op_g = OperatorSet.from_grids(grids, D)
    assert op_g.stats()["periodic_x"] == 1
    Ts = [assembly.assemble_transport_block(g) for g in grids]
    To = [asm.assemble_transport_block(g) for g in grids]
    for A, B in zip(Ts, To):
        assert abs(A - B).max() <= 1e-12 * abs(B).max()
    op_c = OperatorSet.from_shared(Ts, D, g0.nz, g0.ny, g0.nx)
    assert op_c.stats()["periodic_x"] == 1
    for prec in ("fp64", "mixed"):
        og, ig = op_g.solve_forward(e, tol=1e-10, precision=prec)
        oc, _ = op_c.solve_forward(e, tol=1e-10, precision=prec)
        assert np.array_equal(og, oc), prec
        assert all(i["converged"] for i in ig)
    og, _ = op_g.solve_forward(e, tol=1e-11, precision="fp64")
    lg, _ = op_g.solve_adjoint(e, tol=1e-11, precision="fp64")
    for s in range(e.shape[0]):
        lus = [orb.factor_one_bin((To[t] + sp.diags(D[t, s])).tocsr()) for t in range(8)]
        ref, _ = orb.solve_orbit_one_species(lus, list(e[s]), tol=1e-12, maxiter=500)
        assert _rel(og[s, 0], ref[0]) < 1e-8, s
        lref, _ = orb.solve_orbit_adjoint_one_species(lus, list(e[s]), tol=1e-12, maxiter=500)
        assert _rel(lg[s, 0], lref[0]) < 1e-8, s
    # omega closes the air budget through the wrap face as well
    mb = assembly.air_mass_preflight(op_g, grids)
    assert mb["interior_p90"] < 1e-8
    ref = mass_ref.mass_balance_diagnostics(To, asm.density_weights(g0), g0.nz, g0.ny, g0.nx)
    np.testing.assert_allclose(mb["top_p90"], ref["top_p90"], rtol=1e-9)


def test_periodic_odd_nx_month_from_grids():
    """A periodic longitude with an odd nx from grid fields: the device stencil
    equals the reference's golden-checked assembly, the grid-built opset equals
    the CSR-built one bit for bit, and its orbits match the oracle."""
    grids, D, e = synth.grid_workload(11, 6, 4, 2, seed=32, periodic=True)
    A = assembly.assemble_transport_block(grids[2])
    B = asm.assemble_transport_block(grids[2])
    assert abs(A - B).max() <= 1e-12 * abs(B).max()
    op_g = OperatorSet.from_grids(grids, D)
    To = [asm.assemble_transport_block(g) for g in grids]
    op_c = OperatorSet.from_shared([assembly.assemble_transport_block(g) for g in grids], D, 4, 6, 11)
    assert op_g.stats()["periodic_x"] == 1
    for prec in ("fp64", "mixed"):
        og, ig = op_g.solve_forward(e, tol=1e-10, precision=prec)
        oc, _ = op_c.solve_forward(e, tol=1e-10, precision=prec)
        assert np.array_equal(og, oc), prec
        assert all(i["converged"] for i in ig)
    og, _ = op_g.solve_forward(e, tol=1e-11, precision="fp64")
    lg, _ = op_g.solve_adjoint(e, tol=1e-11, precision="fp64")
    for s in range(e.shape[0]):
        lus = [orb.factor_one_bin((To[t] + sp.diags(D[t, s])).tocsr()) for t in range(8)]
        ref, _ = orb.solve_orbit_one_species(lus, list(e[s]), tol=1e-12, maxiter=500)
        lref, _ = orb.solve_orbit_adjoint_one_species(lus, list(e[s]), tol=1e-12, maxiter=500)
        assert _rel(og[s, 0], ref[0]) < 1e-8 and _rel(lg[s, 0], lref[0]) < 1e-8, s


def test_fct_forward_month_on_grid_built_operators_matches_the_oracle():
    """Operators assembled on the device from grid fields drive the forward FCT
    passes (orbit/cli.py:1212-1241) on the device; the oracle runs the same
    passes with splu on its own assembly of the same grids."""
    import torch
    from oracle import drivers_ref
    from orbit_amd import drivers
    grids, D, e = synth.grid_workload(12, 10, 6, 2, seed=51)
    op = OperatorSet.from_grids(grids, D)
    fg = drivers.make_fct_grids(grids, op.ctx)
    orbit, hist = drivers.forward_fct(op, fg, torch.from_numpy(e).cuda(), n_passes=4, tol=1e-10,
                                      precision="fp64")
    To = [asm.assemble_transport_block(g) for g in grids]
    for s in range(2):
        lus = [orb.factor_one_bin((To[t] + sp.diags(D[t, s])).tocsr()) for t in range(8)]
        ref, _ = drivers_ref.forward_fct_passes(lus, list(e[s]), grids, n_passes=4, tol=1e-11, maxiter=500)
        assert _rel(orbit[s, 0].cpu().numpy(), ref[0]) < 1e-8, s
        assert _rel(orbit[s, 8].cpu().numpy(), ref[8]) < 1e-8, s
