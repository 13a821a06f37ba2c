"""GPU parity of the lateral boundary inflow (reference orbit/core/lateral_bc.py)
and of the forward FCT passes that take it (orbit/cli.py:1202-1241).

* orbit_lateral_inflow reproduces the oracle's r_in to 1e-15 on net and
  split winds, a periodic longitude, a single-column row and the reference's
  golden assembly grids (true outer faces or their stand-ins).
* A month with a boundary file for one of two tracers: the reader builds
  s_bc = r_in * c_bc and the ghosts, the batched forward FCT passes on the
  device (one FCT launch for the tracer with a ghost, one for the other)
  match the oracle's passes with the same source and ghost to 1e-8."""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

pytestmark = pytest.mark.gpu

import test_oracle_assembly as toa
from oracle import assembly_ref as asm, drivers_ref, lateral_bc_ref as lbc_ref, orbit_ref as orb
from orbit_amd import drivers, lateral_bc, synth
from orbit_amd.core import OperatorSet


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def test_device_inflow_rate_matches_the_oracle():
    cases = []
    for kw in (dict(), dict(split=True), dict(periodic=True), dict(split=True, periodic=True)):
        cases += synth.make_reference_grids(10, 7, 4, seed=3, **kw)[:2]
    cases += synth.make_reference_grids(1, 5, 3, seed=4)[:1]
    z, names = toa._golden_cases()
    cases += [toa.golden_grid(z, n) for n in names]
    for g in cases:
        got = lateral_bc.compute_lateral_boundary_inflow(g)
        ref = lbc_ref.compute_lateral_boundary_inflow(g)
        # the outer faces' zonal width uses the host's cos on both sides: equal up to its last bit
        np.testing.assert_allclose(got, ref, rtol=1e-15, atol=0,
                                   err_msg=str((g.nx, g.ny, getattr(g, "has_split_fluxes", False))))


def test_forward_fct_with_lateral_inflow_matches_the_oracle(tmp_path):
    grids, D, e = synth.grid_workload(12, 10, 6, 2, seed=61)
    for g in grids:
        g.lon = 60.0 + (np.arange(g.nx) + 0.5) * g.dlon
    rng = np.random.default_rng(62)
    c_bc = rng.uniform(0.5, 2.0, (1, 8, 6, 10, 12)) * 1e-3
    path = tmp_path / "bc_M03.npz"
    np.savez(path, lat=grids[0].lat, lon=grids[0].lon, nz=6, species=np.array(["SO2"]), c_bc=c_bc)
    names = ["PM25", "SO2"]
    src = lateral_bc.load_lateral_bc(str(path), grids, names)
    assert list(src) == [1] and len(src.c_bc[1]) == 8
    # batch order: tracer 0 = SO2 (with inflow), tracer 1 = PM25 (closed edges)
    species = [1, 0]
    op = OperatorSet.from_grids(grids, D)
    fg = drivers.make_fct_grids(grids, op.ctx)
    e2, ghosts = lateral_bc.batch_inputs(src, species, torch.from_numpy(e).cuda())
    assert ghosts[0] is not None and ghosts[1] is None
    orbit, _ = drivers.forward_fct(op, fg, e2, n_passes=4, tol=1e-10, precision="fp64", ghost=ghosts)
    To = [asm.assemble_transport_block(g) for g in grids]
    rates = [lbc_ref.compute_lateral_boundary_inflow(g) for g in grids]
    for q, s in enumerate(species):
        lus = [orb.factor_one_bin((To[t] + sp.diags(D[t, q])).tocsr()) for t in range(8)]
        if s == 1:
            cb = [c_bc[0, t].ravel() for t in range(8)]
            e_list = [e[q, t] + rates[t] * cb[t] for t in range(8)]
            ref, _ = drivers_ref.forward_fct_passes(lus, e_list, grids, n_passes=4, tol=1e-11, maxiter=500,
                                                    ghost_list=cb)
        else:
            ref, _ = drivers_ref.forward_fct_passes(lus, list(e[q]), grids, n_passes=4, tol=1e-11, maxiter=500)
        assert _rel(orbit[q, 0].cpu().numpy(), ref[0]) < 1e-8, s
        assert _rel(orbit[q, 8].cpu().numpy(), ref[8]) < 1e-8, s
    # the tracer without boundary values is the same orbit as with no file at all
    plain, _ = drivers.forward_fct(op, fg, torch.from_numpy(e).cuda(), n_passes=4, tol=1e-10, precision="fp64")
    assert torch.equal(orbit[1], plain[1])
    # the low-order orbit (no FCT: linear, non-negative propagators) only rises with a
    # non-negative inflow source; the limited FCT passes need not be monotone
    lo_bc, _ = op.solve_forward(e2, tol=1e-12, precision="fp64")
    lo, _ = op.solve_forward(torch.from_numpy(e).cuda(), tol=1e-12, precision="fp64")
    diff = lo_bc[0] - lo[0]
    assert float(diff.min()) >= -1e-10 * float(lo[0].abs().max()) and float(diff.max()) > 0
