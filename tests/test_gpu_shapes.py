"""Edge shapes on the device against the oracle: a single column, a single
row of columns (ny = 1) and a single column of rows (nx = 1), odd sizes, one
level (nz = 1) and the largest column (nz = 16), for the propagator, the
forward and adjoint periodic solves in both precisions, the FCT source, the
forward FCT passes (mixed, in the library) and the
transport assembly."""
import types

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import assembly_ref as asm, fct_ref, orbit_ref as orb
from orbit_amd import assembly, fct as hfct, synth
from orbit_amd.core import OperatorSet

SHAPES = [(1, 1, 15), (1, 6, 8), (7, 1, 8), (3, 2, 16), (2, 2, 1), (5, 3, 2)]


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("nx,ny,nz", SHAPES)
def test_solves_on_edge_shapes_match_the_oracle(nx, ny, nz):
    wl = synth.make_workload(nx, ny, nz, 2, seed=[nx, ny, nz], convection=nz > 1)
    e = wl.e.copy()
    e[:, :, : nx * ny] += 1e-6          # every surface cell emits (a single column may draw none)
    op = OperatorSet.from_workload(wl)
    lus = [[orb.factor_one_bin(wl.L(t, s)) for t in range(8)] for s in range(2)]
    b = np.random.default_rng(0).standard_normal(wl.N)
    for tr in (False, True):
        x = op.apply_P(3, 1, b, trans=tr)
        ref = lus[1][3].solve(b, trans="T" if tr else "N")
        assert _rel(x, ref) < 1e-11, tr
    f64, i64 = op.solve_forward(e, tol=1e-11, precision="fp64")
    fmx, imx = op.solve_forward(e, tol=1e-8, precision="mixed")
    adj, iadj = op.solve_adjoint(e, tol=1e-11, precision="fp64")
    for s in range(2):
        ref, _ = orb.solve_orbit_one_species(lus[s], list(e[s]), tol=1e-12, maxiter=500)
        assert i64[s]["converged"] and imx[s]["converged"] and iadj[s]["converged"]
        assert _rel(f64[s, 0], ref[0]) < 1e-8 and _rel(f64[s, 8], ref[8]) < 1e-8
        assert _rel(fmx[s, 0], ref[0]) < 1e-6
        aref, _ = orb.solve_orbit_adjoint_one_species(lus[s], list(e[s]), tol=1e-12, maxiter=500)
        assert _rel(adj[s, 0], aref[0]) < 1e-8


@pytest.mark.parametrize("nx,ny,nz", [(1, 6, 8), (7, 1, 8), (3, 2, 16), (2, 2, 1)])
def test_fct_source_on_edge_shapes(nx, ny, nz):
    wl = synth.make_workload(nx, ny, nz, 1, seed=[9, nx, ny])
    g = wl.grids[1]
    rng = np.random.default_rng(2)
    c = wl.rho * (1.0 + rng.random(wl.N))
    for kw in (dict(), dict(limited=False)):
        ref = fct_ref.compute_horizontal_fct_source(g, c, 10800.0, **kw)
        got = hfct.compute_horizontal_fct_source(g, None, c, 10800.0, **kw)
        np.testing.assert_allclose(got, ref, rtol=1e-11, atol=1e-13 * max(np.abs(ref).max(), 1e-300))


@pytest.mark.parametrize("nx,ny,nz", [(1, 6, 8), (7, 1, 8), (1, 1, 4), (3, 2, 16)])
def test_assembly_on_edge_shapes(nx, ny, nz):
    g = synth.make_reference_grids(nx, ny, nz, seed=[nx, ny, nz])[3]
    st = assembly.transport_stencil(g)
    ref = asm.transport_stencil(g)
    for k in ref:
        assert np.abs(st[k] - ref[k]).max() <= 1e-12 * max(np.abs(ref[k]).max(), 1e-300), k


@pytest.mark.parametrize("nx,ny,nz", [(1, 6, 8), (7, 1, 8), (3, 2, 16), (2, 2, 1)])
def test_forward_fct_passes_on_edge_shapes(nx, ny, nz):
    """The library's forward FCT passes (mixed precision: later passes start from
    the previous orbit plus the propagated change) against the oracle's passes."""
    import torch

    from oracle import drivers_ref
    from orbit_amd import drivers
    wl = synth.make_workload(nx, ny, nz, 2, seed=[11, nx, ny, nz], convection=nz > 1, wind_std=6.0)
    e = wl.e.copy()
    e[:, :, : nx * ny] += 1e-6
    op = OperatorSet.from_workload(wl)
    grids = drivers.make_fct_grids(wl.grids, op.ctx)
    got, hist = drivers.forward_fct(op, grids, torch.from_numpy(e).cuda(), n_passes=3, tol=1e-8, precision="mixed")
    got = got.cpu().numpy()
    for s in range(2):
        lus = [orb.factor_one_bin(wl.L(t, s)) for t in range(8)]
        ref, _ = drivers_ref.forward_fct_passes(lus, list(e[s]), wl.grids, n_passes=3, tol=1e-12, maxiter=500)
        assert all(h[s]["converged"] for h in hist)
        assert _rel(got[s, 0], ref[0]) < 1e-6 and _rel(got[s, 5], ref[5]) < 1e-6, s
