"""GPU parity of the batched FCT source (orbit_fct_source_batch_dev).

The batched source is the one the forward deferred-correction passes call: one
launch over every tracer of a bin, on strided rows (orbit[:, tau+1]), adding a
base (the emissions) in place.  Checked per vector against the CPU oracle
(reference orbit/core/fct.py:459-492) at 1e-11 relative on grids whose tiles of
64 x 8 cells are partial (64x70: 9 row tiles, the last one 6 rows) or several
per row (130x17: 3 column tiles, the last one 2 columns), with and without the
limiter and a ghost; and the overlap rule of the entry point."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from oracle import fct_ref
from orbit_amd import synth
from orbit_amd._lib import OrbitError
from orbit_amd.core import Context
from orbit_amd.fct import FctGrid


def _case(nx, ny, nv, seed):
    wl = synth.make_workload(nx, ny, 15, 1, seed=seed)
    g = wl.grids[3]
    rng = np.random.default_rng(seed)
    c = wl.rho * (1.0 + rng.random((nv, wl.N)))
    for v in range(nv):  # a front per vector, at different places
        c[v].reshape(wl.nz, wl.ny, wl.nx)[:, v:v + ny // 3, 2 * v:2 * v + nx // 3] *= 4.0 + v
    ghost = wl.rho * (1.0 + rng.random((nv, wl.N)))
    base = rng.standard_normal((nv, wl.N))
    return wl, g, c, ghost, base


@pytest.mark.parametrize("nx,ny", [(64, 70), (130, 17)])
def test_batched_source_matches_oracle(nx, ny):
    nv = 3
    wl, g, c, ghost, base = _case(nx, ny, nv, seed=nx + ny)
    ctx = Context(0)
    G = FctGrid(g, ctx)
    dt = 10800.0
    # strided rows, as orbit[:, tau + 1] and e_eff[:, tau] are
    C = torch.zeros((nv, 3, wl.N), dtype=torch.float64, device="cuda")
    C[:, 1] = torch.from_numpy(c).cuda()
    Gh = torch.zeros((nv, 3, wl.N), dtype=torch.float64, device="cuda")
    Gh[:, 1] = torch.from_numpy(ghost).cuda()
    B = torch.from_numpy(base).cuda()
    for limited in (True, False):
        for use_ghost in (False, True):
            out = torch.full((nv, 2, wl.N), float("nan"), dtype=torch.float64, device="cuda")
            G.source_batch(C[:, 1], out[:, 0], base=B, dt=dt, limited=limited,
                           ghost=Gh[:, 1] if use_ghost else None)
            got = out[:, 0].cpu().numpy()
            assert torch.isnan(out[:, 1]).all()  # the rows between stay untouched
            for v in range(nv):
                ref = fct_ref.compute_horizontal_fct_source(g, c[v], dt, limited=limited,
                                                            c_ghost_flat=ghost[v] if use_ghost else None)
                np.testing.assert_allclose(got[v] - base[v], ref, rtol=1e-11,
                                           atol=1e-12 * np.abs(ref).max() + 1e-14 * np.abs(base[v]).max(),
                                           err_msg=f"v={v} limited={limited} ghost={use_ghost}")
    # base may be the output itself (in place)
    out = B.clone()
    G.source_batch(C[:, 1], out, base=out, dt=dt)
    ref = B.clone()
    G.source_batch(C[:, 1], ref, base=B, dt=dt)
    assert torch.equal(out, ref)
    G.close()


def test_batched_source_rejects_overlapping_output():
    wl, g, c, ghost, base = _case(24, 20, 2, seed=5)
    ctx = Context(0)
    G = FctGrid(g, ctx)
    X = torch.from_numpy(c).cuda()
    with pytest.raises(OrbitError, match="overlap"):
        G.source_batch(X, X)  # out == c
    Y = torch.zeros((2, 2, wl.N), dtype=torch.float64, device="cuda")
    Y[:, 0] = X
    with pytest.raises(OrbitError, match="overlap"):
        G.source_batch(Y[:, 0], Y[:, 1])  # interleaved rows of one tensor: spans overlap
    G.close()
