"""Device wrappers refuse tensors the kernels would misread -- host memory, a
narrower dtype, another shape -- with a TypeError or ValueError before any
pointer reaches the library, and the context keeps working afterwards."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from orbit_amd import drivers, synth
from orbit_amd.core import OperatorSet
from orbit_amd.fct import apply_dev


@pytest.fixture(scope="module")
def small():
    wl = synth.make_workload(8, 6, 4, 2, seed=11)
    op = OperatorSet.from_workload(wl, None)
    grids = drivers.make_fct_grids(wl.grids, op.ctx)
    return wl, op, grids


def test_solve_inputs_are_checked(small):
    wl, op, _ = small
    e = torch.from_numpy(wl.e).cuda()
    with pytest.raises(TypeError):
        op.solve_forward(e.float(), precision="fp64")
    with pytest.raises(ValueError):
        op.solve_forward(e[:, :4], precision="fp64")
    # a host warm start is moved to the device, not read as a device pointer
    ref, _ = op.solve_forward(e, tol=1e-10, precision="fp64")
    got, info = op.solve_forward(e, c_warm=ref[:, 0].cpu(), tol=1e-10, precision="fp64")
    assert all(i["converged"] for i in info)
    assert float((got - ref).abs().max() / ref.abs().max()) < 1e-8


def test_fct_wrappers_are_checked(small):
    wl, op, grids = small
    e = torch.from_numpy(wl.e).cuda()
    orbit, _ = op.solve_forward(e, tol=1e-10, precision="fp64")
    c = orbit[:, 1]
    out = torch.empty((wl.n_tracers, wl.N), dtype=torch.float64, device="cuda")
    G = grids[0]
    with pytest.raises(ValueError):
        G.source_batch(c.cpu(), out)
    with pytest.raises(TypeError):
        G.source_batch(c.float(), out)
    with pytest.raises(ValueError):
        G.source_batch(c, out[:1])
    with pytest.raises(ValueError):
        G.source_batch(c, out, ghost=torch.zeros((wl.n_tracers, wl.N), dtype=torch.float64))
    with pytest.raises(ValueError):
        G.linop(c[0].cpu())
    L = G.linop(c[0].contiguous())
    y = torch.empty(wl.N, dtype=torch.float64, device="cuda")
    with pytest.raises(TypeError):
        apply_dev(L, c[0].contiguous().float(), y)
    with pytest.raises(ValueError):
        apply_dev(L, c[0].contiguous(), y[:-1])
    with pytest.raises(ValueError):
        drivers.forward_fct(op, grids, e.cpu(), n_passes=2)
    with pytest.raises(TypeError):
        drivers.forward_fct(op, grids, e.float(), n_passes=2)
    # everything still runs after the refusals
    G.source_batch(c, out, dt=op.dtau)
    apply_dev(L, c[0].contiguous(), y)
    assert torch.isfinite(out).all() and torch.isfinite(y).all()


def test_coupling_block_checks(small):
    import scipy.sparse as sp
    from orbit_amd.species import CouplingBlock
    wl, op, _ = small
    K = CouplingBlock(sp.identity(wl.N, format="csr"), op.ctx)
    x = torch.ones(wl.N, dtype=torch.float64, device="cuda")
    y = torch.zeros(wl.N, dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        K.axpy(x.cpu(), y, 1.0)
    with pytest.raises(ValueError):
        K.axpy(x[:-1], y, 1.0)
    K.axpy(x, y, 2.0)
    op.ctx.sync()
    assert np.allclose(y.cpu().numpy(), 2.0)
