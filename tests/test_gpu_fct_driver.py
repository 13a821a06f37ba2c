"""The forward FCT passes inside the library (orbit_forward_fct_dev, reference
orbit/cli.py:1212-1241) against the same passes driven from Python over
orbit_fct_source_batch_dev and orbit_solve_forward_dev: the same launches, so
the same bits -- orbits and every pass's solve information -- with no ghost
and with a ghost for every tracer, in both precisions; one pass is the plain
periodic solve; bad arguments are refused."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from orbit_amd import drivers, synth
from orbit_amd.core import OperatorSet


@pytest.fixture(scope="module")
def month():
    wl = synth.make_workload(12, 10, 6, 3, seed=91, convection=True, wind_std=7.0)
    op = OperatorSet.from_workload(wl)
    grids = drivers.make_fct_grids(wl.grids, op.ctx)
    e = torch.from_numpy(wl.e).cuda()
    ghost = torch.from_numpy(np.abs(np.random.default_rng(2).normal(1.0, 0.3, wl.e.shape)) * 1e-3).cuda()
    return wl, op, grids, e, ghost


@pytest.mark.parametrize("precision,tol", [("mixed", 1e-8), ("fp64", 1e-10)])
@pytest.mark.parametrize("with_ghost", [False, True])
def test_native_passes_equal_python_passes(month, precision, tol, with_ghost):
    wl, op, grids, e, ghost = month
    g = ghost if with_ghost else None
    a, ha = drivers.forward_fct(op, grids, e, n_passes=4, tol=tol, precision=precision, ghost=g, native=True)
    b, hb = drivers.forward_fct(op, grids, e, n_passes=4, tol=tol, precision=precision, ghost=g, native=False)
    assert torch.equal(a, b)
    assert len(ha) == len(hb) == 4
    for pa, pb in zip(ha, hb):
        for ia, ib in zip(pa, pb):
            assert ia == ib
    assert all(i["converged"] for i in ha[-1])


def test_one_pass_is_the_periodic_solve(month):
    wl, op, grids, e, _ = month
    a, h = drivers.forward_fct(op, grids, e, n_passes=1, tol=1e-10, precision="fp64")
    b, ib = op.solve_forward(e, tol=1e-10, precision="fp64")
    assert torch.equal(a, b) and h[0] == ib


def test_native_driver_refuses_bad_arguments(month):
    wl, op, grids, e, _ = month
    with pytest.raises(ValueError):
        drivers.forward_fct(op, grids[:7], e, n_passes=2)
    with pytest.raises(ValueError):
        drivers.forward_fct(op, grids, e[:, :7], n_passes=2)
    with pytest.raises(ValueError):
        drivers.forward_fct(op, grids, e, n_passes=2, ghost=e[:, :, :10])
