"""The forward FCT passes inside the library (orbit_forward_fct_dev, reference
orbit/cli.py:1212-1241) against the same passes driven from Python over
orbit_fct_source_batch_dev and orbit_solve_forward_dev.  FP64: the same
launches, so the same bits -- orbits and every pass's solve information.
Mixed precision: the library starts each pass after the first from the
previous orbit plus the FP32 propagation of the forcing's change, the Python
loop from the previous c0; both converge every pass and agree to the 1e-6 bar,
and the library's passes agree with its FP64 passes to 1e-6; a tracer's passes
are bit-identical alone and in the batch.  With no ghost and with a ghost for
every tracer; one pass is the plain periodic solve; bad arguments are refused."""
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


def _rel(a, b):
    return float((a - b).abs().max() / b.abs().max())


@pytest.mark.parametrize("with_ghost", [False, True])
def test_native_fp64_passes_equal_python_passes(month, with_ghost):
    wl, op, grids, e, ghost = month
    g = ghost if with_ghost else None
    a, ha = drivers.forward_fct(op, grids, e, n_passes=4, tol=1e-10, precision="fp64", ghost=g, native=True)
    b, hb = drivers.forward_fct(op, grids, e, n_passes=4, tol=1e-10, precision="fp64", ghost=g, native=False)
    assert torch.equal(a, b)
    assert len(ha) == len(hb) == 4
    for pa, pb in zip(ha, hb):
        for ia, ib in zip(pa, pb):
            assert ia == ib
    assert all(i["converged"] for i in ha[-1])


@pytest.mark.parametrize("with_ghost", [False, True])
def test_native_mixed_passes_match_python_and_fp64(month, with_ghost):
    wl, op, grids, e, ghost = month
    g = ghost if with_ghost else None
    a, ha = drivers.forward_fct(op, grids, e, n_passes=4, tol=1e-8, precision="mixed", ghost=g, native=True)
    b, hb = drivers.forward_fct(op, grids, e, n_passes=4, tol=1e-8, precision="mixed", ghost=g, native=False)
    f, _ = drivers.forward_fct(op, grids, e, n_passes=4, tol=1e-11, precision="fp64", ghost=g, native=True)
    assert len(ha) == len(hb) == 4
    for h in ha:
        assert all(i["converged"] and i["residual"] <= 1e-8 for i in h)
    for t in range(e.shape[0]):
        assert _rel(a[t, 0], b[t, 0]) <= 1e-6, t
        assert _rel(a[t], f[t]) <= 1e-6, t
    # per-slot arithmetic: the same passes alone and in the batch
    alone, _ = drivers.forward_fct(op, grids, e[1:2], tracers=[1], n_passes=4, tol=1e-8, precision="mixed",
                                   ghost=None if g is None else g[1:2], native=True)
    assert torch.equal(alone[0], a[1])


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


def test_mixed_passes_with_a_zero_tracer_and_a_small_krylov_space(month):
    """A tracer with no emissions stays exactly zero through every pass (its
    forcing change is zero, so is its s~) beside tracers that are not; with a
    Krylov space of 3 the passes still converge, agree with the Python loop to
    the 1e-6 bar and give the zero tracer's neighbours the same bits as
    without it."""
    wl, op, grids, e, _ = month
    ez = e.clone()
    ez[2] = 0.0
    a, ha = drivers.forward_fct(op, grids, ez, n_passes=3, tol=1e-8, precision="mixed", restart=3)
    b, _ = drivers.forward_fct(op, grids, ez, n_passes=3, tol=1e-8, precision="mixed", restart=3, native=False)
    assert not torch.any(a[2])
    for h in ha:
        assert h[2]["zero_source"] == 1
        assert all(i["converged"] for q, i in enumerate(h) if q != 2)
    for t in (0, 1):
        assert _rel(a[t, 0], b[t, 0]) <= 1e-6, t
    pair, _ = drivers.forward_fct(op, grids, ez[[0, 1]], tracers=[0, 1], n_passes=3, tol=1e-8,
                                  precision="mixed", restart=3)
    assert torch.equal(pair, a[[0, 1]])


def test_mixed_pass_not_converging_is_a_warning(month):
    """Iteration caps reached inside the passes give the warning status (the
    reference prints and goes on): finite orbits, converged = 0, no error."""
    wl, op, grids, e, _ = month
    a, h = drivers.forward_fct(op, grids, e, n_passes=3, tol=1e-14, precision="mixed", maxiter=1, restart=2)
    assert torch.isfinite(a).all()
    assert any(not i["converged"] for p in h for i in p)


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("mixed", 1e-8)])
def test_per_tracer_ghosts_run_natively(month, precision, tol):
    """A ghost for some tracers only (the lateral boundary file of some species):
    the library takes one tensor with per-tracer flags and reads the ghost rows
    of the flagged tracers only.  FP64: the same bits as the Python loop's two
    groups of launches; mixed: the 1e-6 bar against it; the tracers without a
    ghost get the same bits as in a run with no ghost at all."""
    wl, op, grids, e, ghost = month
    gl = [None, ghost[1], None]
    a, ha = drivers.forward_fct(op, grids, e, n_passes=3, tol=tol, precision=precision, ghost=gl)
    b, _ = drivers.forward_fct(op, grids, e, n_passes=3, tol=tol, precision=precision, ghost=gl, native=False)
    plain, _ = drivers.forward_fct(op, grids, e, n_passes=3, tol=tol, precision=precision)
    assert all(i["converged"] for h in ha for i in h)
    if precision == "fp64":
        assert torch.equal(a, b)
    else:
        for t in range(3):
            assert _rel(a[t, 0], b[t, 0]) <= 1e-6, t
    assert torch.equal(a[0], plain[0]) and torch.equal(a[2], plain[2])
    assert not torch.equal(a[1], plain[1])


def test_python_loop_warm_orbit_mode(month):
    """warm="orbit" (the Python-driven passes, each started from every bin
    state of the whole previous orbit): converged
    passes within the 1e-6 bar of the library's; and a solve warm-started from
    its own converged orbit stays on it."""
    wl, op, grids, e, _ = month
    a, ha = drivers.forward_fct(op, grids, e, n_passes=3, tol=1e-8, precision="mixed", warm="orbit")
    b, _ = drivers.forward_fct(op, grids, e, n_passes=3, tol=1e-8, precision="mixed")
    assert all(i["converged"] for h in ha for i in h)
    for t in range(e.shape[0]):
        assert _rel(a[t, 0], b[t, 0]) <= 1e-6, t
    c, ic = op.solve_forward(e, tol=1e-8, precision="mixed")
    d, idd = op.solve_forward(e, c_warm=c, tol=1e-8, precision="mixed")
    for t in range(e.shape[0]):
        assert idd[t]["converged"] and idd[t]["gmres_iters"] <= ic[t]["gmres_iters"], t
        assert _rel(d[t], c[t]) <= 1e-7, t
