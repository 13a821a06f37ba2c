"""Bin hand-off (orbit_opset_set_handoff): each bin's last black FP32
half-sweep of a monodromy apply or a residual sweep runs inside the next bin's
first-sweep launch.  The result must carry the bits of the same sweeps
launched one half-sweep at a time: orbits and solve information, forward and
adjoint, one to twelve slots, a permuted tracer list, two to four sweeps per
inner solve (with one, or with four colours, no hand-off applies and the
setting must change nothing), grids whose patches the grid's edge cuts, a
periodic x, and the forward FCT passes of the mixed path (residual sweeps with
epilogues, later passes started from the previous orbit).  The sweeps
approximate the reference's inner solve of (I + dtau L) x = b
(orbit/core/orbit.py:79-96); the one-launch-per-half-sweep path that the other
tests pin against the oracle is the reference here."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from orbit_amd import drivers, synth
from orbit_amd.core import OperatorSet


def _solves(op, wl, R, perm, sweeps):
    inner, check = sweeps
    kw = dict(tol=1e-8, precision="mixed", inner_sweeps=inner, check_sweeps=check)
    out = [op.solve_forward(wl.e, **kw), op.solve_adjoint(R, **kw),
           op.solve_forward(wl.e[:1], tracers=[0], **kw), op.solve_forward(wl.e[perm], tracers=perm, **kw)]
    return out


@pytest.mark.parametrize("nx,ny,nz,periodic", [(10, 8, 6, False), (12, 7, 5, True), (9, 11, 5, False),
                                               (4, 6, 4, True), (7, 6, 4, True), (24, 20, 15, True)])
@pytest.mark.parametrize("sweeps", [(3, 3), (2, 2), (2, 4), (1, 1)])
def test_handoff_gives_the_half_sweeps_bits(nx, ny, nz, periodic, sweeps):
    wl = synth.make_workload(nx, ny, nz, 5, seed=nx * 37 + ny, convection=True, periodic=periodic)
    op = OperatorSet.from_workload(wl)
    print(f"grid {nx}x{ny}x{nz} periodic {periodic} sweeps {sweeps}: patch shape {op.stats()['patch_shape']}",
          flush=True)
    R = np.random.default_rng(5).random(wl.e.shape)
    perm = [3, 0, 4, 1, 2]
    op.set_handoff(False)
    ref = _solves(op, wl, R, perm, sweeps)
    op.set_handoff(True)
    got = _solves(op, wl, R, perm, sweeps)
    for k, ((a, ia), (b, ib)) in enumerate(zip(got, ref)):
        assert np.array_equal(a, b), k
        assert ia == ib, k
    assert all(i["converged"] for i in got[0][1])
    op.close()


def test_handoff_at_the_benchmark_shape_and_fct_passes():
    """64 x 70 x 15, 12 tracers: the mixed forward solve and three forward FCT
    passes (residual sweeps with their epilogues, incremental later passes)
    with and without the hand-off."""
    wl = synth.benchmark_workload("fct-month", n_tracers=12)
    op = OperatorSet.from_workload(wl)
    grids = drivers.make_fct_grids(wl.grids, op.ctx)
    e = torch.from_numpy(wl.e).cuda()
    res = {}
    for on in (False, True):
        op.set_handoff(on)
        print(f"hand-off {on}", flush=True)
        f, info = op.solve_forward(wl.e, tol=1e-8, precision="mixed")
        c, hist = drivers.forward_fct(op, grids, e, n_passes=3, tol=1e-8, precision="mixed")
        res[on] = (f, [i["residual"] for i in info], c.cpu().numpy(), repr(hist))
    assert np.array_equal(res[True][0], res[False][0])
    assert res[True][1] == res[False][1]
    assert np.array_equal(res[True][2], res[False][2])
    assert res[True][3] == res[False][3]
    op.close()


def test_group_handoff_gives_the_half_sweeps_bits():
    """Group solves (slots of several operator sets in one launch) hand off
    too, with each set's own sweep counts: a slot whose count ended its bin's
    solve before the last sweep of the launch sequence keeps the right-hand
    side in memory.  Sets of four sizes (one periodic, one with strong winds
    and so more sweeps), forward and adjoint, the sets' own sweep counts and
    two fixed pairs: bit-identical with the hand-off off, and to each set
    solved alone."""
    from orbit_amd import serve
    jobs = []
    for k, (nx, ny, nt, per, wind) in enumerate(((10, 7, 3, False, 5.0), (12, 9, 2, True, 5.0),
                                                 (6, 5, 4, False, 5.0), (12, 10, 3, False, 8.0))):
        wl = synth.make_workload(nx, ny, 6, nt, seed=[41, k], convection=True, periodic=per, wind_std=wind)
        op = OperatorSet.from_workload(wl)
        st = op.stats()
        print(f"set {nx}x{ny}: default sweeps inner {st['default_inner_sweeps']} check {st['default_check_sweeps']}",
              flush=True)
        jobs.append((wl, op))
    for adjoint in (False, True):
        work = []
        for wl, op in jobs:
            inp = wl.e if not adjoint else np.random.default_rng(op.nx).random(wl.e.shape)
            work.append((op, torch.from_numpy(np.ascontiguousarray(inp)).cuda(), list(range(wl.n_tracers))))
        for sweeps in ((0, 0), (2, 2), (3, 3)):
            print(f"adjoint {adjoint}, sweeps {sweeps}", flush=True)
            kw = dict(precision="mixed", tol=1e-8, adjoint=adjoint, inner_sweeps=sweeps[0], check_sweeps=sweeps[1])
            res = {}
            for on in (False, True):
                for _, op in jobs:
                    op.set_handoff(on)
                res[on] = serve.solve_group(work, **kw)
            for (op, inp, tr), (a, ia), (b, ib) in zip(work, res[True], res[False]):
                assert torch.equal(a, b), (adjoint, sweeps, op.nx)
                assert ia == ib
                fn = op.solve_adjoint if adjoint else op.solve_forward
                alone, ainfo = fn(inp, tr, tol=1e-8, precision="mixed", inner_sweeps=sweeps[0],
                                  check_sweeps=sweeps[1])
                assert torch.equal(a, alone), (adjoint, sweeps, op.nx)
                assert ia == ainfo
    for _, op in jobs:
        op.close()
