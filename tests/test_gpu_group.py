"""Group solves (orbit_solve_group_dev): slots of several operator sets of
different sub-domain sizes in one batch, every launch covering all of them.
Each slot's orbit and solve information is bit-identical to its set solved
alone (forward and adjoint, mixed and FP64), the 144-orbit uneven batch
(12 months x 12 tracers over 24x20 / 48x40 / 64x70 x 15) comes out of one call
equal to the per-set solves, and sets the group path cannot take are refused."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from orbit_amd import _lib, serve, synth
from orbit_amd.core import OperatorSet


@pytest.fixture(scope="module")
def mix():
    jobs = []
    for month, (nx, ny, nt) in enumerate(((24, 20, 5), (12, 10, 3), (48, 40, 6))):
        wl = synth.make_workload(nx, ny, 15, nt, seed=[7, month], convection=True)
        op = OperatorSet.from_workload(wl)
        jobs.append((wl, op))
    return jobs


@pytest.mark.parametrize("precision,tol", [("mixed", 1e-8), ("fp64", 1e-9)])
def test_group_equals_each_set_alone(mix, precision, tol):
    rng = np.random.default_rng(3)
    fwd, adj = [], []
    for wl, op in mix:
        tr = sorted(rng.choice(wl.n_tracers, max(1, wl.n_tracers - 1), replace=False).tolist())
        fwd.append((op, torch.from_numpy(np.ascontiguousarray(wl.e[tr])).cuda(), tr))
        R = np.zeros((len(tr), 8, wl.N))
        for q in range(len(tr)):
            R[q, 0, rng.choice(wl.nx * wl.ny, 20, replace=False)] = rng.random(20)
        adj.append((op, torch.from_numpy(R).cuda(), tr))
    for adjoint, jobs in ((False, fwd), (True, adj)):
        got = serve.solve_group(jobs, precision=precision, tol=tol, adjoint=adjoint)
        for (op, inp, tr), (orbit, info) in zip(jobs, got):
            fn = op.solve_adjoint if adjoint else op.solve_forward
            alone, ainfo = fn(inp, tr, tol=tol, precision=precision)
            print(f"{'adjoint' if adjoint else 'forward'} {precision} {op.nx}x{op.ny} tracers {tr}: "
                  f"iters {[i['gmres_iters'] for i in info]}", flush=True)
            assert torch.equal(orbit, alone), (adjoint, op.nx, tr)
            assert info == ainfo
            assert all(i["converged"] for i in info)


def test_uneven_144_orbits_in_one_call():
    jobs = []
    sizes = ("24x20", "48x40", "64x70")
    for month in range(12):
        wlm = synth.benchmark_workload("uneven-batch", n_tracers=12, month=month, size=sizes[month % 3])
        opm = OperatorSet.from_workload(wlm)
        jobs.append((opm, torch.from_numpy(wlm.e).cuda(), list(range(12))))
    got = serve.solve_group(jobs, precision="mixed", tol=1e-8)
    assert sum(o.shape[0] for o, _ in got) == 144
    for (op, e, tr), (orbit, info) in zip(jobs[:3] + jobs[-2:], got[:3] + got[-2:]):
        alone, _ = op.solve_forward(e, tr, tol=1e-8, precision="mixed")
        assert torch.equal(orbit, alone), op.nx
        assert all(i["converged"] for i in info)


def test_group_refuses_what_it_cannot_take(mix):
    wl, op = mix[0]
    e = torch.from_numpy(wl.e).cuda()
    odd = synth.make_workload(9, 6, 4, 1, seed=5, periodic=True)
    op_odd = OperatorSet.from_workload(odd)
    with pytest.raises(_lib.OrbitError, match="odd nx"):
        serve.solve_group([(op, e[:1], [0]), (op_odd, torch.from_numpy(odd.e).cuda(), [0])])
    with pytest.raises(TypeError):
        serve.solve_group([(op, e[:1].float(), [0])])
    with pytest.raises(_lib.OrbitError, match="tracer index"):
        serve.solve_group([(op, e[:1], [wl.n_tracers])])
    # still usable afterwards
    got = serve.solve_group([(op, e[:2].contiguous(), [0, 1])])
    assert all(i["converged"] for i in got[0][1])
