"""The first sweep of every FP32 inner solve in one launch over column
patches (orbit_opset_set_patch) gives the same bits as its two half-sweep
launches: every patch shape, on grids whose patches the grid's edge cuts in x
and in y, with and without periodic x (with nx = 4 periodic, where a patch's
west and east ring columns are its own columns across the wrap face), one to
twelve slots, forward and adjoint mixed solves and a permuted tracer list,
with three sweeps per inner solve and with one (where the first sweep is
also the last and forms the next bin's right-hand side).  The sweep
approximates the reference's inner solve of (I + dtau L) x = b
(orbit/core/orbit.py:79-96 factors that matrix); the half-sweep path that the
other tests pin against the oracle is the reference here."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import torch

from orbit_amd import serve, synth
from orbit_amd.core import OperatorSet

SHAPES = [1, 2, 3, 4]


@pytest.mark.parametrize("nx,ny,nz,periodic", [(10, 8, 6, False), (12, 7, 5, True), (6, 5, 4, False),
                                               (4, 6, 4, True), (2, 6, 4, False), (9, 11, 5, False), (24, 20, 15, True)])
def test_patch_sweeps_give_the_half_sweeps_bits(nx, ny, nz, periodic):
    wl = synth.make_workload(nx, ny, nz, 5, seed=nx * 100 + ny, convection=True, periodic=periodic)
    op = OperatorSet.from_workload(wl)
    st = op.stats()
    assert st["periodic_x"] == int(periodic)
    print(f"grid {nx}x{ny}x{nz} periodic {periodic}: default patch shape {st['patch_shape']}", flush=True)
    R = np.random.default_rng(3).random(wl.e.shape)
    perm = [3, 0, 4, 1, 2]
    op.set_patch(0)
    ref_f, _ = op.solve_forward(wl.e, tol=1e-8, precision="mixed")
    ref_a, _ = op.solve_adjoint(R, tol=1e-8, precision="mixed")
    ref_1, _ = op.solve_forward(wl.e[:1], tracers=[0], tol=1e-8, precision="mixed")
    ref_p, _ = op.solve_forward(wl.e[perm], tracers=perm, tol=1e-8, precision="mixed")
    ref_s1, _ = op.solve_forward(wl.e, tol=1e-3, precision="mixed", inner_sweeps=1, check_sweeps=1)
    for shape in SHAPES:
        print(f"  patch shape {shape}", flush=True)
        op.set_patch(shape)
        s1, _ = op.solve_forward(wl.e, tol=1e-3, precision="mixed", inner_sweeps=1, check_sweeps=1)
        assert np.array_equal(s1, ref_s1), shape
        f, info = op.solve_forward(wl.e, tol=1e-8, precision="mixed")
        assert all(i["converged"] for i in info)
        assert np.array_equal(f, ref_f), shape
        a, _ = op.solve_adjoint(R, tol=1e-8, precision="mixed")
        assert np.array_equal(a, ref_a), shape
        one, _ = op.solve_forward(wl.e[:1], tracers=[0], tol=1e-8, precision="mixed")
        assert np.array_equal(one, ref_1), shape
        p, _ = op.solve_forward(wl.e[perm], tracers=perm, tol=1e-8, precision="mixed")
        assert np.array_equal(p, ref_p), shape
    op.close()


def test_patch_sweeps_at_the_benchmark_shape_with_twelve_slots():
    """64 x 70 x 15, 12 slots, one bin's worth of solve per shape: the
    headline batch shape (three slot groups of four)."""
    wl = synth.benchmark_workload("fct-month", n_tracers=12)
    op = OperatorSet.from_workload(wl)
    op.set_patch(0)
    ref, _ = op.solve_forward(wl.e, tol=1e-8, precision="mixed")
    for shape in SHAPES:
        print(f"  patch shape {shape}", flush=True)
        op.set_patch(shape)
        f, _ = op.solve_forward(wl.e, tol=1e-8, precision="mixed")
        assert np.array_equal(f, ref), shape
    op.close()


def test_group_first_sweep_with_one_sweep_per_inner_solve():
    """A group solve's first sweep takes its patches from the group's table and
    each slot's sweep count from its set; with one sweep per inner solve it is
    also the last sweep and forms the next bin's right-hand side.  Sets of
    three sizes (one periodic in x, patches cut by the grid's edge), forward
    and adjoint: bit-identical to each set solved alone."""
    jobs = []
    for k, (nx, ny, nt, per) in enumerate(((10, 7, 3, False), (12, 9, 2, True), (6, 5, 4, False))):
        wl = synth.make_workload(nx, ny, 6, nt, seed=[31, k], convection=True, periodic=per)
        op = OperatorSet.from_workload(wl)
        jobs.append((wl, op))
    for adjoint in (False, True):
        work = []
        for wl, op in jobs:
            inp = wl.e if not adjoint else np.random.default_rng(op.nx).random(wl.e.shape)
            work.append((op, torch.from_numpy(np.ascontiguousarray(inp)).cuda(), list(range(wl.n_tracers))))
        for sweeps in (1, 3):
            print(f"adjoint {adjoint}, {sweeps} sweep(s) per inner solve", flush=True)
            got = serve.solve_group(work, precision="mixed", tol=1e-4, adjoint=adjoint, inner_sweeps=sweeps,
                                    check_sweeps=sweeps)
            for (op, inp, tr), (orbit, info) in zip(work, got):
                fn = op.solve_adjoint if adjoint else op.solve_forward
                alone, ainfo = fn(inp, tr, tol=1e-4, precision="mixed", inner_sweeps=sweeps, check_sweeps=sweeps)
                assert torch.equal(orbit, alone), (adjoint, sweeps, op.nx)
                assert info == ainfo
    for _, op in jobs:
        op.close()
