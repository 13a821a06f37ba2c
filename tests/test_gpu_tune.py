"""Sweep launch shapes and the per-size-class autotuner.

Every launch shape of the half-sweep (slots per wave, threads per workgroup,
XCD-contiguous or linear block order) computes the same per-lane arithmetic, so
a solve gives bit-identical results under any shape table; the autotuner only
changes timing.  Checked on forward and adjoint mixed solves of 5 slots (the
4-7 bucket, where the four-slot mapping is eligible) and 1 slot.
"""
import itertools

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from orbit_amd import _lib, serve, synth
from orbit_amd.core import OperatorSet


@pytest.fixture(scope="module")
def small():
    wl = synth.make_workload(14, 9, 6, 3, seed=11, convection=True)
    op = OperatorSet.from_workload(wl)
    return wl, op


def _solve(wl, op):
    e5 = np.concatenate([wl.e, wl.e[:2]])
    tr = [0, 1, 2, 0, 1]
    f5, _ = op.solve_forward(e5, tracers=tr, tol=1e-8, precision="mixed")
    a5, _ = op.solve_adjoint(e5, tracers=tr, tol=1e-8, precision="mixed")
    f1, _ = op.solve_forward(wl.e[2:3], tracers=[2], tol=1e-8, precision="fp64")
    return f5, a5, f1


def test_every_shape_gives_identical_bits(small):
    wl, op = small
    default = op.get_shapes()
    ref = _solve(wl, op)
    for spw, threads, xcd in itertools.product((1, 4), (64, 128, 256), (0, 1)):
        op.set_shapes([(spw, threads, xcd)] * 15)
        got = _solve(wl, op)
        for a, b in zip(got, ref):
            assert np.array_equal(a, b), (spw, threads, xcd)
    op.set_shapes(default)
    assert op.get_shapes() == default


def test_autotune_picks_valid_shapes_and_keeps_results(small):
    wl, op = small
    ref = _solve(wl, op)
    table, us = op.autotune(reps=10)
    assert len(table) == 15 and len(us) == 15
    for spw, threads, xcd in table:
        assert spw in (1, 4) and threads in (64, 128, 256) and xcd in (0, 1)
    # one slot never takes the four-slot mapping
    for p in range(3):
        assert table[p * 5][0] == 1
    assert all(0 < u < 1e4 for u in us)
    assert op.get_shapes() == table
    got = _solve(wl, op)
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)


def test_tuner_shares_a_table_within_a_size_class(small):
    wl, op = small
    op2 = OperatorSet.from_workload(synth.make_workload(14, 9, 6, 2, seed=12, convection=True))
    tuner = serve.ShapeTuner(reps=5)
    t1 = tuner.apply(op)
    t2 = tuner.apply(op2)
    assert t1 == t2 == op2.get_shapes()
    rep = tuner.report()
    assert list(rep) == ["6x9x14"] and set(rep["6x9x14"]) == set(OperatorSet.SWEEP_PRECS)


def test_set_shapes_rejects_bad_shapes(small):
    _, op = small
    with pytest.raises(_lib.OrbitError):
        op.set_shapes([(2, 256, 1)] * 15)
    with pytest.raises(_lib.OrbitError):
        op.set_shapes([(1, 96, 1)] * 15)
    with pytest.raises(ValueError):
        op.set_shapes([(1, 256, 1)] * 3)
