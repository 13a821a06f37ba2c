"""The benchmark workloads by kind: each kind's seed tag is fixed (the same
operators and emissions every run), sizes and tracer counts follow the kind,
and an unknown kind is refused."""
import numpy as np
import pytest

from orbit_amd import synth


def test_kinds_are_seeded_and_sized():
    a = synth.benchmark_workload("single")
    b = synth.make_workload(24, 20, 15, 1, seed=[0, 0, 0], convection=False)
    assert (a.nx, a.ny, a.nz, a.n_tracers) == (24, 20, 15, 1)
    np.testing.assert_array_equal(a.e, b.e)
    np.testing.assert_array_equal(a.D, b.D)
    m = synth.benchmark_workload("uneven-batch", n_tracers=2, month=3, size="24x20")
    assert (m.nx, m.ny, m.n_tracers) == (24, 20, 2)
    tags = synth.WORKLOAD_KINDS
    assert len(set(tags.values())) == len(tags)


def test_unknown_kind_is_refused():
    with pytest.raises(ValueError):
        synth.benchmark_workload("config1")
