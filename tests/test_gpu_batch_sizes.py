"""Large batches: 40 tracers in one solve (the 16+ launch-shape bucket, ten
four-slot groups per wave row), against the same tracers alone and in a small
permuted batch, bit for bit, in both precisions and directions; a tracer listed
twice gives the same result in both slots; and the batch agrees with the CPU
oracle on a sample of tracers."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import orbit_ref as orb
from orbit_amd import synth
from orbit_amd.core import OperatorSet

NT = 40


@pytest.fixture(scope="module")
def many():
    wl = synth.make_workload(12, 10, 6, NT, seed=71, convection=True)
    op = OperatorSet.from_workload(wl)
    return wl, op


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-10), ("mixed", 1e-8)])
def test_forty_tracers_alone_equals_batch(many, prec, tol):
    wl, op = many
    full, info = op.solve_forward(wl.e, tol=tol, precision=prec)
    assert all(i["converged"] for i in info)
    for t in (0, 17, 39):
        alone, _ = op.solve_forward(wl.e[t:t + 1], tracers=[t], tol=tol, precision=prec)
        assert np.array_equal(alone[0], full[t]), (prec, t)
    sub = [39, 3, 17, 3]  # permuted, with a repeat
    part, _ = op.solve_forward(wl.e[sub], tracers=sub, tol=tol, precision=prec)
    for q, t in enumerate(sub):
        assert np.array_equal(part[q], full[t]), (prec, q, t)
    R = np.random.default_rng(3).random((NT, 8, wl.N))
    lam, _ = op.solve_adjoint(R, tol=tol, precision=prec)
    la, _ = op.solve_adjoint(R[[25]], tracers=[25], tol=tol, precision=prec)
    assert np.array_equal(la[0], lam[25]), prec


def test_forty_tracer_batch_matches_oracle(many):
    wl, op = many
    full, _ = op.solve_forward(wl.e, tol=1e-8, precision="mixed")
    for t in (5, 31):
        lus = [orb.factor_one_bin(wl.L(b, t)) for b in range(8)]
        ref, _ = orb.solve_orbit_one_species(lus, list(wl.e[t]), tol=1e-12, maxiter=500)
        assert np.abs(full[t, 0] - ref[0]).max() <= 1e-6 * np.abs(ref[0]).max(), t
