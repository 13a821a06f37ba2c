"""Over-relaxed red-black sweeps: x += omega P (b - A x).

A x = b is the fixed point for any omega, so the relaxation an operator set
chooses at creation changes the sweep counts, not the answer: the orbit with
the chosen omega and with plain block Gauss-Seidel (omega = 1) agree to the
solve tolerance and both match the CPU oracle; choosing again is deterministic
(bit-identical solves); an omega outside [0.5, 1.5] is refused and leaves the
operator set as it was."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import orbit_ref as orb
from orbit_amd import synth
from orbit_amd._lib import OrbitError
from orbit_amd.core import OperatorSet


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def test_relaxation_changes_counts_not_the_orbit():
    wl = synth.make_workload(12, 10, 6, 3, seed=81, convection=True, wind_std=8.0)
    op = OperatorSet.from_workload(wl)
    st = op.stats()
    w = st["relaxation"]
    assert 1.0 <= w <= 1.09
    a, ia = op.solve_forward(wl.e, tol=1e-11, precision="fp64")
    am, _ = op.solve_forward(wl.e, tol=1e-8, precision="mixed")
    gs = op.set_relaxation(1.0)
    assert gs["relaxation"] == 1.0
    if w != 1.0:  # chosen only when it contracts >= 5% better
        assert gs["sweep_contraction"] > st["sweep_contraction"]
        assert gs["default_acc_sweeps"] >= st["default_acc_sweeps"]
    b, ib = op.solve_forward(wl.e, tol=1e-11, precision="fp64")
    for s in range(wl.n_tracers):
        assert ia[s]["converged"] and ib[s]["converged"]
        assert _rel(a[s], b[s]) < 1e-9, s
        lus = [orb.factor_one_bin(wl.L(t, s)) for t in range(8)]
        ref, _ = orb.solve_orbit_one_species(lus, list(wl.e[s]), tol=1e-12, maxiter=500)
        assert _rel(a[s, 0], ref[0]) < 1e-8 and _rel(b[s, 0], ref[0]) < 1e-8, s
        assert _rel(am[s, 0], ref[0]) < 1e-6, s
    # choosing again is deterministic
    again = op.set_relaxation(0.0)
    assert again["relaxation"] == w and again["default_acc_sweeps"] == st["default_acc_sweeps"]
    c, _ = op.solve_forward(wl.e, tol=1e-11, precision="fp64")
    assert np.array_equal(a, c)
    # out of range: refused, the opset unchanged
    with pytest.raises(OrbitError):
        op.set_relaxation(2.0)
    assert op.stats()["relaxation"] == w
    d, _ = op.solve_forward(wl.e, tol=1e-11, precision="fp64")
    assert np.array_equal(a, d)


def test_adjoint_with_chosen_relaxation_matches_oracle():
    wl = synth.make_workload(12, 10, 6, 2, seed=82, convection=True)
    op = OperatorSet.from_workload(wl)
    R = np.random.default_rng(5).random((2, 8, wl.N))
    lam, info = op.solve_adjoint(R, tol=1e-11, precision="fp64")
    for s in range(2):
        lus = [orb.factor_one_bin(wl.L(t, s)) for t in range(8)]
        ref, _ = orb.solve_orbit_adjoint_one_species(lus, list(R[s]), tol=1e-12, maxiter=500)
        assert info[s]["converged"] and _rel(lam[s, 0], ref[0]) < 1e-8, s


def test_calibration_samples_the_weakest_tracer():
    """With more tracers than the calibration measures, the ones with the
    smallest diagonal (the weakest loss contracts slowest) are always among
    them: at a common omega, a set of 20 whose tracer 13 has 1000x less loss
    measures at least the contraction of tracer 13 alone (the same arrays and
    per-slot arithmetic, so the same number when it is sampled), and that
    tracer's mixed solve still matches the oracle."""
    wl = synth.make_workload(12, 10, 6, 20, seed=83, convection=True)
    wl.D[:, 13] *= 1e-3
    op = OperatorSet.from_workload(wl)
    one = OperatorSet.from_shared(wl.T, wl.D[:, 13:14], wl.nz, wl.ny, wl.nx, wl.dtau)
    r_all = op.set_relaxation(1.0)["sweep_contraction"]
    r_one = one.set_relaxation(1.0)["sweep_contraction"]
    assert r_all >= r_one
    op.set_relaxation(0.0)
    orbit, info = op.solve_forward(wl.e[[13]], tracers=[13], tol=1e-8, precision="mixed")
    lus = [orb.factor_one_bin(wl.L(t, 13)) for t in range(8)]
    ref, _ = orb.solve_orbit_one_species(lus, list(wl.e[13]), tol=1e-12, maxiter=500)
    assert info[0]["converged"] and _rel(orbit[0, 0], ref[0]) < 1e-6


def test_changing_relaxation_recaptures_the_sweep_graphs():
    """The mixed path replays captured sweep graphs, which hold the relaxation
    of their launches: after set_relaxation the next solve must use the new
    one.  A set that solved (and captured) with its chosen omega and then
    switched to omega = 1 gives exactly the solve of a fresh set switched to
    omega = 1 before any capture."""
    wl = synth.make_workload(12, 10, 6, 2, seed=84, convection=True, wind_std=8.0)
    a = OperatorSet.from_workload(wl)
    if a.stats()["relaxation"] == 1.0:
        pytest.skip("this operator keeps omega = 1")
    g = dict(tol=1e-8, precision="mixed", use_graph=True)
    a.solve_forward(wl.e, **g)  # captures with the chosen omega
    a.set_relaxation(1.0)
    ra, _ = a.solve_forward(wl.e, **g)
    b = OperatorSet.from_workload(wl)
    b.set_relaxation(1.0)
    rb, _ = b.solve_forward(wl.e, **g)
    assert np.array_equal(ra, rb)
