"""The adjoint mode's end-of-bin receptor pairing (reference
orbit/modes/adjoint.py:192-203; its test tests/test_release_1_0_4.py:376-434)
through the device solver and the reference's call surface (8 HipLU factors,
solve_orbit_one_species, solve_orbit_adjoint_one_species,
gradient_per_emission_bin): with strongly diurnal receptors on the end-of-bin
states, the gradient matches finite differences of J = sum_tau <r_tau, c[tau+1]>
(linear in e) to 1e-6; the pre-1.0.4 pairing misses them by more than 1e-2; and
the mixed-precision gradient of the batched adjoint agrees with FP64 to 1e-6."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from orbit_amd import orbit as horb, orbit_adjoint as hadj, solve as hsolve, synth
from orbit_amd.core import N_BINS, OperatorSet


@pytest.fixture(scope="module")
def system():
    wl = synth.make_workload(12, 10, 6, 1, seed=95, convection=True, wind_std=6.0)
    lus = hsolve.factor_species([wl.L(t, 0) for t in range(N_BINS)], wl.nz, wl.ny, wl.nx, wl.dtau)
    rng = np.random.default_rng(96)
    rec = np.zeros((N_BINS, 1, wl.N))
    for t in range(N_BINS):
        rec[t, 0] = rng.random(wl.N) * (1 + 3 * np.sin(np.pi * t / N_BINS) ** 2)
    e = [rng.standard_normal(wl.N) for _ in range(N_BINS)]
    return wl, lus, rec, e


def _J(lus, rec, e):
    orbit, info = horb.solve_orbit_one_species(lus, e, c_warm=None, tol=1e-12, maxiter=500)
    return float(sum(rec[t, 0] @ orbit[t + 1] for t in range(N_BINS)))


def _grad(lus, R):
    lam, info = hadj.solve_orbit_adjoint_one_species(lus, R, tol=1e-12, maxiter=500)
    return hadj.gradient_per_emission_bin(lam, R)


def test_pairing_matches_finite_differences(system):
    wl, lus, rec, e = system
    g = _grad(lus, hadj.receptor_R_list(rec, 0))
    J0 = _J(lus, rec, e)
    h = 1.0  # J is linear in e
    for t, j in [(0, 3), (2, 170), (5, 81), (7, 300)]:
        ep = [x.copy() for x in e]
        ep[t][j] += h
        fd = (_J(lus, rec, ep) - J0) / h
        assert abs(g[t][j] - fd) <= 1e-6 * abs(fd), (t, j, g[t][j], fd)


def test_old_pairing_is_detectably_wrong(system):
    wl, lus, rec, _ = system
    new = _grad(lus, hadj.receptor_R_list(rec, 0))
    old = _grad(lus, [rec[t, 0].copy() for t in range(N_BINS)])
    rel = max(np.abs(a - b).max() / np.abs(a).max() for a, b in zip(new, old))
    assert rel > 1e-2


def test_batched_mixed_gradient_matches_fp64(system):
    wl, _, rec, _ = system
    op = OperatorSet.from_workload(wl)
    R = np.stack(hadj.receptor_R_list(rec, 0))[None]
    lam64, i64 = op.solve_adjoint(R, tol=1e-10, precision="fp64")
    lamm, im = op.solve_adjoint(R, tol=1e-8, precision="mixed")
    assert i64[0]["converged"] and im[0]["converged"]
    g64 = hadj.gradient_per_emission_bin(lam64[0], R[0], wl.dtau)
    gm = hadj.gradient_per_emission_bin(lamm[0], R[0], wl.dtau)
    for t in range(N_BINS):
        assert np.abs(gm[t] - g64[t]).max() <= 1e-6 * np.abs(g64[t]).max(), t
