"""The adjoint mode's end-of-bin receptor pairing (reference
orbit/modes/adjoint.py:192-203, its test tests/test_release_1_0_4.py:376-434)
on the CPU: with J = sum_tau <r_tau, c[tau+1]> and the oracle's SuperLU
propagators of the reference test's random system (n = 40), the gradient
dtau (lam_tau - R_tau) of the adjoint solve with R = receptor_R_list(...)
matches finite differences of J (J is linear in e) to 1e-6, and the pre-1.0.4
pairing R_tau = r_tau misses them by more than 1e-2."""
import numpy as np
import scipy.sparse as sp

from oracle import orbit_ref as orb
from orbit_amd.core import N_BINS
from orbit_amd.orbit_adjoint import receptor_R_list


def _system():
    n = 40
    rng = np.random.default_rng(3)
    I_N = sp.eye(n, format="csc")
    lus = []
    for _ in range(N_BINS):
        L = sp.random(n, n, density=0.08, format="csc", random_state=rng) * 5e-5
        lus.append(orb.factor_one_bin((L + 1e-4 * I_N).tocsc()))
    rec = np.zeros((N_BINS, 2, n))
    for t in range(N_BINS):
        rec[t, 1] = rng.random(n) * (1 + 3 * np.sin(np.pi * t / N_BINS) ** 2)  # strongly diurnal
    e = [rng.standard_normal(n) for _ in range(N_BINS)]
    return lus, rec, e


def _J(lus, rec, e):
    orbit, _ = orb.solve_orbit_one_species(lus, e, tol=1e-13, maxiter=500)
    return float(sum(rec[t, 1] @ orbit[t + 1] for t in range(N_BINS)))


def _grad(lus, R):
    lam, _ = orb.solve_orbit_adjoint_one_species(lus, R, tol=1e-13, maxiter=500)
    return orb.gradient_per_emission_bin(lam, R)


def test_pairing_matches_finite_differences():
    lus, rec, e = _system()
    g = _grad(lus, receptor_R_list(rec, 1))
    J0 = _J(lus, rec, e)
    h = 1e-3
    for t, j in [(0, 3), (2, 17), (5, 8), (7, 30)]:
        ep = [x.copy() for x in e]
        ep[t][j] += h
        fd = (_J(lus, rec, ep) - J0) / h
        assert abs(g[t][j] - fd) <= 1e-6 * abs(fd), (t, j, g[t][j], fd)


def test_old_pairing_is_detectably_wrong():
    lus, rec, _ = _system()
    new = _grad(lus, receptor_R_list(rec, 1))
    old = _grad(lus, [rec[t, 1].copy() for t in range(N_BINS)])
    rel = max(np.abs(a - b).max() / np.abs(a).max() for a, b in zip(new, old))
    assert rel > 1e-2


def test_pairing_is_a_cyclic_shift():
    rec = np.arange(N_BINS * 3 * 2, dtype=float).reshape(N_BINS, 3, 2)
    R = receptor_R_list(rec, 2)
    assert len(R) == N_BINS
    for tau in range(N_BINS):
        np.testing.assert_array_equal(R[(tau + 1) % N_BINS], rec[tau, 2])
    R[0][0] = -1.0  # a copy: the caller's receptors are not aliased
    assert rec[N_BINS - 1, 2, 0] != -1.0
