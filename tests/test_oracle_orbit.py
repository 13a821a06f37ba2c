"""Pins of the CPU orbit/adjoint oracle (oracle/orbit_ref.py) and the mass-balance oracle.

Each check is chosen so a plausible mistake in the restatement (reversed bin
order, a missing dtau, P instead of P^T, a receptor one bin off) fails it:
brute force on tiny N, closed forms, periodicity, finite differences and the
dot-product identity of the reference's tests/test_orbit_adjoint.py.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import mass_ref, orbit_ref as orb
from orbit_amd import synth

N_BINS = orb.N_BINS
DTAU = orb.DTAU


def _random_lus(n, seed, density=0.08):
    rng = np.random.default_rng(seed)
    Ls, lus = [], []
    for _ in range(N_BINS):
        L = sp.random(n, n, density=density, format="csc", random_state=rng) * 5e-5
        L = (L + 1e-4 * sp.identity(n, format="csc")).tocsc()
        Ls.append(L)
        lus.append(orb.factor_one_bin(L))
    return Ls, lus, rng


def test_factor_solve_matches_dense_both_transposes():
    Ls, lus, rng = _random_lus(25, 1)
    b = rng.standard_normal(25)
    A = np.eye(25) + DTAU * Ls[3].toarray()
    np.testing.assert_allclose(lus[3].solve(b), np.linalg.solve(A, b), rtol=1e-12)
    np.testing.assert_allclose(lus[3].solve(b, trans="T"), np.linalg.solve(A.T, b), rtol=1e-12)
    with pytest.raises(ValueError):
        lus[3].solve(b, trans="H")


def test_monodromy_is_product_in_bin_order():
    Ls, lus, rng = _random_lus(12, 2)
    M = np.eye(12)
    for L in Ls:
        M = np.linalg.solve(np.eye(12) + DTAU * L.toarray(), M)
    np.testing.assert_allclose(orb.dense_monodromy(lus), M, rtol=1e-12, atol=1e-15)
    v = rng.standard_normal(12)
    np.testing.assert_allclose(orb.apply_monodromy_adjoint(v, lus), M.T @ v, rtol=1e-12)


def test_periodic_solve_matches_dense_solve_and_is_periodic():
    Ls, lus, rng = _random_lus(20, 3)
    e = [rng.standard_normal(20) for _ in range(N_BINS)]
    orbit, info = orb.solve_orbit_one_species(lus, e, tol=1e-13, maxiter=500)
    M = orb.dense_monodromy(lus)
    s = orb.compute_s(lus, e)
    c0 = np.linalg.solve(np.eye(20) - M, s)
    np.testing.assert_allclose(orbit[0], c0, rtol=1e-10, atol=1e-12 * np.abs(c0).max())
    assert info["residual"] < 1e-12 and info["periodicity"] < 1e-10
    # recovery: c_{tau+1} = A_tau^-1 (c_tau + dtau e_tau)
    for tau in range(N_BINS):
        A = np.eye(20) + DTAU * Ls[tau].toarray()
        np.testing.assert_allclose(A @ orbit[tau + 1], orbit[tau] + DTAU * e[tau],
                                   rtol=1e-10, atol=1e-10)


def test_steady_state_closed_form():
    """Time-invariant diagonal loss l and emission e give c = e/l for every bin
    (reference tests/test_simulation_modes.py:1239-1270)."""
    n, l = 8, 1.0 / DTAU
    L = l * sp.identity(n, format="csc")
    lus = [orb.factor_one_bin(L) for _ in range(N_BINS)]
    e = [np.full(n, l) for _ in range(N_BINS)]
    orbit, info = orb.solve_orbit_one_species(lus, e, tol=1e-12)
    for c in orbit:
        np.testing.assert_allclose(c, 1.0, rtol=1e-10)


def test_zero_source_gives_zero_orbit():
    _, lus, _ = _random_lus(10, 4)
    orbit, info = orb.solve_orbit_one_species(lus, [np.zeros(10)] * N_BINS)
    assert info["gmres_iters"] == 0 and all(not np.any(c) for c in orbit)
    adj, info = orb.solve_orbit_adjoint_one_species(lus, [np.zeros(10)] * N_BINS)
    assert all(not np.any(c) for c in adj)


def test_adjoint_matches_forward_finite_difference():
    n = 60
    Ls, lus, rng = _random_lus(n, 7, density=0.06)
    s_bins = [rng.standard_normal(n) ** 2 for _ in range(N_BINS)]
    e = [rng.standard_normal(n) for _ in range(N_BINS)]

    def J(e_list):
        orbit, _ = orb.solve_orbit_one_species(lus, e_list, tol=1e-12, maxiter=500)
        return float(sum(np.dot(s_bins[t], orbit[t]) for t in range(N_BINS)))

    adj, info = orb.solve_orbit_adjoint_one_species(lus, s_bins, tol=1e-12, maxiter=500)
    grad = orb.gradient_per_emission_bin(adj, s_bins)
    J0 = J(e)
    h = 1e-6
    for tau in (0, 3, 7):
        for j in rng.choice(n, size=3, replace=False):
            ep = [x.copy() for x in e]
            ep[tau][j] += h
            fd = (J(ep) - J0) / h
            assert abs(fd - grad[tau][j]) < 1e-4 * max(abs(fd), 1.0)
    assert info["periodicity"] < 1e-8


def test_dot_product_identity_on_synthetic_operator():
    wl = synth.make_workload(6, 5, 4, 1, seed=9, convection=True)
    lus = [orb.factor_one_bin(wl.L(t, 0)) for t in range(N_BINS)]
    rng = np.random.default_rng(0)
    s = rng.random(wl.N)
    r = rng.random(wl.N)
    M = orb.dense_monodromy(lus)
    c = np.linalg.solve(np.eye(wl.N) - M, s)
    R = [r] + [np.zeros(wl.N)] * (N_BINS - 1)
    adj, _ = orb.solve_orbit_adjoint_one_species(lus, R, tol=1e-13, maxiter=500)
    assert np.dot(r, c) == pytest.approx(np.dot(adj[0], s), rel=1e-10)


def test_synthetic_operator_is_m_matrix_and_mass_consistent():
    wl = synth.make_workload(8, 6, 5, 2, seed=1, convection=True)
    nz, ny, nx = wl.nz, wl.ny, wl.nx
    for tau in (0, 5):
        A = wl.A(tau, 1)
        off = A - sp.diags(A.diagonal())
        assert A.diagonal().min() > 1.0 and off.max() <= 0.0
        # column convection + vertical + horizontal: every interior row of the
        # pressure-form block closes (mixing ratio 1 is an equilibrium)
        r = (wl.T[tau] @ wl.rho) / wl.rho
        assert np.abs(r.reshape(nz, ny, nx)[:-1, 1:-1, 1:-1]).max() < 1e-15
        # A rho > 0 (nonsingular M-matrix certificate)
        assert np.all(A @ wl.rho > 0)
    st = mass_ref.mass_balance_diagnostics(wl.T, wl.rho, nz, ny, nx)
    assert st["interior_max"] < 1e-9 and mass_ref.mass_balance_gate(st)


def test_mass_balance_sees_a_leak_at_its_rate():
    wl = synth.make_workload(6, 5, 4, 1, seed=2)
    nz, ny, nx = wl.nz, wl.ny, wl.nx
    N = wl.N
    leak = np.zeros(N)
    leak.reshape(nz, ny, nx)[1, 2, 2] = 0.5 / 86400.0     # 0.5 per day sink in one cell
    T = [wl.T[0] + sp.diags(leak)]
    st = mass_ref.mass_balance_diagnostics(T, wl.rho, nz, ny, nx)
    assert st["interior_max"] == pytest.approx(0.5, rel=1e-9)
    inner = (nz - 1) * (ny - 2) * (nx - 2)
    assert st["interior_frac_above_0.1"] == pytest.approx(1.0 / inner)


def test_synthetic_nnz_matches_operator_shapes():
    """7-point operator at 24x20x15 and dense 15x15 columns at 64x70x15."""
    wl0 = synth.benchmark_workload("single")
    assert wl0.L(0, 0).nnz == 48120
    wl1 = synth.make_workload(64, 70, 15, 1, seed=0, convection=True)
    assert wl1.L(0, 0).nnz == 1272780
