"""Pins of the oracle's FCT deferred-correction drivers (oracle/drivers_ref.py).

* With the anti-diffusive term switched off (zero winds for the forward passes,
  L_AD = 0 for the marginal and adjoint drivers) each driver returns the
  low-order solve (reference orbit/cli.py:1212-1241, orbit/modes/marginal.py:627-730,
  orbit/modes/adjoint.py:814-962 all reduce to it).
* The marginal driver's fixed point is the solution of the fully coupled
  block-cyclic system (A_tau + dtau L_AD,tau) dc_{tau+1} = dc_tau + dtau de_tau,
  dc_0 = dc_8, solved directly with splu over all 8N unknowns.
* The adjoint driver's gradient dtau (lam_tau - R_eff_tau) reproduces the
  marginal driver: J(de) = sum_tau <R_tau, dc_tau> is linear in de, so
  J(de) = <grad, de> for any de.
"""
import dataclasses

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import drivers_ref, fct_ref, orbit_ref as orb
from orbit_amd import synth

N_BINS = 8


@pytest.fixture(scope="module")
def case():
    wl = synth.make_workload(6, 5, 4, 1, seed=17, convection=True, wind_std=8.0)
    lus = [orb.factor_one_bin(wl.L(t, 0)) for t in range(N_BINS)]
    e = wl.e[0].copy()
    e[:, : wl.nx * wl.ny] *= 1.0 + 4.0 * (np.arange(wl.nx * wl.ny) % wl.nx > wl.nx // 2)
    base, _ = drivers_ref.forward_fct_passes(lus, list(e), wl.grids, n_passes=2, tol=1e-13)
    L_AD = [fct_ref.assemble_fct_linear_operator(wl.grids[t], base[t + 1], wl.dtau) for t in range(N_BINS)]
    return wl, lus, e, L_AD


def _rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300))


def test_limiter_is_active_in_the_case(case):
    wl, lus, e, L_AD = case
    assert max(abs(L).max() for L in L_AD) > 1e-8


def test_forward_passes_with_zero_winds_are_the_low_order_orbit(case):
    wl, lus, e, _ = case
    calm = [dataclasses.replace(g, UAvg=0 * g.UAvg, VAvg=0 * g.VAvg) for g in wl.grids]
    orbit, hist = drivers_ref.forward_fct_passes(lus, list(e), calm, n_passes=3, tol=1e-13)
    ref, _ = orb.solve_orbit_one_species(lus, list(e), tol=1e-13, maxiter=500)
    for k in (0, 4, 8):
        assert _rel(orbit[k], ref[k]) < 1e-10
    assert len(hist) == 3


def test_forward_passes_change_the_orbit_with_winds(case):
    wl, lus, e, _ = case
    o1, _ = drivers_ref.forward_fct_passes(lus, list(e), wl.grids, n_passes=1, tol=1e-13)
    o3, _ = drivers_ref.forward_fct_passes(lus, list(e), wl.grids, n_passes=3, tol=1e-13)
    assert _rel(o3[0], o1[0]) > 1e-6


def test_drivers_with_zero_L_AD_are_the_low_order_solves(case):
    wl, lus, e, _ = case
    zero = [sp.csr_matrix((wl.N, wl.N)) for _ in range(N_BINS)]
    rng = np.random.default_rng(0)
    de = list(rng.random((N_BINS, wl.N)))
    dc, info = drivers_ref.marginal_fct_gmres(lus, de, zero, tol=1e-13)
    ref, _ = orb.solve_orbit_one_species(lus, de, tol=1e-13, maxiter=500)
    assert _rel(dc[0], ref[0]) < 1e-10 and _rel(dc[5], ref[5]) < 1e-10
    R = list(rng.random((N_BINS, wl.N)))
    lam, Reff, _ = drivers_ref.adjoint_fct_gmres(lus, R, zero, tol=1e-13)
    lref, _ = orb.solve_orbit_adjoint_one_species(lus, R, tol=1e-13, maxiter=500)
    assert _rel(lam[0], lref[0]) < 1e-10 and _rel(lam[3], lref[3]) < 1e-10
    assert all(np.array_equal(a, b) for a, b in zip(Reff, R))


def test_marginal_driver_solves_the_block_cyclic_system(case):
    wl, lus, e, L_AD = case
    N = wl.N
    rng = np.random.default_rng(1)
    de = rng.random((N_BINS, N)) * (rng.random((N_BINS, N)) < 0.3)
    dc, info = drivers_ref.marginal_fct_gmres(lus, list(de), L_AD, tol=1e-13, max_iters=80, rel_tol=1e-11)
    # unknowns dc_1..dc_8 (dc_0 = dc_8); block row tau: (A_tau + dtau L_AD,tau) dc_{tau+1} - dc_tau
    I = sp.identity(N, format="csr")
    blocks = [[None] * N_BINS for _ in range(N_BINS)]
    for t in range(N_BINS):
        blocks[t][t] = (wl.A(t, 0) + wl.dtau * L_AD[t]).tocsr()
        blocks[t][(t - 1) % N_BINS] = -I
    K = sp.bmat(blocks, format="csc")
    x = spla.splu(K).solve(wl.dtau * de.ravel()).reshape(N_BINS, N)
    assert _rel(dc[8], x[7]) < 1e-9
    assert _rel(dc[0], x[7]) < 1e-9
    for t in (0, 3, 6):
        assert _rel(dc[t + 1], x[t]) < 1e-9, t


def test_adjoint_driver_gradient_reproduces_the_marginal_driver(case):
    wl, lus, e, L_AD = case
    N = wl.N
    rng = np.random.default_rng(2)
    R = np.zeros((N_BINS, N))
    R[0, rng.choice(wl.nx * wl.ny, 8, replace=False)] = rng.random(8)
    R[5, rng.choice(N, 6, replace=False)] = rng.random(6)
    lam, Reff, _ = drivers_ref.adjoint_fct_gmres(lus, list(R), L_AD, tol=1e-13, max_iters=80, rel_tol=1e-11)
    grad = np.stack([wl.dtau * (lam[t] - Reff[t]) for t in range(N_BINS)])
    for trial in range(3):
        de = rng.standard_normal((N_BINS, N))
        dc, _ = drivers_ref.marginal_fct_gmres(lus, list(de), L_AD, tol=1e-13, max_iters=80, rel_tol=1e-11)
        J = sum(float(np.dot(R[t], dc[t])) for t in range(N_BINS))
        assert J == pytest.approx(float(np.sum(grad * de)), rel=1e-8, abs=1e-12 * np.abs(grad).sum()), trial


def test_pass_states_split_into_previous_states_plus_the_propagated_change():
    """The identity the mixed forward passes start from (orbit_forward_fct_dev):
    the bin states swept from a fixed c0 are linear in the forcing, so with
    forcing e they equal the states with e_prev plus the sweep of
    dtau (e - e_prev) from zero -- and the periodic residual's source part
    s(e) = s(e_prev) + that sweep's last state.  Checked with SuperLU
    propagators; a sweep of the change from c0 instead of zero, or the change
    without dtau, misses it."""
    wl = synth.make_workload(10, 8, 5, 1, seed=44, convection=True)
    lus = [orb.factor_one_bin(wl.L(t, 0)) for t in range(8)]
    rng = np.random.default_rng(45)
    e_prev = [wl.e[0, t] for t in range(8)]
    e_new = [x + 0.1 * rng.standard_normal(x.shape) * np.abs(x).max() for x in e_prev]
    c0 = rng.random(wl.N)

    def sweep(c, e):
        x = [c]
        for t in range(8):
            x.append(lus[t].solve(x[-1] + wl.dtau * e[t]))
        return x

    old, new = sweep(c0, e_prev), sweep(c0, e_new)
    y = sweep(np.zeros(wl.N), [a - b for a, b in zip(e_new, e_prev)])
    for t in range(1, 9):
        assert _rel(old[t] + y[t], new[t]) < 1e-12, t
    s_new = orb.compute_s(lus, e_new)
    assert _rel(orb.compute_s(lus, e_prev) + y[8], s_new) < 1e-12
    wrong = sweep(c0, [a - b for a, b in zip(e_new, e_prev)])
    assert _rel(old[8] + wrong[8], new[8]) > 1e-3
    y_nodt = sweep(np.zeros(wl.N), [(a - b) / wl.dtau for a, b in zip(e_new, e_prev)])
    assert _rel(old[8] + y_nodt[8], new[8]) > 1e-3
