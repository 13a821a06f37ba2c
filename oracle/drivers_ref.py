"""CPU restatement of the FCT deferred-correction drivers, one species (TEST ORACLE).

Test infrastructure only.

  forward_fct_passes   orbit/cli.py:1212-1241 (d_AD of the previous end-of-bin
                       orbit added to each bin's RHS, recomputed per outer pass)
  marginal_fct_gmres   orbit/modes/marginal.py:627-730 (extra_rhs_tau =
                       -L_AD,tau dc[tau+1]; outer GMRES on (I-K)x = b over the
                       flattened 9-state orbit)
  adjoint_fct_gmres    orbit/modes/adjoint.py:814-962 (mu = lambda - R; receptor
                       nu_tau = -dtau L_AD,(tau-1)^T mu_(tau-1); outer GMRES on mu)

Pinned by tests/test_oracle_drivers.py: with L_AD = 0 each driver returns the
low-order solve; the marginal driver's fixed point satisfies the fully coupled
block-cyclic system (A_tau + dtau L_AD,tau) c_{tau+1} = c_tau + dtau e_tau solved
directly by splu; the adjoint driver's gradient matches finite differences of
the marginal driver.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse.linalg as spla

from . import fct_ref, orbit_ref

N_BINS = orbit_ref.N_BINS
DTAU = orbit_ref.DTAU


def forward_fct_passes(lus, e_list, grids, n_passes=6, tol=1e-8, maxiter=200,
                       dtau=DTAU, ghost_list=None):
    """Pass 0 is the low-order orbit; pass p adds d_AD(orbit_{p-1}[tau+1])."""
    orbit, info = orbit_ref.solve_orbit_one_species(lus, e_list, dtau, tol=tol, maxiter=maxiter)
    history = [info]
    for _ in range(1, n_passes):
        e_eff = []
        for tau in range(N_BINS):
            d = fct_ref.compute_horizontal_fct_source(
                grids[tau], orbit[tau + 1], dtau,
                c_ghost_flat=None if ghost_list is None else ghost_list[tau])
            e_eff.append(np.asarray(e_list[tau]) + d)
        orbit, info = orbit_ref.solve_orbit_one_species(lus, e_eff, dtau, c_warm=orbit[0],
                                                        tol=tol, maxiter=maxiter)
        history.append(info)
    return orbit, history


def _gmres_outer(apply_g, b, max_iters=40, rel_tol=2.0e-3):
    n = b.size
    b_norm = float(np.linalg.norm(b))
    op = spla.LinearOperator((n, n), matvec=lambda v: v - (apply_g(v) - b), dtype=np.float64)
    restart = min(max_iters, 30)
    it = [0]

    def cb(_):
        it[0] += 1

    x, code = spla.gmres(op, b, x0=b.copy(), rtol=1e-30, atol=0.3 * rel_tol * b_norm,
                         restart=restart, maxiter=max(1, -(-max_iters // restart)),
                         callback=cb, callback_type="pr_norm")
    return x, it[0], code


def marginal_fct_gmres(lus, de_list, L_AD_list, tol=1e-10, maxiter=200, dtau=DTAU,
                       max_iters=40, rel_tol=2.0e-3):
    """Tangent-linear orbit with the frozen FCT operator as a deferred RHS source."""
    N = lus[0].shape[0]

    def solve_with(prev_flat):
        e = [np.asarray(de_list[t], dtype=np.float64).copy() for t in range(N_BINS)]
        if prev_flat is not None:
            prev = prev_flat.reshape(N_BINS + 1, N)
            for t in range(N_BINS):
                e[t] = e[t] - L_AD_list[t] @ prev[t + 1]
        orbit, _ = orbit_ref.solve_orbit_one_species(lus, e, dtau, tol=tol, maxiter=maxiter)
        return np.concatenate(orbit)

    b = solve_with(None)
    x, iters, code = _gmres_outer(lambda v: solve_with(v), b, max_iters, rel_tol)
    g = solve_with(x)
    rel = float(np.abs(g - x).max() / max(float(np.abs(g).max()), 1e-30))
    return g.reshape(N_BINS + 1, N), {"outer_iters": iters, "rel_final": rel, "code": code}


def adjoint_fct_gmres(lus, R_list, L_AD_list, tol=1e-10, maxiter=200, dtau=DTAU,
                      max_iters=40, rel_tol=2.0e-3):
    """Adjoint orbit with the transposed frozen FCT operator as a receptor source.

    Returns (lam[9], R_eff[8], info); dJ/de_tau = dtau (lam_tau - R_eff_tau).
    """
    N = lus[0].shape[0]
    R_base = [np.asarray(r, dtype=np.float64) for r in R_list]

    def run(mu_flat):
        R = [r.copy() for r in R_base]
        if mu_flat is not None:
            mu = mu_flat.reshape(N_BINS, N)
            for t in range(N_BINS):
                src = (t - 1) % N_BINS
                R[t] = R[t] - dtau * (L_AD_list[src].T @ mu[src])
        lam, _ = orbit_ref.solve_orbit_adjoint_one_species(lus, R, tol=tol, maxiter=maxiter)
        mu_new = np.concatenate([lam[t] - R[t] for t in range(N_BINS)])
        return mu_new, lam, R

    b, _, _ = run(None)
    x, iters, code = _gmres_outer(lambda v: run(v)[0], b, max_iters, rel_tol)
    g, lam, R = run(x)
    rel = float(np.abs(g - x).max() / max(float(np.abs(g).max()), 1e-30))
    return lam, R, {"outer_iters": iters, "rel_final": rel, "code": code}
