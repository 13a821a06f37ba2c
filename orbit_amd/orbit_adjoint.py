"""Periodic-orbit adjoint with the reference's call surface, on the MI355X.

Mirrors orbit/core/orbit_adjoint.py:36-153: apply_monodromy_adjoint,
compute_s_adjoint, solve_orbit_adjoint_one_species (GMRES on
(I - M^T) lam0 = s_adj, unroll lam_tau = P_tau^T lam_{tau+1} + R_tau, residual
and periodicity diagnostics) and gradient_per_emission_bin
(dJ/de_tau = dtau (lam_tau - R_tau)); and receptor_R_list, the adjoint mode's
pairing of end-of-bin receptors with the recursion (orbit/modes/adjoint.py:192-203).
"""
from __future__ import annotations

import numpy as np

from .core import DTAU, N_BINS
from .orbit import _common_opset, _ref_info


def apply_monodromy_adjoint(v, lu_list_8):
    x = np.array(v, dtype=np.float64, copy=True)
    for lu in reversed(lu_list_8):
        x = lu.solve(x, trans="T")
    return x


def compute_s_adjoint(lu_list_8, R_list):
    x = np.zeros(lu_list_8[0].shape[0])
    for tau in range(N_BINS - 1, -1, -1):
        x = lu_list_8[tau].solve(x, trans="T") + np.asarray(R_list[tau], dtype=np.float64)
    return x


def solve_orbit_adjoint_one_species(lu_list_8, R_list, *, lam_warm=None, tol=1e-6, maxiter=200,
                                    precision="fp64"):
    op, t = _common_opset(lu_list_8)
    if op is None:
        raise TypeError("solve_orbit_adjoint_one_species needs the 8 HipLU factors of one species")
    R = np.stack([np.asarray(r, dtype=np.float64) for r in R_list])[None]
    warm = None if lam_warm is None else np.asarray(lam_warm, dtype=np.float64)[None]
    lam, info = op.solve_adjoint(R, tracers=[t], lam_warm=warm, tol=tol, maxiter=maxiter,
                                 precision=precision)
    inf = info[0]
    if not inf["converged"]:
        print(f"    WARNING: adjoint GMRES did not converge (info=1, iters={inf['gmres_iters']})")
    d = _ref_info(inf)
    d.pop("neg_frac", None)
    return [lam[0, k].copy() for k in range(N_BINS + 1)], d


def gradient_per_emission_bin(adjoint, R_list, dtau=DTAU):
    return [dtau * (np.asarray(adjoint[tau]) - np.asarray(R_list[tau])) for tau in range(N_BINS)]


def receptor_R_list(receptors, sp_idx):
    """R list of one species from per-bin receptors (orbit/modes/adjoint.py:192-203).

    receptors[tau, sp_idx] weights the end-of-bin state c[tau + 1]:
    J = sum_tau <r_tau, c[tau + 1]>.  In lam_tau = P_tau^T lam_{tau+1} + R_tau
    the term R_tau weights c[tau], so r_tau goes to R[(tau + 1) mod N_BINS]
    (c[N_BINS] = c[0] on the periodic orbit)."""
    return [np.array(receptors[(tau - 1) % N_BINS, sp_idx], dtype=np.float64) for tau in range(N_BINS)]
