"""Periodic-orbit adjoint with the reference's call surface, on the MI355X.

Mirrors orbit/core/orbit_adjoint.py:36-153: apply_monodromy_adjoint,
compute_s_adjoint, solve_orbit_adjoint_one_species (GMRES on
(I - M^T) lam0 = s_adj, unroll lam_tau = P_tau^T lam_{tau+1} + R_tau, residual
and periodicity diagnostics) and gradient_per_emission_bin
(dJ/de_tau = dtau (lam_tau - R_tau)).
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
