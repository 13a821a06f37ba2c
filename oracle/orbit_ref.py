"""CPU restatement of the periodic-orbit forward and adjoint solve (TEST ORACLE).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
bench CPU-baseline leg, never by the product path.

Follows the reference file:line by file:line:
  factor_one_bin          orbit/core/orbit.py:79-96 (SuperLU branch, solve.py:361-365)
  apply_monodromy         orbit/core/orbit.py:99-107
  compute_s               orbit/core/orbit.py:110-120
  solve_orbit_one_species orbit/core/orbit.py:123-210
  apply_monodromy_adjoint orbit/core/orbit_adjoint.py:36-47
  compute_s_adjoint       orbit/core/orbit_adjoint.py:50-67
  solve_orbit_adjoint_one_species  orbit/core/orbit_adjoint.py:70-144
  gradient_per_emission_bin        orbit/core/orbit_adjoint.py:147-153

Pinned by (tests/test_oracle_orbit.py): splu against dense numpy.linalg.solve;
the monodromy against an explicitly formed dense M; the periodic solve against
numpy.linalg.solve on (I - M) for tiny N; the steady state c = e/l of a
time-invariant diagonal system (tests/test_simulation_modes.py:1239-1270);
periodicity c_8 = c_0; the adjoint against central finite differences of the
forward (tests/test_orbit_adjoint.py:21-84) and the dot-product identity
<r,(I-M)^-1 s> = <(I-M^T)^-1 r, s>; zero receptor -> zero orbit
(tests/test_orbit_adjoint.py:88-99).  GMRES iteration counts are not pinned
(scipy's trajectory is an implementation detail of the reference's dependency).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

DTAU = 10800.0
N_BINS = 8


class SuperLUFactor:
    """A_tau = I + dtau*L_tau factored by SuperLU; the factor-backend protocol
    of the reference: ``.solve(b, trans)``, ``.shape``, ``.nnz``."""

    def __init__(self, A_csc: sp.csc_matrix):
        self._lu = spla.splu(A_csc)
        self.shape = A_csc.shape
        self.nnz = A_csc.nnz
        self.numeric_bytes = 0.0

    def solve(self, b, trans: str = "N"):
        if trans not in ("N", "T"):
            raise ValueError(f"trans must be 'N' or 'T', got {trans!r}")
        return self._lu.solve(np.asarray(b, dtype=np.float64), trans=trans)


def factor_one_bin(L_tau, dtau: float = DTAU) -> SuperLUFactor:
    """orbit/core/orbit.py:79-96: A = I + L*dtau, CSC, duplicates summed, indices sorted."""
    n = L_tau.shape[0]
    A = (sp.identity(n, format="csc") + dtau * sp.csc_matrix(L_tau)).tocsc()
    A.sum_duplicates()
    A.sort_indices()
    return SuperLUFactor(A)


def apply_monodromy(v, lus):
    """M v = P_7 ... P_0 v: eight sequential solves in bin order."""
    x = np.array(v, dtype=np.float64, copy=True)
    for lu in lus:
        x = lu.solve(x)
    return x


def compute_s(lus, e_list, dtau: float = DTAU):
    """One forced sweep from zero: x_{tau+1} = P_tau (x_tau + dtau e_tau); s = x_8."""
    x = np.zeros(lus[0].shape[0])
    for tau, lu in enumerate(lus):
        x = lu.solve(x + dtau * np.asarray(e_list[tau], dtype=np.float64))
    return x


def _gmres(matvec, b, x0, atol, restart, maxiter):
    """scipy GMRES with the reference's arguments (orbit/core/orbit.py:177-182):
    rtol=1e-30, atol=tol*||s||, restart=30, maxiter restart cycles, pr_norm
    callback counting inner iterations."""
    n = b.shape[0]
    op = spla.LinearOperator((n, n), matvec=matvec, dtype=np.float64)
    count = [0]

    def cb(_):
        count[0] += 1

    x, code = spla.gmres(op, b, x0=x0, rtol=1e-30, atol=atol, restart=restart,
                         maxiter=maxiter, callback=cb, callback_type="pr_norm")
    return x, code, count[0]


def solve_orbit_one_species(lus, e_list, dtau: float = DTAU, c_warm=None,
                            tol: float = 1e-6, maxiter: int = 200, restart: int = 30):
    """orbit/core/orbit.py:123-210.  Returns (orbit[9], info)."""
    n = lus[0].shape[0]
    s = compute_s(lus, e_list, dtau)
    s_norm = float(np.linalg.norm(s))
    if s_norm < 1e-30:
        return ([np.zeros(n) for _ in range(N_BINS + 1)],
                {"gmres_iters": 0, "residual": 0.0, "periodicity": 0.0,
                 "neg_frac": 0.0, "converged": True})

    def matvec(v):
        return v - apply_monodromy(v, lus)

    x0 = np.array(c_warm, dtype=np.float64) if c_warm is not None else s.copy()
    c0, code, iters = _gmres(matvec, s, x0, tol * s_norm, restart, maxiter)
    orbit = [c0.copy()]
    x = c0.copy()
    for tau, lu in enumerate(lus):
        x = lu.solve(x + dtau * np.asarray(e_list[tau], dtype=np.float64))
        orbit.append(x.copy())
    residual = float(np.linalg.norm(matvec(c0) - s) / s_norm)
    periodicity = float(np.linalg.norm(orbit[-1] - c0) / (np.linalg.norm(c0) + 1e-30))
    neg_frac = float(np.mean(np.concatenate(orbit) < 0))
    return orbit, {"gmres_iters": iters, "residual": residual,
                   "periodicity": periodicity, "neg_frac": neg_frac,
                   "converged": code == 0}


def apply_monodromy_adjoint(v, lus):
    """M^T v = P_0^T ... P_7^T v: transposed solves in reverse bin order."""
    x = np.array(v, dtype=np.float64, copy=True)
    for lu in reversed(lus):
        x = lu.solve(x, trans="T")
    return x


def compute_s_adjoint(lus, R_list):
    """Backward sweep from zero: x = P_tau^T x + R_tau for tau = 7..0."""
    x = np.zeros(lus[0].shape[0])
    for tau in range(N_BINS - 1, -1, -1):
        x = lus[tau].solve(x, trans="T") + np.asarray(R_list[tau], dtype=np.float64)
    return x


def solve_orbit_adjoint_one_species(lus, R_list, *, lam_warm=None, tol: float = 1e-6,
                                    maxiter: int = 200, restart: int = 30):
    """orbit/core/orbit_adjoint.py:70-144.  Returns (adjoint[9], info)."""
    n = lus[0].shape[0]
    s_adj = compute_s_adjoint(lus, R_list)
    s_norm = float(np.linalg.norm(s_adj))
    if s_norm < 1e-30:
        return ([np.zeros(n) for _ in range(N_BINS + 1)],
                {"gmres_iters": 0, "residual": 0.0, "periodicity": 0.0, "converged": True})

    def matvec(v):
        return v - apply_monodromy_adjoint(v, lus)

    x0 = np.array(lam_warm, dtype=np.float64) if lam_warm is not None else s_adj.copy()
    lam0, code, iters = _gmres(matvec, s_adj, x0, tol * s_norm, restart, maxiter)
    adj = [None] * (N_BINS + 1)
    adj[0] = lam0.copy()
    adj[N_BINS] = lam0.copy()
    x = lam0.copy()
    for tau in range(N_BINS - 1, 0, -1):
        x = lus[tau].solve(x, trans="T") + np.asarray(R_list[tau], dtype=np.float64)
        adj[tau] = x.copy()
    residual = float(np.linalg.norm(matvec(lam0) - s_adj) / s_norm)
    lam0_check = lus[0].solve(adj[1], trans="T") + np.asarray(R_list[0], dtype=np.float64)
    periodicity = float(np.linalg.norm(lam0_check - lam0) / (np.linalg.norm(lam0) + 1e-30))
    return adj, {"gmres_iters": iters, "residual": residual,
                 "periodicity": periodicity, "converged": code == 0}


def gradient_per_emission_bin(adjoint, R_list, dtau: float = DTAU):
    """dJ/de_tau = dtau (lambda_tau - R_tau) for J = sum_tau <R_tau, c_tau>."""
    return [dtau * (adjoint[tau] - np.asarray(R_list[tau])) for tau in range(N_BINS)]


def dense_monodromy(lus):
    """Explicit dense M for tiny N (brute-force pin)."""
    n = lus[0].shape[0]
    return np.column_stack([apply_monodromy(np.eye(n)[:, j], lus) for j in range(n)])
