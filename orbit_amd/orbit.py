"""Periodic-orbit forward solve with the reference's call surface, on the MI355X.

Mirrors orbit/core/orbit.py:99-210 (apply_monodromy, compute_s,
solve_orbit_one_species): same arguments, same return shapes, same info keys,
same semantics for tol (atol = tol*||s||), maxiter (restart cycles), c_warm
(default x0 = s), the ||s|| < 1e-30 zero orbit, and non-convergence being a
printed warning rather than an error.

When the 8 factors are ``HipLU`` objects of one operator set the whole solve
(compute_s, GMRES(30) with the matrix-free monodromy, orbit recovery and the
diagnostics) runs on the device.  ``solve_orbits`` is the batched form over
tracers of a month that the reference runs one species at a time.
"""
from __future__ import annotations

from typing import Sequence

import numpy as np

from .core import DTAU, N_BINS, OperatorSet
from .solve import HipLU


def _common_opset(lus):
    if all(isinstance(lu, HipLU) for lu in lus):
        op = lus[0].opset
        t = lus[0].tracer
        if all(lu.opset is op and lu.tracer == t for lu in lus) and \
                [lu.bin for lu in lus] == list(range(op.n_bins)):
            return op, t
    return None, None


def apply_monodromy(v, lu_list_8):
    """M v = P_7 ... P_0 v (orbit/core/orbit.py:99-107)."""
    x = np.array(v, dtype=np.float64, copy=True)
    for lu in lu_list_8:
        x = lu.solve(x)
    return x


def compute_s(lu_list_8, e_tau_list, dtau=DTAU):
    """One forced sweep from zero (orbit/core/orbit.py:110-120)."""
    x = np.zeros(lu_list_8[0].shape[0])
    for tau, lu in enumerate(lu_list_8):
        x = lu.solve(x + dtau * np.asarray(e_tau_list[tau], dtype=np.float64))
    return x


def solve_orbit_one_species(lu_list_8, e_tau_list, dtau=DTAU, c_warm=None, tol=1e-6,
                            maxiter=200, precision="fp64"):
    """GMRES solve of (I - M) c0 = s for one species; returns (orbit[9], info)."""
    op, t = _common_opset(lu_list_8)
    if op is None:
        raise TypeError("solve_orbit_one_species needs the 8 HipLU factors of one species "
                        "(orbit_amd.solve.factor_species)")
    if abs(dtau - op.dtau) > 1e-12 * op.dtau:
        raise ValueError(f"dtau {dtau} differs from the factored operators' {op.dtau}")
    e = np.stack([np.asarray(x, dtype=np.float64) for x in e_tau_list])[None]
    warm = None if c_warm is None else np.asarray(c_warm, dtype=np.float64)[None]
    orbit, info = op.solve_forward(e, tracers=[t], c_warm=warm, tol=tol, maxiter=maxiter,
                                   precision=precision)
    inf = info[0]
    if not inf["converged"]:
        print(f"    WARNING: GMRES did not converge (info=1, iters={inf['gmres_iters']})")
    return [orbit[0, k].copy() for k in range(N_BINS + 1)], _ref_info(inf)


def _ref_info(inf):
    return {"gmres_iters": int(inf["gmres_iters"]), "residual": float(inf["residual"]),
            "periodicity": float(inf["periodicity"]), "neg_frac": float(inf["neg_frac"]),
            "converged": bool(inf["converged"]), "rounds": int(inf["rounds"])}


def solve_orbits(opset: OperatorSet, e, tracers: Sequence[int] = None, c_warm=None, tol=1e-8,
                 maxiter=200, precision="mixed", **kw):
    """Batched periodic orbits of several tracers of one operator set.

    e: (nt, n_bins, N) numpy array or CUDA tensor.  Returns (orbits, infos)."""
    return opset.solve_forward(e, tracers=tracers, c_warm=c_warm, tol=tol, maxiter=maxiter,
                               precision=precision, **kw)
