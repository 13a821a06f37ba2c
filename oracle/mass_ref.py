"""CPU restatement of the air-mass-balance pre-flight (TEST ORACLE).

Test infrastructure only.  Follows orbit/core/operator.py:598-652
(mass_balance_diagnostics) with the abort rule of orbit/cli.py:1591-1614
(p90 of |row sum| per day over interior layers above 0.05 aborts).

Pinned by tests/test_oracle_mass.py: a mass-consistent transport block reads
zero in the interior, a block with an unbalanced vertical flux reads the leak
at exactly its rate per day (tests/test_interface_omega_transport.py:216-240),
and percentiles agree with numpy on a hand-made block.
"""
from __future__ import annotations

import numpy as np

DAY = 86400.0


def mass_balance_diagnostics(T_blocks, rho, nz, ny, nx):
    """T_blocks: per-bin concentration-form transport blocks (N x N); rho: (N,)."""
    inner = np.zeros((nz, ny, nx), dtype=bool)
    inner[:, 1:-1, 1:-1] = True
    interior, top, per_bin = [], [], []
    w = np.asarray(rho, dtype=np.float64)
    for T in T_blocks:
        r = (np.asarray(T @ w).ravel() / w).reshape(nz, ny, nx) * DAY
        r = np.where(inner, r, np.nan)
        interior.append(r[:-1].ravel())
        top.append(r[-1].ravel())
        per_bin.append(r)
    ri = np.abs(np.concatenate(interior))
    ri = ri[np.isfinite(ri)]
    rt = np.abs(np.concatenate(top))
    rt = rt[np.isfinite(rt)]
    pl = np.abs(np.stack(per_bin))
    signed = np.stack(per_bin)
    return {
        "interior_median": float(np.median(ri)),
        "interior_p90": float(np.percentile(ri, 90)),
        "interior_p99": float(np.percentile(ri, 99)),
        "interior_max": float(ri.max()),
        "interior_frac_above_0.01": float(np.mean(ri > 0.01)),
        "interior_frac_above_0.1": float(np.mean(ri > 0.1)),
        "top_median": float(np.median(rt)),
        "top_p90": float(np.percentile(rt, 90)),
        "layer_p90": [float(np.nanpercentile(pl[:, k], 90)) for k in range(nz)],
        "layer_mean_signed": [float(np.nanmean(signed[:, k])) for k in range(nz)],
    }


def mass_balance_gate(stats, tol=0.05):
    """True when the pre-flight passes (orbit/cli.py:1604-1614)."""
    return stats["interior_p90"] <= tol
