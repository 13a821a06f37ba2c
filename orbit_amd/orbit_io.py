"""The orbit NPZ interop format on either side of the solve.

Writes and reads the orbit part of the reference's forward outputs and
closure checkpoints (orbit/cli.py:1297-1413 and 3025-3035; MODES.md
"Output files"): c_orbit (S, N_BINS+1, N) along the periodic orbit,
c_mean (S, N) = mean of c_1..c_8, orbit_species_order, the per-species
gmres_iters / gmres_resid / periodicity, month.  Files are plain
np.savez_compressed archives of numeric and string arrays, loaded with
allow_pickle=False, so they interchange with the reference's own files: a
device solve's orbits can seed the reference's marginal / adjoint modes, and
the reference's orbit_M{MM}.npz can warm-start the device solve
(cli.py:1905-1916: the previous month's c_mean).
"""
from __future__ import annotations

from typing import Dict, Optional, Sequence

import numpy as np


def _host(a):
    return a.detach().cpu().numpy() if hasattr(a, "detach") else np.asarray(a)


def orbit_mean(c_orbit):
    """c_mean = mean of states 1..n_bins (cli.py:3031-3034); tensor or array."""
    return c_orbit[:, 1:].mean(axis=1) if isinstance(c_orbit, np.ndarray) else c_orbit[:, 1:].mean(dim=1)


def save_orbit_npz(path: str, c_orbit, species_order: Optional[Sequence[int]] = None,
                   info: Optional[Sequence[dict]] = None, month: Optional[int] = None,
                   iteration: int = 0, converged: Optional[bool] = None, **extra) -> None:
    """c_orbit: (S, n_bins+1, N) device tensor or array; info: per-species dicts
    of the solve (gmres_iters, residual, periodicity).  Every key the
    reference's checkpoint reader takes (orbit/cli.py:1387-1407: c_orbit,
    orbit_species_order, gmres_iters, gmres_resid, periodicity, iter,
    converged) is always written: without info the per-species arrays hold 0
    iterations and NaN residuals; ``iteration`` is the outer closure iteration
    and ``converged`` defaults to every species' converged flag (False
    without info)."""
    c = np.ascontiguousarray(_host(c_orbit), dtype=np.float64)
    if c.ndim != 3:
        raise ValueError("c_orbit must be (S, n_bins+1, N)")
    S = c.shape[0]
    order = np.asarray(range(S) if species_order is None else species_order, dtype=np.int64)
    if order.shape != (S,):
        raise ValueError(f"species_order needs {S} entries, got {order.shape}")
    if info is not None and len(info) != S:
        raise ValueError(f"info needs one dict per species ({S}), got {len(info)}")
    save = {"c_orbit": c, "c_mean": orbit_mean(c), "orbit_species_order": order}
    if info is not None:
        save["gmres_iters"] = np.array([i["gmres_iters"] for i in info], dtype=np.int64)
        save["gmres_resid"] = np.array([i["residual"] for i in info], dtype=np.float64)
        save["periodicity"] = np.array([i["periodicity"] for i in info], dtype=np.float64)
    else:
        save["gmres_iters"] = np.zeros(S, dtype=np.int64)
        save["gmres_resid"] = np.full(S, np.nan)
        save["periodicity"] = np.full(S, np.nan)
    if converged is None:
        converged = info is not None and all(bool(i.get("converged", True)) for i in info)
    save["iter"] = np.array(int(iteration))
    save["converged"] = np.array(bool(converged))
    if month is not None:
        save["month"] = np.array(month)
    for k, v in extra.items():
        save[k] = np.asarray(_host(v))
    np.savez_compressed(path, **save)


def load_orbit_npz(path: str) -> Dict:
    """orbits {species: [c_0, ..., c_n_bins]} (cli.py:1390-1398) plus the arrays."""
    d = np.load(path, allow_pickle=False)
    c = d["c_orbit"]
    order = d["orbit_species_order"] if "orbit_species_order" in d.files else np.arange(c.shape[0])
    out = {k: d[k] for k in d.files}
    out["orbits"] = {int(s): [c[i, t].copy() for t in range(c.shape[1])] for i, s in enumerate(order.tolist())}
    if "c_mean" not in out:
        out["c_mean"] = orbit_mean(c)
    return out


def warm_start_from(path: str, n_species: int):
    """The previous month's orbit mean as a warm start (cli.py:1905-1916):
    c_mean as (n_species, N), or None when the file holds another species count."""
    d = np.load(path, allow_pickle=False)
    cm = d["c_mean"]
    return cm if cm.shape[0] == n_species else None
