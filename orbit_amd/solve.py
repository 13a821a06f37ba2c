"""Factor backend with the reference's LU protocol, on the MI355X.

``factor_species(L_list_8)`` replaces the reference's per-species
``[_factor_one_bin(L_tau, I_N, sym, perm) for tau in range(8)]``
(orbit/core/orbit.py:79-96, 343-347) and returns 8 ``HipLU`` objects with the
duck-typed protocol of ``_UmfpackLU`` (orbit/core/solve.py:210-283):
``.solve(b, trans='N'|'T') -> new (N,) array``, ``.shape``, ``.nnz``,
``.numeric_bytes``, ``.peak_bytes``.  The 8 objects share one device operator
set, so ``orbit.solve_orbit_one_species`` recognises them and runs the whole
periodic GMRES on the device instead of 8 host round trips per iteration.
"""
from __future__ import annotations

from typing import Optional, Sequence

import numpy as np

from .core import DTAU, Context, OperatorSet


class HipLU:
    """A_tau = I + dtau*L_tau held on the device (one bin of one tracer)."""

    def __init__(self, opset: OperatorSet, bin: int, tracer: int, nnz: int = 0):
        self.opset = opset
        self.bin = bin
        self.tracer = tracer
        self.shape = (opset.N, opset.N)
        self.nnz = nnz
        st = opset.stats()
        per = st["device_bytes"] / max(st["n_bins"] * st["n_tracers"], 1)
        self.numeric_bytes = float(per)
        self.peak_bytes = float(per)

    def solve(self, b, trans: str = "N"):
        """x = A^{-1} b ('N') or A^{-T} b ('T'), FP64, to ||b - A x|| <= 1e-14 ||b||."""
        if trans not in ("N", "T"):
            raise ValueError(f"trans must be 'N' or 'T', got {trans!r}")
        return self.opset.apply_P(self.bin, self.tracer, np.asarray(b, dtype=np.float64),
                                  trans=(trans == "T"))


def factor_species(L_list: Sequence, nz: int, ny: int, nx: int, dtau: float = DTAU,
                   ctx: Optional[Context] = None):
    """Factor the 8 per-bin operators of one species; returns 8 HipLU objects."""
    op = OperatorSet.from_csr([[L] for L in L_list], nz, ny, nx, dtau, ctx)
    return [HipLU(op, b, 0, nnz=int(L_list[b].nnz)) for b in range(len(L_list))]


def factor_month(L_species_per_bin: Sequence[Sequence], nz: int, ny: int, nx: int,
                 dtau: float = DTAU, ctx: Optional[Context] = None):
    """All species of a month at once: L_species_per_bin[bin][species].
    Returns {species: [HipLU x n_bins]} sharing one operator set (the LU cache
    of orbit/core/orbit.py:468-806 in one device allocation)."""
    op = OperatorSet.from_csr(L_species_per_bin, nz, ny, nx, dtau, ctx)
    nb = len(L_species_per_bin)
    ns = len(L_species_per_bin[0])
    return {s: [HipLU(op, b, s, nnz=int(L_species_per_bin[b][s].nnz)) for b in range(nb)]
            for s in range(ns)}
