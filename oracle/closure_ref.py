"""CPU oracle of the closure's Anderson(m) acceleration -- TEST INFRASTRUCTURE ONLY.

Restates AndersonAccelerator (reference orbit/core/closure_iteration.py:568-688):
history of iterates X and residuals F = g(x) - x (at most m+1 of each); the
first update is a Picard step x + beta r; later ones solve
gamma = argmin ||r_k - dF gamma|| (numpy lstsq with rcond) and return
x_k + beta r_k - (dX + beta dF) gamma; apply_with_safeguard rejects the step
(restores, clears the history, counts the fall) when ||r_k - dF gamma|| exceeds
safeguard_factor * ||r_k||.

Pinned by tests/test_oracle_closure.py: the first step is Picard; on a linear
fixed point in R^n with full memory the (n+1)-th update is the exact fixed
point (Anderson with full memory spans the same space as GMRES); a safeguard
factor below 1 forces the Picard fallback.
"""
from __future__ import annotations

import numpy as np


class AndersonRef:
    def __init__(self, m=3, beta=1.0, rcond=1e-10, safeguard_factor=1.5):
        if m < 1:
            raise ValueError("m must be >= 1")
        self.m, self.beta, self.rcond, self.safeguard_factor = int(m), float(beta), float(rcond), float(safeguard_factor)
        self.X, self.F = [], []
        self.n_updates = 0
        self.n_safeguard_falls = 0

    def reset(self):
        self.X.clear()
        self.F.clear()

    def _dF(self):
        k = len(self.F) - 1
        return np.column_stack([self.F[i + 1] - self.F[i] for i in range(k)])

    def update(self, x_k, g_x_k):
        x_k = np.asarray(x_k, dtype=np.float64).ravel()
        r_k = np.asarray(g_x_k, dtype=np.float64).ravel() - x_k
        self.X.append(x_k.copy())
        self.F.append(r_k.copy())
        while len(self.X) > self.m + 1:
            self.X.pop(0)
            self.F.pop(0)
        self.n_updates += 1
        if len(self.F) == 1:
            return x_k + self.beta * r_k
        DF = self._dF()
        DX = np.column_stack([self.X[i + 1] - self.X[i] for i in range(len(self.X) - 1)])
        gamma = np.linalg.lstsq(DF, r_k, rcond=self.rcond)[0]
        return x_k + self.beta * r_k - (DX + self.beta * DF) @ gamma

    def apply_with_safeguard(self, x_k, g_x_k):
        x_k = np.asarray(x_k, dtype=np.float64).ravel()
        r_k = np.asarray(g_x_k, dtype=np.float64).ravel() - x_k
        snap = ([x.copy() for x in self.X], [f.copy() for f in self.F])
        x_next = self.update(x_k, g_x_k)
        if len(self.F) <= 1:
            return x_next, True
        DF = self._dF()
        gamma = np.linalg.lstsq(DF, r_k, rcond=self.rcond)[0]
        if np.linalg.norm(r_k - DF @ gamma) > self.safeguard_factor * np.linalg.norm(r_k):
            self.X, self.F = snap
            self.reset()
            self.n_safeguard_falls += 1
            return x_k + self.beta * r_k, False
        return x_next, True
