"""Anderson(m) acceleration of the closure's outer fixed point, on the device.

Replaces AndersonAccelerator (reference orbit/core/closure_iteration.py:568-688),
which accelerates the partitioning/oxidant closure around the orbit solve:
x_next = update(x_k, g(x_k)), where g is one full chemistry + orbit pass.  The
state (the flattened OH field over bins, or any other closure field) stays a
CUDA tensor: the history, the differences and the update are device vector
operations, and the least-squares problem min ||r_k - dF gamma|| is solved from
a thin QR of dF on the device and an SVD of its small m x m factor R on the
host, with the reference's rcond cutoff on the singular values (the singular
values of R are those of dF, so this is the SVD solve numpy's lstsq does).
"""
from __future__ import annotations

from typing import List, Tuple

import numpy as np
import torch


def _lstsq(DF: torch.Tensor, r: torch.Tensor, rcond: float) -> torch.Tensor:
    Q, R = torch.linalg.qr(DF, mode="reduced")
    qtr = (Q.T @ r).cpu().numpy()
    U, s, Vh = np.linalg.svd(R.cpu().numpy())
    keep = s > rcond * s.max() if s.size and s.max() > 0 else np.zeros_like(s, dtype=bool)
    g = Vh.T[:, keep] @ ((U[:, keep].T @ qtr) / s[keep])
    return torch.as_tensor(g, dtype=r.dtype, device=r.device)


class AndersonAccelerator:
    """Anderson(m) on device tensors; same parameters, history rules and
    safeguard as the reference."""

    def __init__(self, m: int = 3, beta: float = 1.0, rcond: float = 1e-10, safeguard_factor: float = 1.5):
        if m < 1:
            raise ValueError(f"the Anderson history keeps at least one past iterate; m = {m} asks for none")
        self.m = int(m)
        self.beta = float(beta)
        self.rcond = float(rcond)
        self.safeguard_factor = float(safeguard_factor)
        self.X: List[torch.Tensor] = []
        self.F: List[torch.Tensor] = []
        self.n_updates = 0
        self.n_safeguard_falls = 0

    def reset(self) -> None:
        self.X.clear()
        self.F.clear()

    def _dF(self) -> torch.Tensor:
        return torch.stack([self.F[i + 1] - self.F[i] for i in range(len(self.F) - 1)], dim=1)

    def update(self, x_k: torch.Tensor, g_x_k: torch.Tensor) -> torch.Tensor:
        x_k = x_k.reshape(-1).to(torch.float64)
        r_k = g_x_k.reshape(-1).to(torch.float64) - x_k
        self.X.append(x_k.clone())
        self.F.append(r_k.clone())
        while len(self.X) > self.m + 1:
            self.X.pop(0)
            self.F.pop(0)
        self.n_updates += 1
        if len(self.F) == 1:
            return x_k + self.beta * r_k
        DF = self._dF()
        DX = torch.stack([self.X[i + 1] - self.X[i] for i in range(len(self.X) - 1)], dim=1)
        gamma = _lstsq(DF, r_k, self.rcond)
        return x_k + self.beta * r_k - (DX + self.beta * DF) @ gamma

    def apply_with_safeguard(self, x_k: torch.Tensor, g_x_k: torch.Tensor) -> Tuple[torch.Tensor, bool]:
        x_k = x_k.reshape(-1).to(torch.float64)
        r_k = g_x_k.reshape(-1).to(torch.float64) - x_k
        snap = ([x.clone() for x in self.X], [f.clone() for f in self.F])
        x_next = self.update(x_k, g_x_k)
        if len(self.F) <= 1:
            return x_next, True
        DF = self._dF()
        gamma = _lstsq(DF, r_k, self.rcond)
        proj = float(torch.linalg.vector_norm(r_k - DF @ gamma))
        if proj > self.safeguard_factor * float(torch.linalg.vector_norm(r_k)):
            self.X, self.F = snap
            self.reset()
            self.n_safeguard_falls += 1
            return x_k + self.beta * r_k, False
        return x_next, True

    def __repr__(self) -> str:
        return (f"AndersonAccelerator(m={self.m}, beta={self.beta}, history_len={len(self.X)}, "
                f"updates={self.n_updates}, safeguard_falls={self.n_safeguard_falls})")
