"""torch.autograd Functions over the device periodic solve.

OrbitSolve: emissions e (nt, n_bins, N) -> orbit states (nt, n_bins+1, N).  The
backward pass is the hand-written adjoint (transposed red-black sweeps in reverse
bin order inside GMRES on (I - M^T)): for output cotangents g_tau,
J = sum_tau <g_tau, c_tau> with c_8 = c_0, so R_tau = g_tau for tau = 0..7 plus
g_8 on R_0, and dJ/de_tau = dtau (lambda_tau - R_tau)
(orbit/core/orbit_adjoint.py:147-153).

PeriodicSolve: s (nt, N) -> c0 (nt, N) for (I - M) c0 = s; backward
lambda0 = (I - M^T)^-1 g, i.e. d(r.c0)/ds = lambda0 for a receptor r.

OrbitSolveFCT: the forward with the FCT deferred correction (fixed passes), and a
backward through the tangent-linear map with the limiter frozen at the forward
state (adjoint outer GMRES with L_AD^T).  The bin states the frozen operators
need are kept (recompute=False) or recomputed by re-running the forward in the
backward pass (recompute=True) when memory asks for it.

No bin state is stored for OrbitSolve / PeriodicSolve: the adjoint needs none.
"""
from __future__ import annotations

import torch

from .core import OperatorSet
from . import drivers


class OrbitSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, e, op: OperatorSet, tracers, tol, precision):
        orbit, info = op.solve_forward(e.detach().to(torch.float64).contiguous(), tracers, tol=tol,
                                       precision=precision)
        ctx.op, ctx.tracers, ctx.tol, ctx.precision = op, tracers, tol, precision
        ctx.infos = info
        return orbit

    @staticmethod
    def backward(ctx, g):
        op = ctx.op
        nb = op.n_bins
        R = g[:, :nb].clone().to(torch.float64)
        R[:, 0] += g[:, nb]
        lam, _ = op.solve_adjoint(R.contiguous(), ctx.tracers, tol=ctx.tol, precision=ctx.precision)
        grad_e = op.dtau * (lam[:, :nb] - R)
        return grad_e, None, None, None, None


class PeriodicSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, s, op: OperatorSet, tracers, tol, precision):
        orbit, info = op.solve_forward(s.detach().to(torch.float64).contiguous(), tracers, tol=tol,
                                       precision=precision, rhs="s")
        ctx.op, ctx.tracers, ctx.tol, ctx.precision = op, tracers, tol, precision
        return orbit[:, 0].clone()

    @staticmethod
    def backward(ctx, g):
        lam, _ = ctx.op.solve_adjoint(g.to(torch.float64).contiguous(), ctx.tracers, tol=ctx.tol,
                                      precision=ctx.precision, rhs="s")
        return lam[:, 0].clone(), None, None, None, None


class OrbitSolveFCT(torch.autograd.Function):
    @staticmethod
    def forward(ctx, e, op: OperatorSet, fct_grids, tracers, n_passes, tol, precision, recompute):
        e64 = e.detach().to(torch.float64).contiguous()
        orbit, hist = drivers.forward_fct(op, fct_grids, e64, tracers, n_passes=n_passes, tol=tol,
                                          precision=precision)
        ctx.op, ctx.grids, ctx.tracers = op, fct_grids, tracers
        ctx.n_passes, ctx.tol, ctx.precision, ctx.recompute = n_passes, tol, precision, recompute
        if recompute:
            ctx.save_for_backward(e64)
        else:
            ctx.save_for_backward(orbit)
        return orbit

    @staticmethod
    def backward(ctx, g):
        op = ctx.op
        nb = op.n_bins
        (saved,) = ctx.saved_tensors
        if ctx.recompute:
            base, _ = drivers.forward_fct(op, ctx.grids, saved, ctx.tracers, n_passes=ctx.n_passes,
                                          tol=ctx.tol, precision=ctx.precision)
        else:
            base = saved
        L = drivers.frozen_operators(op, ctx.grids, base)
        R = g[:, :nb].clone().to(torch.float64)
        R[:, 0] += g[:, nb]
        lam, Re, _ = drivers.adjoint_fct(op, L, R.contiguous(), ctx.tracers, tol=min(ctx.tol, 1e-10),
                                         precision="fp64")
        grad_e = op.dtau * (lam[:, :nb] - Re)
        return grad_e, None, None, None, None, None, None, None


def orbit_solve(e, op, tracers=None, tol=1e-10, precision="fp64"):
    return OrbitSolve.apply(e, op, tracers, tol, precision)


def periodic_solve(s, op, tracers=None, tol=1e-10, precision="fp64"):
    return PeriodicSolve.apply(s, op, tracers, tol, precision)


def orbit_solve_fct(e, op, fct_grids, tracers=None, n_passes=6, tol=1e-10, precision="fp64",
                    recompute=False):
    return OrbitSolveFCT.apply(e, op, fct_grids, tracers, n_passes, tol, precision, recompute)
