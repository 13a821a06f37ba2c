"""torch.autograd Functions over the device periodic solve.

OrbitSolve: emissions e (nt, n_bins, N) -> orbit states (nt, n_bins+1, N).  The
backward pass is the hand-written adjoint (transposed red-black sweeps in reverse
bin order inside GMRES on (I - M^T)): for output cotangents g_tau,
J = sum_tau <g_tau, c_tau> with c_8 = c_0, so R_tau = g_tau for tau = 0..7 plus
g_8 on R_0, and dJ/de_tau = dtau (lambda_tau - R_tau)
(orbit/core/orbit_adjoint.py:147-153).

PeriodicSolve: s (nt, N) -> c0 (nt, N) for (I - M) c0 = s; backward
lambda0 = (I - M^T)^-1 g, i.e. d(r.c0)/ds = lambda0 for a receptor r.

OrbitSolveFCT: the forward with the FCT deferred correction (n_passes passes,
orbit/cli.py:1212-1241), differentiated with the FCT coefficients frozen the
way the reference's adjoint freezes them (the limiter C and the van Leer gain
at the state the source was evaluated at: d_AD(c) = -L_AD c there,
orbit/core/fct.py:495-572).  Two backward modes:

* "passes" (default): the derivative of the forward itself.  Pass p solves
  with e + d_AD(c^(p-1)[tau+1]), so its tangent is
  dc^(p) = S(de - L_AD^(p-1) dc^(p-1)[tau+1]) with L_AD^(p-1) frozen at pass
  p-1's orbit; the backward runs one adjoint solve per pass in reverse and
  carries -L_AD^(p-1)^T of each pass's emission gradient to the previous
  pass.  Exact for the frozen-coefficient tangent of the n passes; the
  forward is homogeneous of degree one in e, so <grad, e> = J(e) exactly.
* "fixed_point": the gradient of the converged FCT fixed point c = S(e +
  d_AD(c)) with L_AD frozen at the forward's last orbit, by the reference's
  outer GMRES on mu (orbit/modes/adjoint.py:814-962; rel_tol, max_iters as
  there).  It is not the derivative of the n-pass iterate: the deferred
  correction map has a dominant eigenvalue about 1.33 in the reference's
  production runs (orbit/modes/adjoint.py:793-802), so the two can differ a
  lot.  A warning is raised when the outer GMRES does not converge.

The states the frozen operators need (every pass's orbit) are kept
(recompute=False) or recomputed by re-running the forward passes in the
backward pass (recompute=True) when memory asks for it.

No bin state is stored for OrbitSolve / PeriodicSolve: the adjoint needs none.
"""
from __future__ import annotations

import warnings

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
    def forward(ctx, e, op: OperatorSet, fct_grids, tracers, n_passes, tol, precision, recompute, mode, outer):
        if mode not in ("passes", "fixed_point"):
            raise ValueError("mode must be 'passes' or 'fixed_point'")
        e64 = e.detach().to(torch.float64).contiguous()
        orbit, hist, passes = drivers.forward_fct(op, fct_grids, e64, tracers, n_passes=n_passes, tol=tol,
                                                  precision=precision, keep_passes=True)
        ctx.op, ctx.grids, ctx.tracers = op, fct_grids, tracers
        ctx.n_passes, ctx.tol, ctx.precision, ctx.recompute = n_passes, tol, precision, recompute
        ctx.mode, ctx.outer = mode, dict(outer or {})
        ctx.infos = hist
        if recompute:
            ctx.save_for_backward(e64)
        else:
            ctx.save_for_backward(orbit, passes)
        return orbit

    @staticmethod
    def _baselines(ctx):
        if ctx.recompute:
            (e64,) = ctx.saved_tensors
            orbit, _, passes = drivers.forward_fct(ctx.op, ctx.grids, e64, ctx.tracers, n_passes=ctx.n_passes,
                                                   tol=ctx.tol, precision=ctx.precision, keep_passes=True)
            return orbit, passes
        return ctx.saved_tensors

    @staticmethod
    def backward(ctx, g):
        op = ctx.op
        nb = op.n_bins
        orbit, passes = OrbitSolveFCT._baselines(ctx)
        cot = g.to(torch.float64)
        if ctx.mode == "fixed_point":
            L = drivers.frozen_operators(op, ctx.grids, orbit)
            R = cot[:, :nb].clone()
            R[:, 0] += cot[:, nb]
            lam, Re, info = drivers.adjoint_fct(op, L, R.contiguous(), ctx.tracers, tol=min(ctx.tol, 1e-10),
                                                precision="fp64", **ctx.outer)
            if not info["converged"]:
                warnings.warn(f"FCT fixed-point gradient: the outer GMRES stopped at relative residual "
                              f"{info['rel_final']:.2e} after {info['outer_iters']} iterations")
            return op.dtau * (lam[:, :nb] - Re), None, None, None, None, None, None, None, None, None
        nt, N = cot.shape[0], op.N
        grad_e = torch.zeros((nt, nb, N), dtype=torch.float64, device=cot.device)
        tmp = torch.empty(N, dtype=torch.float64, device=cot.device)
        for p in reversed(range(ctx.n_passes)):
            R = cot[:, :nb].clone()
            R[:, 0] += cot[:, nb]
            lam, info = op.solve_adjoint(R.contiguous(), ctx.tracers, tol=ctx.tol, precision=ctx.precision)
            if not all(i["converged"] for i in info):
                warnings.warn(f"FCT pass gradient: adjoint solve of pass {p} did not converge")
            ebar = op.dtau * (lam[:, :nb] - R)
            grad_e += ebar
            if p == 0:
                break
            # pass p's source read pass p-1's end-of-bin states: e^(p)_tau =
            # e_tau + d_AD(c^(p-1)_(tau+1)), frozen derivative -L_AD,tau
            L = drivers.frozen_operators(op, ctx.grids, passes[p - 1])
            cot = torch.zeros((nt, nb + 1, N), dtype=torch.float64, device=g.device)
            for tau in range(nb):
                for s in range(nt):
                    drivers.apply_dev(L[tau][s], ebar[s, tau].contiguous(), tmp, trans=True)
                    cot[s, tau + 1] -= tmp
            for row in L:
                for Lop in row:
                    Lop.close()
        return grad_e, None, None, None, None, None, None, None, None, None


def orbit_solve(e, op, tracers=None, tol=1e-10, precision="fp64"):
    return OrbitSolve.apply(e, op, tracers, tol, precision)


def periodic_solve(s, op, tracers=None, tol=1e-10, precision="fp64"):
    return PeriodicSolve.apply(s, op, tracers, tol, precision)


def orbit_solve_fct(e, op, fct_grids, tracers=None, n_passes=6, tol=1e-10, precision="fp64",
                    recompute=False, mode="passes", rel_tol=2.0e-3, max_iters=40):
    """The forward FCT orbit (n_passes passes) with a backward of the given
    mode ("passes": the derivative of these passes; "fixed_point": the
    reference's converged-fixed-point gradient, outer GMRES to rel_tol within
    max_iters iterations).  Coefficients are frozen at the forward states."""
    return OrbitSolveFCT.apply(e, op, fct_grids, tracers, n_passes, tol, precision, recompute, mode,
                               {"rel_tol": rel_tol, "max_iters": max_iters})
