"""FCT deferred-correction drivers around the device periodic solve.

* forward_fct: the forward month's horizontal FCT as a RHS source recomputed
  from the previous pass's end-of-bin orbit, for a fixed number of passes
  (orbit/cli.py:1207-1241; pass 0 is the low-order orbit).
* marginal_fct: tangent-linear orbit with the frozen L_AD as a deferred RHS
  source, extra_rhs_tau = -L_AD,tau dc[tau+1], solved by outer GMRES on
  (I - K)x = b over the flattened 9-state orbits (orbit/modes/marginal.py:627-730).
* adjoint_fct: adjoint orbit with the transposed frozen L_AD as a receptor
  source, nu_tau = -dtau L_AD,(tau-1)^T mu_(tau-1), mu = lambda - R, outer GMRES on
  mu (orbit/modes/adjoint.py:814-962).

K is not a contraction in the reference's production runs (dominant eigenvalue
about 1.33, orbit/modes/adjoint.py:793-802), hence GMRES rather than Picard.
All vectors are CUDA float64 tensors; every inner solve runs on the device.
"""
from __future__ import annotations

import math
from typing import List, Sequence

import torch

from .core import N_BINS, OperatorSet
from .fct import FctGrid, FctLinearOperator, apply_dev


def make_fct_grids(grids: Sequence, ctx=None) -> List[FctGrid]:
    return [FctGrid(g, ctx) for g in grids]


def forward_fct(op: OperatorSet, fct_grids: Sequence[FctGrid], e, tracers=None, n_passes: int = 6,
                tol: float = 1e-8, precision: str = "mixed", ghost=None, maxiter: int = 200,
                warm: str = "c0", ctx=None, native: bool = True, **kw):
    """e: (nt, n_bins, N) CUDA float64.  ghost: optional lateral inflow
    concentration per bin, the FCT value beyond the edge: an (nt, n_bins, N)
    tensor, or a list of nt (n_bins, N) tensors with None for a tracer without
    one (its edge faces stay closed; lateral_bc.batch_inputs builds it).
    warm: "orbit" starts each pass from the whole
    previous orbit (mixed precision), "c0" from its c0 only.  ctx: the context
    (stream) the solves run on, the FCT grids' own (default: the operator set's).
    keep_passes (native only): also return the orbits of passes 0 .. n_passes-2,
    (n_passes - 1, nt, n_bins+1, N), the states each later pass's FCT source
    was evaluated at: (orbit, infos, passes).
    native: with warm "c0", run every pass inside the library
    (orbit_forward_fct_dev, no return to Python between passes; a ghost list
    goes in as one tensor with per-tracer flags; FP64: the same launches as
    the loop here; mixed: each later pass starts every bin state from the
    previous orbit and carries s~ over, so no sweep measures ||s|| again);
    with native=False the passes are driven from here.
    Returns (orbit (nt, n_bins+1, N), [infos per pass])."""
    ctx = ctx or op.ctx
    if any(g.ctx is not ctx for g in fct_grids):
        raise ValueError("FCT grids and the solves must share one context (stream)")
    if native and warm == "c0":
        return _forward_fct_native(op, fct_grids, e, tracers, n_passes, tol, precision, ghost, maxiter, ctx, **kw)
    if kw.pop("keep_passes", False):
        raise ValueError("keep_passes needs the native passes (native=True, warm='c0')")
    kw["ctx"] = ctx
    orbit, info = op.solve_forward(e, tracers, tol=tol, precision=precision, maxiter=maxiter, **kw)
    hist = [info]
    groups = _ghost_groups(ghost, e)
    for _ in range(1, n_passes):
        e_eff = torch.empty_like(e)
        for tau in range(op.n_bins):
            # queued on the solves' stream: the next solve reads e_eff in order
            if groups is None:
                fct_grids[tau].source_batch(orbit[:, tau + 1], e_eff[:, tau], base=e[:, tau], dt=op.dtau,
                                            ghost=None if ghost is None else ghost[:, tau], sync=False)
                continue
            for idx, gh in groups:  # tracers with and without a ghost: one launch each
                out = torch.empty((idx.numel(), e.shape[2]), dtype=e.dtype, device=e.device)
                # sync: torch reads `out` on its own stream next
                fct_grids[tau].source_batch(orbit[idx, tau + 1], out, base=e[idx, tau], dt=op.dtau,
                                            ghost=None if gh is None else gh[:, tau].contiguous(), sync=True)
                e_eff[:, tau].index_copy_(0, idx, out)
        if precision == "mixed" and warm == "orbit":
            # the previous orbit starts every bin state of the refinement
            orbit, info = op.solve_forward(e_eff, tracers, c_warm=orbit, tol=tol, precision=precision,
                                           maxiter=maxiter, **kw)
        else:
            orbit, info = op.solve_forward(e_eff, tracers, c_warm=orbit[:, 0], tol=tol, precision=precision,
                                           maxiter=maxiter, **kw)
        hist.append(info)
    return orbit, hist


def _forward_fct_native(op, fct_grids, e, tracers, n_passes, tol, precision, ghost, maxiter, ctx,
                        restart=30, inner_sweeps=0, acc_sweeps=0, use_graph=None, max_rounds=0, verbose=False,
                        keep_passes=False, check_sweeps=0):
    import ctypes as C
    import numpy as np
    from . import _lib
    from ._lib import check, lib
    tracers = list(range(op.n_tracers)) if tracers is None else list(tracers)
    nt = len(tracers)
    _lib.dev_f64(e, ctx.device, "e")
    if tuple(e.shape) != (nt, op.n_bins, op.N):
        raise ValueError(f"e must have shape {(nt, op.n_bins, op.N)}, got {tuple(e.shape)}")
    if len(fct_grids) != op.n_bins:
        raise ValueError("one FCT grid per bin")
    e = e.contiguous()
    flags = None
    if ghost is not None and not torch.is_tensor(ghost):
        # per-tracer list, None for a tracer without boundary values: one tensor
        # (zeros where there is none) and the flags that say which rows to read
        if len(ghost) != nt:
            raise ValueError("ghost needs one entry per tracer")
        if all(g is None for g in ghost):
            ghost = None
        else:
            flags = np.ascontiguousarray([g is not None for g in ghost], dtype=np.int32)
            ghost = torch.stack([torch.zeros_like(e[q]) if g is None else g.to(e.device, torch.float64)
                                 for q, g in enumerate(ghost)])
    if ghost is not None:
        if tuple(ghost.shape) != tuple(e.shape):
            raise ValueError("ghost must have the shape of e")
        ghost = ghost.to(e.device, torch.float64).contiguous()
    opts = op._opts(tol, maxiter, restart, precision, inner_sweeps, acc_sweeps, use_graph, max_rounds, verbose,
                    check=check_sweeps)
    tr = np.ascontiguousarray(tracers, dtype=np.int32)
    grids = (C.c_void_p * op.n_bins)(*[g.handle.value for g in fct_grids])
    info = (_lib.Info * (n_passes * nt))()
    out = torch.empty((nt, op.n_bins + 1, op.N), dtype=torch.float64, device=e.device)
    torch.cuda.current_stream(e.device).synchronize()  # the library runs on its own stream
    args = (ctx.handle, op.handle, grids, tr.ctypes.data, nt, e.data_ptr(), None if ghost is None else ghost.data_ptr(),
            None if flags is None else flags.ctypes.data, int(n_passes), C.byref(opts), out.data_ptr(), info)
    passes = None
    if keep_passes:
        passes = torch.empty((max(n_passes - 1, 0), nt, op.n_bins + 1, op.N), dtype=torch.float64, device=e.device)
        rc = lib().orbit_forward_fct_passes_dev(*args, passes.data_ptr() if n_passes > 1 else None)
    else:
        rc = lib().orbit_forward_fct_dev(*args)
    check(rc)
    hist = [[info[p * nt + s].as_dict() for s in range(nt)] for p in range(n_passes)]
    return (out, hist, passes) if keep_passes else (out, hist)


def _ghost_groups(ghost, e):
    """None for no ghost or one tensor for all tracers; else [(index tensor,
    stacked ghosts or None)] for the tracers with and those without a ghost."""
    if ghost is None or torch.is_tensor(ghost):
        return None
    if len(ghost) != e.shape[0]:
        raise ValueError("ghost needs one entry per tracer")
    with_g = [q for q, g in enumerate(ghost) if g is not None]
    without = [q for q, g in enumerate(ghost) if g is None]
    out = []
    if without:
        out.append((torch.tensor(without, device=e.device), None))
    if with_g:
        out.append((torch.tensor(with_g, device=e.device),
                    torch.stack([ghost[q].to(e.device, torch.float64) for q in with_g])))
    return out


def frozen_operators(op: OperatorSet, fct_grids: Sequence[FctGrid], orbit_base) -> List[List[FctLinearOperator]]:
    """L_AD[tau][slot] frozen at the baseline end-of-bin state orbit_base[slot, tau+1]
    (orbit/modes/marginal.py:450-473)."""
    out = []
    for tau in range(op.n_bins):
        row = []
        for s in range(orbit_base.shape[0]):
            cb = orbit_base[s, tau + 1].contiguous()
            row.append(fct_grids[tau].linop(cb, dt=op.dtau))
        out.append(row)
    return out


def gmres_torch(matvec, b, x0, atol: float, restart: int, max_cycles: int, callback=None):
    """Restarted GMRES(restart) with modified Gram-Schmidt and Givens rotations on
    device vectors (scipy.sparse.linalg.gmres semantics for rtol=0: stop when the
    residual estimate is <= atol; max_cycles restart cycles).  Returns (x, code, iters)."""
    x = x0.clone()
    iters = 0
    for _ in range(max_cycles):
        r = b - matvec(x)
        beta = float(torch.linalg.vector_norm(r))
        if beta <= atol:
            return x, 0, iters
        V = [r / beta]
        H = torch.zeros((restart + 1, restart), dtype=torch.float64)
        cs = torch.zeros(restart, dtype=torch.float64)
        sn = torch.zeros(restart, dtype=torch.float64)
        g = torch.zeros(restart + 1, dtype=torch.float64)
        g[0] = beta
        k = 0
        for j in range(restart):
            w = matvec(V[j])
            for i in range(j + 1):
                h = float(torch.dot(V[i], w))
                H[i, j] = h
                w = w - h * V[i]
            hn = float(torch.linalg.vector_norm(w))
            H[j + 1, j] = hn
            V.append(w / hn if hn > 0 else w)
            for i in range(j):
                a, bb = float(H[i, j]), float(H[i + 1, j])
                H[i, j] = cs[i] * a + sn[i] * bb
                H[i + 1, j] = -sn[i] * a + cs[i] * bb
            a, bb = float(H[j, j]), float(H[j + 1, j])
            rr = math.hypot(a, bb)
            cs[j], sn[j] = (a / rr, bb / rr) if rr > 0 else (1.0, 0.0)
            H[j, j] = rr
            H[j + 1, j] = 0.0
            g[j + 1] = -sn[j] * g[j]
            g[j] = cs[j] * g[j]
            iters += 1
            k = j + 1
            res = abs(float(g[j + 1]))
            if callback is not None:
                callback(res)
            if res <= atol:
                break
        y = torch.linalg.solve_triangular(H[:k, :k], g[:k, None], upper=True)[:, 0]
        for i in range(k):
            x = x + float(y[i]) * V[i]
    r = b - matvec(x)
    return x, 0 if float(torch.linalg.vector_norm(r)) <= atol else 1, iters


def marginal_fct(op: OperatorSet, L_AD, de, tracers=None, tol: float = 1e-10, precision: str = "fp64",
                 max_iters: int = 40, rel_tol: float = 2.0e-3, verbose: bool = False):
    """Tangent-linear orbits dc (nt, n_bins+1, N) with the frozen FCT source."""
    nt, nb, N = de.shape
    tmp = torch.empty(N, dtype=torch.float64, device=de.device)

    def solve_with(prev):
        e = de.clone()
        if prev is not None:
            prev = prev.view(nt, nb + 1, N)
            for tau in range(nb):
                for s in range(nt):
                    apply_dev(L_AD[tau][s], prev[s, tau + 1].contiguous(), tmp)
                    e[s, tau] -= tmp
        orbit, _ = op.solve_forward(e, tracers, tol=tol, precision=precision)
        return orbit.reshape(-1)

    b = solve_with(None)
    b_norm = float(torch.linalg.vector_norm(b))
    restart = min(max_iters, 30)
    hist = []
    x, code, iters = gmres_torch(lambda v: v - (solve_with(v) - b), b, b, 0.3 * rel_tol * b_norm, restart,
                                 max(1, -(-max_iters // restart)), callback=hist.append)
    g = solve_with(x)
    rel = float((g - x).abs().max() / max(float(g.abs().max()), 1e-30))
    if verbose:
        print(f"  [FCT] marginal outer GMRES: {iters} iters, max rel dc residual {rel:.3e}")
    return g.view(nt, nb + 1, N), {"outer_iters": iters, "rel_final": rel, "converged": rel < rel_tol,
                                   "history": hist}


def adjoint_fct(op: OperatorSet, L_AD, R, tracers=None, tol: float = 1e-10, precision: str = "fp64",
                max_iters: int = 40, rel_tol: float = 2.0e-3, verbose: bool = False):
    """Adjoint orbits with the transposed frozen FCT source.  Returns
    (lam (nt, n_bins+1, N), R_eff (nt, n_bins, N), info); dJ/de_tau = dtau (lam_tau - R_eff_tau)."""
    nt, nb, N = R.shape
    tmp = torch.empty(N, dtype=torch.float64, device=R.device)

    def run(mu):
        Re = R.clone()
        if mu is not None:
            mu = mu.view(nt, nb, N)
            for tau in range(nb):
                src = (tau - 1) % nb
                for s in range(nt):
                    apply_dev(L_AD[src][s], mu[s, src].contiguous(), tmp, trans=True)
                    Re[s, tau] -= op.dtau * tmp
        lam, _ = op.solve_adjoint(Re, tracers, tol=tol, precision=precision)
        return (lam[:, :nb] - Re).reshape(-1), lam, Re

    b = run(None)[0]
    b_norm = float(torch.linalg.vector_norm(b))
    restart = min(max_iters, 30)
    hist = []
    x, code, iters = gmres_torch(lambda v: v - (run(v)[0] - b), b, b, 0.3 * rel_tol * b_norm, restart,
                                 max(1, -(-max_iters // restart)), callback=hist.append)
    g, lam, Re = run(x)
    rel = float((g - x).abs().max() / max(float(g.abs().max()), 1e-30))
    if verbose:
        print(f"  [FCT] adjoint outer GMRES: {iters} iters, max rel mu residual {rel:.3e}")
    return lam, Re, {"outer_iters": iters, "rel_final": rel, "converged": rel < rel_tol, "history": hist}
