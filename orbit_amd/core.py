"""Device context and per-month operator sets on the MI355X.

``Context`` owns one HIP stream (orbit_ctx); ``OperatorSet`` owns the device form
of one month's 8 per-bin operators for a set of tracers (orbit_opset): shared
column blocks and 4-point stencils, per-tracer diagonals and the factored
(inverted) column blocks in FP64 and BF16, plus their transposes for the adjoint.

Replaces the reference's factor backend and LU cache: _factor_one_bin x 8 per
species (orbit/core/orbit.py:79-96, 330-349) and _UmfpackLU
(orbit/core/solve.py:210-283).
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np
import scipy.sparse as sp

from . import _lib
from ._lib import check, lib, ptr

N_BINS = 8
DTAU = 10800.0


class Context:
    """One device and one HIP stream; not thread-safe (one per host thread)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(lib().orbit_ctx_create(device, C.byref(h)))
        self.handle = h
        self.device = device

    def sync(self):
        check(lib().orbit_ctx_sync(self.handle))

    @property
    def stream(self) -> int:
        return lib().orbit_ctx_stream(self.handle) or 0

    def close(self):
        if self.handle:
            lib().orbit_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = {}


def default_context(device: int = 0) -> Context:
    if device not in _default_ctx:
        _default_ctx[device] = Context(device)
    return _default_ctx[device]


def _csr32(m):
    m = sp.csr_matrix(m)
    m.sum_duplicates()
    m.sort_indices()
    return (np.ascontiguousarray(m.indptr, dtype=np.int32),
            np.ascontiguousarray(m.indices, dtype=np.int32),
            np.ascontiguousarray(m.data, dtype=np.float64))


class OperatorSet:
    """Device-resident operators of one month: n_bins x n_tracers.

    Build from per-(bin, tracer) CSR matrices L (``from_csr``) or from the
    shared-transport form T_bin + diag(D_{bin,t}) (``from_shared``).
    """

    def __init__(self, ctx: Context, handle, nz, ny, nx, n_bins, n_tracers, dtau):
        self.ctx = ctx
        self.handle = handle
        self.nz, self.ny, self.nx = nz, ny, nx
        self.n_bins = n_bins
        self.n_tracers = n_tracers
        self.dtau = dtau
        self.N = nz * ny * nx

    @classmethod
    def from_csr(cls, L_per_bin: Sequence[Sequence], nz: int, ny: int, nx: int,
                 dtau: float = DTAU, ctx: Optional[Context] = None) -> "OperatorSet":
        """L_per_bin[bin][tracer]: (N, N) sparse L (A = I + dtau*L)."""
        ctx = ctx or default_context()
        nb = len(L_per_bin)
        nt = len(L_per_bin[0])
        keep, rp, ci, vv = [], [], [], []
        for b in range(nb):
            if len(L_per_bin[b]) != nt:
                raise ValueError("every bin needs the same number of tracers")
            for t in range(nt):
                a = _csr32(L_per_bin[b][t])
                keep.append(a)
                rp.append(ptr(a[0])); ci.append(ptr(a[1])); vv.append(ptr(a[2]))
        arr = lambda xs: (C.c_void_p * len(xs))(*xs)
        h = C.c_void_p()
        check(lib().orbit_opset_create_csr(ctx.handle, nz, ny, nx, nb, nt, float(dtau),
                                           arr(rp), arr(ci), arr(vv), C.byref(h)), allow_warn=False)
        return cls(ctx, h, nz, ny, nx, nb, nt, dtau)

    @classmethod
    def from_shared(cls, T_per_bin: Sequence, D: np.ndarray, nz: int, ny: int, nx: int,
                    dtau: float = DTAU, ctx: Optional[Context] = None) -> "OperatorSet":
        """T_per_bin[bin]: shared (N, N) transport; D: (n_bins, n_tracers, N) diagonal."""
        ctx = ctx or default_context()
        nb = len(T_per_bin)
        D = np.ascontiguousarray(D, dtype=np.float64)
        if D.ndim != 3 or D.shape[0] != nb:
            raise ValueError("D must be (n_bins, n_tracers, N)")
        nt = D.shape[1]
        keep = [_csr32(T) for T in T_per_bin]
        arr = lambda xs: (C.c_void_p * len(xs))(*xs)
        h = C.c_void_p()
        check(lib().orbit_opset_create_shared(
            ctx.handle, nz, ny, nx, nb, nt, float(dtau),
            arr([ptr(a[0]) for a in keep]), arr([ptr(a[1]) for a in keep]),
            arr([ptr(a[2]) for a in keep]), ptr(D), C.byref(h)), allow_warn=False)
        return cls(ctx, h, nz, ny, nx, nb, nt, dtau)

    @classmethod
    def from_grids(cls, grids: Sequence, D: np.ndarray, dtau: float = DTAU,
                   ctx: Optional[Context] = None) -> "OperatorSet":
        """grids[bin]: GridData-like fields of each bin; D: (n_bins, n_tracers, N)
        per-tracer diagonal.  The transport blocks are assembled on the device
        (no host CSR); L_{bin,t} = T_bin + diag(D_{bin,t})."""
        from . import assembly
        ctx = ctx or default_context()
        D = np.ascontiguousarray(D, dtype=np.float64)
        g0 = grids[0]
        N = int(g0.nz) * int(g0.ny) * int(g0.nx)
        if D.ndim != 3 or D.shape[0] != len(grids) or D.shape[2] != N:
            raise ValueError("D must be (n_bins, n_tracers, N)")
        h = assembly.create_opset_handle(ctx, grids, D, dtau)
        return cls(ctx, h, int(g0.nz), int(g0.ny), int(g0.nx), len(grids), D.shape[1], dtau)

    @classmethod
    def from_workload(cls, wl, ctx: Optional[Context] = None) -> "OperatorSet":
        return cls.from_shared(wl.T, wl.D, wl.nz, wl.ny, wl.nx, wl.dtau, ctx)

    def set_relaxation(self, omega: float = 0.0) -> dict:
        """Relaxation of the red-black sweeps' correction x += omega P (b - A x):
        0 chooses it by measuring the sweep contraction (as at creation), 1 is
        plain block Gauss-Seidel, any value in [0.5, 1.5] is measured and used.
        The fixed point A x = b does not depend on omega; the contraction and
        the sweep counts do.  Returns the new stats."""
        check(lib().orbit_opset_set_relaxation(self.ctx.handle, self.handle, float(omega)), allow_warn=False)
        return self.stats()

    def set_mfma(self, on: bool = True) -> None:
        """Block product of the reduced-precision sweeps on the matrix cores
        (on) or on the VALU (the default, 2-3 % faster per launch at
        64x70x15); results are batch-invariant either way and differ between
        the two in the last bits."""
        check(lib().orbit_opset_set_mfma(self.handle, int(bool(on))))

    def set_patch(self, shape: int) -> None:
        """The first sweep (from x = 0) of every FP32 inner solve in one
        launch over column patches of the given shape (1-4, see
        orbit_opset_set_patch; 2 is the default), or 0 for two half-sweep
        launches; the same bits either way."""
        check(lib().orbit_opset_set_patch(self.handle, int(shape)))

    def set_handoff(self, on: bool) -> None:
        """Run each bin's last black FP32 half-sweep inside the next bin's
        first-sweep launch (on, the default; see orbit_opset_set_handoff) or
        as a launch of its own; the same bits either way."""
        check(lib().orbit_opset_set_handoff(self.handle, int(bool(on))))

    def stats(self) -> dict:
        s = _lib.OpsetStats()
        check(lib().orbit_opset_stats_get(self.handle, C.byref(s)))
        return {k: getattr(s, k) for k, _ in s._fields_}

    CHECK_FIELDS = ("V64", "V32", "VT64", "VT32", "H64", "H32", "HT64", "HT32", "d64", "d32",
                    "P64", "P16", "PT64", "PT16")

    def check(self) -> dict:
        """Non-finite entry counts of every stored array (0 everywhere when healthy)."""
        c = (C.c_longlong * 14)()
        check(lib().orbit_opset_check(self.ctx.handle, self.handle, c))
        return dict(zip(self.CHECK_FIELDS, list(c)))

    # -- sweep launch shapes (autotuner) ----------------------------------
    SWEEP_PRECS = ("fp64", "acc_mixed", "f32_inner")
    SLOT_BUCKETS = ("1", "2-3", "4-7", "8-15", "16+")

    @staticmethod
    def _table_to_list(t) -> list:
        return [(e.slots_per_wave, e.threads, e.xcd_map) for e in t]

    def autotune(self, reps: int = 40):
        """Time every half-sweep launch shape per sweep precision and slot bucket on
        this operator set and keep the fastest.  Returns (table, us): table is a
        flat list of (slots_per_wave, threads, xcd_map), 3 precisions x 5 buckets,
        us the microseconds per launch of each chosen shape."""
        t = _lib.ShapeTable()
        us = (C.c_double * len(t))()
        check(lib().orbit_opset_autotune(self.ctx.handle, self.handle, int(reps), t, us))
        return self._table_to_list(t), list(us)

    def get_shapes(self) -> list:
        t = _lib.ShapeTable()
        check(lib().orbit_opset_get_shapes(self.handle, t))
        return self._table_to_list(t)

    def set_shapes(self, table) -> None:
        """Install a shape table (as returned by autotune/get_shapes), e.g. one tuned
        on another operator set of the same size class."""
        table = list(table)
        if len(table) != _lib.N_SWEEP_PRECS * _lib.N_SLOT_BUCKETS:
            raise ValueError("a shape table has 15 entries (3 precisions x 5 slot buckets)")
        t = _lib.ShapeTable(*[_lib.SweepShape(*map(int, e)) for e in table])
        check(lib().orbit_opset_set_shapes(self.handle, t))

    def update_diag(self, D: np.ndarray):
        D = np.ascontiguousarray(D, dtype=np.float64)
        check(lib().orbit_opset_update_diag(self.ctx.handle, self.handle, ptr(D)))

    def apply_P(self, bin: int, tracer: int, b: np.ndarray, trans: bool = False) -> np.ndarray:
        b = np.ascontiguousarray(b, dtype=np.float64)
        if b.shape != (self.N,):
            raise ValueError(f"b must have shape ({self.N},)")
        x = np.empty(self.N)
        check(lib().orbit_apply_P(self.ctx.handle, self.handle, bin, tracer, int(trans), ptr(b), ptr(x)))
        return x

    # -- periodic solves ----------------------------------------------------
    def _solve(self, fn, tracers, inp, warm, opts, device: bool, ctx: Optional[Context] = None):
        ctx = ctx or self.ctx
        tracers = np.ascontiguousarray(tracers, dtype=np.int32)
        nt = len(tracers)
        info = (_lib.Info * nt)()
        in_shape = (nt, self.N) if opts.rhs_is_s else (nt, self.n_bins, self.N)
        if tuple(inp.shape) != in_shape:
            raise ValueError(f"input must have shape {in_shape}, got {tuple(inp.shape)}")
        if device:
            import torch
            inp = _lib.dev_f64(inp, ctx.device, "input").contiguous()
            out = torch.empty((nt, self.n_bins + 1, self.N), dtype=torch.float64, device=inp.device)
            wt = None
            if warm is not None:
                wshape = (nt, self.n_bins + 1, self.N) if opts.warm_orbit else (nt, self.N)
                if not torch.is_tensor(warm):
                    warm = torch.as_tensor(warm)
                # kept alive for the call
                wt = warm.to(inp.device, torch.float64).reshape(wshape).contiguous()
            # the library runs on its own stream: order it after torch's work
            torch.cuda.current_stream(inp.device).synchronize()
            rc = fn(ctx.handle, self.handle, ptr(tracers), nt, inp.data_ptr(),
                    None if wt is None else wt.data_ptr(), C.byref(opts), out.data_ptr(), info)
        else:
            inp = np.ascontiguousarray(inp, dtype=np.float64)
            out = np.empty((nt, self.n_bins + 1, self.N))
            wp = None
            if warm is not None:
                wshape = (nt, self.n_bins + 1, self.N) if opts.warm_orbit else (nt, self.N)
                warm = np.ascontiguousarray(warm, dtype=np.float64).reshape(wshape)
                wp = ptr(warm)
            rc = fn(ctx.handle, self.handle, ptr(tracers), nt, ptr(inp), wp,
                    C.byref(opts), ptr(out), info)
        check(rc)
        return out, [i.as_dict() for i in info]

    def solve_forward(self, e, tracers=None, c_warm=None, tol=1e-8, maxiter=200, restart=30,
                      precision="fp64", inner_sweeps=0, acc_sweeps=0, use_graph=None,
                      max_rounds=0, verbose=False, rhs="e", ctx: Optional[Context] = None,
                      warm_sweeps=0, check_sweeps=0):
        """Periodic orbits of tracers for emissions e (nt, n_bins, N).

        rhs="s": e is s itself, (nt, N), for (I - M)c0 = s.
        c_warm: (nt, N) start for c0, or a whole earlier orbit (nt, n_bins+1, N) whose
        states start every bin state of the mixed refinement (FP64 takes its c0).
        warm_sweeps: accepted for older callers, unused.  check_sweeps (mixed): sweeps per
        inner solve of each refinement round's residual sweep (0: the opset's default).
        Returns (orbit (nt, n_bins+1, N), info list).  e may be a numpy array
        (host) or a CUDA torch tensor (device path, no host copies)."""
        tracers = list(range(self.n_tracers)) if tracers is None else list(tracers)
        opts = self._opts(tol, maxiter, restart, precision, inner_sweeps, acc_sweeps,
                          use_graph, max_rounds, verbose, rhs, check_sweeps)
        if c_warm is not None and len(c_warm.shape) == 3:
            opts.warm_orbit = 1
        dev = hasattr(e, "is_cuda") and e.is_cuda
        fn = lib().orbit_solve_forward_dev if dev else lib().orbit_solve_forward
        return self._solve(fn, tracers, e, c_warm, opts, dev, ctx)

    def solve_adjoint(self, R, tracers=None, lam_warm=None, tol=1e-8, maxiter=200, restart=30,
                      precision="fp64", inner_sweeps=0, acc_sweeps=0, use_graph=None,
                      max_rounds=0, verbose=False, rhs="e", ctx: Optional[Context] = None,
                      check_sweeps=0):
        """Adjoint orbits for receptors R (nt, n_bins, N); returns (lam (nt, n_bins+1, N), info).

        rhs="s": R is s_adj itself, (nt, N), for (I - M^T) lam0 = s_adj."""
        tracers = list(range(self.n_tracers)) if tracers is None else list(tracers)
        opts = self._opts(tol, maxiter, restart, precision, inner_sweeps, acc_sweeps,
                          use_graph, max_rounds, verbose, rhs, check_sweeps)
        dev = hasattr(R, "is_cuda") and R.is_cuda
        fn = lib().orbit_solve_adjoint_dev if dev else lib().orbit_solve_adjoint
        return self._solve(fn, tracers, R, lam_warm, opts, dev, ctx)

    @staticmethod
    def _opts(tol, maxiter, restart, precision, inner, acc, use_graph, max_rounds, verbose, rhs="e", check=0):
        prec = {"fp64": _lib.ORBIT_PREC_FP64, "mixed": _lib.ORBIT_PREC_MIXED}[precision]
        if rhs not in ("e", "s"):
            raise ValueError("rhs must be 'e' or 's'")
        if use_graph is None:
            # direct launches: replaying the inexact monodromy's sweeps from
            # captured graphs measured 23.5 against 23.1 ms per 12-tracer
            # 64x70x15 FCT month step (the same bits either way)
            use_graph = False
        return _lib.default_opts(tol=float(tol), maxiter=int(maxiter), restart=int(restart),
                                 precision=prec, inner_sweeps=int(inner), acc_sweeps=int(acc),
                                 use_graph=int(bool(use_graph)), max_rounds=int(max_rounds),
                                 verbose=int(bool(verbose)), rhs_is_s=int(rhs == "s"), check_sweeps=int(check))

    def mass_balance(self, rho: np.ndarray) -> dict:
        """Air-mass-balance row-sum statistics (orbit/core/operator.py:598-652)."""
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        st = _lib.MbStats()
        check(lib().orbit_mass_balance(self.ctx.handle, self.handle, ptr(rho), C.byref(st)))
        out = {k: getattr(st, k) for k, _ in st._fields_ if not k.startswith("layer")}
        out["interior_frac_above_0.01"] = out.pop("interior_frac_above_0_01")
        out["interior_frac_above_0.1"] = out.pop("interior_frac_above_0_1")
        out["layer_p90"] = list(st.layer_p90[: self.nz])
        out["layer_mean_signed"] = list(st.layer_mean_signed[: self.nz])
        return out

    def close(self):
        if self.handle:
            lib().orbit_opset_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
