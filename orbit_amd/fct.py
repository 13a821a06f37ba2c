"""Horizontal Zalesak FCT on the MI355X with the reference's call surface.

compute_horizontal_fct_source(grid, indexer, c_flat, dt, limited, c_ghost_flat)
mirrors orbit/core/fct.py:459-492 and assemble_fct_linear_operator(grid,
indexer, c_baseline_flat, dt) mirrors orbit/core/fct.py:495-572; ``grid`` is any
object with the reference GridData fields the FCT reads (nz, ny, nx, dx, dy, dP,
Dz, UAvg/VAvg or the split UAvg_plus/minus, VAvg_plus/minus,
has_split_fluxes, concentration_state).  The frozen operator stays on the
device as per-face gains (FctLinearOperator): ``@``/matvec, rmatvec for the
adjoint, and ``tocsr()`` for a host copy.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import scipy.sparse as sp

from . import _lib
from ._lib import check, lib, ptr
from .core import Context, default_context


def _metrics(grid, keep):
    g = _lib.GridMetrics()
    g.nz, g.ny, g.nx = int(grid.nz), int(grid.ny), int(grid.nx)
    split = bool(getattr(grid, "has_split_fluxes", False))
    g.split_fluxes = int(split)
    g.concentration_state = int(bool(getattr(grid, "concentration_state", True)))

    def a(x):
        arr = np.ascontiguousarray(x, dtype=np.float64)
        keep.append(arr)
        return ptr(arr)

    if split:
        g.u_plus, g.u_minus = a(grid.UAvg_plus), a(grid.UAvg_minus)
        g.v_plus, g.v_minus = a(grid.VAvg_plus), a(grid.VAvg_minus)
    else:
        g.u_plus, g.v_plus = a(grid.UAvg), a(grid.VAvg)
        g.u_minus = g.v_minus = None
    g.dx = a(np.broadcast_to(np.asarray(grid.dx, dtype=np.float64), (grid.ny,)))
    g.dy = float(grid.dy)
    g.dP = a(grid.dP)
    Dz = getattr(grid, "Dz", None)
    g.Dz = a(Dz) if Dz is not None else None
    return g


def compute_horizontal_fct_source(grid, indexer, c_flat, dt, limited=True, c_ghost_flat=None,
                                  ctx: Context = None):
    """d_AD = rho*(-div(C*A)) for one species on the device (FP64)."""
    ctx = ctx or default_context()
    keep = []
    g = _metrics(grid, keep)
    N = g.nz * g.ny * g.nx
    c = np.ascontiguousarray(c_flat, dtype=np.float64).reshape(N)
    gh = None
    if c_ghost_flat is not None:
        gh = np.ascontiguousarray(c_ghost_flat, dtype=np.float64).reshape(N)
    d = np.empty(N)
    check(lib().orbit_fct_source(ctx.handle, C.byref(g), ptr(c), float(dt), int(bool(limited)),
                                 None if gh is None else ptr(gh), ptr(d)))
    return d


class FctLinearOperator:
    """Frozen-coefficient L_AD on the device (limiter and van Leer frozen at c_b)."""

    def __init__(self, grid, c_baseline_flat, dt, ctx: Context = None):
        self.ctx = ctx or default_context()
        keep = []
        g = _metrics(grid, keep)
        self.N = g.nz * g.ny * g.nx
        self.shape = (self.N, self.N)
        cb = np.ascontiguousarray(c_baseline_flat, dtype=np.float64).reshape(self.N)
        h = C.c_void_p()
        check(lib().orbit_fct_linop_create(self.ctx.handle, C.byref(g), ptr(cb), float(dt), C.byref(h)))
        self.handle = h

    def matvec(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(self.N)
        y = np.empty(self.N)
        check(lib().orbit_fct_linop_apply(self.ctx.handle, self.handle, 0, ptr(x), ptr(y)))
        return y

    def rmatvec(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(self.N)
        y = np.empty(self.N)
        check(lib().orbit_fct_linop_apply(self.ctx.handle, self.handle, 1, ptr(x), ptr(y)))
        return y

    def __matmul__(self, x):
        return self.matvec(x)

    @property
    def T(self):
        return _Transposed(self)

    def tocsr(self) -> sp.csr_matrix:
        rp = np.empty(self.N + 1, dtype=np.int32)
        ci = np.empty(5 * self.N, dtype=np.int32)
        vv = np.empty(5 * self.N, dtype=np.float64)
        nnz = C.c_int()
        check(lib().orbit_fct_linop_csr(self.ctx.handle, self.handle, ptr(rp), ptr(ci), ptr(vv),
                                        C.byref(nnz)))
        n = nnz.value
        return sp.csr_matrix((vv[:n].copy(), ci[:n].copy(), rp.copy()), shape=self.shape)

    def close(self):
        if getattr(self, "handle", None):
            lib().orbit_fct_linop_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _Transposed:
    def __init__(self, L):
        self.L = L
        self.shape = L.shape

    def __matmul__(self, x):
        return self.L.rmatvec(x)

    def matvec(self, x):
        return self.L.rmatvec(x)


def assemble_fct_linear_operator(grid, indexer, c_baseline_flat, dt, ctx: Context = None):
    return FctLinearOperator(grid, c_baseline_flat, dt, ctx)


class FctGrid:
    """Device-resident metrics of one bin for repeated, batched FCT sources."""

    def __init__(self, grid, ctx: Context = None):
        self.ctx = ctx or default_context()
        keep = []
        g = _metrics(grid, keep)
        self.N = g.nz * g.ny * g.nx
        h = C.c_void_p()
        check(lib().orbit_fct_grid_create(self.ctx.handle, C.byref(g), C.byref(h)))
        self.handle = h

    def source_batch(self, c, out, base=None, dt=10800.0, limited=True, ghost=None, sync=True):
        """out[v] = base[v] + d_AD(c[v]) for the leading index v of 2-D views
        (CUDA float64 tensors whose rows may be strided, e.g. orbit[:, tau+1]).
        sync=False leaves the launch queued on the context's stream (for a
        consumer on that same stream)."""
        nvec = c.shape[0]
        dev = self.ctx.device
        for name, t in (("c", c), ("out", out), ("base", base), ("ghost", ghost)):
            if t is None:
                continue
            _lib_dev_f64(t, dev, name, last=self.N, rows=nvec)
            if t.stride(-1) != 1:
                raise ValueError("FCT batch views need unit-stride rows of length N")
        if ghost is not None and nvec > 1 and ghost.stride(0) != c.stride(0):
            # the entry point reads ghost rows on c's stride: lay them out so,
            # and finish before the copy can be freed
            import torch
            g2 = torch.empty((nvec, c.stride(0)), dtype=ghost.dtype, device=ghost.device)[:, :self.N]
            g2.copy_(ghost)
            ghost = g2
            sync = True
        _torch_sync(c)
        check(lib().orbit_fct_source_batch_dev(
            self.ctx.handle, self.handle, nvec, c.data_ptr(), c.stride(0),
            None if ghost is None else ghost.data_ptr(), float(dt), int(bool(limited)),
            None if base is None else base.data_ptr(), 0 if base is None else base.stride(0),
            out.data_ptr(), out.stride(0)))
        if sync:
            self.ctx.sync()

    def linop(self, c_base, dt=10800.0):
        """Frozen L_AD at the device state c_base (CUDA float64, length N)."""
        L = FctLinearOperator.__new__(FctLinearOperator)
        L.ctx = self.ctx
        L.N = self.N
        L.shape = (self.N, self.N)
        h = C.c_void_p()
        c_base = _lib_dev_f64(c_base, self.ctx.device, "c_base", last=self.N).contiguous()
        if c_base.numel() != self.N:
            raise ValueError(f"c_base must hold N = {self.N} values")
        _torch_sync(c_base)
        check(lib().orbit_fct_linop_create_dev(self.ctx.handle, self.handle, c_base.data_ptr(), float(dt),
                                               C.byref(h)))
        L.handle = h
        return L

    def close(self):
        if getattr(self, "handle", None):
            lib().orbit_fct_grid_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _lib_dev_f64(t, device, name, last=None, rows=None):
    from ._lib import dev_f64
    return dev_f64(t, device, name, last=last, rows=rows)


def _torch_sync(t):
    """The library launches on its own stream: finish torch's pending work on t first."""
    import torch
    torch.cuda.current_stream(t.device).synchronize()


def apply_dev(L: FctLinearOperator, x, out, trans=False):
    """out = L_AD x (or L_AD^T x) for CUDA float64 vectors of length N (synchronous)."""
    for name, t in (("x", x), ("out", out)):
        _lib_dev_f64(t, L.ctx.device, name, last=L.N)
        if t.numel() != L.N or not t.is_contiguous():
            raise ValueError(f"{name} must be a contiguous vector of N = {L.N} values")
    _torch_sync(x)
    check(lib().orbit_fct_linop_apply_dev(L.ctx.handle, L.handle, int(bool(trans)), x.data_ptr(),
                                          out.data_ptr()))
    L.ctx.sync()
