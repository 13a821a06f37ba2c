"""CPU restatement of the Zalesak FCT pieces on the solve path (TEST ORACLE).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
bench CPU-baseline leg, never by the product path.

Reference (orbit/core/fct.py):
  vanleer                      :38-48
  low_order_flux_1d            :54-61     (1-D periodic harness)
  high_order_flux_1d           :64-105
  zalesak_coefficients_1d      :116-164
  upwind_step_1d / fct_step_1d :175-188
  fct_step_2d                  :201-222
  vanleer_phi                  :235-247
  face_fluxes (x and y)        :256-357  (_horizontal_antidiffusive)
  divergence_into_cells        :360-378  (_net_antidiffusive_into_cells)
  zalesak_2d                   :381-447
  compute_horizontal_fct_source:459-492
  assemble_fct_linear_operator :495-572
  density weights rho = dP/(g Dz) (orbit/core/grid_data.py:543-553)

Pinned by: the reference's own harness tests restated in tests/test_oracle_fct.py
(van Leer algebra, square-wave monotone+conservative, Gaussian convergence
orders, 2-D cone, uniform field, superposition, ghost/nested-domain and
no-new-extremum bounds, L_AD c = -d_AD, d_c = rho d_x), and for the exact 2-D
values on a metric + terrain + split-flux grid with a front and a ghost, by the
golden vectors in tests/golden/fct_golden.npz.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

GRAVITY = 9.80665
_EPS = 1e-30


# --------------------------------------------------------------------------- #
# limiter and 1-D periodic harness
# --------------------------------------------------------------------------- #
def vanleer(r):
    """phi(r) = (r + |r|)/(1 + |r|)."""
    r = np.asarray(r, dtype=np.float64)
    a = np.abs(r)
    return (r + a) / (1.0 + a)


def _ratio(num, den):
    """Guarded gradient ratio: |den| < eps -> 1 if |num| < eps else 1e6."""
    flat = np.abs(den) < _EPS
    r = num / np.where(flat, _EPS, den)
    return np.where(flat, np.where(np.abs(num) < _EPS, 1.0, 1.0e6), r)


def low_order_flux_1d(c, u_face):
    c_next = np.roll(c, -1)
    return np.maximum(u_face, 0.0) * c + np.minimum(u_face, 0.0) * c_next


def high_order_flux_1d(c, u_face):
    c_prev, c_next, c_next2 = np.roll(c, 1), np.roll(c, -1), np.roll(c, -2)
    left_up = u_face >= 0.0
    num = np.where(left_up, c - c_prev, c_next2 - c_next)
    phi = vanleer(_ratio(num, c_next - c))
    up = np.where(left_up, c, c_next)
    down = np.where(left_up, c_next, c)
    return u_face * (up + 0.5 * phi * (down - up))


def zalesak_coefficients_1d(c_prev, c_td, A, dx, dt):
    cand = np.vstack([c_prev, c_td, np.roll(c_prev, 1), np.roll(c_prev, -1),
                      np.roll(c_td, 1), np.roll(c_td, -1)])
    hi, lo = cand.max(axis=0), cand.min(axis=0)
    A_in = np.roll(A, 1)            # through the left face, into cell i
    p_pos = np.maximum(A_in, 0.0) + np.maximum(-A, 0.0)
    p_neg = np.maximum(-A_in, 0.0) + np.maximum(A, 0.0)
    q_pos = (hi - c_td) * dx / dt
    q_neg = (c_td - lo) * dx / dt
    r_pos = np.where(p_pos > 0, np.minimum(1.0, q_pos / np.where(p_pos > 0, p_pos, 1.0)), 0.0)
    r_neg = np.where(p_neg > 0, np.minimum(1.0, q_neg / np.where(p_neg > 0, p_neg, 1.0)), 0.0)
    C = np.where(A >= 0.0, np.minimum(np.roll(r_pos, -1), r_neg),
                 np.minimum(np.roll(r_neg, -1), r_pos))
    return np.clip(C, 0.0, 1.0)


def _div_1d(F, dx):
    return (F - np.roll(F, 1)) / dx


def upwind_step_1d(c, u_face, dx, dt):
    return c - dt * _div_1d(low_order_flux_1d(c, u_face), dx)


def fct_step_1d(c, u_face, dx, dt):
    F_lo = low_order_flux_1d(c, u_face)
    c_td = c - dt * _div_1d(F_lo, dx)
    A = high_order_flux_1d(c, u_face) - F_lo
    C = zalesak_coefficients_1d(c, c_td, A, dx, dt)
    return c_td - dt * _div_1d(C * A, dx)


def fct_step_2d(c, ux, uy, dx, dy, dt):
    """x then y split of fct_step_1d on a periodic (ny, nx) field."""
    out = np.array([fct_step_1d(row, np.full(row.shape, ux), dx, dt) for row in c])
    out = np.array([fct_step_1d(col, np.full(col.shape, uy), dy, dt) for col in out.T]).T
    return out


# --------------------------------------------------------------------------- #
# grid-aware horizontal source
# --------------------------------------------------------------------------- #
def vanleer_phi(c_up, c_far, c_down):
    """Frozen limiter value at a face: phi(r), r = (c_up - c_far)/(c_down - c_up)."""
    return vanleer(_ratio(c_up - c_far, c_down - c_up))


def density_weights(grid):
    ok = (grid.dP > 0) & (grid.Dz > 0)
    return np.where(ok, grid.dP / (GRAVITY * np.where(ok, grid.Dz, 1.0)), 1.0).ravel()


def state_weights(grid, N):
    if getattr(grid, "concentration_state", True):
        return density_weights(grid)
    return np.ones(N)


def _face_speeds(grid, axis):
    if axis == 2:
        if grid.has_split_fluxes:
            return grid.UAvg_plus[:, :, 1:], grid.UAvg_minus[:, :, 1:]
        u = grid.UAvg[:, :, 1:]
    else:
        if grid.has_split_fluxes:
            return grid.VAvg_plus[:, 1:, :], grid.VAvg_minus[:, 1:, :]
        u = grid.VAvg[:, 1:, :]
    return np.maximum(u, 0.0), np.maximum(-u, 0.0)


def face_fluxes(grid, x, ghost=None):
    """Anti-diffusive face fluxes and metric weights on interior faces.

    Returns {axis: dict(A, beta, tr_lo, tr_hi, inv, valid)} for axis in
    {'x', 'y'} where 'lo' is the west/south cell and 'hi' the east/north cell.
    """
    nz, ny, nx = grid.nz, grid.ny, grid.nx
    faces = {}
    dx3 = np.broadcast_to(np.asarray(grid.dx, dtype=np.float64)[None, :, None], (nz, ny, nx))
    for name, axis, n in (("x", 2, nx), ("y", 1, ny)):
        if n < 2:
            continue
        sl = lambda a, b: tuple(slice(a, b) if d == axis else slice(None) for d in range(3))
        lo, hi = x[sl(0, n - 1)], x[sl(1, n)]
        # far-upwind neighbours; at the domain edge the edge cell itself (or the ghost)
        lolo = np.concatenate([x[sl(0, 1)] if ghost is None else ghost[sl(0, 1)],
                               x[sl(0, n - 2)]], axis=axis)
        hihi = np.concatenate([x[sl(2, n)],
                               x[sl(n - 1, n)] if ghost is None else ghost[sl(n - 1, n)]],
                              axis=axis)
        phi_lo = vanleer_phi(lo, lolo, hi)        # flow lo -> hi, lo is upwind
        phi_hi = vanleer_phi(hi, hihi, lo)        # flow hi -> lo, hi is upwind
        if ghost is None:
            phi_lo[sl(0, 1)] = 0.0
            phi_hi[sl(n - 2, n - 1)] = 0.0
        sp_, sm_ = _face_speeds(grid, axis)
        beta = 0.5 * (sp_ * phi_lo + sm_ * phi_hi)
        A = beta * (hi - lo)
        dP_lo, dP_hi = grid.dP[sl(0, n - 1)], grid.dP[sl(1, n)]
        dP_f = 0.5 * (dP_lo + dP_hi)
        valid = (dP_lo > 0) & (dP_hi > 0)
        w_lo = dP_f / np.where(dP_lo > 0, dP_lo, 1.0)
        w_hi = dP_f / np.where(dP_hi > 0, dP_hi, 1.0)
        if axis == 2:
            dxf = dx3[sl(1, n)]
            inv = 1.0 / np.where(dxf > 0, dxf, 1.0)
        else:
            dx_lo, dx_hi = dx3[sl(0, n - 1)], dx3[sl(1, n)]
            dx_f = 0.5 * (dx_lo + dx_hi)
            w_lo = w_lo * dx_f / np.where(dx_lo > 0, dx_lo, 1.0)
            w_hi = w_hi * dx_f / np.where(dx_hi > 0, dx_hi, 1.0)
            inv = (1.0 / grid.dy) if grid.dy > 0 else 1.0
        tr_lo = np.where(valid, w_lo, 1.0)
        tr_hi = np.where(valid, w_hi, 1.0)
        faces[name] = dict(A=A, beta=beta, tr_lo=tr_lo, tr_hi=tr_hi, inv=inv,
                           valid=valid, axis=axis, n=n)
    return faces


def _shift(a, axis, n, hi):
    return tuple(slice(1, n) if (d == axis and hi) else
                 slice(0, n - 1) if d == axis else slice(None) for d in range(3))


def divergence_into_cells(shape, faces, C=None):
    """d = -div(C*A) with the terrain/metric weights; C=None means C=1."""
    d = np.zeros(shape)
    for name, f in faces.items():
        c = 1.0 if C is None else C[name]
        flux = c * f["A"] * f["inv"]
        d[_shift(d, f["axis"], f["n"], True)] += np.where(f["valid"], f["tr_hi"] * flux, 0.0)
        d[_shift(d, f["axis"], f["n"], False)] -= np.where(f["valid"], f["tr_lo"] * flux, 0.0)
    return d


def zalesak_2d(x, faces, dt, ghost=None):
    """Per-face limiter C in [0, 1] so x + dt*d makes no new extremum over the
    cell and its 4 neighbours (ghost or clamped beyond the edge)."""
    g = x if ghost is None else ghost
    nbrs = [x,
            np.concatenate([g[:, :, :1], x[:, :, :-1]], axis=2),
            np.concatenate([x[:, :, 1:], g[:, :, -1:]], axis=2),
            np.concatenate([g[:, :1, :], x[:, :-1, :]], axis=1),
            np.concatenate([x[:, 1:, :], g[:, -1:, :]], axis=1)]
    st = np.stack(nbrs)
    x_max, x_min = st.max(axis=0), st.min(axis=0)
    p_pos = np.zeros(x.shape)
    p_neg = np.zeros(x.shape)
    into = {}
    for name, f in faces.items():
        flux = f["A"] * f["inv"]
        in_hi = np.where(f["valid"], f["tr_hi"] * flux, 0.0)
        in_lo = np.where(f["valid"], -f["tr_lo"] * flux, 0.0)
        hs, ls = _shift(x, f["axis"], f["n"], True), _shift(x, f["axis"], f["n"], False)
        p_pos[hs] += np.maximum(in_hi, 0.0)
        p_neg[hs] += np.maximum(-in_hi, 0.0)
        p_pos[ls] += np.maximum(in_lo, 0.0)
        p_neg[ls] += np.maximum(-in_lo, 0.0)
        into[name] = in_hi
    q_pos = (x_max - x) / dt
    q_neg = (x - x_min) / dt
    r_pos = np.where(p_pos > 0, np.minimum(1.0, q_pos / np.where(p_pos > 0, p_pos, 1.0)), 0.0)
    r_neg = np.where(p_neg > 0, np.minimum(1.0, q_neg / np.where(p_neg > 0, p_neg, 1.0)), 0.0)
    C = {}
    for name, f in faces.items():
        hs, ls = _shift(x, f["axis"], f["n"], True), _shift(x, f["axis"], f["n"], False)
        c = np.where(into[name] >= 0.0, np.minimum(r_pos[hs], r_neg[ls]),
                     np.minimum(r_neg[hs], r_pos[ls]))
        C[name] = np.clip(c, 0.0, 1.0)
    return C


def compute_horizontal_fct_source(grid, c_flat, dt, limited=True, c_ghost_flat=None):
    """d_AD = rho * (-div(C*A)) evaluated on x = c/rho (fct.py:459-492)."""
    nz, ny, nx = grid.nz, grid.ny, grid.nx
    N = nz * ny * nx
    rho = state_weights(grid, N)
    x = (np.asarray(c_flat, dtype=np.float64) / rho).reshape(nz, ny, nx)
    ghost = None
    if c_ghost_flat is not None:
        ghost = (np.asarray(c_ghost_flat, dtype=np.float64) / rho).reshape(nz, ny, nx)
    faces = face_fluxes(grid, x, ghost)
    C = zalesak_2d(x, faces, dt, ghost) if limited else None
    return divergence_into_cells(x.shape, faces, C).ravel() * rho


def fct_face_gains(grid, c_baseline_flat, dt):
    """Frozen per-face weights w_lo, w_hi (with C, beta, tr, 1/dx folded in) of the
    tangent-linear operator, for x = c/rho.  Returns {axis: (w_lo, w_hi)}."""
    nz, ny, nx = grid.nz, grid.ny, grid.nx
    N = nz * ny * nx
    rho = state_weights(grid, N)
    x = (np.asarray(c_baseline_flat, dtype=np.float64) / rho).reshape(nz, ny, nx)
    faces = face_fluxes(grid, x)
    C = zalesak_2d(x, faces, dt)
    out = {}
    for name, f in faces.items():
        w = np.where(f["valid"], C[name] * f["beta"] * f["inv"], 0.0)
        out[name] = (f["tr_lo"] * w, f["tr_hi"] * w, f["axis"], f["n"])
    return out, rho


def assemble_fct_linear_operator(grid, c_baseline_flat, dt):
    """Frozen-coefficient tangent-linear L_AD with L_AD c_b = -d_AD(c_b) (fct.py:495-572).

    Row hi: (hi,hi) = -w_hi, (hi,lo) = +w_hi; row lo: (lo,hi) = +w_lo, (lo,lo) = -w_lo;
    then diag(rho) L diag(1/rho) in the concentration form.
    """
    nz, ny, nx = grid.nz, grid.ny, grid.nx
    N = nz * ny * nx
    gains, rho = fct_face_gains(grid, c_baseline_flat, dt)
    idx = np.arange(N).reshape(nz, ny, nx)
    r, c, v = [], [], []
    for name, (w_lo, w_hi, axis, n) in gains.items():
        hs, ls = _shift(idx, axis, n, True), _shift(idx, axis, n, False)
        i_hi, i_lo = idx[hs].ravel(), idx[ls].ravel()
        for rr, cc, vv in ((i_hi, i_hi, -w_hi), (i_hi, i_lo, w_hi),
                           (i_lo, i_hi, w_lo), (i_lo, i_lo, -w_lo)):
            r.append(rr); c.append(cc); v.append(np.asarray(vv).ravel())
    if not r:
        return sp.csc_matrix((N, N))
    L = sp.csc_matrix((np.concatenate(v), (np.concatenate(r), np.concatenate(c))), shape=(N, N))
    if getattr(grid, "concentration_state", True):
        L = (sp.diags(rho) @ L @ sp.diags(1.0 / rho)).tocsc()
    return L
