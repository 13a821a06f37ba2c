"""Seeded synthetic periodic-orbit workloads (grid metrics, per-bin operators, emissions).

The reference assembles its per-bin transport operators from GEOS/MERRA-2 NetCDF
grids (orbit/core/operator.py:281-595, convdiff.py, advection.py, mixing.py,
convection.py, deposition.py).  No such data exists here, so this module draws
seeded grids with the same *conventions*, which is everything the solve path
depends on:

* k-major flat index ``n = k*ny*nx + j*nx + i`` (orbit/core/indexing.py:23-25);
* transport written for the mixing ratio ``x = c/rho`` in the pressure measure and
  moved to the concentration form ``diag(rho) T diag(1/rho)``
  (orbit/core/operator.py:341-353) with ``rho = dP/(g*Dz)``
  (orbit/core/grid_data.py:543-553);
* horizontal first-order upwind face fluxes weighted by ``tr = dP_face/dP_cell``
  (x faces) and ``(dx_face/dx_cell)(dP_face/dP_cell)`` (y faces), which is the
  low-order flux the FCT correction subtracts (orbit/core/fct.py:256-357);
* an interface vertical mass flux built from the horizontal air-mass convergence,
  so every interior row of the pressure-form block sums to zero and only the top
  layer and the lateral edges see clean-air dilution
  (tests/test_interface_omega_transport.py:150-158);
* per-tracer first-order loss and surface dry deposition on the diagonal only:
  ``L_s = T + D_s`` (orbit/core/operator.py:572-585);
* the optional dense column block is a mass-conserving updraft/downdraft
  exchange whose rows and dP-weighted columns both sum to zero, like the
  two-branch CMFMC block (orbit/core/convection.py:25-60).

Every operator produced here is a nonsingular M-matrix after ``I + dtau*L``.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np
import scipy.sparse as sp

GRAVITY = 9.80665
EARTH_RADIUS = 6378137.0
DEG_TO_RAD = np.pi / 180.0
DTAU = 10800.0
N_BINS = 8

# Sub-domain shapes used by the workloads (nx, ny, nz).
SIZE_CLASSES = {
    "24x20": (24, 20, 15),
    "48x40": (48, 40, 15),
    "64x70": (64, 70, 15),
}


@dataclass
class Grid:
    """Metric fields of one bin, in the reference's GridData layout.

    ``UAvg[k, j, i]`` is the velocity on the west face of cell i, so
    ``UAvg[:, :, 1:]`` are the interior x faces (orbit/core/fct.py:296-302);
    ``VAvg`` likewise on south faces.  ``U_east``/``V_north`` hold the outer
    boundary faces that the (nz, ny, nx) arrays have no slot for.
    """
    nz: int
    ny: int
    nx: int
    dx: np.ndarray          # (ny,)
    dy: float
    Dz: np.ndarray          # (nz, ny, nx)
    dP: np.ndarray          # (nz, ny, nx)
    UAvg: np.ndarray        # (nz, ny, nx)
    VAvg: np.ndarray        # (nz, ny, nx)
    U_east: np.ndarray      # (nz, ny)
    V_north: np.ndarray     # (nz, nx)
    Kzz: np.ndarray         # (nz-1, ny, nx) at interior interfaces
    has_split_fluxes: bool = False
    concentration_state: bool = True
    UAvg_plus: Optional[np.ndarray] = None
    UAvg_minus: Optional[np.ndarray] = None
    VAvg_plus: Optional[np.ndarray] = None
    VAvg_minus: Optional[np.ndarray] = None

    @property
    def N(self) -> int:
        return self.nz * self.ny * self.nx


def density_weights(g: Grid) -> np.ndarray:
    """rho = dP/(g*Dz) flat (N,), 1 where a cell is degenerate
    (orbit/core/grid_data.py:543-553)."""
    ok = (g.dP > 0) & (g.Dz > 0)
    rho = np.where(ok, g.dP / (GRAVITY * np.where(ok, g.Dz, 1.0)), 1.0)
    return rho.ravel()


@dataclass
class Workload:
    """One orbit set: 8 bins x T tracers on one sub-domain."""
    nx: int
    ny: int
    nz: int
    n_tracers: int
    grids: List[Grid]                 # 8 bins
    T: List[sp.csr_matrix]            # 8 shared concentration-form transport blocks
    D: np.ndarray                     # (8, T, N) per-tracer diagonal (loss + deposition)
    e: np.ndarray                     # (T, 8, N) emission rates per bin
    rho: np.ndarray                   # (N,)
    dtau: float = DTAU
    tag: dict = field(default_factory=dict)

    @property
    def N(self) -> int:
        return self.nz * self.ny * self.nx

    def L(self, tau: int, t: int) -> sp.csr_matrix:
        """L_{tau,t} = T_tau + diag(D_{tau,t}) as sorted CSR."""
        m = (self.T[tau] + sp.diags(self.D[tau, t])).tocsr()
        m.sum_duplicates()
        m.sort_indices()
        return m

    def A(self, tau: int, t: int) -> sp.csr_matrix:
        """A = I + dtau*L (orbit/core/orbit.py:90)."""
        m = (sp.identity(self.N, format="csr") + self.dtau * self.L(tau, t)).tocsr()
        m.sum_duplicates()
        m.sort_indices()
        return m


def _box3(a: np.ndarray, axis: int) -> np.ndarray:
    """3-cell box filter along one axis, edge cells averaged over what exists."""
    a = np.moveaxis(a, axis, -1)
    s = a.copy()
    cnt = np.ones(a.shape[-1])
    s[..., 1:] += a[..., :-1]
    s[..., :-1] += a[..., 1:]
    cnt[1:] += 1
    cnt[:-1] += 1
    return np.moveaxis(s / cnt, -1, axis)


def make_grid(nx: int, ny: int, nz: int, rng: np.random.Generator, tau: int,
              lon0: float = 60.0, lat0: float = 4.0, dlon: float = 0.625,
              dlat: float = 0.5, wind_std: float = 5.0,
              dP_base: Optional[np.ndarray] = None) -> Grid:
    if lat0 + ny * dlat >= 90.0:
        raise ValueError(f"{ny} rows of {dlat} deg from {lat0} deg N reach the pole; use a smaller dlat")
    lat = lat0 + (np.arange(ny) + 0.5) * dlat
    dy = EARTH_RADIUS * dlat * DEG_TO_RAD
    dx = EARTH_RADIUS * np.cos(lat * DEG_TO_RAD) * dlon * DEG_TO_RAD
    # Layer thickness rising geometrically from 100 m to 2 km.
    dz1 = 100.0 * (2000.0 / 100.0) ** (np.arange(nz) / max(nz - 1, 1))
    Dz = np.broadcast_to(dz1[:, None, None], (nz, ny, nx)).copy()
    if dP_base is None:
        z_mid = np.cumsum(dz1) - 0.5 * dz1
        rho_air = 1.2 * np.exp(-z_mid / 8000.0)
        dP_base = rho_air[:, None, None] * GRAVITY * Dz
    dP = dP_base.copy()
    # Winds: N(0, wind_std) on faces, 3-cell box filter along x and y.
    U = rng.normal(0.0, wind_std, (nz, ny, nx + 1))
    V = rng.normal(0.0, wind_std, (nz, ny + 1, nx))
    U = _box3(_box3(U, 2), 1)
    V = _box3(_box3(V, 2), 1)
    # Kzz: 100 m^2/s at the surface decaying to 1 m^2/s aloft; deeper mixing
    # in the daytime bins.
    zi = np.cumsum(dz1)[:-1]
    depth = 1000.0 * (1.0 + 0.8 * np.sin(np.pi * tau / N_BINS))
    kz1 = 1.0 + 99.0 * np.exp(-zi / depth)
    Kzz = np.broadcast_to(kz1[:, None, None], (nz - 1, ny, nx)).copy()
    Kzz *= np.exp(rng.normal(0.0, 0.2, (1, ny, nx)))
    return Grid(nz=nz, ny=ny, nx=nx, dx=dx, dy=dy, Dz=Dz, dP=dP,
                UAvg=U[:, :, :nx].copy(), VAvg=V[:, :ny, :].copy(),
                U_east=U[:, :, nx].copy(), V_north=V[:, ny, :].copy(), Kzz=Kzz)


def _coo_add(rows, cols, vals, r, c, v):
    rows.append(np.asarray(r).ravel())
    cols.append(np.asarray(c).ravel())
    vals.append(np.asarray(v, dtype=np.float64).ravel())


def transport_pressure_form(g: Grid, Kh: float = 1.0e3,
                            conv: Optional[np.ndarray] = None, periodic: bool = False) -> sp.csr_matrix:
    """Pressure-form transport block T (dx/dt = -T x) for one bin.

    conv: optional (ny, nx, nz, nz) convective exchange rates F[j,i,q,p] >= 0
    (dP/s moving from layer p to layer q, balanced per layer).
    periodic: the x direction wraps (a global longitude band): the face west of
    cell 0 joins cell nx-1 to cell 0 like an interior face, with UAvg[:, :, 0]
    as its wind, in place of the two open x boundaries.
    """
    nz, ny, nx = g.nz, g.ny, g.nx
    N = nz * ny * nx
    n3 = np.arange(N).reshape(nz, ny, nx)
    rows, cols, vals = [], [], []
    dP = g.dP
    G = np.zeros((nz, ny, nx))            # horizontal air-mass convergence (dP/s)

    # ---- x faces: interior (between i-1 and i) ----
    u = g.UAvg[:, :, 1:]
    up, um = np.maximum(u, 0.0), np.maximum(-u, 0.0)
    dxf = np.broadcast_to(g.dx[None, :, None], (nz, ny, nx - 1))
    dPL, dPR = dP[:, :, :-1], dP[:, :, 1:]
    dPf = 0.5 * (dPL + dPR)
    trL, trR = dPf / dPL, dPf / dPR
    kh = Kh / dxf                          # diffusive face conductance (m/s)
    nL, nR = n3[:, :, :-1], n3[:, :, 1:]
    # flux F = (up+kh) x_L - (um+kh) x_R per unit dP_face, over dx
    aL = (up + kh) / dxf
    aR = (um + kh) / dxf
    _coo_add(rows, cols, vals, nL, nL, trL * aL)
    _coo_add(rows, cols, vals, nL, nR, -trL * aR)
    _coo_add(rows, cols, vals, nR, nR, trR * aR)
    _coo_add(rows, cols, vals, nR, nL, -trR * aL)
    air = dPf * u / dxf                    # air mass flux L->R (dP/s)
    G[:, :, 1:] += air
    G[:, :, :-1] -= air
    dxw = np.broadcast_to(g.dx[None, :], (nz, ny))
    if periodic:
        if nx < 3:
            raise ValueError("a periodic x direction needs nx >= 3")
        # wrap face: L = cell nx-1, R = cell 0, wind of the face west of cell 0
        u = g.UAvg[:, :, 0]
        up, um = np.maximum(u, 0.0), np.maximum(-u, 0.0)
        dPL, dPR = dP[:, :, -1], dP[:, :, 0]
        dPf = 0.5 * (dPL + dPR)
        trL, trR = dPf / dPL, dPf / dPR
        kh = Kh / dxw
        aL = (up + kh) / dxw
        aR = (um + kh) / dxw
        nL, nR = n3[:, :, -1], n3[:, :, 0]
        _coo_add(rows, cols, vals, nL, nL, trL * aL)
        _coo_add(rows, cols, vals, nL, nR, -trL * aR)
        _coo_add(rows, cols, vals, nR, nR, trR * aR)
        _coo_add(rows, cols, vals, nR, nL, -trR * aL)
        air = dPf * u / dxw
        G[:, :, 0] += air
        G[:, :, -1] -= air
    else:
        # boundary x faces: outflow only carries tracer, inflow is clean air
        uw = g.UAvg[:, :, 0]
        ue = g.U_east
        _coo_add(rows, cols, vals, n3[:, :, 0], n3[:, :, 0], np.maximum(-uw, 0.0) / dxw)
        _coo_add(rows, cols, vals, n3[:, :, -1], n3[:, :, -1], np.maximum(ue, 0.0) / dxw)
        G[:, :, 0] += dP[:, :, 0] * uw / dxw
        G[:, :, -1] -= dP[:, :, -1] * ue / dxw

    # ---- y faces: interior (between j-1 = S and j = N) ----
    v = g.VAvg[:, 1:, :]
    vp, vm = np.maximum(v, 0.0), np.maximum(-v, 0.0)
    dx3 = np.broadcast_to(g.dx[None, :, None], (nz, ny, nx))
    dxS, dxN = dx3[:, :-1, :], dx3[:, 1:, :]
    dxf = 0.5 * (dxS + dxN)
    dPS, dPN = dP[:, :-1, :], dP[:, 1:, :]
    dPf = 0.5 * (dPS + dPN)
    trS = (dxf / dxS) * (dPf / dPS)
    trN = (dxf / dxN) * (dPf / dPN)
    kh = Kh / g.dy
    aS = (vp + kh) / g.dy
    aN = (vm + kh) / g.dy
    nS, nN = n3[:, :-1, :], n3[:, 1:, :]
    _coo_add(rows, cols, vals, nS, nS, trS * aS)
    _coo_add(rows, cols, vals, nS, nN, -trS * aN)
    _coo_add(rows, cols, vals, nN, nN, trN * aN)
    _coo_add(rows, cols, vals, nN, nS, -trN * aS)
    airy = dPf * dxf * v / g.dy            # air mass flux S->N per unit area of N-cell row
    G[:, 1:, :] += airy / dxN
    G[:, :-1, :] -= airy / dxS
    vs = g.VAvg[:, 0, :]
    vn = g.V_north
    _coo_add(rows, cols, vals, n3[:, 0, :], n3[:, 0, :], np.maximum(-vs, 0.0) / g.dy)
    _coo_add(rows, cols, vals, n3[:, -1, :], n3[:, -1, :], np.maximum(vn, 0.0) / g.dy)
    G[:, 0, :] += dP[:, 0, :] * vs / g.dy
    G[:, -1, :] -= dP[:, -1, :] * vn / g.dy

    # ---- vertical advection from the interface mass flux W (upward, dP/s) ----
    # W_{k+1/2} = W_{k-1/2} + G_k, W_{-1/2} = 0: every layer's air budget closes.
    W = np.cumsum(G, axis=0)               # W[k] = flux through the top of layer k
    for k in range(nz - 1):
        w = W[k]
        wp, wm = np.maximum(w, 0.0), np.maximum(-w, 0.0)
        a, b = n3[k], n3[k + 1]
        _coo_add(rows, cols, vals, a, a, wp / dP[k])
        _coo_add(rows, cols, vals, a, b, -wm / dP[k])
        _coo_add(rows, cols, vals, b, b, wm / dP[k + 1])
        _coo_add(rows, cols, vals, b, a, -wp / dP[k + 1])
    # top interface: upward flux exports tracer, downward brings clean air
    _coo_add(rows, cols, vals, n3[nz - 1], n3[nz - 1], np.maximum(W[nz - 1], 0.0) / dP[nz - 1])

    # ---- vertical diffusion ----
    rho = dP / (GRAVITY * g.Dz)
    for k in range(nz - 1):
        dzf = 0.5 * (g.Dz[k] + g.Dz[k + 1])
        rhof = 0.5 * (rho[k] + rho[k + 1])
        kap = GRAVITY * rhof * g.Kzz[k] / dzf        # dP/s
        a, b = n3[k], n3[k + 1]
        _coo_add(rows, cols, vals, a, a, kap / dP[k])
        _coo_add(rows, cols, vals, a, b, -kap / dP[k])
        _coo_add(rows, cols, vals, b, b, kap / dP[k + 1])
        _coo_add(rows, cols, vals, b, a, -kap / dP[k + 1])

    # ---- dense column convection block ----
    if conv is not None:
        out = conv.sum(axis=2)                         # (ny, nx, nz): total leaving layer p
        for q in range(nz):
            for p in range(nz):
                if p == q:
                    continue
                f = conv[:, :, q, p]
                _coo_add(rows, cols, vals, n3[q], n3[p], -f / dP[q])
        for p in range(nz):
            _coo_add(rows, cols, vals, n3[p], n3[p], out[:, :, p] / dP[p])

    T = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(N, N)).tocsr()
    T.sum_duplicates()
    T.sort_indices()
    return T


def convection_exchange(ny: int, nx: int, nz: int, dP: np.ndarray,
                        rng: np.random.Generator, strength: float = 1.0) -> np.ndarray:
    """Random dense, balanced updraft/downdraft exchange F[j,i,q,p] (dP/s).

    Each ordered pair (p -> q) moves mass directly and returns it through the
    adjacent-layer chain between q and p, so each layer's inflow equals its
    outflow: rows of the pressure-form block sum to zero and dP-weighted
    columns sum to zero (mass conserved).  Every off-diagonal entry is
    strictly negative: a non-local 15x15 block per column.
    """
    intensity = np.exp(rng.normal(0.0, 0.7, (ny, nx))) * strength
    # direct transfer rates scale with the thinner of the two layers' masses
    dPc = np.moveaxis(dP, 0, -1)                       # (ny, nx, nz)
    F = np.zeros((ny, nx, nz, nz))
    base = 0.15 / 10800.0                              # fraction of a layer per bin
    r = rng.uniform(0.05, 1.0, (ny, nx, nz, nz)) / nz
    for p in range(nz):
        for q in range(nz):
            if p == q:
                continue
            f = base * intensity * r[:, :, q, p] * np.minimum(dPc[:, :, p], dPc[:, :, q])
            F[:, :, q, p] += f
            # return chain q -> p through adjacent layers
            step = 1 if p > q else -1
            k = q
            while k != p:
                F[:, :, k + step, k] += f
                k += step
    return F


def to_concentration_form(T: sp.csr_matrix, rho: np.ndarray) -> sp.csr_matrix:
    """diag(rho) T diag(1/rho) (orbit/core/operator.py:341-353)."""
    Tc = (sp.diags(rho) @ T @ sp.diags(1.0 / rho)).tocsr()
    Tc.sum_duplicates()
    Tc.sort_indices()
    return Tc


def make_workload(nx: int, ny: int, nz: int = 15, n_tracers: int = 1, seed=0,
                  convection: bool = False, emis_frac: float = 0.3,
                  wind_std: float = 5.0, dtau: float = DTAU, dlon: float = 0.625,
                  dlat: float = 0.5, periodic: bool = False) -> Workload:
    """Draw one seeded orbit set (8 bins, ``n_tracers`` tracers sharing T).

    ``seed`` may be an int or a sequence (passed to ``np.random.default_rng``).
    dlon, dlat: cell size in degrees (rows start at 4 deg N).  periodic: the
    x direction wraps (transport_pressure_form).
    """
    rng = np.random.default_rng(seed)
    N = nz * ny * nx
    grids, Ts = [], []
    dP_base = None
    for tau in range(N_BINS):
        g = make_grid(nx, ny, nz, rng, tau, wind_std=wind_std, dP_base=dP_base, dlon=dlon, dlat=dlat)
        if dP_base is None:
            # terrain-like +-5% smooth variation of the column air mass
            ps = _box3(_box3(rng.normal(0.0, 1.0, (ny, nx)), 0), 1)
            ps = 1.0 + 0.05 * ps / max(np.abs(ps).max(), 1e-12)
            g.dP = g.dP * ps[None]
            dP_base = g.dP.copy()
        conv = convection_exchange(ny, nx, nz, g.dP, rng) if convection else None
        if periodic:  # the face east of the last cell is the face west of the first
            g.U_east = g.UAvg[:, :, 0].copy()
        Tp = transport_pressure_form(g, conv=conv, periodic=periodic)
        grids.append(g)
        Ts.append(Tp)
    rho = density_weights(grids[0])
    Tc = [to_concentration_form(T, rho) for T in Ts]

    D = np.zeros((N_BINS, n_tracers, N))
    e = np.zeros((n_tracers, N_BINS, N))
    diurnal = 1.0 + 0.5 * np.sin(2.0 * np.pi * np.arange(N_BINS) / N_BINS)
    emis_diurnal = 1.0 + 0.4 * np.cos(2.0 * np.pi * (np.arange(N_BINS) - 4) / N_BINS)
    dz0 = grids[0].Dz[0].ravel()
    ncol = ny * nx
    for t in range(n_tracers):
        k = np.exp(rng.uniform(np.log(1e-6), np.log(1e-4), N))
        vd = rng.uniform(0.001, 0.01, ncol)
        src_cols = rng.random(ncol) < emis_frac
        mag = rng.lognormal(0.0, 1.0, ncol) * src_cols * 1.0e-5
        for tau in range(N_BINS):
            D[tau, t] = k * diurnal[tau]
            D[tau, t, :ncol] += vd / dz0
            e[t, tau, :ncol] = mag * emis_diurnal[tau]
    return Workload(nx=nx, ny=ny, nz=nz, n_tracers=n_tracers, grids=grids, T=Tc,
                    D=D, e=e, rho=rho, dtau=dtau,
                    tag=dict(seed=seed, convection=convection, periodic=periodic))


# Seed tag of each benchmark workload kind (the first seed component).
WORKLOAD_KINDS = {"single": 0, "fct-month": 1, "uneven-batch": 2, "adjoint": 3, "serving": 4}


def benchmark_workload(kind: str, n_tracers: Optional[int] = None, month: int = 0,
                       size: str = "64x70", convection: Optional[bool] = None) -> Workload:
    """Seeded workloads of the benchmark suite, by kind.

    "single": 24x20x15, 1 tracer, no convection.
    "fct-month" / "adjoint": 64x70x15, 12 tracers, dense convection blocks.
    "uneven-batch" / "serving": one (month, size) orbit set of the batched /
    served mixes, 12 tracers by default.
    Seeds are ``[kind tag, month, size_class]``.
    """
    if kind not in WORKLOAD_KINDS:
        raise ValueError(f"unknown workload kind {kind!r}; one of {sorted(WORKLOAD_KINDS)}")
    tag = WORKLOAD_KINDS[kind]
    if kind == "single":
        nx, ny, nz = SIZE_CLASSES["24x20"]
        nt = 1 if n_tracers is None else n_tracers
        conv = False if convection is None else convection
    else:
        nx, ny, nz = SIZE_CLASSES[size]
        nt = 12 if n_tracers is None else n_tracers
        conv = True if convection is None else convection
    size_id = list(SIZE_CLASSES).index(f"{nx}x{ny}")
    return make_workload(nx, ny, nz, nt, seed=[tag, month, size_id], convection=conv)


# --------------------------------------------------------------------------- #
# GridData-like fields for the device-side transport assembly
# --------------------------------------------------------------------------- #
def _mass_divergence(g, U, V, Ue, Vn, Uw=None) -> np.ndarray:
    """Horizontal air-mass divergence per cell (Pa/s) of the Patankar block's
    advective part: dP_n (T_hcd 1)_n = sum over faces of the outgoing
    dP_face u / h, with the y faces weighted by dx_face/dx_cell (the terrain
    weights of orbit/core/convdiff.py:119-130, 285-298).  Uw: the wind of a
    periodic wrap face (i = nx-1 -> 0), in place of the west / east outer faces."""
    nz, ny, nx = g.nz, g.ny, g.nx
    dP = g.dP
    dx3 = np.broadcast_to(g.dx[None, :, None], (nz, ny, nx))
    Dv = np.zeros((nz, ny, nx))
    fx = 0.5 * (dP[:, :, :-1] + dP[:, :, 1:]) * U[:, :, 1:] / dx3[:, :, 1:]    # interior x faces
    Dv[:, :, :-1] += fx
    Dv[:, :, 1:] -= fx
    if Uw is None:
        Dv[:, :, 0] -= dP[:, :, 0] * U[:, :, 0] / dx3[:, :, 0]                 # outer faces
        Dv[:, :, -1] += dP[:, :, -1] * Ue / dx3[:, :, -1]
    else:
        fw = 0.5 * (dP[:, :, -1] + dP[:, :, 0]) * Uw / dx3[:, :, 0]             # wrap face
        Dv[:, :, -1] += fw
        Dv[:, :, 0] -= fw
    dxf = 0.5 * (dx3[:, :-1, :] + dx3[:, 1:, :])
    fy = dxf * 0.5 * (dP[:, :-1, :] + dP[:, 1:, :]) * V[:, 1:, :] / g.dy
    Dv[:, :-1, :] += fy / dx3[:, :-1, :]
    Dv[:, 1:, :] -= fy / dx3[:, 1:, :]
    # outer south / north faces: the zonal width of the face over the edge row's
    # (the ratio of the reference's open-boundary rates, convdiff.py:335-365)
    zonal = lambda la: EARTH_RADIUS * np.cos(la * DEG_TO_RAD) * g.dlon * DEG_TO_RAD  # noqa: E731
    tr_s = zonal(g.lat[0] - g.dlat / 2.0) / (g.dx[0] if g.dx[0] > 0 else 1.0)
    tr_n = zonal(g.lat[-1] + g.dlat / 2.0) / (g.dx[-1] if g.dx[-1] > 0 else 1.0)
    Dv[:, 0, :] -= dP[:, 0, :] * V[:, 0, :] / g.dy * tr_s
    Dv[:, -1, :] += dP[:, -1, :] * Vn / g.dy * tr_n
    return Dv


def make_reference_grids(nx: int, ny: int, nz: int = 15, seed=0, wind_std: float = 5.0,
                         cmfmc: bool = True, interface_omega: bool = True, periodic: bool = False,
                         split: bool = False):
    """Seeded GridData-like fields of the 8 bins, with the reference's attribute
    names (orbit/core/grid_data.py:20-240), for the device-side assembly of the
    transport block: winds on west/south faces plus the true east/north outer
    faces, cell-centred Kxxyy and Kzz, face-centred meander diffusivities,
    an interface mass flux omega_edge diagnosed from the horizontal
    convergence (closed ground, so every interior row of the transport block
    sums to zero: tests/test_interface_omega_transport.py:117-170 of the
    reference), and a CMFMC updraft in a third of the columns.

    interface_omega=False gives the pre-2026-09-26 reading instead: a
    cell-centred omega that does not close the air budget (the air-mass
    pre-flight flags it).  periodic=True wraps the x direction: the face west
    of cell 0 is the wrap face (UAvg_wrap, K_meander_u_wrap), with no west /
    east outer faces (grid_data.py:427, convdiff.py:150-200).  split=True gives
    split fluxes: UAvg_plus - UAvg_minus = UAvg on every face (each side carrying
    a 0.3 m/s counter-flow, as a sub-grid exchange), the outer and wrap faces
    included, so the net air budget is that of the net winds."""
    import types
    rng = np.random.default_rng(seed)
    grids = []
    dP_base = None
    for tau in range(N_BINS):
        b = make_grid(nx, ny, nz, rng, tau, wind_std=wind_std, dP_base=dP_base)
        if dP_base is None:
            ps = _box3(_box3(rng.normal(0.0, 1.0, (ny, nx)), 0), 1)
            ps = 1.0 + 0.05 * ps / max(np.abs(ps).max(), 1e-12)
            b.dP = b.dP * ps[None]
            dP_base = b.dP.copy()
        g = types.SimpleNamespace(nz=nz, ny=ny, nx=nx, dlon=0.625, dlat=0.5)
        g.lat = 4.0 + (np.arange(ny) + 0.5) * g.dlat
        g.dx, g.dy, g.Dz, g.dP = b.dx, b.dy, b.Dz, b.dP
        g.UAvg, g.VAvg = b.UAvg, b.VAvg
        g.U_face_e, g.V_face_n = b.U_east, b.V_north
        g.has_split_fluxes = False
        g.concentration_state = True
        g.periodic_lon = periodic
        g.Kxxyy = 1.0e3 * np.exp(rng.normal(0.0, 0.3, (nz, ny, nx)))
        g.K_meander_u = 50.0 * rng.random((nz, ny, nx))
        g.K_meander_v = 50.0 * rng.random((nz, ny, nx))
        kz = np.concatenate([b.Kzz[:, :, :], b.Kzz[-1:, :, :]], axis=0)   # interfaces -> cells
        g.Kzz = kz
        if periodic:
            g.UAvg_wrap = g.UAvg[:, :, 0].copy()
            g.U_face_e = g.UAvg_wrap
            g.K_meander_u_wrap = g.K_meander_u[:, :, 0].copy()
        if split:
            cf = 0.3
            g.has_split_fluxes = True
            for name, a in (("UAvg", g.UAvg), ("VAvg", g.VAvg), ("U_face_e", g.U_face_e),
                            ("V_face_n", g.V_face_n)):
                base = {"UAvg": "UAvg", "VAvg": "VAvg", "U_face_e": "U", "V_face_n": "V"}[name]
                plus = np.maximum(a, 0.0) + cf
                minus = np.maximum(-a, 0.0) + cf
                if name in ("UAvg", "VAvg"):
                    setattr(g, base + "_plus", plus)
                    setattr(g, base + "_minus", minus)
                else:
                    setattr(g, base + "_plus" + name[1:], plus)
                    setattr(g, base + "_minus" + name[1:], minus)
            if periodic:
                g.UAvg_plus_wrap = g.UAvg_plus[:, :, 0].copy()
                g.UAvg_minus_wrap = g.UAvg_minus[:, :, 0].copy()
        div = _mass_divergence(g, g.UAvg, g.VAvg, g.U_face_e, g.V_face_n,
                               Uw=g.UAvg_wrap if periodic else None)
        om = np.zeros((nz + 1, ny, nx))
        om[1:] = np.cumsum(div, axis=0)        # net downward flux through each bottom face
        if interface_omega:
            g.omega_edge = om
            g.omega_edge_plus = np.maximum(om, 0.0)
            g.omega_edge_minus = np.maximum(-om, 0.0)
        else:
            oc = om[:-1] + rng.normal(0.0, 0.5 * np.abs(om).mean() + 1e-6, (nz, ny, nx))
            g.omega_plus = np.maximum(oc, 0.0)
            g.omega_minus = np.maximum(-oc, 0.0)
        if cmfmc:
            F = np.zeros((nz + 1, ny, nx))
            active = rng.random((ny, nx)) < 1.0 / 3.0
            prof = np.sin(np.pi * np.arange(nz + 1) / nz) ** 2        # 0 at ground and top
            F[:] = 0.02 * prof[:, None, None] * active[None] * rng.random((1, ny, nx))
            g.CMFMC = F
        grids.append(g)
    return grids


def grid_workload(nx: int, ny: int, nz: int = 15, n_tracers: int = 1, seed=0, **kw):
    """Grids of make_reference_grids plus per-tracer diagonals D (n_bins,
    n_tracers, N) and emissions e (n_tracers, n_bins, N) drawn as in
    make_workload."""
    grids = make_reference_grids(nx, ny, nz, seed=seed, **kw)
    rng = np.random.default_rng([7919, *np.atleast_1d(seed).tolist()])
    N = nz * ny * nx
    ncol = nx * ny
    D = np.zeros((N_BINS, n_tracers, N))
    e = np.zeros((n_tracers, N_BINS, N))
    diurnal = 1.0 + 0.5 * np.sin(2.0 * np.pi * np.arange(N_BINS) / N_BINS)
    emis_diurnal = 1.0 + 0.4 * np.cos(2.0 * np.pi * (np.arange(N_BINS) - 4) / N_BINS)
    dz0 = grids[0].Dz[0].ravel()
    for t in range(n_tracers):
        k = np.exp(rng.uniform(np.log(1e-6), np.log(1e-4), N))
        vd = rng.uniform(0.001, 0.01, ncol)
        mag = rng.lognormal(0.0, 1.0, ncol) * (rng.random(ncol) < 0.3) * 1.0e-5
        for tau in range(N_BINS):
            D[tau, t] = k * diurnal[tau]
            D[tau, t, :ncol] += vd / dz0
            e[t, tau, :ncol] = mag * emis_diurnal[tau]
    return grids, D, e
