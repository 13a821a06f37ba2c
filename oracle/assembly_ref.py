"""CPU oracle of the transport-block assembly -- TEST INFRASTRUCTURE ONLY.

Restates, in numpy, how the reference builds the shared transport block of one
bin from its grid fields:

    T = T_hcd + (T_vadv + T_vdiff) + T_conv,  then  diag(rho) T diag(1/rho)

(reference orbit/core/operator.py:281-353).  Every block is a per-cell stencil
on the k-major grid: a diagonal, the four same-level horizontal neighbours
(W, E, S, N) and the two vertical neighbours (down = k-1, up = k+1).  The
oracle computes those stencil arrays face by face and builds scipy CSC from
them; the device assembly returns the same stencil arrays.

Grid objects carry the reference's GridData attribute names (grid_data.py:20-240);
an empty or missing optional field means "absent" as in the reference.

Pinned by: the reference's golden vectors (tests/golden/assembly_golden.npz,
written by tests/golden/make_assembly_golden.py from the reference's own
functions), the Patankar function's closed forms, the exact 1-D exponential
profile, zero-wind / zero-diffusion limits, M-matrix signs, face-by-face
conservation of the dP-area measure, zero row sums of the vertical blocks and
of the concentration form on rho (tests/test_oracle_assembly.py).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

GRAVITY = 9.80665          # grid_data.py:15, mixing.py:11, convection.py:22
EARTH_RADIUS = 6378137.0   # grid_data.py:14
DEG_TO_RAD = np.pi / 180.0  # grid_data.py:16

STENCIL = ("diag", "W", "E", "S", "N", "down", "up")


def _has(g, name) -> bool:
    a = getattr(g, name, None)
    return a is not None and np.size(a) > 0


def patankar_A(pe_abs):
    """A(|Pe|) = |Pe| / (exp|Pe| - 1), 1 - |Pe|/2 below 1e-6, 0 above 500
    (reference orbit/core/convdiff.py:15-39)."""
    pe = np.asarray(pe_abs, dtype=np.float64)
    out = np.empty_like(pe)
    small = pe < 1e-6
    large = pe > 500.0
    mid = ~small & ~large
    out[small] = 1.0 - 0.5 * pe[small]
    out[large] = 0.0
    out[mid] = pe[mid] / np.expm1(pe[mid])
    return out


def edge_face(g, which):
    """Wind on the east / north outer face: the stored true face, else the edge
    cell's west / south face (reference orbit/core/grid_data.py:270-285)."""
    true = {"U": "U_face_e", "U_plus": "U_plus_face_e", "U_minus": "U_minus_face_e",
            "V": "V_face_n", "V_plus": "V_plus_face_n", "V_minus": "V_minus_face_n"}[which]
    if _has(g, true):
        return np.asarray(getattr(g, true), dtype=np.float64)
    proxy = {"U": "UAvg", "U_plus": "UAvg_plus", "U_minus": "UAvg_minus",
             "V": "VAvg", "V_plus": "VAvg_plus", "V_minus": "VAvg_minus"}[which]
    a = np.asarray(getattr(g, proxy), dtype=np.float64)
    return a[:, :, -1] if which.startswith("U") else a[:, -1, :]


def _face_rates(u, K_face, h, split_plus=None, split_minus=None):
    """Patankar rates of a face of width h: (rate_L, rate_R), the loss rates of
    the cell on the low and on the high side (reference convdiff.py:89-119)."""
    safe_h = np.where(h > 0, h, 1.0)
    D = K_face / safe_h
    safe_K = np.where(K_face > 1e-30, K_face, 1e-30)
    pe = np.minimum(np.abs(u * safe_h / safe_K), 500.0)
    diff = D * patankar_A(pe) / safe_h
    if split_plus is not None:
        return diff + split_plus / safe_h, diff + split_minus / safe_h
    return diff + np.maximum(u, 0.0) / safe_h, diff + np.maximum(-u, 0.0) / safe_h


def horizontal_convdiff_stencil(g) -> dict:
    """Patankar exponential horizontal convection-diffusion in the pressure
    measure, as a stencil (reference orbit/core/convdiff.py:42-387).

    W/E at i = 0 / nx-1 hold the wrap couplings when g.periodic_lon."""
    nz, ny, nx = g.nz, g.ny, g.nx
    U, V, dP = (np.asarray(getattr(g, a), dtype=np.float64) for a in ("UAvg", "VAvg", "dP"))
    Kh = np.asarray(g.Kxxyy, dtype=np.float64)
    split = bool(getattr(g, "has_split_fluxes", False))
    periodic = bool(getattr(g, "periodic_lon", False))
    dx = np.asarray(g.dx, dtype=np.float64)
    dx3 = np.broadcast_to(dx[None, :, None], (nz, ny, nx))
    valid = dx3 > 0
    st = {k: np.zeros((nz, ny, nx)) for k in ("diag", "W", "E", "S", "N")}

    def harm(a, b):
        return 2.0 * a * b / (a + b + 1e-30)

    # x faces between i-1 and i, i = 1..nx-1
    if nx > 1:
        Kf = harm(Kh[:, :, :-1], Kh[:, :, 1:])
        if not split:
            Kf = Kf + np.asarray(g.K_meander_u, dtype=np.float64)[:, :, 1:]
        sp_ = (g.UAvg_plus[:, :, 1:], g.UAvg_minus[:, :, 1:]) if split else (None, None)
        rL, rR = _face_rates(U[:, :, 1:], Kf, dx3[:, :, 1:], *sp_)
        dL, dR = dP[:, :, :-1], dP[:, :, 1:]
        ok = (dL > 0) & (dR > 0)
        face = 0.5 * (dL + dR)
        trR = np.where(ok, face / np.where(ok, dR, 1.0), 1.0)
        trL = np.where(ok, face / np.where(ok, dL, 1.0), 1.0)
        v = valid[:, :, 1:] & valid[:, :, :-1]
        st["diag"][:, :, 1:] += np.where(v, rR * trR, 0.0)
        st["W"][:, :, 1:] = np.where(v, -rL * trR, 0.0)
        st["diag"][:, :, :-1] += np.where(v, rL * trL, 0.0)
        st["E"][:, :, :-1] = np.where(v, -rR * trL, 0.0)
    # wrap face between nx-1 (low side) and 0 (high side)
    if periodic and nx > 1:
        Kf = harm(Kh[:, :, -1], Kh[:, :, 0])
        if not split:
            Kf = Kf + np.asarray(g.K_meander_u_wrap, dtype=np.float64)
        sp_ = (g.UAvg_plus_wrap, g.UAvg_minus_wrap) if split else (None, None)
        rL, rR = _face_rates(np.asarray(g.UAvg_wrap, dtype=np.float64), Kf, dx3[:, :, 0], *sp_)
        dL, dR = dP[:, :, -1], dP[:, :, 0]
        ok = (dL > 0) & (dR > 0)
        face = 0.5 * (dL + dR)
        trR = np.where(ok, face / np.where(ok, dR, 1.0), 1.0)
        trL = np.where(ok, face / np.where(ok, dL, 1.0), 1.0)
        v = dx3[:, :, 0] > 0
        st["diag"][:, :, 0] += np.where(v, rR * trR, 0.0)
        st["W"][:, :, 0] = np.where(v, -rL * trR, 0.0)
        st["diag"][:, :, -1] += np.where(v, rL * trL, 0.0)
        st["E"][:, :, -1] = np.where(v, -rR * trL, 0.0)
    # west / east outflow (regional grids)
    if not periodic:
        sdx = np.where(dx3 > 0, dx3, 1.0)
        w0 = np.zeros((nz, ny, nx))
        if split:
            um = g.UAvg_minus[:, :, 0]
            w0[:, :, 0] = np.where(valid[:, :, 0] & (um > 0), um / sdx[:, :, 0], 0.0)
        else:
            w0[:, :, 0] = np.where(valid[:, :, 0] & (U[:, :, 0] < 0), np.abs(U[:, :, 0]) / sdx[:, :, 0], 0.0)
        st["diag"] += w0
        e0 = np.zeros((nz, ny, nx))
        if split:
            up = edge_face(g, "U_plus")
            e0[:, :, -1] = np.where(valid[:, :, -1] & (up > 0), up / sdx[:, :, -1], 0.0)
        else:
            ue = edge_face(g, "U") if nx > 1 else U[:, :, 0]
            e0[:, :, -1] = np.where(valid[:, :, -1] & (ue >= 0), ue / sdx[:, :, -1], 0.0)
        st["diag"] += e0
    # y faces between j-1 and j, j = 1..ny-1
    dy = float(g.dy)
    sdy = dy if dy > 0 else 1.0
    if ny > 1:
        Kf = harm(Kh[:, :-1, :], Kh[:, 1:, :])
        if not split:
            Kf = Kf + np.asarray(g.K_meander_v, dtype=np.float64)[:, 1:, :]
        sp_ = (g.VAvg_plus[:, 1:, :], g.VAvg_minus[:, 1:, :]) if split else (None, None)
        rL, rR = _face_rates(V[:, 1:, :], Kf, np.full(Kf.shape, sdy), *sp_)
        dxS, dxN = dx3[:, :-1, :], dx3[:, 1:, :]
        dxf = 0.5 * (dxS + dxN)
        dS, dN = dP[:, :-1, :], dP[:, 1:, :]
        ok = (dS > 0) & (dN > 0)
        face = 0.5 * (dS + dN)
        trR = np.where(ok, (dxf / np.where(dxN > 0, dxN, 1.0)) * (face / np.where(ok, dN, 1.0)), 1.0)
        trL = np.where(ok, (dxf / np.where(dxS > 0, dxS, 1.0)) * (face / np.where(ok, dS, 1.0)), 1.0)
        v = valid[:, 1:, :] & valid[:, :-1, :]
        st["diag"][:, 1:, :] += np.where(v, rR * trR, 0.0)
        st["S"][:, 1:, :] = np.where(v, -rL * trR, 0.0)
        st["diag"][:, :-1, :] += np.where(v, rL * trL, 0.0)
        st["N"][:, :-1, :] = np.where(v, -rR * trL, 0.0)
    # south / north outflow, with the outer face's zonal width
    lat = np.asarray(g.lat, dtype=np.float64)
    trS = EARTH_RADIUS * np.cos((lat[0] - g.dlat / 2.0) * DEG_TO_RAD) * g.dlon * DEG_TO_RAD / (dx[0] if dx[0] > 0 else 1.0)
    trN = EARTH_RADIUS * np.cos((lat[-1] + g.dlat / 2.0) * DEG_TO_RAD) * g.dlon * DEG_TO_RAD / (dx[-1] if dx[-1] > 0 else 1.0)
    s0 = np.zeros((nz, ny, nx))
    if split:
        vm = g.VAvg_minus[:, 0, :]
        s0[:, 0, :] = np.where(valid[:, 0, :] & (vm > 0), vm / sdy * trS, 0.0)
    else:
        s0[:, 0, :] = np.where(valid[:, 0, :] & (V[:, 0, :] < 0), np.abs(V[:, 0, :]) / sdy * trS, 0.0)
    st["diag"] += s0
    n0 = np.zeros((nz, ny, nx))
    if split:
        vp = edge_face(g, "V_plus")
        n0[:, -1, :] = np.where(valid[:, -1, :] & (vp > 0), vp / sdy * trN, 0.0)
    else:
        vn = edge_face(g, "V") if ny > 1 else V[:, 0, :]
        n0[:, -1, :] = np.where(valid[:, -1, :] & (vn >= 0), vn / sdy * trN, 0.0)
    st["diag"] += n0
    return st


def interface_omega(g):
    """(w_plus, w_minus), (nz+1, ny, nx): downward / upward mass flux through
    the bottom face of each layer; ground closed; from omega_edge_* when the
    grid has it, else the cell-centred omega_* with a closed top
    (reference orbit/core/advection.py:383-412)."""
    nz, ny, nx = g.nz, g.ny, g.nx
    if _has(g, "omega_edge_plus") and np.shape(g.omega_edge_plus)[0] == nz + 1:
        wp = np.array(g.omega_edge_plus, dtype=np.float64)
        wm = np.array(g.omega_edge_minus, dtype=np.float64)
    else:
        wp = np.zeros((nz + 1, ny, nx))
        wm = np.zeros((nz + 1, ny, nx))
        if _has(g, "omega_plus"):
            wp[1:nz] = np.asarray(g.omega_plus, dtype=np.float64)[1:]
            wm[1:nz] = np.asarray(g.omega_minus, dtype=np.float64)[1:]
    wp[0] = 0.0
    wm[0] = 0.0
    return wp, wm


def vertical_advection_stencil(g) -> dict:
    """First-order upwind omega advection on the dP measure
    (reference orbit/core/advection.py:414-525)."""
    nz = g.nz
    dP = np.asarray(g.dP, dtype=np.float64)
    wp, wm = interface_omega(g)
    ok = dP > 0
    sdP = np.where(ok, dP, 1.0)
    below = np.zeros(dP.shape, bool)
    below[1:] = True
    above = np.zeros(dP.shape, bool)
    above[:-1] = True
    diag = np.where(ok & below & (wp[:nz] > 0), wp[:nz] / sdP, 0.0)
    down = np.where(ok & below & (wm[:nz] > 0), -wm[:nz] / sdP, 0.0)
    up = np.where(ok & above & (wp[1:] > 0), -wp[1:] / sdP, 0.0)
    diag = diag + np.where(ok & (wm[1:] > 0), wm[1:] / sdP, 0.0)
    return {"diag": diag, "down": down, "up": up}


def vertical_diffusion_stencil(g) -> dict:
    """Kzz diffusion with the flux divergence on the dP measure
    (reference orbit/core/mixing.py:14-117)."""
    nz = g.nz
    Kzz, Dz, dP = (np.asarray(getattr(g, a), dtype=np.float64) for a in ("Kzz", "Dz", "dP"))
    rho = np.where((Dz > 0) & (dP > 0), dP / (GRAVITY * np.where(Dz > 0, Dz, 1.0)), 0.0)
    diag = np.zeros(Dz.shape)
    down = np.zeros(Dz.shape)
    up = np.zeros(Dz.shape)
    for k in range(nz - 1):
        a, b = Dz[k], Dz[k + 1]
        v = (a > 0) & (b > 0)
        Kf = np.where(v, 2.0 * Kzz[k] * Kzz[k + 1] / (Kzz[k] + Kzz[k + 1] + 1e-30), 0.0)
        dist = np.where(v, 0.5 * (a + b), 1.0)
        rf = 0.5 * (rho[k] + rho[k + 1])
        both = v & (rho[k] > 0) & (rho[k + 1] > 0)
        ra = np.where(both, rf / np.where(rho[k] > 0, rho[k], 1.0), 1.0)
        rb = np.where(both, rf / np.where(rho[k + 1] > 0, rho[k + 1], 1.0), 1.0)
        ca = np.where(v, Kf / (np.where(a > 0, a, 1.0) * dist) * ra, 0.0)
        cb = np.where(v, Kf / (np.where(b > 0, b, 1.0) * dist) * rb, 0.0)
        up[k] = -ca
        down[k + 1] = -cb
        diag[k] += ca
        diag[k + 1] += cb
    return {"diag": diag, "down": down, "up": up}


def cmfmc_stencil(g) -> dict:
    """Convective updraft with its compensating subsidence, on the dP measure
    (reference orbit/core/convection.py:25-140).  Zero without CMFMC."""
    nz = g.nz
    Dz, dP = (np.asarray(getattr(g, a), dtype=np.float64) for a in ("Dz", "dP"))
    diag = np.zeros(Dz.shape)
    down = np.zeros(Dz.shape)
    up = np.zeros(Dz.shape)
    if not _has(g, "CMFMC"):
        return {"diag": diag, "down": down, "up": up}
    F_all = np.asarray(g.CMFMC, dtype=np.float64)
    sDz = np.where(Dz > 0, Dz, 1.0)
    rho = dP / (GRAVITY * sDz)
    srho = np.where(rho > 0, rho, 1.0)
    ok = (Dz > 0) & (dP > 0)
    for k in range(nz):
        F = np.maximum(F_all[k + 1], 0.0)
        act = ok[k] & (F > 0)
        src = np.where(act, F / (srho[k] * sDz[k]), 0.0)
        diag[k] += src
        if k < nz - 1:
            gain = act & ok[k + 1]
            rcv = np.where(gain, F / (srho[k + 1] * np.where(Dz[k + 1] > 0, Dz[k + 1], 1.0)), 0.0)
            down[k + 1] = np.where(gain, -rcv, 0.0)
            diag[k + 1] += rcv
            up[k] = np.where(gain, -src, 0.0)
    return {"diag": diag, "down": down, "up": up}


def density_weights(g):
    """rho = dP / (g Dz), 1 where degenerate, flat (N,)
    (reference orbit/core/grid_data.py:543-553)."""
    dP = np.asarray(g.dP, dtype=np.float64)
    Dz = np.asarray(g.Dz, dtype=np.float64)
    ok = (dP > 0) & (Dz > 0)
    return np.where(ok, dP / (GRAVITY * np.where(ok, Dz, 1.0)), 1.0).ravel()


def stencil_to_csc(st: dict, nz, ny, nx, periodic=False) -> sp.csc_matrix:
    N = nz * ny * nx
    n3 = np.arange(N).reshape(nz, ny, nx)
    rows, cols, vals = [n3.ravel()], [n3.ravel()], [st["diag"].ravel()]
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    nb = {
        "W": (k, j, (i - 1) % nx, (i > 0) | periodic),
        "E": (k, j, (i + 1) % nx, (i < nx - 1) | periodic),
        "S": (k, np.maximum(j - 1, 0), i, j > 0),
        "N": (k, np.minimum(j + 1, ny - 1), i, j < ny - 1),
        "down": (np.maximum(k - 1, 0), j, i, k > 0),
        "up": (np.minimum(k + 1, nz - 1), j, i, k < nz - 1),
    }
    for name, (kk, jj, ii, m) in nb.items():
        a = st.get(name)
        if a is None:
            continue
        m = m & (a != 0)
        rows.append(n3[m])
        cols.append(n3[kk[m], jj[m], ii[m]])
        vals.append(a[m])
    A = sp.csc_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(N, N))
    A.eliminate_zeros()
    return A


def transport_stencil(g, concentration_state=None) -> dict:
    """Stencil of the whole transport block T = T_hcd + (T_vadv + T_vdiff) + T_conv,
    moved to the concentration form diag(rho) T diag(1/rho) when the grid's
    state is a concentration (reference orbit/core/operator.py:281-353)."""
    h = horizontal_convdiff_stencil(g)
    va = vertical_advection_stencil(g)
    vd = vertical_diffusion_stencil(g)
    cv = cmfmc_stencil(g)
    st = {
        "diag": (h["diag"] + (va["diag"] + vd["diag"])) + cv["diag"],
        "W": h["W"], "E": h["E"], "S": h["S"], "N": h["N"],
        "down": (va["down"] + vd["down"]) + cv["down"],
        "up": (va["up"] + vd["up"]) + cv["up"],
    }
    if concentration_state is None:
        concentration_state = bool(getattr(g, "concentration_state", True))
    if not concentration_state:
        return st
    nz, ny, nx = g.nz, g.ny, g.nx
    rho = density_weights(g).reshape(nz, ny, nx)
    inv = 1.0 / rho
    periodic = bool(getattr(g, "periodic_lon", False))
    shift = {
        "diag": inv,
        "W": np.roll(inv, 1, axis=2) if periodic else np.concatenate([inv[:, :, :1], inv[:, :, :-1]], axis=2),
        "E": np.roll(inv, -1, axis=2) if periodic else np.concatenate([inv[:, :, 1:], inv[:, :, -1:]], axis=2),
        "S": np.concatenate([inv[:, :1], inv[:, :-1]], axis=1),
        "N": np.concatenate([inv[:, 1:], inv[:, -1:]], axis=1),
        "down": np.concatenate([inv[:1], inv[:-1]], axis=0),
        "up": np.concatenate([inv[1:], inv[-1:]], axis=0),
    }
    return {k: (rho * st[k]) * shift[k] for k in STENCIL}


def assemble_transport_block(g) -> sp.csc_matrix:
    """The transport block as CSC (reference orbit/core/operator.py:281-347)."""
    st = transport_stencil(g)
    return stencil_to_csc(st, g.nz, g.ny, g.nx, bool(getattr(g, "periodic_lon", False)))


def assemble_horizontal_convdiff(g) -> sp.csc_matrix:
    return stencil_to_csc(horizontal_convdiff_stencil(g), g.nz, g.ny, g.nx, bool(getattr(g, "periodic_lon", False)))


def assemble_vertical_advection(g) -> sp.csc_matrix:
    return stencil_to_csc(vertical_advection_stencil(g), g.nz, g.ny, g.nx)


def assemble_vertical_diffusion(g) -> sp.csc_matrix:
    return stencil_to_csc(vertical_diffusion_stencil(g), g.nz, g.ny, g.nx)


def assemble_cmfmc_transport(g) -> sp.csc_matrix:
    return stencil_to_csc(cmfmc_stencil(g), g.nz, g.ny, g.nx)
