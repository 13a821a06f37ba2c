"""Device-side assembly of the per-bin transport block from grid fields.

Replaces the reference's host CSR assembly of the shared transport block
(orbit/core/operator.py:281-353: Patankar horizontal convection-diffusion,
omega advection, Kzz diffusion, CMFMC convection, concentration form) and its
air-mass pre-flight (operator.py:598-652 with the abort rule of
orbit/cli.py:1591-1614).  Grid objects carry the reference's GridData
attribute names; an empty or missing optional field is absent.

* ``assemble_transport_block(grid)`` -> scipy CSC, computed on the device
  (the reference function's return type, for callers that still want CSR/CSC).
* ``OperatorSet.from_grids(grids, D)`` (in core) builds a month's operator set
  with no host CSR: the kernels write the column blocks, stencils and
  diagonals of the device layout directly.
* ``air_mass_preflight(op, grids)`` runs the mass-balance check on the
  device-assembled blocks and applies the reference's abort threshold.
"""
from __future__ import annotations

import ctypes as C
from typing import List, Sequence

import numpy as np
import scipy.sparse as sp

from . import _lib
from ._lib import check, lib

GRAVITY = 9.80665
MASS_BALANCE_TOL = 0.05   # per day, p90 of interior |row sums| (orbit/cli.py:1604)

_ARRAYS = ("lat", "dx", "dP", "Dz", "UAvg", "VAvg", "UAvg_plus", "UAvg_minus", "VAvg_plus",
           "VAvg_minus", "Kxxyy", "K_meander_u", "K_meander_v", "Kzz", "omega_edge_plus",
           "omega_edge_minus", "omega_plus", "omega_minus", "CMFMC", "U_face_e", "U_plus_face_e",
           "U_minus_face_e", "V_face_n", "V_plus_face_n", "V_minus_face_n", "UAvg_wrap", "UAvg_plus_wrap",
           "UAvg_minus_wrap", "K_meander_u_wrap")


class GridFields(C.Structure):
    _fields_ = ([("nz", C.c_int), ("ny", C.c_int), ("nx", C.c_int), ("split_fluxes", C.c_int),
                 ("concentration_state", C.c_int), ("periodic_lon", C.c_int), ("dy", C.c_double),
                 ("dlon", C.c_double), ("dlat", C.c_double)]
                + [(a, C.c_void_p) for a in _ARRAYS])


def _present(g, name) -> bool:
    a = getattr(g, name, None)
    return a is not None and np.size(a) > 0


def grid_fields(g, keep: List) -> GridFields:
    """ctypes view of a grid's fields; the contiguous float64 copies are appended
    to `keep`, which must outlive the call they are passed to."""
    f = GridFields()
    f.nz, f.ny, f.nx = int(g.nz), int(g.ny), int(g.nx)
    f.split_fluxes = int(bool(getattr(g, "has_split_fluxes", False)))
    f.concentration_state = int(bool(getattr(g, "concentration_state", True)))
    f.periodic_lon = int(bool(getattr(g, "periodic_lon", False)))
    f.dy, f.dlon, f.dlat = float(g.dy), float(g.dlon), float(g.dlat)
    nz = f.nz
    use = set(_ARRAYS)
    if f.split_fluxes:
        use -= {"K_meander_u", "K_meander_v", "K_meander_u_wrap"}
    else:
        use -= {"UAvg_plus", "UAvg_minus", "VAvg_plus", "VAvg_minus", "UAvg_plus_wrap", "UAvg_minus_wrap"}
    if not (f.periodic_lon and f.nx > 1):  # the wrap face (convdiff.py:150-200)
        use -= {"UAvg_wrap", "UAvg_plus_wrap", "UAvg_minus_wrap", "K_meander_u_wrap"}
    # interface omega only with nz+1 levels (advection.py:398-400)
    if not (_present(g, "omega_edge_plus") and np.shape(g.omega_edge_plus)[0] == nz + 1):
        use -= {"omega_edge_plus", "omega_edge_minus"}
    else:
        use -= {"omega_plus", "omega_minus"}
    for name in _ARRAYS:
        if name in use and _present(g, name):
            a = np.ascontiguousarray(getattr(g, name), dtype=np.float64)
            keep.append(a)
            setattr(f, name, a.ctypes.data)
        else:
            setattr(f, name, None)
    return f


def transport_stencil(g, ctx=None) -> dict:
    """Rows of the transport block of one bin from the device: dict of (nz, ny, nx)
    arrays diag, W, E, S, N, down, up (0 where a neighbour does not exist)."""
    from .core import default_context
    ctx = ctx or default_context()
    keep: List = []
    f = grid_fields(g, keep)
    N = f.nz * f.ny * f.nx
    out = np.empty(7 * N)
    check(lib().orbit_transport_assemble(ctx.handle, C.byref(f), out.ctypes.data), allow_warn=False)
    out = out.reshape(7, f.nz, f.ny, f.nx)
    return dict(zip(("diag", "W", "E", "S", "N", "down", "up"), out))


def stencil_to_csc(st: dict, nz: int, ny: int, nx: int, periodic: bool = False) -> sp.csc_matrix:
    """CSC of a 7-point stencil; with periodic, W of i = 0 couples to i = nx-1
    and E of i = nx-1 to i = 0 (summed with the interior coupling when nx = 2)."""
    N = nz * ny * nx
    n3 = np.arange(N).reshape(nz, ny, nx)
    rows, cols, vals = [n3.ravel()], [n3.ravel()], [st["diag"].ravel()]
    sl = {"W": (np.s_[:, :, 1:], np.s_[:, :, :-1]), "E": (np.s_[:, :, :-1], np.s_[:, :, 1:]),
          "S": (np.s_[:, 1:, :], np.s_[:, :-1, :]), "N": (np.s_[:, :-1, :], np.s_[:, 1:, :]),
          "down": (np.s_[1:], np.s_[:-1]), "up": (np.s_[:-1], np.s_[1:])}
    if periodic and nx > 1:
        sl["W_wrap"] = (np.s_[:, :, :1], np.s_[:, :, -1:])
        sl["E_wrap"] = (np.s_[:, :, -1:], np.s_[:, :, :1])
    for name, (own, nbr) in sl.items():
        rows.append(n3[own].ravel())
        cols.append(n3[nbr].ravel())
        vals.append(st[name.split("_")[0]][own].ravel())
    A = sp.csc_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(N, N))
    A.sum_duplicates()
    A.eliminate_zeros()
    return A


def assemble_transport_block(g, ctx=None) -> sp.csc_matrix:
    """The reference's assemble_transport_block (orbit/core/operator.py:281-347),
    computed on the device; returns CSC."""
    return stencil_to_csc(transport_stencil(g, ctx), int(g.nz), int(g.ny), int(g.nx),
                          bool(getattr(g, "periodic_lon", False)))


def density_weights(g) -> np.ndarray:
    """rho = dP / (g Dz), 1 where degenerate, flat (N,) (grid_data.py:543-553)."""
    dP = np.asarray(g.dP, dtype=np.float64)
    Dz = np.asarray(g.Dz, dtype=np.float64)
    ok = (dP > 0) & (Dz > 0)
    return np.where(ok, dP / (GRAVITY * np.where(ok, Dz, 1.0)), 1.0).ravel()


def create_opset_handle(ctx, grids: Sequence, D: np.ndarray, dtau: float):
    """orbit_opset_create_grids: returns the raw handle (core.OperatorSet wraps it)."""
    keep: List = []
    arr = (GridFields * len(grids))(*[grid_fields(g, keep) for g in grids])
    D = np.ascontiguousarray(D, dtype=np.float64)
    h = C.c_void_p()
    check(lib().orbit_opset_create_grids(ctx.handle, len(grids), int(D.shape[1]), float(dtau), arr,
                                         D.ctypes.data, C.byref(h)), allow_warn=False)
    return h


class AirMassImbalance(RuntimeError):
    pass


def air_mass_preflight(op, grids: Sequence, tol: float = MASS_BALANCE_TOL, allow_imbalance: bool = False) -> dict:
    """Mass-balance statistics of the opset's transport blocks (per day, interior
    cells, all bins; orbit/core/operator.py:598-652).  Raises AirMassImbalance when
    the interior p90 exceeds `tol`, as the reference's run aborts
    (orbit/cli.py:1604-1614), unless allow_imbalance."""
    g0 = grids[0]
    if bool(getattr(g0, "concentration_state", True)):
        rho = density_weights(g0)
    else:
        rho = np.ones(int(g0.nz) * int(g0.ny) * int(g0.nx))
    mb = op.mass_balance(rho)
    if mb["interior_p90"] > tol and not allow_imbalance:
        raise AirMassImbalance(
            f"transport block interior row sums p90 = {mb['interior_p90']:.4f} per day exceed the "
            f"air-mass balance tolerance {tol} per day: the operator creates or destroys air")
    return mb
