"""CPU oracle of the lateral boundary inflow -- TEST INFRASTRUCTURE ONLY.

Restates compute_lateral_boundary_inflow (reference
orbit/core/lateral_bc.py:58-92): the first-order upwind inflow rate r_in >= 0
(1/s) of every cell through its boundary faces -- the outflow rule of the
transport block's open boundaries (orbit/core/convdiff.py:204-230, 318-370)
with the wind reversed: the west face of i = 0 and the true east outer face of
i = nx-1 (not on a periodic longitude), the south face of j = 0 and the true
north outer face of j = ny-1 with the outer faces' zonal width over the edge
rows' width; no terrain ratio; split fluxes take the plus / minus winds.  The
source on the orbit right-hand side is r_in * c_bc (lateral_bc.py:95-137).

Pinned by tests/test_oracle_lateral_bc.py: on grids whose interface omega
closes the air budget (net or split winds), T rho = r_in rho in every cell
below the top layer, lateral edges and corners included (the inflow replaces
exactly the clean air that the open boundary lets in); a periodic longitude
has no x inflow; reversing the winds turns r_in into the outflow rates.
"""
from __future__ import annotations

import numpy as np

from oracle.assembly_ref import DEG_TO_RAD, EARTH_RADIUS, edge_face


def compute_lateral_boundary_inflow(g) -> np.ndarray:
    """r_in (N,) for one bin's grid (lateral_bc.py:58-92)."""
    nz, ny, nx = g.nz, g.ny, g.nx
    split = bool(getattr(g, "has_split_fluxes", False))
    dx3 = np.broadcast_to(np.asarray(g.dx, dtype=np.float64)[None, :, None], (nz, ny, nx))
    valid = dx3 > 0
    sdx = np.where(valid, dx3, 1.0)
    rate = np.zeros((nz, ny, nx))
    if not bool(getattr(g, "periodic_lon", False)):
        if split:
            u_w = np.asarray(g.UAvg_plus, dtype=np.float64)[:, :, 0]
            u_e = edge_face(g, "U_minus")
        else:
            u_w = np.maximum(np.asarray(g.UAvg, dtype=np.float64)[:, :, 0], 0.0)
            u_e = np.maximum(-edge_face(g, "U"), 0.0)
        rate[:, :, 0] += np.where(valid[:, :, 0] & (u_w > 0), u_w / sdx[:, :, 0], 0.0)
        rate[:, :, -1] += np.where(valid[:, :, -1] & (u_e > 0), u_e / sdx[:, :, -1], 0.0)
    dy = float(g.dy) if g.dy > 0 else 1.0
    dx = np.asarray(g.dx, dtype=np.float64)
    lat = np.asarray(g.lat, dtype=np.float64)
    zonal = lambda la: EARTH_RADIUS * np.cos(la * DEG_TO_RAD) * g.dlon * DEG_TO_RAD  # noqa: E731
    tr_s = zonal(lat[0] - g.dlat / 2.0) / (dx[0] if dx[0] > 0 else 1.0)
    tr_n = zonal(lat[-1] + g.dlat / 2.0) / (dx[-1] if dx[-1] > 0 else 1.0)
    if split:
        v_s = np.asarray(g.VAvg_plus, dtype=np.float64)[:, 0, :]
        v_n = edge_face(g, "V_minus")
    else:
        v_s = np.maximum(np.asarray(g.VAvg, dtype=np.float64)[:, 0, :], 0.0)
        v_n = np.maximum(-edge_face(g, "V"), 0.0)
    rate[:, 0, :] += np.where(valid[:, 0, :] & (v_s > 0), v_s / dy * tr_s, 0.0)
    rate[:, -1, :] += np.where(valid[:, -1, :] & (v_n > 0), v_n / dy * tr_n, 0.0)
    return rate.ravel()
