"""Write tests/golden/assembly_golden.npz from the reference's own assembly code.

The exact values of the transport block (Patankar horizontal convection-
diffusion with terrain weights, split fluxes, outer faces and the periodic
wrap; interface or cell-centred omega advection; Kzz diffusion on the dP
measure; CMFMC convection; the concentration-form transform) on a grid with
terrain, degenerate cells and random fields are fixed only by the reference
code (orbit/core/operator.py:281-353, convdiff.py:42-387,
advection.py:383-525, mixing.py:14-117, convection.py:25-140).  This script
imports those modules (numpy/scipy only) from the reference checkout in this
container and records inputs and outputs as COO triplets.  Run from the repo
root:

    python tests/golden/make_assembly_golden.py /root/reference

The output is data only; no reference source is stored.
"""
import sys

import numpy as np

FIELDS3 = ("UAvg", "VAvg", "UAvg_plus", "UAvg_minus", "VAvg_plus", "VAvg_minus", "Kxxyy",
           "K_meander_u", "K_meander_v", "Kzz", "dP", "Dz", "omega_plus", "omega_minus",
           "omega_edge_plus", "omega_edge_minus", "CMFMC", "U_face_e", "U_plus_face_e",
           "U_minus_face_e", "V_face_n", "V_plus_face_n", "V_minus_face_n", "UAvg_wrap",
           "UAvg_plus_wrap", "UAvg_minus_wrap", "K_meander_u_wrap", "dx", "lat")


def build_case(GridData, seed, split, interface_omega, edges, periodic, conc, degenerate):
    rng = np.random.default_rng(seed)
    nz, ny, nx = 4, 5, 6
    g = GridData()
    g.nz, g.ny, g.nx = nz, ny, nx
    g.dlon, g.dlat = 0.625, 0.5
    g.lat = 18.0 + g.dlat * np.arange(ny)
    g.lon = 70.0 + g.dlon * np.arange(nx)
    R, d2r = 6378137.0, np.pi / 180.0
    g.dy = R * g.dlat * d2r
    g.dx = R * np.cos(g.lat * d2r) * g.dlon * d2r
    g.Dz = np.array([150.0, 450.0, 1200.0, 2500.0])[:, None, None] * (1.0 + 0.1 * rng.random((nz, ny, nx)))
    g.dP = 1.15 * 9.80665 * g.Dz * (1.0 + 0.1 * rng.standard_normal((1, ny, nx))) \
        * np.array([1.0, 0.93, 0.8, 0.6])[:, None, None]
    if degenerate:
        g.dP[1, 2, 3] = 0.0
        g.Dz[2, 4, 1] = 0.0
    U = rng.normal(0.0, 6.0, (nz, ny, nx))
    V = rng.normal(0.0, 6.0, (nz, ny, nx))
    g.UAvg, g.VAvg = U, V
    g.has_split_fluxes = split
    if split:
        g.UAvg_plus = np.maximum(U, 0.0) + 0.5 * rng.random((nz, ny, nx))
        g.UAvg_minus = np.maximum(-U, 0.0) + 0.5 * rng.random((nz, ny, nx))
        g.VAvg_plus = np.maximum(V, 0.0) + 0.5 * rng.random((nz, ny, nx))
        g.VAvg_minus = np.maximum(-V, 0.0) + 0.5 * rng.random((nz, ny, nx))
    g.Kxxyy = rng.uniform(50.0, 3000.0, (nz, ny, nx))
    g.K_meander_u = rng.uniform(0.0, 500.0, (nz, ny, nx))
    g.K_meander_v = rng.uniform(0.0, 500.0, (nz, ny, nx))
    g.Kzz = 100.0 * np.exp(-np.arange(nz))[:, None, None] * rng.uniform(0.5, 1.5, (nz, ny, nx))
    om = rng.normal(0.0, 0.3, (nz, ny, nx))
    g.omega = om
    g.omega_plus = np.maximum(om, 0.0)
    g.omega_minus = np.maximum(-om, 0.0)
    if interface_omega:
        oe = rng.normal(0.0, 0.3, (nz + 1, ny, nx))
        oe[0] = rng.normal(0.0, 0.3, (ny, nx))   # the ground must be ignored
        g.omega_edge = oe
        g.omega_edge_plus = np.maximum(oe, 0.0)
        g.omega_edge_minus = np.maximum(-oe, 0.0)
        g.has_interface_omega = True
    cm = np.abs(rng.normal(0.0, 0.01, (nz + 1, ny, nx))) * (rng.random((nz + 1, ny, nx)) < 0.6)
    cm[0] = 0.0
    cm[2, 1, 1] = -0.003   # negative flux is clipped
    g.CMFMC = cm
    if edges:
        g.U_face_e = rng.normal(0.0, 6.0, (nz, ny))
        g.V_face_n = rng.normal(0.0, 6.0, (nz, nx))
        g.U_plus_face_e = np.maximum(g.U_face_e, 0.0) + 0.3
        g.U_minus_face_e = np.maximum(-g.U_face_e, 0.0) + 0.3
        g.V_plus_face_n = np.maximum(g.V_face_n, 0.0) + 0.3
        g.V_minus_face_n = np.maximum(-g.V_face_n, 0.0) + 0.3
    g.periodic_lon = periodic
    if periodic:
        g.UAvg_wrap = rng.normal(0.0, 6.0, (nz, ny))
        g.UAvg_plus_wrap = np.maximum(g.UAvg_wrap, 0.0) + 0.2
        g.UAvg_minus_wrap = np.maximum(-g.UAvg_wrap, 0.0) + 0.2
        g.K_meander_u_wrap = rng.uniform(0.0, 500.0, (nz, ny))
    g.concentration_state = conc
    return g


CASES = {
    "plain": dict(seed=1, split=False, interface_omega=False, edges=False, periodic=False, conc=True,
                  degenerate=True),
    "split_edges_iface": dict(seed=2, split=True, interface_omega=True, edges=True, periodic=False,
                              conc=True, degenerate=True),
    "nonsplit_edges_iface": dict(seed=3, split=False, interface_omega=True, edges=True, periodic=False,
                                 conc=True, degenerate=False),
    "pressure_form": dict(seed=4, split=False, interface_omega=True, edges=False, periodic=False,
                          conc=False, degenerate=False),
    "periodic_split": dict(seed=5, split=True, interface_omega=False, edges=True, periodic=True,
                           conc=True, degenerate=False),
    "periodic_nonsplit": dict(seed=6, split=False, interface_omega=True, edges=False, periodic=True,
                              conc=True, degenerate=False),
}


def main(ref_root):
    sys.path.insert(0, ref_root)
    from orbit.core.advection import assemble_vertical_advection
    from orbit.core.convdiff import assemble_horizontal_convdiff
    from orbit.core.convection import assemble_cmfmc_transport
    from orbit.core.grid_data import GridData
    from orbit.core.indexing import CellIndexer
    from orbit.core.mixing import assemble_vertical_diffusion
    from orbit.core.operator import assemble_transport_block

    out = {}
    for name, kw in CASES.items():
        g = build_case(GridData, **kw)
        idx = CellIndexer(g.nz, g.ny, g.nx)
        mats = {
            "T": assemble_transport_block(g, idx),
            "hcd": assemble_horizontal_convdiff(g, idx),
            "vadv": assemble_vertical_advection(g, idx),
            "vdiff": assemble_vertical_diffusion(g, idx),
            "conv": assemble_cmfmc_transport(g, idx),
        }
        out[f"{name}/dims"] = np.array([g.nz, g.ny, g.nx])
        out[f"{name}/flags"] = np.array([int(kw["split"]), int(kw["periodic"]), int(kw["conc"])])
        out[f"{name}/scalars"] = np.array([g.dy, g.dlon, g.dlat])
        for f in FIELDS3:
            a = getattr(g, f, None)
            if a is not None and np.size(a) > 0:
                out[f"{name}/in/{f}"] = np.asarray(a, dtype=np.float64)
        for m, A in mats.items():
            c = A.tocoo()
            out[f"{name}/out/{m}/row"] = c.row.astype(np.int32)
            out[f"{name}/out/{m}/col"] = c.col.astype(np.int32)
            out[f"{name}/out/{m}/val"] = c.data.astype(np.float64)
        print(name, {m: A.nnz for m, A in mats.items()})
    np.savez_compressed("tests/golden/assembly_golden.npz", **out)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
