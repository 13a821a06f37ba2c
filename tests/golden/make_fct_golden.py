"""Write tests/golden/fct_golden.npz from the reference's own FCT module.

The exact 2-D values of the horizontal FCT source and of its frozen
tangent-linear operator on a metric + terrain + split-flux grid are fixed by
nothing but the reference code itself (orbit/core/fct.py:459-572), so this
script imports that module (numpy/scipy only) from the reference checkout in
this container and records inputs and outputs.  Run from the repo root:

    python tests/golden/make_fct_golden.py /root/reference

The output is data only; no reference source is stored.
"""
import sys
import types

import numpy as np


def build_case(seed, split, with_degenerate):
    rng = np.random.default_rng(seed)
    nz, ny, nx = 3, 6, 7
    g = types.SimpleNamespace()
    g.nz, g.ny, g.nx = nz, ny, nx
    g.dx = 5.0e4 * np.cos(np.deg2rad(10.0 + 0.5 * np.arange(ny)))
    g.dy = 5.5e4
    g.Dz = np.broadcast_to(np.array([120.0, 400.0, 1500.0])[:, None, None], (nz, ny, nx)).copy()
    g.Dz *= 1.0 + 0.05 * rng.standard_normal((nz, ny, nx))
    g.dP = 1.2 * 9.80665 * g.Dz * (1.0 + 0.08 * rng.standard_normal((1, ny, nx)))
    if with_degenerate:
        g.dP[1, 2, 3] = 0.0
    g.has_split_fluxes = split
    U = rng.normal(0.0, 4.0, (nz, ny, nx))
    V = rng.normal(0.0, 4.0, (nz, ny, nx))
    if split:
        g.UAvg_plus = np.abs(U) * (rng.random((nz, ny, nx)) < 0.7)
        g.UAvg_minus = np.abs(rng.normal(0.0, 3.0, (nz, ny, nx))) * (rng.random((nz, ny, nx)) < 0.5)
        g.VAvg_plus = np.abs(V) * (rng.random((nz, ny, nx)) < 0.6)
        g.VAvg_minus = np.abs(rng.normal(0.0, 3.0, (nz, ny, nx))) * (rng.random((nz, ny, nx)) < 0.6)
    else:
        g.UAvg = U
        g.VAvg = V
    g.concentration_state = True
    N = nz * ny * nx
    x = 1.0 + rng.random((nz, ny, nx))
    x[:, 2:4, 3:6] *= 5.0                           # a front so the limiter acts
    rho = g.dP / (9.80665 * g.Dz)
    rho = np.where((g.dP > 0) & (g.Dz > 0), rho, 1.0)
    c = (x * rho).ravel()
    ghost = (np.clip(x + rng.normal(0.0, 1.5, x.shape), 0.1, None) * rho).ravel()
    return g, c, ghost, N


def main(ref_root):
    sys.path.insert(0, ref_root)
    from orbit.core import fct  # reference module, numpy/scipy only

    out = {}
    for name, seed, split, degen in (("plain", 11, False, False),
                                     ("split", 12, True, True)):
        g, c, ghost, N = build_case(seed, split, degen)
        dt = 10800.0
        idx = types.SimpleNamespace(N=N)
        d_lim = fct.compute_horizontal_fct_source(g, idx, c, dt)
        d_unl = fct.compute_horizontal_fct_source(g, idx, c, dt, limited=False)
        d_gh = fct.compute_horizontal_fct_source(g, idx, c, dt, c_ghost_flat=ghost)
        d_gh_unl = fct.compute_horizontal_fct_source(g, idx, c, dt, limited=False,
                                                     c_ghost_flat=ghost)
        L = fct.assemble_fct_linear_operator(g, idx, c, dt).tocsr()
        L.sum_duplicates()
        L.sort_indices()
        p = name + "_"
        for key in ("dx", "Dz", "dP"):
            out[p + key] = np.asarray(getattr(g, key))
        out[p + "dy"] = np.array(g.dy)
        if split:
            for key in ("UAvg_plus", "UAvg_minus", "VAvg_plus", "VAvg_minus"):
                out[p + key] = getattr(g, key)
        else:
            out[p + "UAvg"] = g.UAvg
            out[p + "VAvg"] = g.VAvg
        out[p + "shape"] = np.array([g.nz, g.ny, g.nx])
        out[p + "split"] = np.array(split)
        out[p + "c"] = c
        out[p + "ghost"] = ghost
        out[p + "dt"] = np.array(dt)
        out[p + "d_lim"] = d_lim
        out[p + "d_unl"] = d_unl
        out[p + "d_ghost"] = d_gh
        out[p + "d_ghost_unl"] = d_gh_unl
        out[p + "L_indptr"] = L.indptr
        out[p + "L_indices"] = L.indices
        out[p + "L_data"] = L.data
    np.savez_compressed("tests/golden/fct_golden.npz", **out)
    print("wrote tests/golden/fct_golden.npz", {k: v.shape for k, v in out.items() if v.ndim})


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
