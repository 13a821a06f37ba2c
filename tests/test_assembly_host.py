"""Host logic of the device-assembly binding (no GPU): which grid fields are
passed to the library, and the stencil-to-CSC conversion."""
import types

import numpy as np
import scipy.sparse as sp

from oracle import assembly_ref as asm
from orbit_amd import assembly, synth


def _grid(split=False, iface=True):
    g = synth.make_reference_grids(6, 5, 4, seed=1, interface_omega=iface)[0]
    g = types.SimpleNamespace(**vars(g))
    if split:
        g.has_split_fluxes = True
        g.UAvg_plus, g.UAvg_minus = np.maximum(g.UAvg, 0), np.maximum(-g.UAvg, 0)
        g.VAvg_plus, g.VAvg_minus = np.maximum(g.VAvg, 0), np.maximum(-g.VAvg, 0)
    return g


def test_field_selection_follows_the_reference_branches():
    keep = []
    f = assembly.grid_fields(_grid(), keep)
    assert f.K_meander_u and f.K_meander_v and not f.UAvg_plus     # net fluxes
    assert f.omega_edge_plus and not f.omega_plus                  # interface omega preferred
    assert f.CMFMC and f.U_face_e and f.V_face_n and not f.U_plus_face_e
    f = assembly.grid_fields(_grid(split=True, iface=False), keep)
    assert f.UAvg_plus and f.VAvg_minus and not f.K_meander_u      # split fluxes
    assert f.omega_plus and not f.omega_edge_plus                  # cell-centred fallback
    g = _grid()
    g.omega_edge_plus = g.omega_edge_plus[:-1]                     # wrong level count: ignored
    assert not assembly.grid_fields(g, keep).omega_edge_plus
    for a in keep:
        assert a.dtype == np.float64 and a.flags.c_contiguous


def test_stencil_to_csc_agrees_with_the_oracle_layout():
    g = _grid()
    st = asm.transport_stencil(g)
    A = assembly.stencil_to_csc(st, g.nz, g.ny, g.nx)
    B = asm.stencil_to_csc(st, g.nz, g.ny, g.nx)
    assert abs(A - B).max() == 0.0
    assert abs(A - asm.assemble_transport_block(g)).max() == 0.0


def test_synthetic_reference_grids_close_the_air_budget():
    """The diagnosed interface omega makes every interior row of the
    concentration-form block vanish on rho; the cell-centred reading does not."""
    for iface, small in ((True, True), (False, False)):
        g = synth.make_reference_grids(12, 10, 6, seed=3, interface_omega=iface)[2]
        T = asm.assemble_transport_block(g)
        rho = asm.density_weights(g)
        r = (T @ rho / rho).reshape(6, 10, 12) * 86400.0
        p90 = np.percentile(np.abs(r[:-1, 1:-1, 1:-1]), 90)
        assert (p90 < 1e-9) if small else (p90 > 0.05)
        off = T - sp.diags(T.diagonal())
        assert off.max() <= 0.0
