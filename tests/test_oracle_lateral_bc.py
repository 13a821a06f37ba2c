"""CPU checks of the lateral boundary inflow oracle (reference
orbit/core/lateral_bc.py:58-92) and of the boundary-file reader.

What pins r_in: on grids whose interface omega closes the air budget, the
open boundary's clean-air inflow is the only thing that keeps a uniform
mixing ratio from being steady, so T rho = r_in rho in every cell below the
top layer, lateral edges and corners included -- for net and for split winds.
A periodic longitude has no x inflow; reversing the winds turns r_in into
the outflow rate of the transport block's boundary faces."""
import types

import numpy as np
import pytest

from oracle import assembly_ref as asm
from oracle import lateral_bc_ref as lbc
from orbit_amd import synth


@pytest.mark.parametrize("split", [False, True])
def test_inflow_closes_the_air_budget_of_the_edges(split):
    grids, _, _ = synth.grid_workload(9, 7, 5, 1, seed=11, split=split, cmfmc=True)
    for g in grids[:3]:
        T = asm.assemble_transport_block(g).tocsr()
        rho = asm.density_weights(g)
        r_in = lbc.compute_lateral_boundary_inflow(g)
        assert (r_in >= 0).all() and r_in.max() > 0
        resid = ((T @ rho) / rho - r_in).reshape(g.nz, g.ny, g.nx)
        scale = np.abs(r_in).max()
        assert np.abs(resid[:-1]).max() < 1e-10 * scale
        # and the edges do not close without it
        r_no = ((T @ rho) / rho).reshape(g.nz, g.ny, g.nx)
        assert np.abs(r_no[:-1]).max() > 0.1 * scale
        # zero off the boundary
        inner = r_in.reshape(g.nz, g.ny, g.nx)[:, 1:-1, 1:-1]
        assert not inner.any()


def test_periodic_longitude_has_no_x_inflow():
    g = synth.make_reference_grids(8, 6, 4, seed=2, periodic=True)[0]
    r = lbc.compute_lateral_boundary_inflow(g).reshape(g.nz, g.ny, g.nx)
    assert not r[:, 1:-1, :].any()   # x edges get nothing; only the south / north rows
    assert r[:, 0, :].any() or r[:, -1, :].any()
    T = asm.assemble_transport_block(g).tocsr()
    rho = asm.density_weights(g)
    resid = ((T @ rho) / rho - r.ravel()).reshape(g.nz, g.ny, g.nx)
    assert np.abs(resid[:-1]).max() < 1e-10 * np.abs(r).max()


def test_reversed_winds_give_the_outflow_rates():
    g = synth.make_reference_grids(6, 5, 3, seed=4, cmfmc=False)[0]
    rev = types.SimpleNamespace(**vars(g))
    for name in ("UAvg", "VAvg", "U_face_e", "V_face_n"):
        setattr(rev, name, -np.asarray(getattr(g, name)))
    r_out = lbc.compute_lateral_boundary_inflow(rev).reshape(g.nz, g.ny, g.nx)
    # the outflow of the x boundaries (convdiff.py:204-230): u < 0 at the west, u >= 0 at the east
    dx = np.broadcast_to(g.dx[None, :], (g.nz, g.ny))
    west = np.where(g.UAvg[:, :, 0] < 0, -g.UAvg[:, :, 0] / dx, 0.0)
    east = np.where(g.U_face_e > 0, g.U_face_e / dx, 0.0)
    inner_rows = slice(1, -1)
    np.testing.assert_allclose(r_out[:, inner_rows, 0], west[:, inner_rows], rtol=1e-15)
    np.testing.assert_allclose(r_out[:, inner_rows, -1], east[:, inner_rows], rtol=1e-15)


def test_single_column_row_gets_both_x_faces():
    g = synth.make_reference_grids(1, 4, 3, seed=5)[0]
    r = lbc.compute_lateral_boundary_inflow(g).reshape(g.nz, g.ny, g.nx)
    dx = g.dx[None, 1:-1]
    expect = (np.maximum(g.UAvg[:, 1:-1, 0], 0.0) / dx + np.maximum(-g.U_face_e[:, 1:-1], 0.0) / dx)
    np.testing.assert_allclose(r[:, 1:-1, 0], expect, rtol=1e-15)


def _bc_grids(nz=3, ny=4, nx=5):
    grids = synth.make_reference_grids(nx, ny, nz, seed=7)
    for g in grids:
        g.lon = 60.0 + (np.arange(nx) + 0.5) * g.dlon
    return grids


def _write(path, grids, c_bc, species=("SO2",), lat=None, nz=None):
    g = grids[0]
    np.savez(path, lat=g.lat if lat is None else lat, lon=g.lon, nz=g.nz if nz is None else nz,
             species=np.array(species), c_bc=c_bc)


def test_boundary_file_reader_refuses_bad_files(tmp_path):
    """The reader's checks (reference lateral_bc.py:104-127) run before any
    device work: another grid, another c_bc layout, negative or non-finite
    values, species the run does not carry and missing arrays are
    BoundaryFileErrors (ValueErrors)."""
    from orbit_amd import lateral_bc
    grids = _bc_grids()
    good = np.ones((1, 8, 3, 4, 5))
    names = ["PM25", "SO2"]
    cases = {
        "grid mismatch": dict(lat=grids[0].lat + 1.0),
        "nz": dict(nz=4),
    }
    for what, kw in cases.items():
        p = tmp_path / f"{what}.npz"
        _write(p, grids, good, **kw)
        with pytest.raises(ValueError, match="on another grid"):
            lateral_bc.load_lateral_bc(str(p), grids, names)
    p = tmp_path / "shape.npz"
    _write(p, grids, np.ones((1, 7, 3, 4, 5)))
    with pytest.raises(ValueError, match="c_bc has layout"):
        lateral_bc.load_lateral_bc(str(p), grids, names)
    for bad in (np.nan, -1.0):
        c = good.copy()
        c[0, 2, 1, 1, 1] = bad
        p = tmp_path / "bad.npz"
        _write(p, grids, c)
        with pytest.raises(ValueError, match="finite and >= 0"):
            lateral_bc.load_lateral_bc(str(p), grids, names)
    p = tmp_path / "nokey.npz"
    np.savez(p, lat=grids[0].lat, lon=grids[0].lon, nz=grids[0].nz, c_bc=good)
    with pytest.raises(lateral_bc.BoundaryFileError, match="no species"):
        lateral_bc.load_lateral_bc(str(p), grids, names)
    p = tmp_path / "unknown.npz"
    _write(p, grids, good, species=("XYZ",))
    with pytest.raises(ValueError, match="not among the run's"):
        lateral_bc.load_lateral_bc(str(p), grids, names)


def test_merge_into_adds_per_bin():
    from orbit_amd import lateral_bc
    a = {1: [np.ones(3)] * 8}
    src = lateral_bc.LateralBCSource({1: [np.full(3, 2.0)] * 8, 4: [np.zeros(3)] * 8})
    out = lateral_bc.merge_into(a, src)
    assert np.array_equal(out[1][5], np.full(3, 3.0)) and 4 in out
    assert lateral_bc.merge_into(None, None) is None
    assert set(lateral_bc.merge_into(None, src)) == {1, 4}
