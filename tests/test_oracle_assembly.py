"""Pins of the transport-block assembly oracle (oracle/assembly_ref.py).

* Golden vectors: the reference's own assemble_transport_block and its four
  blocks on six small grids (terrain, degenerate cells, split fluxes, true
  outer faces, interface omega, CMFMC, periodic wrap, pressure form), written
  by tests/golden/make_assembly_golden.py.  Bar: 1e-13 relative to each
  matrix's largest entry.
* Closed forms and invariants the reference's tests state
  (tests/test_convdiff.py, test_advection.py, test_mixing.py,
  test_interface_omega_transport.py, test_concentration_form.py): the
  Patankar function's limits and A(1) = 1/(e-1); the exact 1-D exponential
  profile; the zero-wind limit against a separately written harmonic-mean
  diffusion; M-matrix signs; face-by-face conservation of the dP-area measure;
  a uniform mixing ratio as the equilibrium of the concentration form.  Each is
  chosen so that a dropped terrain weight, a swapped upwind side or a wrong
  sign fails it.
"""
import os
import types

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import assembly_ref as asm

GOLD = os.path.join(os.path.dirname(__file__), "golden", "assembly_golden.npz")


def _golden_cases():
    z = np.load(GOLD)
    names = sorted({k.split("/")[0] for k in z.files})
    return z, names


def golden_grid(z, name):
    g = types.SimpleNamespace()
    g.nz, g.ny, g.nx = (int(v) for v in z[f"{name}/dims"])
    split, periodic, conc = (bool(v) for v in z[f"{name}/flags"])
    g.dy, g.dlon, g.dlat = (float(v) for v in z[f"{name}/scalars"])
    g.has_split_fluxes, g.periodic_lon, g.concentration_state = split, periodic, conc
    pre = f"{name}/in/"
    for k in z.files:
        if k.startswith(pre):
            setattr(g, k[len(pre):], z[k])
    return g


def golden_matrix(z, name, m, N):
    p = f"{name}/out/{m}/"
    return sp.csc_matrix((z[p + "val"], (z[p + "row"], z[p + "col"])), shape=(N, N))


def _close(A, B, rtol=1e-13):
    scale = max(abs(B).max(), 1e-300)
    return abs(A - B).max() <= rtol * scale


@pytest.mark.parametrize("name", _golden_cases()[1])
def test_oracle_matches_reference_golden_vectors(name):
    z, _ = _golden_cases()
    g = golden_grid(z, name)
    N = g.nz * g.ny * g.nx
    got = {
        "T": asm.assemble_transport_block(g),
        "hcd": asm.assemble_horizontal_convdiff(g),
        "vadv": asm.assemble_vertical_advection(g),
        "vdiff": asm.assemble_vertical_diffusion(g),
        "conv": asm.assemble_cmfmc_transport(g),
    }
    for m, A in got.items():
        assert _close(A, golden_matrix(z, name, m, N)), (name, m)


def test_golden_vectors_cover_the_branches():
    z, names = _golden_cases()
    flags = np.array([z[f"{n}/flags"] for n in names])
    assert flags[:, 0].any() and (~flags[:, 0].astype(bool)).any()       # split and not
    assert flags[:, 1].any() and (~flags[:, 1].astype(bool)).any()       # periodic and not
    assert (~flags[:, 2].astype(bool)).any()                             # a pressure-form case
    assert any(f"{n}/in/omega_edge_plus" in z.files for n in names)
    assert any(f"{n}/in/U_face_e" in z.files for n in names)
    assert any((z[f"{n}/in/dP"] <= 0).any() for n in names)


# -- closed forms -------------------------------------------------------------
def test_patankar_function_closed_forms():
    A = asm.patankar_A
    assert A(np.array([0.0]))[0] == 1.0
    np.testing.assert_allclose(A(np.array([1.0])), [1.0 / np.expm1(1.0)], rtol=1e-15)
    np.testing.assert_allclose(A(np.array([1e-8, 1e-10])), 1.0 - 0.5 * np.array([1e-8, 1e-10]), rtol=0, atol=1e-15)
    assert (A(np.array([600.0, 1e6])) == 0.0).all()
    pe = np.linspace(0.0, 100.0, 5001)
    assert (np.diff(A(pe)) <= 1e-12).all() and (A(pe) >= 0).all()
    # continuity at both guards
    assert abs(A(np.array([0.9e-6]))[0] - A(np.array([1.1e-6]))[0]) < 1e-6
    assert abs(A(np.array([499.0]))[0] - A(np.array([501.0]))[0]) < 1e-6


def small_grid(nz=3, ny=4, nx=5, seed=0, wind=3.0, terrain=True, split=False):
    rng = np.random.default_rng(seed)
    g = types.SimpleNamespace(nz=nz, ny=ny, nx=nx, dlon=1.0, dlat=1.0)
    g.lat = 10.0 + np.arange(ny, dtype=float)
    g.dy = asm.EARTH_RADIUS * g.dlat * asm.DEG_TO_RAD
    g.dx = asm.EARTH_RADIUS * np.cos(g.lat * asm.DEG_TO_RAD) * g.dlon * asm.DEG_TO_RAD
    Bp = np.linspace(1.0, 0.3, nz + 1)
    Ps = np.full((ny, nx), 101325.0)
    if terrain:
        Ps *= 1.0 - 0.2 * rng.random((ny, nx))
    g.dP = (Bp[:-1] - Bp[1:])[:, None, None] * Ps[None]
    g.Dz = np.full((nz, ny, nx), 500.0) * (1.0 + 0.2 * rng.random((nz, ny, nx)))
    g.UAvg = rng.normal(0.0, wind, (nz, ny, nx))
    g.VAvg = rng.normal(0.0, wind, (nz, ny, nx))
    g.has_split_fluxes = split
    if split:
        g.UAvg_plus, g.UAvg_minus = np.maximum(g.UAvg, 0.0), np.maximum(-g.UAvg, 0.0)
        g.VAvg_plus, g.VAvg_minus = np.maximum(g.VAvg, 0.0), np.maximum(-g.VAvg, 0.0)
    g.Kxxyy = rng.uniform(50.0, 500.0, (nz, ny, nx))
    g.K_meander_u = rng.uniform(0.0, 100.0, (nz, ny, nx))
    g.K_meander_v = rng.uniform(0.0, 100.0, (nz, ny, nx))
    g.Kzz = rng.uniform(1.0, 100.0, (nz, ny, nx))
    g.periodic_lon = False
    return g


def test_exact_1d_exponential_profile():
    """Uniform u and K in 1-D: c = exp(u x / K) is an exact discrete steady
    state of the Patankar scheme at interior cells (reference
    tests/test_convdiff.py:339-394)."""
    nx, u, K = 10, 2.0, 5000.0
    g = types.SimpleNamespace(nz=1, ny=1, nx=nx, dlon=0.625, dlat=0.5, periodic_lon=False,
                              has_split_fluxes=False)
    g.lat = np.array([20.0])
    g.dy = asm.EARTH_RADIUS * g.dlat * asm.DEG_TO_RAD
    g.dx = asm.EARTH_RADIUS * np.cos(g.lat * asm.DEG_TO_RAD) * g.dlon * asm.DEG_TO_RAD
    g.dP = np.full((1, 1, nx), 50662.5)
    g.Dz = np.full((1, 1, nx), 500.0)
    g.UAvg = np.full((1, 1, nx), u)
    g.VAvg = np.zeros((1, 1, nx))
    g.Kxxyy = np.full((1, 1, nx), K)
    g.K_meander_u = np.zeros((1, 1, nx))
    g.K_meander_v = np.zeros((1, 1, nx))
    T = asm.assemble_horizontal_convdiff(g)
    c = np.exp(u * np.arange(nx) * g.dx[0] / K)
    r = T @ c
    assert (np.abs(r[1:-1]) < 1e-12 * c[1:-1] / g.dx[0] * u).all()
    # a wrong upwind side breaks it
    g.UAvg = -g.UAvg
    assert np.abs((asm.assemble_horizontal_convdiff(g) @ c)[1:-1]).max() > 1e-6 * c.max() / g.dx[0]


def test_zero_wind_is_harmonic_mean_diffusion_with_terrain_weights():
    """u = v = 0: each face carries K_face / h^2 times its terrain weight, with
    K_face the harmonic mean plus the meander term (written here from the
    formula, entry by entry)."""
    g = small_grid(wind=0.0)
    g.UAvg[:] = 0.0
    g.VAvg[:] = 0.0
    T = asm.assemble_horizontal_convdiff(g).toarray()
    nz, ny, nx = g.nz, g.ny, g.nx
    n = lambda k, j, i: (k * ny + j) * nx + i
    E = np.zeros_like(T)
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                for dj, di in ((0, 1), (1, 0)):
                    j2, i2 = j + dj, i + di
                    if j2 >= ny or i2 >= nx:
                        continue
                    Ka, Kb = g.Kxxyy[k, j, i], g.Kxxyy[k, j2, i2]
                    Kf = 2 * Ka * Kb / (Ka + Kb + 1e-30)
                    if di:
                        Kf += g.K_meander_u[k, j, i2]
                        h = g.dx[j]
                        wa = wb = 1.0
                    else:
                        Kf += g.K_meander_v[k, j2, i]
                        h = g.dy
                        xf = 0.5 * (g.dx[j] + g.dx[j2])
                        wa, wb = xf / g.dx[j], xf / g.dx[j2]
                    pf = 0.5 * (g.dP[k, j, i] + g.dP[k, j2, i2])
                    ra = Kf / h / h * wa * pf / g.dP[k, j, i]
                    rb = Kf / h / h * wb * pf / g.dP[k, j2, i2]
                    a, b = n(k, j, i), n(k, j2, i2)
                    E[a, a] += ra
                    E[a, b] -= ra
                    E[b, b] += rb
                    E[b, a] -= rb
    np.testing.assert_allclose(T, E, rtol=1e-12, atol=1e-12 * np.abs(E).max())


@pytest.mark.parametrize("split", [False, True])
def test_m_matrix_signs_and_dp_area_conservation(split):
    """Off-diagonals <= 0, diagonal >= 0; sum_n W_n T_nm = 0 at laterally
    interior columns with W = dx dy dP (reference tests/test_convdiff.py:397-455);
    the vertical blocks conserve sum_k dP_k c_k in every closed column."""
    g = small_grid(seed=3, split=split)
    T = asm.assemble_horizontal_convdiff(g).toarray()
    off = T - np.diag(np.diag(T))
    assert (off <= 1e-18).all() and (np.diag(T) >= 0).all()
    W = (np.broadcast_to(g.dx[None, :, None], g.dP.shape) * g.dy * g.dP).ravel()
    cs = (W @ T).reshape(g.dP.shape)
    assert np.abs(cs[:, 1:-1, 1:-1]).max() < 1e-11 * np.abs(W @ np.abs(T)).max()
    # vertical diffusion: dP-weighted symmetric, rows sum to zero
    Vd = asm.assemble_vertical_diffusion(g).toarray()
    Wd = np.diag(g.dP.ravel()) @ Vd
    np.testing.assert_allclose(Wd, Wd.T, rtol=0, atol=1e-12 * np.abs(Wd).max())
    assert np.abs(Vd.sum(axis=1)).max() < 1e-12 * np.abs(Vd).max()


def test_cmfmc_rows_and_mass_measure():
    """Interior rows of the convective block sum to zero, the top row is the
    detrained flux, and sum_k dP_k T_km = 0 below the top
    (reference tests/test_interface_omega_transport.py:173-190)."""
    g = small_grid(seed=4)
    rng = np.random.default_rng(4)
    g.CMFMC = np.abs(rng.normal(0.0, 0.01, (g.nz + 1, g.ny, g.nx)))
    g.CMFMC[0] = 0.0
    C = asm.assemble_cmfmc_transport(g).toarray()
    rs = C.sum(axis=1).reshape(g.dP.shape)
    assert np.abs(rs[:-1]).max() < 1e-12 * np.abs(C).max()
    rho = g.dP / (asm.GRAVITY * g.Dz)
    np.testing.assert_allclose(rs[-1], g.CMFMC[-1] / (rho[-1] * g.Dz[-1]), rtol=1e-12)
    cs = (g.dP.ravel() @ C).reshape(g.dP.shape)
    assert np.abs(cs[:-1]).max() < 1e-12 * np.abs(g.dP.ravel() @ np.abs(C)).max()


def test_interface_omega_closes_the_ground_and_top_is_outflow():
    g = small_grid(seed=5)
    rng = np.random.default_rng(5)
    oe = rng.normal(0.0, 0.3, (g.nz + 1, g.ny, g.nx))
    g.omega_edge_plus, g.omega_edge_minus = np.maximum(oe, 0.0), np.maximum(-oe, 0.0)
    wp, wm = asm.interface_omega(g)
    assert (wp[0] == 0).all() and (wm[0] == 0).all()
    A = asm.assemble_vertical_advection(g).toarray()
    rs = A.sum(axis=1).reshape(g.dP.shape)
    # row sum of layer k = (net outflow through its faces) / dP_k
    expect = (wp[:-1] * (np.arange(g.nz) > 0)[:, None, None] + wm[1:]
              - wm[:-1] * (np.arange(g.nz) > 0)[:, None, None]
              - wp[1:] * (np.arange(g.nz) < g.nz - 1)[:, None, None]) / g.dP
    np.testing.assert_allclose(rs, expect, rtol=1e-12, atol=1e-12 * np.abs(expect).max())
    # cell-centred fallback: the ground and the top are closed
    g2 = small_grid(seed=5)
    g2.omega_plus, g2.omega_minus = np.maximum(oe[:-1], 0.0), np.maximum(-oe[:-1], 0.0)
    wp2, wm2 = asm.interface_omega(g2)
    assert (wp2[0] == 0).all() and (wp2[-1] == 0).all() and (wm2[-1] == 0).all()
    np.testing.assert_array_equal(wp2[1:-1], g2.omega_plus[1:])


def test_concentration_form_keeps_rho_as_equilibrium():
    """Where T 1 = 0 (a uniform mixing ratio is steady), the concentration form
    has T_c rho = 0 (reference tests/test_concentration_form.py:68-79); and
    T_c = diag(rho) T diag(1/rho) entry by entry."""
    g = small_grid(seed=6, wind=0.0)
    g.UAvg[:] = 0.0
    g.VAvg[:] = 0.0
    g.concentration_state = False
    Tp = asm.assemble_transport_block(g)
    g.concentration_state = True
    Tc = asm.assemble_transport_block(g)
    rho = asm.density_weights(g)
    r1 = (Tp @ np.ones(Tp.shape[0])).reshape(g.dP.shape)
    rc = (Tc @ rho).reshape(g.dP.shape) / rho.reshape(g.dP.shape)
    inner = (slice(None), slice(1, -1), slice(1, -1))
    assert np.abs(r1[inner]).max() < 1e-12 * abs(Tp).max()
    assert np.abs(rc[inner]).max() < 1e-12 * abs(Tc).max()
    D = sp.diags(rho) @ Tp @ sp.diags(1.0 / rho)
    assert abs(D - Tc).max() <= 1e-14 * abs(Tc).max()


def test_periodic_reference_grids_close_the_air_budget_through_the_wrap_face():
    """The periodic reference-like grids of synth (wrap face instead of the west
    and east outer faces, omega diagnosed with the wrap flux): every row of the
    oracle's concentration-form block that is not on the south / north edge or
    the top layer has T rho = 0, the x-edge columns included, and the wrap
    couplings are present and of M-matrix sign."""
    from orbit_amd import synth
    grids, _, _ = synth.grid_workload(8, 6, 5, 1, seed=3, periodic=True)
    g = grids[1]
    T = asm.assemble_transport_block(g).tocsr()
    rho = asm.density_weights(g)
    r = (T @ rho / rho).reshape(g.nz, g.ny, g.nx)
    assert np.abs(r[:-1, 1:-1, :]).max() < 1e-12 * abs(T).max()
    n3 = np.arange(g.nz * g.ny * g.nx).reshape(g.nz, g.ny, g.nx)
    w = np.array([T[a, b] for a, b in zip(n3[:, :, 0].ravel(), n3[:, :, -1].ravel())])
    e = np.array([T[a, b] for a, b in zip(n3[:, :, -1].ravel(), n3[:, :, 0].ravel())])
    assert (w < 0).all() and (e < 0).all()
    # the same grid without the wrap has no such couplings
    gp = synth.grid_workload(8, 6, 5, 1, seed=3)[0][1]
    Tn = asm.assemble_transport_block(gp).tocsr()
    assert all(Tn[a, b] == 0 for a, b in zip(n3[:, :, 0].ravel(), n3[:, :, -1].ravel()))
