"""GPU parity of the FCT source and frozen tangent-linear operator.

Against the reference's golden vectors (tests/golden/fct_golden.npz) and the
CPU oracle on a synthetic metric grid, at 1e-12 relative: the kernels compute
in FP64 with the reference's operation order per face."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import fct_ref
from orbit_amd import fct as hfct, synth
from tests.test_oracle_fct import GOLDEN, _golden_grid


@pytest.mark.parametrize("case", ["plain", "split"])
def test_fct_source_matches_golden(case):
    z = np.load(GOLDEN)
    p = case + "_"
    g = _golden_grid(z, p)
    c, gh, dt = z[p + "c"], z[p + "ghost"], float(z[p + "dt"])
    scale = np.abs(z[p + "d_unl"]).max()
    for got, key in ((hfct.compute_horizontal_fct_source(g, None, c, dt), "d_lim"),
                     (hfct.compute_horizontal_fct_source(g, None, c, dt, limited=False), "d_unl"),
                     (hfct.compute_horizontal_fct_source(g, None, c, dt, c_ghost_flat=gh), "d_ghost"),
                     (hfct.compute_horizontal_fct_source(g, None, c, dt, False, gh), "d_ghost_unl")):
        np.testing.assert_allclose(got, z[p + key], rtol=1e-12, atol=1e-13 * scale, err_msg=key)


@pytest.mark.parametrize("case", ["plain", "split"])
def test_fct_linop_matches_golden(case):
    import scipy.sparse as sp
    z = np.load(GOLDEN)
    p = case + "_"
    g = _golden_grid(z, p)
    N = g.nz * g.ny * g.nx
    ref = sp.csr_matrix((z[p + "L_data"], z[p + "L_indices"], z[p + "L_indptr"]), shape=(N, N))
    L = hfct.assemble_fct_linear_operator(g, None, z[p + "c"], float(z[p + "dt"]))
    got = L.tocsr()
    assert abs(got - ref).max() <= 1e-12 * abs(ref).max()
    x = np.random.default_rng(0).standard_normal(N)
    np.testing.assert_allclose(L @ x, ref @ x, rtol=1e-11, atol=1e-13 * np.abs(ref @ x).max())
    np.testing.assert_allclose(L.T @ x, ref.T @ x, rtol=1e-11, atol=1e-13 * np.abs(ref.T @ x).max())


def test_fct_on_synthetic_grid_matches_oracle():
    wl = synth.make_workload(24, 20, 15, 1, seed=1)
    g = wl.grids[2]
    rng = np.random.default_rng(4)
    c = wl.rho * (1.0 + rng.random(wl.N))
    c.reshape(wl.nz, wl.ny, wl.nx)[:, 5:9, 7:13] *= 6.0
    ghost = wl.rho * (1.0 + rng.random(wl.N))
    for kw in (dict(), dict(limited=False), dict(c_ghost_flat=ghost)):
        ref = fct_ref.compute_horizontal_fct_source(g, c, 10800.0, **kw)
        got = hfct.compute_horizontal_fct_source(g, None, c, 10800.0, **kw)
        np.testing.assert_allclose(got, ref, rtol=1e-11, atol=1e-13 * np.abs(ref).max())
    L = hfct.assemble_fct_linear_operator(g, None, c, 10800.0)
    np.testing.assert_allclose(L @ c, -fct_ref.compute_horizontal_fct_source(g, c, 10800.0),
                               rtol=1e-9, atol=1e-12 * np.abs(L @ c).max())
