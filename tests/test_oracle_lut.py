"""Pins of the LUT-query oracle (oracle/lut_ref.py): multilinear functions are
interpolated exactly, in and outside the table; nodes give the table; clips."""
import itertools

import numpy as np

from oracle import lut_ref


def _multilinear(axes_pts, coef):
    """f(x) = sum over subsets S of coef[S] * prod_{i in S} x_i."""
    d = axes_pts.shape[-1]
    out = np.zeros(axes_pts.shape[:-1])
    for k, S in enumerate(itertools.product((0, 1), repeat=d)):
        term = coef[k] * np.ones(axes_pts.shape[:-1])
        for i, s in enumerate(S):
            if s:
                term = term * axes_pts[..., i]
        out += term
    return out


def test_multilinear_functions_are_reproduced_inside_and_outside():
    rng = np.random.default_rng(0)
    t = lut_ref.synthetic_table(1, sizes=(3, 4, 3, 2, 2, 3, 4))
    axes = [t[a] for a in lut_ref.AXES]
    coef = rng.normal(0.0, 1.0, 2 ** 7) * 0.01
    grid = np.stack(np.meshgrid(*axes, indexing="ij"), axis=-1)
    table = _multilinear(grid, coef)
    lo = np.array([a[0] for a in axes]); hi = np.array([a[-1] for a in axes])
    pts = lo + (hi - lo) * rng.uniform(-0.3, 1.3, (400, 7))          # a third outside
    got = lut_ref.interpolate(axes, table, pts)
    np.testing.assert_allclose(got, _multilinear(pts, coef), rtol=1e-9, atol=1e-9 * np.abs(table).max())
    # a non-multilinear function is not reproduced (the check can fail)
    bad = lut_ref.interpolate(axes, table + grid[..., 0] ** 2 * 1e-3, pts)
    assert np.abs(bad - _multilinear(pts, coef)).max() > 1e-4


def test_nodes_return_the_table_and_outputs_are_clipped():
    t = lut_ref.synthetic_table(2)
    axes = [t[a] for a in lut_ref.AXES]
    idx = np.array([[1, 2, 0, 1, 2, 3, 4], [0, 0, 0, 0, 0, 0, 0], [3, 4, 3, 2, 2, 3, 4]])
    pts = np.stack([np.array([axes[d][i[d]] for d in range(7)]) for i in idx])
    v = lut_ref.interpolate(axes, t["ph"], pts)
    np.testing.assert_allclose(v, [t["ph"][tuple(i)] for i in idx], rtol=1e-6)
    f_nh4, f_no3, aw, ph = lut_ref.query(t, ("f_nh4", "f_no3", "aerosol_water", "ph"), *pts.T)
    assert (f_nh4 >= 0).all() and (f_nh4 <= 1).all() and (f_no3 >= 0).all() and (aw >= 0).all()
    f_nh4, f_no3, aw, ph = lut_ref.query(t, ("f_nh4", "f_no3"), *pts.T)
    assert np.isnan(aw).all() and np.isnan(ph).all()
