"""GPU parity of the device LUT query against the oracle (scipy's
RegularGridInterpolator as the reference calls it): in-range, extrapolated,
node and NaN points, the reference's clips, other dimensions, the archive
loader, and an exactly reproduced multilinear function."""
import ctypes as C
import itertools

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from oracle import lut_ref
from orbit_amd import _lib
from orbit_amd.core import default_context
from orbit_amd.lut import ALL_FIELDS, IsorropiaLUT


def _points(t, n, seed):
    rng = np.random.default_rng(seed)
    axes = [t[a] for a in lut_ref.AXES]
    lo = np.array([a[0] for a in axes]); hi = np.array([a[-1] for a in axes])
    p = lo + (hi - lo) * rng.uniform(-0.25, 1.25, (n, 7))
    p[:50] = [[axes[d][rng.integers(len(axes[d]))] for d in range(7)] for _ in range(50)]   # nodes
    p[50, 2] = np.nan
    return p


def test_device_query_matches_the_oracle_on_all_fields():
    t = lut_ref.synthetic_table(5)
    lut = IsorropiaLUT(arrays=t, fields=ALL_FIELDS)
    p = _points(t, 20000, 1)
    got = lut.query(*[torch.from_numpy(p[:, d].copy()).cuda() for d in range(7)])
    ref = lut_ref.query(t, ALL_FIELDS, *p.T)
    for g, r, name in zip(got, ref, ALL_FIELDS):
        g = g.cpu().numpy()
        assert np.array_equal(np.isnan(g), np.isnan(r)), name
        ok = ~np.isnan(r)
        np.testing.assert_allclose(g[ok], r[ok], rtol=1e-12, atol=1e-12, err_msg=name)
    assert np.isnan(got[0][50].item())
    f = got[0].cpu().numpy()
    assert (f[~np.isnan(f)] >= 0).all() and (f[~np.isnan(f)] <= 1).all()


def test_missing_fields_are_nan_and_archive_loads(tmp_path):
    t = lut_ref.synthetic_table(6)
    path = str(tmp_path / "lut.npz")
    np.savez(path, **t)
    lut = IsorropiaLUT(path)                       # f_nh4, f_no3 only (the reference default)
    p = _points(t, 500, 2)
    f_nh4, f_no3, aw, ph = lut.query(*p.T)          # host inputs -> numpy outputs
    r = lut_ref.query(t, ("f_nh4", "f_no3"), *p.T)
    np.testing.assert_allclose(f_no3[51:], r[1][51:], rtol=1e-12, atol=1e-12)
    assert np.isnan(aw).all() and np.isnan(ph).all()
    lut.release()
    assert not lut.loaded
    lut.query(*p[:3].T)
    assert lut.n_loads == 2


def test_other_dimensions_through_the_c_abi_and_multilinear_exactness():
    from scipy.interpolate import RegularGridInterpolator
    rng = np.random.default_rng(7)
    ctx = default_context()
    for nd in (1, 3, 5):
        axes = [np.sort(rng.uniform(0, 10, rng.integers(2, 6))) for _ in range(nd)]
        coef = rng.normal(0, 1, 2 ** nd)

        def f(x):
            out = np.zeros(x.shape[0])
            for k, S in enumerate(itertools.product((0, 1), repeat=nd)):
                out += coef[k] * np.prod(np.where(np.array(S, bool), x, 1.0), axis=1)
            return out
        grid = np.stack(np.meshgrid(*axes, indexing="ij"), -1).reshape(-1, nd)
        tab = f(grid).reshape([len(a) for a in axes]).astype(np.float32)
        pts = rng.uniform(-3, 13, (3000, nd))
        xs = [torch.from_numpy(pts[:, d].copy()).cuda() for d in range(nd)]
        out = torch.empty(3000, dtype=torch.float64, device="cuda")
        h = C.c_void_p()
        _lib.check(_lib.lib().orbit_lut_create(
            ctx.handle, nd, (C.c_int * nd)(*[len(a) for a in axes]),
            (C.c_void_p * nd)(*[a.ctypes.data for a in axes]), 1,
            (C.c_void_p * 1)(tab.ctypes.data), C.byref(h)))
        torch.cuda.synchronize()
        _lib.check(_lib.lib().orbit_lut_query_dev(ctx.handle, h, 3000, (C.c_void_p * nd)(*[x.data_ptr() for x in xs]),
                                                  None, None, (C.c_void_p * 1)(out.data_ptr())))
        ctx.sync()
        _lib.lib().orbit_lut_destroy(h)
        ref = RegularGridInterpolator(axes, tab, bounds_error=False, fill_value=None)(pts)
        np.testing.assert_allclose(out.cpu().numpy(), ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
        # inside the table (weights in [0, 1], summing to 1) the FP32 table holds
        # the multilinear function to float32 rounding of the table
        inside = np.all([(pts[:, d] >= axes[d][0]) & (pts[:, d] <= axes[d][-1]) for d in range(nd)], axis=0)
        np.testing.assert_allclose(out.cpu().numpy()[inside], f(pts)[inside], rtol=0,
                                   atol=4e-7 * np.abs(tab).max())


def test_bad_axes_are_refused():
    t = lut_ref.synthetic_table(8)
    t["t_axis"] = t["t_axis"][::-1].copy()
    with pytest.raises(_lib.OrbitError):
        IsorropiaLUT(arrays=t).query(*np.ones((7, 2)))
