"""The ISORROPIA partitioning update of the oracle (oracle/isorropia_ref.py,
restating reference orbit/core/isorropia_closure.py:75-497), pinned by
golden vectors the reference module wrote itself
(tests/golden/partitioning_golden.npz, tests/golden/make_partitioning_golden.py)
and by closed forms: a constant table makes the marginal fraction the
equilibrium one and every cross partial zero; a table linear in total NH,
f = a + b NH (multilinear interpolation reproduces it), has
d(f NH)/dNH = a + 2 b NH exactly up to the 1e-12 offset of the + step."""
import os
import types

import numpy as np
import pytest

from oracle import isorropia_ref as iso, lut_ref

GOLD = os.path.join(os.path.dirname(__file__), "golden", "partitioning_golden.npz")


def golden_case():
    z = np.load(GOLD, allow_pickle=False)
    table = {k[len("table_"):]: z[k] for k in z.files if k.startswith("table_")}
    grids, orbits = [], {2: [None] * 9, 5: [None] * 9, 6: [None] * 9}
    for tau in range(2):
        T = z[f"T_{tau}"]
        g = types.SimpleNamespace(nz=T.shape[0], ny=T.shape[1], nx=T.shape[2], Temperature=T, RH=z[f"RH_{tau}"],
                                  dust_fine=z[f"dust_{tau}"], sea_salt_fine=z[f"seasalt_{tau}"])
        grids.append(g)
        for s in (2, 5, 6):
            orbits[s][tau + 1] = z[f"state_{s}_{tau}"]
    return z, table, grids, orbits


def test_oracle_matches_the_reference_golden_vectors():
    z, table, grids, orbits = golden_case()
    new = iso.partitioning_per_bin(grids, orbits, table, include_cross_partials=True)
    prev = {tau: {k: np.full_like(v, 0.25 + 0.1 * tau) for k, v in d.items()} for tau, d in new.items()}
    blended = iso.blend(new, prev, 0.5)
    for tau in range(2):
        for k in iso.FIELDS:
            ref = z[f"new_{k}_{tau}"]
            np.testing.assert_allclose(new[tau][k], ref, rtol=1e-12, atol=1e-12 * max(1.0, np.abs(ref).max()),
                                       err_msg=f"{k} bin {tau}")
            np.testing.assert_allclose(blended[tau][k], z[f"blend_{k}_{tau}"], rtol=1e-12, atol=1e-12)
        np.testing.assert_array_equal(z[f"grid_NHPartitioning_{tau}"], z[f"blend_f_nh_marg_{tau}"])
        np.testing.assert_array_equal(z[f"grid_NO3PartitioningEq_{tau}"], z[f"blend_f_no3_eq_{tau}"])
    # the fixture covers the branches: zero states, clipped marginals, extrapolation
    st = np.concatenate([z[f"state_2_{t}"] for t in range(2)])
    assert (st == 0).any() and (st < 0).any()
    m = np.concatenate([z[f"new_f_nh_marg_{t}"].ravel() for t in range(2)])
    assert ((m == 0) | (m == 1)).any() and ((m > 0) & (m < 1)).any()


def _grid(rng, n=(2, 3, 4)):
    return types.SimpleNamespace(nz=n[0], ny=n[1], nx=n[2], Temperature=rng.uniform(270, 300, n),
                                 RH=rng.uniform(30, 90, n), dust_fine=np.array([]), sea_salt_fine=np.array([]))


def test_constant_table_closed_form():
    t = lut_ref.synthetic_table(3)
    t["f_nh4"][:] = 0.3
    t["f_no3"][:] = 0.6
    rng = np.random.default_rng(1)
    g = _grid(rng)
    N = 24
    orbits = {s: [None, rng.uniform(0.5, 20, N)] + [None] * 7 for s in (2, 5, 6)}
    out = iso.partitioning_per_bin([g], orbits, t)[0]
    # the table keeps its FP32 values
    np.testing.assert_allclose(out["f_nh_marg"], float(np.float32(0.3)), rtol=1e-9)
    np.testing.assert_allclose(out["f_no3_marg"], float(np.float32(0.6)), rtol=1e-9)
    np.testing.assert_allclose(out["f_nh_eq"], float(np.float32(0.3)), rtol=1e-12)
    for k in ("f_nh_dno3", "f_no3_dnh", "f_nh_dso4", "f_no3_dso4"):
        assert np.abs(out[k]).max() < 1e-6, k


def test_linear_in_nh_closed_form():
    t = lut_ref.synthetic_table(4)
    a, b = 0.05, 0.012
    nh_axis = t["nh_axis"]
    shape = t["f_nh4"].shape
    t["f_nh4"] = np.broadcast_to((a + b * nh_axis)[None, :, None, None, None, None, None], shape).astype(np.float64)
    rng = np.random.default_rng(2)
    g = _grid(rng)
    N = 24
    nh = rng.uniform(1.0, 25.0, N)  # a + b nh stays inside [0, 1]; a + 2 b nh too
    orbits = {2: [None, nh] + [None] * 7, 5: [None, rng.uniform(1, 30, N)] + [None] * 7,
              6: [None, rng.uniform(1, 20, N)] + [None] * 7}
    out = iso.partitioning_per_bin([g], orbits, t, include_cross_partials=False)[0]
    np.testing.assert_allclose(out["f_nh_eq"].ravel(), a + b * nh, rtol=1e-12)
    np.testing.assert_allclose(out["f_nh_marg"].ravel(), np.clip(a + 2 * b * nh, 0, 1), rtol=1e-9)
    # a wrong factor (the non-derivative f, or a one-sided step) misses it
    assert np.abs(out["f_nh_marg"].ravel() - (a + b * nh)).max() > 1e-2


def test_blend_needs_alpha_in_the_open_interval():
    new = {0: {"x": np.ones(3)}}
    prev = {0: {"x": np.zeros(3)}}
    assert np.array_equal(iso.blend(new, prev, 0.25)[0]["x"], np.full(3, 0.25))
    assert iso.blend(new, prev, 1.0)[0]["x"] is new[0]["x"]
    assert iso.blend(new, None, 0.5)[0]["x"] is new[0]["x"]
