"""CPU oracle of the ISORROPIA lookup-table query -- TEST INFRASTRUCTURE ONLY.

Restates IsorropiaLUT.query (reference orbit/core/isorropia_lut.py:123-147):
a 7-D scipy RegularGridInterpolator per field with method "linear",
bounds_error=False, fill_value=None (linear extrapolation outside the axes),
then f_nh4 and f_no3 clipped to [0, 1] and aerosol water to >= 0, NaN for a
field that was not loaded.  The interpolation itself is scipy's (the
reference's dependency; scipy 1.15.3 here).

Pinned by tests/test_oracle_lut.py: a multilinear function is reproduced
exactly inside and outside the table, grid nodes return the table, and the
clips hold.
"""
from __future__ import annotations

import numpy as np
from scipy.interpolate import RegularGridInterpolator

AXES = ("so4_axis", "nh_axis", "no3_axis", "ca_axis", "na_axis", "t_axis", "rh_axis")


def interpolate(axes, table, pts):
    """scipy RegularGridInterpolator(linear, no bounds error, extrapolating)."""
    return RegularGridInterpolator(tuple(axes), table, bounds_error=False, fill_value=None)(pts)


def query(arrays, fields, total_so4, total_nh, total_no3, crustal_ca, sea_salt_na, T, RH):
    """(f_nh4, f_no3, aerosol_water, ph) as the reference returns them."""
    axes = [arrays[a] for a in AXES]
    pts = np.column_stack([np.atleast_1d(v).ravel() for v in
                           (total_so4, total_nh, total_no3, crustal_ca, sea_salt_na, T, RH)])
    f_nh4 = np.clip(interpolate(axes, arrays["f_nh4"], pts), 0.0, 1.0)
    f_no3 = np.clip(interpolate(axes, arrays["f_no3"], pts), 0.0, 1.0)
    aw = (np.maximum(interpolate(axes, arrays["aerosol_water"], pts), 0.0) if "aerosol_water" in fields
          else np.full_like(f_nh4, np.nan))
    ph = interpolate(axes, arrays["ph"], pts) if "ph" in fields else np.full_like(f_nh4, np.nan)
    return f_nh4, f_no3, aw, ph


def synthetic_table(seed=0, sizes=(4, 5, 4, 3, 3, 4, 5)):
    """A seeded 7-D table with the reference's key names (float32 fields)."""
    rng = np.random.default_rng(seed)
    ranges = [(0.0, 40.0), (0.0, 30.0), (0.0, 30.0), (0.0, 2.0), (0.0, 2.0), (260.0, 310.0), (0.1, 0.99)]
    out = {}
    for name, n, (a, b) in zip(AXES, sizes, ranges):
        x = np.sort(rng.uniform(a, b, n))
        x[0], x[-1] = a, b
        out[name] = x
    shape = tuple(sizes)
    out["f_nh4"] = rng.uniform(-0.1, 1.1, shape).astype(np.float32)
    out["f_no3"] = rng.uniform(-0.1, 1.1, shape).astype(np.float32)
    out["aerosol_water"] = rng.uniform(-5.0, 100.0, shape).astype(np.float32)
    out["ph"] = rng.uniform(-1.0, 7.0, shape).astype(np.float32)
    return out
