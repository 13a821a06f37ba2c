"""CPU oracle of the ISORROPIA partitioning update -- TEST INFRASTRUCTURE ONLY.

Restates partitioning_per_bin and update_grid_partitioning (reference
orbit/core/isorropia_closure.py:75-497) over the oracle's table query
(oracle/lut_ref.py, scipy's RegularGridInterpolator as the reference uses):
per bin and cell, from the end-of-bin states of pSO4 (5), TotalNH (2) and
TotalNO3 (6) clipped at 0, the met fields (T, RH in percent / 100, crustal
Ca = 0.036 * dust_fine, Na = 0.306 * sea_salt_fine, zero when a field is
absent):

  f_nh_eq, f_no3_eq     the table at the state;
  f_nh_marg, f_no3_marg d(f c)/dc by the mean of the +/-1 % one-sided
                        differences (c+ = c (1 + d) + 1e-12, c- = max(c (1 - d), 0);
                        a side with no step gives f itself), clipped to [0, 1];
  cross partials        c_self * (f_marg_self at other + delta - at other - delta)
                        / (other+ - other-), f_marg_self by the same clipped
                        symmetric rule on the shifted slice (isorropia_closure.py:211-417);

then the under-relaxation alpha * new + (1 - alpha) * prev for 0 < alpha < 1.

Pinned by tests/test_oracle_isorropia.py: golden vectors written by the
reference module itself (tests/golden/make_partitioning_golden.py), a constant
table (marginal = equilibrium, no cross terms), a table linear in total NH
(d(f NH)/dNH = a + 2 b NH in closed form), and the blend.
"""
from __future__ import annotations

import numpy as np

from . import lut_ref

IDX_TOTAL_NH, IDX_PSO4, IDX_TOTAL_NO3 = 2, 5, 6
DUST_TO_CA = 0.036      # isorropia_closure.py:40
SEASALT_TO_NA = 0.306   # isorropia_closure.py:42
DELTA = 0.01            # isorropia_closure.py:48
FIELDS = ("f_nh_eq", "f_nh_marg", "f_no3_eq", "f_no3_marg", "f_nh_dno3", "f_no3_dnh", "f_nh_dso4", "f_no3_dso4")


def _q(table, so4, nh, no3, ca, na, T, RH):
    f_nh, f_no3, _, _ = lut_ref.query(table, ("f_nh4", "f_no3"), so4, nh, no3, ca, na, T, RH)
    return f_nh, f_no3


def _marg(f_c, f_p, f_m, c, c_p, c_m):
    """Clipped mean of the one-sided differences of f*c (isorropia_closure.py:163-176)."""
    base = f_c * c
    dp, dm = c_p - c, c - c_m
    up = np.where(dp > 0, (f_p * c_p - base) / np.maximum(dp, 1e-30), f_c)
    dn = np.where(dm > 0, (base - f_m * c_m) / np.maximum(dm, 1e-30), f_c)
    return np.clip(0.5 * (up + dn), 0.0, 1.0)


def partitioning_bin(table, so4, nh, no3, T, RH_percent, dust=None, seasalt=None, cross=True):
    """One bin's fields, flat (N,) each; states already clipped at 0 by the caller."""
    N = nh.size
    ca = dust * DUST_TO_CA if dust is not None and dust.size == N else np.zeros(N)
    na = seasalt * SEASALT_TO_NA if seasalt is not None and seasalt.size == N else np.zeros(N)
    RH = RH_percent / 100.0
    d = DELTA
    met = (ca, na, T, RH)
    f_nh, f_no3 = _q(table, so4, nh, no3, *met)
    nh_p, nh_m = nh * (1.0 + d) + 1e-12, np.maximum(nh * (1.0 - d), 0.0)
    no3_p, no3_m = no3 * (1.0 + d) + 1e-12, np.maximum(no3 * (1.0 - d), 0.0)
    nh_at_nhp, no3_at_nhp = _q(table, so4, nh_p, no3, *met)
    nh_at_nhm, no3_at_nhm = _q(table, so4, nh_m, no3, *met)
    nh_at_no3p, no3_at_no3p = _q(table, so4, nh, no3_p, *met)
    nh_at_no3m, no3_at_no3m = _q(table, so4, nh, no3_m, *met)
    out = {"f_nh_eq": f_nh, "f_nh_marg": _marg(f_nh, nh_at_nhp, nh_at_nhm, nh, nh_p, nh_m),
           "f_no3_eq": f_no3, "f_no3_marg": _marg(f_no3, no3_at_no3p, no3_at_no3m, no3, no3_p, no3_m)}
    if not cross:
        return out
    so4_p, so4_m = so4 * (1.0 + d) + 1e-12, np.maximum(so4 * (1.0 - d), 0.0)
    # NH-NO3 plane corners
    nh_pp, no3_pp = _q(table, so4, nh_p, no3_p, *met)
    nh_pm, no3_pm = _q(table, so4, nh_p, no3_m, *met)
    nh_mp, no3_mp = _q(table, so4, nh_m, no3_p, *met)
    nh_mm, no3_mm = _q(table, so4, nh_m, no3_m, *met)
    # f_marg of NH on the NO3 +/- slices; of NO3 on the NH +/- slices
    mnh_no3p = _marg(nh_at_no3p, nh_pp, nh_mp, nh, nh_p, nh_m)
    mnh_no3m = _marg(nh_at_no3m, nh_pm, nh_mm, nh, nh_p, nh_m)
    out["f_nh_dno3"] = nh * ((mnh_no3p - mnh_no3m) / np.maximum(no3_p - no3_m, 1e-30))
    mno3_nhp = _marg(no3_at_nhp, no3_pp, no3_pm, no3, no3_p, no3_m)
    mno3_nhm = _marg(no3_at_nhm, no3_mp, no3_mm, no3, no3_p, no3_m)
    out["f_no3_dnh"] = no3 * ((mno3_nhp - mno3_nhm) / np.maximum(nh_p - nh_m, 1e-30))
    # SO4 slices: corners in the NH-SO4 and NO3-SO4 planes and the centres
    nh_pp_s, _ = _q(table, so4_p, nh_p, no3, *met)
    nh_pm_s, _ = _q(table, so4_m, nh_p, no3, *met)
    nh_mp_s, _ = _q(table, so4_p, nh_m, no3, *met)
    nh_mm_s, _ = _q(table, so4_m, nh_m, no3, *met)
    _, no3_pp_s = _q(table, so4_p, nh, no3_p, *met)
    _, no3_pm_s = _q(table, so4_m, nh, no3_p, *met)
    _, no3_mp_s = _q(table, so4_p, nh, no3_m, *met)
    _, no3_mm_s = _q(table, so4_m, nh, no3_m, *met)
    nh_c_sp, no3_c_sp = _q(table, so4_p, nh, no3, *met)
    nh_c_sm, no3_c_sm = _q(table, so4_m, nh, no3, *met)
    dso4 = np.maximum(so4_p - so4_m, 1e-30)
    mnh_sp = _marg(nh_c_sp, nh_pp_s, nh_mp_s, nh, nh_p, nh_m)
    mnh_sm = _marg(nh_c_sm, nh_pm_s, nh_mm_s, nh, nh_p, nh_m)
    out["f_nh_dso4"] = nh * ((mnh_sp - mnh_sm) / dso4)
    mno3_sp = _marg(no3_c_sp, no3_pp_s, no3_mp_s, no3, no3_p, no3_m)
    mno3_sm = _marg(no3_c_sm, no3_pm_s, no3_mm_s, no3, no3_p, no3_m)
    out["f_no3_dso4"] = no3 * ((mno3_sp - mno3_sm) / dso4)
    return out


def partitioning_per_bin(grids, orbits, table, include_cross_partials=True):
    """{bin: {field: (nz, ny, nx)}} from orbits {species: [c_0, ..., c_8]}."""
    out = {}
    for tau, g in enumerate(grids):
        shape = (g.nz, g.ny, g.nx)
        st = lambda s: np.maximum(np.asarray(orbits[s][tau + 1], dtype=np.float64), 0.0).ravel()  # noqa: E731
        f = partitioning_bin(table, st(IDX_PSO4), st(IDX_TOTAL_NH), st(IDX_TOTAL_NO3), g.Temperature.ravel(),
                             g.RH.ravel(), getattr(g, "dust_fine", np.array([])).ravel(),
                             getattr(g, "sea_salt_fine", np.array([])).ravel(), include_cross_partials)
        out[tau] = {k: v.reshape(shape) for k, v in f.items()}
    return out


def blend(new_per_bin, prev_per_bin, alpha):
    """Under-relaxation of update_grid_partitioning (isorropia_closure.py:470-478)."""
    out = {}
    for tau, new in new_per_bin.items():
        if prev_per_bin and tau in prev_per_bin and 0.0 < alpha < 1.0:
            out[tau] = {k: alpha * new[k] + (1.0 - alpha) * prev_per_bin[tau][k] for k in new}
        else:
            out[tau] = new
    return out
