"""GPU parity of the ISORROPIA cross-coupling blocks (orbit_amd/iso_coupling.py,
kernel k_iso_cross) against the oracle (oracle/iso_coupling_ref.py): one bin and a
month, with and without the HNO3 field, zero layer thicknesses, disabled blocks;
then the blocks merged into the species layer (summed into a chemistry block,
end-of-bin keys) and the cyclic TotalNH <-> TotalNO3 pair solved by Picard on the
device against the oracle's Picard."""
from types import SimpleNamespace

import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle import iso_coupling_ref as iso
from oracle import orbit_ref as orb
from oracle import species_ref as spc
from orbit_amd import iso_coupling, species, synth
from orbit_amd.core import OperatorSet

N_BINS = 8
NH, PSO4, NO3, SO2 = iso.IDX_TOTAL_NH, iso.IDX_PSO4, iso.IDX_TOTAL_NO3, 3
PARTS = ("f_nh_dno3", "f_no3_dnh", "f_nh_dso4", "f_no3_dso4")


def _grid(rng, nz, ny, nx, hno3=True, zero_dz=False):
    n = (nz, ny, nx)
    g = SimpleNamespace(nz=nz, ny=ny, nx=nx, Dz=rng.uniform(20.0, 200.0, n),
                        particle_wet_dep=rng.uniform(0.0, 3e-4, n), other_gas_wet_dep=rng.uniform(0.0, 1e-4, n),
                        particle_dry_dep=rng.uniform(0.001, 0.004, n), NH3_dry_dep=rng.uniform(0.002, 0.012, n))
    if hno3:
        g.HNO3_dry_dep = rng.uniform(0.01, 0.05, n)
    if zero_dz:
        g.Dz[0, ::2] = 0.0
    return g


def _partials(rng, shape, scale=0.3):
    return {k: [rng.uniform(-scale, scale, shape) for _ in range(N_BINS)] for k in PARTS}


@pytest.mark.parametrize("hno3,zero_dz", [(True, False), (False, False), (True, True)])
def test_one_bin_matches_the_oracle(hno3, zero_dz):
    rng = np.random.default_rng(5)
    g = _grid(rng, 5, 11, 13, hno3=hno3, zero_dz=zero_dz)
    f = {k: rng.uniform(-1, 1, (5, 11, 13)) for k in PARTS}
    got = iso_coupling.assemble_iso_cross_blocks(g, None, **f)
    ref = iso.assemble_iso_cross_blocks(g, disabled=set(), **f)
    assert set(got) == set(ref)
    for key in ref:
        assert isinstance(got[key], sp.csc_matrix) and got[key].shape == ref[key].shape
        np.testing.assert_allclose(got[key].diagonal(), ref[key].diagonal(), rtol=1e-15, atol=0, err_msg=str(key))


def test_month_with_disabled_blocks_matches_the_oracle(monkeypatch):
    rng = np.random.default_rng(6)
    shape = (4, 9, 10)
    grids = [_grid(rng, *shape) for _ in range(N_BINS)]
    parts = _partials(rng, shape)
    monkeypatch.setenv("ORBIT_ISO_DISABLE_BLOCKS", "NO3_SO4")
    N = int(np.prod(shape))
    base = sp.diags(rng.uniform(1e-6, 2e-6, N), 0, format="csc")
    Ks_got = [{(NO3, NH): base, (PSO4, SO2): base} for _ in range(N_BINS)]
    Ks_ref = [dict(k) for k in Ks_got]
    kg, kr = set(), set()
    iso_coupling.add_iso_blocks(Ks_got, grids, parts, kg)
    iso.add_iso_blocks(Ks_ref, grids, parts, kr)
    assert kg == kr == {(NH, NO3), (NH, PSO4), (NO3, NH), (NO3, PSO4)}
    for t in range(N_BINS):
        assert set(Ks_got[t]) == set(Ks_ref[t])
        for key in Ks_ref[t]:
            np.testing.assert_allclose(Ks_got[t][key].diagonal(), Ks_ref[t][key].diagonal(), rtol=1e-15, atol=0,
                                       err_msg=f"bin {t} block {key}")
        assert not Ks_got[t][(NO3, PSO4)].diagonal().any()


def test_device_diagonals_take_strided_bins():
    """(n_bins, N) views of a wider buffer: the kernel honours the bin stride."""
    rng = np.random.default_rng(7)
    shape = (3, 6, 7)
    N = int(np.prod(shape))
    dev = torch.device("cuda", 0)
    grids = [_grid(rng, *shape) for _ in range(N_BINS)]
    parts = _partials(rng, shape)
    names = ("Dz", "particle_dry_dep", "NH3_dry_dep", "HNO3_dry_dep", "particle_wet_dep", "other_gas_wet_dep")
    wide = {k: torch.zeros((N_BINS, N + 64), dtype=torch.float64, device=dev) for k in names + PARTS}
    for k in names:
        wide[k][:, :N] = torch.from_numpy(np.stack([getattr(g, k).ravel() for g in grids])).to(dev)
    for k in PARTS:
        wide[k][:, :N] = torch.from_numpy(np.stack([p.ravel() for p in parts[k]])).to(dev)
    views = {k: wide[k][:, :N] for k in wide}
    out = iso_coupling.iso_cross_diagonals_dev({k: views[k] for k in names}, [views[k] for k in PARTS], shape,
                                               disabled=set())
    for t in range(N_BINS):
        ref = iso.assemble_iso_cross_blocks(grids[t], *(parts[k][t] for k in PARTS), disabled=set())
        for q, (tg, sc, _) in enumerate(iso_coupling.BLOCKS):
            np.testing.assert_allclose(out[q, t].cpu().numpy(), ref[(tg, sc)].diagonal(), rtol=1e-15, atol=0)


def test_iso_coupled_species_solve_matches_the_oracle():
    """7 species on a 9x7x6 grid: SO2 -> pSO4 (start of bin), the iso blocks
    NO3 <-> NH and pSO4 -> both (end of bin), the NH/NO3 cycle by Picard."""
    wl = synth.make_workload(9, 7, 6, 7, seed=43, convection=True)
    N = wl.N
    shape = (wl.nz, wl.ny, wl.nx)
    rng = np.random.default_rng(17)
    grids = [_grid(rng, *shape) for _ in range(N_BINS)]
    parts = _partials(rng, shape, scale=0.2)
    Ks = [{(PSO4, SO2): sp.diags(rng.uniform(0.5, 1.0, N) * 1e-5, 0, format="csr")} for _ in range(N_BINS)]
    Kr = [dict(k) for k in Ks]
    kg, kr = set(), set()
    iso_coupling.add_iso_blocks(Ks, grids, parts, kg, )
    iso.add_iso_blocks(Kr, grids, parts, kr, disabled=set())
    assert kg == kr
    e = np.zeros((N_BINS, 7 * N))
    for s in (NH, SO2, NO3):
        e[:, s * N:(s + 1) * N] = rng.random((N_BINS, N))
    skip = {0, 1, 4}
    cached = {s: [np.zeros(N) for _ in range(N_BINS + 1)] for s in skip}
    kw = dict(tol=1e-12, maxiter=300, skip_solve_species=skip, cached_orbits=cached, endbin_keys=kg,
              solve_order=tuple(range(7)))
    lus = {s: [orb.factor_one_bin(wl.L(t, s)) for t in range(N_BINS)] for s in range(7)}
    ref, its_r, _, ok_r = spc.picard_cyclic(lus, Kr, e, N, (NH, NO3), max_iters=40, rel_tol=1e-11, n_species=7,
                                            **{**kw, "endbin_keys": kr})
    op = OperatorSet.from_workload(wl)
    got, its, _, ok = species.picard_cyclic(None, Ks, e, N, (NH, NO3), max_iters=40, rel_tol=1e-11, opset=op,
                                            verbose=False, **kw)
    assert ok and ok_r and abs(its - its_r) <= 1
    for s in (NH, SO2, PSO4, NO3):
        for j in (0, 4, 8):
            r = ref["orbits"][s][j]
            err = np.abs(got["orbits"][s][j] - r).max() / np.abs(r).max()
            assert err < 1e-8, (s, j, err)
    # the cross coupling is felt: NO3 differs from its solve without iso blocks
    plain = species.solve_orbit_all_species(None, [{(PSO4, SO2): k[(PSO4, SO2)]} for k in Kr], e, N, opset=op,
                                            verbose=False, **{**kw, "endbin_keys": set()})
    d = np.abs(plain["orbits"][NO3][0] - got["orbits"][NO3][0]).max() / np.abs(got["orbits"][NO3][0]).max()
    assert d > 1e-4
