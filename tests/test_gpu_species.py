"""GPU parity of the species layer (orbit_amd/species.py) against the oracle
(oracle/species_ref.py): coupled DAG waves with start- and end-of-bin K blocks,
extra RHS, copy-verbatim skips and zero-RHS species; the cyclic KeyError and the
Picard fixed point (closed form of the reference's NH/NO3 test); the adjoint
species passes, against the oracle and against the device forward layer."""
import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle import orbit_ref as orb
from oracle import species_ref as spc
from orbit_amd import species, synth
from orbit_amd.core import OperatorSet

N_BINS = 8
DTAU = 10800.0


def _rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.fixture(scope="module")
def chain():
    """5 species on a 9x7x6 grid: 0 -> 1 (start of bin), 1 -> 2 (end of bin),
    3 skipped with a cached orbit, 4 with zero emissions and no sources."""
    wl = synth.make_workload(9, 7, 6, 5, seed=41, convection=True)
    N = wl.N
    rng = np.random.default_rng(11)
    K = []
    for t in range(N_BINS):
        d = rng.uniform(0.5, 1.0, N) * 1e-5
        off = sp.random(N, N, density=3.0 / N, random_state=int(rng.integers(1 << 30))) * 2e-6
        K.append((sp.diags(d) + off).tocsr())
    Ks = [{(1, 0): K[t], (2, 1): 0.7 * K[t]} for t in range(N_BINS)]
    op = OperatorSet.from_workload(wl)
    lus = {s: [orb.factor_one_bin(wl.L(t, s)) for t in range(N_BINS)] for s in range(5)}
    e = np.zeros((N_BINS, 5 * N))
    e[:, : 4 * N] = rng.random((N_BINS, 4 * N)) * (rng.random((N_BINS, 4 * N)) < 0.4)
    cached = {3: [rng.random(N) for _ in range(N_BINS + 1)]}
    extra = {2: [rng.random(N) * 1e-3 for _ in range(N_BINS)]}
    return wl, Ks, op, lus, e, cached, extra


def test_coupled_waves_match_the_oracle(chain):
    wl, Ks, op, lus, e, cached, extra = chain
    N = wl.N
    kw = dict(tol=1e-11, maxiter=300, skip_solve_species={3}, cached_orbits=cached,
              extra_rhs_per_bin_per_species=extra, endbin_keys={(2, 1)}, solve_order=(0, 1, 2, 3, 4))
    ref = spc.solve_all_species(lus, Ks, e, N, n_species=5, **kw)
    got = species.solve_orbit_all_species(None, Ks, e, N, opset=op, verbose=False, **kw)
    for s in range(5):
        for j in (0, 3, 8):
            assert _rel(got["orbits"][s][j], ref["orbits"][s][j]) < 1e-8 if np.any(ref["orbits"][s][j]) \
                else not np.any(got["orbits"][s][j]), (s, j)
    assert got["gmres_iters"][3] == 0 and got["gmres_iters"][4] == 0
    assert np.array_equal(got["orbits"][3][5], cached[3][5])
    assert all(got["gmres_resid"][s] <= 1e-11 for s in (0, 1, 2))
    # mixed precision through the same layer
    mixed = species.solve_orbit_all_species(None, Ks, e, N, opset=op, verbose=False, precision="mixed",
                                            **{**kw, "tol": 1e-8})
    for s in (0, 1, 2):
        assert _rel(mixed["orbits"][s][0], ref["orbits"][s][0]) < 1e-6, s


def test_cyclic_pair_needs_seeds_and_picard_matches_the_oracle(chain):
    wl, _, _, _, _, _, _ = chain
    N = wl.N
    rng = np.random.default_rng(12)
    L = [[(wl.L(t, s) + sp.identity(N) / DTAU).tocsr() for s in range(3)] for t in range(N_BINS)]
    K = [(sp.diags(rng.uniform(0.5, 1.0, N) * 5e-5)).tocsr() for _ in range(N_BINS)]
    Ks = [{(0, 1): K[t], (1, 0): K[t], (0, 2): 0.2 * K[t]} for t in range(N_BINS)]
    e = rng.random(3 * N)
    op = OperatorSet.from_csr(L, wl.nz, wl.ny, wl.nx)
    with pytest.raises(KeyError, match="Seed-orbits"):
        species.solve_orbit_all_species(None, Ks, e, N, opset=op, verbose=False, solve_order=(0, 1, 2))
    lus = {s: [orb.factor_one_bin(L[t][s]) for t in range(N_BINS)] for s in range(3)}
    ref, its_r, _, ok_r = spc.picard_cyclic(lus, Ks, e, N, (0, 1), max_iters=40, rel_tol=1e-11, tol=1e-12,
                                            maxiter=300, solve_order=(0, 1, 2), n_species=3)
    got, its, change, ok = species.picard_cyclic(None, Ks, e, N, (0, 1), max_iters=40, rel_tol=1e-11, opset=op,
                                                 tol=1e-12, maxiter=300, verbose=False, solve_order=(0, 1, 2))
    assert ok and ok_r and abs(its - its_r) <= 1
    for s in (0, 1, 2):
        assert _rel(got["orbits"][s][0], ref["orbits"][s][0]) < 1e-8, s


def test_reference_minimal_cyclic_system_reaches_the_closed_form():
    N, l, k = 8, 1.0 / DTAU, 0.1 / DTAU
    NH, NO3 = spc.IDX_TOTAL_NH, spc.IDX_TOTAL_NO3
    L = [[sp.eye(N, format="csr") * l for _ in range(14)] for _ in range(N_BINS)]
    Ks = [{(NH, NO3): sp.eye(N, format="csr") * k, (NO3, NH): sp.eye(N, format="csr") * k} for _ in range(N_BINS)]
    e = np.zeros(14 * N)
    e[NH * N:(NH + 1) * N] = l
    skip = {s for s in range(14) if s not in (NH, NO3)}
    cached = {s: [np.zeros(N) for _ in range(N_BINS + 1)] for s in skip}
    res, its, change, ok = species.picard_cyclic(L, Ks, e, N, (NH, NO3), max_iters=30, rel_tol=1e-12, tol=1e-12,
                                                 maxiter=300, verbose=False, skip_solve_species=skip,
                                                 cached_orbits=cached)
    assert ok
    den = l * l - k * k
    np.testing.assert_allclose(res["orbits"][NH][0], l * l / den, rtol=1e-9)
    np.testing.assert_allclose(res["orbits"][NO3][0], -k * l / den, rtol=1e-9)


def test_adjoint_species_passes_match_the_oracle_and_the_forward_layer(chain):
    wl, Ks, op, lus, _, _, _ = chain
    N = wl.N
    rng = np.random.default_rng(13)
    R = {s: rng.random((N_BINS, N)) * (rng.random((N_BINS, N)) < 0.2) for s in range(3)}
    lam, Reff = species.adjoint_species_passes(op, R, Ks, user_species={0}, endbin_keys={(2, 1)}, tol=1e-12,
                                               maxiter=300, solve_order=(0, 1, 2, 3, 4))
    lr, Rr = spc.adjoint_species_passes(lus, {s: list(R[s]) for s in R}, Ks, user_species={0},
                                        endbin_keys={(2, 1)}, tol=1e-12, maxiter=300,
                                        solve_order=(0, 1, 2, 3, 4))
    assert set(lam) == set(lr) == {0, 1, 2}
    for s in (0, 1, 2):
        assert _rel(lam[s][0].cpu().numpy(), lr[s][0]) < 1e-8, s
        assert _rel(Reff[s].cpu().numpy(), np.stack(Rr[s])) < 1e-8, s
    grad = {s: (DTAU * (lam[s][:N_BINS] - Reff[s])).cpu().numpy() for s in (0, 1, 2)}
    e = np.zeros((N_BINS, 5 * N))
    e[:, : 3 * N] = rng.standard_normal((N_BINS, 3 * N))
    fwd = species.solve_orbit_all_species(None, Ks, e, N, opset=op, tol=1e-12, maxiter=300, verbose=False,
                                          endbin_keys={(2, 1)}, solve_order=(0, 1, 2, 3, 4))
    J = sum(float(np.dot(R[s][t], fwd["orbits"][s][t])) for s in R for t in range(N_BINS))
    lin = sum(float(np.dot(grad[s][t], e[t, s * N:(s + 1) * N])) for s in (0, 1, 2) for t in range(N_BINS))
    assert J == pytest.approx(lin, rel=1e-8)


def test_adjoint_species_passes_with_a_cyclic_pair_match_the_oracle(chain):
    wl = chain[0]
    N = wl.N
    rng = np.random.default_rng(14)
    L = [[(wl.L(t, s) + sp.identity(N) / DTAU).tocsr() for s in range(3)] for t in range(N_BINS)]
    K = [(sp.diags(rng.uniform(0.5, 1.0, N) * 5e-5)).tocsr() for _ in range(N_BINS)]
    Ks = [{(0, 1): K[t], (1, 0): K[t], (0, 2): K[t] * 0.2, (1, 2): K[t] * 0.1} for t in range(N_BINS)]
    op = OperatorSet.from_csr(L, wl.nz, wl.ny, wl.nx)
    lus = {s: [orb.factor_one_bin(L[t][s]) for t in range(N_BINS)] for s in range(3)}
    R = {s: rng.random((N_BINS, N)) for s in range(3)}
    kw = dict(user_species={2}, tol=1e-12, maxiter=300, picard_max_iters=40, picard_rel_tol=1e-12,
              cyclic_species=(0, 1), solve_order=(0, 1, 2))
    lam, Reff = species.adjoint_species_passes(op, R, Ks, **kw)
    lr, Rr = spc.adjoint_species_passes(lus, {s: list(R[s]) for s in R}, Ks, **kw)
    for s in (0, 1, 2):
        assert _rel(lam[s][0].cpu().numpy(), lr[s][0]) < 1e-8, s
        assert _rel(lam[s][5].cpu().numpy(), lr[s][5]) < 1e-8, s
