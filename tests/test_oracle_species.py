"""Pins of the oracle's species layer (oracle/species_ref.py).

The first group restates the reference's own tests of this layer
(tests/test_simulation_modes.py:1158-1580); the second solves the joint
periodic system of coupled species directly; the third checks the adjoint
species passes against the forward layer through J(e) = <grad J, e>.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import orbit_ref as orb
from oracle import species_ref as spc
from orbit_amd import synth

N_BINS = 8
DTAU = 10800.0
NH, NO3 = spc.IDX_TOTAL_NH, spc.IDX_TOTAL_NO3


# -- wave placement (reference tests/test_simulation_modes.py:1523-1580) -----------

def _wave_of(waves):
    return {s: w for w, ss in enumerate(waves) for s in ss}


def test_waves_target_after_higher_index_source():
    w = _wave_of(spc.build_solve_waves({5: [3], 2: [5]}, list(range(9)), set()))
    assert (w[3], w[5], w[2]) == (0, 1, 2)


def test_waves_cyclic_edges_removed_for_placement():
    w = _wave_of(spc.build_solve_waves({5: [3], 2: [5, 6], 6: [5, 2, 7]}, list(range(9)), set()))
    assert (w[5], w[2], w[6]) == (1, 2, 2)


def test_waves_skip_species_in_wave_zero():
    w = _wave_of(spc.build_solve_waves({2: [5], 5: [3]}, list(range(6)), {5}))
    assert (w[5], w[2]) == (0, 1)


def test_reference_dag_waves():
    # forward chemistry DAG of the reference: pSO4 <- SO2, TotalNO3 <- NOx, VBS aging chain
    srcs = {5: [3], 6: [4], 0: [12], 9: [0], 10: [9], 11: [10]}
    waves = spc.build_solve_waves(srcs, list(spc.REF_SOLVE_ORDER), set())
    w = _wave_of(waves)
    assert w[12] == 0 and w[0] == 1 and w[9] == 2 and w[10] == 3 and w[11] == 4
    assert w[5] == 1 and w[6] == 1 and w[3] == 0 and w[13] == 0
    assert sorted(s for ws in waves for s in ws) == list(range(14))


# -- the minimal cyclic NH <-> NO3 system (reference :1158-1370) --------------------

def _minimal(N, l=1.0 / DTAU, k=0.1 / DTAU, e_nh=1.0 / DTAU, e_no3=0.0):
    L = [[sp.eye(N, format="csc") * l for _ in range(spc.N_SPECIES)] for _ in range(N_BINS)]
    K = [{(NH, NO3): sp.eye(N, format="csc") * k, (NO3, NH): sp.eye(N, format="csc") * k}
         for _ in range(N_BINS)]
    e = np.zeros(spc.N_SPECIES * N)
    e[NH * N:(NH + 1) * N] = e_nh
    e[NO3 * N:(NO3 + 1) * N] = e_no3
    skip = {s for s in range(spc.N_SPECIES) if s not in (NH, NO3)}
    cached = {s: [np.zeros(N) for _ in range(N_BINS + 1)] for s in skip}
    lus = {s: [orb.factor_one_bin(L[t][s]) for t in range(N_BINS)] for s in (NH, NO3)}
    return lus, K, e, skip, cached


def test_cyclic_without_seed_raises():
    lus, K, e, skip, cached = _minimal(8)
    with pytest.raises(KeyError, match="Seed-orbits"):
        spc.solve_all_species(lus, K, e, 8, tol=1e-8, skip_solve_species=skip, cached_orbits=cached)


def test_seeded_single_pass_is_the_one_way_answer():
    lus, K, e, skip, cached = _minimal(1)
    seed = {s: [np.zeros(1) for _ in range(N_BINS + 1)] for s in (NH, NO3)}
    res = spc.solve_all_species(lus, K, e, 1, tol=1e-10, skip_solve_species=skip, cached_orbits=cached,
                                seed_orbits=seed)
    assert res["orbits"][NH][0][0] == pytest.approx(1.0, abs=1e-8)
    # NO3 saw the NH seed (zero), so it is the uncoupled zero-emission orbit
    assert abs(res["orbits"][NO3][0][0]) < 1e-12


def test_picard_converges_to_the_closed_form_fixed_point():
    l, k = 1.0 / DTAU, 0.1 / DTAU
    lus, K, e, skip, cached = _minimal(8, l=l, k=k, e_nh=l, e_no3=0.0)
    res, its, change, ok = spc.picard_cyclic(lus, K, e, 8, (NH, NO3), max_iters=30, rel_tol=1e-12,
                                             tol=1e-12, skip_solve_species=skip, cached_orbits=cached)
    assert ok and its < 30
    den = l * l - k * k
    np.testing.assert_allclose(res["orbits"][NH][0], (l * l) / den, rtol=1e-9)
    np.testing.assert_allclose(res["orbits"][NO3][0], -(k * l) / den, rtol=1e-9)


def test_endbin_keys_change_the_target_and_not_the_source():
    N, l, k = 1, 1.0 / DTAU, 0.05 / DTAU
    L = [[sp.eye(N, format="csc") * l for _ in range(spc.N_SPECIES)] for _ in range(N_BINS)]
    K = [{(NH, NO3): sp.eye(N, format="csc") * k} for _ in range(N_BINS)]
    e = np.zeros((N_BINS, spc.N_SPECIES * N))
    e[:, NO3] = np.array([3.0, 1.0, 0.2, 0.0, 0.0, 0.5, 1.5, 2.5]) * l
    skip = {s for s in range(spc.N_SPECIES) if s not in (NH, NO3)}
    cached = {s: [np.zeros(N) for _ in range(N_BINS + 1)] for s in skip}
    lus = {s: [orb.factor_one_bin(L[t][s]) for t in range(N_BINS)] for s in (NH, NO3)}
    kw = dict(tol=1e-12, maxiter=300, skip_solve_species=skip, cached_orbits=cached)
    a = spc.solve_all_species(lus, K, e, N, **kw)
    b = spc.solve_all_species(lus, K, e, N, endbin_keys={(NH, NO3)}, **kw)
    c = spc.solve_all_species(lus, K, e, N, endbin_keys=set(), **kw)
    no3a, no3b = np.array(a["orbits"][NO3]), np.array(b["orbits"][NO3])
    np.testing.assert_allclose(no3a, no3b, rtol=1e-9, atol=1e-12)
    nha, nhb = np.array(a["orbits"][NH]), np.array(b["orbits"][NH])
    assert np.abs(nha - nhb).max() > 1e-6 * np.abs(nha).max()
    for s in (NH, NO3):
        for t in range(N_BINS + 1):
            assert np.array_equal(a["orbits"][s][t], c["orbits"][s][t])


# -- joint periodic system of a coupled pair, solved directly ------------------------

@pytest.fixture(scope="module")
def pair():
    wl = synth.make_workload(5, 4, 3, 3, seed=31, convection=True)
    N = wl.N
    rng = np.random.default_rng(7)
    K = []
    for t in range(N_BINS):
        d = rng.uniform(0.5, 1.0, N) * 1e-5
        off = sp.random(N, N, density=3.0 / N, random_state=int(rng.integers(1 << 30))) * 2e-6
        K.append((sp.diags(d) + off).tocsr())
    lus = {s: [orb.factor_one_bin(wl.L(t, s)) for t in range(N_BINS)] for s in range(3)}
    return wl, K, lus


def _joint_solve(wl, K, e, endbin):
    """Unknowns c^s_1..c^s_8 of species 0 (source) and 1 (target); c_0 = c_8."""
    N = wl.N

    def col(s, j):
        return s * N_BINS + (N_BINS - 1 if j % N_BINS == 0 else j - 1)

    blocks = [[None] * (2 * N_BINS) for _ in range(2 * N_BINS)]
    for s in (0, 1):
        for t in range(N_BINS):
            r = s * N_BINS + t
            blocks[r][col(s, t + 1)] = wl.A(t, s).tocsr()
            blocks[r][col(s, t)] = -sp.identity(N, format="csr")
    for t in range(N_BINS):
        j = t + 1 if endbin else t
        blocks[N_BINS + t][col(0, j)] = DTAU * K[t]
    M = sp.bmat(blocks, format="csc")
    rhs = DTAU * np.concatenate([e[t, s * N:(s + 1) * N] for s in (0, 1) for t in range(N_BINS)])
    x = spla.splu(M).solve(rhs)
    return {s: [x[col(s, j) * N:(col(s, j) + 1) * N] for j in range(N_BINS + 1)] for s in (0, 1)}


@pytest.mark.parametrize("endbin", [False, True])
def test_coupled_pair_matches_the_joint_periodic_system(pair, endbin):
    wl, K, lus = pair
    N = wl.N
    rng = np.random.default_rng(3)
    e = rng.random((N_BINS, 2 * N)) * (rng.random((N_BINS, 2 * N)) < 0.4)
    res = spc.solve_all_species({s: lus[s] for s in (0, 1)}, [{(1, 0): K[t]} for t in range(N_BINS)], e, N,
                                tol=1e-13, maxiter=500, endbin_keys={(1, 0)} if endbin else None,
                                solve_order=(0, 1), n_species=2)
    ref = _joint_solve(wl, K, e, endbin)
    for s in (0, 1):
        for j in (0, 1, 4, 8):
            a, b = res["orbits"][s][j], ref[s][j]
            assert np.abs(a - b).max() <= 1e-9 * np.abs(b).max(), (s, j)


# -- adjoint species passes against the forward layer ------------------------------

def _J(orbits, R):
    return sum(float(np.dot(R[s][t], orbits[s][t])) for s in R for t in range(N_BINS))


def test_adjoint_passes_reproduce_the_forward_chain(pair):
    wl, K, lus = pair
    N = wl.N
    Ks = [{(1, 0): K[t], (2, 1): 0.7 * K[t]} for t in range(N_BINS)]
    endbin = {(2, 1)}
    rng = np.random.default_rng(4)
    R = {s: [rng.random(N) * (rng.random(N) < 0.2) for _ in range(N_BINS)] for s in range(3)}
    lam, Reff = spc.adjoint_species_passes(lus, R, Ks, user_species={0}, endbin_keys=endbin, tol=1e-13,
                                           maxiter=500, solve_order=(0, 1, 2))
    assert set(lam) == {0, 1, 2}
    grad = {s: [DTAU * (lam[s][t] - Reff[s][t]) for t in range(N_BINS)] for s in range(3)}
    for trial in range(2):
        e = rng.standard_normal((N_BINS, 3 * N))
        fwd = spc.solve_all_species(lus, Ks, e, N, tol=1e-13, maxiter=500, endbin_keys=endbin,
                                    solve_order=(0, 1, 2), n_species=3)
        lin = sum(float(np.dot(grad[s][t], e[t, s * N:(s + 1) * N])) for s in range(3) for t in range(N_BINS))
        assert _J(fwd["orbits"], R) == pytest.approx(lin, rel=1e-9), trial


def test_adjoint_passes_with_a_cyclic_pair(pair):
    wl, K, _ = pair
    N = wl.N
    # strong loss on the cyclic pair keeps the Picard iteration contractive
    lus = {s: [orb.factor_one_bin(wl.L(t, s) + sp.identity(N) / DTAU) for t in range(N_BINS)] for s in range(3)}
    Ks = [{(0, 1): K[t] * 5, (1, 0): K[t] * 5, (0, 2): K[t], (1, 2): 0.5 * K[t]} for t in range(N_BINS)]
    rng = np.random.default_rng(5)
    R = {s: [rng.random(N) for _ in range(N_BINS)] for s in range(3)}
    lam, Reff = spc.adjoint_species_passes(lus, R, Ks, user_species={2}, tol=1e-13, maxiter=500,
                                           picard_max_iters=80, picard_rel_tol=1e-13, cyclic_species=(0, 1),
                                           solve_order=(0, 1, 2))
    grad = {s: [DTAU * (lam[s][t] - Reff[s][t]) for t in range(N_BINS)] for s in range(3)}
    e = rng.standard_normal((N_BINS, 3 * N))
    res, its, change, ok = spc.picard_cyclic(lus, Ks, e, N, (0, 1), max_iters=80, rel_tol=1e-13, tol=1e-13,
                                             maxiter=500, solve_order=(0, 1, 2), n_species=3)
    assert ok
    lin = sum(float(np.dot(grad[s][t], e[t, s * N:(s + 1) * N])) for s in range(3) for t in range(N_BINS))
    assert _J(res["orbits"], R) == pytest.approx(lin, rel=1e-8)
