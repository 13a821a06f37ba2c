"""GPU parity of the periodic-orbit solve against the CPU oracle (splu + scipy GMRES).

Bars: the LU shim reproduces splu to 1e-12; the FP64 periodic solve matches
the oracle's c0 to 1e-8 relative (both driven to tol 1e-11/1e-12); the mixed
FP32/BF16 solve with FP64 refinement matches to 1e-6 relative per tracer at
tol 1e-8 (the bar of the production workload).  All sizes here take seconds
on the oracle.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import orbit_ref as orb
from orbit_amd import synth
from orbit_amd.core import OperatorSet

N_BINS = 8


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.fixture(scope="module")
def small():
    wl = synth.make_workload(12, 10, 6, 3, seed=5, convection=True)
    op = OperatorSet.from_workload(wl)
    lus = [[orb.factor_one_bin(wl.L(t, s)) for t in range(N_BINS)] for s in range(wl.n_tracers)]
    return wl, op, lus


@pytest.fixture(scope="module")
def single24():
    wl = synth.benchmark_workload("single")
    op = OperatorSet.from_workload(wl)
    lus = [orb.factor_one_bin(wl.L(t, 0)) for t in range(N_BINS)]
    return wl, op, lus


def test_apply_P_matches_splu_both_transposes(small):
    wl, op, lus = small
    rng = np.random.default_rng(0)
    for tau, s in ((0, 0), (5, 2)):
        b = rng.standard_normal(wl.N)
        np.testing.assert_allclose(op.apply_P(tau, s, b), lus[s][tau].solve(b), rtol=0,
                                   atol=1e-12 * np.abs(lus[s][tau].solve(b)).max())
        np.testing.assert_allclose(op.apply_P(tau, s, b, trans=True), lus[s][tau].solve(b, trans="T"),
                                   rtol=0, atol=1e-12 * np.abs(lus[s][tau].solve(b, trans="T")).max())


def test_apply_P_single_tracer_nonnegative_for_nonnegative_rhs(single24):
    """A_tau is an M-matrix: a non-negative source gives a non-negative state
    (reference tests/test_solve.py:19-52)."""
    wl, op, lus = single24
    b = np.zeros(wl.N)
    b[: wl.nx * wl.ny] = np.random.default_rng(1).random(wl.nx * wl.ny)
    x = op.apply_P(3, 0, b)
    assert x.min() >= -1e-15 * x.max()
    np.testing.assert_allclose(x, lus[3].solve(b), rtol=0, atol=1e-12 * np.abs(x).max())


def test_forward_fp64_matches_oracle(small):
    wl, op, lus = small
    orbit, info = op.solve_forward(wl.e, tol=1e-11, precision="fp64")
    for s in range(wl.n_tracers):
        ref, rinfo = orb.solve_orbit_one_species(lus[s], list(wl.e[s]), tol=1e-12, maxiter=500)
        assert info[s]["converged"] and info[s]["residual"] <= 1e-11
        for k in range(N_BINS + 1):
            assert _rel(orbit[s, k], ref[k]) < 1e-8, (s, k)
        assert info[s]["periodicity"] < 1e-9


def test_forward_mixed_matches_oracle(small):
    wl, op, lus = small
    orbit, info = op.solve_forward(wl.e, tol=1e-8, precision="mixed")
    for s in range(wl.n_tracers):
        ref, _ = orb.solve_orbit_one_species(lus[s], list(wl.e[s]), tol=1e-12, maxiter=500)
        assert info[s]["converged"] and info[s]["residual"] <= 1e-8
        assert _rel(orbit[s, 0], ref[0]) < 1e-6, s


def test_forward_single_tracer_both_precisions(single24):
    wl, op, lus = single24
    ref, _ = orb.solve_orbit_one_species(lus, list(wl.e[0]), tol=1e-12, maxiter=500)
    for prec, tol, bar in (("fp64", 1e-10, 1e-8), ("mixed", 1e-8, 1e-6)):
        orbit, info = op.solve_forward(wl.e, tol=tol, precision=prec)
        assert info[0]["converged"], (prec, info)
        assert _rel(orbit[0, 0], ref[0]) < bar, prec
        assert _rel(orbit[0, 5], ref[5]) < bar, prec


def test_warm_start_and_zero_source(small):
    wl, op, lus = small
    e = wl.e.copy()
    e[1] = 0.0
    orbit, info = op.solve_forward(e, tol=1e-10)
    assert info[1]["zero_source"] == 1 and not np.any(orbit[1])
    ref, _ = orb.solve_orbit_one_species(lus[0], list(e[0]), tol=1e-12, maxiter=500)
    warm = np.stack([ref[0] * 1.01, np.zeros(wl.N), ref[0]])
    orbit2, info2 = op.solve_forward(e, c_warm=warm, tol=1e-10)
    assert _rel(orbit2[0, 0], ref[0]) < 1e-8
    assert info2[0]["gmres_iters"] <= info[0]["gmres_iters"]


def test_adjoint_zero_receptor_gives_zero_orbit(small):
    """Reference tests/test_orbit_adjoint.py:87-99 and :102-120: an empty
    receptor gives an identically zero adjoint orbit (here beside a non-zero
    slot of the same batch, in both precisions), and a non-zero one is
    periodic."""
    wl, op, _ = small
    R = np.random.default_rng(4).random((wl.n_tracers, N_BINS, wl.N))
    R[1] = 0.0
    for prec, tol in (("fp64", 1e-10), ("mixed", 1e-8)):
        lam, info = op.solve_adjoint(R, tol=tol, precision=prec)
        assert info[1]["zero_source"] == 1 and not np.any(lam[1]), prec
        assert info[0]["converged"] and np.any(lam[0]), prec
        assert info[0]["periodicity"] < 1e-8, prec
        assert _rel(lam[0, N_BINS], lam[0, 0]) < 1e-8, prec


def test_not_converged_is_a_warning_not_an_error(small):
    wl, op, _ = small
    orbit, info = op.solve_forward(wl.e, tol=1e-14, maxiter=1, restart=2)
    assert not info[0]["converged"] and info[0]["gmres_iters"] == 2
    assert np.all(np.isfinite(orbit))


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
def test_non_finite_input_stops_its_slot_unconverged(small, prec):
    """A NaN in one tracer's emissions makes its residual non-finite: that slot
    stops at once and reports converged = 0 (the call returns the warning
    status, which the wrapper accepts, not an error), and the other slots of
    the batch are the same bits as without it (every decision is per slot)."""
    import torch
    wl, op, _ = small
    e = torch.from_numpy(wl.e).cuda()
    ref, rinfo = op.solve_forward(e, tol=1e-8, precision=prec)
    bad = e.clone()
    bad[1, 3, 17] = float("nan")
    out, info = op.solve_forward(bad, tol=1e-8, precision=prec)
    assert not info[1]["converged"] and not np.isfinite(info[1]["residual"])
    for s in (0, 2):
        assert info[s]["converged"] and info[s] == rinfo[s]
        assert torch.equal(out[s], ref[s])


def test_adjoint_matches_oracle_and_dot_identity(small):
    wl, op, lus = small
    rng = np.random.default_rng(3)
    R = rng.random((wl.n_tracers, N_BINS, wl.N))
    for prec, tol, bar in (("fp64", 1e-11, 1e-8), ("mixed", 1e-8, 1e-6)):
        lam, info = op.solve_adjoint(R, tol=tol, precision=prec)
        for s in range(wl.n_tracers):
            ref, _ = orb.solve_orbit_adjoint_one_species(lus[s], list(R[s]), tol=1e-12, maxiter=500)
            assert info[s]["converged"], (prec, s)
            for k in (0, 1, 4, 7, 8):
                assert _rel(lam[s, k], ref[k]) < bar, (prec, s, k)
    # <r, (I-M)^-1 s> = <(I-M^T)^-1 r, s> with R_0 = r, R_{t>0} = 0
    r = rng.random(wl.N)
    Rr = np.zeros((1, N_BINS, wl.N))
    Rr[0, 0] = r
    lam, _ = op.solve_adjoint(Rr, tracers=[2], tol=1e-12)
    orbit, _ = op.solve_forward(wl.e[2:3], tracers=[2], tol=1e-12)
    s = orb.compute_s(lus[2], list(wl.e[2]))
    assert np.dot(r, orbit[0, 0]) == pytest.approx(np.dot(lam[0, 0], s), rel=1e-9)


def test_alone_and_in_a_batch_give_identical_results(small):
    wl, op, _ = small
    both, _ = op.solve_forward(wl.e, tol=1e-8, precision="mixed")
    alone, _ = op.solve_forward(wl.e[1:2], tracers=[1], tol=1e-8, precision="mixed")
    assert np.array_equal(both[1], alone[0])
    rev, _ = op.solve_forward(wl.e[::-1].copy(), tracers=[2, 1, 0], tol=1e-8, precision="mixed")
    assert np.array_equal(both[0], rev[2])
    # five slots take the four-slots-per-column sweep mapping: same bits again
    five, _ = op.solve_forward(np.concatenate([wl.e, wl.e[:2]]), tracers=[0, 1, 2, 0, 1], tol=1e-8,
                               precision="mixed")
    for k, t in enumerate((0, 1, 2, 0, 1)):
        assert np.array_equal(five[k], both[t]), k
    lam3, _ = op.solve_adjoint(wl.e, tol=1e-8, precision="mixed")
    lam5, _ = op.solve_adjoint(np.concatenate([wl.e, wl.e[:2]]), tracers=[0, 1, 2, 0, 1], tol=1e-8,
                               precision="mixed")
    assert np.array_equal(lam5[4], lam3[1])


def test_lu_protocol_call_surface(single24):
    from orbit_amd import solve as hsolve, orbit as horbit, orbit_adjoint as hadj
    wl, _, lus = single24
    hl = hsolve.factor_species([wl.L(t, 0) for t in range(N_BINS)], wl.nz, wl.ny, wl.nx)
    assert hl[0].shape == (wl.N, wl.N)
    b = np.random.default_rng(2).random(wl.N)
    np.testing.assert_allclose(horbit.apply_monodromy(b, hl), orb.apply_monodromy(b, lus), rtol=1e-10)
    np.testing.assert_allclose(horbit.compute_s(hl, list(wl.e[0])), orb.compute_s(lus, list(wl.e[0])),
                               rtol=1e-10)
    orbit, info = horbit.solve_orbit_one_species(hl, list(wl.e[0]), tol=1e-10)
    ref, _ = orb.solve_orbit_one_species(lus, list(wl.e[0]), tol=1e-12, maxiter=500)
    assert _rel(orbit[0], ref[0]) < 1e-8 and set(info) >= {"gmres_iters", "residual", "periodicity", "neg_frac"}
    R = [np.random.default_rng(t).random(wl.N) for t in range(N_BINS)]
    adj, ainfo = hadj.solve_orbit_adjoint_one_species(hl, R, tol=1e-10)
    aref, _ = orb.solve_orbit_adjoint_one_species(lus, R, tol=1e-12, maxiter=500)
    g = hadj.gradient_per_emission_bin(adj, R)
    gr = orb.gradient_per_emission_bin(aref, R)
    assert max(_rel(a, b) for a, b in zip(g, gr)) < 1e-8
    with pytest.raises(ValueError):
        hl[0].solve(b, trans="H")


def test_structure_error_for_non_grid_pattern():
    import scipy.sparse as sp
    from orbit_amd._lib import OrbitError
    wl = synth.make_workload(5, 4, 3, 1, seed=0)
    L = wl.L(0, 0).tolil()
    L[0, wl.N - 1] = -1e-6     # couples two distant cells
    with pytest.raises(OrbitError):
        OperatorSet.from_csr([[L.tocsr()]] * N_BINS, wl.nz, wl.ny, wl.nx)


def test_mass_balance_matches_oracle():
    from oracle import mass_ref
    wl = synth.make_workload(9, 8, 5, 1, seed=4, convection=True)
    op = OperatorSet.from_workload(wl)
    got = op.mass_balance(wl.rho)
    ref = mass_ref.mass_balance_diagnostics(wl.T, wl.rho, wl.nz, wl.ny, wl.nx)
    for k in ("interior_median", "interior_p90", "interior_p99", "interior_max", "top_median", "top_p90"):
        assert got[k] == pytest.approx(ref[k], rel=1e-9, abs=1e-12), k
    np.testing.assert_allclose(got["layer_p90"], ref["layer_p90"], rtol=1e-9, atol=1e-12)
    assert got["interior_max"] < 1e-9


def test_shared_context_many_opsets_stay_finite():
    """Operator sets of different sizes share one context's scratch: a solve must
    never read scratch it did not write (stale values, even NaN, are harmless)."""
    import torch
    outs = []
    for month, size, nt in ((0, "24x20", 3), (1, "48x40", 3), (0, "24x20", 4), (2, "48x40", 2)):
        wl = synth.benchmark_workload("uneven-batch", n_tracers=nt, month=month, size=size)
        op = OperatorSet.from_workload(wl)
        assert not any(op.check().values())
        e = torch.from_numpy(wl.e).cuda()
        for prec in ("mixed", "fp64"):
            orbit, info = op.solve_forward(e, tol=1e-8, precision=prec)
            assert torch.isfinite(orbit).all(), (size, nt, prec)
            assert all(i["converged"] and np.isfinite(i["s_norm"]) for i in info)
            outs.append(orbit)
        lam, ainfo = op.solve_adjoint(e, tol=1e-8, precision="mixed")
        assert torch.isfinite(lam).all() and all(i["converged"] for i in ainfo)


def test_consecutive_and_permuted_tracer_lists_replay_the_right_graphs(small):
    """A batch of consecutive tracers derives each slot's tracer arithmetically
    and its captured sweeps carry that base; a permuted list of the same size on
    the same context reads the slot map instead.  Each tracer's orbit must be
    the same bits either way, in either order of the calls, and the same as
    direct launches without graphs."""
    wl, op, _ = small
    g = dict(tol=1e-8, precision="mixed", use_graph=True)
    a, _ = op.solve_forward(wl.e, tracers=[0, 1, 2], **g)
    perm = [2, 0, 1]
    b, _ = op.solve_forward(wl.e[perm], tracers=perm, **g)
    c, _ = op.solve_forward(wl.e, tracers=[0, 1, 2], **g)
    for k, t in enumerate(perm):
        assert np.array_equal(b[k], a[t]), t
    assert np.array_equal(c, a)
    d, _ = op.solve_forward(wl.e, tracers=[0, 1, 2], tol=1e-8, precision="mixed", use_graph=False)
    assert np.array_equal(d, a)  # direct launches: the same bits
    la, _ = op.solve_adjoint(wl.e, tracers=[0, 1, 2], **g)
    lb, _ = op.solve_adjoint(wl.e[perm], tracers=perm, **g)
    for k, t in enumerate(perm):
        assert np.array_equal(lb[k], la[t]), t
