"""GPU parity of the FCT deferred-correction drivers and the autograd functions.

Oracle: oracle/drivers_ref.py (SuperLU + scipy GMRES restatement of
orbit/cli.py:1212-1241, orbit/modes/marginal.py:627-730,
orbit/modes/adjoint.py:814-962) and oracle/orbit_ref.py, on a 12x10x6 grid with
dense convective columns.  Outer GMRES tolerances are tightened in these tests
so both sides converge far below the comparison bar.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle import drivers_ref, fct_ref, orbit_ref as orb
from orbit_amd import autograd as hag, drivers, synth
from orbit_amd.core import OperatorSet

N_BINS = 8


def _rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.fixture(scope="module")
def setup():
    wl = synth.make_workload(12, 10, 6, 2, seed=21, convection=True, wind_std=8.0)
    op = OperatorSet.from_workload(wl)
    grids = drivers.make_fct_grids(wl.grids, op.ctx)
    lus = [[orb.factor_one_bin(wl.L(t, s)) for t in range(N_BINS)] for s in range(wl.n_tracers)]
    # a front in the emissions so the limiter is active
    e = wl.e.copy()
    e[:, :, : wl.nx * wl.ny] *= 1.0 + 4.0 * (np.arange(wl.nx * wl.ny) % wl.nx > wl.nx // 2)
    return wl, op, grids, lus, e


def test_forward_fct_passes_match_oracle(setup):
    wl, op, grids, lus, e = setup
    ed = torch.from_numpy(e).cuda()
    orbit, hist = drivers.forward_fct(op, grids, ed, n_passes=3, tol=1e-11, precision="fp64")
    for s in range(wl.n_tracers):
        ref, _ = drivers_ref.forward_fct_passes(lus[s], list(e[s]), wl.grids, n_passes=3, tol=1e-12)
        o = orbit[s].cpu().numpy()
        for k in (0, 3, 8):
            assert _rel(o[k], ref[k]) < 1e-8, (s, k)
    orbit_m, _ = drivers.forward_fct(op, grids, ed, n_passes=3, tol=1e-8, precision="mixed")
    assert _rel(orbit_m[:, 0].cpu().numpy(), orbit[:, 0].cpu().numpy()) < 1e-6


def _oracle_LAD(wl, orbit_s):
    return [fct_ref.assemble_fct_linear_operator(wl.grids[t], orbit_s[t + 1], wl.dtau) for t in range(N_BINS)]


def test_marginal_fct_matches_oracle(setup):
    wl, op, grids, lus, e = setup
    ed = torch.from_numpy(e).cuda()
    base, _ = drivers.forward_fct(op, grids, ed, n_passes=2, tol=1e-11, precision="fp64")
    L = drivers.frozen_operators(op, grids, base[:1])
    de = 0.01 * e[:1] * np.random.default_rng(0).lognormal(0.0, 1.0, e[:1].shape)
    dc, info = drivers.marginal_fct(op, L, torch.from_numpy(de).cuda(), tracers=[0], tol=1e-12,
                                    max_iters=60, rel_tol=1e-9)
    Lref = _oracle_LAD(wl, base[0].cpu().numpy())
    # the device frozen operator equals the oracle's
    for t in (0, 5):
        assert abs(L[t][0].tocsr() - Lref[t]).max() <= 1e-12 * abs(Lref[t]).max()
    ref, rinfo = drivers_ref.marginal_fct_gmres(lus[0], list(de[0]), Lref, tol=1e-12, max_iters=60, rel_tol=1e-9)
    assert _rel(dc[0].cpu().numpy(), ref) < 1e-7


def test_adjoint_fct_matches_oracle_and_fd_of_marginal(setup):
    wl, op, grids, lus, e = setup
    ed = torch.from_numpy(e).cuda()
    base, _ = drivers.forward_fct(op, grids, ed, n_passes=2, tol=1e-11, precision="fp64")
    L = drivers.frozen_operators(op, grids, base[:1])
    rng = np.random.default_rng(1)
    R = np.zeros((1, N_BINS, wl.N))
    R[0, 0, rng.choice(wl.nx * wl.ny, 10, replace=False)] = rng.random(10)
    lam, Re, info = drivers.adjoint_fct(op, L, torch.from_numpy(R).cuda(), tracers=[0], tol=1e-12,
                                        max_iters=60, rel_tol=1e-9)
    grad = wl.dtau * (lam[0, :N_BINS] - Re[0]).cpu().numpy()
    Lref = _oracle_LAD(wl, base[0].cpu().numpy())
    lr, Rr, _ = drivers_ref.adjoint_fct_gmres(lus[0], list(R[0]), Lref, tol=1e-12, max_iters=60, rel_tol=1e-9)
    gr = np.stack([wl.dtau * (lr[t] - Rr[t]) for t in range(N_BINS)])
    assert _rel(grad, gr) < 1e-7
    # J(de) = <R_0, dc_0> is linear in de: FD of the marginal driver equals the gradient
    cells = rng.choice(wl.N, 4, replace=False)
    for tau, j in zip((0, 3, 7, 2), cells):
        de = np.zeros((1, N_BINS, wl.N))
        de[0, tau, j] = 1.0
        dc, _ = drivers.marginal_fct(op, L, torch.from_numpy(de).cuda(), tracers=[0], tol=1e-12,
                                     max_iters=60, rel_tol=1e-10)
        fd = float(np.dot(R[0, 0], dc[0, 0].cpu().numpy()))
        assert fd == pytest.approx(grad[tau, j], rel=1e-6, abs=1e-12 * np.abs(grad).max())


def test_orbit_solve_autograd_matches_fd_and_oracle(setup):
    wl, op, grids, lus, e = setup
    et = torch.from_numpy(e).cuda().requires_grad_(True)
    orbit = hag.orbit_solve(et, op, tol=1e-12)
    rng = np.random.default_rng(2)
    r = torch.zeros_like(orbit)
    cells = rng.choice(wl.nx * wl.ny, 20, replace=False)
    r[:, 0, cells] = torch.from_numpy(rng.random((wl.n_tracers, 20))).cuda()
    J = (r * orbit).sum()
    J.backward()
    g = et.grad.cpu().numpy()
    # oracle adjoint with R_0 = r
    for s in range(wl.n_tracers):
        R = [r[s, 0].cpu().numpy()] + [np.zeros(wl.N)] * (N_BINS - 1)
        adj, _ = orb.solve_orbit_adjoint_one_species(lus[s], R, tol=1e-12, maxiter=500)
        gref = np.stack(orb.gradient_per_emission_bin(adj, R))
        assert _rel(g[s], gref) < 1e-8
    # central finite differences (eps = 1e-4 relative) on emission cells
    for s, tau, j in ((0, 0, int(cells[0])), (1, 4, int(cells[1])), (0, 7, 5)):
        h = 1e-4 * max(abs(e[s, tau, j]), np.abs(e[s]).max())
        ep, em = e.copy(), e.copy()
        ep[s, tau, j] += h
        em[s, tau, j] -= h
        op_, _ = op.solve_forward(ep, tol=1e-12, maxiter=20)
        om_, _ = op.solve_forward(em, tol=1e-12, maxiter=20)
        rn = r.cpu().numpy()
        fd = (np.sum(rn * op_) - np.sum(rn * om_)) / (2 * h)
        assert fd == pytest.approx(g[s, tau, j], rel=1e-5, abs=1e-9 * np.abs(g).max())


def test_periodic_solve_gradient_is_lambda0(setup):
    wl, op, grids, lus, e = setup
    s_ref = orb.compute_s(lus[0], list(e[0]))
    st = torch.from_numpy(np.stack([s_ref, 2.0 * s_ref])).cuda().requires_grad_(True)
    c0 = hag.periodic_solve(st, op, tracers=[0, 0], tol=1e-12)
    orbit_ref, _ = orb.solve_orbit_one_species(lus[0], list(e[0]), tol=1e-12, maxiter=500)
    assert _rel(c0[0].detach().cpu().numpy(), orbit_ref[0]) < 1e-9
    rng = np.random.default_rng(3)
    r = np.zeros((2, wl.N))
    r[:, rng.choice(wl.N, 50, replace=False)] = rng.random((2, 50))
    (torch.from_numpy(r).cuda() * c0).sum().backward()
    for s in range(2):
        adj, _ = orb.solve_orbit_adjoint_one_species(lus[0], [r[s]] + [np.zeros(wl.N)] * 7, tol=1e-12,
                                                     maxiter=500)
        assert _rel(st.grad[s].cpu().numpy(), adj[0]) < 1e-8


def test_fct_autograd_recompute_equals_stored(setup):
    wl, op, grids, lus, e = setup
    gs = []
    for recompute in (False, True):
        et = torch.from_numpy(e[:1]).cuda().requires_grad_(True)
        orbit = hag.orbit_solve_fct(et, op, grids, tracers=[0], n_passes=2, tol=1e-11, recompute=recompute)
        orbit[:, 0, :20].sum().backward()
        gs.append(et.grad.cpu().numpy())
    assert _rel(gs[1], gs[0]) < 1e-9


def test_fct_autograd_mixed_recompute_is_bit_identical(setup):
    """Mixed precision: the forward passes (later passes started from the
    previous orbit plus the propagated change) are deterministic, so recomputing
    them in the backward gives the stored orbit bit for bit, hence the same
    gradient; and it is within the 1e-6 bar of the FP64 gradient."""
    wl, op, grids, lus, e = setup
    gs = []
    for recompute, precision in ((False, "mixed"), (True, "mixed"), (False, "fp64")):
        et = torch.from_numpy(e[:1]).cuda().requires_grad_(True)
        orbit = hag.orbit_solve_fct(et, op, grids, tracers=[0], n_passes=3, tol=1e-8 if precision == "mixed"
                                    else 1e-11, precision=precision, recompute=recompute)
        orbit[:, 0, :20].sum().backward()
        gs.append(et.grad.cpu().numpy())
    assert np.array_equal(gs[0], gs[1])
    assert _rel(gs[0], gs[2]) < 1e-6


def test_marginal_fct_mixed_matches_oracle(setup):
    """The marginal driver's outer GMRES over mixed-precision periodic solves
    (each matvec a cold batched solve) against the oracle at the 1e-6 bar."""
    wl, op, grids, lus, e = setup
    ed = torch.from_numpy(e).cuda()
    base, _ = drivers.forward_fct(op, grids, ed, n_passes=2, tol=1e-11, precision="fp64")
    L = drivers.frozen_operators(op, grids, base[:1])
    de = 0.01 * e[:1] * np.random.default_rng(0).lognormal(0.0, 1.0, e[:1].shape)
    dc, info = drivers.marginal_fct(op, L, torch.from_numpy(de).cuda(), tracers=[0], tol=1e-8,
                                    precision="mixed", max_iters=60, rel_tol=1e-8)
    Lref = _oracle_LAD(wl, base[0].cpu().numpy())
    ref, _ = drivers_ref.marginal_fct_gmres(lus[0], list(de[0]), Lref, tol=1e-12, max_iters=60, rel_tol=1e-9)
    assert _rel(dc[0].cpu().numpy(), ref) < 1e-6


def test_adjoint_fct_mixed_matches_oracle(setup):
    """The adjoint FCT driver's outer GMRES over mixed-precision adjoint solves:
    the emission gradient against the oracle's at the 1e-6 bar."""
    wl, op, grids, lus, e = setup
    ed = torch.from_numpy(e).cuda()
    base, _ = drivers.forward_fct(op, grids, ed, n_passes=2, tol=1e-11, precision="fp64")
    L = drivers.frozen_operators(op, grids, base[:1])
    rng = np.random.default_rng(1)
    R = np.zeros((1, N_BINS, wl.N))
    R[0, 0, rng.choice(wl.nx * wl.ny, 10, replace=False)] = rng.random(10)
    lam, Re, info = drivers.adjoint_fct(op, L, torch.from_numpy(R).cuda(), tracers=[0], tol=1e-8,
                                        precision="mixed", max_iters=60, rel_tol=1e-8)
    grad = wl.dtau * (lam[0, :N_BINS] - Re[0]).cpu().numpy()
    Lref = _oracle_LAD(wl, base[0].cpu().numpy())
    lr, Rr, _ = drivers_ref.adjoint_fct_gmres(lus[0], list(R[0]), Lref, tol=1e-12, max_iters=60, rel_tol=1e-9)
    gr = np.stack([wl.dtau * (lr[t] - Rr[t]) for t in range(N_BINS)])
    assert _rel(grad, gr) < 1e-6


def _frozen_tangent(op, grids, passes, v, tol):
    """dc of the n-pass recursion with the FCT coefficients frozen at each
    pass's input orbit: dc^(0) = S(v), dc^(p) = S(v - L_AD^(p-1) dc^(p-1)[tau+1])."""
    from orbit_amd.fct import apply_dev
    dc, _ = op.solve_forward(v, [0], tol=tol, precision="fp64")
    tmp = torch.empty(op.N, dtype=torch.float64, device="cuda")
    for p in range(1, passes.shape[0] + 1):
        L = drivers.frozen_operators(op, grids, passes[p - 1])
        rhs = v.clone()
        for tau in range(N_BINS):
            apply_dev(L[tau][0], dc[0, tau + 1].contiguous(), tmp)
            rhs[0, tau] -= tmp
        dc, _ = op.solve_forward(rhs, [0], tol=tol, precision="fp64")
    return dc


def test_fct_autograd_is_the_derivative_of_its_passes(setup):
    """orbit_solve_fct's backward ("passes") is the derivative of the forward it
    ran: (1) a central finite difference of the true forward along e itself
    (the passes are homogeneous of degree one in e, so the limiter's
    switching does not enter) equals <grad, e> and J(e); (2) along a random
    direction the gradient is the exact adjoint of the frozen-coefficient
    tangent of the same passes (duality to 1e-8).  The converged-fixed-point
    gradient ("fixed_point", the reference's outer GMRES) is another quantity
    (the derivative of the limit of the passes); it runs and is printed
    beside the others."""
    wl, op, grids, lus, e = setup
    n = 4
    rng = np.random.default_rng(17)
    gmask = torch.from_numpy(rng.random((1, N_BINS + 1, wl.N))).cuda()
    et = torch.from_numpy(e[:1].copy()).cuda().requires_grad_(True)
    orbit = hag.orbit_solve_fct(et, op, grids, tracers=[0], n_passes=n, tol=1e-12)
    J = float((gmask * orbit.detach()).sum())
    (gmask * orbit).sum().backward()
    grad = et.grad.clone()
    ed = torch.from_numpy(e[:1].copy()).cuda()
    eps = 1e-3
    fwd = lambda x: float((gmask * drivers.forward_fct(op, grids, x, tracers=[0], n_passes=n, tol=1e-12,  # noqa
                                                       precision="fp64")[0]).sum())
    fd = (fwd(ed * (1 + eps)) - fwd(ed * (1 - eps))) / (2 * eps)
    ge = float((grad * ed).sum())
    print(f"J {J:.12e}  FD along e {fd:.12e}  <grad, e> {ge:.12e}", flush=True)
    assert abs(fd - J) <= 1e-9 * abs(J)
    assert abs(ge - fd) <= 1e-8 * abs(fd)
    # duality with the frozen tangent along a random direction
    _, _, passes = drivers.forward_fct(op, grids, ed, tracers=[0], n_passes=n, tol=1e-12, precision="fp64",
                                       keep_passes=True)
    v = torch.from_numpy(rng.standard_normal((1, N_BINS, wl.N)) * np.abs(e[:1]).mean()).cuda()
    dc = _frozen_tangent(op, grids, passes, v, 1e-12)
    lhs, rhs = float((gmask * dc).sum()), float((grad * v).sum())
    print(f"<g, dc(v)> {lhs:.12e}  <grad, v> {rhs:.12e}", flush=True)
    assert abs(lhs - rhs) <= 1e-8 * max(abs(lhs), float((gmask.abs() * dc.abs()).sum()))
    # the fixed-point gradient: the derivative of the limit of the passes
    et2 = torch.from_numpy(e[:1].copy()).cuda().requires_grad_(True)
    o2 = hag.orbit_solve_fct(et2, op, grids, tracers=[0], n_passes=n, tol=1e-12, mode="fixed_point",
                             rel_tol=1e-10, max_iters=80)
    (gmask * o2).sum().backward()
    gf = float((et2.grad * ed).sum())
    print(f"fixed-point <grad, e> {gf:.12e} (rel. to the passes' {abs(gf - fd) / abs(fd):.2e})", flush=True)
    assert np.isfinite(gf)
