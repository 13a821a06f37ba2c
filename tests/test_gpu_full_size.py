"""The FCT month (64x70x15, 12 tracers) and an uneven mix of month sets at
full size, through size-independent properties
(the oracle does not finish in seconds here): the mixed-precision FCT month
against the plain FP64 path per tracer at the 1e-6 bar with converged
residuals and periodic orbits; a tracer's orbit bit-identical alone and inside
the 12-tracer batch; an uneven 24x20 / 48x40 / 64x70 mix of month sets solved
concurrently equal, bit for bit, to each set solved alone."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from orbit_amd import drivers, serve, synth
from orbit_amd.core import OperatorSet


def test_fct_month_mixed_vs_fp64_and_alone_vs_batch():
    wl = synth.benchmark_workload("fct-month")
    op = OperatorSet.from_workload(wl)
    grids = drivers.make_fct_grids(wl.grids, op.ctx)
    e = torch.from_numpy(wl.e).cuda()
    om, hm = drivers.forward_fct(op, grids, e, n_passes=6, tol=1e-8, precision="mixed")
    for h in hm:
        assert all(i["converged"] and i["residual"] <= 1e-8 for i in h)
    assert all(i["periodicity"] < 1e-7 for i in hm[-1])
    for t in (0, 7):
        of, hf = drivers.forward_fct(op, grids, e[t:t + 1], tracers=[t], n_passes=6, tol=1e-8,
                                     precision="fp64")
        err = float((om[t, 0] - of[0, 0]).abs().max() / of[0, 0].abs().max())
        assert err <= 1e-6, (t, err)
    alone, _ = drivers.forward_fct(op, grids, e[5:6], tracers=[5], n_passes=6, tol=1e-8, precision="mixed")
    assert torch.equal(alone[0], om[5])


def test_uneven_mix_equals_each_set_alone():
    jobs = []
    for month, size in enumerate(("24x20", "48x40", "64x70")):
        wlm = synth.benchmark_workload("uneven-batch", n_tracers=12, month=month, size=size)
        opm = OperatorSet.from_workload(wlm)
        jobs.append((opm, torch.from_numpy(wlm.e).cuda(), list(range(12))))
    res, _ = serve.solve_many(jobs, n_workers=3)
    for (opm, e, tr), (orbit, info) in zip(jobs, res):
        alone, ia = opm.solve_forward(e, tr, tol=1e-8, precision="mixed")
        assert torch.equal(orbit, alone)
        assert all(i["converged"] for i in info)


# ---------------------------------------------------------------------------
# against the CPU oracle at the production shape (configs 1 and 3):
# tests/golden/fullsize_oracle.npz holds the oracle's FCT-month c0 and adjoint
# lambda0 (SuperLU + scipy GMRES to 1e-12) of tracers 0 and 7, written by
# tests/golden/make_fullsize_golden.py from the same seeded workloads
GOLDEN = __import__("os").path.join(__import__("os").path.dirname(__file__), "golden", "fullsize_oracle.npz")


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def _receptors(wl, seed=303):
    """The fixture's receptors: 50 seeded surface cells per tracer, U(0,1) weights, bin 0."""
    rng = np.random.default_rng(seed)
    ncol = wl.nx * wl.ny
    R = np.zeros((wl.n_tracers, 8, wl.N))
    for t in range(wl.n_tracers):
        cells = rng.choice(ncol, 50, replace=False)
        R[t, 0, cells] = rng.random(50)
    return R


def test_fct_month_against_the_cpu_oracle():
    gold = np.load(GOLDEN, allow_pickle=False)
    tracers = [int(t) for t in gold["tracers"]]
    wl = synth.benchmark_workload("fct-month")
    op = OperatorSet.from_workload(wl)
    grids = drivers.make_fct_grids(wl.grids, op.ctx)
    e = torch.from_numpy(wl.e).cuda()
    # the headline solve: all 12 tracers batched, mixed precision, tol 1e-8
    om, hm = drivers.forward_fct(op, grids, e, n_passes=6, tol=1e-8, precision="mixed")
    et = torch.from_numpy(np.ascontiguousarray(wl.e[tracers])).cuda()
    of, hf = drivers.forward_fct(op, grids, et, tracers=tracers, n_passes=6, tol=1e-10, precision="fp64")
    for q, t in enumerate(tracers):
        ref0, ref8 = gold[f"forward_c0_{t}"], gold[f"forward_c8_{t}"]
        em = _rel(om[t, 0].cpu().numpy(), ref0)
        ef = _rel(of[q, 0].cpu().numpy(), ref0)
        print(f"FCT month tracer {t}: mixed {em:.2e}, fp64 {ef:.2e} vs the oracle", flush=True)
        assert em <= 1e-6 and ef <= 1e-8, (t, em, ef)
        assert _rel(of[q, 8].cpu().numpy(), ref8) <= 1e-8


def test_adjoint_against_the_cpu_oracle():
    gold = np.load(GOLDEN, allow_pickle=False)
    tracers = [int(t) for t in gold["tracers"]]
    wl = synth.benchmark_workload("adjoint")
    R = _receptors(wl)
    assert abs(R.sum() - gold["adjoint_fingerprint"][-1]) <= 1e-12 * abs(R.sum())
    op = OperatorSet.from_workload(wl)
    Rd = torch.from_numpy(R).cuda()
    lm, im = op.solve_adjoint(Rd, tol=1e-8, precision="mixed")
    Rt = torch.from_numpy(np.ascontiguousarray(R[tracers])).cuda()
    lf, inf = op.solve_adjoint(Rt, tracers=tracers, tol=1e-10, precision="fp64")
    assert all(i["converged"] for i in im + inf)
    for q, t in enumerate(tracers):
        ref = gold[f"adjoint_lam0_{t}"]
        em = _rel(lm[t, 0].cpu().numpy(), ref)
        ef = _rel(lf[q, 0].cpu().numpy(), ref)
        print(f"adjoint tracer {t}: mixed {em:.2e}, fp64 {ef:.2e} vs the oracle", flush=True)
        assert em <= 1e-6 and ef <= 1e-8, (t, em, ef)
