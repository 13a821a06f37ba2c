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
