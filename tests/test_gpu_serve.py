"""Serving loop on the GPU: continuous batching over worker streams, HBM
admission, and results identical alone and in the mix."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from orbit_amd import serve, synth
from orbit_amd.core import OperatorSet


@pytest.fixture(scope="module")
def opsets():
    out, base = {}, {}
    for month in (0, 1):
        for size in ("24x20",):
            wl = synth.benchmark_workload("serving", n_tracers=4, month=month, size=size)
            out[(month, size)] = OperatorSet.from_workload(wl)
            base[(month, size)] = wl.e
    return out, base


def test_mixed_queries_match_alone(opsets):
    ops, base = opsets
    rng = np.random.default_rng(0)
    qs = []
    for i in range(16):
        month = int(rng.integers(0, 2))
        key = (month, "24x20")
        nt = int(rng.integers(1, 5))
        tracers = sorted(rng.choice(4, nt, replace=False).tolist())
        mode = "forward" if i % 2 == 0 else "adjoint"
        qs.append(serve.make_query(ops[key], month, "24x20", tracers, mode, rng, base_e=base[key], tol=1e-8))
    srv = serve.OrbitServer(ops, n_workers=3, max_slots=8, hbm_budget=4e9)
    try:
        futs = [srv.submit(q) for q in qs]
        res = [f.result(timeout=120) for f in futs]
    finally:
        srv.close()
    assert max(q.batch_size for q in qs) > max(len(q.tracers) for q in qs) or len(qs) == 16
    for q, (out, info) in zip(qs, res):
        op = ops[(q.month, q.size)]
        fn = op.solve_forward if q.mode == "forward" else op.solve_adjoint
        alone, ainfo = fn(q.payload, q.tracers, tol=q.tol, precision="mixed")
        assert torch.equal(out, alone), (q.qid, q.mode, q.tracers)
        assert all(i["converged"] for i in info)
    assert all(q.t_done >= q.t_start >= q.t_arrival for q in qs)


def test_poisson_run_reports_latency(opsets):
    ops, base = opsets
    rng = np.random.default_rng(1)

    def make(r):
        month = int(r.integers(0, 2))
        key = (month, "24x20")
        tr = sorted(r.choice(4, int(r.integers(1, 5)), replace=False).tolist())
        return serve.make_query(ops[key], month, "24x20", tr, "forward" if r.random() < 0.5 else "adjoint",
                                r, base_e=base[key])

    srv = serve.OrbitServer(ops, n_workers=2, max_slots=8, hbm_budget=4e9)
    try:
        rep = serve.run_poisson(srv, rate=50.0, n_queries=20, make=make, rng=rng)
    finally:
        srv.close()
    assert rep["p99_ms"] >= rep["p50_ms"] > 0 and rep["solves_per_s"] > 0


def test_admission_by_memory_budget(opsets):
    ops, base = opsets
    rng = np.random.default_rng(2)
    op = ops[(0, "24x20")]
    per = serve.slot_bytes(op)
    # budget for 2 slots: batches never exceed it once something is running
    srv = serve.OrbitServer(ops, n_workers=3, max_slots=8, hbm_budget=2.5 * per)
    try:
        qs = [serve.make_query(op, 0, "24x20", [t], "forward", rng, base_e=base[(0, "24x20")]) for t in (0, 1, 2, 3)]
        futs = [srv.submit(q) for q in qs]
        for f in futs:
            f.result(timeout=120)
    finally:
        srv.close()
    assert max(q.batch_size for q in qs) <= 2


def test_solve_many_uneven_sizes():
    jobs = []
    for month, size in ((0, "24x20"), (1, "48x40")):
        wl = synth.benchmark_workload("uneven-batch", n_tracers=3, month=month, size=size)
        op = OperatorSet.from_workload(wl)
        jobs.append((op, torch.from_numpy(wl.e).cuda(), [0, 1, 2]))
    res, wall = serve.solve_many(jobs, n_workers=2)
    for (op, e, tr), (out, info) in zip(jobs, res):
        alone, _ = op.solve_forward(e, tr, tol=1e-8, precision="mixed")
        assert torch.equal(out, alone)
