"""Serving loop on the GPU: continuous batching over worker streams, HBM
admission, and results identical alone and in the mix.

test_config4_mix_matches_alone_and_the_oracle serves the benchmark's serving
workload: resident month sets at 24x20, 48x40 and 64x70 (x15, 12 tracers,
four months), queries of 1-12 tracers, forward-marginal (1 % lognormal de)
and adjoint (10-100-cell receptor), tol 1e-6, admitted by the device's free
memory.  Every served result equals the same query solved alone, bit for bit,
and the pinned 24x20 and 48x40 queries of both modes match the CPU oracle
(SuperLU + GMRES, tests/golden/serving_oracle.npz, written by
tests/golden/make_serving_golden.py) within 1e-6 relative in c0 / lambda0 --
what the reference caller solve_orbit_one_species (orbit/core/orbit.py:123-210)
would get."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from orbit_amd import serve, synth
from orbit_amd.core import OperatorSet

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "serving_oracle.npz")
MIX = ((0, "24x20"), (1, "48x40"), (2, "64x70"), (3, "24x20"))
PIN_SEEDS = {"forward": 40, "adjoint": 50}


@pytest.fixture(scope="module")
def opsets():
    out, base = {}, {}
    for month in (0, 1):
        for size in ("24x20",):
            wl = synth.benchmark_workload("serving", n_tracers=4, month=month, size=size)
            out[(month, size)] = OperatorSet.from_workload(wl)
            base[(month, size)] = wl.e
    return out, base


@pytest.fixture(scope="module")
def mix():
    ops, base = {}, {}
    for month, size in MIX:
        wl = synth.benchmark_workload("serving", n_tracers=12, month=month, size=size)
        ops[(month, size)] = OperatorSet.from_workload(wl)
        base[(month, size)] = wl.e
    return ops, base


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def test_config4_mix_matches_alone_and_the_oracle(mix):
    ops, base = mix
    gold = np.load(GOLDEN, allow_pickle=False)
    pin_tr = [int(t) for t in gold["tracers"]]
    queries, pinned = [], []
    for month, size in MIX[:2]:
        op = ops[(month, size)]
        for mode, seed in PIN_SEEDS.items():
            rng = np.random.default_rng([seed, month])
            q = serve.make_query(op, month, size, pin_tr, mode, rng, base_e=base[(month, size)], tol=1e-6)
            # the inputs the oracle answered
            assert abs(float(q.payload.sum()) - float(gold[f"{mode}_{month}_payload_sum"])) <= \
                1e-12 * abs(float(gold[f"{mode}_{month}_payload_sum"]))
            pinned.append((len(queries), gold[f"{mode}_{month}_x0"]))
            queries.append(q)
    rng = np.random.default_rng(4)
    for i in range(20):
        month, size = MIX[int(rng.integers(0, len(MIX)))]
        nt = int(rng.integers(1, 13))
        tracers = sorted(rng.choice(12, nt, replace=False).tolist())
        mode = "forward" if rng.random() < 0.5 else "adjoint"
        queries.append(serve.make_query(ops[(month, size)], month, size, tracers, mode, rng,
                                        base_e=base[(month, size)], tol=1e-6))
    order = np.random.default_rng(5).permutation(len(queries))
    srv = serve.OrbitServer(ops, n_workers=4, max_slots=16)
    assert srv.budget_source == "hipMemGetInfo" and srv.budget > 0
    try:
        futs = {int(k): srv.submit(queries[int(k)]) for k in order}
        res = {k: f.result(timeout=300) for k, f in futs.items()}
    finally:
        srv.close()
    assert max(q.batch_size for q in queries) > min(len(q.tracers) for q in queries)  # something was batched
    for k, q in enumerate(queries):
        out, info = res[k]
        print(f"query {k}: month {q.month} {q.size} {q.mode} tracers {q.tracers} batch {q.batch_size}", flush=True)
        assert all(i["converged"] and i["residual"] <= q.tol for i in info), info
        op = ops[(q.month, q.size)]
        fn = op.solve_forward if q.mode == "forward" else op.solve_adjoint
        alone, ainfo = fn(q.payload, q.tracers, tol=q.tol, precision="mixed")
        assert torch.equal(out, alone), (k, q.mode, q.tracers)
        assert [i["gmres_iters"] for i in info] == [i["gmres_iters"] for i in ainfo]
    for k, ref in pinned:
        q = queries[k]
        out = res[k][0]
        for i in range(len(q.tracers)):
            err = _rel(out[i, 0].cpu().numpy(), ref[i])
            print(f"pinned {q.mode} month {q.month} tracer {q.tracers[i]}: rel err vs oracle {err:.2e}", flush=True)
            assert err <= 1e-6, (q.mode, q.month, q.tracers[i], err)


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
