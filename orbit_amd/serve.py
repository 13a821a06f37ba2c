"""Serving loop: continuous batching of forward-marginal and adjoint orbit queries.

The reference has no serving path (batch/CLI only); this is the request loop
the periodic-orbit solve is wrapped in for interactive marginal-emission
queries.  Every month's operator set stays resident on the device
(OperatorSet: shared column blocks, stencils, factored BF16/FP64 column blocks).

* A query names a month, a sub-domain size class, 1-12 tracers of that month's
  set and a mode: "forward" (marginal: the linear response dc to a perturbation
  de, a forward periodic solve) or "adjoint" (receptor weights R at c_0, an
  adjoint periodic solve giving dJ/de for every bin).
* Admission: a query is admitted when the device memory its batch needs
  (scratch of the solve + its inputs/outputs) fits in the free part of the HBM
  budget; otherwise it waits.  The budget is the device's free memory
  (hipMemGetInfo) once the month operator sets are resident, less a reserve,
  unless the caller gives one.
* Continuous batching: each of n_workers host threads owns a HIP stream
  (Context).  A free worker takes the oldest waiting query and every other
  waiting query of the same mode and tolerance -- of any month and sub-domain
  (a group solve over their operator sets, orbit_solve_group_dev), or of the
  same month only when a set cannot join a group -- up to max_slots tracers
  and within the budget, and solves them as one batch; queries that arrive
  while batches are on the device are picked up by the next free worker, so
  work of different batches overlaps on the GPU.
* Results are independent of batching: every kernel and reduction works per
  tracer slot in a fixed order, so a query's answer is bit-identical alone and
  in any mix (tests/test_gpu_serve.py).
"""
from __future__ import annotations

import dataclasses
import threading
import time
from collections import deque
from concurrent.futures import Future
from typing import Dict, List, Optional, Tuple

import numpy as np

from .core import Context, OperatorSet


@dataclasses.dataclass
class Query:
    month: int
    size: str
    tracers: List[int]
    mode: str                      # "forward" | "adjoint"
    payload: object                # CUDA float64 (nt, n_bins, N): de (forward) or R (adjoint)
    tol: float = 1e-6
    qid: int = -1
    t_arrival: float = 0.0
    t_start: float = 0.0
    t_done: float = 0.0
    batch_size: int = 0
    future: Optional[Future] = None


def slot_bytes(op: OperatorSet, restart: int = 30) -> int:
    """Device bytes one tracer slot of a solve needs (scratch + input + output)."""
    n16 = op.ny * op.nx * 16
    nb = op.n_bins
    doubles = n16 * (nb + (nb + 1) + (restart + 1) + 6) + op.N * (nb + (nb + 1))
    return int(doubles * 8 + n16 * 3 * 4 + 64 * 1024)


class ShapeTuner:
    """Autotuner of the sweep launch shape per sub-domain size class.

    The first operator set of a size class (nz, ny, nx) is timed over every
    half-sweep launch shape (slots per wave, threads per workgroup, XCD
    mapping) for each sweep precision and slot bucket
    (``OperatorSet.autotune``); every later set of that class gets the same
    table (``set_shapes``).  Shapes change timing only: every shape computes
    the same per-lane arithmetic, so results are bit-identical."""

    def __init__(self, reps: int = 40):
        self.reps = reps
        self.tables: Dict[Tuple[int, int, int], list] = {}
        self.us: Dict[Tuple[int, int, int], list] = {}

    def apply(self, op: OperatorSet) -> list:
        key = (op.nz, op.ny, op.nx)
        if key not in self.tables:
            self.tables[key], self.us[key] = op.autotune(self.reps)
        else:
            op.set_shapes(self.tables[key])
        return self.tables[key]

    def report(self) -> dict:
        """{"NZxNYxNX": {precision: {bucket: [slots_per_wave, threads, xcd_map, us]}}}"""
        out = {}
        for key, tab in self.tables.items():
            us = self.us[key]
            d = {}
            for p, pname in enumerate(OperatorSet.SWEEP_PRECS):
                d[pname] = {b: list(tab[p * 5 + k]) + [round(us[p * 5 + k], 3)]
                            for k, b in enumerate(OperatorSet.SLOT_BUCKETS)}
            out["x".join(map(str, key))] = d
        return out


def device_free_bytes(device: int = 0) -> int:
    """Free device memory now (hipMemGetInfo), after whatever is resident."""
    import torch
    return int(torch.cuda.mem_get_info(device)[0])


class OrbitServer:
    """hbm_budget: device bytes the in-flight batches may use.  None (the
    default) takes it from the device itself: the free memory reported by
    hipMemGetInfo once the month operator sets are resident, less a reserve
    (``reserve_frac`` of it, for torch's allocator and the libraries)."""

    def __init__(self, opsets: Dict[Tuple[int, str], OperatorSet], n_workers: int = 4,
                 max_slots: int = 16, hbm_budget: Optional[float] = None, precision: str = "mixed",
                 device: int = 0, tuner: Optional[ShapeTuner] = None, reserve_frac: float = 0.1,
                 across_months: bool = True):
        if tuner is not None:  # before any worker launches on these sets
            for op in opsets.values():
                tuner.apply(op)
        self.opsets = opsets
        self.max_slots = max_slots
        if hbm_budget is None:
            self.free_at_start = device_free_bytes(device)
            hbm_budget = (1.0 - reserve_frac) * self.free_at_start
            self.budget_source = "hipMemGetInfo"
        else:
            self.free_at_start = None
            self.budget_source = "caller"
        self.budget = float(hbm_budget)
        self.in_use = 0.0
        # queries of different months / sub-domains share a batch (a group
        # solve) when their sets allow it: one shared transport per bin and
        # two sweep colours
        self.across = across_months
        self.groupable = {}
        for key, op in opsets.items():
            st = op.stats()
            self.groupable[key] = st["n_groups"] == 1 and not (st["periodic_x"] and op.nx >= 3 and op.nx % 2)
        self.precision = precision
        self.queue: deque = deque()
        self.cv = threading.Condition()
        self.stop = False
        self.done: List[Query] = []
        self.next_id = 0
        self.ctxs = [Context(device) for _ in range(n_workers)]
        self.workers = [threading.Thread(target=self._work, args=(c,), daemon=True) for c in self.ctxs]
        for w in self.workers:
            w.start()

    # -- client side ------------------------------------------------------
    def submit(self, q: Query) -> Future:
        key = (q.month, q.size)
        if key not in self.opsets:
            raise KeyError(f"no resident operator set for month {q.month}, size {q.size}")
        op = self.opsets[key]
        if q.mode not in ("forward", "adjoint"):
            raise ValueError("mode must be 'forward' or 'adjoint'")
        if any(t < 0 or t >= op.n_tracers for t in q.tracers):
            raise ValueError("tracer index out of range")
        if tuple(q.payload.shape) != (len(q.tracers), op.n_bins, op.N):
            raise ValueError("payload must be (n_tracers, n_bins, N)")
        q.future = Future()
        with self.cv:
            q.qid = self.next_id
            self.next_id += 1
            q.t_arrival = time.perf_counter()
            self.queue.append(q)
            self.cv.notify()
        return q.future

    def close(self):
        with self.cv:
            self.stop = True
            self.cv.notify_all()
        for w in self.workers:
            w.join()
        for c in self.ctxs:
            c.close()

    # -- worker side ------------------------------------------------------
    def _compatible(self, q, head) -> bool:
        if (q.mode, q.tol) != (head.mode, head.tol):
            return False
        if (q.month, q.size) == (head.month, head.size):
            return True
        return self.across and self.groupable[(q.month, q.size)] and self.groupable[(head.month, head.size)]

    def _take_batch(self):
        """Oldest waiting query plus compatible ones that fit (called with cv held)."""
        while self.queue:
            head = self.queue[0]
            per_head = slot_bytes(self.opsets[(head.month, head.size)])
            if self.in_use + per_head * len(head.tracers) > self.budget and self.in_use > 0:
                return None, 0.0                      # wait for memory to free up
            batch, slots, need = [], 0, 0.0
            for q in list(self.queue):
                if not self._compatible(q, head):
                    continue
                n = len(q.tracers)
                b = slot_bytes(self.opsets[(q.month, q.size)]) * n
                if batch and (slots + n > self.max_slots or self.in_use + need + b > self.budget):
                    continue
                batch.append(q)
                slots += n
                need += b
            for q in batch:
                self.queue.remove(q)
            self.in_use += need
            return batch, need
        return None, 0.0

    def _work(self, ctx: Context):
        import torch
        while True:
            with self.cv:
                batch = None
                while not self.stop:
                    batch, need = self._take_batch()
                    if batch:
                        break
                    self.cv.wait(timeout=0.05)
                if self.stop and not batch:
                    return
            t0 = time.perf_counter()
            try:
                head = batch[0]
                nslots = sum(len(q.tracers) for q in batch)
                if len({(q.month, q.size) for q in batch}) > 1:
                    # several months / sub-domains: one group solve over all their slots
                    res = solve_group([(self.opsets[(q.month, q.size)], q.payload, q.tracers) for q in batch],
                                      precision=self.precision, tol=head.tol, adjoint=head.mode == "adjoint",
                                      ctx=ctx)
                else:
                    op = self.opsets[(head.month, head.size)]
                    tracers = [t for q in batch for t in q.tracers]
                    inp = torch.cat([q.payload for q in batch], dim=0).contiguous()
                    fn = op.solve_forward if head.mode == "forward" else op.solve_adjoint
                    out, info = fn(inp, tracers, tol=head.tol, precision=self.precision, ctx=ctx)
                    res, k = [], 0
                    for q in batch:
                        n = len(q.tracers)
                        res.append((out[k:k + n], info[k:k + n]))
                        k += n
                t1 = time.perf_counter()
                for q, r in zip(batch, res):
                    q.t_start, q.t_done, q.batch_size = t0, t1, nslots
                    q.future.set_result(r)
            except Exception as exc:  # surfaced to every query of the batch
                for q in batch:
                    q.future.set_exception(exc)
            finally:
                with self.cv:
                    self.in_use -= need
                    self.done.extend(batch)
                    self.cv.notify_all()


def query_payload(n_bins: int, N: int, ncol: int, tracers, mode: str, rng, base_e=None) -> np.ndarray:
    """Synthetic query input (nt, n_bins, N), host float64: "forward" = a 1 %
    lognormal perturbation of the month's emissions of the chosen tracers
    (the marginal forcing de); "adjoint" = a receptor on 10-100 random surface
    cells with U(0,1) weights at c_0."""
    nt = len(tracers)
    if mode == "forward":
        e = np.asarray(base_e)[list(tracers)]
        return 0.01 * e * rng.lognormal(0.0, 1.0, e.shape)
    if mode != "adjoint":
        raise ValueError("mode must be 'forward' or 'adjoint'")
    R = np.zeros((nt, n_bins, N))
    for i in range(nt):
        m = int(rng.integers(10, 101))
        cells = rng.choice(ncol, m, replace=False)
        R[i, 0, cells] = rng.random(m)
    return R


def make_query(op: OperatorSet, month: int, size: str, tracers, mode: str, rng, base_e=None,
               tol: float = 1e-6):
    """A Query whose payload (``query_payload``) is on the operator set's device."""
    import torch
    payload = query_payload(op.n_bins, op.N, op.nx * op.ny, tracers, mode, rng, base_e)
    return Query(month=month, size=size, tracers=list(tracers), mode=mode,
                 payload=torch.from_numpy(np.ascontiguousarray(payload)).to(f"cuda:{op.ctx.device}"), tol=tol)


def run_poisson(server: OrbitServer, rate: float, n_queries: int, make, rng) -> dict:
    """Submit n_queries with exponential inter-arrival times at `rate` per second;
    return latency percentiles and throughput."""
    queries = [make(rng) for _ in range(n_queries)]
    gaps = rng.exponential(1.0 / rate, n_queries)
    t0 = time.perf_counter()
    futs = []
    t_next = t0
    for q, g in zip(queries, gaps):
        t_next += g
        dt = t_next - time.perf_counter()
        if dt > 0:
            time.sleep(dt)
        futs.append(server.submit(q))
    for f in futs:
        f.result()
    wall = time.perf_counter() - t0
    lat = np.array([q.t_done - q.t_arrival for q in queries]) * 1e3
    solves = sum(len(q.tracers) for q in queries)
    return {"rate": rate, "queries": n_queries, "orbit_solves": solves, "wall_s": wall,
            "solves_per_s": solves / wall, "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)), "mean_batch": float(np.mean([q.batch_size for q in queries]))}


def solve_group(jobs, precision: str = "mixed", tol: float = 1e-8, adjoint: bool = False, maxiter: int = 200,
                restart: int = 30, ctx: Optional[Context] = None, inner_sweeps: int = 0, check_sweeps: int = 0):
    """Every orbit of every job in one batched solve (orbit_solve_group_dev):
    jobs = [(OperatorSet, input (nt, n_bins, N) CUDA float64, tracers)], the
    input emissions (forward) or receptors (adjoint).  Each launch covers all
    slots of all sets, whatever their sub-domain sizes.  Returns [(orbit
    (nt, n_bins+1, N), infos)] per job, each bit-identical to the job solved
    alone.  inner_sweeps / check_sweeps (mixed): sweeps per inner solve of the
    monodromy and of the residual sweep, 0 = each set's own."""
    import ctypes as C
    import torch
    from . import _lib
    if not jobs:
        return []
    ctx = ctx or jobs[0][0].ctx
    sets, set_idx = [], {}
    slot_set, tracers, ins, outs, results, keep = [], [], [], [], [], []
    for op, inp, tr in jobs:
        tr = list(tr)
        if id(op) not in set_idx:
            set_idx[id(op)] = len(sets)
            sets.append(op)
        _lib.dev_f64(inp, ctx.device, "input")
        if tuple(inp.shape) != (len(tr), op.n_bins, op.N):
            raise ValueError(f"input must have shape {(len(tr), op.n_bins, op.N)}, got {tuple(inp.shape)}")
        inp = inp.contiguous()
        out = torch.empty((len(tr), op.n_bins + 1, op.N), dtype=torch.float64, device=inp.device)
        keep.append(inp)
        results.append(out)
        for q, t in enumerate(tr):
            slot_set.append(set_idx[id(op)])
            tracers.append(int(t))
            ins.append(inp[q].data_ptr())
            outs.append(out[q].data_ptr())
    nt = len(tracers)
    opts = OperatorSet._opts(tol, maxiter, restart, precision, inner_sweeps, 0, False, 0, False,
                             check=check_sweeps)
    info = (_lib.Info * nt)()
    arr_i = lambda v: (C.c_int * len(v))(*v)  # noqa: E731
    arr_p = lambda v: (C.c_void_p * len(v))(*v)  # noqa: E731
    torch.cuda.current_stream(results[0].device).synchronize()
    _lib.check(_lib.lib().orbit_solve_group_dev(ctx.handle, len(sets), arr_p([s.handle.value for s in sets]), nt,
                                                arr_i(slot_set), arr_i(tracers), arr_p(ins), int(bool(adjoint)),
                                                C.byref(opts), arr_p(outs), info))
    ctx.sync()
    infos = [info[k].as_dict() for k in range(nt)]
    out, k = [], 0
    for (op, inp, tr), res in zip(jobs, results):
        out.append((res, infos[k:k + len(tr)]))
        k += len(tr)
    return out


def solve_many(jobs, n_workers: int = 4, precision: str = "mixed", tol: float = 1e-8,
               tuner: Optional[ShapeTuner] = None):
    """Batched uneven orbits: jobs = [(OperatorSet, e (nt, n_bins, N) CUDA, tracers)],
    run as concurrent batches on n_workers streams (launch shapes per size class
    from `tuner` when given).  Returns (results, wall seconds)."""
    opsets = {(i, "job"): j[0] for i, j in enumerate(jobs)}
    srv = OrbitServer(opsets, n_workers=n_workers, max_slots=max(len(j[2]) for j in jobs),
                      precision=precision, tuner=tuner)
    try:
        t0 = time.perf_counter()
        futs = [srv.submit(Query(month=i, size="job", tracers=list(j[2]), mode="forward", payload=j[1], tol=tol))
                for i, j in enumerate(jobs)]
        res = [f.result() for f in futs]
        wall = time.perf_counter() - t0
    finally:
        srv.close()
    return res, wall
