"""Benchmark: periodic-orbit solves/s at 64x70x15 with 12 tracers on one MI355X.

Headline (the production configuration, FCT on): one step = the forward
periodic orbits of the 12 tracers of a month's operator set with the horizontal
FCT deferred correction (6 passes, each a periodic solve whose RHS carries the
anti-diffusive source of the previous pass's end-of-bin orbit), solved as one
batch in mixed precision (FP32 red-black column sweeps with BF16 column blocks
inside FP64 GMRES, FP64 refinement to 1e-8 relative residual), inputs resident
in HBM.  value = orbit solves/s = 12 / step time.

Beside it: the plain FP64 GPU solve of the same work (same algorithm in FP64,
one orbit at a time, no batching, no graphs) and the same 12 orbits batched in
FP64 (so the speed-up splits into batching and precision), every tracer's c0
against its plain FP64 solve and tracers 0 and 7 against the CPU oracle's FCT
month (tests/golden/fullsize_oracle.npz), p50/p99 per-step latency, the same
without FCT, the adjoint at the same size (50-cell receptors), 144 uneven
batched orbits (12 months x 12 tracers over 24x20 / 48x40 / 64x70), a Poisson
serving sweep over those months admitted by the device's free memory, the
roofline of the dominant half-sweep variant from the launches' own start/stop
events over one headline step, and the CPU baseline: the oracle (SuperLU
factors + scipy GMRES + numpy FCT) as min(cores, 12) single-threaded processes,
each one complete FCT-month orbit, plus one process alone (single core) and the
complete 24x20x15 orbit of config 0.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--quick]
"""
import argparse
import ctypes as C
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "orbit solves/s (64×70×15, 12 tracers) vs plain FP64 GPU; p50/p99 latency"
HBM_PEAK_GBS = 8000.0
N_PASSES = 6
# sweep kernel with the most device time per headline step (kernel-trace profile)
DOMINANT = "inner"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


CPU_THREAD_VARS = ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS")
BOX_CPU_SHARE = 16  # host cores a one-GPU box gives a job


def _cpu_worker(spec):
    """One single-threaded oracle orbit (the CPU baseline's unit of work),
    printed as JSON: "fct-month:<t>" = tracer t of the 64x70x15 FCT month (8
    SuperLU factors + 6 deferred-correction passes of scipy GMRES(30) to tol
    1e-8); "single" = the 24x20x15 one-tracer orbit without FCT (tol 1e-8)."""
    from oracle import drivers_ref, orbit_ref as orb
    from orbit_amd import synth
    kind, _, t = spec.partition(":")
    wl = synth.benchmark_workload(kind)
    t = int(t or 0)
    t0 = time.perf_counter()
    lus = [orb.factor_one_bin(wl.L(b, t)) for b in range(8)]
    t1 = time.perf_counter()
    if kind == "fct-month":
        orbit, hist = drivers_ref.forward_fct_passes(lus, list(wl.e[t]), wl.grids, n_passes=N_PASSES, tol=1e-8)
        iters = [h["gmres_iters"] for h in hist]
        ok = all(h["converged"] for h in hist)
    else:
        orbit, info = orb.solve_orbit_one_species(lus, list(wl.e[t]), tol=1e-8)
        iters, ok = [info["gmres_iters"]], info["converged"]
    t2 = time.perf_counter()
    print(json.dumps({"spec": spec, "factor_s": t1 - t0, "solve_s": t2 - t1, "iters": iters, "converged": bool(ok)}))


def _run_workers(specs):
    import subprocess
    env = dict(os.environ)
    for v in CPU_THREAD_VARS:
        env[v] = "1"
    procs = [subprocess.Popen([sys.executable, os.path.abspath(__file__), "--cpu-worker", sp], env=env,
                              stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True) for sp in specs]
    out = []
    for p in procs:
        o, e = p.communicate()
        if p.returncode != 0:
            raise RuntimeError(f"CPU baseline worker failed ({p.returncode}): {e[-2000:]}")
        out.append(json.loads(o.strip().splitlines()[-1]))
    return out


def cpu_baseline(n_tracers):
    """The oracle on the host: single-threaded processes, one orbit each.
    First one process alone (the single-core figure: the 24x20x15 orbit and
    one complete FCT-month orbit), then min(cores, tracers) processes at once,
    each the complete FCT-month orbit of its own tracer; throughput = orbits /
    the slowest process's time, with and without the SuperLU factorisations
    (the reference factors per solve)."""
    cores = min(len(os.sched_getaffinity(0)), BOX_CPU_SHARE)
    P = max(1, min(cores, n_tracers))
    single = _run_workers(["single"])[0]
    one = _run_workers(["fct-month:0"])[0]
    many = _run_workers([f"fct-month:{t}" for t in range(P)])
    slow_total = max(w["factor_s"] + w["solve_s"] for w in many)
    slow_solve = max(w["solve_s"] for w in many)
    return {
        "value": round(P / slow_total, 5), "unit": "orbit solves/s", "cores": P, "kind": "port",
        "solve_only_value": round(P / slow_solve, 5),
        "single_core_value": round(1.0 / (one["factor_s"] + one["solve_s"]), 5),
        "single_core_solve_only_value": round(1.0 / one["solve_s"], 5),
        "host_cores_visible": len(os.sched_getaffinity(0)),
        "config0_single_core": {"orbit_s": round(single["factor_s"] + single["solve_s"], 3),
                                "factor_s": round(single["factor_s"], 3), "gmres_iters": single["iters"],
                                "converged": single["converged"]},
        "workers": [{k: (round(v, 3) if isinstance(v, float) else v) for k, v in w.items()} for w in many],
        "sample": (f"oracle (SuperLU per bin + scipy GMRES(30), numpy FCT) on the GPU box's host: {P} "
                   f"single-threaded processes at once, each one complete 64x70x15 FCT-month orbit ({N_PASSES} "
                   "passes, tol 1e-8) of its own tracer, 8 factorisations included (solve_only_value excludes "
                   "them); single_core_* = tracer 0's orbit in one process alone; config0_single_core = the "
                   "complete 24x20x15 one-tracer orbit"),
    }


def from_profile(name, what="traffic_bytes_per_launch"):
    """A figure of a sweep kernel variant from the committed rocprofv3 summary
    (HBM bytes per launch from FETCH_SIZE + WRITE_SIZE scaled by a calibration
    kernel, or the kernel trace's average duration), or None."""
    path = os.path.join(ROOT, "profiles", "halfsweep_64x70x15_12tracers.md")
    if not os.path.exists(path):
        return None
    m = re.search(name + "_" + what + r":\s*([0-9.eE+]+)", open(path).read())
    return float(m.group(1)) if m else None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--quick", action="store_true", help="headline only (profiling runs)")
    ap.add_argument("--tracers", type=int, default=12)
    ap.add_argument("--plain-tracers", type=int, default=12,
                    help="tracers timed one at a time with the plain FP64 solve (each one's c0 is "
                         "also the reference its mixed-precision c0 is compared with)")
    ap.add_argument("--roof-reps", type=int, default=400,
                    help="back-to-back half-sweep launches timed for the roofline")
    ap.add_argument("--calib-copy", action="store_true",
                    help="run one 256 MiB device copy first (scales profiler byte counters)")
    ap.add_argument("--mfma", action="store_true",
                    help="column-block products of the reduced-precision sweeps on the matrix cores instead of "
                         "the VALU loop (the default) for the headline")
    ap.add_argument("--cpu-worker", default=None, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_worker:
        _cpu_worker(args.cpu_worker)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: there is no multi-GPU path here; the orbit "
              "solves of different months and tracers are independent and would shard "
              "over GPUs with no exchange", file=sys.stderr)
        sys.exit(2)

    import torch
    from orbit_amd import _lib, drivers, serve, synth
    from orbit_amd.core import Context, OperatorSet

    T = args.tracers
    t0 = time.perf_counter()
    wl = synth.benchmark_workload("fct-month", n_tracers=T)
    ctx = Context(0)
    op = OperatorSet.from_workload(wl, ctx)
    grids = drivers.make_fct_grids(wl.grids, ctx)
    if args.mfma:
        op.set_mfma(True)
    log(f"headline: 64x70x15 x {T} tracers, operator set {op.stats()['device_bytes'] / 1e9:.2f} GB "
        f"on device, setup {time.perf_counter() - t0:.1f}s")
    e_dev = torch.from_numpy(wl.e).to("cuda:0")
    if args.calib_copy:
        a = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    def step():
        return drivers.forward_fct(op, grids, e_dev, n_passes=N_PASSES, tol=1e-8, precision="mixed")

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    lat = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t1 = time.perf_counter()
        orbit_mixed, hist = step()
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t1)
    wall = time.perf_counter() - t_start
    ms_per_step = 1e3 * wall / args.steps
    value = T * args.steps / wall
    iters_m = [[i["gmres_iters"] for i in h] for h in hist]
    log(f"mixed + FCT: {ms_per_step:.2f} ms/step, {value:.1f} orbit solves/s; iters per pass {iters_m[0]}..{iters_m[-1]}")

    # one more step of the same work with every half-sweep timed by its own
    # start/stop events (hipExtLaunchKernelGGL on the library's stream; graphs
    # off, so every launch is seen): the kernel durations behind `roofline`
    _lib.check(_lib.lib().orbit_timing_start(ctx.handle, 20000))
    t1 = time.perf_counter()
    step()
    torch.cuda.synchronize()
    t_instr = time.perf_counter() - t1
    tab = (_lib.SweepTiming * 32)()
    nvar, untimed = C.c_int(), C.c_longlong()
    _lib.check(_lib.lib().orbit_timing_stop(ctx.handle, tab, 32, C.byref(nvar), C.byref(untimed)))
    variants = [tab[i] for i in range(min(nvar.value, 32))]
    sweep_ms = sum(v.total_ms for v in variants)
    sweep_bytes = sum(v.total_bytes for v in variants)
    log(f"instrumented step: {t_instr * 1e3:.1f} ms, {sum(v.launches for v in variants)} sweep launches, "
        f"{sweep_ms:.2f} ms of sweep kernels, {sweep_bytes / 1e9:.1f} GB algorithmic, {untimed.value} untimed")

    def timed_variants(fn):
        _lib.check(_lib.lib().orbit_timing_start(ctx.handle, 20000))
        fn()
        torch.cuda.synchronize()
        t = (_lib.SweepTiming * 32)()
        n, u = C.c_int(), C.c_longlong()
        _lib.check(_lib.lib().orbit_timing_stop(ctx.handle, t, 32, C.byref(n), C.byref(u)))
        return {f"prec{t[i].precision}_cs{t[i].slots_per_wave}_mf{t[i].mfma}_oz{t[i].own_zero}_nz{t[i].nbr_zero}"
                + ("_handoff" if t[i].first_sweep == 2 else "_first" if t[i].first_sweep else ""):
                round(1e3 * t[i].total_ms / t[i].launches, 3) for i in range(min(n.value, 32))}

    # the other block product (matrix cores against the VALU loop): the same
    # step with the setting flipped, timed the same way
    mfma_ab = None
    if not args.quick:
        # alternating rounds (5 steps after one untimed step each), medians
        ms = {False: [], True: []}
        for _ in range(3):
            for on in (not args.mfma, args.mfma):
                op.set_mfma(on)
                step()
                torch.cuda.synchronize()
                t1 = time.perf_counter()
                for _ in range(5):
                    step()
                torch.cuda.synchronize()
                ms[on].append(1e3 * (time.perf_counter() - t1) / 5)
        us = {}
        for on in (not args.mfma, args.mfma):
            op.set_mfma(on)
            step()
            us[on] = timed_variants(step)
        op.set_mfma(args.mfma)
        step()
        torch.cuda.synchronize()
        mfma_ab = {("mfma" if on else "valu"): {"ms_per_step_rounds": [round(x, 3) for x in ms[on]],
                                                "ms_per_step_median": round(float(np.median(ms[on])), 3),
                                                "us_per_launch": us[on]} for on in (False, True)}
        log(f"block product A/B: {mfma_ab}")

    # plain FP64 with the same FCT passes: one orbit at a time, no graphs
    npl = min(args.plain_tracers, T)
    drivers.forward_fct(op, grids, e_dev[:1], tracers=[0], n_passes=1, tol=1e-8, precision="fp64")  # warm
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    plain = []
    for s in range(npl):
        o, h = drivers.forward_fct(op, grids, e_dev[s:s + 1], tracers=[s], n_passes=N_PASSES, tol=1e-8,
                                   precision="fp64")
        plain.append((o, h))
    torch.cuda.synchronize()
    t_plain = (time.perf_counter() - t1) / npl
    plain_value = 1.0 / t_plain
    errs = [float((orbit_mixed[s, 0] - plain[s][0][0, 0]).abs().max() / plain[s][0][0, 0].abs().max())
            for s in range(npl)]
    log(f"plain fp64 + FCT: {t_plain * 1e3:.1f} ms/orbit, {plain_value:.2f} orbit solves/s; "
        f"mixed-vs-fp64 c0 rel err {errs}")
    # the same 12 orbits batched in FP64 (no graphs apply to FP64 sweeps): how
    # much of the speed-up is batching and how much precision
    drivers.forward_fct(op, grids, e_dev, n_passes=1, tol=1e-8, precision="fp64")
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    orbit_b64, _ = drivers.forward_fct(op, grids, e_dev, n_passes=N_PASSES, tol=1e-8, precision="fp64")
    torch.cuda.synchronize()
    batched_fp64_value = T / (time.perf_counter() - t1)
    log(f"batched fp64 + FCT: {batched_fp64_value:.1f} orbit solves/s")
    # c0 against the CPU oracle's FCT month (SuperLU + GMRES to 1e-12,
    # tests/golden/fullsize_oracle.npz) for the tracers it holds
    oracle_err = None
    gold_path = os.path.join(ROOT, "tests", "golden", "fullsize_oracle.npz")
    if os.path.exists(gold_path):
        gold = np.load(gold_path, allow_pickle=False)
        oracle_err = {}
        for t in [int(x) for x in gold["tracers"] if int(x) < T]:
            ref = gold[f"forward_c0_{t}"]
            rel = lambda a: float(np.abs(a.cpu().numpy() - ref).max() / np.abs(ref).max())  # noqa: E731
            oracle_err[str(t)] = {"mixed": float(f"{rel(orbit_mixed[t, 0]):.3e}"),
                                  "plain_fp64": float(f"{rel(plain[t][0][0, 0]):.3e}") if t < npl else None,
                                  "batched_fp64": float(f"{rel(orbit_b64[t, 0]):.3e}")}
        log(f"c0 vs the CPU oracle: {oracle_err}")

    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "orbit solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+bf16 (inner sweeps), f64 (GMRES, residuals, refinement, FCT)",
        "data": "synthetic (seeded; random-init operators of the production shape)",
        "config": {"workload": "64x70x15 grid, 12 tracers, 8 bins, dense 15x15 convective columns, "
                               f"FCT deferred correction ({N_PASSES} passes); batched mixed-precision "
                               "periodic forward solve to 1e-8 relative residual",
                   "global_batch": T, "seq_len": 8, "parallelism": "single GPU, 12 orbits batched",
                   "fct_passes": N_PASSES},
        "p50_ms": round(float(np.percentile(np.array(lat) * 1e3, 50)), 3),
        "p99_ms": round(float(np.percentile(np.array(lat) * 1e3, 99)), 3),
        "plain_fp64_value": round(plain_value, 3),
        "speedup_vs_plain_fp64": round(value / plain_value, 2),
        "batched_fp64_value": round(batched_fp64_value, 3),
        "speedup_from_batching": round(batched_fp64_value / plain_value, 2),
        "speedup_from_precision": round(value / batched_fp64_value, 2),
        "rel_err_c0_vs_cpu_oracle": oracle_err,
        "block_product_ab": mfma_ab,
        "max_rel_err_c0_vs_fp64": max(errs),
        "rel_err_c0_vs_fp64_per_tracer": [float(f"{x:.3e}") for x in errs],
        "gmres_iters_mixed_per_pass": iters_m[-1] if iters_m else None,
        # the operator set's calibrated sweeps: relaxation omega of the red-black
        # correction, measured contraction (with its 20% margin), sweep counts
        "sweeps": {k: (round(v, 4) if isinstance(v, float) else v) for k, v in op.stats().items()
                   if k in ("relaxation", "sweep_contraction", "default_acc_sweeps", "default_inner_sweeps",
                            "default_check_sweeps", "mfma")},
    }

    if not args.quick:
        # same without FCT (one periodic solve per orbit)
        op.solve_forward(e_dev, tol=1e-8, precision="mixed")
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(3):
            op.solve_forward(e_dev, tol=1e-8, precision="mixed")
        nofct_mixed = 3 * T / (time.perf_counter() - t1)
        t1 = time.perf_counter()
        for s in range(npl):
            op.solve_forward(e_dev[s:s + 1], tracers=[s], tol=1e-8, precision="fp64")
        nofct_plain = npl / (time.perf_counter() - t1)
        out["no_fct"] = {"mixed_value": round(nofct_mixed, 2), "plain_fp64_value": round(nofct_plain, 2),
                         "speedup": round(nofct_mixed / nofct_plain, 2)}
        log(f"no FCT: mixed {nofct_mixed:.1f}/s, plain {nofct_plain:.1f}/s")

        # adjoint: receptor on 50 surface cells, batched mixed vs FP64 one at a time
        rng = np.random.default_rng(3)
        R = np.zeros((T, 8, wl.N))
        for s in range(T):
            cells = rng.choice(wl.nx * wl.ny, 50, replace=False)
            R[s, 0, cells] = rng.random(50)
        R_dev = torch.from_numpy(R).cuda()
        op.solve_adjoint(R_dev, tol=1e-8, precision="mixed")
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        lam_m, inf_m = op.solve_adjoint(R_dev, tol=1e-8, precision="mixed")
        adj_mixed = T / (time.perf_counter() - t1)
        t1 = time.perf_counter()
        lam_p = []
        for s in range(npl):
            lp, _ = op.solve_adjoint(R_dev[s:s + 1], tracers=[s], tol=1e-8, precision="fp64")
            lam_p.append(lp)
        adj_plain = npl / (time.perf_counter() - t1)
        aerr = [float((lam_m[s, 0] - lam_p[s][0, 0]).abs().max() / lam_p[s][0, 0].abs().max()) for s in range(npl)]
        out["adjoint"] = {"mixed_value": round(adj_mixed, 2), "plain_fp64_value": round(adj_plain, 2),
                          "max_rel_err_lambda0_vs_fp64": max(aerr)}
        log(f"adjoint: mixed {adj_mixed:.1f}/s, plain {adj_plain:.1f}/s, err {max(aerr):.2e}")

        # operator sets built on the device from grid fields (transport assembly,
        # factors, calibration) against the host-CSR path, and the air-mass pre-flight
        from orbit_amd import assembly
        grids_r, D_r, _ = synth.grid_workload(wl.nx, wl.ny, wl.nz, T, seed=[1, 0, 2])
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        op_r = OperatorSet.from_grids(grids_r, D_r, ctx=ctx)
        t_grid = time.perf_counter() - t1
        t1 = time.perf_counter()
        mb = assembly.air_mass_preflight(op_r, grids_r)
        t_mb = time.perf_counter() - t1
        Ts_r = [assembly.assemble_transport_block(g, ctx) for g in grids_r]
        t1 = time.perf_counter()
        op_r2 = OperatorSet.from_shared(Ts_r, D_r, wl.nz, wl.ny, wl.nx, ctx=ctx)
        t_csr = time.perf_counter() - t1
        out["grid_assembly"] = {"opset_from_grids_s": round(t_grid, 3), "opset_from_host_csr_s": round(t_csr, 3),
                                "air_mass_preflight_s": round(t_mb, 3),
                                "interior_p90_per_day": mb["interior_p90"], "top_p90_per_day": mb["top_p90"],
                                "shape": f"{wl.nx}x{wl.ny}x{wl.nz}, {T} tracers, 8 bins"}
        log(f"opset from grids {t_grid:.3f} s (host CSR {t_csr:.3f} s), pre-flight {t_mb:.3f} s, "
            f"interior p90 {mb['interior_p90']:.2e}/day")
        op_r.close()
        op_r2.close()

        # 144 uneven orbits = 12 months x 12 tracers, sizes 1:1:1
        jobs, base_e = [], []
        sizes = ["24x20", "48x40", "64x70"]
        for month in range(12):
            wlm = synth.benchmark_workload("uneven-batch", n_tracers=12, month=month, size=sizes[month % 3])
            opm = OperatorSet.from_workload(wlm, ctx)
            jobs.append((opm, torch.from_numpy(wlm.e).cuda(), list(range(12))))
            base_e.append(wlm.e)
        # one group solve: every launch covers all 144 orbits of the 12 sets
        def group_run():
            t = time.perf_counter()
            serve.solve_group(jobs, precision="mixed", tol=1e-8, ctx=ctx)
            torch.cuda.synchronize()
            return time.perf_counter() - t

        group_run()  # scratch and tables warm
        wall_grp = float(np.median([group_run() for _ in range(5)]))
        ctx_t = Context(0)
        _lib.check(_lib.lib().orbit_timing_start(ctx_t.handle, 20000))
        serve.solve_group(jobs, precision="mixed", tol=1e-8, ctx=ctx_t)
        torch.cuda.synchronize()
        tg = (_lib.SweepTiming * 32)()
        ng, ug = C.c_int(), C.c_longlong()
        _lib.check(_lib.lib().orbit_timing_stop(ctx_t.handle, tg, 32, C.byref(ng), C.byref(ug)))
        gl = [tg[i] for i in range(min(ng.value, 32))]
        grp_sweeps = {"sweep_launches": int(sum(v.launches for v in gl)), "kernel_ms": round(sum(v.total_ms for v in gl), 3),
                      "us_per_launch": {f"prec{v.precision}_oz{v.own_zero}_nz{v.nbr_zero}":
                                        round(1e3 * v.total_ms / v.launches, 3) for v in gl},
                      "achieved_gbs": round(sum(v.total_bytes for v in gl) / (sum(v.total_ms for v in gl) * 1e-3) / 1e9, 1)}
        ctx_t.close()
        log(f"uneven batch, one group solve: 144 orbits in {wall_grp * 1e3:.1f} ms = {144 / wall_grp:.1f}/s; "
            f"{grp_sweeps}")
        # the same as concurrent per-set batches on 4 streams
        serve.solve_many(jobs[:3], n_workers=4)  # graphs and scratch warm
        # each wall time is the median of 3 runs: one 144-orbit run is ~60 ms
        wall_def = float(np.median([serve.solve_many(jobs, n_workers=4)[1] for _ in range(3)]))
        # autotuner: launch shapes per size class, tuned on the first month of each
        tuner = serve.ShapeTuner(reps=40)
        t1 = time.perf_counter()
        for opm, _, _ in jobs:
            tuner.apply(opm)
        t_tune = time.perf_counter() - t1
        serve.solve_many(jobs[:3], n_workers=4)  # graphs of the new shapes
        wall2 = float(np.median([serve.solve_many(jobs, n_workers=4)[1] for _ in range(3)]))
        out["batched_144"] = {"orbits": 144, "wall_ms": round(wall_grp * 1e3, 2),
                              "value": round(144 / wall_grp, 2), "unit": "orbit solves/s",
                              "how": "one group solve (orbit_solve_group_dev): every launch covers all 144 orbits",
                              "group_sweeps": grp_sweeps,
                              "streams_value": round(144 / wall2, 2), "streams_untuned_value": round(144 / wall_def, 2),
                              "streams": 4, "autotune_s": round(t_tune, 3), "shapes": tuner.report(),
                              "sizes": "24x20, 48x40, 64x70 (x15) in 1:1:1"}
        log(f"uneven batch on 4 streams: 144 orbits in {wall2 * 1e3:.1f} ms = {144 / wall2:.1f}/s tuned "
            f"({144 / wall_def:.1f}/s untuned; autotune {t_tune:.2f} s)")

        # serving: Poisson sweep over resident month operator sets
        ops = {(m, sizes[m % 3]): jobs[m][0] for m in range(12)}
        sweeps = []
        srv = serve.OrbitServer(ops, n_workers=4, max_slots=16)  # budget: free HBM (hipMemGetInfo)
        try:
            for rate, nq in ((5.0, 25), (50.0, 60), (200.0, 120)):
                rq = np.random.default_rng(int(rate))

                def make(r):
                    m = int(r.integers(0, 12))
                    key = (m, sizes[m % 3])
                    opk = ops[key]
                    nt = int(r.integers(1, 13))
                    tr = sorted(r.choice(12, nt, replace=False).tolist())
                    mode = "forward" if r.random() < 0.5 else "adjoint"
                    return serve.make_query(opk, m, key[1], tr, mode, r, base_e=base_e[m], tol=1e-6)

                rep = serve.run_poisson(srv, rate, nq, make, rq)
                sweeps.append({k: (round(v, 3) if isinstance(v, float) else v) for k, v in rep.items()})
                log(f"serving at {rate}/s: p50 {rep['p50_ms']:.1f} ms, p99 {rep['p99_ms']:.1f} ms, "
                    f"{rep['solves_per_s']:.1f} solves/s, mean batch {rep['mean_batch']:.1f}")
        finally:
            srv.close()
        out["serving"] = sweeps
        out["serving_admission"] = {"budget_gb": round(srv.budget / 1e9, 2), "source": srv.budget_source,
                                    "free_at_start_gb": round(srv.free_at_start / 1e9, 2)}

    # roofline: the half-sweep variant with the most device time in the
    # instrumented headline step; achieved = its algorithmic bytes / its
    # launches' own durations (start/stop events around each launch)
    def vname(v):
        base = {_lib.ORBIT_PREC_FP64: "fp64", _lib.ORBIT_SWEEP_ACC_MIXED: "accurate",
                _lib.ORBIT_PREC_MIXED: "inner"}[v.precision]
        if v.first_sweep == 2:
            return base + "_handoff"
        if v.first_sweep:
            return base + "_first"
        return base + ("_ownzero" if v.own_zero else "") + ("_nbrzero" if v.nbr_zero else "")

    def kname(v):
        xt = {_lib.ORBIT_PREC_FP64: "double,double,double,double", _lib.ORBIT_SWEEP_ACC_MIXED:
              "double,double,unsigned short,double", _lib.ORBIT_PREC_MIXED: "float,float,unsigned short,float"}
        if v.first_sweep == 2:
            return (f"k_first_sweep, hand-off form (the previous bin's last black half-sweep and both colours of "
                    f"this bin's first sweep from x = 0 in one launch over column patches, shape {op.stats()['patch_shape']})")
        if v.first_sweep:
            return (f"k_first_sweep (both colours of the first sweep from x = 0 in one launch over column patches, "
                    f"shape {op.stats()['patch_shape']})")
        return (f"k_halfsweep<{xt[v.precision]},{1 if v.slots_per_wave == 4 else 0},"
                f"{'true' if v.own_zero else 'false'},{'true' if v.nbr_zero else 'false'},"
                f"{'true' if v.mfma else 'false'},false>")

    def roof(v):
        us = 1e3 * v.total_ms / v.launches
        bpl = v.total_bytes / v.launches
        achieved = bpl / (us * 1e-6) / 1e9
        tr = from_profile(vname(v))
        prof_us = from_profile(vname(v), "rocprof_us_per_launch")
        return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": tr,
                "kernel": f"{kname(v)} ({T} slots, 64x70x15, one colour)", "variant": vname(v),
                "us_per_launch": round(us, 3), "algorithmic_bytes_per_launch": round(bpl),
                "launches_per_step": int(v.launches), "ms_per_step": round(v.total_ms, 3),
                "traffic_over_algorithmic": round(tr / bpl, 3) if tr else None,
                "rocprof_us_per_launch": prof_us,
                "frac_from_rocprof_duration": round(bpl / (prof_us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4) if prof_us else None,
                "timing": "kernel start/stop events (hipExtLaunchKernelGGL) on the library's stream over one "
                          "headline step (the timed work; graphs off so every launch is timed)",
                "residency": "a bin's working set (40-60 MB) fits the 256 MB Infinity Cache: bytes leaving L2, "
                             "not necessarily HBM"}

    out["roofline"] = roof(variants[0])
    out["roofline_other"] = [roof(v) for v in variants[1:4]]
    out["step_sweeps"] = {"sweep_launches": int(sum(v.launches for v in variants)),
                          "kernel_ms": round(sweep_ms, 3), "algorithmic_gb": round(sweep_bytes / 1e9, 3),
                          "achieved_gbs_over_step": round(sweep_bytes / (ms_per_step * 1e-3) / 1e9, 1),
                          "frac_over_step": round(sweep_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    # the isolated figure: `roof_reps` back-to-back launches on one bin (HIP
    # events around the whole run, gaps between dispatches included)
    iso = {}
    for name, prec in (("inner", _lib.ORBIT_PREC_MIXED), ("accurate", _lib.ORBIT_SWEEP_ACC_MIXED)):
        ms = C.c_double()
        by = C.c_double()
        _lib.check(_lib.lib().orbit_bench_halfsweep(ctx.handle, op.handle, 0, prec, T, args.roof_reps,
                                                    C.byref(ms), C.byref(by)))
        iso[name] = {"us_per_launch": round(ms.value * 1e3, 3), "algorithmic_bytes": by.value,
                     "achieved_gbs": round(by.value / (ms.value * 1e-3) / 1e9, 1)}
    out["isolated_back_to_back"] = iso
    log(f"roofline {out['roofline']['variant']}: {out['roofline']['us_per_launch']} us/launch, "
        f"{out['roofline']['achieved']} GB/s ({out['roofline']['frac']}); isolated {iso}")

    cpu = None
    if not args.no_cpu and not args.quick:
        del orbit_mixed, plain, orbit_b64
        cpu = cpu_baseline(T)
        log(f"cpu baseline: {cpu['value']:.4f} orbit solves/s on {cpu['cores']} cores (solve only "
            f"{cpu['solve_only_value']:.4f}); one core {cpu['single_core_value']:.4f}; config 0 "
            f"{cpu['config0_single_core']['orbit_s']} s per orbit")
    out["cpu_baseline"] = cpu
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
