"""Benchmark: periodic-orbit solves/s at 64x70x15 with 12 tracers on one MI355X.

One step = one batched mixed-precision periodic forward solve of the 12
tracers of a month's operator set (FP32 red-black column sweeps with BF16
column blocks inside FP64 GMRES, FP64 refinement to 1e-8 relative residual),
inputs resident in HBM.  value = orbit solves/s = 12 / step time.

Also reported: the plain FP64 GPU solve (same algorithm in FP64, one orbit at
a time, no batching, no graphs), the max relative c0 difference of the mixed
solve from it per tracer, p50/p99 per-step latency, the roofline of the
dominant kernel (red-black half-sweep) from live HIP-event timing, and a CPU
baseline (the oracle: SuperLU factors + scipy GMRES) on a bounded sample.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "orbit solves/s (64×70×15, 12 tracers) vs plain FP64 GPU; p50/p99 latency"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(wl, iters_hint):
    """Oracle on the host: SuperLU factor of bins 0-1 of tracer 0 and 16 back-solves,
    extrapolated to 8 bins and 8*(iters+4) solves of one periodic solve."""
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    from oracle import orbit_ref as orb
    t0 = time.perf_counter()
    lus = [orb.factor_one_bin(wl.L(b, 0)) for b in range(2)]
    t_fac = (time.perf_counter() - t0) / 2
    rng = np.random.default_rng(0)
    vs = [rng.random(wl.N) for _ in range(16)]
    t0 = time.perf_counter()
    for k, v in enumerate(vs):
        lus[k % 2].solve(v)
    t_sol = (time.perf_counter() - t0) / 16
    n_solves = 8 * (iters_hint + 4)
    t_orbit = 8 * t_fac + n_solves * t_sol
    return {
        "value": 1.0 / t_orbit, "unit": "orbit solves/s", "cores": 1, "kind": "port",
        "solve_only_value": 1.0 / (n_solves * t_sol),
        "sample": (f"SuperLU factor of 2 of 8 bins ({t_fac:.2f} s each) and 16 back-solves "
                   f"({t_sol * 1e3:.1f} ms each) of tracer 0 at 64x70x15, extrapolated to "
                   f"8 factors + {n_solves} solves per orbit (GMRES(30) iterations of the FP64 "
                   f"solve + s, restart residual and recovery sweeps)"),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tracers", type=int, default=12)
    ap.add_argument("--plain-tracers", type=int, default=3,
                    help="tracers timed one at a time with the plain FP64 solve")
    ap.add_argument("--roof-reps", type=int, default=400,
                    help="back-to-back half-sweep launches timed for the roofline")
    ap.add_argument("--calib-copy", action="store_true",
                    help="run one 256 MiB device copy first (scales profiler byte counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(orbits shard with no exchange; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import torch
    from orbit_amd import synth
    from orbit_amd import _lib
    from orbit_amd.core import Context, OperatorSet
    import ctypes as C

    t0 = time.perf_counter()
    wl = synth.config_workload(1, n_tracers=args.tracers)
    log(f"workload 64x70x15 x {args.tracers} tracers generated in {time.perf_counter() - t0:.1f}s")
    ctx = Context(0)
    t0 = time.perf_counter()
    op = OperatorSet.from_workload(wl, ctx)
    log(f"operator set on device in {time.perf_counter() - t0:.2f}s, "
        f"{op.stats()['device_bytes'] / 1e9:.2f} GB")
    e_dev = torch.from_numpy(wl.e).to("cuda:0")
    T = args.tracers
    if args.calib_copy:
        a = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    def step():
        return op.solve_forward(e_dev, tol=1e-8, precision="mixed")

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    lat = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t1 = time.perf_counter()
        orbit_mixed, info_mixed = step()
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t1)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    ms_per_step = 1e3 * wall / args.steps
    value = T * args.steps / wall
    log(f"mixed: {ms_per_step:.2f} ms/step, {value:.1f} orbit solves/s; iters "
        f"{[i['gmres_iters'] for i in info_mixed]} rounds {[i['rounds'] for i in info_mixed]}")

    # plain FP64: one orbit at a time, no graphs
    npl = min(args.plain_tracers, T)
    op.solve_forward(e_dev[:1], tracers=[0], tol=1e-8, precision="fp64")  # warm
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    plain = []
    for s in range(npl):
        o, inf = op.solve_forward(e_dev[s:s + 1], tracers=[s], tol=1e-8, precision="fp64")
        plain.append((o, inf[0]))
    torch.cuda.synchronize()
    t_plain = (time.perf_counter() - t1) / npl
    plain_value = 1.0 / t_plain
    errs = []
    for s in range(npl):
        a = orbit_mixed[s, 0].double()
        b = plain[s][0][0, 0].double()
        errs.append(float((a - b).abs().max() / b.abs().max()))
    log(f"plain fp64: {t_plain * 1e3:.1f} ms/orbit, {plain_value:.2f} orbit solves/s; "
        f"iters {[p[1]['gmres_iters'] for p in plain]}; mixed-vs-fp64 c0 rel err {errs}")

    # roofline of the dominant kernel: mixed half-sweep on all slots, live HIP events
    ms = C.c_double()
    by = C.c_double()
    _lib.check(_lib.lib().orbit_bench_halfsweep(ctx.handle, op.handle, 0, _lib.ORBIT_PREC_MIXED, T, args.roof_reps,
                                                C.byref(ms), C.byref(by)))
    achieved = by.value / (ms.value * 1e-3) / 1e9
    log(f"halfsweep mixed: {ms.value * 1e3:.2f} us/launch, {by.value / 1e6:.2f} MB, {achieved:.0f} GB/s")
    ms64 = C.c_double()
    by64 = C.c_double()
    _lib.check(_lib.lib().orbit_bench_halfsweep(ctx.handle, op.handle, 0, _lib.ORBIT_PREC_FP64, 1, args.roof_reps,
                                                C.byref(ms64), C.byref(by64)))
    log(f"halfsweep fp64 1 slot: {ms64.value * 1e3:.2f} us/launch, "
        f"{by64.value / (ms64.value * 1e-3) / 1e9:.0f} GB/s")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(wl, int(plain[0][1]["gmres_iters"]))
        log(f"cpu baseline: {cpu['value']:.4f} orbit solves/s (solve only {cpu['solve_only_value']:.3f})")

    lat_ms = np.array(lat) * 1e3
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "orbit solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32/bf16 inner sweeps, f64 GMRES+refinement", "data": "synthetic (seeded config 1)",
        "config": {"workload": "64x70x15 grid, 12 tracers, 8 bins, dense 15x15 convective columns; "
                               "batched mixed-precision periodic forward solve to 1e-8 rel. residual",
                   "global_batch": T, "seq_len": 8, "parallelism": "single GPU, 12 orbits batched",
                   "fct_passes": 0},
        "p50_ms": round(float(np.percentile(lat_ms, 50)), 3),
        "p99_ms": round(float(np.percentile(lat_ms, 99)), 3),
        "plain_fp64_value": round(plain_value, 3),
        "speedup_vs_plain_fp64": round(value / plain_value, 2),
        "max_rel_err_c0_vs_fp64": max(errs),
        "gmres_iters_mixed": [int(i["gmres_iters"]) for i in info_mixed],
        "residual_mixed_max": max(float(i["residual"]) for i in info_mixed),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "k_halfsweep<float,float,bf16,float>",
                     "us_per_launch": round(ms.value * 1e3, 3),
                     "algorithmic_bytes_per_launch": by.value},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
