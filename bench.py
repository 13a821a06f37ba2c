"""Benchmark: periodic-orbit solves/s at 64x70x15 with 12 tracers on one MI355X.

Headline (the production configuration, FCT on): one step = the forward
periodic orbits of the 12 tracers of a month's operator set with the horizontal
FCT deferred correction (6 passes, each a periodic solve whose RHS carries the
anti-diffusive source of the previous pass's end-of-bin orbit), solved as one
batch in mixed precision (FP32 red-black column sweeps with BF16 column blocks
inside FP64 GMRES, FP64 refinement to 1e-8 relative residual), inputs resident
in HBM.  value = orbit solves/s = 12 / step time.

Beside it: the plain FP64 GPU solve of the same work (same algorithm in FP64,
one orbit at a time, no batching, no graphs), the per-tracer c0 difference,
p50/p99 per-step latency, the same without FCT, the adjoint at the same size
(50-cell receptors), 144 uneven batched orbits (12 months x 12 tracers over
24x20 / 48x40 / 64x70), a Poisson serving sweep over those months, the
roofline of the dominant kernel from live HIP-event timing and a CPU baseline
(the oracle: SuperLU factors + scipy GMRES) on a bounded sample.

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


def cpu_baseline(wl, iters_per_pass):
    """Oracle on the host, one core: SuperLU factors of bins 0-1 of tracer 0, 16
    back-solves and the FCT source of one bin, extrapolated to one FCT-on orbit
    (8 factors; per pass 8*(iters+4) solves and 8 FCT sources)."""
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    from oracle import fct_ref, orbit_ref as orb
    t0 = time.perf_counter()
    lus = [orb.factor_one_bin(wl.L(b, 0)) for b in range(2)]
    t_fac = (time.perf_counter() - t0) / 2
    rng = np.random.default_rng(0)
    vs = [rng.random(wl.N) for _ in range(16)]
    t0 = time.perf_counter()
    for k, v in enumerate(vs):
        lus[k % 2].solve(v)
    t_sol = (time.perf_counter() - t0) / 16
    c = rng.random(wl.N)
    t0 = time.perf_counter()
    fct_ref.compute_horizontal_fct_source(wl.grids[0], c, wl.dtau, True, None)
    t_fct = time.perf_counter() - t0
    n_solves = sum(8 * (it + 4) for it in iters_per_pass)
    n_pass = len(iters_per_pass)
    t_orbit = 8 * t_fac + n_solves * t_sol + 8 * n_pass * t_fct
    return {
        "value": 1.0 / t_orbit, "unit": "orbit solves/s", "cores": 1, "kind": "port",
        "solve_only_value": 1.0 / (n_solves * t_sol),
        "sample": (f"oracle (SuperLU + numpy FCT) on tracer 0 at 64x70x15: factor of 2 of 8 bins "
                   f"({t_fac:.2f} s each), 16 back-solves ({t_sol * 1e3:.1f} ms each), FCT source of "
                   f"one bin ({t_fct * 1e3:.1f} ms); extrapolated to one orbit with {n_pass} FCT passes: "
                   f"8 factors + {n_solves} solves (GMRES iterations per pass of the FP64 GPU solve "
                   f"{iters_per_pass}, +4 for s, restart residual and recovery) + {8 * n_pass} FCT "
                   f"sources; 1 core"),
    }


def traffic_from_profile(name):
    """HBM bytes per launch of a sweep kernel from the committed rocprofv3 PMC
    summary (FETCH_SIZE + WRITE_SIZE scaled by a calibration copy), or None."""
    path = os.path.join(ROOT, "profiles", "halfsweep_64x70x15_12tracers.md")
    if not os.path.exists(path):
        return None
    m = re.search(name + r"_traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(path).read())
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
    args = ap.parse_args()
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
        "max_rel_err_c0_vs_fp64": max(errs),
        "rel_err_c0_vs_fp64_per_tracer": [float(f"{x:.3e}") for x in errs],
        "gmres_iters_mixed_per_pass": iters_m[-1] if iters_m else None,
        # the operator set's calibrated sweeps: relaxation omega of the red-black
        # correction, measured contraction (with its 20% margin), sweep counts
        "sweeps": {k: (round(v, 4) if isinstance(v, float) else v) for k, v in op.stats().items()
                   if k in ("relaxation", "sweep_contraction", "default_acc_sweeps", "default_inner_sweeps")},
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
        out["batched_144"] = {"orbits": 144, "wall_ms": round(wall2 * 1e3, 2),
                              "value": round(144 / wall2, 2), "unit": "orbit solves/s",
                              "untuned_value": round(144 / wall_def, 2),
                              "autotune_s": round(t_tune, 3), "shapes": tuner.report(),
                              "sizes": "24x20, 48x40, 64x70 (x15) in 1:1:1", "streams": 4}
        log(f"uneven batch: 144 orbits in {wall2 * 1e3:.1f} ms = {144 / wall2:.1f}/s tuned "
            f"({144 / wall_def:.1f}/s untuned; autotune {t_tune:.2f} s)")

        # serving: Poisson sweep over resident month operator sets
        ops = {(m, sizes[m % 3]): jobs[m][0] for m in range(12)}
        sweeps = []
        srv = serve.OrbitServer(ops, n_workers=4, max_slots=16, hbm_budget=64e9)
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

    # roofline of the two sweep kernels of the mixed solve (all slots, bin 0), live
    # HIP events on the context's stream: the FP32/BF16 inner sweep and the
    # FP64/BF16 accurate sweep; the dominant one (more device time per step in the
    # kernel-trace profile of this run) is reported as `roofline`.
    roofs = {}
    for name, prec, kern in (("inner", _lib.ORBIT_PREC_MIXED, "k_halfsweep<float,float,bf16,float>"),
                             ("accurate", _lib.ORBIT_SWEEP_ACC_MIXED, "k_halfsweep<double,double,bf16,double>")):
        ms = C.c_double()
        by = C.c_double()
        _lib.check(_lib.lib().orbit_bench_halfsweep(ctx.handle, op.handle, 0, prec, T, args.roof_reps,
                                                    C.byref(ms), C.byref(by)))
        achieved = by.value / (ms.value * 1e-3) / 1e9
        log(f"halfsweep {name}: {ms.value * 1e3:.2f} us/launch, {by.value / 1e6:.2f} MB, {achieved:.0f} GB/s")
        roofs[name] = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profile(name),
                       "kernel": f"{kern} ({T} slots, 64x70x15, one colour)",
                       "us_per_launch": round(ms.value * 1e3, 3), "algorithmic_bytes_per_launch": by.value}
    out["roofline"] = roofs[DOMINANT]
    out["roofline_other"] = roofs["inner" if DOMINANT == "accurate" else "accurate"]

    cpu = None
    if not args.no_cpu:
        ipp = [int(h[0]["gmres_iters"]) for h in plain[0][1]] if plain else [6] * N_PASSES
        cpu = cpu_baseline(wl, ipp)
        log(f"cpu baseline: {cpu['value']:.4f} orbit solves/s (solve only {cpu['solve_only_value']:.3f})")
    out["cpu_baseline"] = cpu
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
