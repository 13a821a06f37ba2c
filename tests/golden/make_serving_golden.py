"""Write tests/golden/serving_oracle.npz: the CPU oracle's answers to the
served queries that the serving test pins (the 24x20 and 48x40 month sets of
the served mix; a forward-marginal and an adjoint query of three tracers
each), solved by the oracle's SuperLU factors and scipy GMRES(30) to tol
1e-12 (oracle.orbit_ref).  The test rebuilds the same seeded operator sets and
query inputs (orbit_amd.synth, orbit_amd.serve.query_payload) and serves them
among other queries; the file keeps the inputs' sums so a change of the
generator fails loudly.  Oracle only (no reference code); about 2 minutes.

    python tests/golden/make_serving_golden.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import orbit_ref  # noqa: E402
from orbit_amd import serve, synth  # noqa: E402

PINNED = ((0, "24x20"), (1, "48x40"))   # (month, size class)
TRACERS = (2, 5, 11)
SEEDS = {"forward": 40, "adjoint": 50}


def pinned_payload(wl, month, mode):
    rng = np.random.default_rng([SEEDS[mode], month])
    return serve.query_payload(8, wl.N, wl.nx * wl.ny, TRACERS, mode, rng, base_e=wl.e)


def main():
    out = {"tracers": np.array(TRACERS), "months": np.array([m for m, _ in PINNED])}
    for month, size in PINNED:
        wl = synth.benchmark_workload("serving", n_tracers=12, month=month, size=size)
        pays = {mode: pinned_payload(wl, month, mode) for mode in SEEDS}
        for mode, p in pays.items():
            out[f"{mode}_{month}_payload_sum"] = np.array(p.sum())
        res = {mode: [] for mode in SEEDS}
        t0 = time.perf_counter()
        for i, t in enumerate(TRACERS):
            lus = [orbit_ref.factor_one_bin(wl.L(b, t)) for b in range(8)]
            fo, fi = orbit_ref.solve_orbit_one_species(lus, list(pays["forward"][i]), tol=1e-12, maxiter=500)
            ao, ai = orbit_ref.solve_orbit_adjoint_one_species(lus, list(pays["adjoint"][i]), tol=1e-12, maxiter=500)
            assert fi["converged"] and ai["converged"], (fi, ai)
            res["forward"].append(fo[0])
            res["adjoint"].append(ao[0])
        for mode in SEEDS:
            out[f"{mode}_{month}_x0"] = np.stack(res[mode])
        print(f"month {month} ({size}): {time.perf_counter() - t0:.1f} s", flush=True)
    path = os.path.join(HERE, "serving_oracle.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path) / 1e6:.2f} MB)")


if __name__ == "__main__":
    main()
