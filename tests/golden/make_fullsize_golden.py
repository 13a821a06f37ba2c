"""Write tests/golden/fullsize_oracle.npz: the CPU oracle's answers at the
production shape (64x70x15, dense convective columns), for the parity tests
of the benchmark configurations that the oracle cannot run in seconds.

* forward: c0 of the FCT month (benchmark workload "fct-month"), tracers 0
  and 7, 6 deferred-correction passes (orbit/cli.py:1212-1241) solved by the
  oracle's SuperLU factors and scipy GMRES(30) to tol 1e-12
  (oracle.drivers_ref.forward_fct_passes).
* adjoint: lambda0 of the benchmark workload "adjoint", tracers 0 and 7, for a
  receptor on 50 seeded surface cells with U(0,1) weights in bin 0's state
  (R_0 = r, R_tau>0 = 0, so lambda0 = d(r . c0)/ds), to tol 1e-12
  (oracle.orbit_ref.solve_orbit_adjoint_one_species).

Beside the vectors the file keeps fingerprints of the inputs (sums of e, of
the receptors and of the operator diagonals), so a change of the synthetic
generator that would silently change the inputs fails a CPU test instead.
Runs the oracle only (no reference code); about 3 minutes on one core.

    python tests/golden/make_fullsize_golden.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import drivers_ref, orbit_ref  # noqa: E402
from orbit_amd import synth  # noqa: E402

TRACERS = (0, 7)
N_PASSES = 6
RECEPTOR_SEED = 303


def receptors(wl, seed=RECEPTOR_SEED):
    """R (T, 8, N): 50 random surface cells per tracer with U(0,1) weights in bin 0."""
    rng = np.random.default_rng(seed)
    ncol = wl.nx * wl.ny
    R = np.zeros((wl.n_tracers, 8, wl.N))
    for t in range(wl.n_tracers):
        cells = rng.choice(ncol, 50, replace=False)
        R[t, 0, cells] = rng.random(50)
    return R


def fingerprint(wl):
    return np.array([wl.e.sum(), np.abs(wl.e).max(), wl.D.sum(), sum(float(T.sum()) for T in wl.T)])


def main():
    out = {"tracers": np.array(TRACERS), "n_passes": np.array(N_PASSES), "tol": np.array(1e-12)}
    wl = synth.benchmark_workload("fct-month")
    out["forward_fingerprint"] = fingerprint(wl)
    for t in TRACERS:
        t0 = time.perf_counter()
        lus = [orbit_ref.factor_one_bin(wl.L(b, t)) for b in range(8)]
        orbit, hist = drivers_ref.forward_fct_passes(lus, list(wl.e[t]), wl.grids, n_passes=N_PASSES, tol=1e-12,
                                                     maxiter=500)
        out[f"forward_c0_{t}"] = np.asarray(orbit[0])
        out[f"forward_c8_{t}"] = np.asarray(orbit[8])
        out[f"forward_residual_{t}"] = np.array([h["residual"] for h in hist])
        print(f"forward tracer {t}: {time.perf_counter() - t0:.1f} s, residuals {out[f'forward_residual_{t}']}",
              flush=True)
        del lus
    wl = synth.benchmark_workload("adjoint")
    R = receptors(wl)
    out["adjoint_fingerprint"] = np.concatenate([fingerprint(wl), [R.sum()]])
    for t in TRACERS:
        t0 = time.perf_counter()
        lus = [orbit_ref.factor_one_bin(wl.L(b, t)) for b in range(8)]
        lam, info = orbit_ref.solve_orbit_adjoint_one_species(lus, list(R[t]), tol=1e-12, maxiter=500)
        out[f"adjoint_lam0_{t}"] = np.asarray(lam[0])
        out[f"adjoint_residual_{t}"] = np.array(info["residual"])
        print(f"adjoint tracer {t}: {time.perf_counter() - t0:.1f} s, residual {info['residual']:.2e}", flush=True)
        del lus
    path = os.path.join(HERE, "fullsize_oracle.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path) / 1e6:.2f} MB)")


if __name__ == "__main__":
    main()
