"""CPU restatement of the species layer around the periodic-orbit solve (TEST ORACLE).

Test infrastructure only: the product never imports this module.

  build_solve_waves       orbit/core/orbit.py:397-465 (cyclic edges dropped for
                          placement; wave = 1 + max wave of the remaining sources,
                          iterated to a fixed point; skipped species in wave 0)
  species_rhs             orbit/core/orbit.py:249-312 (e_s - sum_src K_(s,src),tau
                          c_src[tau], or c_src[tau+1] for endbin keys, + extra RHS;
                          KeyError when a source orbit is missing)
  solve_all_species       orbit/core/orbit.py:468-806, wave-barrier path (orbits
                          snapshot at each wave start, seed orbits, copy-verbatim
                          skip, zero-RHS short-circuit)
  picard_cyclic           orbit/modes/marginal.py:573-612 (zero seeds, repeat until
                          the max relative change of the cyclic orbits < rel_tol)
  adjoint_plan            orbit/modes/adjoint.py:49-69, 496-515
  adjoint_species_passes  orbit/modes/adjoint.py:634-783 (receptor R_src -=
                          dtau K^T (lam_r - R_r) from solved receivers; reverse-DAG
                          first pass, Picard on the cyclic pair, refresh of the
                          waves after it)

Pinned by tests/test_oracle_species.py: the reference's own tests of this layer
(tests/test_simulation_modes.py:1158-1580: wave placement, the cyclic KeyError,
seeded one-way pass, Picard to the closed-form NH/NO3 fixed point, endbin
convention) restated, a direct sparse solve of the joint periodic system of a
coupled pair, and the adjoint passes' gradient against the forward layer.
"""
from __future__ import annotations

import numpy as np

from . import orbit_ref

N_BINS = orbit_ref.N_BINS
DTAU = orbit_ref.DTAU

# orbit/core/deposition.py:57-77 and orbit/core/orbit.py:70
IDX_TOTAL_NH, IDX_SO2, IDX_NOX, IDX_PSO4, IDX_TOTAL_NO3 = 2, 3, 4, 5, 6
N_SPECIES = 14
REF_SOLVE_ORDER = (0, 1, 13, 2, 3, 4, 5, 6, 7, 8, 12, 9, 10, 11)


def sources_of(K_sources_bin0):
    out = {}
    for (tgt, src) in K_sources_bin0:
        out.setdefault(tgt, []).append(src)
    return out


def build_solve_waves(sources_for_target, solve_order, skip_species):
    cyclic = {(t, s) for t, ss in sources_for_target.items() for s in ss
              if t in sources_for_target.get(s, ())}
    deps = {t: [s for s in ss if (t, s) not in cyclic] for t, ss in sources_for_target.items()}
    wave = {}
    for _ in range(len(solve_order) + 2):
        moved = False
        for s in solve_order:
            if s in skip_species or not deps.get(s):
                w = 0
            else:
                w = 1 + max(wave.get(d, 0) for d in deps[s])
            if wave.get(s) != w:
                wave[s] = w
                moved = True
        if not moved:
            break
    out = [[] for _ in range(1 + max(wave.values(), default=0))]
    for s in solve_order:
        out[wave[s]].append(s)
    return out


def _base_emissions(emissions_SN, s, N):
    e = np.asarray(emissions_SN, dtype=np.float64)
    if e.ndim == 1:
        return [e[s * N:(s + 1) * N].copy() for _ in range(N_BINS)]
    return [e[t, s * N:(s + 1) * N].copy() for t in range(N_BINS)]


def species_rhs(s, emissions_SN, N, K_sources_per_bin, sources, orbits, extra_rhs_s=None,
                endbin_keys=None):
    e = _base_emissions(emissions_SN, s, N)
    for t in range(N_BINS):
        for src in sources:
            if src not in orbits:
                raise KeyError(f"orbits[{src}] missing for target s={s} at tau={t}: "
                               "Seed-orbits machinery is needed for cyclic source DAGs")
            lvl = t + 1 if endbin_keys and (s, src) in endbin_keys else t
            e[t] = e[t] - K_sources_per_bin[t][(s, src)] @ orbits[src][lvl]
        if extra_rhs_s is not None:
            e[t] = e[t] + np.asarray(extra_rhs_s[t], dtype=np.float64)
    return e


def solve_all_species(lus, K_sources_per_bin, emissions_SN, N, tol=1e-8, maxiter=200,
                      c_warm_SN=None, skip_solve_species=None, cached_orbits=None,
                      extra_rhs_per_bin_per_species=None, seed_orbits=None, endbin_keys=None,
                      solve_order=REF_SOLVE_ORDER, n_species=N_SPECIES):
    """lus: {species: [8 factors]} (the LU cache)."""
    skip = set(skip_solve_species or ())
    orbits = {s: [c.copy() for c in o] for s, o in (seed_orbits or {}).items()}
    iters = np.zeros(n_species, dtype=int)
    resid = np.zeros(n_species)
    period = np.zeros(n_species)
    srcs = sources_of(K_sources_per_bin[0])
    for wave in build_solve_waves(srcs, list(solve_order), skip):
        snap = dict(orbits)
        done = {}
        for s in wave:
            if s in skip and cached_orbits is not None and s in cached_orbits:
                done[s] = ([c.copy() for c in cached_orbits[s]], 0, 0.0, 0.0)
                continue
            extra = (extra_rhs_per_bin_per_species or {}).get(s)
            e = species_rhs(s, emissions_SN, N, K_sources_per_bin, srcs.get(s, []), snap, extra,
                            endbin_keys)
            if all(not np.any(x) for x in e):
                done[s] = ([np.zeros(N) for _ in range(N_BINS + 1)], 0, 0.0, 0.0)
                continue
            warm = None if c_warm_SN is None else np.asarray(c_warm_SN)[s * N:(s + 1) * N].copy()
            orbit, info = orbit_ref.solve_orbit_one_species(lus[s], e, DTAU, c_warm=warm, tol=tol,
                                                            maxiter=maxiter)
            done[s] = (orbit, info["gmres_iters"], info["residual"], info["periodicity"])
        for s, (orbit, it, r, p) in done.items():
            orbits[s] = orbit
            iters[s], resid[s], period[s] = it, r, p
    return {"orbits": orbits, "gmres_iters": iters, "gmres_resid": resid, "periodicity": period}


def _max_rel_change(prev, curr):
    a = np.stack(prev[:N_BINS])
    b = np.stack(curr[:N_BINS])
    den = float(np.abs(b).max())
    d = float(np.abs(b - a).max())
    return d if den < 1e-30 else d / den


def picard_cyclic(lus, K_sources_per_bin, emissions_SN, N, cyclic_species, max_iters=30,
                  rel_tol=1e-8, **kw):
    """Returns (result, iterations, last max relative change, converged)."""
    seed = {s: [np.zeros(N) for _ in range(N_BINS + 1)] for s in cyclic_species}
    res, change = None, 0.0
    for it in range(max_iters):
        res = solve_all_species(lus, K_sources_per_bin, emissions_SN, N, seed_orbits=seed, **kw)
        change = max(_max_rel_change(seed[s], res["orbits"][s]) for s in cyclic_species)
        if it > 0 and change < rel_tol:
            return res, it + 1, change, True
        seed = {s: [c.copy() for c in res["orbits"][s]] for s in cyclic_species}
    return res, max_iters, change, False


def adjoint_plan(K_sources_bin0, user_species, solve_order=REF_SOLVE_ORDER):
    """(receivers_for_source, to_solve, waves, adjoint order)."""
    recv = {}
    for (tgt, src) in K_sources_bin0:
        recv.setdefault(src, []).append(tgt)
    to_solve = set(user_species)
    stack = list(user_species)
    while stack:
        s = stack.pop()
        for r in recv.get(s, []):
            if r not in to_solve:
                to_solve.add(r)
                stack.append(r)
    waves = build_solve_waves(recv, list(solve_order), set())
    order = [s for w in waves for s in w if s in to_solve]
    return recv, to_solve, waves, order


def adjoint_species_passes(lus, R_base, K_sources_per_bin, user_species, endbin_keys=(),
                           tol=1e-8, maxiter=200, picard_max_iters=10, picard_rel_tol=1e-4,
                           cyclic_species=(IDX_TOTAL_NH, IDX_TOTAL_NO3), extra_recep=None,
                           solve_order=REF_SOLVE_ORDER):
    """R_base: {species: [8 receptors]}.  Returns (lam, R_eff) per species; the
    gradient is dJ/de_(s,tau) = dtau (lam_s[tau] - R_eff_s[tau])."""
    recv, to_solve, waves, order = adjoint_plan(K_sources_per_bin[0], user_species, solve_order)
    endbin = set(endbin_keys or ())
    lam, Reff = {}, {}

    def receptor(s):
        N = lus[s][0].shape[0]
        R = [np.asarray(R_base[s][t], dtype=np.float64).copy() if s in R_base else np.zeros(N)
             for t in range(N_BINS)]
        for r in recv.get(s, []):
            if r not in lam:
                continue
            for t in range(N_BINS):
                K = K_sources_per_bin[t].get((r, s))
                if K is None:
                    continue
                tb = (t + 1) % N_BINS if (r, s) in endbin else t
                R[tb] = R[tb] - DTAU * (K.T @ (lam[r][t] - Reff[r][t]))
        if extra_recep is not None and s in extra_recep:
            for t in range(N_BINS):
                R[t] = R[t] + extra_recep[s][t]
        return R

    def solve(s):
        R = receptor(s)
        out, _ = orbit_ref.solve_orbit_adjoint_one_species(lus[s], R, tol=tol, maxiter=maxiter)
        lam[s], Reff[s] = out, R

    for s in order:
        solve(s)
    K0 = K_sources_per_bin[0]
    a, b = cyclic_species
    active = a in lam and b in lam and ((a, b) in K0 or (b, a) in K0)
    if active:
        for _ in range(picard_max_iters):
            prev = {s: [x.copy() for x in lam[s]] for s in cyclic_species}
            for s in cyclic_species:
                solve(s)
            change = 0.0
            for s in cyclic_species:
                for t in range(N_BINS):
                    d = float(np.linalg.norm(lam[s][t] - prev[s][t]))
                    change = max(change, d / (float(np.linalg.norm(lam[s][t])) + 1e-30))
            if change < picard_rel_tol:
                break
        w_cyc = next(i for i, w in enumerate(waves) if a in w or b in w)
        for s in [s for w in waves[w_cyc + 1:] for s in w if s in to_solve]:
            solve(s)
    return lam, Reff
