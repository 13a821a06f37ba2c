"""Species layer on the MI355X: coupled species of a month as DAG waves of batched solves.

Mirrors
  * ``solve_orbit_all_species`` / ``_solve_one_species`` / ``_build_solve_waves``
    (orbit/core/orbit.py:213-806): per target species the effective RHS
    e_s,tau - sum_src K_(s,src),tau c_src[tau] (c_src[tau+1] for ``endbin_keys``)
    + extra RHS; species placed in dependency waves with cyclic edges removed;
    each wave reads the orbits as they were at the wave's start (seed orbits
    stand in for cyclic partners); copy-verbatim skips; the zero-RHS
    short-circuit; non-convergence is a warning;
  * the marginal driver's cyclic Picard loop (orbit/modes/marginal.py:573-612);
  * the adjoint species passes (orbit/modes/adjoint.py:496-515, 634-783):
    receptors R_src -= dtau K_(r,src)^T (lam_r - R_r) from solved receivers,
    reverse-DAG first pass, Picard on the cyclic pair, refresh of later waves.

MI355X form: all species of a month are the tracers of one device operator set
(the reference's LU cache, factored once); a wave's species are solved as one
batched periodic solve instead of one thread per species; the K blocks live on
the device (``orbit_spmat``) and are applied to the device orbits directly, so
nothing leaves HBM between waves.  Species are tracer indices of the operator set.
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, Iterable, List, Optional, Sequence

import numpy as np
import scipy.sparse as sp

from ._lib import check, lib, ptr
from .core import DTAU, N_BINS, Context, OperatorSet

# orbit/core/deposition.py:57-77, orbit/core/orbit.py:70
IDX_TOTAL_NH, IDX_TOTAL_NO3 = 2, 6
REF_N_SPECIES = 14
REF_SOLVE_ORDER = (0, 1, 13, 2, 3, 4, 5, 6, 7, 8, 12, 9, 10, 11)


class CouplingBlock:
    """One N x N coupling block K on the device (CSR and the CSR of K^T)."""

    def __init__(self, K, ctx: Context):
        K = sp.csr_matrix(K, dtype=np.float64)
        K.sum_duplicates()
        K.sort_indices()
        if K.shape[0] != K.shape[1]:
            raise ValueError("coupling blocks must be square")
        self.ctx = ctx
        self.n = K.shape[0]
        rp = np.ascontiguousarray(K.indptr, dtype=np.int32)
        ci = np.ascontiguousarray(K.indices, dtype=np.int32)
        va = np.ascontiguousarray(K.data, dtype=np.float64)
        h = C.c_void_p()
        check(lib().orbit_spmat_create(ctx.handle, self.n, ptr(rp), ptr(ci), ptr(va), C.byref(h)))
        self.handle = h

    def axpy(self, x, y, alpha: float, trans: bool = False):
        """y += alpha K x (or K^T x) on the context's stream; x, y CUDA float64 (n,)."""
        from ._lib import dev_f64
        for name, t in (("x", x), ("y", y)):
            dev_f64(t, self.ctx.device, name, last=self.n)
            if t.numel() != self.n or not t.is_contiguous():
                raise ValueError(f"{name} must be a contiguous vector of n = {self.n} values")
        check(lib().orbit_spmat_apply_dev(self.ctx.handle, self.handle, int(bool(trans)), float(alpha),
                                          x.data_ptr(), y.data_ptr()))

    def close(self):
        if getattr(self, "handle", None):
            lib().orbit_spmat_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SpeciesCouplings:
    """All K blocks of a month on the device: ``K_sources_per_bin[tau][(target, source)]``."""

    def __init__(self, K_sources_per_bin: Sequence[dict], ctx: Context):
        if len(K_sources_per_bin) != N_BINS:
            raise ValueError(f"K_sources_per_bin must have {N_BINS} bins")
        keys = set(K_sources_per_bin[0].keys())
        for t, d in enumerate(K_sources_per_bin):
            if set(d.keys()) != keys:
                raise ValueError(f"bin {t} has a different set of (target, source) blocks than bin 0")
        self.keys = sorted(keys)
        self.blocks = [{k: CouplingBlock(K_sources_per_bin[t][k], ctx) for k in self.keys} for t in range(N_BINS)]
        self.sources: Dict[int, List[int]] = {}
        self.receivers: Dict[int, List[int]] = {}
        for (tgt, src) in K_sources_per_bin[0].keys():   # the reference's dict order
            self.sources.setdefault(tgt, []).append(src)
            self.receivers.setdefault(src, []).append(tgt)

    def close(self):
        for d in self.blocks:
            for b in d.values():
                b.close()


def build_solve_waves(sources_for_target: dict, solve_order: Sequence[int], skip_species: Iterable[int]):
    """Dependency waves (orbit/core/orbit.py:397-465): cyclic edges (a <- b and b <- a)
    are ignored for placement, a species sits one wave after its latest source,
    skipped species sit in wave 0; placement is iterated to a fixed point."""
    skip = set(skip_species)
    cyc = {(t, s) for t, ss in sources_for_target.items() for s in ss if t in sources_for_target.get(s, ())}
    deps = {t: [s for s in ss if (t, s) not in cyc] for t, ss in sources_for_target.items()}
    wave: Dict[int, int] = {}
    for _ in range(len(solve_order) + 2):
        moved = False
        for s in solve_order:
            w = 0 if (s in skip or not deps.get(s)) else 1 + max(wave.get(d, 0) for d in deps[s])
            if wave.get(s) != w:
                wave[s] = w
                moved = True
        if not moved:
            break
    out: List[List[int]] = [[] for _ in range(1 + max(wave.values(), default=0))]
    for s in solve_order:
        out[wave[s]].append(s)
    return out


def _torch():
    import torch
    return torch


def _dev(x, device):
    torch = _torch()
    if hasattr(x, "is_cuda"):
        return x.to(device=device, dtype=torch.float64)
    return torch.as_tensor(np.asarray(x, dtype=np.float64), device=device)


def _as_orbit(o, N, device):
    torch = _torch()
    if hasattr(o, "is_cuda"):
        return o.to(device=device, dtype=torch.float64).reshape(N_BINS + 1, N).clone()
    return _dev(np.stack([np.asarray(c, dtype=np.float64) for c in o]), device).reshape(N_BINS + 1, N)


def _shape_for(N, shape):
    if shape is not None:
        nz, ny, nx = shape
        if nz * ny * nx != N:
            raise ValueError(f"shape {shape} does not hold N = {N} cells")
        return nz, ny, nx
    if N <= 16:
        return N, 1, 1
    raise ValueError("pass shape=(nz, ny, nx) (or a prebuilt opset) for N > 16")


def _prepare(L_species_per_bin, K_sources_per_bin, N, opset, lu_cache, shape, ctx):
    if opset is None and isinstance(lu_cache, dict) and lu_cache:
        first = next(iter(lu_cache.values()))
        if first and hasattr(first[0], "opset"):
            opset = first[0].opset
    if opset is None:
        if L_species_per_bin is None:
            raise ValueError("need L_species_per_bin, an opset or an lu_cache of HipLU objects")
        nz, ny, nx = _shape_for(N, shape)
        opset = OperatorSet.from_csr(L_species_per_bin, nz, ny, nx, DTAU, ctx)
    if opset.N != N:
        raise ValueError(f"operator set has N = {opset.N}, expected {N}")
    cpl = K_sources_per_bin if isinstance(K_sources_per_bin, SpeciesCouplings) else \
        SpeciesCouplings(K_sources_per_bin, opset.ctx)
    return opset, cpl


def solve_orbit_all_species(L_species_per_bin, K_sources_per_bin, emissions_SN, N, umfpack_sym=None,
                            perm=None, tol=1e-6, maxiter=200, verbose=True, c_warm_SN=None, lu_cache=None,
                            skip_solve_species=None, cached_orbits=None, return_lu_cache=False,
                            keep_lu_species=None, extra_rhs_per_bin_per_species=None, seed_orbits=None,
                            endbin_keys=None, *, opset: Optional[OperatorSet] = None, shape=None,
                            precision: str = "fp64", solve_order: Optional[Sequence[int]] = None,
                            device_out: bool = False, ctx: Optional[Context] = None):
    """Periodic orbits of all species of a month (orbit/core/orbit.py:468-806).

    Same arguments and result as the reference (``umfpack_sym``, ``perm`` and
    ``keep_lu_species`` are accepted and unused: the operator set is factored
    once on the device).  ``K_sources_per_bin`` may be a prebuilt
    ``SpeciesCouplings``.  Keyword-only: ``opset`` (species = tracers of a prebuilt
    operator set), ``shape`` (nz, ny, nx) when building one from
    ``L_species_per_bin``, ``precision`` ("fp64" or "mixed"), ``solve_order``
    (default: the reference's order for 14 species, else 0..S-1) and
    ``device_out`` (orbits as CUDA tensors (9, N) instead of lists of arrays).
    ``L_species_per_bin`` is not modified."""
    torch = _torch()
    op, cpl = _prepare(L_species_per_bin, K_sources_per_bin, N, opset, lu_cache, shape, ctx)
    S = op.n_tracers
    order = list(solve_order) if solve_order is not None else (
        list(REF_SOLVE_ORDER) if S == REF_N_SPECIES else list(range(S)))
    if sorted(order) != sorted(set(order)) or any(s < 0 or s >= S for s in order):
        raise ValueError("solve_order must list distinct species of the operator set")
    dev = torch.device("cuda", op.ctx.device)
    skip = set(skip_solve_species or ())
    endbin = set(endbin_keys or ())
    extra = extra_rhs_per_bin_per_species or {}
    e_host = emissions_SN if hasattr(emissions_SN, "is_cuda") else np.asarray(emissions_SN, dtype=np.float64)
    if e_host.ndim not in (1, 2) or e_host.shape[-1] != S * N or (e_host.ndim == 2 and e_host.shape[0] != N_BINS):
        raise ValueError(f"emissions_SN must be ({S * N},) or ({N_BINS}, {S * N})")
    e_all = _dev(e_host, dev)
    e_all = e_all.reshape(1, S, N).expand(N_BINS, S, N) if e_all.ndim == 1 else e_all.reshape(N_BINS, S, N)

    orbits: Dict[int, object] = {s: _as_orbit(o, N, dev) for s, o in (seed_orbits or {}).items()}
    iters = np.zeros(S, dtype=int)
    resid = np.zeros(S)
    period = np.zeros(S)
    for wave in build_solve_waves(cpl.sources, order, skip):
        snap = dict(orbits)
        done = {}
        todo = []
        for s in wave:
            if s in skip and cached_orbits is not None and s in cached_orbits:
                done[s] = (_as_orbit(cached_orbits[s], N, dev), 0, 0.0, 0.0)
            else:
                todo.append(s)
        if todo:
            E = e_all[:, todo].permute(1, 0, 2).contiguous()        # (nw, 8, N)
            for i, s in enumerate(todo):
                if s in extra:
                    E[i] += _dev(np.stack([np.asarray(x) for x in extra[s]]) if not hasattr(extra[s], "is_cuda")
                                 else extra[s], dev).reshape(N_BINS, N)
            torch.cuda.current_stream(dev).synchronize()
            for i, s in enumerate(todo):
                for src in cpl.sources.get(s, []):
                    if src not in snap:
                        raise KeyError(f"orbits[{src}] missing for target s={s}: Seed-orbits machinery "
                                       "(Picard iteration) must be used for cyclic source DAGs")
                    for t in range(N_BINS):
                        lvl = t + 1 if (s, src) in endbin else t
                        cpl.blocks[t][(s, src)].axpy(snap[src][lvl], E[i, t], -1.0)
            op.ctx.sync()
            nonzero = (E.reshape(len(todo), -1).abs().amax(dim=1) > 0).tolist()
            live = [i for i, nz in enumerate(nonzero) if nz]
            for i, s in enumerate(todo):
                if not nonzero[i]:
                    done[s] = (torch.zeros((N_BINS + 1, N), dtype=torch.float64, device=dev), 0, 0.0, 0.0)
            if live:
                warm = None
                if c_warm_SN is not None:
                    cw = _dev(c_warm_SN, dev).reshape(S, N)
                    warm = cw[[todo[i] for i in live]].contiguous()
                out, info = op.solve_forward(E[live].contiguous(), tracers=[todo[i] for i in live], c_warm=warm,
                                             tol=tol, maxiter=maxiter, precision=precision)
                for k, i in enumerate(live):
                    inf = info[k]
                    if not inf["converged"] and verbose:
                        print(f"    WARNING: GMRES did not converge for species {todo[i]} "
                              f"(iters={inf['gmres_iters']})")
                    done[todo[i]] = (out[k], inf["gmres_iters"], inf["residual"], inf["periodicity"])
        for s, (o, it, r, p) in done.items():
            orbits[s] = o
            iters[s], resid[s], period[s] = it, r, p
    if device_out:
        out_orbits = orbits
    else:
        out_orbits = {s: [c.copy() for c in o.cpu().numpy()] for s, o in orbits.items()}
    res = {"orbits": out_orbits, "gmres_iters": iters, "gmres_resid": resid, "periodicity": period}
    if return_lu_cache:
        from .solve import HipLU
        res["lu_cache"] = {s: [HipLU(op, b, s) for b in range(op.n_bins)] for s in order}
    return res


def _max_rel_change(prev, curr):
    d = float((curr[:N_BINS] - prev[:N_BINS]).abs().max())
    den = float(curr[:N_BINS].abs().max())
    return d if den < 1e-30 else d / den


def picard_cyclic(L_species_per_bin, K_sources_per_bin, emissions_SN, N, cyclic_species, max_iters: int = 30,
                  rel_tol: float = 1e-8, **kw):
    """The marginal driver's Picard loop for cyclic species (orbit/modes/marginal.py:573-612):
    zero seeds, repeated all-species solves until the largest relative change of the
    cyclic orbits falls below rel_tol (not tested on the first pass).
    Returns (result, iterations, last change, converged); orbits as in ``device_out``."""
    torch = _torch()
    op, cpl = _prepare(L_species_per_bin, K_sources_per_bin, N, kw.pop("opset", None), kw.get("lu_cache"),
                       kw.pop("shape", None), kw.pop("ctx", None))
    want_dev = kw.pop("device_out", False)
    dev = torch.device("cuda", op.ctx.device)
    seed = {s: torch.zeros((N_BINS + 1, N), dtype=torch.float64, device=dev) for s in cyclic_species}
    res, change = None, 0.0
    for it in range(max_iters):
        res = solve_orbit_all_species(None, cpl, emissions_SN, N, opset=op, seed_orbits=seed, device_out=True,
                                      **kw)
        change = max(_max_rel_change(seed[s], res["orbits"][s]) for s in cyclic_species)
        if it > 0 and change < rel_tol:
            break
        seed = {s: res["orbits"][s].clone() for s in cyclic_species}
    converged = bool(it > 0 and change < rel_tol)
    if not want_dev:
        res["orbits"] = {s: [c.copy() for c in o.cpu().numpy()] for s, o in res["orbits"].items()}
    return res, it + 1, change, converged


def adjoint_plan(receivers_for_source: dict, user_species: Iterable[int], solve_order: Sequence[int]):
    """Species to solve (the user's and every receiver downstream of them), the
    adjoint waves and the adjoint order (orbit/modes/adjoint.py:49-69, 496-515)."""
    to_solve = set(user_species)
    stack = list(to_solve)
    while stack:
        s = stack.pop()
        for r in receivers_for_source.get(s, []):
            if r not in to_solve:
                to_solve.add(r)
                stack.append(r)
    waves = build_solve_waves(receivers_for_source, list(solve_order), set())
    return to_solve, waves, [s for w in waves for s in w if s in to_solve]


def adjoint_species_passes(opset: OperatorSet, R_base: dict, K_sources_per_bin, user_species: Iterable[int],
                           endbin_keys=(), tol: float = 1e-8, maxiter: int = 200, precision: str = "fp64",
                           picard_max_iters: int = 10, picard_rel_tol: float = 1e-4,
                           cyclic_species=(IDX_TOTAL_NH, IDX_TOTAL_NO3), extra_recep: Optional[dict] = None,
                           solve_order: Optional[Sequence[int]] = None):
    """Adjoint orbits of coupled species (orbit/modes/adjoint.py:634-783).

    R_base: {species: (8, N) receptors} (numpy or CUDA; missing species = 0).
    Returns (lam, R_eff): {species: CUDA (9, N)} and {species: CUDA (8, N)}; the
    gradient is dJ/de_(s,tau) = dtau (lam_s[tau] - R_eff_s[tau]).

    The reference solves the species one at a time in adjoint order, each seeing
    every receiver solved before it.  Here consecutive species with no receiver
    among themselves are solved as one batch, which gives the same result: a
    species starts a new batch exactly when one of its receivers is in the
    current one (a cyclic partner in the same wave)."""
    torch = _torch()
    op = opset
    cpl = K_sources_per_bin if isinstance(K_sources_per_bin, SpeciesCouplings) else \
        SpeciesCouplings(K_sources_per_bin, op.ctx)
    S, N = op.n_tracers, op.N
    order_ref = list(solve_order) if solve_order is not None else (
        list(REF_SOLVE_ORDER) if S == REF_N_SPECIES else list(range(S)))
    to_solve, waves, order = adjoint_plan(cpl.receivers, user_species, order_ref)
    endbin = set(endbin_keys or ())
    dev = torch.device("cuda", op.ctx.device)
    lam: Dict[int, object] = {}
    Reff: Dict[int, object] = {}

    def receptor(s):
        R = _dev(R_base[s], dev).reshape(N_BINS, N).clone() if s in R_base else \
            torch.zeros((N_BINS, N), dtype=torch.float64, device=dev)
        if extra_recep is not None and s in extra_recep:
            extra = _dev(extra_recep[s], dev).reshape(N_BINS, N)
        else:
            extra = None
        diffs = {r: (lam[r][:N_BINS] - Reff[r]).contiguous() for r in cpl.receivers.get(s, []) if r in lam}
        torch.cuda.current_stream(dev).synchronize()
        for r, diff in diffs.items():
            for t in range(N_BINS):
                tb = (t + 1) % N_BINS if (r, s) in endbin else t
                cpl.blocks[t][(r, s)].axpy(diff[t], R[tb], -DTAU, trans=True)
        op.ctx.sync()
        if extra is not None:
            R += extra
        return R

    def solve_batch(batch):
        Rs = [receptor(s) for s in batch]
        out, _ = op.solve_adjoint(torch.stack(Rs).contiguous(), tracers=list(batch), tol=tol, maxiter=maxiter,
                                  precision=precision)
        for k, s in enumerate(batch):
            lam[s] = out[k]
            Reff[s] = Rs[k]

    def run(seq):
        batch: List[int] = []
        for s in seq:
            if any(r in batch for r in cpl.receivers.get(s, [])):
                solve_batch(batch)
                batch = []
            batch.append(s)
        if batch:
            solve_batch(batch)

    run(order)
    a, b = cyclic_species
    K0 = set(cpl.keys)
    if a in lam and b in lam and ((a, b) in K0 or (b, a) in K0):
        for _ in range(picard_max_iters):
            prev = {s: lam[s].clone() for s in cyclic_species}
            for s in cyclic_species:
                solve_batch([s])
            change = 0.0
            for s in cyclic_species:
                d = torch.linalg.vector_norm(lam[s][:N_BINS] - prev[s][:N_BINS], dim=1)
                base = torch.linalg.vector_norm(lam[s][:N_BINS], dim=1)
                change = max(change, float((d / (base + 1e-30)).max()))
            if change < picard_rel_tol:
                break
        w_cyc = next(i for i, w in enumerate(waves) if a in w or b in w)
        run([s for w in waves[w_cyc + 1:] for s in w if s in to_solve])
    return lam, Reff
