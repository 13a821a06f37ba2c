"""The mixed solve's orbit-state refinement on operator sets of shapes the
benchmark does not reach, against exact dense solves: tracers with transports
of their own (no shared column blocks, so the residual launch and the sweeps
read each tracer's rows), and one, two, three and eight bins (no, one, two
and seven intermediate states), forward and adjoint, both precisions.

The reference solves (I - M) c0 = s with M = P_{nb-1} ... P_0 and propagates
the orbit from c0 (orbit/core/orbit.py:99-210); the adjoint solves
(I - M^T) lam0 = s_adj with lam_t = P_t^T lam_{t+1} + R_t
(orbit/core/orbit_adjoint.py:36-153).  Here both are formed densely on a
10 x 8 x 6 grid (N = 480) and solved with numpy: the answer to 1e-13, with
no iteration of its own.  Every bin's recursion except the closing one is
checked against the host CSR operators (the closing one carries the periodic
residual, as in the reference)."""
import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu

from orbit_amd import synth
from orbit_amd.core import OperatorSet


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _operators(nb, distinct):
    """Per-bin CSR L of two tracers; with distinct transports tracer 1 takes
    another bin's wind field (T_{(t + 3) mod 8}) with its own diagonal."""
    wl = synth.make_workload(10, 8, 6, 2, seed=91, convection=True)
    L = []
    for t in range(nb):
        row = [wl.L(t, 0)]
        if distinct:
            m = (wl.T[(t + 3) % 8] + sp.diags(wl.D[t, 1])).tocsr()
            m.sum_duplicates()
            m.sort_indices()
            row.append(m)
        else:
            row.append(wl.L(t, 1))
        L.append(row)
    return wl, L


def _dense_orbits(L, dtau, e, R):
    """Exact forward c0 and adjoint lam0 for every tracer (dense inverses)."""
    nb, nt = len(L), len(L[0])
    n = L[0][0].shape[0]
    eye = np.eye(n)
    fwd, adj = [], []
    for s in range(nt):
        P = [np.linalg.inv(eye + dtau * L[t][s].toarray()) for t in range(nb)]
        M = eye.copy()
        sv = np.zeros(n)
        for t in range(nb):
            M = P[t] @ M
            sv = P[t] @ (sv + dtau * e[s, t])
        c0 = np.linalg.solve(eye - M, sv)
        x = np.zeros(n)
        for t in range(nb - 1, -1, -1):
            x = P[t].T @ x + R[s, t]
        lam0 = np.linalg.solve(eye - M.T, x)
        fwd.append(c0)
        adj.append(lam0)
    return fwd, adj


@pytest.mark.parametrize("nb,distinct", [(1, False), (2, False), (3, True), (8, True)])
def test_refinement_shapes_match_dense_solves(nb, distinct):
    wl, L = _operators(nb, distinct)
    op = OperatorSet.from_csr(L, wl.nz, wl.ny, wl.nx)
    assert op.stats()["n_groups"] == (2 if distinct else 1)
    e = np.ascontiguousarray(wl.e[:, :nb])
    R = np.random.default_rng(5).random((2, nb, wl.N))
    fwd, adj = _dense_orbits(L, wl.dtau, e, R)
    for prec, tol, bar in (("mixed", 1e-8, 1e-6), ("fp64", 1e-10, 1e-8)):
        print(f"n_bins {nb}, distinct transports {distinct}, {prec}", flush=True)
        orbit, info = op.solve_forward(e, tol=tol, precision=prec)
        lam, ainfo = op.solve_adjoint(R, tol=tol, precision=prec)
        for s in range(2):
            assert info[s]["converged"] and info[s]["residual"] <= tol, (prec, s, info[s])
            assert ainfo[s]["converged"] and ainfo[s]["residual"] <= tol, (prec, s, ainfo[s])
            assert _rel(orbit[s, 0], fwd[s]) < bar, (prec, s)
            assert _rel(lam[s, 0], adj[s]) < bar, (prec, s)
            assert _rel(lam[s, nb], lam[s, 0]) == 0.0  # lam_nb is lam0 itself
            # every bin but the closing one: forward bins 0 .. nb-2, adjoint bins 1 .. nb-1
            for t in range(nb):
                A = sp.identity(wl.N, format="csr") + wl.dtau * L[t][s]
                if t < nb - 1:
                    r = A @ orbit[s, t + 1] - (orbit[s, t] + wl.dtau * e[s, t])
                    assert np.linalg.norm(r) <= 1e-10 * np.linalg.norm(orbit[s, t + 1]), (prec, s, t)
                if t > 0:
                    ra = A.T @ (lam[s, t] - R[s, t]) - lam[s, t + 1]
                    assert np.linalg.norm(ra) <= 1e-10 * np.linalg.norm(lam[s, t + 1]), (prec, s, t)
    op.close()
