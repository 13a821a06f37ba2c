"""The adjoint gradient at the benchmark size (64x70x15, 12 tracers):
d(r . c0)/ds through the autograd function, with r an indicator on 50 random
surface cells weighted U(0,1) per tracer.  At this size the oracle does not
finish in seconds, so the checks are size-independent: the duality
<r, c0(s)> = <lambda0, s> of the linear periodic solve, central finite
differences (eps = 1e-4 relative) on 20 random emission cells, and the mixed
precision gradient against FP64 at the 1e-6 bar."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from orbit_amd import autograd as hag, synth
from orbit_amd.core import OperatorSet


@pytest.fixture(scope="module")
def full():
    wl = synth.benchmark_workload("adjoint")
    op = OperatorSet.from_workload(wl)
    rng = np.random.default_rng(33)
    ncol = wl.nx * wl.ny
    s = np.zeros((wl.n_tracers, wl.N))
    src = rng.random(ncol) < 0.3
    s[:, :ncol] = rng.lognormal(0.0, 1.0, (wl.n_tracers, ncol)) * src * 1e-3
    r = np.zeros((wl.n_tracers, wl.N))
    for t in range(wl.n_tracers):
        cells = rng.choice(ncol, 50, replace=False)
        r[t, cells] = rng.random(50)
    return wl, op, s, r, np.flatnonzero(src), rng


def _grad(op, s, r, precision):
    st = torch.from_numpy(s).cuda().requires_grad_(True)
    c0 = hag.periodic_solve(st, op, tol=1e-12 if precision == "fp64" else 1e-8, precision=precision)
    J = (torch.from_numpy(r).cuda() * c0).sum()
    J.backward()
    return float(J.detach()), st.grad.cpu().numpy()


def test_full_size_gradient_duality_fd_and_mixed(full):
    wl, op, s, r, src, rng = full
    J, g = _grad(op, s, r, "fp64")
    # duality of the linear solve: r . (I - M)^-1 s = lambda0 . s
    assert np.sum(g * s) == pytest.approx(J, rel=1e-9)
    # central finite differences on 20 emission cells, all in one batched solve
    picks = [(int(rng.integers(wl.n_tracers)), int(rng.choice(src))) for _ in range(20)]
    S, tr = [], []
    hs = []
    for t, j in picks:
        h = 1e-4 * s[t, j]
        for sign in (1.0, -1.0):
            v = s[t].copy()
            v[j] += sign * h
            S.append(v)
            tr.append(t)
        hs.append(h)
    out, info = op.solve_forward(np.stack(S), tracers=tr, tol=1e-12, precision="fp64", rhs="s")
    assert all(i["converged"] for i in info)
    for k, (t, j) in enumerate(picks):
        fd = (r[t] @ out[2 * k, 0] - r[t] @ out[2 * k + 1, 0]) / (2 * hs[k])
        assert fd == pytest.approx(g[t, j], rel=1e-6, abs=1e-10 * np.abs(g).max()), (t, j)
    # mixed precision gradient at the 1e-6 bar, per tracer
    _, gm = _grad(op, s, r, "mixed")
    for t in range(wl.n_tracers):
        assert np.abs(gm[t] - g[t]).max() <= 1e-6 * np.abs(g[t]).max(), t
