"""The shared column block's product on the matrix cores (v_mfma 16x16x4 in
FP32 for the mixed solve's sweeps, the four slots of a wave as B's columns)
against the VALU loop it replaces: the same orbits to the solve tolerance in
both directions, the residual the mixed solve reports (its exact FP64 bin
residuals propagated by these sweeps) still the true one (oracle), and a
tracer's result bit-identical alone and in any batch with the matrix cores on
(every slot count uses the same four-slot mapping)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from oracle import orbit_ref as orb
from orbit_amd import synth
from orbit_amd.core import OperatorSet


@pytest.fixture(scope="module")
def month():
    wl = synth.make_workload(16, 12, 15, 5, seed=131, convection=True, wind_std=6.0)
    op = OperatorSet.from_workload(wl)
    assert op.stats()["mfma"] == 0  # the VALU loop is the default
    op.set_mfma(True)
    return wl, op


def _rel(a, b):
    return float((a - b).abs().max() / b.abs().max())


def test_matrix_cores_match_the_valu_loop(month):
    wl, op = month
    e = torch.from_numpy(wl.e).cuda()
    R = torch.from_numpy(np.random.default_rng(1).random((wl.n_tracers, 8, wl.N))).cuda()
    assert op.stats()["mfma"] == 1
    on_f, i_on = op.solve_forward(e, tol=1e-10, precision="mixed")
    on_a, _ = op.solve_adjoint(R, tol=1e-10, precision="mixed")
    op.set_mfma(False)
    try:
        off_f, i_off = op.solve_forward(e, tol=1e-10, precision="mixed")
        off_a, _ = op.solve_adjoint(R, tol=1e-10, precision="mixed")
    finally:
        op.set_mfma(True)
    assert all(i["converged"] for i in i_on + i_off)
    assert not torch.equal(on_f, off_f)  # the two products round differently...
    assert _rel(on_f, off_f) < 1e-8 and _rel(on_a, off_a) < 1e-8  # ...inside the tolerance


def test_matrix_core_residual_is_the_true_one(month):
    wl, op = month
    orbit, info = op.solve_forward(wl.e, tol=1e-8, precision="mixed")
    for s in (0, 3):
        lus = [orb.factor_one_bin(wl.L(t, s)) for t in range(8)]
        sv = orb.compute_s(lus, list(wl.e[s]))
        c0 = orbit[s, 0]
        true = np.linalg.norm(sv - (c0 - orb.apply_monodromy(c0, lus))) / np.linalg.norm(sv)
        assert true <= 1.05e-8 and abs(info[s]["residual"] - true) <= 5e-10, (s, true, info[s]["residual"])


def test_matrix_cores_are_batch_invariant(month):
    wl, op = month
    e = torch.from_numpy(wl.e).cuda()
    full, _ = op.solve_forward(e, tol=1e-8, precision="mixed")
    for tr in ([2], [4, 1], [0, 2, 3], [3, 0, 4, 1, 2]):
        part, _ = op.solve_forward(e[tr].contiguous(), tracers=tr, tol=1e-8, precision="mixed")
        for q, t in enumerate(tr):
            assert torch.equal(part[q], full[t]), (tr, t)
