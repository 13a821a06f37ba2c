"""The outer closure around the orbit solve on the device: Anderson(m) on CUDA
tensors against the oracle, and a partitioning fixed point -- per-tracer loss
depending on the orbit-mean state through the LUT, refactored with
update_diag, accelerated by Anderson -- against the same loop on the oracle
(splu + GMRES, scipy LUT, numpy Anderson)."""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

pytestmark = pytest.mark.gpu

from oracle import assembly_ref as asm, lut_ref, orbit_ref as orb
from oracle.closure_ref import AndersonRef
from orbit_amd import synth
from orbit_amd.closure import AndersonAccelerator
from orbit_amd.core import OperatorSet
from orbit_amd.lut import IsorropiaLUT


def test_anderson_on_device_tensors_follows_the_oracle():
    rng = np.random.default_rng(4)
    n = 300
    W = rng.normal(0.0, 0.25, (n, n)) / np.sqrt(n)
    b = rng.normal(0.0, 1.0, n)
    ref, dev = AndersonRef(m=3), AndersonAccelerator(m=3)
    Wd, bd = torch.from_numpy(W).cuda(), torch.from_numpy(b).cuda()
    x_r, x_d = np.zeros(n), torch.zeros(n, dtype=torch.float64, device="cuda")
    for it in range(12):
        x_r, _ = ref.apply_with_safeguard(x_r, np.tanh(W @ x_r) + b)
        x_d, _ = dev.apply_with_safeguard(x_d, torch.tanh(Wd @ x_d) + bd)
        np.testing.assert_allclose(x_d.cpu().numpy(), x_r, rtol=1e-10, atol=1e-12)
    assert np.abs(np.tanh(W @ x_r) + b - x_r).max() < 1e-6   # converging (12 steps)


def test_partitioning_closure_matches_the_oracle_loop():
    grids, D0, e = synth.grid_workload(12, 10, 6, 1, seed=31)
    N = D0.shape[2]
    op = OperatorSet.from_grids(grids, D0)
    t = lut_ref.synthetic_table(9)
    lut = IsorropiaLUT(arrays=t)
    rng = np.random.default_rng(9)
    fixed = [rng.uniform(0, 30, N), None, rng.uniform(0, 20, N), rng.uniform(0, 1, N), rng.uniform(0, 1, N),
             rng.uniform(270, 300, N), rng.uniform(0.3, 0.9, N)]
    k_part = 3e-5            # extra loss of the particle fraction, 1/s
    To = [asm.assemble_transport_block(g) for g in grids]

    def D_of(f):
        return D0 + k_part * f[None, None, :]

    def scale(cm):          # orbit-mean concentration -> total NH axis units
        return 30.0 * cm / cm.max()

    f_d = torch.full((N,), 0.5, dtype=torch.float64, device="cuda")
    f_r = np.full(N, 0.5)
    acc_d, acc_r = AndersonAccelerator(m=3), AndersonRef(m=3)
    e_d = torch.from_numpy(e).cuda()
    for it in range(5):
        # device: refactor with the new loss, solve, query, accelerate
        op.update_diag(D_of(f_d.cpu().numpy()))
        o, _ = op.solve_forward(e_d, tol=1e-11, precision="fp64")
        cm = o[0, 1:].mean(dim=0)
        x = [torch.from_numpy(v).cuda() if v is not None else scale(cm) for v in fixed]
        g_d = lut.query(*x)[0]
        f_d, _ = acc_d.apply_with_safeguard(f_d, g_d)
        # oracle
        Dr = D_of(f_r)
        lus = [orb.factor_one_bin((To[b] + sp.diags(Dr[b, 0])).tocsr()) for b in range(8)]
        ref, _ = orb.solve_orbit_one_species(lus, list(e[0]), tol=1e-12, maxiter=500)
        cmr = np.mean(ref[1:], axis=0)
        xr = [v if v is not None else scale(cmr) for v in fixed]
        g_r = lut_ref.query(t, ("f_nh4", "f_no3"), *xr)[0]
        f_r, _ = acc_r.apply_with_safeguard(f_r, g_r)
        np.testing.assert_allclose(f_d.cpu().numpy(), f_r, rtol=1e-7, atol=1e-9, err_msg=f"iteration {it}")
    assert acc_d.n_updates == 5
