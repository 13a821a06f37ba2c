"""Anderson(m) acceleration: the oracle (oracle/closure_ref.py) pinned by
closed forms, and the device class's host logic checked against it on CPU
tensors (the arithmetic is torch's on either device)."""
import numpy as np
import pytest
import torch

from oracle.closure_ref import AndersonRef
from orbit_amd.closure import AndersonAccelerator


def _linear(n, seed):
    rng = np.random.default_rng(seed)
    A = rng.normal(0.0, 1.0, (n, n))
    A *= 0.9 / np.abs(np.linalg.eigvals(A)).max()
    b = rng.normal(0.0, 1.0, n)
    return A, b, np.linalg.solve(np.eye(n) - A, b)


def test_first_update_is_picard():
    acc = AndersonRef(m=3, beta=0.7)
    x = np.array([1.0, 2.0, 3.0])
    g = np.array([2.0, 0.0, 3.5])
    np.testing.assert_array_equal(acc.update(x, g), x + 0.7 * (g - x))


def test_full_memory_reaches_a_linear_fixed_point_in_n_plus_1_updates():
    """Anderson with memory >= n on g(x) = A x + b: the (n+1)-th update lands
    on the fixed point (Walker & Ni 2011, equivalence with GMRES)."""
    n = 6
    A, b, xs = _linear(n, 1)
    acc = AndersonRef(m=n + 2)
    x = np.zeros(n)
    errs = []
    for _ in range(n + 1):
        x = acc.update(x, A @ x + b)
        errs.append(np.abs(x - xs).max())
    assert errs[-1] < 1e-9 * np.abs(xs).max()
    assert errs[0] > 1e-3          # Picard alone is far off
    # memory 1 does not get there in the same number of steps
    acc1 = AndersonRef(m=1)
    x = np.zeros(n)
    for _ in range(n + 1):
        x = acc1.update(x, A @ x + b)
    assert np.abs(x - xs).max() > 1e-6


def test_safeguard_below_one_forces_picard_and_clears_history():
    A, b, _ = _linear(4, 2)
    acc = AndersonRef(m=3, safeguard_factor=1e-3)
    x = np.zeros(4)
    x1, ok1 = acc.apply_with_safeguard(x, A @ x + b)
    x2, ok2 = acc.apply_with_safeguard(x1, A @ x1 + b)
    assert ok1 and not ok2 and acc.n_safeguard_falls == 1 and acc.X == []
    np.testing.assert_allclose(x2, x1 + (A @ x1 + b - x1))


@pytest.mark.parametrize("m,beta,factor", [(3, 1.0, 1.5), (2, 0.6, 1.5), (3, 1.0, 0.2)])
def test_device_class_follows_the_oracle(m, beta, factor):
    rng = np.random.default_rng(3)
    n = 40
    W = rng.normal(0.0, 0.3, (n, n)) / np.sqrt(n)
    b = rng.normal(0.0, 1.0, n)
    g = lambda x: np.tanh(W @ x) + b
    ref = AndersonRef(m=m, beta=beta, safeguard_factor=factor)
    dev = AndersonAccelerator(m=m, beta=beta, safeguard_factor=factor)
    x_r = np.zeros(n)
    x_d = torch.zeros(n, dtype=torch.float64)
    for it in range(10):
        x_r, ok_r = ref.apply_with_safeguard(x_r, g(x_r))
        x_d, ok_d = dev.apply_with_safeguard(x_d, torch.from_numpy(g(x_d.numpy())))
        assert ok_r == ok_d, it
        np.testing.assert_allclose(x_d.numpy(), x_r, rtol=1e-10, atol=1e-12)
    assert dev.n_safeguard_falls == ref.n_safeguard_falls and dev.n_updates == ref.n_updates
    # rank-deficient history (a repeated residual) is cut by rcond like lstsq
    dev2, ref2 = AndersonAccelerator(m=3), AndersonRef(m=3)
    x = np.ones(n)
    for _ in range(3):
        xr = ref2.update(x, x + 1.0)
        xd = dev2.update(torch.from_numpy(x), torch.from_numpy(x + 1.0))
        np.testing.assert_allclose(xd.numpy(), xr, rtol=1e-12)
