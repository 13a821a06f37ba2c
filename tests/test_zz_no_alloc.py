"""The allocation-free forward solve (orbit_solve_forward_dev_noalloc) against
orbit_solve_forward_dev, both called through the C ABI: the same orbits and
solve information bit for bit on two calls in a row, the second on a smaller
input (fewer tracers of a smaller grid), in both precisions; a request larger
than the fixed scratch array is refused before anything runs."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from orbit_amd import _lib, synth
from orbit_amd.core import OperatorSet


def _call(fn, op, tracers, e, precision, restart=30):
    tr = np.ascontiguousarray(tracers, dtype=np.int32)
    nt = len(tr)
    out = torch.full((nt, op.n_bins + 1, op.N), float("nan"), dtype=torch.float64, device="cuda")
    info = (_lib.Info * nt)()
    opts = op._opts(1e-8, 200, restart, precision, 0, 0, None, 0, False)
    torch.cuda.synchronize()
    rc = fn(op.ctx.handle, op.handle, tr.ctypes.data, nt, e.data_ptr(), None, C.byref(opts), out.data_ptr(), info)
    op.ctx.sync()
    return rc, out, [i.as_dict() for i in info]


@pytest.fixture(scope="module")
def sets():
    big = synth.make_workload(48, 40, 15, 6, seed=71, convection=True)
    small = synth.make_workload(24, 20, 15, 2, seed=72, convection=True)
    ob = OperatorSet.from_workload(big, None)
    os_ = OperatorSet.from_workload(small, ob.ctx)
    return (big, ob), (small, os_)


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_noalloc_matches_the_allocating_entry_point(sets, precision):
    lib = _lib.lib()
    for (wl, op), tracers in ((sets[0], [0, 1, 2, 3, 4, 5]), (sets[1], [1])):
        print(f"noalloc vs dev: {wl.nx}x{wl.ny}x{wl.nz}, tracers {tracers}, {precision}", flush=True)
        e = torch.from_numpy(np.ascontiguousarray(wl.e[tracers])).cuda()
        rc_a, a, ia = _call(lib.orbit_solve_forward_dev, op, tracers, e, precision)
        rc_b, b, ib = _call(lib.orbit_solve_forward_dev_noalloc, op, tracers, e, precision)
        assert rc_a == 0 and rc_b == 0, (rc_a, rc_b, _lib.last_error())
        assert torch.isfinite(b).all()
        assert torch.equal(a, b)
        assert ia == ib
        assert all(i["converged"] for i in ib)


def test_noalloc_refuses_more_than_the_fixed_array(sets):
    wl, op = sets[0]
    tracers = list(range(6)) * 8  # 48 slots at 48x40x15 in mixed precision: about 1.3 GB
    e = torch.from_numpy(np.ascontiguousarray(wl.e[tracers])).cuda()
    rc, out, _ = _call(_lib.lib().orbit_solve_forward_dev_noalloc, op, tracers, e, "mixed", restart=64)
    assert rc == _lib.ORBIT_ERR_INVALID
    assert "fixed scratch array" in _lib.last_error()
    assert torch.isnan(out).all()  # refused before anything ran
    # the array is still usable afterwards
    rc, _, info = _call(_lib.lib().orbit_solve_forward_dev_noalloc, op, [0, 1], e[:2], "mixed")
    assert rc == 0 and all(i["converged"] for i in info)
