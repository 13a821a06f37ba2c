"""The committed oracle answers at full size (tests/golden/fullsize_oracle.npz,
serving_oracle.npz) were computed from seeded synthetic inputs that the GPU
tests rebuild; these CPU checks fail loudly if the generator drifts from the
inputs the oracle saw (sums of the emissions, diagonals, transport blocks and
query inputs)."""
import os

import numpy as np
import pytest

from orbit_amd import serve, synth

HERE = os.path.join(os.path.dirname(__file__), "golden")


def _fingerprint(wl):
    return np.array([wl.e.sum(), np.abs(wl.e).max(), wl.D.sum(), sum(float(T.sum()) for T in wl.T)])


@pytest.mark.parametrize("kind", ["fct-month", "adjoint"])
def test_fullsize_inputs_unchanged(kind):
    gold = np.load(os.path.join(HERE, "fullsize_oracle.npz"), allow_pickle=False)
    key = "forward_fingerprint" if kind == "fct-month" else "adjoint_fingerprint"
    np.testing.assert_allclose(_fingerprint(synth.benchmark_workload(kind)), gold[key][:4], rtol=1e-12)


def test_serving_inputs_unchanged():
    gold = np.load(os.path.join(HERE, "serving_oracle.npz"), allow_pickle=False)
    tracers = [int(t) for t in gold["tracers"]]
    for month, size in ((0, "24x20"), (1, "48x40")):
        wl = synth.benchmark_workload("serving", n_tracers=12, month=month, size=size)
        for mode, seed in (("forward", 40), ("adjoint", 50)):
            p = serve.query_payload(8, wl.N, wl.nx * wl.ny, tracers, mode, np.random.default_rng([seed, month]),
                                    base_e=wl.e)
            np.testing.assert_allclose(p.sum(), gold[f"{mode}_{month}_payload_sum"], rtol=1e-12)
