"""The orbit NPZ interop format (reference orbit/cli.py:1297-1413, 3025-3035,
1905-1916): keys, c_mean = mean of c_1..c_8, round trip, warm start."""
import numpy as np

from orbit_amd import orbit_io


def test_round_trip_keys_and_orbit_mean(tmp_path):
    rng = np.random.default_rng(0)
    c = rng.random((3, 9, 50))
    info = [{"gmres_iters": 7 + i, "residual": 1e-9 * (i + 1), "periodicity": 1e-12} for i in range(3)]
    p = str(tmp_path / "orbit_M01.npz")
    orbit_io.save_orbit_npz(p, c, species_order=[2, 5, 9], info=info, month=1)
    d = np.load(p, allow_pickle=False)
    assert set(d.files) >= {"c_orbit", "c_mean", "orbit_species_order", "gmres_iters", "gmres_resid",
                            "periodicity", "month"}
    np.testing.assert_array_equal(d["c_orbit"], c)
    np.testing.assert_array_equal(d["c_mean"], c[:, 1:].mean(axis=1))
    out = orbit_io.load_orbit_npz(p)
    assert sorted(out["orbits"]) == [2, 5, 9] and len(out["orbits"][5]) == 9
    np.testing.assert_array_equal(out["orbits"][9][4], c[2, 4])
    np.testing.assert_array_equal(out["gmres_iters"], [7, 8, 9])
    np.testing.assert_array_equal(orbit_io.warm_start_from(p, 3), c[:, 1:].mean(axis=1))
    assert orbit_io.warm_start_from(p, 14) is None


# keys the reference's checkpoint reader takes (orbit/cli.py:1387-1407)
CHECKPOINT_KEYS = ("c_orbit", "orbit_species_order", "gmres_iters", "gmres_resid", "periodicity",
                   "iter", "converged")


def test_every_checkpoint_key_is_written(tmp_path):
    rng = np.random.default_rng(1)
    c = rng.random((2, 9, 20))
    for info, kw in ((None, {}), ([{"gmres_iters": 3, "residual": 1e-9, "periodicity": 1e-12, "converged": 1},
                                   {"gmres_iters": 4, "residual": 2e-9, "periodicity": 1e-12, "converged": 0}],
                                  {"iteration": 5})):
        p = str(tmp_path / "ck.npz")
        orbit_io.save_orbit_npz(p, c, info=info, **kw)
        d = np.load(p, allow_pickle=False)
        missing = [k for k in CHECKPOINT_KEYS if k not in d.files]
        assert not missing, missing
        # what the reader does with them
        assert int(d["iter"]) == kw.get("iteration", 0)
        assert bool(d["converged"]) is False  # no info, or a species that did not converge
        assert d["gmres_iters"].shape == (2,) and d["gmres_resid"].shape == (2,)
    info = [{"gmres_iters": 3, "residual": 1e-9, "periodicity": 1e-12, "converged": 1}] * 2
    orbit_io.save_orbit_npz(p, c, info=info)
    assert bool(np.load(p, allow_pickle=False)["converged"]) is True
