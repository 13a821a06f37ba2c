"""CPU oracle for the periodic-orbit solve: TEST INFRASTRUCTURE ONLY.

This package restates the reference algorithm (SumilThakr/orbit v1.0.4) on the
CPU with numpy/scipy so the HIP path can be checked against it.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may import
it, and only as the checker or as the timed CPU baseline; the product path in
``orbit_amd`` never calls into it.

Modules:
  orbit_ref  -- per-bin splu factors, monodromy, compute_s, GMRES(30) periodic
                solve, adjoint solve and gradient (orbit/core/orbit.py:79-210,
                orbit/core/orbit_adjoint.py:36-153)
  fct_ref    -- 1-D FCT harness and grid-aware horizontal FCT source and frozen
                tangent-linear operator (orbit/core/fct.py:38-572)
  mass_ref   -- air-mass-balance row-sum diagnostics (orbit/core/operator.py:598-652)
  drivers_ref -- FCT deferred-correction drivers (orbit/cli.py:1212-1241,
                orbit/modes/marginal.py:627-730, orbit/modes/adjoint.py:814-962)

Parity status of each function is listed in the module docstrings.
"""
