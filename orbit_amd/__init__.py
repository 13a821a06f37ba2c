"""orbit_amd: the ORBIT periodic-orbit transport solve on the AMD MI355X.

The product path is liborbit_hip.so (hand-written HIP for gfx950, C ABI in
include/orbit_hip.h) driven from Python through ctypes; there is no CPU
fallback.  Modules mirror the reference's call surface:
  orbit_amd.solve          factor backend (HipLU: .solve(b, trans))
  orbit_amd.orbit          apply_monodromy, compute_s, solve_orbit_one_species, solve_orbits
  orbit_amd.orbit_adjoint  solve_orbit_adjoint_one_species, gradient_per_emission_bin
  orbit_amd.fct            compute_horizontal_fct_source, assemble_fct_linear_operator
  orbit_amd.core           Context, OperatorSet (device-resident month operators)
  orbit_amd.synth          seeded synthetic workloads
"""
__version__ = "0.1.0"
