"""MI355X-native hot path of the HEXAT topology-optimisation loop.

The numeric work runs in ``libhexat.so`` (hand-written gfx950 HIP kernels
behind the C ABI of ``include/hexat.h``); the modules here mirror the
reference's Julia entry points on top of it:

* ``solver.solve_system`` / ``solver.solve_system_gpu``   -- Solver.jl:38, GPUSolver.jl:611
* ``gpu_helmholtz.apply_gpu_filter``                       -- GPUHelmholtz.jl:344
* ``stress.compute_stress_field``                          -- Stress.jl:57
* ``topopt.update_density`` (stress-ratio, parity mode)    -- TopOpt.jl:139
* ``topopt.SimpOptimizer`` (SIMP/OC, device-resident loop)
* ``driver.run_main`` (YAML-driven TO loop)               -- Main.jl:175-476
"""
from . import _lib
from ._lib import Context, HexatError, device_count, element_stiffness, scalar_matrices

__all__ = ["_lib", "Context", "HexatError", "device_count", "element_stiffness", "scalar_matrices"]
