"""Elasticity solve entry points (mirror of Solver.jl:38-164,
IterativeSolver.jl:20-43 and GPUSolver.jl:274-627) on libhexat's HIP PCG.

The whole solve policy runs in the library (``hx_solve_system``), on device
data:
* active elements: rho >= min_stiffness_threshold, modulus E*rho (linear,
  GPUSolver.jl:128-147); with ``prune_voids`` and fewer than 99 % of the
  elements above 1.01*threshold, voids and the nodes they orphan leave the
  system (Solver.jl:60-91, MeshPruner.jl:18-138): their moduli are 0 and the
  orphan DOFs carry no stiffness, so they are not solved for and return 0; a
  warm start keeps u_prev on the kept nodes (Solver.jl:71-83);
* shift s = shift_factor * max(diag over free DOFs) (GPUSolver.jl:391);
* three attempts: acceptance of a stagnated attempt below 5e-2, otherwise the
  shift factor x10 (hard failure) or x5 (stagnation, its best iterate kept),
  and the best kept iterate (or zeros) when all fail (GPUSolver.jl:531-606);
  no AMG (the CPU AMG branch is out of scope; "amg" and "jacobi_amg" run as
  Jacobi);
* zero RHS returns zeros, 0.0, "None" (GPUSolver.jl:373).

``solve_system`` adds the adaptive tolerance 10^((1-t)(-5) + t log10(tol))
with the GPU floor 5e-7 (Solver.jl:99-120) in ``parity`` mode; ``northstar``
mode uses the configured tolerance as is.

Results are float64 (the reference returns Float32).
"""
from __future__ import annotations

import math
import threading
import warnings

import numpy as np

from . import _lib
from . import boundary as _bnd
from . import mesh as _mesh

_ctx_lock = threading.Lock()
_ctx = None


def get_context(device: int = 0) -> _lib.Context:
    """Process-wide HIP context (one per process, device 0 like Main.jl:179)."""
    global _ctx
    with _ctx_lock:
        if _ctx is None:
            _ctx = _lib.Context(device)
        return _ctx


def _bind_grid(ctx, nodes, elements, nu):
    nx, ny, nz, dx, dy, dz = _mesh.grid_of(nodes, elements)
    if ctx.grid != (nx, ny, nz, dx, dy, dz, float(nu)):
        ctx.set_grid(nx, ny, nz, dx, dy, dz, nu)
    return nx, ny, nz, dx, dy, dz


def _precision_of(config):
    sp = (config or {}).get("solver_parameters", {})
    p = str(sp.get("precision", "fp64")).lower()
    return _lib.HX_FP32_IR if p in ("fp32_mfma_ir", "fp32_ir", "fp32") else _lib.HX_FP64


def _precond_of(config):
    """solver_parameters.preconditioner: "jacobi" (the reference's; "amg" and
    "jacobi_amg" also run Jacobi, the CPU AMG branch being out of scope),
    "multilevel" or "vcycle" (added here)."""
    sp = (config or {}).get("solver_parameters", {})
    p = str(sp.get("preconditioner", "jacobi")).lower()
    kinds = {"jacobi": (_lib.HX_PC_JACOBI, "Jacobi"), "jacobi_amg": (_lib.HX_PC_JACOBI, "Jacobi"),
             "amg": (_lib.HX_PC_JACOBI, "Jacobi"), "multilevel": (_lib.HX_PC_MULTILEVEL, "Multilevel"),
             "vcycle": (_lib.HX_PC_VCYCLE, "V-cycle")}
    if p not in kinds:
        raise ValueError(f"solver_parameters.preconditioner: {p!r} (jacobi | multilevel | vcycle)")
    return kinds[p]


def _solve(nodes, elements, E, nu, bc_indicator, f, density, *, max_iter, tol, shift_factor, min_stiffness_threshold,
           u_guess, config, prune_voids):
    ctx = get_context()
    _bind_grid(ctx, nodes, elements, nu)
    ctx.set_fixed(_bnd.fixed_dof_mask(bc_indicator))
    pc, pc_name = _precond_of(config)
    u0 = None
    if u_guess is not None and len(u_guess) == ctx.ndof:
        u0 = np.asarray(u_guess, dtype=np.float64)
    u, st = ctx.solve_system(np.asarray(density, dtype=np.float64), np.asarray(f, dtype=np.float64), u0,
                             E=float(E), min_stiffness_threshold=float(min_stiffness_threshold),
                             prune_voids=bool(prune_voids), tol=float(tol), max_iter=int(max_iter),
                             shift_factor=float(shift_factor), precision=_precision_of(config), precond=pc)
    if st.outcome == _lib.HX_SOLVE_ZERO_RHS:  # GPUSolver.jl:372-373
        return np.zeros(ctx.ndof), 0.0, "None"
    if st.outcome == _lib.HX_SOLVE_FAILED:
        warnings.warn("All solver attempts failed. Returning best stagnated result.")
    return u, st.rel_res, pc_name


def solve_system_gpu(nodes, elements, E, nu, bc_indicator, f, density, *, max_iter=40000, tol=1e-6,
                     method="native", solver="cg", use_precond=True, shift_factor=1.0e-6,
                     min_stiffness_threshold=1.0e-3, u_guess=None, config=None):
    """GPUSolver.jl:611-627 -> (x_full, final_rel_res, prec_name).  Raises
    RuntimeError when no element is active (GPUSolver.jl:132-134)."""
    if _lib.device_count() == 0:
        raise RuntimeError("CUDA not found.")  # GPUSolver.jl:620 -- no CPU fallback here
    return _solve(nodes, elements, E, nu, bc_indicator, f, density, max_iter=max_iter, tol=tol,
                  shift_factor=shift_factor, min_stiffness_threshold=min_stiffness_threshold, u_guess=u_guess,
                  config=config, prune_voids=False)


def adaptive_tolerance(config, solver_type="gpu"):
    """Solver.jl:99-120"""
    sp = config["solver_parameters"]
    base = float(sp.get("tolerance", 1.0e-6))
    it = config.get("current_outer_iter", 1)
    n = config.get("number_of_iterations", 30)
    t = min(max(float(it) / float(n), 0.0), 1.0)
    tol = 10.0 ** ((1.0 - t) * -5.0 + t * math.log10(base))
    tol = float(np.float32(tol))
    if solver_type == "gpu" and tol < 5.0e-7:
        tol = 5.0e-7
    return tol


def choose_solver(nNodes, nElem, config):
    """Solver.jl:13-36 (only the GPU path exists in this package)."""
    t = str(config["solver_parameters"].get("solver_type", "direct")).lower()
    if t == "gpu":
        if _lib.device_count() > 0:
            return "gpu"
        raise RuntimeError("CUDA not found.")  # GPUSolver.jl:620 -- no CPU fallback here
    raise NotImplementedError(f"solver_type '{t}': only the GPU path is implemented (CPU solvers are out of scope)")


def solve_system(nodes, elements, E, nu, bc_indicator, F, *, density=None, config=None,
                 min_stiffness_threshold=1.0e-3, prune_voids=True, u_prev=None):
    """Solver.jl:38-164 -> (U_full, res, prec)."""
    config = config or {"solver_parameters": {"solver_type": "gpu"}}
    ne = elements.shape[0]
    if density is None:
        density = np.ones(ne, dtype=np.float32)
    sp = config["solver_parameters"]
    stype = choose_solver(nodes.shape[0], ne, config)
    mode = str(sp.get("mode", "parity")).lower()
    tol = adaptive_tolerance(config, stype) if mode == "parity" else float(sp.get("tolerance", 1e-8))
    return _solve(nodes, elements, np.float32(E), np.float32(nu), bc_indicator, F, density,
                  max_iter=int(sp.get("max_iterations", 1000)), tol=tol,
                  shift_factor=float(np.float32(sp.get("diagonal_shift_factor", 1.0e-6))),
                  min_stiffness_threshold=float(np.float32(min_stiffness_threshold)), u_guess=u_prev, config=config,
                  prune_voids=prune_voids)
