"""YAML configuration in the reference grammar (mirror of Configuration.jl:17-173,
Helpers.calculate_element_distribution and Main.apply_hardware_profile!).

Keys added by this package (all optional, old configs keep their meaning):

    optimization_parameters:
      method: stress_ratio | simp_oc      # default stress_ratio (reference)
      penal: 3.0          # SIMP exponent p
      volfrac: 0.3        # OC volume fraction
      E_min: 1.0e-9       # void modulus (relative to material.E)
      move: 0.2           # OC move limit
      filter_radius: 1.5  # PDE filter radius r (R = r / (2 sqrt 3))
    solver_parameters:
      precision: fp64 | fp32_mfma_ir      # elasticity solve form
      preconditioner: jacobi | jacobi_amg | learned
      mode: parity | northstar            # reference quirks on / off
    geometry:
      nElem_x/nElem_y/nElem_z: explicit grid (overrides target_elem_count)
"""
from __future__ import annotations

import numpy as np
import yaml

from . import mesh as _mesh


def load_configuration(path):
    """Configuration.jl:17-23 (safe YAML loader)."""
    with open(path) as fh:
        return yaml.safe_load(fh)


def _as_int(v):
    return int(str(v).replace("_", "")) if isinstance(v, str) else int(v)


def calculate_element_distribution(length_x, length_y, length_z, target_elem_count):
    """Helpers.jl:163-180"""
    vol = length_x * length_y * length_z
    c = vol ** (1.0 / 3.0)
    base = target_elem_count ** (1.0 / 3.0)
    rnd = lambda v: int(round(v))  # Julia round(Int, v): nearest, ties to even
    nx = max(1, rnd(base * length_x / c))
    ny = max(1, rnd(base * length_y / c))
    nz = max(1, rnd(base * length_z / c))
    return nx, ny, nz, np.float32(length_x / nx), np.float32(length_y / ny), np.float32(length_z / nz), nx * ny * nz


def setup_geometry(config):
    """Configuration.jl:84-139 (+ explicit nElem_* keys)."""
    g = config["geometry"]
    lx, ly, lz = g["length_x"], g["length_y"], g["length_z"]
    skip = {"length_x", "length_y", "length_z", "target_elem_count", "shape_notes", "nElem_x", "nElem_y",
            "nElem_z", "nElem_x_computed", "nElem_y_computed", "nElem_z_computed", "dx_computed", "dy_computed",
            "dz_computed", "max_domain_dim"}
    shapes = [v for k, v in g.items() if k not in skip and isinstance(v, dict) and "type" in v]
    if all(k in g for k in ("nElem_x", "nElem_y", "nElem_z")):
        nx, ny, nz = (_as_int(g[k]) for k in ("nElem_x", "nElem_y", "nElem_z"))
        dx, dy, dz = np.float32(lx / nx), np.float32(ly / ny), np.float32(lz / nz)
        count = nx * ny * nz
    else:
        nx, ny, nz, dx, dy, dz, count = calculate_element_distribution(lx, ly, lz, _as_int(g["target_elem_count"]))
    return dict(nElem_x=nx, nElem_y=ny, nElem_z=nz, dx=dx, dy=dy, dz=dz, shapes=shapes, actual_elem_count=count,
                max_domain_dim=np.float32(max(lx, ly, lz)))


def initialize_density_field(nodes, elements, shapes, config):
    """Configuration.jl:147-173 -> (density, original_density, protected_mask, alpha_field)"""
    min_density = np.float32(config["optimization_parameters"].get("min_density", 1e-3))
    ne = elements.shape[0]
    density = np.ones(ne, dtype=np.float32)
    alpha = np.zeros(ne, dtype=np.float32)
    _mesh.apply_geometric_modifiers(density, alpha, nodes, elements, shapes, min_density)
    original = density.copy()
    protected = original != np.float32(1.0)
    return density, original, protected, alpha


def apply_hardware_profile(config):
    """Main.jl:118-155: the profile overrides tolerance/shift/solver type.
    "MI355X" is the MI355X profile: FP64 tolerance 1e-8, no shift cap."""
    gpu = str(config.get("gpu_profile", "RTX")).upper()
    growth = config.setdefault("growth_settings", {})
    solver = config.setdefault("solver_parameters", {})
    if gpu in ("H", "H200", "H100"):
        growth.update(max_background_elements=1_500_000_000, max_growth_rate=1.5, gpu_solver_safety_factor=0.98)
        solver.update(tolerance=1.0e-12, diagonal_shift_factor=1.0e-10, solver_type="gpu")
    elif gpu == "V100":
        growth.update(max_background_elements=150_000_000, max_growth_rate=1.3, gpu_solver_safety_factor=0.95)
        solver.update(tolerance=1.0e-10, diagonal_shift_factor=1.0e-9, solver_type="gpu")
    elif gpu == "RTX":
        growth.update(max_background_elements=200_000_000, max_growth_rate=1.2, gpu_solver_safety_factor=0.90)
        solver.update(tolerance=1.0e-6, solver_type="gpu")
    elif gpu == "MI355X":
        growth.update(max_background_elements=2_000_000_000, max_growth_rate=1.5)
        solver.setdefault("tolerance", 1.0e-8)
        solver["solver_type"] = "gpu"
    else:
        growth["max_background_elements"] = 100_000_000
    config["hardware_profile_applied"] = gpu
    return config
