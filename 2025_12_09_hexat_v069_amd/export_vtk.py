"""Legacy binary VTK of the solution (mirror of ExportVTK.export_solution,
ExportVTK.jl:20-172).

The data blocks -- node coordinates, displacement, external force,
boundary-condition indicator and displacement magnitude for every node; the
cells, L1 stress, von Mises stress and density of the elements at or above
the threshold -- are formed on the device (``hx_vtk_pack``) as big-endian
FP32/Int32 words, the byte order the file stores; this module writes the
ASCII section headers around them, byte for byte as the reference does.
"""
from __future__ import annotations

import numpy as np

from . import boundary as _bnd
from . import mesh as _mesh
from .solver import get_context


def _vtk_name(filename):
    low = filename.lower()
    if low.endswith(".vtu"):
        return filename[:-4] + ".vtk"
    if not low.endswith(".vtk"):
        return filename + ".vtk"
    return filename


def export_solution(nodes, elements, U_full, F, bc_indicator, principal_field, vonmises_field, full_stress_voigt,
                    l1_stress_norm_field, principal_dir_field, *, density=None, scale=1.0, threshold=0.01,
                    filename="solution_output.vtu", ctx=None):
    """ExportVTK.jl:20-172.  Returns the file name written."""
    ctx = ctx or get_context()
    grid = _mesh.grid_of(nodes, elements)
    if ctx.grid is None or tuple(ctx.grid[:6]) != tuple(grid):
        ctx.set_grid(*grid, ctx.grid[6] if ctx.grid else 0.3)
    ne, nn = ctx.ne, ctx.nn
    fixed = _bnd.fixed_dof_mask(bc_indicator) if bc_indicator is not None else None
    if density is None:  # every element is written (ExportVTK.jl:63-64)
        rho, thr = np.zeros(ne, np.float32), -np.inf
    else:
        rho, thr = np.asarray(density, dtype=np.float32), np.float32(threshold)
    w = ctx.vtk_pack(U_full, F, fixed, rho, l1_stress_norm_field, vonmises_field, thr)
    n = w["count"]
    pts = w["points"]
    with open(_vtk_name(filename), "wb") as fh:
        wr = lambda t: fh.write(t.encode("ascii"))
        wr("# vtk DataFile Version 3.0\n")
        wr("HEXA FEM Solution (Undeformed)\n")
        wr("BINARY\n")
        wr("DATASET UNSTRUCTURED_GRID\n")
        wr(f"POINTS {nn} float\n")
        fh.write(pts[:3 * nn].tobytes())
        wr(f"\nCELLS {n} {n * 9}\n")
        fh.write(w["cells"].tobytes())
        wr(f"\nCELL_TYPES {n}\n")
        fh.write(np.full(n, 12, dtype=">i4").tobytes())
        wr(f"\nPOINT_DATA {nn}\n")
        wr("VECTORS Displacement float\n")
        fh.write(pts[3 * nn:6 * nn].tobytes())
        wr("\nVECTORS External_Forces float\n")
        fh.write(pts[6 * nn:9 * nn].tobytes())
        wr("\nVECTORS Boundary_Conditions float\n")
        fh.write(pts[9 * nn:12 * nn].tobytes())
        wr("\nSCALARS Displacement_Magnitude float 1\n")
        wr("LOOKUP_TABLE default\n")
        fh.write(pts[12 * nn:13 * nn].tobytes())
        wr(f"\nCELL_DATA {n}\n")
        wr("SCALARS l1_stress_norm float 1\n")
        wr("LOOKUP_TABLE default\n")
        fh.write(w["l1"].tobytes())
        wr("\nSCALARS Von_Mises float 1\n")
        wr("LOOKUP_TABLE default\n")
        fh.write(w["vm"].tobytes())
        if density is not None:
            wr("\nSCALARS Element_Density float 1\n")
            wr("LOOKUP_TABLE default\n")
            fh.write(w["density"].tobytes())
    return _vtk_name(filename)
