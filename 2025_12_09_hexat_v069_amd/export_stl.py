"""Isosurface STL of the density (mirror of
Postprocessing.export_smooth_watertight_stl, Postprocessing.jl:53-148).

The nodal field, its refinement and the isosurface triangles are formed on
the device (``hx_stl_pack``); this module writes the binary STL: an 80-byte
header, the triangle count, and per triangle the unit normal, three vertices
(FP32, little-endian) and a zero attribute word.  No triangle: no file, as in
the reference (:97-100).  The reference extracts the surface with
MarchingCubes.jl; the triangles here come from a generated crack-free case
table, so the surface's position is the same but its triangulation differs.
"""
from __future__ import annotations

import os
import struct
import warnings

import numpy as np

from .solver import get_context


def _g(geom, k):
    return geom[k] if isinstance(geom, dict) else getattr(geom, k)


def stl_bytes(triangles, threshold):
    """The file: header, count, 50 bytes per triangle."""
    head = ("Binary STL generated by HEXA TopOpt. Threshold: %s" % str(np.float32(threshold))).encode()[:80]
    rec = np.zeros(len(triangles), dtype=[("f", "<f4", 12), ("attr", "<u2")])
    rec["f"] = triangles
    return head.ljust(80, b"\0") + struct.pack("<I", len(triangles)) + rec.tobytes()


def export_smooth_watertight_stl(density, geom, threshold, filename, subdivision_level=1, ctx=None):
    """Postprocessing.jl:53-148.  Returns the number of triangles written."""
    ctx = ctx or get_context()
    nx, ny, nz = int(_g(geom, "nElem_x")), int(_g(geom, "nElem_y")), int(_g(geom, "nElem_z"))
    if len(density) != nx * ny * nz:  # :66
        return 0
    key = (nx, ny, nz, float(_g(geom, "dx")), float(_g(geom, "dy")), float(_g(geom, "dz")))
    if ctx.grid is None or tuple(ctx.grid[:6]) != key:
        ctx.set_grid(*key, nu=ctx.grid[6] if ctx.grid else 0.3)
    tris = ctx.stl_pack(density, threshold, subdivision_level)
    if len(tris) == 0:
        warnings.warn(f"No isosurface found at threshold {np.float32(threshold)}.")
        return 0
    d = os.path.dirname(filename)
    if d:
        os.makedirs(d, exist_ok=True)
    with open(filename, "wb") as fh:
        fh.write(stl_bytes(tris, threshold))
    return len(tris)
