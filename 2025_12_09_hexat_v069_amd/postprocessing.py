"""Viewer/checkpoint file (mirror of Postprocessing.export_binary_for_web,
Postprocessing.jl:159-244, and Configuration.load_checkpoint,
Configuration.jl:31-77).

Layout, little-endian, as the reference writes it and its viewer parses it
(PRE_POST/post24.html:596-654):

    u32 magic 0x48455841 ("HEXA")   u32 version 2
    i32 iteration   f32 radius   f32 threshold
    u32 count       f32 dx   f32 dy   f32 dz
    f32 centroids[count][3]   f32 density[count]   f32 signed_l1[count]
    u32 json_len    u8 json[json_len]

The packing (which elements, their centroids, the signed L1 stress) runs on
the device (``hx_web_pack``); this module writes and reads the bytes.

The JSON block holds the same keys as the reference's, but not the same
bytes: Julia's Dict iteration order is hash order, so no byte-identical key
order can be asked of it.

Restart: the reference's file holds only the ``count`` elements at or above
the threshold and its reader returns that many densities
(Configuration.jl:55-57), so a restart on the full mesh is refused unless
every element was kept (Main.jl:259-261).  Files written here carry, after
the JSON block, a restart section the viewer and the reference's reader never
reach (both stop at the JSON block):

    u32 magic 0x53525848 ("HXRS")   u32 version 1   u32 nx   u32 ny   u32 nz
    f32 density[nx*ny*nz]           (the whole field, element order)

``load_checkpoint`` returns that whole field when the section is there, and
the ``count`` densities of a file without it (``restore_density`` then keeps
the reference's refusal); ``full_density_from_kept`` rebuilds a whole field
from a file without it, placing each kept element by its centroid and giving
the others a fill value (the culled values are not in such a file).
"""
from __future__ import annotations

import copy
import json
import os
import struct

import numpy as np

from . import mesh as _mesh
from .solver import get_context

MAGIC = 0x48455841
VERSION = 2
_HEAD = struct.Struct("<IIiffIfff")
RS_MAGIC = 0x53525848
_RS_HEAD = struct.Struct("<IIIII")


def _f32(v):
    """A Float32 value as the shortest decimal that reads back to it."""
    return float(str(np.float32(v)))


def _meta(config, iteration, radius, threshold):
    """Postprocessing.jl:198-232: a copy of the configuration plus the
    iteration state, 'action' synthesised from 'stiffness_ratio' for the
    viewer, and the loads/bcs/settings aliases."""
    meta = copy.deepcopy(config) if config is not None else {}
    meta["iteration"] = int(iteration)
    meta["radius"] = _f32(radius)
    meta["threshold"] = _f32(threshold)
    geo = meta.get("geometry")
    if isinstance(geo, dict):
        for shape in geo.values():
            if isinstance(shape, dict) and "type" in shape and "action" not in shape and "stiffness_ratio" in shape:
                shape["action"] = "add" if float(np.float32(shape["stiffness_ratio"])) > 0 else "remove"
    meta["loads"] = (config or {}).get("external_forces", [])
    meta["bcs"] = (config or {}).get("boundary_conditions", [])
    meta["settings"] = (config or {}).get("optimization_parameters", {})
    return meta


def write_bin(filename, iteration, radius, threshold, dx, dy, dz, centroids, density, signed_l1, meta,
              full_density=None, grid=None):
    """Write the byte layout of Postprocessing.jl:236-258, then (with
    full_density and grid = (nx, ny, nz)) the restart section."""
    count = int(len(density))
    js = json.dumps(meta, separators=(",", ":")).encode("utf-8")
    with open(filename, "wb") as fh:
        fh.write(_HEAD.pack(MAGIC, VERSION, int(iteration), np.float32(radius), np.float32(threshold), count,
                            np.float32(dx), np.float32(dy), np.float32(dz)))
        fh.write(np.ascontiguousarray(centroids, dtype="<f4").reshape(count * 3).tobytes())
        fh.write(np.ascontiguousarray(density, dtype="<f4").tobytes())
        fh.write(np.ascontiguousarray(signed_l1, dtype="<f4").tobytes())
        fh.write(struct.pack("<I", len(js)))
        fh.write(js)
        if full_density is not None:
            nx, ny, nz = (int(v) for v in grid)
            fd = np.ascontiguousarray(full_density, dtype="<f4").ravel()
            if fd.size != nx * ny * nz:
                raise ValueError("restart section: the density must cover the grid")
            fh.write(_RS_HEAD.pack(RS_MAGIC, 1, nx, ny, nz))
            fh.write(fd.tobytes())


def export_binary_for_web(filename, nodes, elements, density, l1_stress, principal_field, geom, threshold,
                          iteration, current_radius, config, *, ctx=None):
    """Postprocessing.export_binary_for_web (Postprocessing.jl:159-259).

    ``principal_field`` is (3, Ne) as ``stress.compute_stress_field`` returns
    it; ``geom`` carries dx, dy, dz (attributes or keys).  Returns the number
    of elements written; with none at or above the threshold no file is
    written (Postprocessing.jl:174)."""
    grid = _mesh.grid_of(nodes, elements)
    gd = (lambda k: geom[k]) if isinstance(geom, dict) else (lambda k: getattr(geom, k))
    dx, dy, dz = float(gd("dx")), float(gd("dy")), float(gd("dz"))
    if ctx is None:
        ctx = get_context()
    if ctx.grid is None or tuple(ctx.grid[:6]) != tuple(grid):  # the packing does not depend on nu: keep it
        ctx.set_grid(*grid, ctx.grid[6] if ctx.grid else 0.3)
    pr = np.asarray(principal_field, dtype=np.float32)
    if pr.shape == (3, ctx.ne):
        pr = pr.T
    cent, dens, sl1 = ctx.web_pack(density, l1_stress, pr, threshold)
    if dens.size == 0:
        return 0
    write_bin(filename, iteration, current_radius, threshold, dx, dy, dz, cent, dens, sl1,
              _meta(config, iteration, current_radius, threshold), full_density=density, grid=grid[:3])
    return int(dens.size)


def read_bin(filename):
    """Every field of the file, as the viewer reads it (post24.html:596-654):
    dict with iteration, radius, threshold, dx, dy, dz, centroids (count x 3),
    density, signed_l1 and config."""
    with open(filename, "rb") as fh:
        buf = fh.read()
    magic, version, it, radius, thr, count, dx, dy, dz = _HEAD.unpack_from(buf, 0)
    if magic != MAGIC:
        raise ValueError("Invalid file format. Not a HEXA checkpoint.")
    off = _HEAD.size
    need = off + 20 * count + 4
    if len(buf) < need:
        raise ValueError(f"truncated checkpoint: {len(buf)} bytes, {need} needed for {count} elements")
    cent = np.frombuffer(buf, "<f4", 3 * count, off).reshape(count, 3).copy()
    off += 12 * count
    dens = np.frombuffer(buf, "<f4", count, off).copy()
    off += 4 * count
    sl1 = np.frombuffer(buf, "<f4", count, off).copy()
    off += 4 * count
    (jlen,) = struct.unpack_from("<I", buf, off)
    off += 4
    if len(buf) < off + jlen:
        raise ValueError("truncated checkpoint: JSON block cut short")
    cfg = json.loads(buf[off:off + jlen].decode("utf-8"))
    off += jlen
    out = {"version": version, "iteration": it, "radius": radius, "threshold": thr, "dx": dx, "dy": dy, "dz": dz,
           "centroids": cent, "density": dens, "signed_l1": sl1, "config": cfg, "full_density": None, "grid": None}
    if len(buf) >= off + _RS_HEAD.size:
        rs, rv, nx, ny, nz = _RS_HEAD.unpack_from(buf, off)
        n = nx * ny * nz
        if rs == RS_MAGIC and rv == 1 and len(buf) >= off + _RS_HEAD.size + 4 * n:
            out["full_density"] = np.frombuffer(buf, "<f4", n, off + _RS_HEAD.size).copy()
            out["grid"] = (nx, ny, nz)
    return out


def full_density_from_kept(d, nx, ny, nz, fill):
    """A whole density field from a file without the restart section: each
    kept element at the cell its centroid falls in (i = floor(x / dx), ...),
    every other element `fill`."""
    rho = np.full(nx * ny * nz, np.float32(fill), dtype=np.float32)
    c = d["centroids"].astype(np.float64)
    i = np.clip(np.floor(c[:, 0] / d["dx"]).astype(int), 0, nx - 1)
    j = np.clip(np.floor(c[:, 1] / d["dy"]).astype(int), 0, ny - 1)
    k = np.clip(np.floor(c[:, 2] / d["dz"]).astype(int), 0, nz - 1)
    rho[i + nx * (j + ny * k)] = d["density"]
    return rho


def load_checkpoint(filename):
    """Configuration.load_checkpoint (Configuration.jl:31-77):
    (config, density, restart_iter, restart_radius, restart_threshold); the
    density is the whole field when the file has the restart section."""
    if not os.path.isfile(filename):
        raise FileNotFoundError("Checkpoint file not found.")
    d = read_bin(filename)
    rho = d["full_density"] if d["full_density"] is not None else d["density"]
    return d["config"], rho, int(d["iteration"]), float(d["radius"]), float(d["threshold"])


def restore_density(density, n_elements):
    """Main.jl:259-261: a restart density must cover the whole mesh."""
    if len(density) != n_elements:
        raise ValueError(f"Restart Mismatch: Checkpoint density size ({len(density)}) != "
                         f"Generated Mesh size ({n_elements}).")
    return np.asarray(density, dtype=np.float32)
