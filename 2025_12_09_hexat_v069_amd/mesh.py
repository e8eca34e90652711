"""Structured hexahedral mesh (mirror of Mesh.jl, MeshUtilities.jl and the
geometric modifiers of MeshShapeProcessing.jl).

Arrays follow the reference's layout: ``nodes`` is (nNodes, 3) float32,
``elements`` is (nElem, 8) int64 with **1-based** node ids in the order
n1..n8 of Mesh.jl:45-52.  The HIP kernels never receive these arrays: they
derive connectivity from (i, j, k); ``grid_of`` recovers (nx, ny, nz, dx,
dy, dz) from a structured mesh and refuses anything else.
"""
from __future__ import annotations

import numpy as np


def node_index(i, j, k, nNodes_x, nNodes_y):
    """Mesh.jl:13-15 (1-based)."""
    return i + (j - 1) * nNodes_x + (k - 1) * nNodes_x * nNodes_y


def generate_mesh(nElem_x: int, nElem_y: int, nElem_z: int, dx=1.0, dy=1.0, dz=1.0):
    """Mesh.jl:23-60 -> (nodes, elements, dims)."""
    NX, NY, NZ = nElem_x + 1, nElem_y + 1, nElem_z + 1
    k, j, i = np.meshgrid(np.arange(NZ), np.arange(NY), np.arange(NX), indexing="ij")
    nodes = np.stack([i.ravel() * np.float32(dx), j.ravel() * np.float32(dy), k.ravel() * np.float32(dz)],
                     axis=1).astype(np.float32)
    k, j, i = np.meshgrid(np.arange(1, nElem_z + 1), np.arange(1, nElem_y + 1), np.arange(1, nElem_x + 1),
                          indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    n = lambda a, b, c: node_index(a, b, c, NX, NY)
    elements = np.stack([n(i, j, k), n(i + 1, j, k), n(i + 1, j + 1, k), n(i, j + 1, k), n(i, j, k + 1),
                         n(i + 1, j, k + 1), n(i + 1, j + 1, k + 1), n(i, j + 1, k + 1)], axis=1).astype(np.int64)
    return nodes, elements, (NX, NY, NZ)


def grid_of(nodes, elements):
    """Recover (nx, ny, nz, dx, dy, dz) of a structured mesh made by
    generate_mesh.  Raises ValueError for any other mesh (pruned or irregular
    meshes are expressed with element moduli and DOF masks instead)."""
    nodes = np.asarray(nodes)
    elements = np.asarray(elements)
    xs, ys, zs = (np.unique(nodes[:, d]) for d in range(3))
    NX, NY, NZ = xs.size, ys.size, zs.size
    if NX * NY * NZ != nodes.shape[0] or elements.shape != ((NX - 1) * (NY - 1) * (NZ - 1), 8):
        raise ValueError("mesh is not a full structured grid")
    dx = float(xs[1] - xs[0]) if NX > 1 else 1.0
    dy = float(ys[1] - ys[0]) if NY > 1 else 1.0
    dz = float(zs[1] - zs[0]) if NZ > 1 else 1.0
    # spot-check connectivity against the generator (first, middle, last element)
    nx, ny, nz = NX - 1, NY - 1, NZ - 1
    for e in {0, elements.shape[0] // 2, elements.shape[0] - 1}:
        i, j, k = e % nx + 1, (e // nx) % ny + 1, e // (nx * ny) + 1
        want = [node_index(i + a, j + b, k + c, NX, NY)
                for (a, b, c) in [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1),
                                  (0, 1, 1)]]
        if list(elements[e]) != want:
            raise ValueError("element connectivity differs from Mesh.generate_mesh")
    return nx, ny, nz, dx, dy, dz


def element_centroids(nodes, elements):
    """MeshUtilities.jl:18-33 (vectorised)."""
    return nodes[elements - 1].mean(axis=1)


def apply_geometric_modifiers(density, alpha_field, nodes, elements, shapes, min_density):
    """MeshShapeProcessing.jl:16-84: sphere/box regions by element centroid;
    stiffness_ratio 0 -> void, < 0 -> |ratio| with thermal flag, > 0 -> ratio."""
    if not shapes:
        return
    c = element_centroids(nodes, elements)
    for shape in shapes:
        typ = str(shape.get("type", "")).lower()
        inside = np.zeros(c.shape[0], bool)
        if typ == "sphere" and "center" in shape and "diameter" in shape:
            ctr = np.asarray(shape["center"], dtype=np.float32)
            r = np.float32(shape["diameter"]) / 2
            inside = np.sum((c - ctr) ** 2, axis=1) <= r * r
        elif typ == "box" and "center" in shape:
            ctr = np.asarray(shape["center"], dtype=np.float32)
            if "size" in shape and len(shape["size"]) >= 3:
                half = np.asarray(shape["size"][:3], dtype=np.float32) / 2
            elif "side" in shape:
                half = np.full(3, np.float32(shape["side"]) / 2)
            else:
                continue
            inside = np.all(np.abs(c - ctr) <= half, axis=1)
        if not inside.any():
            continue
        ratio = np.float32(shape.get("stiffness_ratio", 0.0))
        if ratio == 0:
            density[inside] = min_density
            alpha_field[inside] = 0
        elif ratio < 0:
            density[inside] = abs(ratio)
            alpha_field[inside] = 1
        else:
            density[inside] = ratio
            alpha_field[inside] = 0


def refine_mesh_and_fields(ctx, dims, density, alpha, target_active, lengths, max_growth_rate=1.2,
                           hard_element_limit=800_000_000, free_bytes=None):
    """MeshRefiner.refine_mesh_and_fields (MeshRefiner.jl:52-156) on a
    structured grid: the new element count from the active ratio, capped by
    the growth rate, the memory left (free_bytes; device memory by default)
    and the hard limit; no change below 5 % growth.  The field map runs on the
    device (hx_regrid).  Returns (new_dims, density, alpha) -- dims as element
    counts (nx, ny, nz)."""
    from .configuration import calculate_element_distribution
    nx, ny, nz = dims
    n_old = nx * ny * nz
    density = np.asarray(density, dtype=np.float32)
    ratio = max(0.01, int((density > np.float32(0.05)).sum()) / n_old)
    ideal = int(round(target_active / ratio))
    rate = int(round(n_old * max_growth_rate))
    if free_bytes is None:
        used = ctx.memory_used_gb()
        free_bytes = max(0.0, ctx.device_info()["mem_gb"] - used) * 1024 ** 3
    per = 180 if hard_element_limit > 500_000_000 else 250
    buf = max(free_bytes * (0.05 if hard_element_limit > 500_000_000 else 0.20), (2.0 if hard_element_limit >
                                                                                 500_000_000 else 4.0) * 1024 ** 3)
    ram = int((free_bytes - buf) // per) if free_bytes > buf else 1_000_000
    new_total = min(ideal, rate, ram, hard_element_limit)
    if new_total < n_old * 1.05:
        return dims, density, np.asarray(alpha, dtype=np.float32)
    lx, ly, lz = (np.float32(v) for v in lengths)
    nnx, nny, nnz, *_ = calculate_element_distribution(lx, ly, lz, new_total)
    out = ctx.regrid((nx, ny, nz), (nnx, nny, nnz), (lx, ly, lz), np.stack([density, np.asarray(alpha, np.float32)]))
    return (nnx, nny, nnz), out[0], out[1]
