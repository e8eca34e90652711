"""Boundary conditions and loads in the reference's YAML grammar (mirror of
Boundary.jl:18-291).

A ``location`` is three components, each ``":"`` (all), a number (values in
[0, 1] are fractions of the extent, others absolute coordinates) or a
percentage string.  Matching uses the tolerance 1e-4 * max(1, extent).
Forces given by ``location`` are split evenly over the matched nodes; a point
location that matches no node falls back to the nearest node (ties go to the
larger x, then y, then z).
"""
from __future__ import annotations

import numpy as np


def resolve_coordinate_value(spec, cmin, cmax):
    """Boundary.jl:117-130"""
    if isinstance(spec, bool):
        raise ValueError(f"Invalid location component: {spec}")
    if isinstance(spec, (int, float, np.floating, np.integer)):
        if 0.0 <= spec <= 1.0:
            return np.float32(cmin + spec * (cmax - cmin))
        return np.float32(spec)
    if isinstance(spec, str) and spec.endswith("%"):
        frac = min(max(float(spec[:-1]) / 100.0, 0.0), 1.0)
        return np.float32(cmin + frac * (cmax - cmin))
    return np.float32(cmin)


def interpret_location_component(spec, coords, cmin, cmax):
    """Boundary.jl:79-115"""
    tol = np.float32(1e-4) * max(np.float32(1.0), abs(cmax - cmin))
    if isinstance(spec, str) and spec == ":":
        return np.ones(coords.shape[0], bool)
    if isinstance(spec, (int, float, np.floating, np.integer)) and not isinstance(spec, bool):
        return np.abs(coords - resolve_coordinate_value(spec, cmin, cmax)) <= tol
    if isinstance(spec, str) and spec.endswith("%"):
        return np.abs(coords - resolve_coordinate_value(spec, cmin, cmax)) <= tol
    raise ValueError(f"Invalid location component: {spec}")


def get_nodes_by_location(loc, nodes):
    """Boundary.jl:56-74 -> 0-based node indices"""
    m = np.ones(nodes.shape[0], bool)
    for d in range(3):
        col = nodes[:, d]
        m &= interpret_location_component(loc[d], col, col.min(), col.max())
    return np.nonzero(m)[0]


def get_affected_nodes(spec, nodes):
    """Boundary.jl:18-49 (explicit 'node' ids are 1-based, clamped)"""
    n = nodes.shape[0]
    if "node" in spec:
        raw = spec["node"]
        if isinstance(raw, (int, np.integer)):
            return np.array([min(max(int(raw), 1), n) - 1])
        if isinstance(raw, (list, tuple)):
            return np.unique([min(max(int(r), 1), n) - 1 for r in raw])
        raise ValueError("'node' must be an integer or an array of integers")
    if "location" in spec:
        if len(spec["location"]) < 3:
            raise ValueError("Location specification must have at least 3 components (x,y,z)")
        return get_nodes_by_location(spec["location"], nodes)
    raise ValueError("Specification must include either 'node' or 'location'")


def get_bc_indicator(nNodes, nodes, bc_data):
    """Boundary.jl:135-160 -> (nNodes, 3) float32, 1 = fixed"""
    ind = np.zeros((nNodes, 3), dtype=np.float32)
    for bc in bc_data:
        dofs = list(bc["DoFs"])
        for d in dofs:
            if d < 1 or d > 3:
                raise ValueError(f"Invalid DoF index: {d} (must be 1..3).")
        nodes_hit = get_affected_nodes(bc, nodes)
        for d in dofs:
            ind[nodes_hit, d - 1] = 1.0
    return ind


def find_nearest_node(target, nodes):
    """Boundary.jl:195-224 (0-based)"""
    d2 = np.sum((nodes.astype(np.float64) - np.asarray(target, np.float64)) ** 2, axis=1)
    best = int(np.argmin(d2))
    near = np.nonzero(np.abs(d2 - d2[best]) <= 1e-9)[0]
    if near.size > 1:  # tie -> larger x, then y, then z
        keys = [tuple(nodes[i]) for i in near]
        best = int(near[max(range(near.size), key=lambda t: keys[t])])
    return best


def apply_external_forces(F, forces_data, nodes, elements=None):
    """Boundary.jl:232-291: adds the loads into F (length 3*nNodes) in place."""
    lo, hi = nodes.min(axis=0), nodes.max(axis=0)
    for force in forces_data:
        hit = get_affected_nodes(force, nodes)
        if hit.size == 0 and "location" in force:
            loc = force["location"]
            if all(not (isinstance(c, str) and c == ":") for c in loc):
                tgt = [resolve_coordinate_value(loc[d], lo[d], hi[d]) for d in range(3)]
                hit = np.array([find_nearest_node(tgt, nodes)])
        if hit.size == 0:
            continue
        fv = np.zeros(3, dtype=F.dtype)
        raw = list(force["F"])[:3]
        fv[:len(raw)] = raw
        scale = 1.0 / hit.size if "location" in force else 1.0
        for d in range(3):
            np.add.at(F, 3 * hit + d, scale * fv[d])
    return F


def fixed_dof_mask(bc_indicator):
    """GPUSolver.jl:112-122 as a per-DOF mask (node-major)."""
    return (np.asarray(bc_indicator) > 0).astype(np.uint8).ravel()


# ---------------------------------------------------------------------------
# the same mapping on the device for a structured grid (hx_bc_indicator,
# hx_external_forces): the host resolves each location to FP32 values as
# above, the library matches nodes, splits loads and finds nearest nodes
# ---------------------------------------------------------------------------

def _location(loc, nodes):
    lo, hi = nodes.min(axis=0), nodes.max(axis=0)
    all3, val3 = [], []
    for d in range(3):
        spec = loc[d]
        if isinstance(spec, str) and spec == ":":
            all3.append(1)
            val3.append(0.0)
        elif isinstance(spec, (int, float, np.floating, np.integer)) and not isinstance(spec, bool) or (
                isinstance(spec, str) and spec.endswith("%")):
            all3.append(0)
            val3.append(float(resolve_coordinate_value(spec, lo[d], hi[d])))
        else:
            raise ValueError(f"Invalid location component: {spec}")
    return all3, val3


def _all_located(specs):
    return all("location" in s and "node" not in s and len(s["location"]) >= 3 for s in specs)


def get_bc_indicator_device(ctx, nodes, bc_data):
    """get_bc_indicator on the device (the context's grid is the mesh's);
    the mask is also installed in the context.  Specs by explicit node ids
    take the host path."""
    if not _all_located(bc_data):
        ind = get_bc_indicator(nodes.shape[0], nodes, bc_data)
        ctx.set_fixed(fixed_dof_mask(ind))
        return ind
    locs, bits = [], []
    for bc in bc_data:
        b = 0
        for d in bc["DoFs"]:
            if d < 1 or d > 3:
                raise ValueError(f"Invalid DoF index: {d} (must be 1..3).")
            b |= 1 << (d - 1)
        locs.append(_location(bc["location"], nodes))
        bits.append(b)
    fixed = ctx.bc_indicator(locs, bits)
    return fixed.reshape(-1, 3).astype(np.float32)


def apply_external_forces_device(ctx, F, forces_data, nodes):
    """apply_external_forces on the device; returns the new F."""
    if not _all_located(forces_data):
        return apply_external_forces(F, forces_data, nodes)
    locs, fv, point = [], [], []
    for f in forces_data:
        locs.append(_location(f["location"], nodes))
        v = np.zeros(3)
        raw = list(f["F"])[:3]
        v[:len(raw)] = raw
        fv.append(v)
        point.append(all(not (isinstance(c, str) and c == ":") for c in f["location"]))
    out, _ = ctx.external_forces(locs, np.array(fv), point, F)
    return out
