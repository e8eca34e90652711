"""Synthetic inputs of the measured configurations, in the reference's
location grammar (Boundary.jl:56-291) on structured grids of unit cubes.

* ``cantilever``: clamp [0, ':', ':'] DoFs 1-3; total load (0, -1, 0) split
  over the nodes of [100%, 0, ':'] (the 60x20x20 and 256x128x128 cases);
* ``mbb_half``: half MBB beam -- symmetry rollers [0, ':', ':'] DoF 1, support
  line [100%, 0, ':'] DoF 2, z-symmetry [':', ':', 0] DoF 3, unit load
  [0, 100%, 50%] F (0, -1, 0) (the 192x96x96 solver-only case);
* ``block_loads``: 8 unit loads, directions uniform on the sphere, at 8
  distinct seeded nodes of the x = nx face (the 128x64x64 batched case);
* ``smoothed_density``: rho ~ U(0.01, 1) with one PDE-filter pass of radius r
  on the device, clamped to [0.01, 1] (seeded density fields).
"""
from __future__ import annotations

import math

import numpy as np


def _ijk(nx, ny, nz):
    NX, NY, NZ = nx + 1, ny + 1, nz + 1
    k, j, i = np.meshgrid(np.arange(NZ), np.arange(NY), np.arange(NX), indexing="ij")
    return i.ravel(), j.ravel(), k.ravel()


def cantilever(nx, ny, nz):
    """-> (fixed DOF mask uint8, load vector)"""
    i, j, k = _ijk(nx, ny, nz)
    fixed = np.repeat((i == 0)[:, None], 3, axis=1).ravel().astype(np.uint8)
    f = np.zeros(3 * i.size)
    edge = np.nonzero((i == nx) & (j == 0))[0]
    f[3 * edge + 1] = -1.0 / edge.size
    return fixed, f


def mbb_half(nx, ny, nz):
    """-> (fixed DOF mask uint8, load vector)"""
    i, j, k = _ijk(nx, ny, nz)
    fixed = np.zeros((i.size, 3), dtype=np.uint8)
    fixed[i == 0, 0] = 1
    fixed[(i == nx) & (j == 0), 1] = 1
    fixed[k == 0, 2] = 1
    f = np.zeros(3 * i.size)
    top = int(np.nonzero((i == 0) & (j == ny) & (k == nz // 2))[0][0])
    f[3 * top + 1] = -1.0
    return fixed.ravel(), f


def block_loads(nx, ny, nz, nrhs=8, seed=8):
    """-> (nrhs x ndof) loads: unit forces in directions N(0, I3)/|.| at nrhs
    distinct nodes of the x = nx face drawn without replacement."""
    NX, NY, NZ = nx + 1, ny + 1, nz + 1
    rng = np.random.default_rng(seed)
    F = np.zeros((nrhs, 3 * NX * NY * NZ))
    picks = rng.choice(NY * NZ, size=nrhs, replace=False)
    for c in range(nrhs):
        j, k = int(picks[c] % NY), int(picks[c] // NY)
        d = rng.standard_normal(3)
        node = nx + NX * (j + NY * k)
        F[c, 3 * node:3 * node + 3] = d / np.linalg.norm(d)
    return F


def smoothed_density(ctx, ne, seed, r):
    """rho ~ U(0.01, 1) filtered once by the PDE filter of radius r (R = r/(2 sqrt 3)), clamped to [0.01, 1]."""
    rho = np.random.default_rng(seed).uniform(0.01, 1.0, ne)
    out, _ = ctx.helmholtz(r / (2 * math.sqrt(3)), rho, quirk=0, tol=1e-8, max_iter=2000, check_every=1)
    return np.clip(out, 0.01, 1.0)
