"""Helmholtz PDE density filter (mirror of GPUHelmholtz.jl:344-376).

``apply_gpu_filter`` keeps the reference's contract: it returns
``(filtered, time, iterations, relres)`` and signals failure with an empty
array instead of raising (GPUHelmholtz.jl:365-375).  The kernels solve the
nodal system (R^2 Ks + Ms) x = sum_e rho_e vol/8 with Jacobi-PCG from x0 = 0
and return the element mean of x.  ``quirk=True`` (default, the reference's
element) puts the scalar element's top face at z = +dz, so a constant field
comes back scaled by exactly 2/3; ``quirk=False`` is the mass-preserving
filter used by the SIMP loop.  The reference's memory-driven blocked variant
(:231-341) is not needed: the whole grid fits in HBM.
"""
from __future__ import annotations

import time

import numpy as np

from . import _lib
from .solver import get_context


def apply_gpu_filter(density, nElem_x, nElem_y, nElem_z, dx, dy, dz, radius, elements=None, tol=1.0e-5, *,
                     quirk=True, max_iter=200, check_every=10, nu=0.3):
    try:
        ctx = get_context()
        key = (int(nElem_x), int(nElem_y), int(nElem_z), float(dx), float(dy), float(dz))
        if ctx.grid is None or ctx.grid[:6] != key:
            ctx.set_grid(*key, nu=ctx.grid[6] if ctx.grid else nu)
        t0 = time.perf_counter()
        out, st = ctx.helmholtz(float(radius), np.asarray(density, dtype=np.float64), quirk=int(bool(quirk)),
                                tol=float(tol), max_iter=int(max_iter), check_every=int(check_every))
        out = out.astype(np.asarray(density).dtype, copy=False)
        return out, time.perf_counter() - t0, st.iterations, st.rel_res
    except (_lib.HexatError, ValueError) as e:  # GPUHelmholtz.jl:365-375
        print(f"  [GPU Filter] CRITICAL ERROR: {e}.")
        return np.zeros(0, dtype=np.float32), 0.0, 0, 0.0
