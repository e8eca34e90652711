"""Element stress field (mirror of Stress.jl:57-100) on the HIP kernel.

Returns the reference's five fields, in its (component, element) layout:
principal (3, Ne) descending, von Mises (Ne), Voigt (6, Ne) in the order
xx, yy, zz, xy, yz, xz, L1 = sum |principal| (Ne) and the unit direction of
the largest principal stress (3, Ne; sign arbitrary).  Stress is evaluated at
the element centroid for modulus E*rho_e; elements with rho_e <= 1e-6 give
zeros (Stress.jl:85).
"""
from __future__ import annotations

import numpy as np

from .solver import _bind_grid, get_context


def compute_stress_field(nodes, elements, U, E, nu, density, *, moduli=None):
    ctx = get_context()
    _bind_grid(ctx, nodes, elements, nu)
    density = np.asarray(density, dtype=np.float64)
    Ee = moduli if moduli is not None else np.where(density > 1e-6, float(E) * density, 0.0)
    out = ctx.element_stress(Ee, np.asarray(U, dtype=np.float64))
    return out["principal"].T.copy(), out["vm"], out["voigt"].T.copy(), out["l1"], out["dir"].T.copy()
