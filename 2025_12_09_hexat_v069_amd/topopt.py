"""Density updates.

``update_density`` mirrors TopOpt.update_density! (TopOpt.jl:139-405), the
reference's stress-ratio heuristic, step for step: optional stress
pre-filter with R = ratio * mean(h) (:176-211), the radius schedule with the
1.2*h guard (:216-252), R_eff = R/2.5 (:255), proposal
clamp(L1/allowable/E) (:262-271), Helmholtz density filter (:276-322, the
reference's element, i.e. ``quirk``), clamp and max change (:330-343),
progressive culling of the weakest candidates (:350-395) and the protected
restore (:398-402).  The two filter solves run on the GPU; the element-wise
steps are cheap host loops in the reference too and stay vectorised numpy on
the caller's arrays, which the function mutates in place like the original.

``SimpOptimizer`` is the SIMP/OC path: SIMP E(rho) = Emin + rho^p (E0-Emin),
compliance sensitivities, PDE sensitivity filter, OC update and PDE density
filter, all device-resident in libhexat (one ``hx_to_step`` per iteration).
"""
from __future__ import annotations

import math
import time

import numpy as np

from . import _lib
from .gpu_helmholtz import apply_gpu_filter
from .solver import get_context


def radius_schedule(opt, iter_, n_iter, max_domain_dim, avg_h):
    """TopOpt.jl:216-255 -> (R_length, R_effective)"""
    f = lambda k, d: np.float32(opt.get(k, d))
    Ri = f("filter_R_init_perc", 0.0) / np.float32(100) * np.float32(max_domain_dim)
    Rm = f("filter_R_interm_perc", 0.0) / np.float32(100) * np.float32(max_domain_dim)
    Rf = f("filter_R_final_perc", 0.0) / np.float32(100) * np.float32(max_domain_dim)
    frac = f("filter_R_interm_iter_perc", 50.0) / np.float32(100)
    it_m = max(1, int(round(float(frac) * n_iter)))
    c = min(iter_, n_iter)
    if c <= it_m:
        t = np.float32(c - 1) / np.float32(it_m - 1) if it_m > 1 else np.float32(0)
        R = Ri * (1 - t) + Rm * t
    else:
        t = np.float32(c - it_m) / np.float32(n_iter - it_m) if n_iter > it_m else np.float32(0)
        R = Rm * (1 - t) + Rf * t
    R_safe = np.float32(1.2) * np.float32(avg_h)
    if R < R_safe:
        R = R_safe
    return float(R), float(R / np.float32(2.5))


def _filter_or_exact(field, nx, ny, nz, dx, dy, dz, R, elements, ftol, what):
    """The GPU filter; when it signals failure (the empty vector), the
    reference falls back to a direct CHOLMOD solve of the same system
    (TopOpt.jl:195-205, 302-321).  Here the fallback solves that system again
    on the device to the exact answer's precision (relative residual 1e-12,
    residual checked every iteration, no iteration cap to speak of); if that
    fails too, the update stops with an error."""
    out = apply_gpu_filter(field, nx, ny, nz, dx, dy, dz, R, elements, ftol)
    if out[0].size:
        return out
    t0 = time.perf_counter()
    ex, _, it, res = apply_gpu_filter(field, nx, ny, nz, dx, dy, dz, R, elements, 1e-12, max_iter=100000,
                                      check_every=1)
    if ex.size == 0:
        raise RuntimeError(f"{what} filter failed on the device (direct-solve fallback failed too)")
    return ex, time.perf_counter() - t0, 1, 0.0


def update_density(density, l1_stress_norm_field, protected_elements_mask, E, l1_stress_allowable, iter_,
                   number_of_iterations, original_density, min_density, max_density, config, elements=None,
                   is_annealing=False):
    """TopOpt.jl:139-405 -> (max_change, R_length, threshold, filter_time, filter_iters, filter_res)"""
    opt = config["optimization_parameters"]
    geo = config["geometry"]
    sp = config["solver_parameters"]
    nx, ny, nz = int(geo["nElem_x_computed"]), int(geo["nElem_y_computed"]), int(geo["nElem_z_computed"])
    dx, dy, dz = float(geo["dx_computed"]), float(geo["dy_computed"]), float(geo["dz_computed"])
    max_dim = float(geo["max_domain_dim"])
    ftol = float(sp.get("filter_tolerance", 1.0e-5))
    avg_h = (np.float32(dx) + np.float32(dy) + np.float32(dz)) / np.float32(3)
    prot = np.asarray(protected_elements_mask, dtype=bool)

    ratio = float(opt.get("stress_regularization_ratio", 0.0))
    if ratio > 1e-6:  # :176-211
        fs, _, _, _ = _filter_or_exact(l1_stress_norm_field, nx, ny, nz, dx, dy, dz, ratio * float(avg_h),
                                       elements, ftol, "stress")
        l1_stress_norm_field[:] = fs

    R_len, R_eff = radius_schedule(opt, iter_, number_of_iterations, max_dim, avg_h)
    prop = np.where(prot, original_density,
                    np.clip(l1_stress_norm_field / np.float32(l1_stress_allowable) / np.float32(E),
                            min_density, max_density)).astype(density.dtype)
    filt, f_time, f_it, f_res = prop, 0.0, 0, 0.0
    if R_eff > 1e-4:  # :281-322
        filt, f_time, f_it, f_res = _filter_or_exact(prop, nx, ny, nz, dx, dy, dz, R_eff, elements, ftol,
                                                     "density")
    new = np.clip(filt, min_density, max_density).astype(density.dtype)
    diff = np.where(prot, 0.0, np.abs(new.astype(np.float64) - density))
    max_change = float(diff.max()) if diff.size else 0.0
    density[:] = np.where(prot, density, new)

    fin = np.float32(opt.get("final_density_threshold", 0.95))
    mcr = np.float32(opt.get("max_culling_ratio", 0.1))
    thr = fin if iter_ > number_of_iterations else fin * (np.float32(iter_) / np.float32(number_of_iterations))
    active = np.where(prot, original_density > min_density, density > min_density)
    cand = np.nonzero((~prot) & (density > min_density) & (density < thr))[0]
    max_cull = int(math.floor(int(active.sum()) * float(mcr)))
    if cand.size > max_cull:  # weakest first, ties by element order (stable sort)
        order = np.argsort(density[cand], kind="stable")
        density[cand[order[:max_cull]]] = min_density
    else:
        density[cand] = min_density
    density[prot] = original_density[prot]
    return max_change, R_len, float(thr), f_time, f_it, f_res


class SimpOptimizer:
    """Device-resident SIMP/OC loop on a structured grid (libhexat hx_to_*)."""

    def __init__(self, nx, ny, nz, f, fixed, *, dx=1.0, dy=1.0, dz=1.0, nu=0.3, E0=1.0, Emin=1e-9, penal=3.0,
                 volfrac=0.3, rmin=1.5, move=0.2, xmin=0.0, pcg_tol=1e-8, pcg_max_iter=200000, precision="fp64",
                 precond="jacobi", filter_tol=1e-10, filter_max_iter=5000, oc_rel_tol=1e-9, with_stress=True,
                 x0=None, ctx=None):
        self.ctx = ctx or get_context()
        self.ctx.set_grid(nx, ny, nz, dx, dy, dz, nu)
        self.ctx.set_fixed(fixed)
        prec = _lib.HX_FP32_IR if str(precision).lower() in ("fp32_mfma_ir", "fp32_ir", "fp32") else _lib.HX_FP64
        pc = _lib.HX_PC_LEARNED if str(precond).lower() == "learned" else _lib.HX_PC_JACOBI
        self.params = _lib.TOParams(E0=E0, Emin=Emin, nu=nu, penal=penal, volfrac=volfrac, move=move, xmin=xmin,
                                    rmin=rmin, pcg_tol=pcg_tol, pcg_max_iter=pcg_max_iter, precision=prec,
                                    precond=pc, filter_tol=filter_tol, filter_max_iter=filter_max_iter,
                                    oc_rel_tol=oc_rel_tol, with_stress=1 if with_stress else 0)
        x0 = np.full(nx * ny * nz, volfrac) if x0 is None else x0
        self.ctx.to_init(self.params, f, x0)
        self.history = []

    def step(self):
        t0 = time.perf_counter()
        s = self.ctx.to_step()
        rec = {"iter": s.iter, "compliance": s.compliance, "change": s.change, "volume": s.volume,
               "pcg_iterations": s.pcg_iterations, "pcg_rel_res": s.pcg_rel_res, "pcg_status": s.pcg_status,
               "filter_iterations": s.filter_iterations, "max_vm": s.max_vm, "device_s": s.seconds,
               "wall_s": time.perf_counter() - t0}
        self.history.append(rec)
        return rec

    def run(self, n_iter):
        return [self.step() for _ in range(n_iter)]

    def field(self, name):
        ids = {"x": _lib.HX_FIELD_X, "xphys": _lib.HX_FIELD_XPHYS, "u": _lib.HX_FIELD_U, "vm": _lib.HX_FIELD_VM,
               "dc": _lib.HX_FIELD_DC}
        return self.ctx.to_read(ids[name])
