"""CPU oracle for the HEXAT topology-optimisation hot path.

TEST INFRASTRUCTURE ONLY.  This module is the reference that the HIP path is
compared against; it may be imported only by ``tests/``, by
``__graft_entry__.smoke()`` and by the ``cpu_baseline`` leg of ``bench.py``.
The product package never imports it.

Numeric kernels live in ``hexat_oracle.c`` (FP64 restatement of the
reference's Julia algorithms, built by ``oracle/Makefile`` into
``liboracle.so``); this file wraps them with ctypes and restates the host-side
pieces in numpy:

* ``parity_update_density`` -- TopOpt.jl:139-405 (stress-ratio update with the
  radius schedule, Helmholtz filter, clamp, culling and protected restore).
* ``bc_indicator`` / ``external_forces`` -- Boundary.jl:56-291 location
  grammar on a structured grid.
* ``simp_*`` -- textbook SIMP/OC (Andreassen et al. 2011 "top88", Liu & Tovar
  2014 "top3d", Lazarov & Sigmund 2011 PDE filter); no reference code exists.
* ``ml_apply`` / ``ml_pcg`` -- additive multilevel diagonal scaling (the BPX
  form of Bramble, Pasciak & Xu 1990 with Jacobi level scalings) in separable
  numpy passes; no reference code exists.
* ``vc_apply`` / ``vc_pcg`` -- the same hierarchy as a symmetric V-cycle with
  l1-Jacobi smoothing (absolute row sums, the smoother of Baker, Falgout,
  Kolev & Yang 2011) on every level above the coarsest; no reference code
  exists.
* ``web_pack`` -- Postprocessing.jl:173-195 (kept elements, centroids, signed
  L1 of the viewer/checkpoint file) in FP32 numpy.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


class OxGrid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        P = C.POINTER
        d, i = C.c_double, C.c_int
        pd = P(C.c_double)
        pu8 = P(C.c_uint8)
        g = P(OxGrid)
        L.ox_material_matrix.argtypes = [d, d, pd]
        L.ox_ke_hex.argtypes = [d, d, d, d, d, pd]
        L.ox_scalar_ke_me.argtypes = [d, d, d, i, pd, pd]
        L.ox_matvec.argtypes = [g, pd, pd, pd, pd]
        L.ox_diag.argtypes = [g, pd, pd, pd]
        L.ox_pcg.argtypes = [g, pd, pd, pu8, pd, pd, d, d, i, pd, P(C.c_int)]
        L.ox_pcg.restype = C.c_int
        L.ox_helmholtz.argtypes = [g, d, i, pd, pd, d, i, i, pd, P(C.c_int)]
        L.ox_helmholtz.restype = C.c_int
        L.ox_helm_matvec.argtypes = [g, pd, pd, pd]
        L.ox_stress.argtypes = [g, d, d, pd, pd, pd, pd, pd, pd, pd]
        L.ox_element_energy.argtypes = [g, pd, pd, pd]
        L.ox_eig3.argtypes = [pd, pd, pd]
        L.ox_oc.argtypes = [i, pd, pd, pd, d, d, d, d, pd, pd]
        L.ox_oc.restype = C.c_int
        pf = P(C.c_float)
        L.ox_matvec_f32.argtypes = [g, pf, pf, pf, pf]
        L.ox_pcg_f32.argtypes = [g, pf, pf, pu8, pf, pf, C.c_float, d, i, pd, P(C.c_int)]
        L.ox_pcg_f32.restype = C.c_int
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Grid:
    """Structured grid of nx*ny*nz box elements (Mesh.jl:23-60)."""

    def __init__(self, nx, ny, nz, dx=1.0, dy=1.0, dz=1.0):
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.dx, self.dy, self.dz = float(dx), float(dy), float(dz)
        self.c = OxGrid(self.nx, self.ny, self.nz, self.dx, self.dy, self.dz)

    @property
    def ne(self):
        return self.nx * self.ny * self.nz

    @property
    def nn(self):
        return (self.nx + 1) * (self.ny + 1) * (self.nz + 1)

    @property
    def ndof(self):
        return 3 * self.nn

    def node_coords(self):
        NX, NY, NZ = self.nx + 1, self.ny + 1, self.nz + 1
        k, j, i = np.meshgrid(np.arange(NZ), np.arange(NY), np.arange(NX), indexing="ij")
        return np.stack([i.ravel() * self.dx, j.ravel() * self.dy, k.ravel() * self.dz], axis=1)

    def elements(self):
        """0-based Ne x 8 connectivity in reference node order n1..n8."""
        NX, NY = self.nx + 1, self.ny + 1
        k, j, i = np.meshgrid(np.arange(self.nz), np.arange(self.ny), np.arange(self.nx), indexing="ij")
        i, j, k = i.ravel(), j.ravel(), k.ravel()
        nid = lambda a, b, c: a + NX * (b + NY * c)
        return np.stack([nid(i, j, k), nid(i + 1, j, k), nid(i + 1, j + 1, k), nid(i, j + 1, k),
                         nid(i, j, k + 1), nid(i + 1, j, k + 1), nid(i + 1, j + 1, k + 1),
                         nid(i, j + 1, k + 1)], axis=1)


def material_matrix(E, nu):
    D = np.zeros(36)
    lib().ox_material_matrix(E, nu, _p(D))
    return D.reshape(6, 6)


def ke_hex(dx, dy, dz, nu, E=1.0):
    K = np.zeros(576)
    lib().ox_ke_hex(dx, dy, dz, nu, E, _p(K))
    return K.reshape(24, 24)


def scalar_ke_me(dx, dy, dz, quirk):
    Ks, Ms = np.zeros(64), np.zeros(64)
    lib().ox_scalar_ke_me(dx, dy, dz, int(quirk), _p(Ks), _p(Ms))
    return Ks.reshape(8, 8), Ms.reshape(8, 8)


def matvec(g: Grid, Ke, Ee, x):
    Ke, Ee, x = _f64(Ke), _f64(Ee), _f64(x)
    y = np.empty(g.ndof)
    lib().ox_matvec(C.byref(g.c), _p(Ke), _p(Ee), _p(x), _p(y))
    return y


def diag(g: Grid, Ke, Ee):
    Ke, Ee = _f64(Ke), _f64(Ee)
    y = np.empty(g.ndof)
    lib().ox_diag(C.byref(g.c), _p(Ke), _p(Ee), _p(y))
    return y


def pcg(g: Grid, Ke, Ee, fixed, b, x0=None, shift=0.0, tol=1e-8, max_iter=100000):
    Ke, Ee, b = _f64(Ke), _f64(Ee), _f64(b)
    fixed = np.ascontiguousarray(fixed, dtype=np.uint8)
    x = np.zeros(g.ndof) if x0 is None else _f64(x0).copy()
    rel = C.c_double(0.0)
    it = C.c_int(0)
    st = lib().ox_pcg(C.byref(g.c), _p(Ke), _p(Ee), fixed.ctypes.data_as(C.POINTER(C.c_uint8)),
                      _p(b), _p(x), shift, tol, int(max_iter), C.byref(rel), C.byref(it))
    return x, st, rel.value, it.value


def pcg_f32(g: Grid, Ke, Ee, fixed, b, x0=None, shift=0.0, tol=1e-6, max_iter=40000):
    """FP32 mirror of the reference's GPU Jacobi-PCG (GPUSolver.jl, Float32)."""
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    pf = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
    Ke, Ee, b = f32(Ke), f32(Ee), f32(b)
    fixed = np.ascontiguousarray(fixed, dtype=np.uint8)
    x = np.zeros(g.ndof, np.float32) if x0 is None else f32(x0).copy()
    rel = C.c_double(0.0)
    it = C.c_int(0)
    st = lib().ox_pcg_f32(C.byref(g.c), pf(Ke), pf(Ee), fixed.ctypes.data_as(C.POINTER(C.c_uint8)), pf(b), pf(x),
                          float(shift), float(tol), int(max_iter), C.byref(rel), C.byref(it))
    return x, st, rel.value, it.value


def helmholtz(g: Grid, R, quirk, rho, tol=1e-5, max_iter=200, check_every=10):
    rho = _f64(rho)
    out = np.empty(g.ne)
    rel = C.c_double(0.0)
    it = C.c_int(0)
    st = lib().ox_helmholtz(C.byref(g.c), R, int(quirk), _p(rho), _p(out), tol, int(max_iter),
                            int(check_every), C.byref(rel), C.byref(it))
    return out, st, rel.value, it.value


def helm_matvec(g: Grid, Ae, x):
    Ae, x = _f64(Ae), _f64(x)
    y = np.empty(g.nn)
    lib().ox_helm_matvec(C.byref(g.c), _p(Ae), _p(x), _p(y))
    return y


def stress(g: Grid, nu, E, rho, u):
    rho, u = _f64(rho), _f64(u)
    ne = g.ne
    pr, vm, sv, l1, dr = np.empty(3 * ne), np.empty(ne), np.empty(6 * ne), np.empty(ne), np.empty(3 * ne)
    lib().ox_stress(C.byref(g.c), nu, E, _p(rho), _p(u), _p(pr), _p(vm), _p(sv), _p(l1), _p(dr))
    return pr.reshape(ne, 3), vm, sv.reshape(ne, 6), l1, dr.reshape(ne, 3)


def element_energy(g: Grid, Ke, u):
    Ke, u = _f64(Ke), _f64(u)
    ce = np.empty(g.ne)
    lib().ox_element_energy(C.byref(g.c), _p(Ke), _p(u), _p(ce))
    return ce


def eig3(A):
    A = _f64(A).reshape(9)
    v, V = np.empty(3), np.empty(9)
    lib().ox_eig3(_p(A), _p(v), _p(V))
    return v, V.reshape(3, 3)


def oc(x, dc, dv, volfrac, move=0.2, xmin=0.0, rel_tol=1e-3):
    x, dc, dv = _f64(x), _f64(dc), _f64(dv)
    xn = np.empty_like(x)
    lm = C.c_double(0.0)
    steps = lib().ox_oc(x.size, _p(x), _p(dc), _p(dv), volfrac, move, xmin, rel_tol, _p(xn), C.byref(lm))
    return xn, lm.value, steps


# ---------------------------------------------------------------------------
# Boundary grammar (Boundary.jl:56-291) on a structured grid
# ---------------------------------------------------------------------------

def _resolve(spec, cmin, cmax):
    """Boundary.jl:117-130"""
    if isinstance(spec, (int, float)) and not isinstance(spec, bool):
        if 0.0 <= spec <= 1.0:
            return cmin + spec * (cmax - cmin)
        return float(spec)
    if isinstance(spec, str) and spec.endswith("%"):
        frac = min(max(float(spec[:-1]) / 100.0, 0.0), 1.0)
        return cmin + frac * (cmax - cmin)
    return cmin


def _mask(spec, coords, cmin, cmax):
    """Boundary.jl:79-115"""
    tol = 1e-4 * max(1.0, abs(cmax - cmin))
    if spec == ":":
        return np.ones(coords.shape[0], bool)
    return np.abs(coords - _resolve(spec, cmin, cmax)) <= tol


def nodes_by_location(loc, X):
    m = np.ones(X.shape[0], bool)
    for d in range(3):
        m &= _mask(loc[d], X[:, d], X[:, d].min(), X[:, d].max())
    return np.nonzero(m)[0]


def bc_indicator(X, bc_data):
    """Boundary.jl:135-160 -> (nNodes, 3) float mask"""
    ind = np.zeros((X.shape[0], 3))
    for bc in bc_data:
        nodes = nodes_by_location(bc["location"], X)
        for d in bc["DoFs"]:
            ind[nodes, d - 1] = 1.0
    return ind


def nearest_node(target, X):
    """Boundary.jl:195-224 (ties -> larger x, then y, then z)"""
    best, best_d = -1, math.inf
    for i in range(X.shape[0]):
        d2 = float(np.sum((X[i] - target) ** 2))
        if d2 < best_d - 1e-9:
            best, best_d = i, d2
        elif abs(d2 - best_d) <= 1e-9 and best >= 0:
            if tuple(X[i]) > tuple(X[best]):
                best = i
    return best


def external_forces(X, forces):
    """Boundary.jl:232-291"""
    F = np.zeros(3 * X.shape[0])
    for f in forces:
        nodes = nodes_by_location(f["location"], X)
        if nodes.size == 0 and all(c != ":" for c in f["location"]):
            tgt = np.array([_resolve(f["location"][d], X[:, d].min(), X[:, d].max()) for d in range(3)])
            nodes = np.array([nearest_node(tgt, X)])
        if nodes.size == 0:
            continue
        fv = np.zeros(3)
        fr = list(f["F"])[:3]
        fv[:len(fr)] = fr
        s = 1.0 / nodes.size
        for d in range(3):
            F[3 * nodes + d] += s * fv[d]
    return F


# ---------------------------------------------------------------------------
# Parity-mode density update (TopOpt.jl:139-405), numpy restatement
# ---------------------------------------------------------------------------

def radius_schedule(iter_, n_iter, R_init_perc, R_interm_perc, R_final_perc, interm_iter_perc,
                    max_dim, avg_h):
    """TopOpt.jl:216-255 -> (R_length, R_effective)"""
    Ri, Rm, Rf = (R_init_perc / 100 * max_dim, R_interm_perc / 100 * max_dim, R_final_perc / 100 * max_dim)
    it_m = max(1, int(round(interm_iter_perc / 100 * n_iter)))
    c = min(iter_, n_iter)
    if c <= it_m:
        t = (c - 1) / (it_m - 1) if it_m > 1 else 0.0
        R = Ri * (1 - t) + Rm * t
    else:
        t = (c - it_m) / (n_iter - it_m) if n_iter > it_m else 0.0
        R = Rm * (1 - t) + Rf * t
    R = max(R, 1.2 * avg_h)
    return R, R / 2.5


def parity_update_density(g: Grid, density, l1, protected, E, l1_allowable, iter_, n_iter,
                          original, min_density, max_density, opt, max_dim, filter_tol=1e-5):
    """TopOpt.jl:139-405.  Mutates density and l1 in place; returns
    (max_change, R_length, threshold)."""
    avg_h = (g.dx + g.dy + g.dz) / 3.0
    ratio = float(opt.get("stress_regularization_ratio", 0.0))
    if ratio > 1e-6:  # :176-211
        fs, _, _, _ = helmholtz(g, ratio * avg_h, 1, l1, tol=filter_tol)
        l1[:] = fs
    R_len, R_eff = radius_schedule(iter_, n_iter, float(opt.get("filter_R_init_perc", 0.0)),
                                   float(opt.get("filter_R_interm_perc", 0.0)),
                                   float(opt.get("filter_R_final_perc", 0.0)),
                                   float(opt.get("filter_R_interm_iter_perc", 50.0)), max_dim, avg_h)
    prop = np.where(protected, original, np.clip(l1 / l1_allowable / E, min_density, max_density))
    filt = prop
    if R_eff > 1e-4:  # :281-322
        filt, _, _, _ = helmholtz(g, R_eff, 1, prop, tol=filter_tol)
    new = np.clip(filt, min_density, max_density)
    diff = np.where(protected, 0.0, np.abs(new - density))
    max_change = float(diff.max()) if diff.size else 0.0
    density[:] = np.where(protected, density, new)
    fin = float(opt.get("final_density_threshold", 0.95))
    mcr = float(opt.get("max_culling_ratio", 0.1))
    thr = fin if iter_ > n_iter else fin * (iter_ / n_iter)  # :354-359
    act = np.where(protected, original > min_density, density > min_density)
    cand = np.nonzero((~protected) & (density > min_density) & (density < thr))[0]
    max_cull = int(math.floor(int(act.sum()) * mcr))
    if cand.size > max_cull:  # :383-395 (stable sort by density)
        order = np.argsort(density[cand], kind="stable")
        density[cand[order[:max_cull]]] = min_density
    else:
        density[cand] = min_density
    density[protected] = original[protected]
    return max_change, R_len, thr


def used_dofs(g: Grid, Ee):
    """DOF mask of the nodes touched by an element with non-zero modulus
    (MeshPruner.jl:39-48: the nodes an active element uses)."""
    on = (np.asarray(Ee) != 0).reshape(g.nz, g.ny, g.nx)
    used = np.zeros((g.nz + 1, g.ny + 1, g.nx + 1), bool)
    for dk in (0, 1):
        for dj in (0, 1):
            for di in (0, 1):
                used[dk:dk + g.nz, dj:dj + g.ny, di:di + g.nx] |= on
    return np.repeat(used.ravel(), 3)


def parity_tolerance(base_tol, iter_, n_iter, gpu=True):
    """Solver.jl:99-120: 10^((1-t)(-5) + t log10(tol)) in Float32, GPU floor 5e-7."""
    t = min(max(np.float32(iter_) / np.float32(n_iter), 0.0), 1.0)
    tol = float(np.float32(10.0 ** ((1.0 - t) * -5.0 + t * math.log10(float(np.float32(base_tol))))))
    return max(tol, float(np.float32(5.0e-7))) if gpu else tol


def parity_solve(g: Grid, nu, E, density, fixed, F, threshold, prune_voids=True, tol=1e-6, max_iter=40000,
                 shift_factor=1e-6, u_prev=None, fp32=False):
    """One attempt of Solver.solve_system (Solver.jl:38-164) on the structured
    grid: active elements rho >= threshold with modulus E*rho
    (GPUSolver.jl:128-147); with prune_voids and fewer than 99 % of the elements
    above the FP32 product 1.01*threshold, only those stay and the nodes they
    do not touch leave the system (MeshPruner.jl:18-138; here: those DOFs are
    held at 0); shift = shift_factor * max free diagonal (GPUSolver.jl:315-316,
    391); warm start from u_prev (Solver.jl:71-83).  Returns
    (u, status, rel, iterations, info)."""
    density = np.asarray(density, dtype=np.float64)
    thr_keep = float(np.float32(threshold) * np.float32(1.01))
    active = density >= threshold
    keep = density > thr_keep
    pruned = bool(prune_voids and keep.sum() < 0.99 * density.size)
    sel = keep if pruned else active
    if not sel.any():
        raise RuntimeError("No active elements found")
    Ee = np.where(sel, E * density, 0.0)
    Ke = ke_hex(g.dx, g.dy, g.dz, nu)
    fixed = np.asarray(fixed, dtype=np.uint8)
    d = diag(g, Ke, Ee)
    free = fixed == 0
    shift = shift_factor * float(np.abs(d[free]).max()) if free.any() else 0.0
    fixed_eff = (fixed.astype(bool) | ~used_dofs(g, Ee)).astype(np.uint8)
    x0 = None if u_prev is None else np.where(fixed_eff == 0, u_prev, 0.0)
    if fp32:  # the reference's own arithmetic (FP32 throughout, recursive residual)
        u, st, rel, it = pcg_f32(g, Ke, Ee, fixed_eff, F, x0=x0, shift=shift, tol=tol, max_iter=max_iter)
        u = u.astype(np.float64)
    else:
        u, st, rel, it = pcg(g, Ke, Ee, fixed_eff, F, x0=x0, shift=shift, tol=tol, max_iter=max_iter)
    return u, st, rel, it, {"pruned": pruned, "n_active": int(active.sum()), "n_kept": int(sel.sum()),
                            "shift": shift, "Ee": Ee}


def parity_loop(g: Grid, nu, E, fixed, F, n_iter, opt, n_nominal, max_dim, l1_allowable=1.0, base_tol=1e-6,
                max_iter=40000, shift_factor=1e-6, filter_tol=1e-5, fixed_tol=None, fp32=False):
    """The stress-ratio TO loop of Main.jl:331-474 without refinement, export,
    gravity or thermal loads, from rho = 1 (no protected shapes): solve
    (parity_solve), compliance F'u, centroid stress (Stress.jl:57-100), update
    (parity_update_density with n_nominal + 10, Main.jl:446).  fixed_tol
    replaces the adaptive schedule.  Returns per-iteration records and the
    density after each update."""
    min_density = float(opt.get("min_density", 1e-3))
    max_density = float(opt.get("density_clamp_max", 1.0))
    density = np.ones(g.ne)
    original = np.ones(g.ne)
    protected = np.zeros(g.ne, bool)
    u = np.zeros(g.ndof)
    hist, dens = [], []
    for it in range(1, n_iter + 1):
        tol = fixed_tol if fixed_tol is not None else parity_tolerance(base_tol, it, n_nominal)
        u, st, rel, its, info = parity_solve(g, nu, E, density, fixed, F, min_density, True, tol, max_iter,
                                             shift_factor, u_prev=u, fp32=fp32)
        comp = float(F @ u)
        _, _, _, l1, _ = stress(g, nu, E, density, u)
        act = density > 0.1
        avg_l1 = float(l1[act].mean()) if act.any() else 0.0
        vf = float(density.sum() / density.size)
        n_active = int((density > min_density).sum())
        mc, R, thr = parity_update_density(g, density, l1, protected, E, l1_allowable, it, n_nominal + 10, original,
                                           min_density, max_density, opt, max_dim, filter_tol)
        hist.append({"iter": it, "compliance": comp, "avg_l1": avg_l1, "vol_frac": vf, "active": n_active,
                     "delta_rho": mc, "radius": R, "cutoff": thr, "pcg_status": st, "pcg_rel": rel,
                     "pcg_iterations": its, "tol": tol, "pruned": info["pruned"]})
        dens.append(density.copy())
    return hist, dens


# ---------------------------------------------------------------------------
# Textbook SIMP pieces
# ---------------------------------------------------------------------------

def simp_modulus(xphys, penal, E0, Emin):
    return Emin + xphys ** penal * (E0 - Emin)


def simp_dc(xphys, ce, penal, E0, Emin):
    return -penal * xphys ** (penal - 1) * (E0 - Emin) * ce


def simp_to_loop(g: Grid, f, fixed, n_iter, nu=0.3, E0=1.0, Emin=1e-9, penal=3.0, volfrac=0.3, rmin=1.5,
                 move=0.2, xmin=0.0, pcg_tol=1e-8, filter_tol=1e-10, oc_rel_tol=1e-9, x0=None,
                 pcg_max_iter=200000, filter_max_iter=5000):
    """Textbook SIMP/OC loop (top88/top3d order, PDE filter of Lazarov & Sigmund
    with R = rmin/(2 sqrt 3)) on the oracle kernels.  Returns a list of per-
    iteration dicts and the final (x, xphys, u)."""
    Ke = ke_hex(g.dx, g.dy, g.dz, nu)
    R = rmin / (2.0 * math.sqrt(3.0))
    x = np.full(g.ne, volfrac) if x0 is None else np.array(x0, dtype=np.float64)
    filt = lambda v: helmholtz(g, R, 0, v, tol=filter_tol, max_iter=filter_max_iter, check_every=1)[0]
    xphys = filt(x)
    dv = filt(np.ones(g.ne))
    u = np.zeros(g.ndof)
    hist = []
    for it in range(1, n_iter + 1):
        E = simp_modulus(xphys, penal, E0, Emin)
        u, st, rel, its = pcg(g, Ke, E, fixed, f, x0=u, tol=pcg_tol, max_iter=pcg_max_iter)
        c = float(f @ u)
        ce = element_energy(g, Ke, u)
        dcf = filt(simp_dc(xphys, ce, penal, E0, Emin))
        xnew, lmid, steps = oc(x, dcf, dv, volfrac, move, xmin, oc_rel_tol)
        change = float(np.abs(xnew - x).max())
        xphys = filt(xnew)
        x = xnew
        hist.append({"iter": it, "compliance": c, "change": change, "pcg_iterations": its, "pcg_rel": rel,
                     "volume": float(xphys.mean()), "lmid": lmid})
    return hist, x, xphys, u


# ---------------------------------------------------------------------------------------------
# multilevel diagonal scaling:  z = D0^-1 r + sum_l P_l D_l^-1 P_l^T r
# (no reference counterpart; the library's HX_PC_MULTILEVEL)

def ml_levels(g: Grid, max_levels=12):
    """2:1 coarsenings while every element count is even."""
    L, n = 0, [g.nx, g.ny, g.nz]
    while L < max_levels and all(v % 2 == 0 for v in n):
        n = [v // 2 for v in n]
        L += 1
    return L


def _restrict_axis(a, axis):
    """full weighting along one node axis: c[I] = f[2I] + (f[2I-1] + f[2I+1]) / 2"""
    f = np.moveaxis(a, axis, 0)
    c = f[0::2].copy()
    c[:-1] += 0.5 * f[1::2]
    c[1:] += 0.5 * f[1::2]
    return np.moveaxis(c, 0, axis)


def _prolong_axis(a, axis):
    """linear interpolation along one node axis: f[2I] = c[I], f[2I+1] = (c[I] + c[I+1]) / 2"""
    c = np.moveaxis(a, axis, 0)
    f = np.empty((2 * c.shape[0] - 1,) + c.shape[1:])
    f[0::2] = c
    f[1::2] = 0.5 * (c[:-1] + c[1:])
    return np.moveaxis(f, 0, axis)


def _coarsen_E(E, nx, ny, nz):
    e = E.reshape(nz, ny, nx)
    return 0.125 * sum(e[c::2, b::2, a::2] for c in (0, 1) for b in (0, 1) for a in (0, 1)).ravel()


def _node_sum(E, nx, ny, nz):
    e = E.reshape(nz, ny, nx)
    s = np.zeros((nz + 1, ny + 1, nx + 1))
    for c in (0, 1):
        for b in (0, 1):
            for a in (0, 1):
                s[c:c + nz, b:b + ny, a:a + nx] += e
    return s


ML_EXACT_MAX = 192  # DOFs of a coarsest level solved exactly


def ml_hierarchy(g: Grid, Ke, Ee, fixed, max_levels=12, exact=True):
    """Per level (l = 0..L): (element counts, inverse diagonal (NZ, NY, NX, 3), A_L^-1 or None).
    With `exact`, the first level with at most ML_EXACT_MAX DOFs ends the hierarchy and is solved
    exactly: its stiffness assembled from the averaged moduli and h Ke, DOFs fixed where the fine DOF
    at the same place is (or without stiffness) as identity rows and columns, inverted, those rows and
    columns of the inverse zeroed."""
    Ke = _f64(Ke).reshape(24, 24)
    kd = np.diag(Ke)[:3]
    fx = np.asarray(fixed, dtype=bool).reshape(g.nz + 1, g.ny + 1, g.nx + 1, 3)
    L = ml_levels(g, max_levels)
    if exact:
        n = (g.nx, g.ny, g.nz)
        for l in range(1, L + 1):
            n = tuple(v // 2 for v in n)
            if 3 * (n[0] + 1) * (n[1] + 1) * (n[2] + 1) <= ML_EXACT_MAX:
                L = l
                break
        else:
            exact = False
    levels = []
    E, n = _f64(Ee), (g.nx, g.ny, g.nz)
    for l in range(L + 1):
        if l > 0:
            E = _coarsen_E(E, *n)
            n = tuple(v // 2 for v in n)
        h = 2 ** l
        S = _node_sum(E, *n)[..., None]
        d = h * kd[None, None, None, :] * S
        free = ~fx[::h, ::h, ::h] & (S > 0)
        ainv = None
        if exact and l == L and l > 0:
            gc = Grid(*n)
            N = gc.ndof
            A = np.zeros((N, N))
            dof = (3 * gc.elements()[:, :, None] + np.arange(3)[None, None, :]).reshape(-1, 24)
            for e in range(gc.ne):
                A[np.ix_(dof[e], dof[e])] += E[e] * h * Ke
            keep = (~fx[::h, ::h, ::h]).reshape(-1) & (np.diag(A) > 0)
            A[~keep, :] = 0.0
            A[:, ~keep] = 0.0
            A[~keep, ~keep] = 1.0
            ainv = np.linalg.inv(A)
            ainv[~keep, :] = 0.0
            ainv[:, ~keep] = 0.0
        levels.append((n, np.where(free, 1.0 / np.where(free, d, 1.0), 0.0), ainv))
    return levels


def ml_apply(g: Grid, Ke, Ee, fixed, r, max_levels=12, levels=None, exact=True):
    """z = M^-1 r of the multilevel preconditioner (masked where D0^-1 = 0)."""
    levels = levels or ml_hierarchy(g, Ke, Ee, fixed, max_levels, exact)
    L = len(levels) - 1
    rs = [_f64(r).reshape(g.nz + 1, g.ny + 1, g.nx + 1, 3)]
    for l in range(1, L + 1):
        a = rs[-1]
        for ax in (0, 1, 2):
            a = _restrict_axis(a, ax)
        rs.append(a)
    if levels[L][2] is not None:
        y = (levels[L][2] @ rs[L].ravel()).reshape(rs[L].shape)
    else:
        y = levels[L][1] * rs[L]
    for l in range(L - 1, -1, -1):
        p = y
        for ax in (0, 1, 2):
            p = _prolong_axis(p, ax)
        m = levels[l][1]
        y = m * rs[l] + (p if l > 0 else np.where(m != 0.0, p, 0.0))
    if L == 0:
        y = levels[0][1] * rs[0]
    return y.ravel()


def ml_pcg(g: Grid, Ke, Ee, fixed, b, tol=1e-8, max_iter=100000, max_levels=12, exact=True):
    """Plain PCG with the multilevel preconditioner (free DOFs; fixed stay 0)."""
    Ke, Ee, b = _f64(Ke), _f64(Ee), _f64(b)
    levels = ml_hierarchy(g, Ke, Ee, fixed, max_levels, exact)
    mask = levels[0][1].ravel() != 0.0
    b = np.where(mask, b, 0.0)
    x = np.zeros(g.ndof)
    r = b.copy()
    z = ml_apply(g, Ke, Ee, fixed, r, levels=levels)
    p = z.copy()
    rz = r @ z
    nb = np.linalg.norm(b)
    for it in range(1, max_iter + 1):
        q = np.where(mask, matvec(g, Ke, Ee, p), 0.0)
        a = rz / (p @ q)
        x += a * p
        r -= a * q
        rel = np.linalg.norm(r) / nb
        if rel < tol:
            return x, rel, it
        z = ml_apply(g, Ke, Ee, fixed, r, levels=levels)
        rzn = r @ z
        p = z + (rzn / rz) * p
        rz = rzn
    return x, rel, max_iter


# ---------------------------------------------------------------------------------------------
# manufactured displacement of the learned preconditioner's training (no reference counterpart)

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def mix64(z):
    """splitmix64's output function of state z + golden gamma (Steele, Lea & Flood 2014), on uint64 arrays"""
    z = np.asarray(z, dtype=np.uint64) + _GOLDEN
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def lattice_normal(seed, q):
    """standard normal value of lattice key q: Box-Muller on two hashed uniforms"""
    h1 = mix64(np.uint64(seed) ^ mix64(q))
    h2 = mix64(h1 ^ np.uint64(0xD1B54A32D192ED03))
    u1 = ((h1 >> np.uint64(11)).astype(np.float64) + 0.5) * 2.0 ** -53
    u2 = (h2 >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def smooth_field(g: Grid, seed, levels):
    """sum over l = 1..levels of 2^(l-levels) times the trilinear interpolation of lattice_normal on
    the lattice of spacing 2^l (key 3*((l << 56) ^ lattice index) + component), node-major (ndof,)"""
    NX, NY, NZ = g.nx + 1, g.ny + 1, g.nz + 1
    k, j, i = np.meshgrid(np.arange(NZ), np.arange(NY), np.arange(NX), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    out = np.zeros((i.size, 3))
    with np.errstate(over="ignore"):
        for l in range(1, levels + 1):
            h = 1 << l
            LX, LY = (NX + h - 1) // h + 1, (NY + h - 1) // h + 1
            I, J, K = i >> l, j >> l, k >> l
            fx, fy, fz = (i & (h - 1)) / h, (j & (h - 1)) / h, (k & (h - 1)) / h
            w = 2.0 ** (l - levels)
            for c8 in range(8):
                ox, oy, oz = c8 & 1, (c8 >> 1) & 1, c8 >> 2
                wt = w * (fx if ox else 1 - fx) * (fy if oy else 1 - fy) * (fz if oz else 1 - fz)
                pt = (np.uint64(l) << np.uint64(56)) ^ (((K + oz) * LY + (J + oy)) * LX + (I + ox)).astype(np.uint64)
                for c in range(3):
                    out[:, c] += wt * lattice_normal(seed, np.uint64(3) * pt + np.uint64(c))
    return out.ravel()


# ---------------------------------------------------------------------------------------------
# V-cycle form (the library's HX_PC_VCYCLE): V-cycle on levels 0..L-1, the coarsest level L exactly
# (or by its diagonal when the hierarchy has no exact level)

def vc_levels(levels):
    """Smoothed levels: every level above the coarsest, L (0 without coarse levels)."""
    return len(levels) - 1


def _ke_binary(Ke):
    """24x24 Ke with the element's nodes in binary order L = bx + 2 by + 4 bz."""
    e = Grid(1, 1, 1).elements()[0]
    ref = [int(np.flatnonzero(e == (L & 1) + 2 * ((L >> 1) & 1) + 4 * (L >> 2))[0]) for L in range(8)]
    idx = np.array([3 * ref[L] + c for L in range(8) for c in range(3)])
    return _f64(Ke).reshape(24, 24)[np.ix_(idx, idx)]


def vc_smoother(n, E, h, Ke, fixed_l, omega):
    """omega / sum_j |A_ij| over the solved DOFs j of the level stiffness (moduli E, element
    stiffness h Ke), 0 where DOF i is not solved for; as (NZ, NY, NX, 3).  The 3x3 coupling of
    node (i,j,k) with each of its 27 neighbours is summed over their shared elements first."""
    nx, ny, nz = n
    Kb = _ke_binary(Ke)
    Ep = np.zeros((nz + 2, ny + 2, nx + 2))
    Ep[1:-1, 1:-1, 1:-1] = h * _f64(E).reshape(nz, ny, nx)
    NZ, NY, NX = nz + 1, ny + 1, nx + 1
    B = np.zeros((3, 3, 3, NZ, NY, NX, 3, 3))  # [dk+1, dj+1, di+1, node, a, b]
    for ez in (0, 1):
        for ey in (0, 1):
            for ex in (0, 1):
                # element (i-1+ex, j-1+ey, k-1+ez) of node (i,j,k): padded index (k+ez, j+ey, i+ex)
                Ee = Ep[ez:ez + NZ, ey:ey + NY, ex:ex + NX]
                La = (1 - ex) + 2 * (1 - ey) + 4 * (1 - ez)
                for bz in (0, 1):
                    for by in (0, 1):
                        for bx in (0, 1):
                            Lb = bx + 2 * by + 4 * bz
                            di, dj, dk = bx - (1 - ex), by - (1 - ey), bz - (1 - ez)
                            blk = Kb[3 * La:3 * La + 3, 3 * Lb:3 * Lb + 3]
                            B[dk + 1, dj + 1, di + 1] += Ee[..., None, None] * blk[None, None, None]
    fx = np.asarray(fixed_l, dtype=bool)
    fxp = np.ones((NZ + 2, NY + 2, NX + 2, 3), dtype=bool)
    fxp[1:-1, 1:-1, 1:-1] = fx
    l1 = np.zeros((NZ, NY, NX, 3))
    for dk in (-1, 0, 1):
        for dj in (-1, 0, 1):
            for di in (-1, 0, 1):
                solved = ~fxp[1 + dk:1 + dk + NZ, 1 + dj:1 + dj + NY, 1 + di:1 + di + NX]  # (..., b)
                l1 += (np.abs(B[dk + 1, dj + 1, di + 1]) * solved[..., None, :]).sum(axis=-1)
    diag = np.stack([B[1, 1, 1][..., a, a] for a in range(3)], axis=-1)
    keep = ~fx & (diag > 0)
    return np.where(keep, omega / np.where(keep, l1, 1.0), 0.0)


def _restrict3(a):
    for ax in (0, 1, 2):
        a = _restrict_axis(a, ax)
    return a


def _prolong3(a):
    for ax in (0, 1, 2):
        a = _prolong_axis(a, ax)
    return a


def _additive_from(levels, c, rc):
    """y_c of the additive form started at level c: D_l^-1 r_l + P y_{l+1} (unmasked P), the
    coarsest level by its inverse (exact) or its diagonal."""
    L = len(levels) - 1
    rs = [rc]
    for l in range(c + 1, L + 1):
        rs.append(_restrict3(rs[-1]))
    ainv = levels[L][2]
    y = (ainv @ rs[-1].ravel()).reshape(rs[-1].shape) if ainv is not None else levels[L][1] * rs[-1]
    for l in range(L - 1, c - 1, -1):
        y = levels[l][1] * rs[l - c] + _prolong3(y)
    return y


VC_W_NODES = 512  # levels of at most this many nodes: the library's one-workgroup pass


def vc_w_levels(levels):
    """The levels that visit the next coarser level twice: L-1 when the coarsest level L is solved
    exactly, L-1 >= 1 and level L-1 has at most VC_W_NODES nodes (the library's one-workgroup pass holds
    it); with it L-2 when L-2 >= 1 and it has more (the library's last multi-block level)."""
    L = len(levels) - 1
    if L < 2 or levels[L][2] is None:
        return set()
    nn = [(n[0] + 1) * (n[1] + 1) * (n[2] + 1) for n, *_ in levels]
    if nn[L - 1] > VC_W_NODES:
        return set()
    return {L - 1} | ({L - 2} if L - 2 >= 1 and nn[L - 2] > VC_W_NODES else set())


def vc_apply(g: Grid, Ke, Ee, fixed, r, omega=1.95, max_levels=12, exact=True, state=None):
    """z = M^-1 r of the V-cycle form: on levels l < L (vc_levels), with S_l = vc_smoother and A_l
    the level operator (moduli 2^l E_l averaged over children, the fine Ke),
      x = S r;  res = mask(r - A x);  e = V_{l+1}(R res);  x += mask(P e);  z = x + S (r - A x)
    and the coarsest level L by its inverse (exact) or its diagonal (_additive_from at L).  The levels
    vc_w_levels names repeat the correction and the smoothing after it once (smoothing, correction,
    smoothing, correction, smoothing: still symmetric).  `state` caches the levels between calls."""
    Ke = _f64(Ke)
    if state is None or "levels" not in state:
        levels = ml_hierarchy(g, Ke, Ee, fixed, max_levels, exact)
        c = vc_levels(levels)
        fx = np.asarray(fixed, dtype=bool).reshape(g.nz + 1, g.ny + 1, g.nx + 1, 3)
        E, n, Es, Ss = _f64(Ee), (g.nx, g.ny, g.nz), [], []
        for l in range(c):
            if l > 0:
                E = _coarsen_E(E, *n)
                n = tuple(v // 2 for v in n)
            h = 2 ** l
            Es.append(h * E)
            Ss.append(vc_smoother(n, E, h, Ke, fx[::h, ::h, ::h], omega))
        st = {"levels": levels, "c": c, "E": Es, "S": Ss}
        if state is not None:
            state.update(st)
    else:
        st = state
    levels, c = st["levels"], st["c"]
    r0 = _f64(r).reshape(g.nz + 1, g.ny + 1, g.nx + 1, 3)
    if c == 0:
        return (levels[0][1] * r0).ravel()

    def A(l, x):
        n = levels[l][0]
        return matvec(Grid(*n), Ke, st["E"][l], x.ravel()).reshape(x.shape)

    w = vc_w_levels(levels)

    def V(l, rl):
        if l == c:
            return _additive_from(levels, c, rl)
        S = st["S"][l]
        x = S * rl
        for _ in range(2 if l in w else 1):
            res = np.where(S != 0.0, rl - A(l, x), 0.0)
            e = V(l + 1, _restrict3(res))
            x = x + np.where(S != 0.0, _prolong3(e), 0.0)
            x = x + S * (rl - A(l, x))
        return x

    return V(0, r0).ravel()


def vc_pcg(g: Grid, Ke, Ee, fixed, b, tol=1e-8, max_iter=100000, omega=1.95, max_levels=12, exact=True):
    """Plain PCG with the V-cycle preconditioner (free DOFs; fixed stay 0)."""
    Ke, Ee, b = _f64(Ke), _f64(Ee), _f64(b)
    state = {}
    z = vc_apply(g, Ke, Ee, fixed, b, omega, max_levels, exact, state=state)
    mask = state["levels"][0][1].ravel() != 0.0
    b = np.where(mask, b, 0.0)
    x = np.zeros(g.ndof)
    r = b.copy()
    z = vc_apply(g, Ke, Ee, fixed, r, state=state)
    p = z.copy()
    rz = r @ z
    nb = np.linalg.norm(b)
    rel = 1.0
    for it in range(1, max_iter + 1):
        q = np.where(mask, matvec(g, Ke, Ee, p), 0.0)
        a = rz / (p @ q)
        x += a * p
        r -= a * q
        rel = np.linalg.norm(r) / nb
        if rel < tol:
            return x, rel, it
        z = vc_apply(g, Ke, Ee, fixed, r, state=state)
        rzn = r @ z
        p = z + (rzn / rz) * p
        rz = rzn
    return x, rel, max_iter


# ----------------------------------------------------------------------------
# viewer/checkpoint payload
# ----------------------------------------------------------------------------
def web_pack(nx, ny, nz, dx, dy, dz, density, l1, principal, threshold):
    """Packing loop of Postprocessing.export_binary_for_web
    (Postprocessing.jl:173-195) on the mesh of Mesh.generate_mesh (Mesh.jl:23-60),
    in FP32 as the reference computes it:

    * kept = findall(rho >= threshold), in element order (:173);
    * centroid = MeshUtilities.element_centroid (MeshUtilities.jl:18-32): the
      FP32 node coordinates Float32(i-1)*dx of the 8 nodes summed in node order
      from 0f0, times 0.125f0;
    * signed L1 = l1 * sign of the principal stress of largest magnitude, the
      first one on ties (strict >, :187-195); sign(0) = 0.

    principal is (Ne, 3).  Returns (centroids (count, 3), density, signed_l1)
    as float32 arrays."""
    f32 = np.float32
    NX, NY = nx + 1, ny + 1
    rho = np.asarray(density, dtype=f32).ravel()
    keep = np.nonzero(rho >= f32(threshold))[0]
    # node coordinates as generate_mesh stores them
    xs = np.arange(nx + 1, dtype=f32) * f32(dx)
    ys = np.arange(ny + 1, dtype=f32) * f32(dy)
    zs = np.arange(nz + 1, dtype=f32) * f32(dz)
    i, j, k = keep % nx, (keep // nx) % ny, keep // (nx * ny)
    order = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    cx = np.zeros(keep.size, f32)
    cy = np.zeros(keep.size, f32)
    cz = np.zeros(keep.size, f32)
    for a, b, c in order:  # element row n1..n8 of Mesh.jl:45-53
        node = (i + a) + NX * ((j + b) + NY * (k + c))
        cx = cx + xs[node % NX]
        cy = cy + ys[(node // NX) % NY]
        cz = cz + zs[node // (NX * NY)]
    inv8 = f32(0.125)
    cent = np.stack([cx * inv8, cy * inv8, cz * inv8], axis=1).astype(f32)
    pr = np.asarray(principal, dtype=f32).reshape(-1, 3)[keep]
    s1, s2, s3 = pr[:, 0], pr[:, 1], pr[:, 2]
    amax, sg = np.abs(s1), np.sign(s1)
    t = np.abs(s2) > amax
    amax, sg = np.where(t, np.abs(s2), amax), np.where(t, np.sign(s2), sg)
    t = np.abs(s3) > amax
    sg = np.where(t, np.sign(s3), sg).astype(f32)
    sl1 = (np.asarray(l1, dtype=f32).ravel()[keep] * sg).astype(f32)
    return cent, rho[keep].copy(), sl1


# ----------------------------------------------------------------------------
# legacy VTK file (ExportVTK.export_solution, ExportVTK.jl:20-172), restated on
# explicit node and element arrays
# ----------------------------------------------------------------------------

def _sanitize32(a):
    """ExportVTK.jl:35-42 on Float32 data"""
    a = np.asarray(a, dtype=np.float32).copy()
    a[~np.isfinite(a)] = np.float32(0.0)
    if a.size and np.abs(a).max() > np.float32(1.0e10):
        a = np.clip(a, np.float32(-1.0e10), np.float32(1.0e10))
    return a


def vtk_bytes(nodes, elements, U, F, bc_indicator, vonmises, l1, density=None, threshold=0.01):
    """The bytes ExportVTK.export_solution writes.  nodes: (nNodes, 3) Float32,
    elements: (nElem, 8) 1-based node ids, U and F: 3 nNodes, bc_indicator:
    (nNodes, 3); every value is taken as Float32 as the reference holds it."""
    U = _sanitize32(U)
    Fs = _sanitize32(F)
    bc = _sanitize32(bc_indicator)
    nodes = _sanitize32(nodes)
    l1 = _sanitize32(l1)
    nN, nE = nodes.shape[0], elements.shape[0]
    if density is not None:
        valid = [e for e in range(nE) if density[e] >= np.float32(threshold)]
    else:
        valid = list(range(nE))
    n = len(valid)
    disp = U.reshape(nN, 3)
    forces = Fs.reshape(nN, 3)
    bcs = bc.reshape(nN, 3)
    mag = np.sqrt((disp[:, 0] * disp[:, 0] + disp[:, 1] * disp[:, 1]) + disp[:, 2] * disp[:, 2]).astype(np.float32)
    cells = np.empty((n, 9), dtype=np.int32)
    for t, e in enumerate(valid):
        cells[t, 0] = 8
        cells[t, 1:] = np.asarray(elements[e], dtype=np.int64) - 1
    be32 = lambda a: np.asarray(a, dtype=np.float32).astype(">f4").tobytes()
    bei = lambda a: np.asarray(a, dtype=np.int32).astype(">i4").tobytes()
    out = bytearray()
    out += b"# vtk DataFile Version 3.0\nHEXA FEM Solution (Undeformed)\nBINARY\nDATASET UNSTRUCTURED_GRID\n"
    out += f"POINTS {nN} float\n".encode() + be32(nodes.reshape(-1))
    out += f"\nCELLS {n} {n * 9}\n".encode() + bei(cells.reshape(-1))
    out += f"\nCELL_TYPES {n}\n".encode() + bei(np.full(n, 12))
    out += f"\nPOINT_DATA {nN}\n".encode() + b"VECTORS Displacement float\n" + be32(disp.reshape(-1))
    out += b"\nVECTORS External_Forces float\n" + be32(forces.reshape(-1))
    out += b"\nVECTORS Boundary_Conditions float\n" + be32(bcs.reshape(-1))
    out += b"\nSCALARS Displacement_Magnitude float 1\nLOOKUP_TABLE default\n" + be32(mag)
    out += f"\nCELL_DATA {n}\n".encode() + b"SCALARS l1_stress_norm float 1\nLOOKUP_TABLE default\n" + be32(l1[valid])
    out += b"\nSCALARS Von_Mises float 1\nLOOKUP_TABLE default\n" + be32(np.asarray(vonmises, np.float32)[valid])
    if density is not None:
        out += b"\nSCALARS Element_Density float 1\nLOOKUP_TABLE default\n" + be32(
            np.asarray(density, np.float32)[valid])
    return bytes(out)


# ----------------------------------------------------------------------------
# STL sampling grid (Postprocessing.jl:19-92): nodal mean density, trilinear
# refinement, one layer of zeros; FP32 in the reference's operation order
# ----------------------------------------------------------------------------

def stl_field(density, nx, ny, nz, sub):
    rho = np.asarray(density, dtype=np.float32).reshape(nz, ny, nx)
    NX, NY, NZ = nx + 1, ny + 1, nz + 1
    sums = np.zeros((NZ, NY, NX), np.float32)
    cnt = np.zeros((NZ, NY, NX), np.int64)
    # the reference loops over elements in order (x fastest), adding each
    # density to its 8 nodes: a node receives its elements (i-1..i, j-1..j,
    # k-1..k) in that order
    for dk in (-1, 0):
        for dj in (-1, 0):
            for di in (-1, 0):
                kk, jj, ii = np.meshgrid(np.arange(NZ) + dk, np.arange(NY) + dj, np.arange(NX) + di, indexing="ij")
                ok = (ii >= 0) & (jj >= 0) & (kk >= 0) & (ii < nx) & (jj < ny) & (kk < nz)
                v = np.where(ok, rho[np.clip(kk, 0, nz - 1), np.clip(jj, 0, ny - 1), np.clip(ii, 0, nx - 1)],
                             np.float32(0))
                sums = np.where(ok, (sums + v).astype(np.float32), sums)
                cnt += ok
    nod = np.where(cnt > 0, (sums / cnt.astype(np.float32)).astype(np.float32), np.float32(0))
    FX, FY, FZ = sub * nx + 3, sub * ny + 3, sub * nz + 3
    fine = np.zeros((FZ, FY, FX), np.float32)
    one = np.float32(1)
    for c in range(1, FZ - 1):
        iz = c - 1
        cz = min(iz // sub, nz - 1)
        zd = np.float32(iz - cz * sub) / np.float32(sub)
        for b in range(1, FY - 1):
            iy = b - 1
            cy = min(iy // sub, ny - 1)
            yd = np.float32(iy - cy * sub) / np.float32(sub)
            ix = np.arange(FX - 2)
            cx = np.minimum(ix // sub, nx - 1)
            xd = (ix - cx * sub).astype(np.float32) / np.float32(sub)
            v1, v2 = nod[cz, cy, cx], nod[cz, cy, cx + 1]
            v3, v4 = nod[cz, cy + 1, cx + 1], nod[cz, cy + 1, cx]
            v5, v6 = nod[cz + 1, cy, cx], nod[cz + 1, cy, cx + 1]
            v7, v8 = nod[cz + 1, cy + 1, cx + 1], nod[cz + 1, cy + 1, cx]
            c00 = v1 * (one - xd) + v2 * xd
            c01 = v4 * (one - xd) + v3 * xd
            c10 = v5 * (one - xd) + v6 * xd
            c11 = v8 * (one - xd) + v7 * xd
            c0 = c00 * (one - yd) + c01 * yd
            c1 = c10 * (one - yd) + c11 * yd
            fine[c, b, 1:FX - 1] = c0 * (one - zd) + c1 * zd
    return fine


def trilinear_volume(fine, iso, h, samples=4):
    """Volume of {field >= iso} for the trilinear interpolant of a sampling
    grid (spacing h), by midpoint sampling at `samples` points per cell edge."""
    FZ, FY, FX = fine.shape
    t = (np.arange(samples) + 0.5) / samples
    f = fine.astype(np.float64)
    tot = 0
    for tz in t:
        for ty in t:
            for tx in t:
                a = f[:-1, :-1, :-1] * (1 - tx) + f[:-1, :-1, 1:] * tx
                b = f[:-1, 1:, :-1] * (1 - tx) + f[:-1, 1:, 1:] * tx
                c = f[1:, :-1, :-1] * (1 - tx) + f[1:, :-1, 1:] * tx
                d = f[1:, 1:, :-1] * (1 - tx) + f[1:, 1:, 1:] * tx
                v = (a * (1 - ty) + b * ty) * (1 - tz) + (c * (1 - ty) + d * ty) * tz
                tot += int((v >= iso).sum())
    return tot / samples ** 3 * h[0] * h[1] * h[2]


# ----------------------------------------------------------------------------
# regrid field map (MeshRefiner.jl:131-150), FP32 as the reference
# ----------------------------------------------------------------------------

def regrid_map(old_dims, new_dims, lengths, field):
    onx, ony, onz = old_dims
    nnx, nny, nnz = new_dims
    lx, ly, lz = (np.float32(v) for v in lengths)
    ndx, ndy, ndz = lx / np.float32(nnx), ly / np.float32(nny), lz / np.float32(nnz)
    odx, ody, odz = lx / np.float32(onx), ly / np.float32(ony), lz / np.float32(onz)
    e = np.arange(nnx * nny * nnz)
    iz = e // (nnx * nny) + 1
    rem = e % (nnx * nny)
    iy = rem // nnx + 1
    ix = rem % nnx + 1
    cx = (ix.astype(np.float32) - np.float32(0.5)) * ndx
    cy = (iy.astype(np.float32) - np.float32(0.5)) * ndy
    cz = (iz.astype(np.float32) - np.float32(0.5)) * ndz
    ox_ = np.clip(np.floor(cx / odx).astype(np.int64) + 1, 1, onx)
    oy_ = np.clip(np.floor(cy / ody).astype(np.int64) + 1, 1, ony)
    oz_ = np.clip(np.floor(cz / odz).astype(np.int64) + 1, 1, onz)
    lin = ox_ + (oy_ - 1) * onx + (oz_ - 1) * onx * ony
    return np.asarray(field)[lin - 1]
