"""ctypes binding of libhexat.so (the C ABI declared in include/hexat.h).

The library is the product: every numeric call below runs hand-written gfx950
HIP kernels.  If the library is missing or no HIP device is visible the calls
raise -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhexat.so")

HX_OK = 0
HX_ERR_NO_DEVICE = 1
HX_ERR_BAD_ARG = 2
HX_ERR_NOT_CONVERGED = 3
HX_ERR_BREAKDOWN = 4
HX_ERR_DIVERGED = 5
HX_ERR_NO_ACTIVE = 6
HX_ERR_HIP = 7
HX_ERR_STATE = 8
HX_ERR_OOM = 9
HX_ZERO_RHS = 10

HX_FP64, HX_FP32_IR = 0, 1
HX_PC_JACOBI, HX_PC_LEARNED, HX_PC_MULTILEVEL, HX_PC_VCYCLE = 0, 1, 2, 3
HX_MV_AUTO, HX_MV_PARITY, HX_MV_DENSE, HX_MV_GATHER, HX_MV_MFMA = 0, 1, 2, 3, 4
HX_K_MATVEC, HX_K_UPDATE, HX_K_PUPDATE, HX_K_HELM, HX_K_ELEM, HX_K_MATVEC32, HX_K_PRECOND = range(7)
HX_FIELD_X, HX_FIELD_XPHYS, HX_FIELD_U, HX_FIELD_VM, HX_FIELD_DC = range(5)

STATUS_NAMES = {0: "ok", 1: "no device", 2: "bad argument", 3: "not converged", 4: "breakdown",
                5: "diverged", 6: "no active element", 7: "HIP error", 8: "bad state", 9: "out of memory",
                10: "zero rhs"}

#: every symbol declared in include/hexat.h
EXPORTS = [
    "hx_version", "hx_device_count", "hx_ctx_create", "hx_ctx_destroy", "hx_last_error", "hx_grid_set",
    "hx_set_fixed", "hx_element_stiffness", "hx_scalar_matrices", "hx_matvec", "hx_diagonal", "hx_pcg_solve",
    "hx_solve_system",
    "hx_block_pcg_solve", "hx_helmholtz_filter", "hx_element_stress", "hx_web_pack", "hx_compliance_sens",
    "hx_oc_update",
    "hx_to_init", "hx_to_step", "hx_to_read", "hx_learned_train", "hx_learned_weights", "hx_timing_enable",
    "hx_timing_read", "hx_device_info", "hx_autotune", "hx_tune_set", "hx_time_operator", "hx_time_operator_pc",
    "hx_time_smoothing", "hx_device_memory", "hx_web_pack_device", "hx_vtk_pack",
    "hx_stl_pack", "hx_bc_indicator", "hx_external_forces", "hx_regrid",
]
#: test hooks declared in include/hexat_testing.h
TEST_EXPORTS = ["hx_test_unet_nparam", "hx_test_unet_init", "hx_test_unet_forward", "hx_test_unet_backward",
                "hx_test_copy", "hx_test_ml_apply", "hx_test_vc_apply", "hx_test_device_alloc",
                "hx_test_device_free", "hx_test_memcpy", "hx_test_oc_search", "hx_test_smooth_field"]


class HexatError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


class PcgOpts(C.Structure):
    _fields_ = [("tol", C.c_double), ("max_iter", C.c_int), ("shift", C.c_double), ("precision", C.c_int),
                ("precond", C.c_int), ("warm_start", C.c_int), ("matvec", C.c_int), ("batch", C.c_int)]


class PcgStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("status", C.c_int), ("rel_res", C.c_double), ("seconds", C.c_double),
                ("refinements", C.c_int), ("fallbacks", C.c_int)]


HX_SOLVE_CONVERGED, HX_SOLVE_ACCEPTED, HX_SOLVE_FAILED, HX_SOLVE_ZERO_RHS = range(4)


class SolveOpts(C.Structure):
    _fields_ = [("E", C.c_double), ("min_stiffness_threshold", C.c_double), ("prune_voids", C.c_int),
                ("tol", C.c_double), ("max_iter", C.c_int), ("shift_factor", C.c_double), ("precision", C.c_int),
                ("precond", C.c_int), ("max_attempts", C.c_int), ("warm_start", C.c_int)]


class SolveStats(C.Structure):
    _fields_ = [("status", C.c_int), ("outcome", C.c_int), ("attempts", C.c_int), ("iterations", C.c_int),
                ("rel_res", C.c_double), ("true_rel", C.c_double), ("best_rel", C.c_double), ("shift", C.c_double),
                ("n_active", C.c_long), ("n_kept", C.c_long), ("pruned", C.c_int), ("seconds", C.c_double)]


class Location(C.Structure):
    _fields_ = [("all", C.c_int * 3), ("value", C.c_float * 3)]


class FilterOpts(C.Structure):
    _fields_ = [("quirk", C.c_int), ("tol", C.c_double), ("max_iter", C.c_int), ("check_every", C.c_int)]


class FilterStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("status", C.c_int), ("rel_res", C.c_double), ("seconds", C.c_double)]


class TOParams(C.Structure):
    _fields_ = [("E0", C.c_double), ("Emin", C.c_double), ("nu", C.c_double), ("penal", C.c_double),
                ("volfrac", C.c_double), ("move", C.c_double), ("xmin", C.c_double), ("rmin", C.c_double),
                ("pcg_tol", C.c_double), ("pcg_max_iter", C.c_int), ("precision", C.c_int), ("precond", C.c_int),
                ("filter_tol", C.c_double), ("filter_max_iter", C.c_int), ("oc_rel_tol", C.c_double),
                ("with_stress", C.c_int)]


class TOStats(C.Structure):
    _fields_ = [("iter", C.c_int), ("compliance", C.c_double), ("change", C.c_double), ("volume", C.c_double),
                ("pcg_iterations", C.c_int), ("pcg_rel_res", C.c_double), ("pcg_status", C.c_int),
                ("filter_iterations", C.c_int), ("max_vm", C.c_double), ("seconds", C.c_double)]


class TuneResult(C.Structure):
    _fields_ = [("tile", C.c_int), ("kchunk", C.c_int), ("variant64", C.c_int), ("variant32", C.c_int),
                ("ms64", C.c_double), ("ms32", C.c_double), ("ms32_parity", C.c_double), ("ms32_dense", C.c_double),
                ("ms32_mfma", C.c_double), ("tile32", C.c_int), ("kchunk32", C.c_int), ("pdepth", C.c_int),
                ("pdepth32", C.c_int), ("tile_pre", C.c_int), ("kchunk_pre", C.c_int), ("pdepth_pre", C.c_int),
                ("tile_post", C.c_int), ("kchunk_post", C.c_int), ("pdepth_post", C.c_int), ("ms_pre", C.c_double),
                ("ms_post", C.c_double)]


class LearnStats(C.Structure):
    _fields_ = [("samples", C.c_int), ("epochs", C.c_int), ("lr", C.c_double), ("seed", C.c_int),
                ("final_loss", C.c_double), ("seconds", C.c_double), ("jacobi_loss", C.c_double),
                ("manufactured", C.c_int)]


_lib = None
_lock = threading.Lock()


def load():
    """Load libhexat.so (raises if it has not been built)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python {_HERE}/build.py`")
        L = C.CDLL(LIB_PATH)
        P, d, i = C.POINTER, C.c_double, C.c_int
        pd = P(C.c_double)
        vp = C.c_void_p
        L.hx_version.restype = i
        L.hx_device_count.argtypes = [P(C.c_int)]
        L.hx_ctx_create.argtypes = [i, P(vp)]
        L.hx_ctx_destroy.argtypes = [vp]
        L.hx_ctx_destroy.restype = None
        L.hx_last_error.argtypes = [vp]
        L.hx_last_error.restype = C.c_char_p
        L.hx_grid_set.argtypes = [vp, i, i, i, d, d, d, d]
        L.hx_set_fixed.argtypes = [vp, P(C.c_uint8)]
        L.hx_element_stiffness.argtypes = [d, d, d, d, d, pd]
        L.hx_scalar_matrices.argtypes = [d, d, d, i, pd, pd]
        L.hx_matvec.argtypes = [vp, pd, pd, pd, i]
        L.hx_diagonal.argtypes = [vp, pd, pd]
        L.hx_pcg_solve.argtypes = [vp, pd, pd, pd, P(PcgOpts), P(PcgStats)]
        L.hx_block_pcg_solve.argtypes = [vp, i, pd, pd, pd, P(PcgOpts), P(PcgStats)]
        L.hx_solve_system.argtypes = [vp, pd, pd, pd, P(SolveOpts), P(SolveStats)]
        L.hx_helmholtz_filter.argtypes = [vp, d, P(FilterOpts), pd, pd, P(FilterStats)]
        L.hx_element_stress.argtypes = [vp, pd, pd, pd, pd, pd, pd, pd]
        L.hx_web_pack.argtypes = [vp, P(C.c_float), P(C.c_float), P(C.c_float), C.c_float, P(C.c_float),
                                  P(C.c_float), P(C.c_float), P(C.c_uint32)]
        L.hx_compliance_sens.argtypes = [vp, pd, d, d, d, pd, pd, pd, pd, pd]
        L.hx_web_pack_device.argtypes = [vp, vp, vp, vp, C.c_float, vp, vp, vp, vp]
        pu32 = P(C.c_uint32)
        L.hx_bc_indicator.argtypes = [vp, i, P(Location), P(C.c_int), P(C.c_uint8)]
        L.hx_external_forces.argtypes = [vp, i, P(Location), pd, P(C.c_int), pd, P(C.c_uint32)]
        L.hx_regrid.argtypes = [vp, i, i, i, i, i, i, C.c_float, C.c_float, C.c_float, i, P(C.c_float),
                                P(C.c_float)]
        L.hx_stl_pack.argtypes = [vp, P(C.c_float), i, C.c_float, P(C.c_float), C.c_long, P(C.c_long), P(C.c_float)]
        L.hx_vtk_pack.argtypes = [vp, pd, pd, P(C.c_uint8), P(C.c_float), P(C.c_float), P(C.c_float), C.c_float, pu32,
                                  pu32, pu32, pu32]
        L.hx_oc_update.argtypes = [vp, pd, pd, pd, d, d, d, d, pd, pd, P(C.c_int), pd]
        L.hx_to_init.argtypes = [vp, P(TOParams), pd, pd]
        L.hx_to_step.argtypes = [vp, P(TOStats)]
        L.hx_to_read.argtypes = [vp, i, pd]
        L.hx_learned_train.argtypes = [vp, pd, pd, P(LearnStats)]
        L.hx_learned_weights.argtypes = [vp, P(C.c_uint16), P(C.c_int)]
        L.hx_timing_enable.argtypes = [vp, i]
        L.hx_timing_read.argtypes = [vp, i, pd, P(C.c_long)]
        L.hx_device_info.argtypes = [vp, P(C.c_int), P(C.c_int), pd]
        L.hx_device_memory.argtypes = [vp, pd, pd]
        L.hx_autotune.argtypes = [vp, i, P(TuneResult)]
        L.hx_tune_set.argtypes = [vp, P(TuneResult)]
        L.hx_time_operator.argtypes = [vp, i, i, i, P(C.c_double)]
        L.hx_time_operator_pc.argtypes = [vp, i, i, i, i, P(C.c_double)]
        L.hx_time_smoothing.argtypes = [vp, i, i, i, P(C.c_double)]
        L.hx_test_copy.argtypes = [vp, C.c_long, i, P(C.c_double)]
        L.hx_test_device_alloc.argtypes = [vp, C.c_long, P(vp)]
        L.hx_test_device_free.argtypes = [vp, vp]
        L.hx_test_memcpy.argtypes = [vp, vp, vp, C.c_long, i]
        L.hx_test_ml_apply.argtypes = [vp, pd, pd, pd, i, P(C.c_int)]
        L.hx_test_oc_search.argtypes = [vp, pd, pd, pd, d, d, d, d, d, pd, pd, P(C.c_int)]
        L.hx_test_smooth_field.argtypes = [vp, C.c_ulonglong, i, pd]
        L.hx_test_vc_apply.argtypes = [vp, pd, pd, pd, i, P(C.c_int), P(C.c_int)]
        pf = P(C.c_float)
        L.hx_test_unet_nparam.argtypes = [vp, P(C.c_long)]
        L.hx_test_unet_init.argtypes = [vp, i, pf]
        L.hx_test_unet_forward.argtypes = [vp, P(C.c_uint16), pf]
        L.hx_test_unet_backward.argtypes = [vp, pf, pf]
        for name in EXPORTS + TEST_EXPORTS:
            f = getattr(L, name)
            if name not in ("hx_ctx_destroy", "hx_last_error"):
                f.restype = C.c_int
        _lib = L
        return L


def dptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(C.c_double))


def f64(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float64).ravel()
    if n is not None and a.size != n:
        raise ValueError(f"expected {n} values, got {a.size}")
    return a


def device_count() -> int:
    n = C.c_int(0)
    load().hx_device_count(C.byref(n))
    return n.value


def element_stiffness(dx, dy, dz, nu, E=1.0):
    K = np.empty(576)
    rc = load().hx_element_stiffness(float(dx), float(dy), float(dz), float(nu), float(E), dptr(K))
    if rc:
        raise HexatError(rc, "hx_element_stiffness")
    return K.reshape(24, 24)


def scalar_matrices(dx, dy, dz, quirk):
    Ks, Ms = np.empty(64), np.empty(64)
    rc = load().hx_scalar_matrices(float(dx), float(dy), float(dz), int(quirk), dptr(Ks), dptr(Ms))
    if rc:
        raise HexatError(rc, "hx_scalar_matrices")
    return Ks.reshape(8, 8), Ms.reshape(8, 8)


class Context:
    """One HIP device context with a structured grid (libhexat hx_ctx)."""

    def __init__(self, device: int = 0):
        self.L = load()
        h = C.c_void_p()
        rc = self.L.hx_ctx_create(int(device), C.byref(h))
        if rc:
            raise HexatError(rc, "hx_ctx_create: no usable HIP device" if rc == HX_ERR_NO_DEVICE else "hx_ctx_create")
        self.h = h
        self.grid = None

    def close(self):
        if getattr(self, "h", None):
            self.L.hx_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def check(self, rc, what, ok=(HX_OK,)):
        if rc not in ok:
            msg = self.L.hx_last_error(self.h)
            raise HexatError(rc, f"{what}: {msg.decode() if msg else ''}")
        return rc

    # ---- grid -------------------------------------------------------------
    def set_grid(self, nx, ny, nz, dx=1.0, dy=1.0, dz=1.0, nu=0.3):
        self.check(self.L.hx_grid_set(self.h, int(nx), int(ny), int(nz), float(dx), float(dy), float(dz),
                                      float(nu)), "hx_grid_set")
        self.grid = (int(nx), int(ny), int(nz), float(dx), float(dy), float(dz), float(nu))
        return self

    @property
    def ne(self):
        nx, ny, nz = self.grid[:3]
        return nx * ny * nz

    @property
    def nn(self):
        nx, ny, nz = self.grid[:3]
        return (nx + 1) * (ny + 1) * (nz + 1)

    @property
    def ndof(self):
        return 3 * self.nn

    def set_fixed(self, fixed):
        if fixed is None:
            self.check(self.L.hx_set_fixed(self.h, None), "hx_set_fixed")
            return
        m = np.ascontiguousarray(fixed, dtype=np.uint8).ravel()
        if m.size != self.ndof:
            raise ValueError("fixed mask must have 3*nNodes entries")
        self._fixed = m
        self.check(self.L.hx_set_fixed(self.h, m.ctypes.data_as(C.POINTER(C.c_uint8))), "hx_set_fixed")

    # ---- elasticity ---------------------------------------------------------
    def matvec(self, Ee, x, variant=HX_MV_AUTO):
        Ee, x = f64(Ee, self.ne), f64(x, self.ndof)
        y = np.empty(self.ndof)
        self.check(self.L.hx_matvec(self.h, dptr(Ee), dptr(x), dptr(y), int(variant)), "hx_matvec")
        return y

    def diagonal(self, Ee):
        Ee = f64(Ee, self.ne)
        d = np.empty(self.ndof)
        self.check(self.L.hx_diagonal(self.h, dptr(Ee), dptr(d)), "hx_diagonal")
        return d

    def pcg(self, Ee, f, u0=None, tol=1e-8, max_iter=100000, shift=0.0, precision=HX_FP64,
            precond=HX_PC_JACOBI, matvec=HX_MV_AUTO, batch=0, raise_on_fail=False):
        Ee, f = f64(Ee, self.ne), f64(f, self.ndof)
        u = np.zeros(self.ndof) if u0 is None else f64(u0, self.ndof).copy()
        o = PcgOpts(float(tol), int(max_iter), float(shift), int(precision), int(precond),
                    0 if u0 is None else 1, int(matvec), int(batch))
        s = PcgStats()
        rc = self.L.hx_pcg_solve(self.h, dptr(Ee), dptr(f), dptr(u), C.byref(o), C.byref(s))
        ok = (HX_OK,) if raise_on_fail else (HX_OK, HX_ERR_NOT_CONVERGED, HX_ERR_BREAKDOWN, HX_ERR_DIVERGED)
        self.check(rc, "hx_pcg_solve", ok)
        return u, s

    def solve_system(self, density, f, u_prev=None, *, E=1.0, min_stiffness_threshold=1e-3, prune_voids=True,
                     tol=1e-6, max_iter=40000, shift_factor=1e-6, precision=HX_FP64, precond=HX_PC_JACOBI,
                     max_attempts=3):
        """hx_solve_system: the reference's solve policy (pruning, shift, retries,
        best iterate) on the device.  Returns (u, SolveStats); raises only on
        errors other than "all attempts failed" (stats.outcome tells)."""
        rho, f = f64(density, self.ne), f64(f, self.ndof)
        u = np.zeros(self.ndof) if u_prev is None else f64(u_prev, self.ndof).copy()
        o = SolveOpts(float(E), float(min_stiffness_threshold), 1 if prune_voids else 0, float(tol), int(max_iter),
                      float(shift_factor), int(precision), int(precond), int(max_attempts),
                      0 if u_prev is None else 1)
        st = SolveStats()
        rc = self.L.hx_solve_system(self.h, dptr(rho), dptr(f), dptr(u), C.byref(o), C.byref(st))
        self.check(rc, "hx_solve_system", (HX_OK, HX_ERR_NOT_CONVERGED))
        return u, st

    def block_pcg(self, Ee, F, tol=1e-8, max_iter=100000, shift=0.0, precision=HX_FP64, matvec=HX_MV_AUTO,
                  precond=HX_PC_JACOBI):
        Ee = f64(Ee, self.ne)
        F = np.ascontiguousarray(F, dtype=np.float64)
        nrhs = F.shape[0]
        if F.shape != (nrhs, self.ndof):
            raise ValueError("F must be nrhs x ndof")
        U = np.zeros_like(F)
        o = PcgOpts(float(tol), int(max_iter), float(shift), int(precision), int(precond), 0, int(matvec), 0)
        S = (PcgStats * nrhs)()
        rc = self.L.hx_block_pcg_solve(self.h, nrhs, dptr(Ee), dptr(F), dptr(U), C.byref(o), S)
        self.check(rc, "hx_block_pcg_solve", (HX_OK, HX_ERR_NOT_CONVERGED))
        return U, list(S)

    # ---- filter ---------------------------------------------------------------
    def helmholtz(self, R, rho, quirk=0, tol=1e-5, max_iter=200, check_every=10):
        rho = f64(rho, self.ne)
        out = np.empty(self.ne)
        o = FilterOpts(int(quirk), float(tol), int(max_iter), int(check_every))
        s = FilterStats()
        self.check(self.L.hx_helmholtz_filter(self.h, float(R), C.byref(o), dptr(rho), dptr(out), C.byref(s)),
                   "hx_helmholtz_filter")
        return out, s

    # ---- element post ----------------------------------------------------------
    def element_stress(self, Ee, u, principal=True, vm=True, voigt=True, l1=True, direction=True):
        Ee, u = f64(Ee, self.ne), f64(u, self.ndof)
        ne = self.ne
        out = {"principal": np.empty((ne, 3)) if principal else None, "vm": np.empty(ne) if vm else None,
               "voigt": np.empty((ne, 6)) if voigt else None, "l1": np.empty(ne) if l1 else None,
               "dir": np.empty((ne, 3)) if direction else None}
        self.check(self.L.hx_element_stress(self.h, dptr(Ee), dptr(u), dptr(out["principal"]), dptr(out["vm"]),
                                            dptr(out["voigt"]), dptr(out["l1"]), dptr(out["dir"])),
                   "hx_element_stress")
        return out

    def web_pack(self, density, l1, principal, threshold):
        """Kept elements of the viewer/checkpoint file (density >= threshold,
        element order): (centroids count x 3, densities, signed L1), FP32."""
        ne = self.ne
        rho = np.ascontiguousarray(density, dtype=np.float32).reshape(ne)
        l1v = np.ascontiguousarray(l1, dtype=np.float32).reshape(ne)
        pr = np.ascontiguousarray(principal, dtype=np.float32).reshape(ne, 3)
        cent = np.empty((ne, 3), np.float32)
        dout = np.empty(ne, np.float32)
        sl1 = np.empty(ne, np.float32)
        n = C.c_uint32(0)
        fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
        self.check(self.L.hx_web_pack(self.h, fp(rho), fp(l1v), fp(pr), C.c_float(np.float32(threshold)),
                                      fp(cent), fp(dout), fp(sl1), C.byref(n)), "hx_web_pack")
        k = int(n.value)
        return cent[:k].copy(), dout[:k].copy(), sl1[:k].copy()

    def vtk_pack(self, u, f, fixed, density, l1, vm, threshold):
        """Big-endian data blocks of the legacy VTK file (hx_vtk_pack):
        dict of uint32 arrays points (13 nn), cells (9 n), l1, vm, density (n)."""
        nn, ne = self.nn, self.ne
        fp = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float32).reshape(ne)
        rho, l1v, vmv = fp(density), fp(l1), fp(vm)
        uu = None if u is None else f64(u, self.ndof)
        ff = None if f is None else f64(f, self.ndof)
        fx = None if fixed is None else np.ascontiguousarray(fixed, dtype=np.uint8).reshape(self.ndof)
        pts = np.empty(13 * nn, np.uint32)
        cells = np.empty(9 * ne, np.uint32)
        cs = np.empty(3 * ne, np.uint32)
        n = C.c_uint32(0)
        pf = lambda a: None if a is None else a.ctypes.data_as(C.POINTER(C.c_float))
        pu = lambda a: a.ctypes.data_as(C.POINTER(C.c_uint32))
        self.check(self.L.hx_vtk_pack(self.h, dptr(uu), dptr(ff),
                                      None if fx is None else fx.ctypes.data_as(C.POINTER(C.c_uint8)), pf(rho),
                                      pf(l1v), pf(vmv), C.c_float(np.float32(threshold)), pu(pts), pu(cells), pu(cs),
                                      C.byref(n)), "hx_vtk_pack")
        k = int(n.value)
        return {"points": pts, "cells": cells[:9 * k].copy(), "l1": cs[:k].copy(), "vm": cs[k:2 * k].copy(),
                "density": cs[2 * k:3 * k].copy(), "count": k}

    def bc_indicator(self, locs, dof_bits):
        """hx_bc_indicator: locs = [(all3, value3)], dof_bits per spec -> fixed mask (3 nn uint8)."""
        n = len(locs)
        L = (Location * max(n, 1))(*[Location((C.c_int * 3)(*a), (C.c_float * 3)(*v)) for a, v in locs])
        bits = (C.c_int * max(n, 1))(*dof_bits)
        out = np.empty(self.ndof, np.uint8)
        self.check(self.L.hx_bc_indicator(self.h, n, L, bits, out.ctypes.data_as(C.POINTER(C.c_uint8))),
                   "hx_bc_indicator")
        return out

    def external_forces(self, locs, forces, point_fallback, F):
        """hx_external_forces: adds into F (ndof float64, returned) -> (F, matched)."""
        n = len(locs)
        L = (Location * max(n, 1))(*[Location((C.c_int * 3)(*a), (C.c_float * 3)(*v)) for a, v in locs])
        fv = np.ascontiguousarray(forces, dtype=np.float64).reshape(max(n, 0) * 3)
        pf = (C.c_int * max(n, 1))(*[1 if p else 0 for p in point_fallback])
        F = f64(F, self.ndof).copy()
        m = np.zeros(max(n, 1), np.uint32)
        self.check(self.L.hx_external_forces(self.h, n, L, dptr(fv), pf, dptr(F),
                                             m.ctypes.data_as(C.POINTER(C.c_uint32))), "hx_external_forces")
        return F, m[:n]

    def regrid(self, old_dims, new_dims, lengths, fields):
        """hx_regrid: fields (nf x old ne, float32) -> (nf x new ne)."""
        fo = np.ascontiguousarray(fields, dtype=np.float32)
        nf = fo.shape[0] if fo.ndim == 2 else 1
        nne = int(np.prod(new_dims))
        out = np.empty((nf, nne), np.float32)
        pf = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
        self.check(self.L.hx_regrid(self.h, *[int(v) for v in old_dims], *[int(v) for v in new_dims],
                                    *[C.c_float(np.float32(v)) for v in lengths], nf, pf(fo), pf(out)), "hx_regrid")
        return out

    def stl_pack(self, density, threshold, subdivision=1, field=False):
        """Isosurface triangles (count x 12 FP32: normal, v1, v2, v3) of
        hx_stl_pack, and the sampling grid when field is True."""
        rho = np.ascontiguousarray(density, dtype=np.float32).reshape(self.ne)
        nx, ny, nz = self.grid[:3]
        s = int(subdivision)
        fld = np.empty((s * nz + 3, s * ny + 3, s * nx + 3), np.float32) if field else None
        n = C.c_long(0)
        pf = lambda a: None if a is None else a.ctypes.data_as(C.POINTER(C.c_float))
        self.check(self.L.hx_stl_pack(self.h, pf(rho), s, C.c_float(np.float32(threshold)), None, 0, C.byref(n),
                                      pf(fld)), "hx_stl_pack")
        tris = np.empty((n.value, 12), np.float32)
        if n.value:
            self.check(self.L.hx_stl_pack(self.h, pf(rho), s, C.c_float(np.float32(threshold)), pf(tris), n.value,
                                          C.byref(n), None), "hx_stl_pack")
        return (tris, fld) if field else tris

    def compliance_sens(self, u, xphys=None, penal=3.0, E0=1.0, Emin=1e-9, f=None):
        u = f64(u, self.ndof)
        xp = None if xphys is None else f64(xphys, self.ne)
        ff = None if f is None else f64(f, self.ndof)
        ce = np.empty(self.ne)
        dc = np.empty(self.ne) if xp is not None else None
        comp = np.zeros(1) if ff is not None else None
        self.check(self.L.hx_compliance_sens(self.h, dptr(xp), float(penal), float(E0), float(Emin), dptr(u),
                                             dptr(ff), dptr(ce), dptr(dc), dptr(comp)), "hx_compliance_sens")
        return ce, dc, (None if comp is None else float(comp[0]))

    def oc_update(self, x, dc, dv, volfrac, move=0.2, xmin=0.0, rel_tol=1e-3):
        x, dc, dv = f64(x, self.ne), f64(dc, self.ne), f64(dv, self.ne)
        xn = np.empty(self.ne)
        lm, ch = C.c_double(0), C.c_double(0)
        st = C.c_int(0)
        self.check(self.L.hx_oc_update(self.h, dptr(x), dptr(dc), dptr(dv), float(volfrac), float(move),
                                       float(xmin), float(rel_tol), dptr(xn), C.byref(lm), C.byref(st),
                                       C.byref(ch)), "hx_oc_update")
        return xn, lm.value, st.value, ch.value

    def test_oc_search(self, x, dc, dv, volfrac, move=0.2, xmin=0.0, rel_tol=1e-9, lguess=0.0):
        """hx_test_oc_search: the SIMP loop's 8-point OC search -> (x_new, multiplier, passes)"""
        x, dc, dv = f64(x, self.ne), f64(dc, self.ne), f64(dv, self.ne)
        xn = np.empty(self.ne)
        lm, st = C.c_double(0), C.c_int(0)
        self.check(self.L.hx_test_oc_search(self.h, dptr(x), dptr(dc), dptr(dv), float(volfrac), float(move),
                                            float(xmin), float(rel_tol), float(lguess), dptr(xn), C.byref(lm),
                                            C.byref(st)), "hx_test_oc_search")
        return xn, lm.value, st.value

    def test_smooth_field(self, seed, levels):
        """hx_test_smooth_field: the training's manufactured displacement (ndof values)"""
        e = np.empty(self.ndof)
        self.check(self.L.hx_test_smooth_field(self.h, C.c_ulonglong(int(seed)), int(levels), dptr(e)),
                   "hx_test_smooth_field")
        return e

    # ---- device-resident SIMP loop ------------------------------------------------
    def to_init(self, params: TOParams, f, x0):
        f, x0 = f64(f, self.ndof), f64(x0, self.ne)
        self.check(self.L.hx_to_init(self.h, C.byref(params), dptr(f), dptr(x0)), "hx_to_init")

    def to_step(self) -> TOStats:
        s = TOStats()
        self.check(self.L.hx_to_step(self.h, C.byref(s)), "hx_to_step")
        return s

    def to_read(self, field):
        n = self.ndof if field == HX_FIELD_U else self.ne
        out = np.empty(n)
        self.check(self.L.hx_to_read(self.h, int(field), dptr(out)), "hx_to_read")
        return out

    # ---- learned preconditioner -----------------------------------------------------
    def learned_train(self, Ee=None, f=None, samples=64, epochs=4, lr=1e-3, seed=0, manufactured=0):
        Ee = None if Ee is None else f64(Ee, self.ne)
        f = None if f is None else f64(f, self.ndof)
        s = LearnStats(int(samples), int(epochs), float(lr), int(seed), 0.0, 0.0, 0.0, int(manufactured))
        self.check(self.L.hx_learned_train(self.h, dptr(Ee), dptr(f), C.byref(s)), "hx_learned_train")
        return s

    def learned_weights(self):
        n = C.c_int(0)
        self.check(self.L.hx_learned_weights(self.h, None, C.byref(n)), "hx_learned_weights")
        w = np.empty(n.value, dtype=np.uint16)
        self.check(self.L.hx_learned_weights(self.h, w.ctypes.data_as(C.POINTER(C.c_uint16)), C.byref(n)),
                   "hx_learned_weights")
        return w

    # ---- network test hooks (include/hexat_testing.h) ---------------------------
    def unet_nparam(self):
        n = C.c_long(0)
        self.check(self.L.hx_test_unet_nparam(self.h, C.byref(n)), "hx_test_unet_nparam")
        return n.value

    def unet_init(self, seed=0):
        p = np.empty(self.unet_nparam(), dtype=np.float32)
        self.check(self.L.hx_test_unet_init(self.h, int(seed), p.ctypes.data_as(C.POINTER(C.c_float))),
                   "hx_test_unet_init")
        return p

    def unet_forward(self, x0_bf16):
        x0 = np.ascontiguousarray(x0_bf16, dtype=np.uint16)
        if x0.size != self.nn * 8:
            raise ValueError("x0 must be nNodes x 8 BF16")
        y = np.empty((self.nn, 3), dtype=np.float32)
        self.check(self.L.hx_test_unet_forward(self.h, x0.ctypes.data_as(C.POINTER(C.c_uint16)),
                                               y.ctypes.data_as(C.POINTER(C.c_float))), "hx_test_unet_forward")
        return y

    def unet_backward(self, gy):
        gy = np.ascontiguousarray(gy, dtype=np.float32)
        g = np.empty(self.unet_nparam(), dtype=np.float32)
        self.check(self.L.hx_test_unet_backward(self.h, gy.ctypes.data_as(C.POINTER(C.c_float)),
                                                g.ctypes.data_as(C.POINTER(C.c_float))), "hx_test_unet_backward")
        return g

    def autotune(self, nrhs=1):
        r = TuneResult()
        self.check(self.L.hx_autotune(self.h, int(nrhs), C.byref(r)), "hx_autotune")
        return {k: getattr(r, k) for k, _ in TuneResult._fields_}

    def tune_set(self, tune: dict):
        r = TuneResult(**{k: tune[k] for k, _ in TuneResult._fields_ if k in tune})
        self.check(self.L.hx_tune_set(self.h, C.byref(r)), "hx_tune_set")

    def time_operator(self, precision=HX_FP64, nrhs=1, reps=20, precond=HX_PC_JACOBI) -> float:
        """Average ms of the PCG operator launch (back to back, rotating vector sets) in the form the
        preconditioner uses (fused D^-1 r + beta p for Jacobi, z + beta p for the multilevel one)."""
        ms = C.c_double(0)
        self.check(self.L.hx_time_operator_pc(self.h, int(precision), int(precond), int(nrhs), int(reps),
                                              C.byref(ms)), "hx_time_operator_pc")
        return ms.value

    def time_smoothing(self, precision=HX_FP32_IR, form=4, reps=20) -> float:
        """Average ms of the fine level's V-cycle smoothing operator (form 3 pre-, 4 post-smoothing with
        the fused prolongation), back to back on rotating vector sets."""
        ms = C.c_double(0)
        self.check(self.L.hx_time_smoothing(self.h, int(precision), int(form), int(reps), C.byref(ms)),
                   "hx_time_smoothing")
        return ms.value

    def test_copy(self, nbytes: int, reps: int = 4) -> float:
        ms = C.c_double(0)
        self.check(self.L.hx_test_copy(self.h, int(nbytes), int(reps), C.byref(ms)), "hx_test_copy")
        return ms.value

    def test_ml_apply(self, Ee, r, precision=HX_FP64):
        """z = M^-1 r of the multilevel preconditioner; returns (z, coarse levels)."""
        Ee, r = f64(Ee, self.ne), f64(r, self.ndof)
        z = np.empty(self.ndof)
        lv = C.c_int(0)
        self.check(self.L.hx_test_ml_apply(self.h, dptr(Ee), dptr(r), dptr(z), int(precision), C.byref(lv)),
                   "hx_test_ml_apply")
        return z, lv.value

    def test_vc_apply(self, Ee, r, precision=HX_FP64):
        """z = M^-1 r of the V-cycle form; returns (z, coarse levels, smoothed levels)."""
        Ee, r = f64(Ee, self.ne), f64(r, self.ndof)
        z = np.empty(self.ndof)
        lv, vl = C.c_int(0), C.c_int(0)
        self.check(self.L.hx_test_vc_apply(self.h, dptr(Ee), dptr(r), dptr(z), int(precision), C.byref(lv),
                                           C.byref(vl)), "hx_test_vc_apply")
        return z, lv.value, vl.value

    # ---- timing -----------------------------------------------------------------------
    def timing(self, on: bool):
        self.check(self.L.hx_timing_enable(self.h, 1 if on else 0), "hx_timing_enable")

    def timing_read(self, kernel):
        ms, cnt = C.c_double(0), C.c_long(0)
        self.check(self.L.hx_timing_read(self.h, int(kernel), C.byref(ms), C.byref(cnt)), "hx_timing_read")
        return ms.value, cnt.value

    def memory_used_gb(self):
        used, tot = C.c_double(0), C.c_double(0)
        self.check(self.L.hx_device_memory(self.h, C.byref(used), C.byref(tot)), "hx_device_memory")
        return used.value

    def device_info(self):
        cu, clk = C.c_int(0), C.c_int(0)
        mem = C.c_double(0)
        self.check(self.L.hx_device_info(self.h, C.byref(cu), C.byref(clk), C.byref(mem)), "hx_device_info")
        return {"cu": cu.value, "clock_khz": clk.value, "mem_gb": mem.value}
