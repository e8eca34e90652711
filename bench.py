"""Benchmark: seconds per topology-optimisation iteration at 12.6M DOFs.

Workload (one GPU): cantilever on a 256x128x128 grid of unit hex elements
(4,194,304 elements, 12,830,211 DOFs), clamped x=0 face, unit downward line
load split over the bottom edge at x=256, SIMP p=3, E0=1, Emin=1e-9, nu=0.3,
volfrac 0.15, PDE filter r=2.5, PCG to relative residual 1e-8, von Mises
stress at every element centroid each iteration.  A "step" is one full TO
iteration on the device: SIMP moduli, elasticity PCG (warm start),
compliance, sensitivities + von Mises, sensitivity filter, OC, density filter.

Two forms run from the same start: the plain FP64 Jacobi-PCG and the
precision form (FP32 element contraction and vectors with FP64 residual
replacement, dot products and refinement).  ``value`` is the precision form's
s/iteration (lower is better); the plain form's number and the speed-up are
reported beside it.  All inputs are resident in HBM during the timed region.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--grid 256x128x128]
"""
from __future__ import annotations

import argparse
import csv
import glob
import importlib
import json
import math
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

_wl = importlib.import_module("2025_12_09_hexat_v069_amd.workloads")
cantilever, mbb_half, block_loads, smoothed_density = (_wl.cantilever, _wl.mbb_half, _wl.block_loads,
                                                       _wl.smoothed_density)

METRIC = "s/TO-iteration at 12.6M DOFs: plain FP64 PCG vs FP32-MFMA+learned-PC"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def other_workload(args, hx, L, ctx):
    """Solver-only measurements of the other configurations (not the headline
    metric): 'mbb' -- FP64 vs precision-form PCG on the 192x96x96 half MBB
    beam; 'block8' -- 8 load cases on 128x64x64 as one block solve vs eight
    single solves (precision form)."""
    out = {"workload": args.workload}
    if args.workload == "mbb":
        nx, ny, nz = 192, 96, 96
        fixed, f = mbb_half(nx, ny, nz)
        ctx.set_grid(nx, ny, nz, 1.0, 1.0, 1.0, 0.3)
        ctx.set_fixed(fixed)
        ctx.tune_set(ctx.autotune(1))
        Ee = 1e-9 + smoothed_density(ctx, nx * ny * nz, 1, 2.0) ** 3
        for name, prec in (("plain", L.HX_FP64), ("precision", L.HX_FP32_IR)):
            ctx.pcg(Ee, f, tol=1e-8, precision=prec)  # warm-up (graphs)
            t0 = time.perf_counter()
            u, st = ctx.pcg(Ee, f, tol=1e-8, precision=prec)
            out[name] = {"s": time.perf_counter() - t0, "iterations": st.iterations, "rel_res": st.rel_res,
                         "compliance": float(f @ u), "refinements": st.refinements}
            log(f"  {name}: {out[name]}")
        out["speedup"] = out["plain"]["s"] / out["precision"]["s"]
        out["compliance_rel_diff"] = abs(out["plain"]["compliance"] - out["precision"]["compliance"]) / abs(
            out["plain"]["compliance"])
        out["config"] = "half MBB 192x96x96 (5,447,811 DOFs), smoothed random density, SIMP p=3, tol 1e-8"
    elif args.workload == "cfg4":
        # the learned-preconditioner configuration: the network trained on
        # manufactured-solution pairs streamed on a G^3 grid, then deployed in
        # the flexible CG at 320x160x160, beside the main form and plain FP64
        # Jacobi; solver only, then one TO iteration per form
        nx, ny, nz = 320, 160, 160
        learned = None
        if args.learn_samples > 0:
            G = args.learn_grid
            tfix, _ = cantilever(G, G, G)
            ctx.set_grid(G, G, G, 1.0, 1.0, 1.0, 0.3)
            ctx.set_fixed(tfix)
            chunk, done, tt, loss, jl = 2000, 0, 0.0, [], []
            while done < args.learn_samples:
                ls = ctx.learned_train(samples=min(chunk, args.learn_samples - done), epochs=1, lr=1e-3,
                                       seed=done // chunk, manufactured=1)
                done += ls.samples
                tt += ls.seconds
                loss.append(ls.final_loss)
                jl.append(ls.jacobi_loss)
                log(f"  cfg4 training on {G}^3: {done} pairs, loss {ls.final_loss:.4f} (D^-1 r alone "
                    f"{ls.jacobi_loss:.4f}), {tt:.1f} s")
            learned = {"train_grid": f"{G}x{G}x{G}", "pairs": done, "train_s": round(tt, 2),
                       "loss_per_chunk": [round(v, 4) for v in loss], "jacobi_loss": round(float(np.mean(jl)), 4),
                       "source": "manufactured solutions: smooth random displacement e*, load A e* on random "
                                 "thresholded densities; (r, e) = (A e*, e*) streamed into training"}
        fixed, f = cantilever(nx, ny, nz)
        ctx.set_grid(nx, ny, nz, 1.0, 1.0, 1.0, 0.3)
        ctx.set_fixed(fixed)
        ctx.tune_set(ctx.autotune(1))
        Ee = 1e-9 + smoothed_density(ctx, nx * ny * nz, 4, 2.0) ** 3
        res = {}
        forms = [("precision_vc", L.HX_FP32_IR, L.HX_PC_VCYCLE), ("plain", L.HX_FP64, L.HX_PC_JACOBI)]
        if learned is not None:
            forms.append(("learned", L.HX_FP64, L.HX_PC_LEARNED))
        for name, prec, pc in forms:
            if pc != L.HX_PC_LEARNED:
                ctx.pcg(Ee, f, tol=1e-8, precision=prec, precond=pc)  # warm-up (graphs)
            t0 = time.perf_counter()
            u, st = ctx.pcg(Ee, f, tol=1e-8, precision=prec, precond=pc)
            res[name] = (time.perf_counter() - t0, st, u)
            log(f"  cfg4 {name}: {res[name][0]:.4f} s, {st.iterations} iterations, rel {st.rel_res:.3e}")
        (tm, sm, um), (tp, sp, up) = res["precision_vc"], res["plain"]
        cm, cp = float(f @ um), float(f @ up)
        # one TO iteration per form from uniform density (the main configuration's
        # volume fraction and filter radius), its results against plain FP64
        to = {}
        for name, prec, pc in forms:
            ctx.to_init(to_params(L, args, prec, pc), f, np.full(nx * ny * nz, args.volfrac))
            s1 = ctx.to_step()
            to[name] = {"s_per_iter": round(s1.seconds, 4), "pcg_iterations": s1.pcg_iterations,
                        "pcg_rel_res": s1.pcg_rel_res, "compliance": s1.compliance,
                        "xphys": ctx.to_read(L.HX_FIELD_XPHYS)}
            log(f"  cfg4 TO iteration {name}: {s1.seconds:.3f} s, {s1.pcg_iterations} PCG iterations, "
                f"c={s1.compliance:.6e}")
        for name in to:
            to[name]["compliance_rel_diff_vs_plain"] = abs(to[name]["compliance"] - to["plain"]["compliance"]) / abs(
                to["plain"]["compliance"])
            to[name]["xphys_max_abs_diff_vs_plain"] = float(np.abs(to[name]["xphys"] - to["plain"]["xphys"]).max())
        for name in to:
            del to[name]["xphys"]
        out = {"metric": "s per solve at 24.96M DOFs (320x160x160, tol 1e-8)", "value": round(tm, 5), "unit": "s",
               "n_gpus": 1, "higher_is_better": False, "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
               "config": {"workload": "320x160x160 cantilever (8,192,000 elements, 24,961,923 DOFs), smoothed random "
                                      "density (seed 4, r 2), SIMP p=3, solver only, tol 1e-8; main form FP32 "
                                      "operator and vectors with FP64 reliable updates, V-cycle"},
               "iterations": sm.iterations, "rel_res": sm.rel_res,
               "plain_fp64": {"s": round(tp, 5), "iterations": sp.iterations, "rel_res": sp.rel_res},
               "speedup_vs_plain_fp64": round(tp / tm, 2),
               "compliance_rel_diff": abs(cm - cp) / cp,
               "displacement_rel_diff": float(np.linalg.norm(um - up) / np.linalg.norm(up)),
               "to_iteration": to}
        if learned is not None:
            tl, sl, ul = res["learned"]
            learned.update({"s": round(tl, 5), "iterations": sl.iterations, "rel_res": sl.rel_res,
                            "fallbacks": sl.fallbacks, "restarts": sl.refinements, "status": sl.status,
                            "iterations_vs_plain_jacobi": round(sl.iterations / sp.iterations, 3),
                            "time_vs_plain_jacobi": round(tl / tp, 3),
                            "compliance_rel_diff_vs_plain": abs(float(f @ ul) - cp) / cp,
                            "displacement_rel_diff_vs_plain": float(np.linalg.norm(ul - up) / np.linalg.norm(up))})
            out["learned"] = learned
    elif args.workload == "block8":
        # 8 load cases in one block solve (every launch carries the 8 columns;
        # the V-cycle's one-workgroup pass runs one workgroup per column)
        # against 8 single solves, precision form, Jacobi and V-cycle
        nx, ny, nz = 128, 64, 64
        fixed, _ = cantilever(nx, ny, nz)
        ctx.set_grid(nx, ny, nz, 1.0, 1.0, 1.0, 0.3)
        ctx.set_fixed(fixed)
        F = block_loads(nx, ny, nz, 8, seed=8)
        Ee = 1e-9 + smoothed_density(ctx, nx * ny * nz, 3, 2.0) ** 3
        tune8, tune1 = ctx.autotune(8), ctx.autotune(1)
        forms = {}
        for name, pc in (("jacobi", L.HX_PC_JACOBI), ("vcycle", L.HX_PC_VCYCLE)):
            ctx.tune_set(tune8)
            ctx.block_pcg(Ee, F, tol=1e-8, precision=L.HX_FP32_IR, precond=pc)  # warm-up (graphs)
            t0 = time.perf_counter()
            U, S = ctx.block_pcg(Ee, F, tol=1e-8, precision=L.HX_FP32_IR, precond=pc)
            tb = time.perf_counter() - t0
            ctx.tune_set(tune1)
            ctx.pcg(Ee, F[0], tol=1e-8, precision=L.HX_FP32_IR, precond=pc)  # warm-up
            t0 = time.perf_counter()
            cs, its = [], []
            for c in range(8):
                u, st = ctx.pcg(Ee, F[c], tol=1e-8, precision=L.HX_FP32_IR, precond=pc)
                cs.append(float(F[c] @ u))
                its.append(st.iterations)
            ts = time.perf_counter() - t0
            wc_block = sum(float(F[c] @ U[c]) for c in range(8)) / 8
            forms[name] = {"block_s": round(tb, 5), "single_s": round(ts, 5), "speedup_vs_8_singles": round(ts / tb, 3),
                           "block_iterations": [s.iterations for s in S], "single_iterations": its,
                           "weighted_compliance": wc_block,
                           "weighted_compliance_rel_diff": abs(wc_block - sum(cs) / 8) / abs(wc_block),
                           "max_rel_res": max(s.rel_res for s in S)}
            log(f"  block8 {name}: {forms[name]}")
        main = forms["vcycle"]
        out = {"metric": "s per 8-load-case solve (block PCG, 128x64x64)", "value": main["block_s"], "unit": "s",
               "n_gpus": 1, "higher_is_better": False, "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
               "config": {"workload": "block8: 128x64x64 cantilever (1,635,075 DOFs per load case), 8 unit loads on "
                                      "the x=128 face (seed 8), smoothed random density, FP32 operator and vectors "
                                      "with FP64 reliable updates, V-cycle, tol 1e-8"},
               "speedup_vs_8_single_solves": main["speedup_vs_8_singles"], "forms": forms}
    print(json.dumps(out), flush=True)


def level_dims(nx, ny, nz):
    """Element counts of the preconditioner's levels: 2:1 coarsenings while every count is even,
    ending at the first level with at most 192 DOFs (solved exactly)."""
    dims = [(nx, ny, nz)]
    while all(d % 2 == 0 for d in dims[-1]):
        dims.append(tuple(d // 2 for d in dims[-1]))
        a, b, c = dims[-1]
        if 3 * (a + 1) * (b + 1) * (c + 1) <= 192:
            break
    return dims


def smoothed_levels(nx, ny, nz, word):
    """Levels the V-cycle smooths with multi-block launches: those below the
    first level (>= 1) of the one-workgroup pass, which holds at most 512
    nodes and whose LDS (the finer level's residual or the element outputs,
    each pass level's moduli and smoother diagonal, r, x, t of every level
    from it down) fits 128 KiB -- the library's rule."""
    dims = level_dims(nx, ny, nz)
    nn = [(a + 1) * (b + 1) * (c + 1) for a, b, c in dims]
    ne = [a * b * c for a, b, c in dims]
    L = len(nn) - 1
    if L == 0:
        return 0

    def lds_values(s0):
        w = max([3 * nn[s0 - 1]] + [24 * ne[l] for l in range(s0, L)])
        return w + sum(ne[l] + 3 * nn[l] for l in range(s0, L)) + sum(9 * nn[l] for l in range(s0, L + 1))

    s0 = 1
    while s0 <= L and (nn[s0] > 512 or lds_values(s0) * word > 128 * 1024):
        s0 += 1
    return min(s0, L)


def ml_apply_bytes(nx, ny, nz, word):
    """Algorithmic HBM bytes of one multilevel preconditioner application:
    every restriction reads its level and writes the next; every
    prolongation reads r_l, D_l^-1 and y_{l+1} and writes y_l (z on the fine
    grid); the coarsest level reads D_L^-1 and writes y_L."""
    dims = level_dims(nx, ny, nz)
    n = [3 * (a + 1) * (b + 1) * (c + 1) for a, b, c in dims]
    L = len(n) - 1
    v = sum(n[l] + n[l + 1] for l in range(L)) + sum(3 * n[l] + n[l + 1] for l in range(L)) + 2 * n[L]
    return float(v * word)


def vc_apply_bytes(nx, ny, nz, word, smoothed):
    """Algorithmic HBM bytes of one V-cycle application: on each smoothed level
    the pre-smoothing operator reads r, S and the moduli and writes x and the
    residual, the restriction reads the residual and writes the next level's
    r, the prolongation reads x, S and the coarse correction and writes x, the
    post-smoothing operator reads x, r, S and the moduli and writes its output;
    the additive levels below read D_l^-1 and write y_l."""
    dims = level_dims(nx, ny, nz)
    n = [3 * (a + 1) * (b + 1) * (c + 1) for a, b, c in dims]
    ne = [a * b * c for a, b, c in dims]
    v = sum(12 * n[l] + 2 * ne[l] + 2 * n[l + 1] for l in range(smoothed))
    v += sum(2 * n[l] for l in range(smoothed, len(n)))
    return float(v * word)


# form -> (precision, preconditioner)
FORMS = {"plain": ("HX_FP64", "HX_PC_JACOBI"), "precision": ("HX_FP32_IR", "HX_PC_JACOBI"),
         "plain_ml": ("HX_FP64", "HX_PC_MULTILEVEL"), "precision_ml": ("HX_FP32_IR", "HX_PC_MULTILEVEL"),
         "plain_vc": ("HX_FP64", "HX_PC_VCYCLE"), "precision_vc": ("HX_FP32_IR", "HX_PC_VCYCLE")}
EXPLICIT_PC = ("_ml", "_vc")  # forms whose operator is the z-form (p = z + beta p_old)


def run_form(hx, L, ctx, nx, ny, nz, f, fixed, precision, steps, warmup, args, tune=None, precond=0):
    ctx.set_grid(nx, ny, nz, 1.0, 1.0, 1.0, 0.3)
    ctx.set_fixed(fixed)
    if tune is not None:  # the shapes the traffic passes measured
        ctx.tune_set(tune)
        log(f"  tuning (from the traffic pass): {tune}")
    elif not args.no_autotune:
        t = ctx.autotune(1)
        log(f"  autotune: {t}")
    ctx.to_init(to_params(L, args, precision, precond), f, np.full(nx * ny * nz, args.volfrac))
    for w in range(warmup):
        s = ctx.to_step()
        log(f"  warmup {w + 1}: c={s.compliance:.6e} pcg={s.pcg_iterations} t={s.seconds:.5f}s")
    # timed region: production launches (hipGraph batches, no per-kernel events)
    t0 = time.perf_counter()
    recs = []
    for k in range(steps):
        s = ctx.to_step()
        recs.append(s)
        log(f"  step {k + 1}: c={s.compliance:.6e} change={s.change:.4f} pcg={s.pcg_iterations} "
            f"rel={s.pcg_rel_res:.2e} filt={s.filter_iterations} vm_max={s.max_vm:.3e} t={s.seconds:.5f}s")
    wall = time.perf_counter() - t0  # to_step synchronises the stream at its end
    # operator kernel alone: back-to-back launches at the production shape
    op_ms = ctx.time_operator(precision, 1, 40, precond if precond in (L.HX_PC_MULTILEVEL, L.HX_PC_VCYCLE)
                              else L.HX_PC_JACOBI)
    # kernel breakdown: one more TO iteration with HIP events around every hot
    # launch (direct launches, so its wall time is not the reported value)
    ctx.timing(True)
    s = ctx.to_step()
    kern = {}
    names = {L.HX_K_MATVEC: "matvec_fp64", L.HX_K_MATVEC32: "matvec_fp32", L.HX_K_UPDATE: "pcg_update",
             L.HX_K_PUPDATE: "pcg_pupdate", L.HX_K_HELM: "helm_matvec", L.HX_K_ELEM: "element_post",
             L.HX_K_PRECOND: "ml_precond"}
    for kid, nm in names.items():
        ms, cnt = ctx.timing_read(kid)
        kern[nm] = {"avg_ms": ms, "count": cnt, "total_ms": ms * cnt}
    ctx.timing(False)
    kern["operator_back_to_back"] = {"avg_ms": op_ms, "count": 40, "total_ms": 40 * op_ms}
    if precond == L.HX_PC_VCYCLE:  # the fine level's smoothing operators, back to back, once per application
        napp = kern["ml_precond"]["count"]
        for form, nm in ((3, "vc_pre_smoothing"), (4, "vc_post_smoothing")):
            ms = ctx.time_smoothing(precision, form, 40)
            kern[nm] = {"avg_ms": ms, "count": napp, "total_ms": napp * ms}
    log(f"  breakdown step: pcg={s.pcg_iterations} t={s.seconds:.3f}s; operator kernel {op_ms:.4f} ms")
    return wall / steps, recs, kern, ctx.to_read(L.HX_FIELD_X)


def to_params(L, args, precision, precond):
    return L.TOParams(E0=1.0, Emin=1e-9, nu=0.3, penal=3.0, volfrac=args.volfrac, move=0.2, xmin=0.0,
                      rmin=args.rmin, pcg_tol=args.tol, pcg_max_iter=args.max_iter, precision=precision,
                      precond=precond, filter_tol=1e-10, filter_max_iter=2000, oc_rel_tol=1e-9, with_stress=1)


def per_result_error(L, ctx, nx, ny, nz, f, fixed, x, main, tune, args):
    """One TO iteration from the same design x in the main form and in the
    plain FP64 form (fresh sessions, the main form's launch shapes): the
    difference of each result -- compliance, physical densities after the
    update, von Mises stress -- is the main form's error against FP64."""
    out = {}
    for form in ("plain", main):
        prec, pc = (getattr(L, nm) for nm in FORMS[form])
        ctx.set_grid(nx, ny, nz, 1.0, 1.0, 1.0, 0.3)
        ctx.set_fixed(fixed)
        if tune is not None:
            ctx.tune_set(tune)
        ctx.to_init(to_params(L, args, prec, pc), f, x)
        s = ctx.to_step()
        out[form] = (s, ctx.to_read(L.HX_FIELD_XPHYS), ctx.to_read(L.HX_FIELD_VM), ctx.to_read(L.HX_FIELD_U))
        log(f"  per-result check, {form}: c={s.compliance:.9e} pcg={s.pcg_iterations} rel={s.pcg_rel_res:.2e}")
    (s0, x0, v0, u0), (s1, x1, v1, u1) = out["plain"], out[main]
    return {"compliance_rel": abs(s1.compliance - s0.compliance) / abs(s0.compliance),
            "xphys_max_abs": float(np.abs(x1 - x0).max()),
            "von_mises_max_rel": float(np.abs(v1 - v0).max() / np.abs(v0).max()),
            "displacement_max_rel": float(np.abs(u1 - u0).max() / np.abs(u0).max()),
            "volume_main": float(x1.mean()), "pcg_rel_res_main": s1.pcg_rel_res, "pcg_rel_res_plain": s0.pcg_rel_res,
            "from": "one TO iteration of each form from the main form's design after its timed steps"}


def full_loop(L, ctx, nx, ny, nz, f, fixed, form, n_iter, tune, args):
    """The whole TO loop of the configuration (n_iter iterations from uniform
    rho) in one form: wall time and solver record of every iteration."""
    prec, pc = (getattr(L, nm) for nm in FORMS[form])
    ctx.set_grid(nx, ny, nz, 1.0, 1.0, 1.0, 0.3)
    ctx.set_fixed(fixed)
    if tune is not None:
        ctx.tune_set(tune)
    ctx.to_init(to_params(L, args, prec, pc), f, np.full(nx * ny * nz, args.volfrac))
    t, its, ch, c = [], [], [], []
    for k in range(n_iter):
        t0 = time.perf_counter()
        s = ctx.to_step()
        t.append(time.perf_counter() - t0)
        its.append(s.pcg_iterations)
        ch.append(round(s.change, 5))
        c.append(s.compliance)
        if k % 10 == 9:
            log(f"  full loop {form}: iteration {k + 1}: c={s.compliance:.6e} change={s.change:.4f} "
                f"pcg={s.pcg_iterations} t={t[-1]:.4f}s")
    return {"form": form, "iterations": n_iter, "s_per_iter_mean": round(float(np.mean(t)), 6),
            "s_per_iter_first10": round(float(np.mean(t[:10])), 6), "s_per_iter_last10": round(float(np.mean(t[-10:])), 6),
            "s_per_iteration": [round(v, 5) for v in t], "pcg_iterations": its, "change": ch,
            "compliance_final": c[-1]}


COPY_BYTES = 1 << 30
SAMPLE_LAUNCHES = 40


def kernel_sample(args):
    """Child process run under ``rocprofv3 --pmc``: autotune on the workload's
    grid, a plain 1 GiB device copy (the counter calibration), then the
    dominant kernel of the given form SAMPLE_LAUNCHES times back to back at its
    production shape, on vector sets larger together than the last-level
    cache -- the launches whose average duration is the bench's
    avg_launch_ms, so counters and time come from one regime.  Prints the
    tuning result as JSON."""
    nx, ny, nz = (int(v) for v in args.grid.lower().split("x"))
    hx = importlib.import_module("2025_12_09_hexat_v069_amd")
    L = hx._lib
    ctx = hx.Context(0)
    fixed, f = cantilever(nx, ny, nz)
    ctx.set_grid(nx, ny, nz, 1.0, 1.0, 1.0, 0.3)
    ctx.set_fixed(fixed)
    if args.kernel_sample == "tune":
        print(json.dumps({"tune": ctx.autotune(1)}), flush=True)
        return
    tune = json.loads(args.tune_json)
    ctx.tune_set(tune)
    ctx.test_copy(COPY_BYTES, 4)
    Ee = np.full(nx * ny * nz, 1e-9 + args.volfrac ** 3)
    prec, pc = (getattr(L, nm) for nm in FORMS[args.kernel_sample])
    ctx.pcg(Ee, f, tol=1e-30, max_iter=1, precision=prec)  # moduli on the device (Jacobi: no smoothing launches)
    # the dominant kernel through the same back-to-back hook that gives the bench its launch time
    if args.kernel_sample.endswith("_vc"):
        ctx.time_smoothing(prec, 4, SAMPLE_LAUNCHES)
    else:
        ctx.time_operator(prec, 1, SAMPLE_LAUNCHES, pc)
    print(json.dumps({"copy_bytes": COPY_BYTES}), flush=True)


def _counter_rows(outdir):
    rows = []
    for path in glob.glob(os.path.join(outdir, "**", "*counter_collection.csv"), recursive=True):
        with open(path, newline="") as fh:
            rows.extend(csv.DictReader(fh))
    return rows


def measure_traffic(args, form, tune=None):
    """HBM bytes per launch of the dominant kernel from two rocprofv3 --pmc
    passes (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch) over a child
    process (kernel_sample), each scaled by what the same counter reads for a
    plain copy of known size in the same process.  Runs before this process
    touches the GPU.  Returns (traffic dict, tuning result or None)."""
    exe = shutil.which("rocprofv3")
    if exe is None:
        return {"error": "rocprofv3 not found"}, None
    base = [sys.executable, os.path.abspath(__file__), "--grid", args.grid, "--volfrac", str(args.volfrac)]
    if tune is None:
        try:  # tune once, unprofiled; the counter passes and this process reuse the result
            cp = subprocess.run(base + ["--kernel-sample", "tune"], capture_output=True, text=True, timeout=240)
            tune = json.loads([ln for ln in cp.stdout.splitlines() if ln.startswith("{")][-1])["tune"]
        except (subprocess.TimeoutExpired, IndexError, ValueError, KeyError):
            return {"error": "tuning child failed"}, None
    out = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="hx_pmc_")
        cmd = [exe, "--pmc", counter, "--output-format", "csv", "-d", d, "-o", "run", "--"] + base + [
            "--kernel-sample", form, "--tune-json", json.dumps(tune)]
        log(f"traffic: {counter} pass ...")
        try:
            cp = subprocess.run(cmd, capture_output=True, text=True, timeout=240)
        except subprocess.TimeoutExpired:
            shutil.rmtree(d, ignore_errors=True)
            return {"error": f"{counter} pass timed out"}, tune
        lines = [ln for ln in cp.stdout.splitlines() if ln.startswith("{")]
        if cp.returncode != 0 or not lines:
            shutil.rmtree(d, ignore_errors=True)
            return {"error": f"{counter} pass failed (rc {cp.returncode}): {cp.stderr[-300:]}"}, tune
        info = json.loads(lines[-1])
        rows = _counter_rows(d)
        shutil.rmtree(d, ignore_errors=True)

        def kname(r):
            return r.get("Kernel_Name") or r.get("kernel_name") or ""

        def val(r):
            return float(r.get("Counter_Value") or r.get("counter_value") or 0.0)

        copy = [val(r) for r in rows if "k_copy16" in kname(r)]
        fp32 = FORMS[form][0] == "HX_FP32_IR"
        variant = tune["variant32"] if fp32 else tune["variant64"]
        tname = "float" if fp32 else "double"
        # the explicit-preconditioner forms always run the parity contraction: the multilevel form's
        # dominant kernel is its z + beta p operator (template argument FUSE = 2), the V-cycle's the
        # fine level's post-smoothing operator (FUSE = 4, the largest grid)
        explicit = form.endswith(EXPLICIT_PC)
        fuse = "4" if form.endswith("_vc") else "2"
        key = f"k_matvec_sf<{tname}" if (variant == 1 or explicit) else f"k_matvec_zmarch<{tname}"
        cand = {}
        for r in rows:
            if key in kname(r):
                m = re.search(r"k_matvec_sf<\w+, \d+, \d+, (\d), \d>", kname(r))
                if explicit and (m is None or m.group(1) != fuse):
                    continue
                cand.setdefault((kname(r), r.get("Grid_Size") or r.get("grid_size")), []).append(val(r))
        if explicit and cand:  # the fine level: the largest grid among the kept launches
            top = max(int(float(g or 0)) for _, g in cand)
            cand = {k: v for k, v in cand.items() if int(float(k[1] or 0)) == top}
        if not copy or not cand:
            return {"error": f"{counter}: kernels not found in the counter output ({len(rows)} rows)"}, tune
        (name, _), vals = max(cand.items(), key=lambda kv: len(kv[1]))
        copy_kib = sum(copy[1:]) / max(len(copy) - 1, 1)  # the first launch is the warm-up
        scale = (info["copy_bytes"] / 1024.0) / copy_kib
        m = re.search(r"k_\w+<[^()]*>", name)
        out[counter] = {"kib_per_launch": sum(vals) / len(vals), "launches": len(vals), "copy_kib": copy_kib,
                        "scale": scale, "kernel": m.group(0) if m else name[:80]}
    f, w = out["FETCH_SIZE"], out["WRITE_SIZE"]
    bytes_ = 1024.0 * (f["kib_per_launch"] * f["scale"] + w["kib_per_launch"] * w["scale"])
    return {"bytes_per_launch": bytes_, "read_bytes": 1024.0 * f["kib_per_launch"] * f["scale"],
            "write_bytes": 1024.0 * w["kib_per_launch"] * w["scale"], "fetch_scale": round(f["scale"], 4),
            "write_scale": round(w["scale"], 4), "kernel": f["kernel"], "launches": f["launches"]}, tune


def cpu_baseline(nx, ny, nz, f, fixed, pcg_iters_per_step, filt_iters_per_step, args):
    """Oracle (FP64 C/OpenMP restatement of the reference CPU path) on a
    bounded sample of the same workload, extrapolated per TO iteration."""
    from oracle import oracle as ox
    g = ox.Grid(nx, ny, nz)
    Ke = ox.ke_hex(1, 1, 1, 0.3)
    rng = np.random.default_rng(0)
    xphys = rng.uniform(0.05, 1.0, g.ne)
    Ee = 1e-9 + xphys ** 3
    n_pcg, n_filt = args.cpu_pcg_iters, args.cpu_filter_iters
    t0 = time.perf_counter()
    ox.pcg(g, Ke, Ee, fixed, f, tol=1e-30, max_iter=n_pcg)
    t_pcg = (time.perf_counter() - t0) / n_pcg
    t0 = time.perf_counter()
    ox.helmholtz(g, args.rmin / (2 * math.sqrt(3)), 0, xphys, tol=1e-30, max_iter=n_filt, check_every=1)
    t_filt = (time.perf_counter() - t0) / n_filt
    u = rng.standard_normal(g.ndof)
    t0 = time.perf_counter()
    ox.element_energy(g, Ke, u)
    ox.stress(g, 0.3, 1.0, xphys, u)
    t_elem = time.perf_counter() - t0
    per_step = t_pcg * pcg_iters_per_step + t_filt * filt_iters_per_step + t_elem
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            model = next((ln.split(":", 1)[1].strip() for ln in fh if ln.startswith("model name")), None)
    except OSError:
        pass
    return {"value": per_step, "unit": "s/TO-iteration", "cores": cores, "kind": "port", "cpu_model": model,
            "nproc": os.cpu_count(), "extrapolated": True,
            "sample": (f"{n_pcg} FP64 Jacobi-PCG iterations ({t_pcg:.3f} s each), {n_filt} Helmholtz PCG iterations "
                       f"({t_filt:.4f} s each) and one sensitivity+stress pass ({t_elem:.3f} s) on the "
                       f"{nx}x{ny}x{nz} grid with the oracle, extrapolated with the GPU run's mean "
                       f"{pcg_iters_per_step:.0f} PCG and {filt_iters_per_step:.0f} filter iterations per TO "
                       f"iteration")}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--grid", default="256x128x128")
    ap.add_argument("--volfrac", type=float, default=0.15)
    ap.add_argument("--rmin", type=float, default=2.5)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--max-iter", type=int, default=100000)
    ap.add_argument("--forms", default="plain,precision_vc",
                    help="comma-separated: plain (FP64 Jacobi-PCG), precision (FP32 operator + FP64 refinement), "
                         "plain_ml / precision_ml (the same with the additive multilevel preconditioner), "
                         "plain_vc / precision_vc (with the V-cycle)")
    ap.add_argument("--main-form", default=None,
                    help="the form `value` reports (default: precision_vc when run, else precision_ml, else "
                         "precision, else the first form)")
    ap.add_argument("--full-loop", type=int, default=50,
                    help="after the timed steps, run the main form's whole loop of this many TO iterations from "
                         "the start and report every iteration's time (0: skip)")
    ap.add_argument("--no-error-check", action="store_true",
                    help="skip the per-result comparison of the main form against plain FP64")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--cpu-pcg-iters", type=int, default=20)
    ap.add_argument("--cpu-filter-iters", type=int, default=20)
    ap.add_argument("--learn-samples", type=int, default=20000,
                    help="cfg4: manufactured-solution pairs to train the learned preconditioner on (0: skip it)")
    ap.add_argument("--learn-grid", type=int, default=96, help="cfg4: edge of the cubic training grid")
    ap.add_argument("--workload", choices=("cantilever", "mbb", "block8", "cfg4"), default="cantilever",
                    help="cantilever: the headline TO-iteration bench; mbb / block8 / cfg4: solver-only "
                         "measurements of the precision (192x96x96), batching (8 load cases, 128x64x64) and "
                         "large-grid (320x160x160) configurations, each its own JSON line")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the rocprofv3 FETCH_SIZE / WRITE_SIZE passes (roofline.traffic = null)")
    ap.add_argument("--kernel-sample", choices=("tune",) + tuple(FORMS), default=None,
                    help=argparse.SUPPRESS)  # child modes of the traffic passes
    ap.add_argument("--tune-json", default=None,
                    help="apply this tuning result (the 'tune' object of an earlier run's JSON line) instead of "
                         "autotuning, e.g. to profile the same launch shapes")
    args = ap.parse_args()
    if args.kernel_sample:
        kernel_sample(args)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "one GPU only: --gpus N>1 is not supported (the path is designed to shard by "
                                   "x-slabs with a halo exchange of one node plane per neighbour)"}))
        sys.exit(2)
    if args.workload != "cantilever":
        hx = importlib.import_module("2025_12_09_hexat_v069_amd")
        ctx = hx.Context(0)
        other_workload(args, hx, hx._lib, ctx)
        return
    nx, ny, nz = (int(v) for v in args.grid.lower().split("x"))
    forms = [s.strip() for s in args.forms.split(",") if s.strip()]
    for fm in forms:
        if fm not in FORMS:
            sys.exit(f"unknown form {fm!r} (choose from {', '.join(FORMS)})")
    main_form = args.main_form or next((fm for fm in ("precision_vc", "precision_ml", "precision") if fm in forms),
                                       forms[0])
    if main_form not in forms:
        sys.exit(f"--main-form {main_form!r} is not among --forms")
    traffic, tune = (None, json.loads(args.tune_json) if args.tune_json else None)
    if not args.no_traffic:  # before this process initialises the GPU
        traffic, tune = measure_traffic(args, main_form, tune)
        log(f"traffic: {traffic}")
    hx = importlib.import_module("2025_12_09_hexat_v069_amd")
    L = hx._lib
    ctx = hx.Context(0)
    info = ctx.device_info()
    log(f"device: {info}")
    fixed, f = cantilever(nx, ny, nz)
    ne, ndof = nx * ny * nz, 3 * (nx + 1) * (ny + 1) * (nz + 1)
    res = {}
    for form in forms:
        prec, pc = (getattr(L, nm) for nm in FORMS[form])
        log(f"form {form}:")
        sps, recs, kern, xfinal = run_form(hx, L, ctx, nx, ny, nz, f, fixed, prec, args.steps, args.warmup, args,
                                           tune, precond=pc)
        res[form] = {"s_per_iter": sps, "recs": recs, "kern": kern, "x": xfinal}
    r = res[main_form]
    err = None
    if not args.no_error_check and main_form != "plain":
        log("per-result error of the main form against plain FP64:")
        err = per_result_error(L, ctx, nx, ny, nz, f, fixed, r["x"], main_form, tune, args)
        log(f"  {err}")
    loop = None
    if args.full_loop > 0:
        log(f"full loop of {args.full_loop} TO iterations, {main_form}:")
        loop = full_loop(L, ctx, nx, ny, nz, f, fixed, main_form, args.full_loop, tune, args)
    # roofline of the dominant kernel of the main form.  Algorithmic bytes per
    # launch: the PCG matvec reads r, D^-1, p_old (the p-update is fused into
    # its plane loads) and E, and writes q and p_new; the update reads x, p, r,
    # q, D^-1 and writes x, r.
    # (explicit-preconditioner forms: the operator reads p_old and z = M^-1 r instead of r and D^-1)
    ml_main = main_form.endswith(EXPLICIT_PC)
    mv_bytes64 = (32.0 if ml_main else 40.0) * ndof + 8.0 * ne
    mv_bytes32 = (16.0 if ml_main else 20.0) * ndof + 4.0 * ne
    fp64_main = FORMS[main_form][0] == "HX_FP64"
    upd_bytes = 56.0 * ndof if fp64_main else 28.0 * ndof
    pup_bytes = 32.0 * ndof if fp64_main else 16.0 * ndof
    # V-cycle: the fine level's smoothing operators (r, S and the moduli in; the post-smoothing form
    # also reads the coarse correction on the first coarsening's nodes; one output vector)
    word = 8.0 if fp64_main else 4.0
    ndof1 = 3 * (nx // 2 + 1) * (ny // 2 + 1) * (nz // 2 + 1)
    pre_bytes = 3 * word * ndof + word * ne
    post_bytes = 3 * word * ndof + word * ne + word * ndof1
    cand = []
    for nm, by in (("matvec_fp64", mv_bytes64), ("matvec_fp32", mv_bytes32), ("pcg_update", upd_bytes),
                   ("pcg_pupdate", pup_bytes), ("vc_pre_smoothing", pre_bytes), ("vc_post_smoothing", post_bytes)):
        k = r["kern"].get(nm)
        if k and k["count"]:
            cand.append((k["total_ms"], nm, by, k["avg_ms"]))
    cand.sort(reverse=True)
    roof = None
    if cand:
        _, nm, by, avg = cand[0]
        if nm.startswith("matvec"):  # the kernel alone, back to back, at the production shape
            avg = r["kern"]["operator_back_to_back"]["avg_ms"]
        ach = by / (avg * 1e-3) / 1e9
        tr = None
        measured = "vc_post_smoothing" if main_form.endswith("_vc") else ("matvec_fp64" if fp64_main else "matvec_fp32")
        if traffic and "bytes_per_launch" in traffic and nm == measured:
            tr = round(traffic["bytes_per_launch"])
        roof = {"bound": "hbm", "kernel": nm, "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": tr, "bytes_per_launch": by,
                "avg_launch_ms": avg, "traffic_detail": traffic}
    if ml_main and r["kern"].get("ml_precond", {}).get("count"):
        k = r["kern"]["ml_precond"]
        if main_form.endswith("_vc"):
            sl = smoothed_levels(nx, ny, nz, 8 if fp64_main else 4)
            by = vc_apply_bytes(nx, ny, nz, 8 if fp64_main else 4, sl)
            kn = ("k_matvec_sf forms 3, 4, 6 and 7, k_ml_restrict_zm, k_vc_coarse, k_fin_rz "
                  f"({sl} smoothed levels, W visits on the two above an exact coarsest level; one application; "
                  "bytes of a single visit per level)")
        else:
            by = ml_apply_bytes(nx, ny, nz, 8 if fp64_main else 4)
            kn = "k_ml_restrict_zm, k_ml_coarse, k_ml_prolong, k_fin_rz (one application)"
        roof_ml = {"kernels": kn,
                   "avg_ms": k["avg_ms"], "bytes_per_apply": by, "achieved": round(by / (k["avg_ms"] * 1e-3) / 1e9, 1),
                   "unit": "GB/s", "frac": round(by / (k["avg_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    else:
        roof_ml = None
    pcg_mean = float(np.mean([s.pcg_iterations for s in res.get("plain", r)["recs"]]))
    filt_mean = float(np.mean([s.filter_iterations for s in res.get("plain", r)["recs"]]))
    cpu = None
    if not args.no_cpu:
        log("cpu baseline (oracle) ...")
        cpu = cpu_baseline(nx, ny, nz, f, fixed, pcg_mean, filt_mean, args)
    out = {
        "metric": METRIC, "value": round(r["s_per_iter"], 6), "unit": "s/TO-iteration", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * r["s_per_iter"], 3),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if fp64_main else "f32+f64", "data": "synthetic",
        "config": {"workload": f"cantilever {nx}x{ny}x{nz} ({ne} elements, {ndof} DOFs), SIMP p=3, volfrac "
                               f"{args.volfrac}, PDE filter r={args.rmin}, PCG rel. residual {args.tol}, von Mises "
                               f"each iteration", "form": main_form,
                   "preconditioner": {"_ml": "multilevel", "_vc": "vcycle"}.get(main_form[-3:], "jacobi"),
                   "iterations_timed": f"TO iterations {args.warmup + 1}..{args.warmup + args.steps} from uniform "
                                       f"rho={args.volfrac}",
                   "contraction": "FP32 parity-basis sum factorisation on VALU (the FP32 MFMA contraction is built "
                                  "and tested, and slower)" if FORMS[main_form][0] == "HX_FP32_IR"
                                  else "FP64 parity-basis sum factorisation on VALU",
                   "tune": tune},
        "per_result_error_vs_fp64": err,
        "full_loop": loop,
        "roofline": roof,
        "preconditioner_apply": roof_ml,
        "cpu_baseline": cpu,
        "forms": {k: {"s_per_iter": round(v["s_per_iter"], 6),
                      "pcg_iterations": [s.pcg_iterations for s in v["recs"]],
                      "compliance": [s.compliance for s in v["recs"]],
                      "pcg_rel_res": [s.pcg_rel_res for s in v["recs"]],
                      "kernels": {n: {"avg_ms": round(x["avg_ms"], 5), "count": x["count"]}
                                  for n, x in v["kern"].items() if x["count"]}} for k, v in res.items()},
    }
    if "plain" in res and main_form != "plain":
        out["speedup_vs_plain_fp64"] = round(res["plain"]["s_per_iter"] / res[main_form]["s_per_iter"], 3)
        cp, cq = res["plain"]["recs"][-1].compliance, res[main_form]["recs"][-1].compliance
        out["compliance_rel_diff_vs_fp64"] = abs(cq - cp) / abs(cp)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
