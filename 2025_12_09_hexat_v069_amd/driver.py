"""YAML-driven TO loop (mirror of Main._run_safe, Main.jl:175-476, and the
iteration log of Diagnostics.jl:11-83).

    python -m <package>.driver config.yaml [--iters N]

``optimization_parameters.method: stress_ratio`` (default) runs the
reference's loop: solve -> compliance f'u -> stress field -> stress-ratio
update, with the viewer/checkpoint file, the VTK solution and the isosurface
STL at output_settings' frequencies and restarts from a checkpoint.
``simp_oc`` runs the device-resident SIMP/OC loop.  Mesh refinement and
gravity/thermal loads are outside this package's hot path (a config that asks
for gravity or thermal load is rejected rather than silently ignored).
"""
from __future__ import annotations

import argparse
import datetime
import os
import time

import numpy as np

from . import boundary, configuration, export_stl, export_vtk, mesh, postprocessing, stress, topopt
from .solver import solve_system

LOG_HEADER = ("| Iter | Mesh Size | Total El | Active El | Radius | Cutoff | Compliance | Strain Energy | "
              "Avg L1 Stress | Vol Frac | Delta Rho | Refine? | Lin Res  | Prec   | Time (s) | Wall Time | VRAM |\n"
              "|------|-----------|----------|-----------|--------|--------|------------|---------------|"
              "---------------|----------|-----------|---------|----------|--------|----------|-----------|------|")


def format_log_line(it, dims, n_total, n_active, R, thr, comp, se, avg_l1, vf, drho, status, t, res, prec,
                    wall=None, vram=None):
    """Diagnostics.jl:52-83 line format (wall: HH:MM:SS, now by default; vram:
    device memory in use, "%.1fG")."""
    wall = wall if wall is not None else datetime.datetime.now().strftime("%H:%M:%S")
    vram = vram if vram is not None else _vram_str()
    return ("| %4d | %9s | %8d | %9d | %6.3f | %6.3f | %10.3e | %13.3e | %13.3e | %8.4f | %8.2f%% | %7s | %8.1e "
            "| %-6s | %8.2f | %9s | %4s |" % (it, dims, n_total, n_active, R, thr, comp, se, avg_l1, vf, drho * 100,
                                              status, res, prec, t, wall, vram))


def _vram_str():
    """Diagnostics.jl:32-39: device memory in use."""
    try:
        from .solver import get_context
        used = get_context().memory_used_gb()
        return "%.1fG" % used
    except Exception:
        return "CPU"


def init_log_file(path, config):
    """Diagnostics.jl:41-49"""
    with open(path, "w") as fh:
        fh.write("HEXA FEM TOPOLOGY OPTIMIZATION LOG\n")
        fh.write("Start Date: %s\n" % datetime.datetime.now().isoformat(timespec="milliseconds"))
        fh.write("Config Geometry: %s\n" % (config.get("geometry"),))
        fh.write("=" * 200 + "\n")
        fh.write(LOG_HEADER + "\n")


def build_problem(config):
    geo = configuration.setup_geometry(config)
    nodes, elements, dims = mesh.generate_mesh(geo["nElem_x"], geo["nElem_y"], geo["nElem_z"], geo["dx"], geo["dy"],
                                               geo["dz"])
    g = config["geometry"]
    g.update(nElem_x_computed=geo["nElem_x"], nElem_y_computed=geo["nElem_y"], nElem_z_computed=geo["nElem_z"],
             dx_computed=float(geo["dx"]), dy_computed=float(geo["dy"]), dz_computed=float(geo["dz"]),
             max_domain_dim=float(geo["max_domain_dim"]))
    # boundary conditions and loads mapped on the device (the structured grid of the run)
    from .solver import get_context
    ctx = get_context()
    key = (geo["nElem_x"], geo["nElem_y"], geo["nElem_z"], float(geo["dx"]), float(geo["dy"]), float(geo["dz"]))
    if ctx.grid is None or tuple(ctx.grid[:6]) != key:
        ctx.set_grid(*key, nu=float(config["material"]["nu"]))
    bc = boundary.get_bc_indicator_device(ctx, nodes, list(config["boundary_conditions"]))
    F = boundary.apply_external_forces_device(ctx, np.zeros(3 * nodes.shape[0], dtype=np.float64),
                                              list(config["external_forces"]), nodes)
    mat = config["material"]
    if float(mat.get("gravity_acceleration", 0.0)) != 0.0 and float(mat.get("material_density", 0.0)) > 1e-9:
        raise NotImplementedError("self-weight loads are not supported by libhexat")
    if abs(float(mat.get("delta_temperature", 0.0))) > 1e-6:
        raise NotImplementedError("thermal loads are not supported by libhexat")
    return geo, nodes, elements, bc, F


def _export(results_dir, base, it, nodes, elements, geo, density, l1, principal, R, config, iso, what, U=None, F=None,
            bc=None, vm=None):
    """Main.jl:433-466, Postprocessing.export_iteration_results (:264-312):
    the viewer/checkpoint file, the VTK solution and the isosurface STL of
    this iteration, as `what` (a set of "bin", "vtk", "stl") asks."""
    os.makedirs(results_dir, exist_ok=True)
    pre = os.path.join(results_dir, f"iter_{it}_{base}")
    out = {}
    if "bin" in what:
        out["bin"] = pre + "_webdata.bin"
        postprocessing.export_binary_for_web(out["bin"], nodes, elements, density, l1, principal, geo,
                                             np.float32(iso), it, np.float32(R), config)
    if "vtk" in what:
        out["vtk"] = export_vtk.export_solution(nodes, elements, U, F, bc, principal, vm, None, l1, None,
                                                density=density, threshold=np.float32(iso),
                                                filename=pre + "_solution")
    if "stl" in what:
        out["stl"] = pre + "_isosurface.stl"
        export_stl.export_smooth_watertight_stl(density, geo, np.float32(iso), out["stl"], subdivision_level=1)
    return out


def _restart(config, n_elements):
    """Main.jl:198-227, 259-261: (density, start_iter, radius, threshold) of
    an enabled restart whose file exists, else None; the configuration stored
    in the file is merged under the current one."""
    rc = config.get("restart_configuration", {}) or {}
    path = rc.get("file_path", "")
    if not rc.get("enable_restart", False):
        return None
    if not os.path.isfile(path):
        print(f"!!! WARNING: Restart requested but file not found: '{path}'\n!!! Falling back to FRESH START.")
        return None
    saved, rho, it, R, thr = postprocessing.load_checkpoint(path)
    merged = dict(saved)
    merged.update(config)
    config.clear()
    config.update(merged)
    return postprocessing.restore_density(rho, n_elements), it + 1, R, thr


def run_stress_ratio(config, n_iter=None, log_path=None, verbose=True, keep_density=False, results_dir=None,
                     base_name="run"):
    """Main.jl:331-474 without refinement.  keep_density: each iteration's
    record also holds a copy of the density after its update.  results_dir:
    write the viewer/checkpoint files at output_settings.save_bin_frequency
    (export_frequency, default 5; iteration 0 at the start).  A restart
    (restart_configuration.enable_restart / file_path) continues from the
    file's iteration + 1 with its density, radius and threshold."""
    geo, nodes, elements, bc, F = build_problem(config)
    restart = _restart(config, elements.shape[0])
    E = np.float32(config["material"]["E"])
    nu = np.float32(config["material"]["nu"])
    opt = config["optimization_parameters"]
    min_density = np.float32(opt.get("min_density", 1e-3))
    max_density = np.float32(opt.get("density_clamp_max", 1.0))
    density, original, protected, _ = configuration.initialize_density_field(nodes, elements, geo["shapes"], config)
    start, R, thr = 1, 0.0, 0.0
    if restart is not None:
        density, start, R, thr = restart
        density = density.copy()
    allow = np.float32(config.get("l1_stress_allowable", 1.0)) or np.float32(1.0)
    nominal = int(config.get("number_of_iterations", 30))
    n_iter = n_iter or nominal
    out = config.get("output_settings", {}) or {}
    dfreq = int(out.get("export_frequency", 5))  # Main.jl:235-238
    freq = {"bin": int(out.get("save_bin_frequency", dfreq)), "stl": int(out.get("save_STL_frequency", dfreq)),
            "vtk": int(out.get("save_VTK_frequency", dfreq))}
    iso = float(out.get("iso_surface_threshold", 0.8))
    U = np.zeros(3 * nodes.shape[0])
    dims = f"{geo['nElem_x']}x{geo['nElem_y']}x{geo['nElem_z']}"
    hist = []
    if log_path and restart is None:
        init_log_file(log_path, config)
    for it in range(start, n_iter + 1):
        t0 = time.perf_counter()
        config["current_outer_iter"] = it
        U, res, prec = solve_system(nodes, elements, E, nu, bc, F, density=density, config=config,
                                    min_stiffness_threshold=min_density, prune_voids=True, u_prev=U)
        comp = float(F @ U)
        principal, vm, _, l1, _ = stress.compute_stress_field(nodes, elements, U, E, nu, density)
        l1 = l1.astype(np.float32)
        if results_dir and it == 1:  # the initial state as iteration 0 (Main.jl:433-439)
            what = {k for k, v in freq.items() if v > 0}
            if what:
                _export(results_dir, base_name, 0, nodes, elements, geo, density, l1, principal, R, config, iso, what,
                        U, F, bc, vm)
        act = density > np.float32(0.1)
        avg_l1 = float(l1[act].mean()) if act.any() else 0.0
        vf = float(density.astype(np.float64).sum() / density.size)
        n_active = int((density > min_density).sum())
        mc, R, thr, *_ = topopt.update_density(density, l1, protected, E, allow, it, nominal + 10, original,
                                               min_density, max_density, config, elements, False)
        dt = time.perf_counter() - t0
        rec = dict(iter=it, compliance=comp, strain_energy=0.5 * comp, avg_l1=avg_l1, vol_frac=vf, delta_rho=mc,
                   radius=R, cutoff=thr, lin_res=res, prec=prec, time=dt, active=n_active)
        if keep_density:
            rec["density"] = density.copy()
        hist.append(rec)
        line = format_log_line(it, dims, density.size, n_active, R, thr, comp, 0.5 * comp, avg_l1, vf, mc,
                               "Nominal", dt, res, prec)
        if verbose:
            if it == 1 or it % 10 == 0:  # Diagnostics.jl:77-79
                print(LOG_HEADER)
            print(line, flush=True)
        if log_path:
            with open(log_path, "a") as fh:
                fh.write(line + "\n")
        what = {k for k, v in freq.items() if v > 0 and it % v == 0}
        if results_dir and what:
            rec.update(_export(results_dir, base_name, it, nodes, elements, geo, density, l1, principal, R, config,
                               iso, what, U, F, bc, vm))
    return hist, density, U


def run_simp(config, n_iter=None, verbose=True):
    geo, nodes, elements, bc, F = build_problem(config)
    opt = config["optimization_parameters"]
    sp = config.get("solver_parameters", {})
    fixed = boundary.fixed_dof_mask(bc)
    o = topopt.SimpOptimizer(geo["nElem_x"], geo["nElem_y"], geo["nElem_z"], F, fixed, dx=float(geo["dx"]),
                             dy=float(geo["dy"]), dz=float(geo["dz"]), nu=float(config["material"]["nu"]),
                             E0=float(config["material"]["E"]),
                             Emin=float(opt.get("E_min", 1e-9)) * float(config["material"]["E"]),
                             penal=float(opt.get("penal", 3.0)), volfrac=float(opt.get("volfrac", 0.3)),
                             rmin=float(opt.get("filter_radius", 1.5)), move=float(opt.get("move", 0.2)),
                             pcg_tol=float(sp.get("tolerance", 1e-8)),
                             precision=sp.get("precision", "fp64"), precond=sp.get("preconditioner", "jacobi"))
    n_iter = n_iter or int(config.get("number_of_iterations", 20))
    for _ in range(n_iter):
        r = o.step()
        if verbose:
            print("iter %3d  c=%.6e  change=%.4f  vol=%.4f  pcg=%d (%.1e)  vm_max=%.3e  %.3fs" % (
                r["iter"], r["compliance"], r["change"], r["volume"], r["pcg_iterations"], r["pcg_rel_res"],
                r["max_vm"], r["device_s"]), flush=True)
    return o.history, o


def run_main(config_file, n_iter=None, log_path=None, verbose=True):
    config = configuration.apply_hardware_profile(configuration.load_configuration(config_file))
    method = str(config.get("optimization_parameters", {}).get("method", "stress_ratio")).lower()
    if method == "simp_oc":
        return run_simp(config, n_iter, verbose)
    return run_stress_ratio(config, n_iter, log_path, verbose)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("config")
    ap.add_argument("--iters", type=int, default=None)
    ap.add_argument("--log", default=None)
    a = ap.parse_args(argv)
    run_main(a.config, a.iters, a.log)


if __name__ == "__main__":
    main()
