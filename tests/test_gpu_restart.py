"""Checkpoint export and restart through the driver (Main.jl:198-261,
433-466): a run that stops after iteration 2 and restarts from its
iteration-2 file continues exactly like an uninterrupted run (the file holds
the whole density field, the radius and the threshold; only the warm start is
lost, and both solves run to 1e-10)."""
import sys

import numpy as np
import pytest
import yaml

pytestmark = pytest.mark.gpu


def _cfg(tmp=None):
    cfg = yaml.safe_load(open(__import__("os").path.join(__import__("os").path.dirname(__file__), "golden",
                                                          "cantilever_4loads.yaml")))
    cfg["geometry"]["target_elem_count"] = 12000
    cfg["solver_parameters"].update(mode="northstar", tolerance=1e-10)
    cfg["output_settings"] = {"save_bin_frequency": 1, "save_VTK_frequency": 2, "save_STL_frequency": 2,
                              "iso_surface_threshold": 0.3}
    return cfg


def test_restart_continues_the_run(hx, tmp_path):
    from importlib import import_module
    drv = import_module("2025_12_09_hexat_v069_amd.driver")
    conf = import_module("2025_12_09_hexat_v069_amd.configuration")
    pp = import_module("2025_12_09_hexat_v069_amd.postprocessing")
    full, _, _ = drv.run_stress_ratio(conf.apply_hardware_profile(_cfg()), n_iter=3, verbose=False,
                                      keep_density=True, results_dir=str(tmp_path / "a"), base_name="c")
    f2 = full[1]["bin"]
    import os
    assert os.path.getsize(full[1]["vtk"]) > 0 and os.path.getsize(full[1]["stl"]) > 84  # iteration 2 of 3
    assert "vtk" not in full[0] and "stl" not in full[2]
    d = pp.read_bin(f2)
    assert d["iteration"] == 2 and d["full_density"] is not None
    assert np.array_equal(d["full_density"], full[1]["density"])
    assert d["density"].size == int((full[1]["density"] >= np.float32(0.3)).sum()) < d["full_density"].size
    cfg = conf.apply_hardware_profile(_cfg())
    cfg["restart_configuration"] = {"enable_restart": True, "file_path": f2}
    rest, dens, _ = drv.run_stress_ratio(cfg, n_iter=3, verbose=False, keep_density=True,
                                         results_dir=str(tmp_path / "b"), base_name="c")
    print("restart", [r["iter"] for r in rest], rest[0]["compliance"], full[2]["compliance"], flush=True)
    sys.stdout.flush()
    assert [r["iter"] for r in rest] == [3]
    assert abs(rest[0]["compliance"] / full[2]["compliance"] - 1) < 1e-7
    assert rest[0]["radius"] == pytest.approx(full[2]["radius"]) and rest[0]["cutoff"] == full[2]["cutoff"]
    assert np.abs(rest[0]["density"] - full[2]["density"]).max() < 1e-6
