"""Boundary conditions, loads and the regrid map on the device against the
oracle's restatements of Boundary.jl:56-291 and MeshRefiner.jl:131-150:
fixed masks and load vectors equal, bit for bit, including the even split
over matched nodes and the nearest-node fallback with its tie-break (the
reference log's mid-span load lies exactly between two node columns)."""
import numpy as np
import pytest
import yaml

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mods(hx):
    from importlib import import_module
    return {k: import_module(f"2025_12_09_hexat_v069_amd.{k}") for k in ("boundary", "mesh")}


BCS = [{"location": [0, ":", ":"], "DoFs": [1, 2, 3]}, {"location": ["100%", 0, ":"], "DoFs": [2]},
       {"location": [":", ":", 0], "DoFs": [3]}, {"location": [0.5, "25%", ":"], "DoFs": [1, 3]},
       {"location": [3.15, ":", "100%"], "DoFs": [2]}]
LOADS = [{"location": ["100%", 0, ":"], "F": [0, -1, 0]}, {"location": ["100%", ":", ":"], "F": [0.5, 0, 2]},
         {"location": ["50%", 0, "50%"], "F": [0, 15, 0]}, {"location": [2.0, 1.234, 0.777], "F": [1, 2, 3]},
         {"location": [":", "100%", 0.5], "F": [0, 0, -4]}, {"location": [40.0, 0, 0], "F": [1, 1, 1]}]


@pytest.mark.parametrize("dims,h", [((13, 7, 5), (0.7, 1.3, 0.45)), ((97, 32, 32), (60 / 97, 0.625, 0.625))])
def test_bc_and_loads_match_oracle(ox, hx, mods, dims, h):
    nodes, elements, _ = mods["mesh"].generate_mesh(*dims, *(np.float32(v) for v in h))
    ctx = hx.Context(0)
    try:
        ctx.set_grid(*dims, *(float(np.float32(v)) for v in h), nu=0.3)
        ind = mods["boundary"].get_bc_indicator_device(ctx, nodes, BCS)
        assert np.array_equal(ind, ox.bc_indicator(nodes, BCS).astype(np.float32))
        F = mods["boundary"].apply_external_forces_device(ctx, np.zeros(3 * nodes.shape[0]), LOADS, nodes)
        Fo = ox.external_forces(nodes, LOADS)
        assert np.array_equal(F, Fo)
        if dims == (97, 32, 32):  # the log's loads: mid-span point on the larger-x node column
            cfg = yaml.safe_load(open(__import__("os").path.join(__import__("os").path.dirname(__file__), "golden",
                                                                  "cantilever_4loads.yaml")))
            F4 = mods["boundary"].apply_external_forces_device(ctx, np.zeros(3 * nodes.shape[0]),
                                                               cfg["external_forces"], nodes)
            assert np.array_equal(F4, ox.external_forces(nodes, cfg["external_forces"]))
            mid = [d for d in np.nonzero(F4)[0] if abs(F4[d] - 15.0) < 1e-12][0] // 3
            assert nodes[mid, 0] == np.float32(49) * np.float32(60 / 97)
    finally:
        ctx.close()


def test_regrid_matches_oracle(ox, hx):
    ctx = hx.Context(0)
    try:
        rng = np.random.default_rng(3)
        for od, nd, L in (((97, 32, 32), (100, 33, 33), (60.0, 20.0, 20.0)), ((12, 7, 5), (19, 11, 8), (3.3, 1.7, 1.1)),
                          ((8, 8, 8), (5, 6, 7), (1.0, 1.0, 1.0))):
            f = rng.uniform(0, 1, (2, int(np.prod(od)))).astype(np.float32)
            g = ctx.regrid(od, nd, L, f)
            for q in range(2):
                assert np.array_equal(g[q], ox.regrid_map(od, nd, L, f[q]))
    finally:
        ctx.close()
