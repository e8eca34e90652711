"""hx_web_pack (the viewer/checkpoint payload) against the oracle, bit for bit,
through the C ABI; and export_binary_for_web end to end."""
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
pp = importlib.import_module("2025_12_09_hexat_v069_amd.postprocessing")


@pytest.fixture(scope="module")
def ctx(hx):
    c = hx.Context(0)
    yield c
    c.close()


def _case(rng, ne, thr):
    rho = rng.uniform(0, 1, ne).astype(np.float32)
    rho[rng.integers(0, ne, max(1, ne // 50))] = np.float32(thr)  # equality is kept
    rho[rng.integers(0, ne, max(1, ne // 97))] = np.nan           # never kept
    l1 = rng.uniform(0, 10, ne).astype(np.float32)
    pr = rng.standard_normal((ne, 3)).astype(np.float32)
    t = rng.integers(0, ne, max(1, ne // 31))
    pr[t, 1] = -pr[t, 0]                                         # magnitude ties
    pr[rng.integers(0, ne, max(1, ne // 53))] = 0.0              # sign 0
    return rho, l1, pr


def _same(a, b):
    return a.shape == b.shape and np.array_equal(a, b, equal_nan=True)


@pytest.mark.parametrize("grid", [(1, 1, 1, 1.0, 1.0, 1.0), (3, 2, 1, 0.5, 1.0, 2.0), (13, 7, 5, 0.37, 1.3, 0.071),
                                  (64, 33, 17, 1.0, 1.0, 1.0), (60, 20, 20, 1.0, 1.0, 1.0)])
@pytest.mark.parametrize("thr", [0.5, 0.0, 0.8, 2.0])
def test_web_pack_matches_oracle(ctx, ox, grid, thr):
    nx, ny, nz, dx, dy, dz = grid
    print(f"web_pack grid={grid} thr={thr}", flush=True)
    ctx.set_grid(nx, ny, nz, dx, dy, dz, 0.3)
    rho, l1, pr = _case(np.random.default_rng(nx * 1000 + ny), nx * ny * nz, thr)
    c, d, s = ctx.web_pack(rho, l1, pr, thr)
    co, do, so = ox.web_pack(nx, ny, nz, dx, dy, dz, rho, l1, pr, thr)
    assert _same(c, co) and _same(d, do) and _same(s, so)


def test_web_pack_full_size(ctx, ox):
    """256x128x128 (4,194,304 elements, 2048 tiles): the production size."""
    nx, ny, nz = 256, 128, 128
    print("web_pack grid=256x128x128", flush=True)
    ctx.set_grid(nx, ny, nz, 1.0, 1.0, 1.0, 0.3)
    rho, l1, pr = _case(np.random.default_rng(7), nx * ny * nz, 0.3)
    for _ in range(3):  # repeated calls give the same bytes
        c, d, s = ctx.web_pack(rho, l1, pr, 0.3)
    co, do, so = ox.web_pack(nx, ny, nz, 1.0, 1.0, 1.0, rho, l1, pr, 0.3)
    assert d.size == np.count_nonzero(rho >= np.float32(0.3))
    assert _same(c, co) and _same(d, do) and _same(s, so)


def test_export_binary_for_web_end_to_end(hx, ox, tmp_path):
    mesh = importlib.import_module("2025_12_09_hexat_v069_amd.mesh")
    stress = importlib.import_module("2025_12_09_hexat_v069_amd.stress")
    nx, ny, nz = 12, 4, 4
    print("export_binary_for_web 12x4x4", flush=True)
    nodes, elements, _ = mesh.generate_mesh(nx, ny, nz)
    rng = np.random.default_rng(3)
    rho = rng.uniform(0.01, 1, nx * ny * nz).astype(np.float32)
    U = rng.standard_normal(nodes.shape[0] * 3) * 1e-2
    pr, vm, voigt, l1, _ = stress.compute_stress_field(nodes, elements, U, 1.0, 0.3, rho)
    cfg = {"geometry": {"hole": {"type": "sphere", "stiffness_ratio": 0}}, "external_forces": [], "boundary_conditions": []}
    fn = tmp_path / "it.bin"
    n = pp.export_binary_for_web(fn, nodes, elements, rho, l1, pr, {"dx": 1.0, "dy": 1.0, "dz": 1.0}, 0.4, 9, 1.5,
                                 cfg)
    co, do, so = ox.web_pack(nx, ny, nz, 1, 1, 1, rho, np.float32(l1), np.float32(pr.T), 0.4)
    d = pp.read_bin(fn)
    assert n == do.size == d["density"].size
    assert _same(d["centroids"], co) and _same(d["density"], do) and _same(d["signed_l1"], so)
    assert d["iteration"] == 9 and d["config"]["geometry"]["hole"]["action"] == "remove"
    assert pp.export_binary_for_web(tmp_path / "none.bin", nodes, elements, rho, l1, pr,
                                    {"dx": 1.0, "dy": 1.0, "dz": 1.0}, 5.0, 9, 1.5, cfg) == 0
    assert not (tmp_path / "none.bin").exists()
