"""V-cycle preconditioner (HX_PC_VCYCLE) through the C ABI against the CPU
oracle: one application z = M^-1 r (FP64 round-off; FP32 vectors to single
precision) with the hierarchy depth and the number of smoothed levels, PCG and
the SIMP loop to the solver tolerance, graph replay, and the property the
form exists for (fewer iterations than the additive one)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx(hx):
    c = hx.Context(0)
    yield c
    c.close()


def _setup(ox, ctx, dims, h=(1.0, 1.0, 1.0)):
    g = ox.Grid(*dims, *h)
    ctx.set_grid(*dims, *h, nu=0.3)
    X = g.node_coords()
    fixed = np.repeat((X[:, 0] == 0.0)[:, None], 3, axis=1).ravel().astype(np.uint8)
    fixed[3 * np.flatnonzero(X[:, 2] == 0.0) + 2] = 1  # a partially fixed face too
    ctx.set_fixed(fixed)
    f = ox.external_forces(X, [{"location": ["100%", "50%", "50%"], "F": [0, 0, -1]}])
    return g, fixed, f


# one smoothed level; anisotropic elements; two smoothed levels (level 1 beyond one
# workgroup); the coarsest level by its diagonal below a smoothed one (50x26x26: one
# coarsening, 5096 nodes); no coarse level at all (odd count: Jacobi)
APPLY_CASES = [((8, 4, 4), (1.0, 1.0, 1.0)), ((16, 8, 8), (1.0, 0.5, 2.0)), ((64, 32, 32), (1.0, 1.0, 1.0)),
               ((50, 26, 26), (1.0, 1.0, 1.0)), ((7, 4, 4), (1.0, 1.0, 1.0))]


# Coarse levels' launch width 1 (HEXAT_VC_MINBLOCKS, read when a context is
# created): every restriction then marches 8 coarse planes per workgroup,
# carrying plane 2K+1's xy sum to K+1 (at the default width only production
# grids do), and the level operators keep their full z chunks.
@pytest.mark.parametrize("dims", [(64, 32, 32), (50, 26, 26), (16, 8, 8)])
@pytest.mark.parametrize("precision", [0, 1])
def test_deep_restriction_chunks_match_oracle(ox, hx, monkeypatch, dims, precision):
    _say("deep restriction chunks", dims, precision)
    monkeypatch.setenv("HEXAT_VC_MINBLOCKS", "1")
    c = hx.Context(0)
    try:
        g, fixed, _ = _setup(ox, c, dims)
        rng = np.random.default_rng(7 * sum(dims) + precision)
        Ee = 1e-9 + rng.uniform(0, 1, g.ne) ** 3
        Ke = ox.ke_hex(1.0, 1.0, 1.0, 0.3)
        m0 = ox.ml_hierarchy(g, Ke, Ee, fixed, 0)[0][1].ravel()
        r = np.where(m0 != 0, rng.standard_normal(g.ndof), 0.0)
        tol = 1e-10 if precision == 0 else 1e-5
        z, L, _ = c.test_vc_apply(Ee, r, precision)
        zo = ox.vc_apply(g, Ke, Ee, fixed, r)
        assert L >= 1 and np.allclose(z, zo, rtol=tol, atol=tol * np.abs(zo).max())
        z, L = c.test_ml_apply(Ee, r, precision)
        zo = ox.ml_apply(g, Ke, Ee, fixed, r)
        assert L >= 1 and np.allclose(z, zo, rtol=tol, atol=tol * np.abs(zo).max())
    finally:
        c.close()


@pytest.mark.parametrize("dims,h", APPLY_CASES)
@pytest.mark.parametrize("precision", [0, 1])
def test_vc_apply_matches_oracle(ox, hx, ctx, dims, h, precision):
    _say("vc apply", dims, h, precision)
    g, fixed, _ = _setup(ox, ctx, dims, h)
    rng = np.random.default_rng(sum(dims) + precision)
    Ee = 1e-9 + rng.uniform(0, 1, g.ne) ** 3
    Ee[rng.random(g.ne) < 0.05] = 0.0
    Ke = ox.ke_hex(*h, 0.3)
    m0 = ox.ml_hierarchy(g, Ke, Ee, fixed, 0)[0][1].ravel()
    r = np.where(m0 != 0, rng.standard_normal(g.ndof), 0.0)
    z, L, c = ctx.test_vc_apply(Ee, r, precision)
    state = {}
    zo = ox.vc_apply(g, Ke, Ee, fixed, r, state=state)
    assert L == len(state["levels"]) - 1 and c == state["c"]
    tol = 1e-10 if precision == 0 else 1e-5
    assert np.allclose(z, zo, rtol=tol, atol=tol * np.abs(zo).max())


@pytest.mark.parametrize("dims", [(16, 8, 8), (64, 32, 32)])
@pytest.mark.parametrize("precision", [0, 1])
def test_vc_pcg_matches_oracle_with_fewer_iterations(ox, hx, ctx, dims, precision):
    _say("vc pcg", dims, precision)
    g, fixed, f = _setup(ox, ctx, dims)
    rng = np.random.default_rng(3)
    Ee = 1e-9 + rng.uniform(0.05, 1.0, g.ne) ** 3
    Ke = ox.ke_hex(1, 1, 1, 0.3)
    u, s = ctx.pcg(Ee, f, tol=1e-10, precision=precision, precond=hx._lib.HX_PC_VCYCLE)
    um, sm = ctx.pcg(Ee, f, tol=1e-10, precision=precision, precond=hx._lib.HX_PC_MULTILEVEL)
    _say("  iterations vc", s.iterations, "multilevel", sm.iterations, "rel", s.rel_res)
    assert s.status == 0 and s.rel_res < 1e-10
    assert s.iterations < sm.iterations
    if dims[0] <= 16:
        uo, rel_o, it_o = ox.vc_pcg(g, Ke, Ee, fixed, f, tol=1e-10)
        c, co = f @ u, f @ uo
        assert abs(c - co) <= (1e-9 if precision == 0 else 1e-7) * abs(co)
        if precision == 0:
            assert abs(s.iterations - it_o) <= 2  # same algorithm, FP64 both sides
    else:
        assert abs(f @ u - f @ um) <= 1e-8 * abs(f @ um)


def test_vc_graph_replay_matches_plain_launches(ox, hx, ctx, monkeypatch):
    g, fixed, f = _setup(ox, ctx, (16, 8, 8))
    Ee = 1e-9 + np.random.default_rng(12).uniform(0.05, 1.0, g.ne) ** 3
    VC = hx._lib.HX_PC_VCYCLE
    u1, s1 = ctx.pcg(Ee, f, tol=1e-10, batch=4, precond=VC)
    u3, s3 = ctx.pcg(Ee, f, tol=1e-10, batch=4, precision=1, precond=VC)
    monkeypatch.setenv("HEXAT_NO_GRAPH", "1")
    c2 = hx.Context(0)
    try:
        _setup(ox, c2, (16, 8, 8))
        u2, s2 = c2.pcg(Ee, f, tol=1e-10, batch=4, precond=VC)
        u4, s4 = c2.pcg(Ee, f, tol=1e-10, batch=4, precision=1, precond=VC)
    finally:
        c2.close()
    assert s1.iterations == s2.iterations and np.array_equal(u1, u2)
    assert s3.iterations == s4.iterations and np.array_equal(u3, u4)


@pytest.mark.parametrize("precision", [0, 1])
def test_vc_simp_loop_matches_oracle(ox, hx, ctx, precision):
    _say("vc simp loop", precision)
    g = ox.Grid(24, 8, 8)
    ctx.set_grid(24, 8, 8, nu=0.3)
    X = g.node_coords()
    fixed = np.repeat((X[:, 0] == 0.0)[:, None], 3, axis=1).ravel().astype(np.uint8)
    ctx.set_fixed(fixed)
    f = ox.external_forces(X, [{"location": ["100%", "50%", 0], "F": [0, 0, -1]}])
    kw = dict(nu=0.3, E0=1.0, Emin=1e-9, penal=3.0, volfrac=0.3, rmin=1.5, move=0.2, xmin=0.0, pcg_tol=1e-10,
              filter_tol=1e-12, oc_rel_tol=1e-10)
    hist, xo, xpo, uo = ox.simp_to_loop(g, f, fixed, 3, **kw)
    L = hx._lib
    p = L.TOParams(E0=1.0, Emin=1e-9, nu=0.3, penal=3.0, volfrac=0.3, move=0.2, xmin=0.0, rmin=1.5,
                   pcg_tol=1e-10, pcg_max_iter=200000, precision=precision, precond=L.HX_PC_VCYCLE,
                   filter_tol=1e-12, filter_max_iter=5000, oc_rel_tol=1e-10, with_stress=1)
    ctx.to_init(p, f, np.full(g.ne, 0.3))
    for h in hist:
        s = ctx.to_step()
        _say("  it", s.iter, "c", s.compliance, "oracle", h["compliance"], "pcg", s.pcg_iterations, h["pcg_iterations"])
        assert s.pcg_status == 0
        assert abs(s.compliance - h["compliance"]) <= 1e-6 * abs(h["compliance"])
        assert s.pcg_iterations < h["pcg_iterations"]
    assert np.allclose(ctx.to_read(L.HX_FIELD_XPHYS), xpo, rtol=1e-6, atol=1e-7)


# Block solves: the columns share the hierarchy and run every V-cycle launch
# with blockIdx.y = column (the one-workgroup pass as one workgroup per
# column); each column must take its single solve's iterations and bits.
# 16x8x8: levels in the one-workgroup pass only; 64x32x32: multi-block levels
# first (whose coarse z chunks follow the column count).
@pytest.mark.parametrize("dims", [(16, 8, 8), (64, 32, 32)])
@pytest.mark.parametrize("precision", [0, 1])
def test_vc_block_columns_equal_single_solves(ox, hx, ctx, dims, precision):
    _say("vc block", dims, precision)
    g, fixed, f = _setup(ox, ctx, dims)
    rng = np.random.default_rng(31)
    Ee = 1e-9 + rng.uniform(0.05, 1.0, g.ne) ** 3
    F = np.stack([f, -2.0 * f, np.where(fixed == 0, rng.standard_normal(g.ndof), 0.0)])
    VC = hx._lib.HX_PC_VCYCLE
    U, S = ctx.block_pcg(Ee, F, tol=1e-10, precision=precision, precond=VC)
    Ke = ox.ke_hex(1, 1, 1, 0.3)
    for c in range(F.shape[0]):
        u, s = ctx.pcg(Ee, F[c], tol=1e-10, precision=precision, precond=VC)
        r = np.where(fixed == 0, F[c] - ox.matvec(g, Ke, Ee, U[c]), 0.0)
        true_rel = np.linalg.norm(r) / np.linalg.norm(F[c])
        _say("  column", c, "its", S[c].iterations, s.iterations, "true rel", true_rel)
        assert S[c].status == 0 and S[c].iterations == s.iterations
        assert np.array_equal(U[c], u)
        assert true_rel < 1.05e-10


def test_block_solve_rejects_what_it_cannot_run(ox, hx, ctx):
    """V-cycle on a grid without a coarse level; additive multilevel and
    learned preconditioners with more than one column."""
    g, fixed, f = _setup(ox, ctx, (7, 4, 4))  # odd element count: no coarse level
    Ee = np.ones(g.ne)
    F = np.stack([f, f])
    L = hx._lib
    for pc in (L.HX_PC_VCYCLE, L.HX_PC_MULTILEVEL, L.HX_PC_LEARNED):
        o = L.PcgOpts(1e-8, 100, 0.0, 0, pc, 0, 0, 0)
        S = (L.PcgStats * 2)()
        U = np.zeros_like(F)
        rc = ctx.L.hx_block_pcg_solve(ctx.h, 2, L.dptr(Ee), L.dptr(F), L.dptr(U), L.C.byref(o), S)
        assert rc == L.HX_ERR_BAD_ARG


def test_vc_smoothing_timing_hook(ox, hx, ctx):
    """Back-to-back timing of the fine level's smoothing operators (both forms, both precisions);
    a form other than 3 or 4 and a grid without a coarsening are rejected."""
    _setup(ox, ctx, (32, 16, 16))
    for precision in (0, 1):
        for form in (3, 4):
            assert ctx.time_smoothing(precision, form, 4) > 0.0
    with pytest.raises(hx.HexatError):
        ctx.time_smoothing(1, 2, 4)
    _setup(ox, ctx, (7, 4, 4))
    with pytest.raises(hx.HexatError):
        ctx.time_smoothing(1, 4, 4)


@pytest.mark.parametrize("pre,post", [((2, 16, 1), (0, 16, 2)), ((0, 8, 2), (1, 8, 1)), ((1, 16, 2), (2, 4, 1)),
                                      ((2, 12, 1), (2, 12, 1))])
def test_vc_smoothing_shapes_match_oracle(ox, hx, ctx, pre, post):
    """Tuned launch shapes of the smoothing forms (restored by hx_tune_set) leave the application
    at the oracle's value; the autotuner reports shapes for both forms."""
    g, fixed, _ = _setup(ox, ctx, (64, 32, 32))
    t = {"tile": 2, "kchunk": 8, "pdepth": 1, "tile32": 2, "kchunk32": 8, "pdepth32": 1, "variant64": 1,
         "variant32": 1, "tile_pre": pre[0], "kchunk_pre": pre[1], "pdepth_pre": pre[2], "tile_post": post[0],
         "kchunk_post": post[1], "pdepth_post": post[2]}
    ctx.tune_set(t)
    rng = np.random.default_rng(5)
    Ee = 1e-9 + rng.uniform(0, 1, g.ne) ** 3
    Ke = ox.ke_hex(1, 1, 1, 0.3)
    r = np.where(fixed == 0, rng.standard_normal(g.ndof), 0.0)
    for precision, tol in ((0, 1e-10), (1, 1e-5)):
        z, _, _ = ctx.test_vc_apply(Ee, r, precision)
        zo = ox.vc_apply(g, Ke, Ee, fixed, r)
        assert np.allclose(z, zo, rtol=tol, atol=tol * np.abs(zo).max())
    tuned = ctx.autotune(1)
    assert tuned["kchunk_pre"] > 0 and tuned["kchunk_post"] > 0 and tuned["ms_post"] > 0.0
    with pytest.raises(hx.HexatError):
        ctx.tune_set(dict(t, kchunk_post=4, pdepth_post=3))
