"""Parity of the HIP path (through the C ABI) against the CPU oracle.

Sizes are ones the oracle finishes in seconds; the bar is FP64 round-off for
the operators (rtol 1e-12) and the solver tolerance for iterative results.
"""
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


def _grid_setup(ox, ctx, dims, h=(1.0, 1.0, 1.0), nu=0.3):
    g = ox.Grid(*dims, *h)
    ctx.set_grid(*dims, *h, nu=nu)
    return g


def _cantilever(ox, g):
    X = g.node_coords()
    fixed = np.repeat((X[:, 0] == 0.0)[:, None], 3, axis=1).ravel().astype(np.uint8)
    f = ox.external_forces(X, [{"location": ["100%", 0, ":"], "F": [0, -1, 0]}])
    return fixed, f


MV_CASES = [((1, 1, 1), (1.0, 1.0, 1.0)), ((3, 2, 1), (1.0, 0.5, 2.0)), ((5, 4, 3), (0.7, 1.3, 0.9)),
            ((17, 16, 18), (1.0, 1.0, 1.0)), ((60, 20, 20), (1.0, 1.0, 1.0)), ((31, 9, 40), (1.0, 2.0, 0.5))]


@pytest.mark.parametrize("dims,h", MV_CASES)
@pytest.mark.parametrize("variant", [1, 2, 3])
def test_matvec_matches_oracle(ox, hx, ctx, dims, h, variant):
    _say("matvec", dims, h, variant)
    g = _grid_setup(ox, ctx, dims, h)
    rng = np.random.default_rng(hash((dims, variant)) % 2**32)
    Ee = rng.uniform(0.01, 2.0, g.ne)
    Ee[rng.random(g.ne) < 0.1] = 0.0
    x = rng.standard_normal(g.ndof)
    y = ctx.matvec(Ee, x, variant)
    yo = ox.matvec(g, ox.ke_hex(*h, 0.3), Ee, x)
    assert np.allclose(y, yo, rtol=1e-12, atol=1e-12 * np.abs(yo).max())


def test_diagonal_matches_oracle(ox, ctx):
    g = _grid_setup(ox, ctx, (7, 5, 4), (1.0, 0.8, 1.2))
    Ee = np.random.default_rng(1).uniform(0, 1, g.ne)
    assert np.allclose(ctx.diagonal(Ee), ox.diag(g, ox.ke_hex(1.0, 0.8, 1.2, 0.3), Ee), rtol=1e-13)


def _true_rel(ox, g, Ke, Ee, fixed, f, u, shift=0.0):
    r = f - ox.matvec(g, Ke, Ee, u) - shift * u
    r[fixed.astype(bool)] = 0
    b = f.copy()
    b[fixed.astype(bool)] = 0
    return np.linalg.norm(r) / np.linalg.norm(b)


@pytest.mark.parametrize("dims", [(6, 3, 2), (24, 8, 8), (60, 20, 20)])
@pytest.mark.parametrize("precision", [0, 1])
def test_pcg_matches_oracle(ox, hx, ctx, dims, precision):
    _say("pcg", dims, precision)
    g = _grid_setup(ox, ctx, dims)
    fixed, f = _cantilever(ox, g)
    ctx.set_fixed(fixed)
    rng = np.random.default_rng(7)
    xphys = np.clip(rng.uniform(0.0, 1.0, g.ne), 0.01, 1.0)
    Ee = 1e-9 + xphys ** 3
    Ke = ox.ke_hex(1, 1, 1, 0.3)
    tol = 1e-8
    u, s = ctx.pcg(Ee, f, tol=tol, precision=precision)
    uo, st, rel, it = ox.pcg(g, Ke, Ee, fixed, f, tol=tol)
    _say("  gpu it", s.iterations, "rel", s.rel_res, "refinements", s.refinements, "oracle it", it)
    assert s.status == 0 and st == 0
    assert s.rel_res < tol
    assert _true_rel(ox, g, Ke, Ee, fixed, f, u) < 1.05 * tol
    assert not u[fixed.astype(bool)].any()
    c, co = f @ u, f @ uo
    assert abs(c - co) <= 1e-6 * abs(co)
    if precision == 0:
        # same algorithm in FP64: identical iteration count up to round-off
        assert abs(s.iterations - it) <= max(2, it // 50)


def test_pcg_warm_start_shift_and_zero_rhs(ox, hx, ctx):
    g = _grid_setup(ox, ctx, (8, 4, 4))
    fixed, f = _cantilever(ox, g)
    ctx.set_fixed(fixed)
    Ee = np.ones(g.ne)
    Ke = ox.ke_hex(1, 1, 1, 0.3)
    s_ = 1e-3
    u, s = ctx.pcg(Ee, f, tol=1e-12, shift=s_)
    uo, *_ = ox.pcg(g, Ke, Ee, fixed, f, shift=s_, tol=1e-12)
    assert np.allclose(u, uo, rtol=1e-8, atol=1e-10 * np.abs(uo).max())
    u2, s2 = ctx.pcg(Ee, f, u0=u, tol=1e-6, shift=s_)
    assert s2.iterations <= 1
    u3, s3 = ctx.pcg(Ee, np.zeros(g.ndof), tol=1e-8)
    assert s3.status == hx._lib.HX_ZERO_RHS and not u3.any()


def test_pcg_not_converged_returns_best_iterate(ox, hx, ctx):
    """GPUSolver.jl:441, 492-499, 603-606: a solve that stops at its iteration
    cap returns its best iterate among those with a residual below 1e-2, the
    start vector if there is none."""
    g = _grid_setup(ox, ctx, (12, 4, 4))
    fixed, f = _cantilever(ox, g)
    ctx.set_fixed(fixed)
    Ke = ox.ke_hex(1, 1, 1, 0.3)
    u, s = ctx.pcg(np.ones(g.ne), f, tol=1e-14, max_iter=5)
    assert s.status == hx._lib.HX_ERR_NOT_CONVERGED and s.iterations == 5
    uo, st, rel, it = ox.pcg(g, Ke, np.ones(g.ne), fixed, f, tol=1e-14, max_iter=5)
    assert rel > 1e-2 and np.all(u == 0.0)  # no iterate below 1e-2 after 5 iterations: the zero start
    # a longer cap: the best iterate is one of the oracle's iterates (the one of least residual)
    best = None
    for k in range(60, 81):
        uk, _, rk, _ = ox.pcg(g, Ke, np.ones(g.ne), fixed, f, tol=1e-14, max_iter=k)
        if rk < 1e-2 and (best is None or rk < best[1]):
            best = (uk, rk)
    u, s = ctx.pcg(np.ones(g.ne), f, tol=1e-14, max_iter=80)
    assert best is not None
    assert np.allclose(u, best[0], rtol=1e-9, atol=1e-9 * np.abs(best[0]).max())


@pytest.mark.parametrize("quirk", [0, 1])
@pytest.mark.parametrize("dims,R", [((6, 5, 4), 1.2), ((60, 20, 20), 0.48), ((33, 17, 9), 2.5 / (2 * 3 ** 0.5))])
def test_helmholtz_matches_oracle(ox, ctx, quirk, dims, R):
    _say("helmholtz", dims, R, quirk)
    g = _grid_setup(ox, ctx, dims)
    rho = np.random.default_rng(3).uniform(0, 1, g.ne)
    out, s = ctx.helmholtz(R, rho, quirk=quirk, tol=1e-12, max_iter=5000, check_every=1)
    oo, st, rel, it = ox.helmholtz(g, R, quirk, rho, tol=1e-12, max_iter=5000, check_every=1)
    assert np.allclose(out, oo, rtol=1e-9, atol=1e-11)
    # reference settings (tol 1e-5, residual every 10 its, <= 200 its): same stopping iteration
    out, s = ctx.helmholtz(R, rho, quirk=quirk, tol=1e-5, max_iter=200, check_every=10)
    oo, st, rel, it = ox.helmholtz(g, R, quirk, rho, tol=1e-5, max_iter=200, check_every=10)
    assert s.iterations == it
    assert np.allclose(out, oo, rtol=1e-8, atol=1e-10)


def test_helmholtz_constant_two_thirds(ctx, ox):
    _grid_setup(ox, ctx, (10, 6, 5))
    c = np.full(300, 0.6)
    out, _ = ctx.helmholtz(0.9, c, quirk=1, tol=1e-13, max_iter=2000, check_every=1)
    assert np.allclose(out, 0.4, rtol=1e-9)
    out, _ = ctx.helmholtz(0.9, c, quirk=0, tol=1e-13, max_iter=2000, check_every=1)
    assert np.allclose(out, 0.6, rtol=1e-9)


def test_stress_matches_oracle(ox, ctx):
    g = _grid_setup(ox, ctx, (9, 7, 5), (1.0, 0.8, 1.3))
    rng = np.random.default_rng(5)
    u = rng.standard_normal(g.ndof)
    rho = rng.uniform(0, 1, g.ne)
    rho[::7] = 1e-7
    E = 2.0
    Ee = np.where(rho > 1e-6, E * rho, 0.0)
    out = ctx.element_stress(Ee, u)
    pr, vm, sv, l1, dr = ox.stress(g, 0.3, E, rho, u)
    sc = np.abs(sv).max()
    assert np.allclose(out["voigt"], sv, rtol=1e-11, atol=1e-12 * sc)
    assert np.allclose(out["vm"], vm, rtol=1e-11, atol=1e-12 * sc)
    assert np.allclose(out["principal"], pr, rtol=1e-10, atol=1e-11 * sc)
    assert np.allclose(out["l1"], l1, rtol=1e-10, atol=1e-11 * sc)
    live = rho > 1e-6
    dots = np.abs(np.sum(out["dir"][live] * dr[live], axis=1))
    gap = np.min(np.abs(np.diff(pr[live], axis=1)), axis=1)
    ok = gap > 1e-6 * sc  # direction defined only for a distinct largest eigenvalue
    assert np.all(np.abs(dots[ok] - 1) < 1e-8)


def test_compliance_sens_matches_oracle(ox, ctx):
    g = _grid_setup(ox, ctx, (11, 6, 5))
    rng = np.random.default_rng(6)
    u = rng.standard_normal(g.ndof)
    f = rng.standard_normal(g.ndof)
    xp = rng.uniform(0.05, 1, g.ne)
    ce, dc, comp = ctx.compliance_sens(u, xphys=xp, penal=3.0, E0=1.0, Emin=1e-9, f=f)
    ceo = ox.element_energy(g, ox.ke_hex(1, 1, 1, 0.3), u)
    assert np.allclose(ce, ceo, rtol=1e-11, atol=1e-12 * np.abs(ceo).max())
    assert np.allclose(dc, ox.simp_dc(xp, ceo, 3.0, 1.0, 1e-9), rtol=1e-10)
    assert abs(comp - f @ u) < 1e-11 * np.abs(f * u).sum()


def test_oc_matches_oracle(ox, ctx):
    g = _grid_setup(ox, ctx, (20, 10, 10))
    rng = np.random.default_rng(8)
    x = rng.uniform(0.05, 0.9, g.ne)
    dc = -rng.uniform(0.01, 5.0, g.ne)
    dv = rng.uniform(0.9, 1.1, g.ne)
    xn, lm, steps, ch = ctx.oc_update(x, dc, dv, 0.3, move=0.2, xmin=0.0, rel_tol=1e-9)
    xo, lo, so = ox.oc(x, dc, dv, 0.3, move=0.2, xmin=0.0, rel_tol=1e-9)
    assert steps == so
    assert abs(lm - lo) <= 1e-12 * lo
    assert np.allclose(xn, xo, rtol=1e-12, atol=1e-14)
    assert abs(ch - np.abs(xo - x).max()) < 1e-14


def test_oc_search_from_a_guessed_bracket(ox, ctx):
    """The SIMP loop's 8-point search finds top88's multiplier from [0, 1e9]
    and from a guessed bracket [g/16, 16 g] around any guess g: inside it,
    and with the crossing below or above it, where the first pass has to
    widen the bracket to 0 or to 1e9.  A close guess needs fewer passes."""
    g = _grid_setup(ox, ctx, (20, 10, 10))
    rng = np.random.default_rng(9)
    x = rng.uniform(0.05, 0.9, g.ne)
    dc = -rng.uniform(0.01, 5.0, g.ne)
    dv = rng.uniform(0.9, 1.1, g.ne)
    xo, lo, _ = ox.oc(x, dc, dv, 0.3, move=0.2, xmin=0.0, rel_tol=1e-12)
    _, _, cold = ctx.test_oc_search(x, dc, dv, 0.3, rel_tol=1e-9)
    for guess in (0.0, 1.3 * lo, lo / 7.0, 300.0 * lo, lo / 300.0):
        xn, lm, passes = ctx.test_oc_search(x, dc, dv, 0.3, rel_tol=1e-9, lguess=guess)
        _say("oc search guess", guess / lo, "passes", passes, "cold", cold)
        assert abs(lm - lo) <= 2e-9 * lo, (guess, lm, lo)
        assert np.allclose(xn, xo, rtol=2e-9, atol=1e-12)
        assert abs(xn.mean() - 0.3) < 1e-8
        if 0.1 < guess / lo < 10.0:
            assert passes < cold


@pytest.mark.parametrize("dims,precision,n_iter", [((24, 8, 8), 0, 4), ((24, 8, 8), 1, 4), ((60, 20, 20), 0, 3)])
def test_simp_loop_matches_oracle(ox, hx, ctx, dims, precision, n_iter):
    _say("simp loop", dims, precision, n_iter)
    g = _grid_setup(ox, ctx, dims)
    fixed, f = _cantilever(ox, g)
    ctx.set_fixed(fixed)
    kw = dict(nu=0.3, E0=1.0, Emin=1e-9, penal=3.0, volfrac=0.3, rmin=1.5, move=0.2, xmin=0.0, pcg_tol=1e-10,
              filter_tol=1e-12, oc_rel_tol=1e-10)
    hist, xo, xpo, uo = ox.simp_to_loop(g, f, fixed, n_iter, **kw)
    L = hx._lib
    p = L.TOParams(E0=1.0, Emin=1e-9, nu=0.3, penal=3.0, volfrac=0.3, move=0.2, xmin=0.0, rmin=1.5,
                   pcg_tol=1e-10, pcg_max_iter=200000, precision=precision, precond=0, filter_tol=1e-12,
                   filter_max_iter=5000, oc_rel_tol=1e-10, with_stress=1)
    ctx.to_init(p, f, np.full(g.ne, 0.3))
    for h in hist:
        s = ctx.to_step()
        _say("  it", s.iter, "c", s.compliance, "oracle", h["compliance"], "pcg", s.pcg_iterations, h["pcg_iterations"])
        assert abs(s.compliance - h["compliance"]) <= 1e-6 * abs(h["compliance"])
        assert abs(s.change - h["change"]) <= 1e-6
    xp = ctx.to_read(L.HX_FIELD_XPHYS)
    x = ctx.to_read(L.HX_FIELD_X)
    assert np.allclose(x, xo, rtol=1e-6, atol=1e-7)
    assert np.allclose(xp, xpo, rtol=1e-6, atol=1e-7)


@pytest.mark.parametrize("precision", [0, 1])
def test_block_pcg_columns_equal_single_solves(ox, hx, ctx, precision):
    """Eight load cases in one fused solve: every column is bit-identical to
    the single-RHS solve (same kernels, same reduction order per column)."""
    _say("block pcg", precision)
    g = _grid_setup(ox, ctx, (16, 8, 8))
    fixed, _ = _cantilever(ox, g)
    ctx.set_fixed(fixed)
    rng = np.random.default_rng(11)
    X = g.node_coords()
    face = np.nonzero(X[:, 0] == g.nx)[0]
    F = np.zeros((8, g.ndof))
    for k in range(8):
        d = rng.standard_normal(3)
        F[k, 3 * face[rng.integers(face.size)] + np.arange(3)] = d / np.linalg.norm(d)
    Ee = 1e-9 + rng.uniform(0.05, 1.0, g.ne) ** 3
    U, S = ctx.block_pcg(Ee, F, tol=1e-9, precision=precision)
    Ke = ox.ke_hex(1, 1, 1, 0.3)
    for k in range(8):
        u, s = ctx.pcg(Ee, F[k], tol=1e-9, precision=precision)
        assert S[k].status == 0 and S[k].iterations == s.iterations
        assert np.array_equal(U[k], u)
        assert _true_rel(ox, g, Ke, Ee, fixed, F[k], U[k]) < 1.05e-9


def test_graph_replay_matches_plain_launches(ox, hx, ctx, monkeypatch):
    """Captured iteration batches (hipGraph, replayed many times) give the
    same iterates as plain launches."""
    g = _grid_setup(ox, ctx, (20, 6, 6))
    fixed, f = _cantilever(ox, g)
    ctx.set_fixed(fixed)
    Ee = 1e-9 + np.random.default_rng(12).uniform(0.05, 1.0, g.ne) ** 3
    u1, s1 = ctx.pcg(Ee, f, tol=1e-10, batch=4)
    monkeypatch.setenv("HEXAT_NO_GRAPH", "1")
    c2 = hx.Context(0)
    try:
        c2.set_grid(20, 6, 6)
        c2.set_fixed(fixed)
        u2, s2 = c2.pcg(Ee, f, tol=1e-10, batch=4)
        u3, s3 = c2.pcg(Ee, f, tol=1e-10, batch=4, precision=1)
    finally:
        c2.close()
    u4, s4 = ctx.pcg(Ee, f, tol=1e-10, batch=4, precision=1)
    assert s1.iterations == s2.iterations and np.array_equal(u1, u2)
    assert s3.iterations == s4.iterations and np.array_equal(u3, u4)


@pytest.mark.parametrize("variant", [1, 2, 4])
def test_fp32_contraction_variants_solve_to_fp64_tolerance(ox, hx, ctx, variant):
    """The precision form with each FP32 element contraction (parity-basis
    VALU, dense VALU, dense MFMA) reaches the FP64 true residual 1e-8 and the
    FP64 oracle's compliance."""
    _say("fp32 variant", variant)
    g = _grid_setup(ox, ctx, (24, 8, 8))
    fixed, f = _cantilever(ox, g)
    ctx.set_fixed(fixed)
    Ee = 1e-9 + np.random.default_rng(13).uniform(0.05, 1.0, g.ne) ** 3
    Ke = ox.ke_hex(1, 1, 1, 0.3)
    u, s = ctx.pcg(Ee, f, tol=1e-8, precision=1, matvec=variant, max_iter=20000)
    uo, *_ = ox.pcg(g, Ke, Ee, fixed, f, tol=1e-10)
    _say("  its", s.iterations, "refinements", s.refinements, "rel", s.rel_res)
    assert s.status == 0
    assert _true_rel(ox, g, Ke, Ee, fixed, f, u) < 1.05e-8
    assert abs(f @ u - f @ uo) < 1e-7 * abs(f @ uo)


SHAPES = [(0, 8, 1), (0, 16, 2), (0, 32, 1), (1, 8, 1), (1, 16, 2), (2, 8, 1), (2, 16, 2), (2, 32, 1), (2, 2, 2),
          (2, 1, 1), (0, 1, 2)]


def _tune(tile, kchunk, pdepth):
    return {"tile": tile, "kchunk": kchunk, "pdepth": pdepth, "tile32": tile, "kchunk32": kchunk,
            "pdepth32": pdepth, "variant64": 1, "variant32": 1}


@pytest.mark.parametrize("shape", SHAPES)
def test_operator_launch_shapes_match_oracle(ox, hx, ctx, shape):
    """Every launch shape the autotuner can pick (16x16 / 16x32 / 32x8 tiles,
    z-chunks, one or two prefetched planes) applies the operator to FP64
    round-off and drives both PCG forms (fused p-update), the multilevel
    form (fused z + beta p) and the V-cycle (its smoothing forms) to the FP64
    tolerance."""
    _say("operator shape", shape)
    for dims in ((17, 16, 18), (31, 9, 40)):
        g = _grid_setup(ox, ctx, dims)
        ctx.tune_set(_tune(*shape))
        rng = np.random.default_rng(11)
        Ee = rng.uniform(0.01, 2.0, g.ne)
        Ee[rng.random(g.ne) < 0.1] = 0.0
        x = rng.standard_normal(g.ndof)
        yo = ox.matvec(g, ox.ke_hex(1, 1, 1, 0.3), Ee, x)
        assert np.allclose(ctx.matvec(Ee, x, 1), yo, rtol=1e-12, atol=1e-12 * np.abs(yo).max())
    g = _grid_setup(ox, ctx, (24, 8, 8))
    ctx.tune_set(_tune(*shape))
    fixed, f = _cantilever(ox, g)
    ctx.set_fixed(fixed)
    Ee = 1e-9 + np.clip(np.random.default_rng(3).uniform(0, 1, g.ne), 0.01, 1.0) ** 3
    Ke = ox.ke_hex(1, 1, 1, 0.3)
    for precision in (0, 1):
        for pc in (hx._lib.HX_PC_JACOBI, hx._lib.HX_PC_MULTILEVEL, hx._lib.HX_PC_VCYCLE):
            u, st = ctx.pcg(Ee, f, tol=1e-8, precision=precision, precond=pc)
            assert st.status == 0 and _true_rel(ox, g, Ke, Ee, fixed, f, u) < 1.05e-8, (precision, pc, st.rel_res)


def test_tune_set_rejects_invalid_shapes(hx, ctx):
    with pytest.raises(hx.HexatError):
        ctx.tune_set(_tune(5, 16, 1))
    with pytest.raises(hx.HexatError):
        ctx.tune_set(_tune(0, 16, 3))


def test_timing_hooks(ox, hx, ctx):
    """Operator timing (both precisions, cache-defeating buffer sets) and the
    copy-kernel calibration run and report positive durations."""
    _grid_setup(ox, ctx, (32, 16, 16))
    ctx.tune_set(_tune(0, 8, 1))
    assert ctx.time_operator(hx._lib.HX_FP64, 1, 4) > 0.0
    assert ctx.time_operator(hx._lib.HX_FP32_IR, 2, 4) > 0.0
    assert ctx.test_copy(1 << 20, 2) > 0.0
    with pytest.raises(hx.HexatError):
        ctx.test_copy(1000, 1)  # not a multiple of 16 bytes
