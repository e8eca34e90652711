"""Pins the oracle's V-cycle preconditioner (CPU only) against an independent
construction: level stiffness matrices assembled with scipy from explicit
loops over children for the coarse moduli, trilinear prolongations as
Kronecker products of 1-D interpolation matrices, the smoother diagonal as the
absolute row sums of the assembled matrix with its unsolved DOFs removed, and
the cycle applied with sparse products.  A mistake in a level's moduli scale,
a smoother mask, the coupling sum before the absolute value, a restriction or
prolongation, or the order of the smoothing steps fails here; the dense form
on a small grid is symmetric positive definite."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import oracle as ox
from tests.test_multilevel_oracle import _assemble, _coarse_E_loops, _prolong


def _fixed(g, partial=True):
    X = g.node_coords()
    fixed = np.repeat((X[:, 0] == 0.0)[:, None], 3, axis=1).ravel().astype(np.uint8)
    if partial:
        fixed[3 * np.flatnonzero(X[:, 2] == 0.0) + 2] = 1
    return fixed


def _levels(nx, ny, nz, dx, dy, dz, E, fixed, omega):
    """Per level: n, A_l (h Ke, averaged moduli), solved mask, smoother S_l, Jacobi D_l^-1, and the
    2:1 prolongation to the level above."""
    Ke = ox.ke_hex(dx, dy, dz, 0.3)
    g = ox.Grid(nx, ny, nz, dx, dy, dz)
    hier = ox.ml_hierarchy(g, Ke, E, fixed)
    L = len(hier) - 1
    fx = np.asarray(fixed, dtype=bool).reshape(nz + 1, ny + 1, nx + 1, 3)
    out, El, n = [], np.asarray(E, dtype=float), (nx, ny, nz)
    for l in range(L + 1):
        if l > 0:
            El = _coarse_E_loops(El, *n)
            n = tuple(v // 2 for v in n)
        h = 2 ** l
        A = _assemble(*n, El, h * Ke).tocsr()
        solved = ~fx[::h, ::h, ::h].reshape(-1) & (A.diagonal() > 0)
        Ak = sp.diags(solved.astype(float)) @ A @ sp.diags(solved.astype(float))
        l1 = np.asarray(abs(Ak).sum(axis=1)).ravel()
        S = np.where(solved, omega / np.where(solved, l1, 1.0), 0.0)
        Dinv = np.where(solved, 1.0 / np.where(solved, A.diagonal(), 1.0), 0.0)
        P = _prolong(*n) if l < L else None
        out.append(dict(n=n, A=A, solved=solved, S=S, Dinv=Dinv, P=P))
    return out, hier


def _vcycle_sparse(lv, hier, r):
    L = len(lv) - 1
    c = L  # every level above the coarsest is smoothed

    def additive(l, rl):
        if l == L:
            return hier[L][2] @ rl if hier[L][2] is not None else lv[L]["Dinv"] * rl
        P = lv[l]["P"]
        return lv[l]["Dinv"] * rl + P @ additive(l + 1, P.T @ rl)

    # the level above an exact coarsest level, when small, visits it twice, and so does the level
    # above that when it is not small (and not the fine level)
    nn = [(n[0] + 1) * (n[1] + 1) * (n[2] + 1) for n in (v["n"] for v in lv)]
    w = set()
    if L >= 2 and hier[L][2] is not None and nn[L - 1] <= 512:
        w = {L - 1} | ({L - 2} if L - 2 >= 1 and nn[L - 2] > 512 else set())

    def V(l, rl):
        if l == c:
            return additive(l, rl)
        A, S, m, P = lv[l]["A"], lv[l]["S"], lv[l]["solved"], lv[l]["P"]
        x = S * rl
        for _ in range(2 if l in w else 1):
            res = np.where(m, rl - A @ x, 0.0)
            x = x + np.where(m, P @ V(l + 1, P.T @ res), 0.0)
            x = x + S * (rl - A @ x)
        return x

    return V(0, r), c, w


SHAPES = [(8, 4, 4, 1.0, 1.0, 1.0), (16, 8, 8, 1.0, 0.5, 2.0), (48, 24, 24, 1.0, 1.0, 1.0), (7, 4, 4, 1.0, 1.0, 1.0),
          (32, 16, 16, 1.0, 1.0, 1.0)]


@pytest.mark.parametrize("shape", SHAPES)
def test_smoother_is_absolute_row_sum(shape):
    nx, ny, nz, dx, dy, dz = shape
    g = ox.Grid(nx, ny, nz, dx, dy, dz)
    rng = np.random.default_rng(nx + ny)
    E = 1e-3 + rng.uniform(0, 1, g.ne) ** 3
    E[rng.uniform(0, 1, g.ne) < 0.1] = 0.0
    fixed = _fixed(g)
    lv, _ = _levels(nx, ny, nz, dx, dy, dz, E, fixed, 1.7)
    S0 = ox.vc_smoother((nx, ny, nz), E, 1.0, ox.ke_hex(dx, dy, dz, 0.3), fixed.reshape(nz + 1, ny + 1, nx + 1, 3),
                        1.7)
    assert np.allclose(S0.ravel(), lv[0]["S"], rtol=1e-12, atol=0.0)
    if len(lv) > 1:  # a coarse level: moduli 2 x averaged, fixed where the fine DOF is
        E1 = _coarse_E_loops(E, nx, ny, nz)
        fx1 = fixed.reshape(nz + 1, ny + 1, nx + 1, 3)[::2, ::2, ::2]
        S1 = ox.vc_smoother(lv[1]["n"], E1, 2.0, ox.ke_hex(dx, dy, dz, 0.3), fx1, 1.7)
        assert np.allclose(S1.ravel(), lv[1]["S"], rtol=1e-12, atol=0.0)


@pytest.mark.parametrize("shape", SHAPES)
def test_vc_apply_matches_sparse_cycle(shape):
    nx, ny, nz, dx, dy, dz = shape
    g = ox.Grid(nx, ny, nz, dx, dy, dz)
    rng = np.random.default_rng(nx * 7 + nz)
    E = 1e-3 + rng.uniform(0, 1, g.ne) ** 3
    E[rng.uniform(0, 1, g.ne) < 0.1] = 0.0
    fixed = _fixed(g)
    lv, hier = _levels(nx, ny, nz, dx, dy, dz, E, fixed, 1.95)
    Ke = ox.ke_hex(dx, dy, dz, 0.3)
    state = {}
    for s in range(2):
        r = np.where(lv[0]["solved"], rng.standard_normal(g.ndof), 0.0)
        z = ox.vc_apply(g, Ke, E, fixed, r, state=state)
        zr, c, w = _vcycle_sparse(lv, hier, r)
        assert c == state["c"] and w == ox.vc_w_levels(state["levels"])
        assert np.allclose(z, zr, rtol=1e-10, atol=1e-10 * np.abs(zr).max())
    if shape[0] == 48:
        assert state["c"] == 3  # three smoothed levels above a coarsest level solved by its diagonal
    if shape[0] == 32:  # levels 32x16x16, 16x8x8 (1377 nodes), 8x4x4 (225), 4x2x2 exact
        assert state["c"] == 3 and ox.vc_w_levels(state["levels"]) == {1, 2}


def test_vc_dense_form_is_spd():
    nx, ny, nz = 8, 4, 4
    g = ox.Grid(nx, ny, nz)
    rng = np.random.default_rng(2)
    E = 1e-6 + rng.uniform(0, 1, g.ne) ** 3
    fixed = _fixed(g)
    Ke = ox.ke_hex(1, 1, 1, 0.3)
    state = {}
    M = np.stack([ox.vc_apply(g, Ke, E, fixed, e, state=state) for e in np.eye(g.ndof)], axis=1)
    free = np.diag(M) != 0.0
    Mf = M[np.ix_(free, free)]
    assert np.allclose(Mf, Mf.T, atol=1e-12 * np.abs(Mf).max())
    assert np.linalg.eigvalsh(0.5 * (Mf + Mf.T)).min() > 0.0


def test_vc_symmetric_with_two_smoothed_levels():
    nx, ny, nz = 48, 24, 24
    g = ox.Grid(nx, ny, nz)
    rng = np.random.default_rng(4)
    E = 1e-9 + rng.uniform(0, 1, g.ne) ** 3
    fixed = _fixed(g, partial=False)
    Ke = ox.ke_hex(1, 1, 1, 0.3)
    state = {}
    mask = ox.ml_hierarchy(g, Ke, E, fixed, 0)[0][1].ravel() != 0
    u, v = (np.where(mask, rng.standard_normal(g.ndof), 0.0) for _ in range(2))
    Mu, Mv = ox.vc_apply(g, Ke, E, fixed, u, state=state), ox.vc_apply(g, Ke, E, fixed, v, state=state)
    assert state["c"] == 3
    assert abs(v @ Mu - u @ Mv) < 1e-10 * abs(v @ Mu)
    assert u @ Mu > 0 and v @ Mv > 0


def test_vc_symmetric_with_the_w_visit():
    """16x8x8: levels 16x8x8, 8x4x4 (225 nodes, visits the exact 4x2x2 level twice), 4x2x2."""
    nx, ny, nz = 16, 8, 8
    g = ox.Grid(nx, ny, nz)
    rng = np.random.default_rng(6)
    E = 1e-9 + rng.uniform(0, 1, g.ne) ** 3
    fixed = _fixed(g)
    Ke = ox.ke_hex(1, 1, 1, 0.3)
    state = {}
    mask = ox.ml_hierarchy(g, Ke, E, fixed, 0)[0][1].ravel() != 0
    u, v = (np.where(mask, rng.standard_normal(g.ndof), 0.0) for _ in range(2))
    Mu, Mv = ox.vc_apply(g, Ke, E, fixed, u, state=state), ox.vc_apply(g, Ke, E, fixed, v, state=state)
    assert state["c"] == 2 and ox.vc_w_levels(state["levels"]) == {1}
    assert abs(v @ Mu - u @ Mv) < 1e-10 * abs(v @ Mu)
    assert u @ Mu > 0 and v @ Mv > 0


def test_vc_pcg_matches_direct_solve_and_beats_additive():
    nx, ny, nz = 16, 8, 8
    g = ox.Grid(nx, ny, nz)
    rng = np.random.default_rng(11)
    E = 1e-6 + rng.uniform(0.05, 1.0, g.ne) ** 3
    fixed = _fixed(g, partial=False)
    f = ox.external_forces(g.node_coords(), [{"location": ["100%", "50%", 0], "F": [0, 0, -1]}])
    Ke = ox.ke_hex(1, 1, 1, 0.3)
    x, rel, it_vc = ox.vc_pcg(g, Ke, E, fixed, f, tol=1e-10)
    A = _assemble(nx, ny, nz, E, Ke)
    fr = fixed == 0
    xd = np.zeros(g.ndof)
    xd[fr] = sp.linalg.spsolve(A[fr][:, fr].tocsc(), f[fr])
    assert rel < 1e-10
    assert abs(f @ x - f @ xd) < 1e-8 * abs(f @ xd)
    _, _, it_ml = ox.ml_pcg(g, Ke, E, fixed, f, tol=1e-10)
    assert it_vc < it_ml
