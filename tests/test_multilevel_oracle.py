"""Pins the oracle's multilevel preconditioner (CPU only) against an
independent matrix construction: trilinear prolongations as Kronecker
products of 1-D interpolation matrices, coarse diagonals from scipy assembly
of the coarse-grid stiffness, the additive sum formed with sparse products.
A mistake in a weight, an index, a level scaling or a mask fails here."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import oracle as ox


def _interp_1d(n_el):
    """(n_el+1) x (n_el/2+1) linear interpolation, coarse nodes at even fine nodes."""
    nc = n_el // 2 + 1
    P = sp.lil_matrix((n_el + 1, nc))
    for i in range(n_el + 1):
        if i % 2 == 0:
            P[i, i // 2] = 1.0
        else:
            P[i, i // 2] = P[i, i // 2 + 1] = 0.5
    return P.tocsr()


def _prolong(nx, ny, nz):
    """2:1 trilinear prolongation for node-major DOFs 3 (i + NX (j + NY k)) + c."""
    P = sp.kron(_interp_1d(nz), sp.kron(_interp_1d(ny), _interp_1d(nx)))
    return sp.kron(P, sp.eye(3)).tocsr()


def _assemble(nx, ny, nz, E, Ke):
    g = ox.Grid(nx, ny, nz)
    dof = (3 * g.elements()[:, :, None] + np.arange(3)[None, None, :]).reshape(-1, 24)
    rows = np.repeat(dof, 24, axis=1).ravel()
    cols = np.tile(dof, (1, 24)).ravel()
    vals = (np.asarray(E)[:, None] * Ke.ravel()[None, :]).ravel()
    return sp.csr_matrix((vals, (rows, cols)), shape=(g.ndof, g.ndof))


def _coarse_E_loops(E, nx, ny, nz):
    Ec = np.zeros((nx // 2) * (ny // 2) * (nz // 2))
    for K in range(nz // 2):
        for J in range(ny // 2):
            for I in range(nx // 2):
                s = 0.0
                for c in (0, 1):
                    for b in (0, 1):
                        for a in (0, 1):
                            s += E[(2 * I + a) + nx * ((2 * J + b) + ny * (2 * K + c))]
                Ec[I + (nx // 2) * (J + (ny // 2) * K)] = s / 8
    return Ec


def _matrix_form(nx, ny, nz, dx, dy, dz, E, fixed, exact=True):
    """M^-1 as a dense matrix from the independent construction.  With `exact`,
    the first level with at most 192 DOFs is the last and enters through the
    inverse of its assembled stiffness (boundary and stiffness-free DOFs
    removed) instead of its diagonal."""
    Ke = ox.ke_hex(dx, dy, dz, 0.3)
    g = ox.Grid(nx, ny, nz, dx, dy, dz)
    L = ox.ml_levels(g)
    last_exact = False
    if exact:
        for l in range(1, L + 1):
            if 3 * (nx // 2 ** l + 1) * (ny // 2 ** l + 1) * (nz // 2 ** l + 1) <= 192:
                L, last_exact = l, True
                break
    fx = np.asarray(fixed, dtype=bool)
    A0 = _assemble(nx, ny, nz, E, Ke)
    d0 = A0.diagonal()
    free0 = ~fx & (d0 > 0)
    Minv = sp.diags(np.where(free0, 1.0 / np.where(free0, d0, 1.0), 0.0)).toarray()
    P = sp.eye(g.ndof, format="csr")
    El, n = np.asarray(E, dtype=float), (nx, ny, nz)
    NX, NY = nx + 1, ny + 1
    for l in range(1, L + 1):
        P = (P @ _prolong(*n)).tocsr()
        El = _coarse_E_loops(El, *n)
        n = tuple(v // 2 for v in n)
        h = 2 ** l
        Al = _assemble(*n, El, ox.ke_hex(h * dx, h * dy, h * dz, 0.3))
        dl = Al.diagonal()
        # coarse node (I,J,K) sits on fine node (hI, hJ, hK)
        cn = np.arange((n[0] + 1) * (n[1] + 1) * (n[2] + 1))
        I, J, K = cn % (n[0] + 1), (cn // (n[0] + 1)) % (n[1] + 1), cn // ((n[0] + 1) * (n[1] + 1))
        fnode = h * I + NX * (h * J + NY * (h * K))
        cfix = fx.reshape(-1, 3)[fnode].ravel()
        freel = ~cfix & (dl > 0)
        if last_exact and l == L:
            Ad = Al.toarray()[np.ix_(freel, freel)]
            Dd = np.zeros((Al.shape[0], Al.shape[0]))
            Dd[np.ix_(freel, freel)] = np.linalg.inv(Ad)
            Dl = sp.csr_matrix(Dd)
        else:
            Dl = sp.diags(np.where(freel, 1.0 / np.where(freel, dl, 1.0), 0.0))
        Pm = sp.diags(free0.astype(float)) @ P  # the fine result is masked
        Minv += (Pm @ Dl @ P.T).toarray()
    return Minv, L


@pytest.mark.parametrize("exact", [True, False])
@pytest.mark.parametrize("shape", [(8, 4, 4, 1.0, 1.0, 1.0), (8, 8, 4, 1.0, 0.5, 2.0), (12, 4, 8, 1.0, 1.0, 1.0),
                                   (16, 8, 8, 1.0, 1.0, 1.0)])
def test_ml_apply_matches_matrix_form(shape, exact):
    nx, ny, nz, dx, dy, dz = shape
    g = ox.Grid(nx, ny, nz, dx, dy, dz)
    rng = np.random.default_rng(nx * 100 + nz)
    E = 1e-3 + rng.uniform(0, 1, g.ne) ** 3
    E[rng.uniform(0, 1, g.ne) < 0.1] = 0.0  # removed elements
    X = g.node_coords()
    fixed = np.repeat((X[:, 0] == 0.0)[:, None], 3, axis=1).ravel().astype(np.uint8)
    fixed[3 * np.flatnonzero(X[:, 2] == 0.0) + 2] = 1  # a partially fixed face too
    Minv, L = _matrix_form(nx, ny, nz, dx, dy, dz, E, fixed, exact)
    Ke = ox.ke_hex(dx, dy, dz, 0.3)
    assert L == len(ox.ml_hierarchy(g, Ke, E, fixed, exact=exact)) - 1 and L >= 1
    mask = np.diag(Minv) != 0.0
    for s in range(3):
        r = np.where(mask, rng.standard_normal(g.ndof), 0.0)
        z = ox.ml_apply(g, Ke, E, fixed, r, exact=exact)
        zr = Minv @ r
        assert np.allclose(z, zr, rtol=1e-10, atol=1e-10 * np.abs(zr).max())


def test_ml_operator_is_spd_on_free_dofs():
    nx, ny, nz = 8, 4, 4
    g = ox.Grid(nx, ny, nz)
    rng = np.random.default_rng(5)
    E = 1e-3 + rng.uniform(0, 1, g.ne) ** 3
    X = g.node_coords()
    fixed = np.repeat((X[:, 0] == 0.0)[:, None], 3, axis=1).ravel().astype(np.uint8)
    Minv, _ = _matrix_form(nx, ny, nz, 1.0, 1.0, 1.0, E, fixed)
    free = np.diag(Minv) != 0.0
    Mf = Minv[np.ix_(free, free)]
    assert np.allclose(Mf, Mf.T, atol=1e-14 * np.abs(Mf).max())
    assert np.linalg.eigvalsh(0.5 * (Mf + Mf.T)).min() > 0.0


def test_ml_pcg_matches_direct_solve_and_beats_jacobi():
    nx, ny, nz = 16, 8, 8
    g = ox.Grid(nx, ny, nz)
    rng = np.random.default_rng(11)
    E = 1e-6 + rng.uniform(0.05, 1.0, g.ne) ** 3
    X = g.node_coords()
    fixed = np.repeat((X[:, 0] == 0.0)[:, None], 3, axis=1).ravel().astype(np.uint8)
    f = ox.external_forces(X, [{"location": ["100%", "50%", 0], "F": [0, 0, -1]}])
    Ke = ox.ke_hex(1, 1, 1, 0.3)
    x, rel, it_ml = ox.ml_pcg(g, Ke, E, fixed, f, tol=1e-10)
    A = _assemble(nx, ny, nz, E, Ke)
    fr = fixed == 0
    xd = np.zeros(g.ndof)
    xd[fr] = sp.linalg.spsolve(A[fr][:, fr].tocsc(), f[fr])
    assert rel < 1e-10
    assert abs(f @ x - f @ xd) < 1e-8 * abs(f @ xd)
    _, _, _, it_j = ox.pcg(g, Ke, E, fixed, f, tol=1e-10)
    assert it_ml < it_j
