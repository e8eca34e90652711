"""Pins the CPU oracle against independent closed forms, scipy and finite
differences (CPU only).  None of these checks re-types the oracle's own
formulas: stiffness/mass come from exact polynomial integration of the
Lame bilinear form, global operators from scipy assembly and direct solves,
sensitivities from finite differences of a direct solve."""
import itertools

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

NAT = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]])


def _int1d(lo, hi):
    """Exact 1-D integrals of products of the two linear hat functions on [lo,hi]:
    returns dicts m[(s,t)] = int phi_s phi_t, k[(s,t)] = int phi_s' phi_t',
    c[(s,t)] = int phi_s' phi_t  (s,t in {-1,+1}, phi_-1 at lo, phi_+1 at hi)."""
    L = hi - lo
    m, k, c = {}, {}, {}
    for s in (-1, 1):
        for t in (-1, 1):
            m[(s, t)] = L / 3 if s == t else L / 6
            k[(s, t)] = (s * t) / L
            c[(s, t)] = s * 0.5  # phi_s' = s/L, int phi_t = L/2
    return m, k, c


def exact_ke(dx, dy, dz, nu, E=1.0):
    """Lame form a(u,v) = lam div u div v + 2 mu eps(u):eps(v), integrated exactly."""
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    I = [_int1d(-dx / 2, dx / 2), _int1d(-dy / 2, dy / 2), _int1d(-dz / 2, dz / 2)]

    def integ(A, B, da, db):
        # int d_da N_A * d_db N_B (da/db = None for no derivative)
        v = 1.0
        for ax in range(3):
            m, k, c = I[ax]
            s, t = NAT[A][ax], NAT[B][ax]
            if da == ax and db == ax:
                v *= k[(s, t)]
            elif da == ax:
                v *= c[(s, t)]
            elif db == ax:
                v *= c[(t, s)]
            else:
                v *= m[(s, t)]
        return v

    K = np.zeros((24, 24))
    for A in range(8):
        for B in range(8):
            gg = sum(integ(A, B, d, d) for d in range(3))
            for i in range(3):
                for j in range(3):
                    K[3 * A + i, 3 * B + j] = (lam * integ(A, B, i, j) + mu * integ(A, B, j, i)
                                               + (mu * gg if i == j else 0.0))
    return K


def exact_scalar(dx, dy, dz, quirk):
    ztop = dz if quirk else dz / 2
    I = [_int1d(-dx / 2, dx / 2), _int1d(-dy / 2, dy / 2), _int1d(-dz / 2, ztop)]
    Ks, Ms = np.zeros((8, 8)), np.zeros((8, 8))
    for A in range(8):
        for B in range(8):
            mm = 1.0
            for ax in range(3):
                mm *= I[ax][0][(NAT[A][ax], NAT[B][ax])]
            Ms[A, B] = mm
            s = 0.0
            for d in range(3):
                v = 1.0
                for ax in range(3):
                    key = (NAT[A][ax], NAT[B][ax])
                    v *= I[ax][1][key] if ax == d else I[ax][0][key]
                s += v
            Ks[A, B] = s
    return Ks, Ms


def assemble(ox, g, Ke, Ee):
    el = g.elements()
    dofs = (3 * el[:, :, None] + np.arange(3)[None, None, :]).reshape(g.ne, 24)
    rows = np.repeat(dofs, 24, axis=1).ravel()
    cols = np.tile(dofs, (1, 24)).ravel()
    vals = (Ee[:, None, None] * Ke[None, :, :]).reshape(g.ne, 576).ravel()
    return sp.csr_matrix((vals, (rows, cols)), shape=(g.ndof, g.ndof))


@pytest.mark.parametrize("dims,nu", [((1, 1, 1), 0.3), ((1.0, 2.0, 0.5), 0.3), ((0.3, 0.7, 1.1), 0.45),
                                     ((2.5, 1.0, 1.0), 0.0)])
def test_ke_matches_exact_integration(ox, dims, nu):
    K = ox.ke_hex(*dims, nu)
    Kx = exact_ke(*dims, nu)
    assert np.allclose(K, Kx, rtol=1e-12, atol=1e-13 * np.abs(Kx).max())


def test_ke_known_entry_and_rigid_modes(ox):
    K = ox.ke_hex(1.0, 1.0, 1.0, 0.3)
    # unit cube, nu=0.3: Ke11 = 0.2350427 (value quoted from the reference's element)
    assert abs(K[0, 0] - 0.2350427) < 5e-8
    assert np.allclose(K, K.T, atol=1e-15)
    ev = np.linalg.eigvalsh(K)
    assert np.sum(np.abs(ev) < 1e-12) == 6 and ev.min() > -1e-12
    X = NAT * 0.5
    modes = []
    for d in range(3):
        t = np.zeros((8, 3)); t[:, d] = 1; modes.append(t.ravel())
    for (a, b) in [(0, 1), (1, 2), (0, 2)]:
        r = np.zeros((8, 3)); r[:, a] = -X[:, b]; r[:, b] = X[:, a]; modes.append(r.ravel())
    for m in modes:
        assert np.abs(K @ m).max() < 1e-13


def test_patch_energy_uniform_strain(ox):
    """u = eps . x gives u^T K u = vol * eps_v^T D eps_v exactly."""
    dx, dy, dz, nu, E = 1.3, 0.8, 0.6, 0.27, 2.0
    K = ox.ke_hex(dx, dy, dz, nu, E)
    D = ox.material_matrix(E, nu)
    X = NAT * np.array([dx, dy, dz]) / 2
    rng = np.random.default_rng(1)
    for _ in range(5):
        eps = rng.standard_normal((3, 3)); eps = 0.5 * (eps + eps.T)
        u = (X @ eps.T).ravel()
        ev = np.array([eps[0, 0], eps[1, 1], eps[2, 2], 2 * eps[0, 1], 2 * eps[1, 2], 2 * eps[0, 2]])
        assert np.isclose(u @ K @ u, dx * dy * dz * ev @ D @ ev, rtol=1e-12)


@pytest.mark.parametrize("quirk", [0, 1])
def test_scalar_matrices_exact(ox, quirk):
    Ks, Ms = ox.scalar_ke_me(1.0, 0.7, 1.3, quirk)
    Kx, Mx = exact_scalar(1.0, 0.7, 1.3, quirk)
    assert np.allclose(Ks, Kx, rtol=1e-12, atol=1e-14)
    assert np.allclose(Ms, Mx, rtol=1e-12, atol=1e-14)


def test_matvec_matches_scipy_assembly(ox):
    g = ox.Grid(3, 2, 2, 1.0, 0.5, 2.0)
    Ke = ox.ke_hex(g.dx, g.dy, g.dz, 0.3)
    rng = np.random.default_rng(2)
    Ee = rng.uniform(0.1, 2.0, g.ne)
    Ee[3] = 0.0
    x = rng.standard_normal(g.ndof)
    K = assemble(ox, g, Ke, Ee)
    y = ox.matvec(g, Ke, Ee, x)
    assert np.allclose(y, K @ x, rtol=1e-13, atol=1e-13 * np.abs(y).max())
    d = ox.diag(g, Ke, Ee)
    assert np.allclose(d, K.diagonal(), rtol=1e-14)


def _cantilever(ox, nx, ny, nz):
    g = ox.Grid(nx, ny, nz)
    X = g.node_coords()
    fixed = np.repeat((X[:, 0] == 0.0)[:, None], 3, axis=1).ravel().astype(np.uint8)
    f = ox.external_forces(X, [{"location": ["100%", 0, ":"], "F": [0, -1, 0]}])
    return g, fixed, f


def test_pcg_matches_direct_solve(ox):
    g, fixed, f = _cantilever(ox, 6, 3, 2)
    Ke = ox.ke_hex(1, 1, 1, 0.3)
    rng = np.random.default_rng(3)
    Ee = 1e-9 + rng.uniform(0.05, 1.0, g.ne) ** 3
    K = assemble(ox, g, Ke, Ee)
    free = np.nonzero(fixed == 0)[0]
    ud = np.zeros(g.ndof)
    ud[free] = spla.spsolve(K[free][:, free].tocsc(), f[free])
    u, st, rel, it = ox.pcg(g, Ke, Ee, fixed, f, tol=1e-12)
    assert st == 0 and rel < 1e-12
    assert np.allclose(u, ud, rtol=1e-9, atol=1e-10 * np.abs(ud).max())
    assert abs(f @ u - f @ ud) <= 1e-10 * abs(f @ ud)
    # a warm start at the solution converges at once and stays put
    u2, st2, rel2, it2 = ox.pcg(g, Ke, Ee, fixed, f, x0=ud, tol=1e-6)
    assert st2 == 0 and it2 <= 1 and np.allclose(u2, ud, rtol=1e-6)


def test_pcg_zero_rhs_and_shift(ox):
    g, fixed, f = _cantilever(ox, 4, 2, 2)
    Ke = ox.ke_hex(1, 1, 1, 0.3)
    Ee = np.ones(g.ne)
    u, st, rel, it = ox.pcg(g, Ke, Ee, fixed, np.zeros(g.ndof))
    assert st == 3 and not u.any()
    # shifted system (K_ff + s I) against scipy
    s = 1e-3
    K = assemble(ox, g, Ke, Ee)
    free = np.nonzero(fixed == 0)[0]
    A = (K[free][:, free] + s * sp.eye(free.size)).tocsc()
    ud = np.zeros(g.ndof); ud[free] = spla.spsolve(A, f[free])
    u, st, rel, it = ox.pcg(g, Ke, Ee, fixed, f, shift=s, tol=1e-12)
    assert np.allclose(u, ud, rtol=1e-9, atol=1e-12)


def test_uniaxial_tension_exact(ox):
    """Bar with symmetry rollers on three faces under uniform traction: the
    FE solution is exactly linear, u_x = s x / E, u_y = -nu s y / E."""
    nx, ny, nz, nu = 4, 3, 2, 0.3
    g = ox.Grid(nx, ny, nz)
    X = g.node_coords()
    bc = [{"location": [0, ":", ":"], "DoFs": [1]}, {"location": [":", 0, ":"], "DoFs": [2]},
          {"location": [":", ":", 0], "DoFs": [3]}]
    fixed = ox.bc_indicator(X, bc).ravel().astype(np.uint8)
    # consistent nodal loads of a uniform traction s on the x=nx face (bilinear quads)
    s = 0.7
    F = np.zeros(g.ndof)
    face = np.nonzero(X[:, 0] == nx)[0]
    for n in face:
        y, z = X[n, 1], X[n, 2]
        wy = 0.5 if y in (0, ny) else 1.0
        wz = 0.5 if z in (0, nz) else 1.0
        F[3 * n] = s * wy * wz
    Ke = ox.ke_hex(1, 1, 1, nu)
    u, st, rel, it = ox.pcg(g, Ke, np.ones(g.ne), fixed, F, tol=1e-13)
    assert st == 0
    assert np.allclose(u[0::3], s * X[:, 0], atol=1e-10)
    assert np.allclose(u[1::3], -nu * s * X[:, 1], atol=1e-10)
    assert np.allclose(u[2::3], -nu * s * X[:, 2], atol=1e-10)


def _helm_assemble(g, Ae):
    el = g.elements()
    rows = np.repeat(el, 8, axis=1).ravel()
    cols = np.tile(el, (1, 8)).ravel()
    vals = np.tile(Ae.ravel(), g.ne)
    return sp.csc_matrix((vals, (rows, cols)), shape=(g.nn, g.nn))


@pytest.mark.parametrize("quirk", [0, 1])
def test_helmholtz_matches_direct(ox, quirk):
    g = ox.Grid(5, 4, 3, 1.0, 1.0, 1.0)
    R = 0.8
    rng = np.random.default_rng(4)
    rho = rng.uniform(0, 1, g.ne)
    Ks, Ms = ox.scalar_ke_me(g.dx, g.dy, g.dz, quirk)
    A = _helm_assemble(g, R * R * Ks + Ms)
    el = g.elements()
    b = np.zeros(g.nn)
    np.add.at(b, el.ravel(), np.repeat(rho * g.dx * g.dy * g.dz / 8, 8))
    xn = spla.spsolve(A, b)
    ref = xn[el].mean(axis=1)
    out, st, rel, it = ox.helmholtz(g, R, quirk, rho, tol=1e-13, max_iter=1000, check_every=1)
    assert st == 0
    assert np.allclose(out, ref, rtol=1e-10, atol=1e-12)
    y = ox.helm_matvec(g, R * R * Ks + Ms, xn)
    assert np.allclose(y, A @ xn, rtol=1e-12)


def test_helmholtz_constant_field(ox):
    """Quirk (top face at +dz) scales a constant by exactly 2/3; the corrected
    mass-preserving filter returns it unchanged."""
    g = ox.Grid(6, 5, 4)
    c = np.full(g.ne, 0.5)
    out_q, *_ = ox.helmholtz(g, 1.2, 1, c, tol=1e-13, max_iter=2000, check_every=1)
    out_c, *_ = ox.helmholtz(g, 1.2, 0, c, tol=1e-13, max_iter=2000, check_every=1)
    assert np.allclose(out_q, 0.5 * 2 / 3, rtol=1e-9)
    assert np.allclose(out_c, 0.5, rtol=1e-9)
    # volume preservation of the corrected filter on a random field
    rho = np.random.default_rng(5).uniform(0, 1, g.ne)
    out, *_ = ox.helmholtz(g, 1.2, 0, rho, tol=1e-13, max_iter=2000, check_every=1)
    assert abs(out.sum() - rho.sum()) < 1e-9 * rho.sum()


def test_stress_uniform_strain_and_principal(ox):
    g = ox.Grid(3, 2, 2, 1.0, 0.8, 1.2)
    X = g.node_coords()
    nu, E = 0.3, 2.0
    eps = np.array([[1e-3, 2e-4, -3e-4], [2e-4, -5e-4, 1e-4], [-3e-4, 1e-4, 7e-4]])
    u = (X @ eps.T).ravel()
    rho = np.linspace(0.2, 1.0, g.ne)
    rho[1] = 1e-7  # below the reference's 1e-6 cut (Stress.jl:85)
    pr, vm, sv, l1, dr = ox.stress(g, nu, E, rho, u)
    ev = np.array([eps[0, 0], eps[1, 1], eps[2, 2], 2 * eps[0, 1], 2 * eps[1, 2], 2 * eps[0, 2]])
    D1 = ox.material_matrix(1.0, nu)
    for e in range(g.ne):
        if e == 1:
            assert not sv[e].any() and vm[e] == 0 and l1[e] == 0
            continue
        s = E * rho[e] * (D1 @ ev)
        assert np.allclose(sv[e], s, rtol=1e-11)
        S = np.array([[s[0], s[3], s[5]], [s[3], s[1], s[4]], [s[5], s[4], s[2]]])
        w, V = np.linalg.eigh(S)
        assert np.allclose(pr[e], w[::-1], rtol=1e-10, atol=1e-14)
        assert np.isclose(l1[e], np.abs(w).sum(), rtol=1e-10)
        assert abs(abs(dr[e] @ V[:, -1]) - 1) < 1e-9
        dev = S - np.trace(S) / 3 * np.eye(3)
        assert np.isclose(vm[e], np.sqrt(1.5 * np.sum(dev * dev)), rtol=1e-10)


def test_eig3_repeated_and_diagonal(ox):
    v, V = ox.eig3(np.diag([3.0, -1.0, 2.0]))
    assert np.allclose(v, [3, 2, -1])
    v, V = ox.eig3(np.full((3, 3), 1.0))
    assert np.allclose(v, [3, 0, 0], atol=1e-14)
    assert np.allclose(np.abs(V[:, 0]), 1 / np.sqrt(3))


def test_sensitivity_finite_difference(ox):
    """dc/drho_e = -p rho^(p-1) (E0-Emin) u_e^T Ke u_e against central differences
    of c = f^T K(rho)^-1 f from scipy direct solves."""
    g, fixed, f = _cantilever(ox, 3, 2, 2)
    Ke = ox.ke_hex(1, 1, 1, 0.3)
    p, E0, Emin = 3.0, 1.0, 1e-9
    rho = np.random.default_rng(6).uniform(0.3, 1.0, g.ne)
    free = np.nonzero(fixed == 0)[0]

    def comp(r):
        K = assemble(ox, g, Ke, ox.simp_modulus(r, p, E0, Emin))
        u = np.zeros(g.ndof)
        u[free] = spla.spsolve(K[free][:, free].tocsc(), f[free])
        return f @ u, u

    c0, u0 = comp(rho)
    dc = ox.simp_dc(rho, ox.element_energy(g, Ke, u0), p, E0, Emin)
    h = 1e-6
    for e in range(g.ne):
        rp, rm = rho.copy(), rho.copy()
        rp[e] += h; rm[e] -= h
        fd = (comp(rp)[0] - comp(rm)[0]) / (2 * h)
        assert abs(fd - dc[e]) <= 1e-5 * abs(dc).max()


def test_oc_volume_and_bounds(ox):
    rng = np.random.default_rng(7)
    n = 500
    x = rng.uniform(0.05, 0.9, n)
    dc = -rng.uniform(0.01, 5.0, n)
    xn, lm, steps = ox.oc(x, dc, np.ones(n), 0.3, move=0.2, xmin=0.0, rel_tol=1e-12)
    assert abs(xn.mean() - 0.3) < 1e-6
    assert np.all(xn >= np.maximum(0, x - 0.2) - 1e-15) and np.all(xn <= np.minimum(1, x + 0.2) + 1e-15)
    # at fixed lambda the update is monotone in -dc for equal x
    xe = np.full(4, 0.5)
    xn2, _, _ = ox.oc(xe, np.array([-1.0, -2.0, -3.0, -4.0]), np.ones(4), 0.5, move=0.2, rel_tol=1e-12)
    assert np.all(np.diff(xn2) >= 0)


def test_regrid_map_closed_forms(ox):
    """MeshRefiner.jl:131-150: same grid -> same field; each dimension doubled
    -> every old value on its 2x2x2 children; a coarser grid samples the old
    element under each new centroid."""
    rng = np.random.default_rng(0)
    f = rng.uniform(0, 1, 5 * 4 * 3).astype(np.float32)
    assert np.array_equal(ox.regrid_map((5, 4, 3), (5, 4, 3), (2.5, 4.0, 0.9), f), f)
    g = ox.regrid_map((5, 4, 3), (10, 8, 6), (2.5, 4.0, 0.9), f).reshape(6, 8, 10)
    assert np.array_equal(g, np.repeat(np.repeat(np.repeat(f.reshape(3, 4, 5), 2, 0), 2, 1), 2, 2))
    h = ox.regrid_map((6, 4, 2), (3, 2, 1), (6.0, 4.0, 2.0), np.arange(48, dtype=np.float32)).astype(int)
    # new centroid (2i + 1) * 1.0 in x lies in old element 2i + 1 (0-based), likewise y, z: (2i+1, 2j+1, 1)
    want = [(2 * i + 1) + 6 * ((2 * j + 1) + 4 * 1) for j in range(2) for i in range(3)]
    assert list(h) == want


def test_smooth_field_generator(ox):
    """The manufactured displacement of the learned preconditioner's training:
    splitmix64 against its published first outputs from seed 0, the
    Box-Muller values' moments over many lattice keys, and the structure of
    the sum: with one level the field equals the lattice values at even nodes
    and their mean at odd ones (linear interpolation along each axis)."""
    z = np.array([0, 0x9E3779B97F4A7C15, (2 * 0x9E3779B97F4A7C15) % 2 ** 64], dtype=np.uint64)
    assert [int(v) for v in ox.mix64(z)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    v = ox.lattice_normal(7, np.arange(200000, dtype=np.uint64))
    assert abs(v.mean()) < 0.01 and abs(v.std() - 1.0) < 0.01
    assert abs((v ** 3).mean()) < 0.03 and abs((v ** 4).mean() - 3.0) < 0.06
    g = ox.Grid(8, 6, 4)
    e = ox.smooth_field(g, 11, 1).reshape(5, 7, 9, 3)  # (k, j, i, c)
    even = e[::2, ::2, ::2]
    LX, LY = (9 + 1) // 2 + 1, (7 + 1) // 2 + 1
    K, J, I = np.meshgrid(np.arange(3), np.arange(4), np.arange(5), indexing="ij")
    pt = (np.uint64(1) << np.uint64(56)) ^ ((K * LY + J) * LX + I).astype(np.uint64)
    for c in range(3):
        assert np.allclose(even[..., c], ox.lattice_normal(11, np.uint64(3) * pt + np.uint64(c)), rtol=0, atol=1e-15)
    assert np.allclose(e[:, :, 1::2], 0.5 * (e[:, :, 0:-1:2] + e[:, :, 2::2]), atol=1e-14)
    assert np.allclose(e[:, 1::2], 0.5 * (e[:, 0:-1:2] + e[:, 2::2]), atol=1e-14)
    assert np.allclose(e[1::2], 0.5 * (e[0:-1:2] + e[2::2]), atol=1e-14)
