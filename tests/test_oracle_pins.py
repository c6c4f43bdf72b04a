"""Pins the FP64 oracle against closed forms, invariants and brute force.

The reference ships no tests or golden vectors for this path, so every oracle
function is checked against mathematics chosen so that a plausible slip (sign,
index, transposition, dropped term) fails.
"""
import math

import numpy as np
import pytest
import scipy.sparse as sp

import smx_oracle as O
from smx import mesh as M


@pytest.fixture(scope="module")
def c1():
    m = M.cube_c1()
    return m, O.boundary_data(m.X, m.tets)


# ---------------------------------------------------------------- solid angle
def test_solid_angle_octant_and_sign():
    x = np.zeros(3)
    e = np.eye(3)
    # triangle (e1, e2, e3) seen from the origin spans one octant: 4 pi / 8
    w = O.solid_angle(x, e[0], e[1], e[2])
    assert abs(w - math.pi / 2) < 1e-14
    # reversed orientation flips the sign
    assert abs(O.solid_angle(x, e[0], e[2], e[1]) + math.pi / 2) < 1e-14


def test_solid_angle_cube_faces_sum_to_4pi():
    # unit cube centred at the origin, 12 outward triangles; from the centre
    # every face subtends 4pi/6 and the closed surface 4pi
    m = M.kuhn_box(1, 1, 1, 1.0, 1.0, 1.0, origin=(-0.5, -0.5, -0.5))
    bd = O.boundary_data(m.X, m.tets)
    om = O.solid_angle(np.zeros(3), bd.tri_pts[:, 0], bd.tri_pts[:, 1], bd.tri_pts[:, 2])
    assert abs(om.sum() - 4 * math.pi) < 1e-13
    face = np.round(bd.ntri, 12)
    for n in np.unique(face, axis=0):
        sel = np.all(face == n, axis=1)
        assert abs(om[sel].sum() - 4 * math.pi / 6) < 1e-13


# ----------------------------------------------------------- Lindholm weights
def _one_triangle(p0, p1, p2):
    tri = np.array([p0, p1, p2], dtype=float)
    n = np.cross(tri[1] - tri[0], tri[2] - tri[0])
    area = 0.5 * np.linalg.norm(n)
    n = n / np.linalg.norm(n)
    e = [tri[1] - tri[0], tri[2] - tri[1], tri[0] - tri[2]]
    s = np.array([np.linalg.norm(v) + 1e-30 for v in e])
    xi = np.array([e[i] / s[i] for i in range(3)])
    eta = np.array([np.cross(n, xi[i]) for i in range(3)])
    gamma = np.array([[xi[(i + 1) % 3] @ xi[j] for j in range(3)] for i in range(3)])
    return tri, n, area, s, eta, gamma


def _quadrature_weights(x, tri, n, order=12, sub=6):
    """(1/4pi) int_T phi_k(y) (x-y).n / |x-y|^3 dS_y by Gauss-Legendre on a
    collapsed square, on sub x sub sub-triangles."""
    gx, gw = np.polynomial.legendre.leggauss(order)
    u = 0.5 * (gx + 1)
    wu = 0.5 * gw
    out = np.zeros(3)
    for i in range(sub):
        for j in range(sub - i):
            for up in (False, True):
                if up and j == sub - i - 1:
                    continue
                # barycentric corners of the sub-triangle (in (l1, l2) coordinates)
                if not up:
                    c = np.array([[i, j], [i + 1, j], [i, j + 1]], dtype=float) / sub
                else:
                    c = np.array([[i + 1, j], [i + 1, j + 1], [i, j + 1]], dtype=float) / sub
                for a, wa in zip(u, wu):
                    for b, wb in zip(u, wu):
                        s_, t_ = a, b * (1 - a)
                        jac = (1 - a)
                        l12 = c[0] + s_ * (c[1] - c[0]) + t_ * (c[2] - c[0])
                        l1, l2 = l12
                        lam = np.array([1 - l1 - l2, l1, l2])
                        y = lam @ tri
                        d1, d2 = c[1] - c[0], c[2] - c[0]
                        area_sub = abs(d1[0] * d2[1] - d1[1] * d2[0])
                        r = x - y
                        k = r @ n / np.linalg.norm(r) ** 3
                        out += wa * wb * jac * area_sub * k * lam
    area = 0.5 * np.linalg.norm(np.cross(tri[1] - tri[0], tri[2] - tri[0]))
    # d(area) = 2 |T| d(l1) d(l2)
    return out * 2.0 * area / (4 * math.pi)


@pytest.mark.parametrize("x", [(0.3, 0.2, 1.7), (-1.1, 2.0, -0.6), (2.5, -0.4, 0.9)])
def test_lindholm_weights_brute_force(x):
    tri, n, area, s, eta, gamma = _one_triangle((0.0, 0.0, 0.0), (1.0, 0.1, 0.0), (0.2, 0.9, 0.3))
    x = np.array(x)
    w = np.array(O.lindholm_weights(x, tri[0], tri[1], tri[2], n, area, s, eta, gamma))
    q = _quadrature_weights(x, tri, n)
    assert np.abs(w - q).max() <= 1e-11 * max(np.abs(q).max(), 1e-3)


def test_lindholm_weights_sum_and_scale():
    tri, n, area, s, eta, gamma = _one_triangle((0.0, 0.0, 0.0), (1.0, 0.0, 0.2), (0.3, 1.2, 0.0))
    rng = np.random.default_rng(0)
    for _ in range(20):
        x = rng.normal(size=3) * 2
        w = np.array(O.lindholm_weights(x, tri[0], tri[1], tri[2], n, area, s, eta, gamma))
        om = O.solid_angle(x, tri[0], tri[1], tri[2])
        assert abs(w.sum() + om / (4 * math.pi)) < 1e-14
        # scale invariance (nm <-> m): weights are dimensionless
        k = 1e-9
        t2, n2, a2, s2, e2, g2 = _one_triangle(*(tri * k))
        w2 = np.array(O.lindholm_weights(x * k, t2[0], t2[1], t2[2], n2, a2, s2, e2, g2))
        assert np.abs(w2 - w).max() < 1e-12


# -------------------------------------------------------------- boundary data
def test_boundary_counts_and_omega(c1):
    m, bd = c1
    assert (bd.Nb, bd.Nt) == (602, 1200)
    assert m.num_nodes == 1331
    vals = np.unique(np.round(bd.omega_sum / math.pi, 12))
    assert set(vals.tolist()) == {0.5, 1.0, 2.0}
    assert (np.round(bd.omega_sum / math.pi, 12) == 0.5).sum() == 8
    # outward normals: sum_t n_t A_t = 0 and sum_t (c_t . n_t) A_t = 3 V
    c = bd.tri_pts.mean(axis=1)
    assert np.abs((bd.ntri * bd.area[:, None]).sum(axis=0)).max() < 1e-9
    assert abs((np.einsum("ij,ij->i", c, bd.ntri) * bd.area).sum() / 3 - 40.0 ** 3) < 1e-6


def test_dense_B_rows_sum_to_minus_one(c1):
    _, bd = c1
    B = O.dense_B_rows(bd)
    assert np.abs(B.sum(axis=1) + 1.0).max() < 1e-13
    # block fill = off-diagonal restriction of the dense rows
    rows = np.arange(0, 602, 7)
    cols = np.arange(3, 602, 5)
    blk = O.block_B(bd, rows, cols)
    ref = B[np.ix_(rows, cols)].copy()
    ref[rows[:, None] == cols[None, :]] = 0.0
    assert np.abs(blk - ref).max() < 1e-15


@pytest.mark.slow
def test_boundary_counts_c2_c3():
    from smx.boundary import build_boundary_geometry
    for mk, nb, nt in ((M.cube_c2, 5402, 10800),):
        m = mk()
        bd = O.boundary_data(m.X, m.tets)
        assert (bd.Nb, bd.Nt) == (nb, nt)
    m3 = M.film_c3()
    g3 = build_boundary_geometry(m3.X, m3.tets)
    assert (m3.num_nodes, g3.Nb, g3.Nt) == (202005, 83202, 166400)


# ---------------------------------------------------------------- FEM operators
def test_fem_invariants(c1):
    m, _ = c1
    f = O.fem_operators(m.X, m.tets, 8e5)
    N = f.N
    assert np.abs(f.A @ np.ones(N)).max() < 1e-12
    assert abs(f.V.sum() - 40.0 ** 3) < 1e-8
    # interior row of a Kuhn grid: h * (6, -1 x 6)
    h = 4.0
    i = 5 + 11 * (5 + 11 * 5)
    row = f.A.getrow(i).toarray().ravel()
    nz = np.abs(row) > 1e-12
    assert nz.sum() == 7 and abs(row[i] - 6 * h) < 1e-12
    for d in (1, 11, 121):
        assert abs(row[i + d] + h) < 1e-12 and abs(row[i - d] + h) < 1e-12
    # G (linear phi) = -grad(phi) V_i for every node
    g = np.array([0.3, -1.2, 0.7])
    phi = m.X @ g + 2.0
    Gphi = (f.G @ phi).reshape(-1, 3)
    assert np.abs(Gphi + g[None, :] * f.V[:, None]).max() < 1e-9
    # Div = -Ms G^T
    assert abs(f.Div + 8e5 * f.G.T).max() < 1e-6
    # Dirichlet matrix: identity rows on the boundary
    Abc = f.A_bc.tolil()
    for b in f.bnodes[:20]:
        r = Abc.getrow(b).toarray().ravel()
        assert r[b] == 1.0 and np.abs(r).sum() == 1.0


def test_demag_patch_tests(c1):
    m, bd = c1
    dem = O.DemagOracle(m.X, m.tets, 8e5, bd=bd)
    # u1 for uniform m is Ms m.x up to a constant
    mz = np.tile([0.0, 0.0, 1.0], m.num_nodes)
    u1 = dem.solve_u1(mz)
    z = m.X[:, 2]
    assert np.abs(u1 - 8e5 * (z - z.mean())).max() < 1e-6 * 8e5
    # u2 with linear Dirichlet data is that linear function
    L = 0.5 * m.X[:, 0] - 0.25 * m.X[:, 1] + 3.0
    g = np.zeros(m.num_nodes)
    g[dem.fem.bnodes] = L[dem.fem.bnodes]
    u2 = dem.solve_u2(g)
    assert np.abs(u2 - L).max() < 1e-10
    # constant shift of u1 leaves H unchanged (B 1 = -1)
    H = dem.compute(mz)
    u1s = dem.solve_u1(mz) + 5.0
    H2 = (dem.fem.G @ (u1s + dem.solve_u2(dem.boundary_rhs(u1s)))) / np.repeat(dem.fem.V, 3)
    assert np.linalg.norm(H2 - H) / np.linalg.norm(H) < 1e-10
    # uniformly magnetized cube: volume average close to -Ms/3
    avg = (H.reshape(-1, 3)[:, 2] * dem.fem.V).sum() / dem.fem.V.sum() / 8e5
    assert -0.40 < avg < -0.28


# ------------------------------------------------------------ exchange / DMI
def _linear_m(X):
    A = np.array([[0.01, -0.02, 0.005], [0.003, 0.01, -0.01], [-0.004, 0.002, 0.015]])
    return (X @ A.T + np.array([0.1, 0.2, 0.9])).ravel(), A


def test_exchange_kernel_and_laplacian(c1):
    m, _ = c1
    f = O.fem_operators(m.X, m.tets, 1.0)
    Ms, A = 8e5, 1.3e-11
    K = O.exchange_matrix(m.X, m.tets, f.V, A, Ms)
    assert np.abs(K @ np.tile([0.3, -0.4, 0.5], m.num_nodes)).max() < 1e-6
    rng = np.random.default_rng(0)
    mv = rng.normal(size=3 * m.num_nodes)
    H = (K @ mv).reshape(-1, 3)
    mm = mv.reshape(-1, 3)
    i = 4 + 11 * (6 + 11 * 3)
    lap = (mm[i + 1] + mm[i - 1] + mm[i + 11] + mm[i - 11] + mm[i + 121] + mm[i - 121] - 6 * mm[i]) / 16.0
    expect = 2 * A / (O.MU0 * Ms) * lap / 1e-18
    assert np.abs(H[i] - expect).max() < 1e-9 * np.abs(expect).max()


def test_dmi_linear_field_interior(c1):
    m, _ = c1
    f = O.fem_operators(m.X, m.tets, 1.0)
    Ms, D = 8e5, 3e-3
    mv, G = _linear_m(m.X)  # dm_i/dx_j = G[i, j]
    nrm = np.array([0.0, 0.0, 1.0])
    Ki = O.dmi_interfacial_matrix(m.X, m.tets, f.V, D, Ms, nrm)
    Kb = O.dmi_bulk_matrix(m.X, m.tets, f.V, D, Ms)
    Hi = (Ki @ mv).reshape(-1, 3)
    Hb = (Kb @ mv).reshape(-1, 3)
    div = np.trace(G)
    grad_mn = G.T @ nrm  # grad(m.n)
    Hi_exp = -2 * D / (O.MU0 * Ms) * (nrm * div - grad_mn) * 1e9
    curl = np.array([G[2, 1] - G[1, 2], G[0, 2] - G[2, 0], G[1, 0] - G[0, 1]])
    Hb_exp = -2 * D / (O.MU0 * Ms) * curl * 1e9
    interior = np.setdiff1d(np.arange(m.num_nodes), f.bnodes)
    assert np.abs(Hi[interior] - Hi_exp).max() < 1e-9 * np.abs(Hi_exp).max()
    assert np.abs(Hb[interior] - Hb_exp).max() < 1e-9 * np.abs(Hb_exp).max()


def test_anisotropy_closed_form():
    mv = np.array([0.6, 0.0, 0.8, 0.0, 1.0, 0.0])
    H = O.anisotropy_field(mv, (0.0, 0.0, 2.0), 5e5, 8e5).reshape(-1, 3)
    pref = 2 * 5e5 / (O.MU0 * 8e5)
    assert np.allclose(H[0], [0, 0, pref * 0.8], rtol=1e-15)
    assert np.allclose(H[1], 0.0)


# ------------------------------------------------------------------ LLG RHS
def test_llg_rhs_invariants():
    rng = np.random.default_rng(1)
    mv = rng.normal(size=30)
    H = rng.normal(size=30) * 1e5
    g, a, Ms = 2.211e5, 0.5, 8e5
    r = O.llg_rhs(mv, H, g, a, Ms).reshape(-1, 3)
    m3 = mv.reshape(-1, 3)
    n2 = (m3 ** 2).sum(axis=1)
    S = Ms * g / (1 + a * a) * 0.5
    assert np.allclose((m3 * r).sum(axis=1), S * (1 - n2) * n2, rtol=1e-10, atol=1e-3)
    mu = m3 / np.sqrt(n2)[:, None]
    r2 = O.llg_rhs(mu.ravel(), (mu * 3e4).ravel(), g, a, Ms)
    assert np.abs(r2).max() < 1e-13 * g * 3e4
    # pure precession: |dm/dt| = gamma/(1+a^2) |m x H| for a = 0, do_precess = 1
    r0 = O.llg_rhs(mu.ravel(), H, g, 0.0, Ms).reshape(-1, 3)
    assert np.allclose(np.linalg.norm(r0, axis=1), g * np.linalg.norm(np.cross(mu, H.reshape(-1, 3)), axis=1))


def test_jac_vec_matches_finite_differences():
    rng = np.random.default_rng(2)
    n = 12
    K = rng.normal(size=(3 * n, 3 * n)) * 1e4
    mv = rng.normal(size=3 * n)
    v = rng.normal(size=3 * n)
    g, a, Ms = 2.211e5, 0.3, 8e5
    f = lambda x: O.llg_rhs(x, K @ x, g, a, Ms)
    h = 1e-6
    fd = (f(mv + h * v) - f(mv - h * v)) / (2 * h)
    jv = O.jac_vec(mv, K @ mv, v, K @ v, g, a, Ms)
    assert np.linalg.norm(jv - fd) / np.linalg.norm(fd) < 1e-8


def test_dp5_macrospin_analytic():
    """Single spin in a static field: tan(theta/2) = tan(theta0/2) exp(-a g' H t),
    phi = phi0 + g' H t with g' = gamma/(1+a^2)."""
    g, a, Ms, Hz = 2.211e5, 0.1, 8e5, 1e5
    H = np.array([0.0, 0.0, Hz])
    th0, ph0 = 1.0, 0.3
    m0 = np.array([math.sin(th0) * math.cos(ph0), math.sin(th0) * math.sin(ph0), math.cos(th0)])
    f = lambda t, y: O.llg_rhs(y, H, g, a, Ms)
    y = m0.copy()
    dt, nst = 2e-13, 200
    k1 = None
    for _ in range(nst):
        y, _, k1 = O.dp5_step(f, 0.0, y, dt, k1)
    T = dt * nst
    gp = g / (1 + a * a)
    th = 2 * math.atan(math.tan(th0 / 2) * math.exp(-a * gp * Hz * T))
    ph = ph0 + gp * Hz * T
    exact = np.array([math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)])
    assert np.abs(y - exact).max() < 1e-9


# ------------------------------------------------------ tree / ACA / H-matrix
def test_cluster_tree_invariants(c1):
    _, bd = c1
    perm, root = O.cluster_tree(bd.Xb, leaf_size=16)
    assert np.array_equal(np.sort(perm), np.arange(bd.Nb))
    stack = [root]
    while stack:
        c = stack.pop()
        pts = bd.Xb[perm[c.start:c.stop]]
        assert np.allclose(pts.min(axis=0), c.bmin) and np.allclose(pts.max(axis=0), c.bmax)
        if c.is_leaf:
            assert c.size <= 16
        else:
            assert c.left.start == c.start and c.left.stop == c.right.start and c.right.stop == c.stop
            assert c.left.size == c.size // 2
            stack += [c.left, c.right]


def test_admissibility_brute_force(c1):
    _, bd = c1
    perm, root = O.cluster_tree(bd.Xb, leaf_size=32)
    leaves = []
    stack = [root]
    while stack:
        c = stack.pop()
        (leaves.append(c) if c.is_leaf else stack.extend([c.left, c.right]))
    for a in leaves[:6]:
        for b in leaves[::3]:
            pa, pb = bd.Xb[perm[a.start:a.stop]], bd.Xb[perm[b.start:b.stop]]
            dmin = np.sqrt(((pa[:, None] - pb[None]) ** 2).sum(-1)).min()
            gap = np.maximum(np.maximum(a.bmin - b.bmax, b.bmin - a.bmax), 0)
            d = np.linalg.norm(gap)
            assert d <= dmin + 1e-12
            exp = d > 0 and max(np.linalg.norm(a.bmax - a.bmin), np.linalg.norm(b.bmax - b.bmin)) <= 2.0 * d
            assert O.admissible(a, b, 2.0) == exp


def test_fullaca_exact_rank_and_tolerance(c1):
    rng = np.random.default_rng(0)
    A = rng.normal(size=(40, 3)) @ rng.normal(size=(3, 50))
    U, V, rel = O.fullaca_dense(A, rtol=1e-12)
    assert U.shape[1] == 3 and np.abs(U @ V - A).max() < 1e-12 * np.abs(A).max()
    _, bd = c1
    perm, root = O.cluster_tree(bd.Xb, leaf_size=64)
    blk = O.block_B(bd, perm[root.left.left.start:root.left.left.stop],
                    perm[root.right.right.start:root.right.right.stop])
    U, V, rel = O.fullaca_dense(blk, rtol=1e-6)
    assert np.linalg.norm(blk - U @ V) <= 1e-6 * np.linalg.norm(blk) * (1 + 1e-12)
    assert U.shape[1] < min(blk.shape)


def test_hmatrix_validate_against_dense(c1):
    """validate_against_dense protocol (integration.py:305-329): seed 1234, 3 trials."""
    _, bd = c1
    B = O.dense_B_rows(bd)
    H = O.build_hmatrix(bd, eps=1e-6, eta=2.0, leaf_size=32)
    assert H.stats["low_rank_blocks"] > 0
    rng = np.random.default_rng(1234)
    for _ in range(3):
        x = rng.standard_normal(bd.Nb)
        assert np.linalg.norm(H.matvec(x) - B @ x) / np.linalg.norm(B @ x) < 1e-5
