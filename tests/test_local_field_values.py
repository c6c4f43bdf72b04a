"""Local-field operators: the product's node-graph values against the oracle's
assembled matrices (host, CPU), and pins of the oracle's bulk-DMI form that
involve its boundary term.

Bulk DMI (DMI_Bulk_GPU.py:43-86, 248-264): K(u, v) = -int curl u . v dx
+ 1/2 int (nu x u) . v ds.  With the identity
int curl u . v = int u . curl v + int_S (nu x u) . v, the form equals
-1/2 (int curl u . v + int u . curl v): symmetric, and K(m, m) = -int m . curl m.
For a linear m both sides are exact in P1, and int m . curl m has a closed form
on a box.  A dropped or sign-flipped boundary term breaks both checks (shown
below by planting it in a copy).
"""
import numpy as np
import pytest
import scipy.sparse as sp

import smx_oracle as O
from smx import mesh as M
from smx.fem import assemble_fem, dmi_bulk_values, dmi_interfacial_values, exchange_values

MS, D, A = 8e5, 3e-3, 1.3e-11


@pytest.fixture(scope="module", params=["cube", "jitter"])
def case(request):
    m = M.kuhn_box(6, 5, 4, 24.0, 20.0, 16.0)
    X = m.X.copy()
    if request.param == "jitter":
        rng = np.random.default_rng(5)
        lo, hi = X.min(axis=0), X.max(axis=0)
        inner = np.all((X > lo) & (X < hi), axis=1)
        X[inner] += rng.uniform(-0.6, 0.6, size=(int(inner.sum()), 3))
    fem = assemble_fem(X, m.tets)
    return X, m.tets, fem


def _graph_matrix(fem, blocks):
    """(nnz, 3, 3) node-graph blocks -> 3N x 3N CSR."""
    g = fem.graph
    r = np.repeat(np.arange(fem.N), np.diff(g.indptr))
    rows = (3 * r[:, None, None] + np.arange(3)[None, :, None]) * np.ones((1, 1, 3), dtype=np.int64)
    cols = (3 * g.indices[:, None, None] + np.arange(3)[None, None, :]) * np.ones((1, 3, 1), dtype=np.int64)
    return sp.csr_matrix((blocks.ravel(), (rows.ravel(), cols.ravel())), shape=(3 * fem.N, 3 * fem.N))


def _rel(Kp, Ko):
    d = (Kp - Ko).tocsr()
    return (np.abs(d.data).max() if d.nnz else 0.0) / np.abs(Ko.data).max()


def test_dmi_bulk_values_match_oracle(case):
    X, tets, fem = case
    Kp = _graph_matrix(fem, dmi_bulk_values(X, tets, fem, D, MS))
    Ko = O.dmi_bulk_matrix(X, tets, fem.volume, D, MS)
    assert _rel(Kp, Ko) <= 1e-13


def test_dmi_interfacial_values_match_oracle(case):
    X, tets, fem = case
    n0 = (0.0, 0.0, 1.0)
    Kp = _graph_matrix(fem, dmi_interfacial_values(X, tets, fem, D, MS, n0))
    Ko = O.dmi_interfacial_matrix(X, tets, fem.volume, D, MS, n0)
    assert _rel(Kp, Ko) <= 1e-13


def test_exchange_values_match_oracle(case):
    X, tets, fem = case
    ex = exchange_values(fem, A, MS)
    Kp = _graph_matrix(fem, ex[:, None, None] * np.eye(3)[None])
    Ko = O.exchange_matrix(X, tets, fem.volume, A, MS)
    assert _rel(Kp, Ko) <= 1e-13


def _unscaled_bulk(X, tets, V, boundary_sign=1.0):
    """The oracle's bulk-DMI form without the 2D/(mu0 Ms) 1e9 / V_i scaling;
    boundary_sign plants a mistake in the boundary term (0: dropped, -1:
    flipped) by recombining the oracle's volume and boundary parts."""
    K = O.dmi_bulk_matrix(X, tets, V, D, MS)
    pref = np.repeat(2.0 * D / (O.MU0 * MS) * 1e9 / V, 3)
    K0 = sp.diags(1.0 / pref) @ K
    if boundary_sign == 1.0:
        return K0.tocsr()
    # the volume part alone: the form on a mesh with no boundary facets is
    # obtained from the symmetric identity; K_vol = K0 - K_bnd, K_bnd = 1/2 int (nu x u).v
    faces, nu, farea = O._boundary_tris_outward(X, tets)
    mass = (np.ones((3, 3)) + np.eye(3)) / 12.0
    nxd = np.einsum("cid,fi->fcd", O._EPS, nu)
    Kb = 0.5 * np.einsum("f,ab,fcd->fabcd", farea, mass, nxd)
    rows = np.broadcast_to(3 * faces[:, :, None, None, None] + np.arange(3)[None, None, None, :, None], Kb.shape)
    cols = np.broadcast_to(3 * faces[:, None, :, None, None] + np.arange(3)[None, None, None, None, :], Kb.shape)
    n3 = 3 * X.shape[0]
    Kbnd = sp.csr_matrix((Kb.ravel(), (rows.ravel(), cols.ravel())), shape=(n3, n3))
    return (K0 - Kbnd + boundary_sign * Kbnd).tocsr()


def _linear_field(X, seed):
    rng = np.random.default_rng(seed)
    c = rng.normal(size=3)
    G = rng.normal(size=(3, 3)) / 10.0
    return (c[None, :] + X @ G.T).ravel(), c, G


def _m_dot_curl_m(X, c, G):
    """int_box m . curl m for m = c + G x on the axis-aligned box of X."""
    lo, hi = X.min(axis=0), X.max(axis=0)
    vol = float(np.prod(hi - lo))
    w = np.array([G[2, 1] - G[1, 2], G[0, 2] - G[2, 0], G[1, 0] - G[0, 1]])
    xbar = 0.5 * (lo + hi)
    return float(w @ (c + G @ xbar)) * vol


def test_bulk_dmi_symmetric_and_energy_closed_form():
    m = M.kuhn_box(6, 5, 4, 24.0, 20.0, 16.0)
    fem = assemble_fem(m.X, m.tets)
    K0 = _unscaled_bulk(m.X, m.tets, fem.volume)
    asym = abs(K0 - K0.T).max()
    assert asym <= 1e-13 * abs(K0).max()
    for seed in range(3):
        mv, c, G = _linear_field(m.X, seed)
        exact = _m_dot_curl_m(m.X, c, G)
        assert float(mv @ (K0 @ mv)) == pytest.approx(-exact, rel=1e-11)


@pytest.mark.parametrize("sign", [0.0, -1.0])
def test_bulk_dmi_pins_catch_a_boundary_slip(sign):
    m = M.kuhn_box(6, 5, 4, 24.0, 20.0, 16.0)
    fem = assemble_fem(m.X, m.tets)
    K0 = _unscaled_bulk(m.X, m.tets, fem.volume, boundary_sign=sign)
    mv, c, G = _linear_field(m.X, 0)
    exact = _m_dot_curl_m(m.X, c, G)
    sym_ok = abs(K0 - K0.T).max() <= 1e-13 * abs(K0).max()
    energy_ok = abs(float(mv @ (K0 @ mv)) + exact) <= 1e-11 * abs(exact)
    assert not (sym_ok and energy_ok)
