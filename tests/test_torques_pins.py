"""Pins of the oracle's cubic anisotropy, Zhang-Li / spin-orbit torques and
minimizer step, and the product's host-side Zhang-Li values against it.

* cubic field = -(1/(mu0 Ms)) dE/dm by central differences of the cubic
  energy density (pins the field formula against the energy formula);
  closed forms at m along an axis and along a face diagonal; the JVP by
  central differences of the field;
* Zhang-Li operator: for a linear m and uniform J, (K_J m)/V = (J.grad) m
  exactly at every node (P1 exactness); with a nodal J the degree-2 rule
  equals the element closed form |T|/20 (sum_c J_c + J_a).grad phi_b; the
  product's zhang_li_values equal the oracle matrix;
* torques: orthogonal to m for |m| = 1; a damping-like spin-orbit torque with
  H_dl > 0 turns a macrospin toward p; m parallel to p feels no SOT;
* Cayley step: |y| = |m| and y is the rotation of m by 2 atan|c| about c
  (Rodrigues' formula); y = m + step d + O(step^2) for tangent d.
"""
import math

import numpy as np
import pytest
import scipy.sparse as sp

import smx_oracle as O
from smx import mesh as M

MU0 = 4e-7 * math.pi


def _axes():
    return O.cubic_axes_oracle([1.0, 1.0, 0.0], [0.0, 0.3, 1.0])


def test_cubic_field_is_minus_energy_gradient():
    rng = np.random.default_rng(0)
    axes = _axes()
    K1, Ms = 4.5e4, 8e5
    m = rng.standard_normal((20, 3))
    V = np.ones(20) * 1e27          # unit volume after the 1e-27 scale
    H = O.cubic_field(m, axes, K1, Ms).reshape(-1, 3)
    h = 1e-6
    for i in range(5):
        for k in range(3):
            mp, mm = m.copy(), m.copy()
            mp[i, k] += h
            mm[i, k] -= h
            dE = (O.cubic_energy(mp, axes, K1, V) - O.cubic_energy(mm, axes, K1, V)) / (2 * h)
            # E = -K1 sum S, H = -(1/(mu0 Ms)) dE/dm
            assert abs(H[i, k] - (-dE / (MU0 * Ms))) <= 1e-7 * (abs(H[i, k]) + 1e-3)


def test_cubic_closed_forms_and_jvp():
    axes = _axes()
    c1, c2, c3 = (a[0] for a in axes)
    K1, Ms = 4.5e4, 8e5
    pref = 2 * K1 / (MU0 * Ms)
    assert np.allclose(O.cubic_field(c1, axes, K1, Ms), 0.0, atol=1e-12 * pref)
    m = (c1 + c2) / math.sqrt(2)
    # a1 = a2 = 1/sqrt2, a3 = 0: s1 = s2 = 1/(2 sqrt2), s3 = 0
    assert np.allclose(O.cubic_field(m, axes, K1, Ms), pref / (2 * math.sqrt(2)) * (c1 + c2), rtol=1e-14)
    assert abs(O.cubic_energy(m, axes, K1, np.array([1e27])) + K1 / 4) <= 1e-14 * K1
    rng = np.random.default_rng(1)
    mm, v = rng.standard_normal((7, 3)), rng.standard_normal((7, 3))
    h = 1e-6
    fd = (O.cubic_field(mm + h * v, axes, K1, Ms) - O.cubic_field(mm - h * v, axes, K1, Ms)) / (2 * h)
    jv = O.cubic_field_jvp(mm, v, axes, K1, Ms)
    assert np.linalg.norm(jv - fd) <= 1e-8 * np.linalg.norm(fd)


@pytest.mark.parametrize("jitter", [0.0, 0.4])
def test_zhang_li_linear_field_is_exact(jitter):
    m = M.kuhn_box(6, 5, 4, 18.0, 15.0, 12.0, name="b")
    X = m.X + (np.random.default_rng(2).uniform(-jitter, jitter, m.X.shape) if jitter else 0.0)
    fem = O.fem_operators(X, m.tets, 1.0)
    J = np.array([0.3, -1.2, 0.7])
    A = np.array([[0.2, -0.1, 0.4], [1.0, 0.3, -0.2], [-0.5, 0.6, 0.1]])   # m(x) = A x + b
    mv = X @ A.T + np.array([0.1, 0.2, 0.3])
    K = O.zhang_li_matrix(X, m.tets, fem.V, J)
    Z = np.column_stack([K @ mv[:, k] for k in range(3)])
    assert np.allclose(Z, np.broadcast_to(A @ J, Z.shape), rtol=0, atol=1e-12 * np.abs(A @ J).max())


def test_zhang_li_nodal_current_and_product_values():
    from smx.fem import assemble_fem, zhang_li_values
    m = M.kuhn_box(5, 4, 3, 15.0, 12.0, 9.0, name="b")
    X = m.X + np.random.default_rng(3).uniform(-0.3, 0.3, m.X.shape)
    fem = assemble_fem(X, m.tets)
    Jn = np.random.default_rng(4).standard_normal((fem.N, 3))
    K = O.zhang_li_matrix(X, m.tets, fem.volume, Jn).tocsr()
    # element closed form, assembled independently of the quadrature rule
    vol, g = O.element_geometry(X, m.tets)
    Jt = Jn[m.tets]
    w = (vol / 20.0)[:, None, None] * (Jt.sum(axis=1)[:, None, :] + Jt)
    elem = np.einsum("tak,tbk->tab", w, g)
    rows = np.repeat(m.tets[:, :, None], 4, axis=2)
    cols = np.repeat(m.tets[:, None, :], 4, axis=1)
    K2 = sp.diags(1.0 / fem.volume) @ sp.coo_matrix((elem.ravel(), (rows.ravel(), cols.ravel())),
                                                    shape=(fem.N, fem.N)).tocsr()
    assert abs(K - K2).max() <= 1e-13 * abs(K).max()
    # the product's graph values
    gv = zhang_li_values(fem, m.tets, Jn)
    r = np.repeat(np.arange(fem.N), np.diff(fem.graph.indptr))
    Kp = sp.csr_matrix((gv, (r, fem.graph.indices)), shape=(fem.N, fem.N))
    assert abs(Kp - K).max() <= 1e-13 * abs(K).max()


def test_torques_orthogonal_and_sot_turns_toward_p():
    rng = np.random.default_rng(5)
    m = rng.standard_normal((30, 3))
    m /= np.linalg.norm(m, axis=1)[:, None]
    H = np.zeros_like(m)
    Z = rng.standard_normal((30, 3))
    p = np.array([0.0, 1.0, 0.0])
    r = O.llg_rhs_torques(m.ravel(), H.ravel(), 2.211e5, 0.3, 8e5, Z=Z, stt=(1e12, 0.6, 0.05),
                          p=p, H_dl=1e4, H_fl=3e3).reshape(-1, 3)
    assert np.abs(np.einsum("ij,ij->i", m, r)).max() <= 1e-9 * np.abs(r).max()
    # damping-like SOT only: dm/dt = c_dl m x (m x p) = -c_dl (p - (m.p) m), c_dl < 0
    r = O.llg_rhs_torques(m.ravel(), H.ravel(), 2.211e5, 0.0, 8e5, p=p, H_dl=1e4).reshape(-1, 3)
    tang = p[None, :] - (m @ p)[:, None] * m
    assert np.all(np.einsum("ij,ij->i", r, tang) > 0.0)
    assert np.allclose(r, 2.211e5 * 1e4 * tang, rtol=1e-12, atol=1e-3)
    mp = np.tile(p, (3, 1))
    r = O.llg_rhs_torques(mp.ravel(), np.zeros(9), 2.211e5, 0.2, 8e5, p=p, H_dl=1e4, H_fl=2e3)
    assert np.abs(r).max() <= 1e-9


def test_cayley_step_is_a_rodrigues_rotation():
    rng = np.random.default_rng(6)
    m = rng.standard_normal((25, 3))
    d = rng.standard_normal((25, 3))
    step = 0.37
    y = O.cayley_step(m.ravel(), d.ravel(), step).reshape(-1, 3)
    assert np.allclose(np.linalg.norm(y, axis=1), np.linalg.norm(m, axis=1), rtol=1e-14)
    c = 0.5 * step * np.cross(m, d)
    for i in range(25):
        cn = np.linalg.norm(c[i])
        k = c[i] / cn
        th = 2 * math.atan(cn)
        rot = m[i] * math.cos(th) + np.cross(k, m[i]) * math.sin(th) + k * (k @ m[i]) * (1 - math.cos(th))
        assert np.allclose(y[i], rot, rtol=1e-13, atol=1e-13)
    # small step along a tangent direction: y = m + step d + O(step^2)
    mu = m / np.linalg.norm(m, axis=1)[:, None]
    dt = O.projected_direction(mu.ravel(), d.ravel()).reshape(-1, 3)
    s = 1e-5
    ys = O.cayley_step(mu.ravel(), dt.ravel(), s).reshape(-1, 3)
    assert np.abs(ys - (mu + s * dt)).max() <= 10 * s * s * np.abs(dt).max() ** 2
