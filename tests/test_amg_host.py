"""Smoothed-aggregation hierarchy (host set-up of the AMG-CG solves).

* the folded V-cycle M^-1 b = S b + Q M_c^-1 R b equals a textbook V-cycle
  written out step by step here (damped-Jacobi pre-smoothing from zero,
  residual, restriction by P^T, Galerkin coarse operator P^T A P, coarse
  correction, prolongation, damped-Jacobi post-smoothing) on every level;
* M^-1 is symmetric positive definite (CG needs it);
* aggregation: every node with a strong neighbour is aggregated, Dirichlet
  identity rows are not, aggregates are connected in the strength graph;
* the constants stay in the nullspace of every coarse Neumann operator;
* host AMG-PCG converges in a handful of iterations where Jacobi needs many;
* the solver entry point rejects inconsistent level shapes before touching
  the device (pure host check, so it runs without a GPU).
"""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import connected_components

from smx import mesh as M
from smx.amg import (aggregate, apply_vcycle, build_levels, lanes_per_row, sell_layout, sell_matvec,
                     spectral_radius_dinv_a, strength_graph)
from smx.fem import assemble_fem, compact_csr


def _op(nx, ny, nz, bc=False, jitter=0.0):
    m = M.kuhn_box(nx, ny, nz, 3.0 * nx, 3.0 * ny, 3.0 * nz, name="b")
    X = m.X
    if jitter:
        X = X + np.random.default_rng(3).uniform(-jitter, jitter, X.shape)
    fem = assemble_fem(X, m.tets)
    vals = fem.stiffness_bc() if bc else fem.stiff
    ip, ix, vv = compact_csr(fem.graph.indptr, fem.graph.indices, vals)
    A = sp.csr_matrix((vv, ix, ip), shape=(fem.N, fem.N))
    A.sort_indices()
    return A, fem


def _tentative(agg, na, B):
    """T with column j = B on aggregate j / its norm (loop form)."""
    n = agg.size
    T = np.zeros((n, na))
    Bc = np.zeros(na)
    for j in range(na):
        idx = np.nonzero(agg == j)[0]
        Bc[j] = np.linalg.norm(B[idx])
        T[idx, j] = B[idx] / Bc[j]
    return sp.csr_matrix(T), Bc


def _textbook_vcycle(A, b, depth, theta=None, B=None, level=0):
    """V(nu,nu) cycle with damped Jacobi, written in its usual form: weight
    SMOOTH_WEIGHT / rho, one sweep before and after the coarse correction on
    the finest level and COARSE_SWEEPS on the coarser ones, the prolongator
    smoothed with 4/3 / rho, `depth` coarsenings and a minimum-norm
    least-squares coarse solve."""
    from smx.amg import COARSE_SWEEPS, SMOOTH_WEIGHT, THETA
    theta = THETA if theta is None else theta
    n = A.shape[0]
    if depth == 0:
        return np.linalg.lstsq(A.toarray(), b, rcond=1e-11)[0]
    if B is None:
        B = np.ones(n)
    agg, na = aggregate(strength_graph(A, theta))
    T, Bc = _tentative(agg, na, B)
    rho = spectral_radius_dinv_a(A)
    w = SMOOTH_WEIGHT / rho
    dinv = 1.0 / A.diagonal()
    P = T - (4.0 / (3.0 * rho)) * sp.diags(dinv) @ (A @ T)
    nu = 1 if level == 0 else COARSE_SWEEPS
    x = np.zeros(n)
    for _ in range(nu):                               # pre-smoothing from x = 0
        x = x + w * dinv * (b - A @ x)
    r = b - A @ x                                     # residual
    Ac = (P.T @ A @ P).tocsr()
    Ac = (Ac + Ac.T) * 0.5
    e = _textbook_vcycle(Ac, P.T @ r, depth - 1, theta, Bc, level + 1)
    x = x + P @ e                                     # coarse correction
    for _ in range(nu):                               # post-smoothing
        x = x + w * dinv * (b - A @ x)
    return x


@pytest.mark.parametrize("bc", [False, True])
@pytest.mark.parametrize("jitter", [0.0, 0.6])
def test_folded_vcycle_is_the_textbook_vcycle(bc, jitter):
    A, _ = _op(16, 14, 10, bc=bc, jitter=jitter)
    levels = build_levels(A, max_coarse=20)
    assert len(levels) >= (3 if jitter == 0.0 else 2)
    rng = np.random.default_rng(0)
    for _ in range(3):
        b = rng.standard_normal(A.shape[0])
        got = apply_vcycle(levels, b)
        ref = _textbook_vcycle(A, b, len(levels) - 1)
        assert np.linalg.norm(got - ref) <= 1e-11 * np.linalg.norm(ref)


def test_folded_operators_shapes_and_transpose():
    A, _ = _op(12, 10, 8)
    levels = build_levels(A, max_coarse=40)
    for l, L in enumerate(levels[:-1]):
        nxt = levels[l + 1].n
        assert L.S.shape == (L.n, L.n) and L.Q.shape == (L.n, nxt) and L.R.shape == (nxt, L.n)
        assert abs(L.R - L.Q.T).max() == 0.0
        assert abs(L.S - L.S.T).max() <= 1e-15 * abs(L.S).max()
    c = levels[-1]
    assert c.Q is None and c.S.nnz == c.n * c.n       # dense coarse rows


@pytest.mark.parametrize("bc", [False, True])
def test_preconditioner_is_spd(bc):
    A, _ = _op(9, 8, 7, bc=bc, jitter=0.5)
    levels = build_levels(A, max_coarse=30)
    n = A.shape[0]
    Mi = np.column_stack([apply_vcycle(levels, e) for e in np.eye(n)])
    assert np.abs(Mi - Mi.T).max() <= 1e-12 * np.abs(Mi).max()
    ev = np.linalg.eigvalsh(0.5 * (Mi + Mi.T))
    assert ev.min() > 0.0


def test_aggregation_properties():
    A, fem = _op(12, 10, 8, bc=True)
    S = strength_graph(A)
    agg, na = aggregate(S)
    isolated = np.diff(S.indptr) == 0
    # Dirichlet identity rows have no strong neighbour and stay out
    assert np.array_equal(isolated, fem.is_boundary)
    assert np.all(agg[isolated] == -1) and np.all(agg[~isolated] >= 0)
    assert np.array_equal(np.unique(agg[~isolated]), np.arange(na))
    for a in range(na):
        idx = np.nonzero(agg == a)[0]
        ncomp, _ = connected_components(S[idx][:, idx], directed=False)
        assert ncomp == 1
    assert na < (~isolated).sum() / 5                 # 3-D coarsening (small box: many faces)


def test_coarse_operators_keep_constants_in_the_nullspace():
    from smx.amg import THETA
    A, _ = _op(16, 14, 10)
    n = A.shape[0]
    assert abs(A @ np.ones(n)).max() <= 1e-12 * abs(A).max()
    Al = A
    B = np.ones(n)
    nlev = 0
    while Al.shape[0] > 20:
        agg, na = aggregate(strength_graph(Al, THETA))
        T, c = _tentative(agg, na, B)
        assert np.allclose(T @ c, B)                  # T interpolates the near-nullspace
        w = 4.0 / (3.0 * spectral_radius_dinv_a(Al))
        P = T - w * sp.diags(1.0 / Al.diagonal()) @ (Al @ T)
        Al = (P.T @ Al @ P).tocsr()
        assert abs(Al @ c).max() <= 1e-11 * abs(Al).max() * np.linalg.norm(c)
        B = c
        nlev += 1
    assert nlev >= 2
    levels = build_levels(A, max_coarse=20)
    assert len(levels) == nlev + 1
    Z = levels[-1].S.toarray()
    Ac_dense = Al.toarray()
    assert np.allclose(Z @ Ac_dense @ Z, Z, atol=1e-9 * abs(Z).max())
    assert abs(Z @ B).max() <= 1e-8 * abs(Z).max() * np.linalg.norm(B)   # pinv kills the nullspace


def _pcg(A, b, apply, rtol):
    x = np.zeros_like(b)
    r = b.copy()
    z = apply(r)
    p = z.copy()
    rz = r @ z
    for it in range(1, 5000):
        q = A @ p
        a = rz / (p @ q)
        x += a * p
        r -= a * q
        if np.linalg.norm(r) <= rtol * np.linalg.norm(b):
            return x, it
        z = apply(r)
        rz2 = r @ z
        p = z + (rz2 / rz) * p
        rz = rz2
    return x, 5000


@pytest.mark.parametrize("bc", [False, True])
def test_host_amg_pcg_iterations(bc):
    A, _ = _op(24, 20, 8, bc=bc)
    levels = build_levels(A, max_coarse=64)
    b = np.random.default_rng(1).standard_normal(A.shape[0])
    if not bc:
        b -= b.mean()
    x, its = _pcg(A, b, lambda r: apply_vcycle(levels, r), 1e-8)
    _, its_j = _pcg(A, b, lambda r: r / A.diagonal(), 1e-8)
    assert its <= 20 < its_j
    assert np.linalg.norm(b - A @ x) <= 1e-8 * np.linalg.norm(b)


def test_spectral_radius_matches_dense():
    A, _ = _op(6, 5, 4, jitter=0.4)
    d = 1.0 / np.sqrt(A.diagonal())
    lam = np.linalg.eigvalsh((sp.diags(d) @ A @ sp.diags(d)).toarray()).max()
    est = spectral_radius_dinv_a(A, iters=200)
    assert abs(est - lam) <= 1e-6 * lam


def test_lanes_per_row():
    eye = sp.identity(100, format="csr")
    assert lanes_per_row(eye) == 1
    dense = sp.csr_matrix(np.ones((10, 300)))
    assert lanes_per_row(dense) == 64
    assert lanes_per_row(sp.csr_matrix(np.ones((10, 7))), None) == 4


def test_solver_rejects_inconsistent_levels_on_the_host():
    """Shape validation happens before any device call (fake, never
    dereferenced addresses)."""
    from smx import _capi
    from smx.amg import AmgLevel, AmgOp as CsrOp
    L = _capi.lib()
    lv = (AmgLevel * 2)()
    fake = 0x1000
    lv[0].n = 100
    lv[0].S = CsrOp(fake, fake, fake, 100)
    lv[0].Q = CsrOp(fake, fake, fake, 100)
    lv[0].R = CsrOp(fake, fake, fake, 7)              # level 1 has 10 unknowns
    lv[0].s_tpr = lv[0].r_tpr = 4
    lv[1].n = 10
    lv[1].S = CsrOp(fake, fake, fake, 10)
    lv[1].s_tpr = 8
    lv[1].b = lv[1].x = fake
    its, rn, sums = C.c_int32(0), C.c_double(0.0), (C.c_double * 1)()
    args = [100, 0, 7, 128, fake, fake, 0, fake, None, 0, 0, 0, 0, None, None, None, None, fake, 0, None,
            C.addressof(lv), 2, _capi.SMX_F64, 1e-6, 0.0, 10, 8, 0, fake, C.byref(its), C.byref(rn), sums, None]
    assert L.smx_pcg_amg(*args) == _capi.SMX_ERR_SHAPE
    assert "level 0" in L.smx_last_error().decode()
    lv[0].R = CsrOp(fake, fake, fake, 10)
    lv[0].r_tpr = 3
    assert L.smx_pcg_amg(*args) == _capi.SMX_ERR_VALUE
    lv[0].r_tpr = 4
    args[22] = _capi.SMX_BF16
    assert L.smx_pcg_amg(*args) == _capi.SMX_ERR_TYPE
    args[22] = _capi.SMX_F64
    # trailing identity rows: mode 0 only without the mean projection, n >= 1
    args[1], args[10] = 5, 1
    assert L.smx_pcg_amg(*args) == _capi.SMX_ERR_VALUE
    assert "project" in L.smx_last_error().decode()
    args[0], args[10] = 0, 0
    assert L.smx_pcg_amg(*args) == _capi.SMX_ERR_SHAPE
    assert L.smx_amg_apply(C.addressof(lv), 2, _capi.SMX_F64, fake, None, fake, None) == _capi.SMX_ERR_VALUE


@pytest.mark.parametrize("tpr", [1, 4, 64])
def test_sliced_ell_layout(tpr):
    """The device slot layout, restated on the host, reproduces the CSR
    product for ragged rows, empty rows and a partial last slice."""
    rng = np.random.default_rng(0)
    n, m = 700, 300
    lens = rng.integers(0, 40, n)
    lens[5] = 0
    rows = np.repeat(np.arange(n), lens)
    cols = np.concatenate([np.sort(rng.choice(m, k, replace=False)) for k in lens])
    M = sp.csr_matrix((rng.standard_normal(rows.size), (rows, cols)), shape=(n, m))
    off, col, val = sell_layout(M, tpr)
    rps = 256 // tpr
    assert len(off) == (n + rps - 1) // rps + 1 and off[0] == 0
    assert np.all(np.diff(off) % rps == 0)
    v = rng.standard_normal(m)
    assert np.allclose(sell_matvec(off, col, val, n, tpr, v), M @ v, rtol=1e-14, atol=1e-14)
    D = sp.csr_matrix(rng.standard_normal((50, 50)))
    off, col, val = sell_layout(D, tpr, dense=True)
    assert col is None
    v = rng.standard_normal(50)
    assert np.allclose(sell_matvec(off, None, val, 50, tpr, v), D @ v, rtol=1e-14, atol=1e-14)
    with pytest.raises(ValueError):
        sell_layout(M, tpr, dense=True)
