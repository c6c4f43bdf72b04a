"""Pins of the oracle's implicit-path restatements.

* llg_local_jacobians (vectorized) equals llg_local_jacobian node by node;
* gmres_right solves nonsymmetric systems to its tolerance against numpy's
  dense solve, with restarts (restart 5) and a diagonal right
  preconditioner;
* bdf_integrate with fixed steps on y' = A y (A nonnormal, stable) converges
  at second order against the matrix exponential (error ratio ~4 per
  halving of dt);
* the adaptive run meets its tolerance on y' = A y and on the damped
  macrospin, whose closed form is tan(theta/2) = tan(theta0/2)
  exp(-alpha g' H t), phi = phi0 + g' H t;
* llg_bdf_relax on the 40 nm cube agrees with the oracle's fixed-step
  Dormand-Prince solution of the same problem.
"""
import math

import numpy as np
import pytest
import scipy.linalg as sla

import smx_oracle as O


def test_vectorized_local_jacobians():
    rng = np.random.default_rng(0)
    n = 40
    m = rng.standard_normal(3 * n) * 0.7
    h = rng.standard_normal(3 * n) * 1e5
    kap = rng.uniform(-3e5, 3e5, n)
    for stab in (True, False):
        J = O.llg_local_jacobians(m, h, kap, 2.211e5, 0.3, 8e5, include_stab=stab)
        for i in range(n):
            Ji = O.llg_local_jacobian(m[3 * i:3 * i + 3], h[3 * i:3 * i + 3], kap[i], 2.211e5, 0.3, 8e5,
                                      include_stab=stab)
            assert np.abs(J[i] - Ji).max() <= 1e-12 * np.abs(Ji).max()


@pytest.mark.parametrize("restart,diag_pc", [(30, False), (5, False), (5, True)])
def test_gmres_right(restart, diag_pc):
    rng = np.random.default_rng(1)
    n = 60
    A = np.eye(n) * 4 + rng.standard_normal((n, n)) * 0.3 + np.diag(rng.uniform(1, 5, n))
    b = rng.standard_normal(n)
    d = np.diag(A)
    pc = (lambda v: v / d) if diag_pc else (lambda v: v.copy())
    x = np.zeros(n)
    its, res = O.gmres_right(lambda v: A @ v, pc, b, x, 1e-10, max_it=500, restart=restart)
    ref = np.linalg.solve(A, b)
    assert res <= 1e-10 * np.linalg.norm(b) * 1.0001
    assert np.linalg.norm(x - ref) <= 1e-8 * np.linalg.norm(ref)
    assert np.linalg.norm(b - A @ x) == pytest.approx(res, rel=1e-6, abs=1e-14)


def _linear_problem():
    A = np.array([[-2.0, 5.0, 0.0], [0.0, -3.0, 1.0], [0.5, 0.0, -1.0]]) * 1e9
    y0 = np.array([1.0, -0.5, 0.25])
    ops = lambda y, shift: ((lambda v: shift * v - A @ v), (lambda v: v.copy()))
    return A, y0, (lambda t, y: A @ y), ops


def test_bdf2_fixed_step_second_order():
    A, y0, f, ops = _linear_problem()
    T = 2e-9
    exact = sla.expm(A * T) @ y0
    errs = []
    for nst in (200, 400, 800):
        # residual entries are ~shift |y| ~ 1e11: an absolute Newton tolerance of 1e-3 is 1e-14 relative
        y, t, st = O.bdf_integrate(f, ops, y0, 0.0, T, T / nst, fixed_dt=True, snes_rtol=1e-12, snes_atol=1e-3,
                                   ksp_rtol=1e-14, dt_max=1.0)
        assert st["steps"] == nst and abs(t - T) <= 1e-12 * T
        errs.append(np.linalg.norm(y - exact))
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 3.6 < r1 < 4.4 and 3.6 < r2 < 4.4, errs


def test_bdf_adaptive_meets_tolerance():
    A, y0, f, ops = _linear_problem()
    T = 5e-9
    exact = sla.expm(A * T) @ y0
    prev = None
    for tol in (1e-5, 1e-7):
        y, t, st = O.bdf_integrate(f, ops, y0, 0.0, T, 1e-12, rtol=tol, atol=tol, snes_rtol=1e-12,
                                   snes_atol=1e-3, ksp_rtol=1e-12, dt_max=1.0)
        err = np.abs(y - exact).max()
        assert err <= 50 * tol and abs(t - T) <= 1e-12 * T
        if prev is not None:
            assert st["steps"] > prev
        prev = st["steps"]


def test_bdf_macrospin_closed_form():
    g, a, Ms, Hz = 2.211e5, 0.1, 8e5, 1e5
    H = np.array([0.0, 0.0, Hz])
    th0, ph0 = 1.0, 0.3
    m0 = np.array([math.sin(th0) * math.cos(ph0), math.sin(th0) * math.sin(ph0), math.cos(th0)])
    f = lambda t, y: O.llg_rhs(y, H, g, a, Ms)

    def ops(y, shift):
        J = O.llg_local_jacobians(y, H, 0.0, g, a, Ms)[0]
        inv = np.linalg.inv(shift * np.eye(3) - J)
        return (lambda v: shift * v - O.jac_vec(y, H, v, np.zeros(3), g, a, Ms)), (lambda v: inv @ v)

    T = 4e-11
    # residual entries are ~shift ~ 1e13-1e14 per second: absolute 1 is ~1e-13 relative
    y, t, st = O.bdf_integrate(f, ops, m0, 0.0, T, 1e-14, rtol=1e-8, atol=1e-8, snes_rtol=1e-10,
                               snes_atol=1.0, ksp_rtol=1e-12)
    gp = g / (1 + a * a)
    th = 2 * math.atan(math.tan(th0 / 2) * math.exp(-a * gp * Hz * T))
    ph = ph0 + gp * Hz * T
    exact = np.array([math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)])
    assert np.abs(y - exact).max() <= 1e-5, (st, np.abs(y - exact).max())


def test_llg_bdf_relax_matches_dp5_on_cube():
    from smx import mesh as M
    m = M.cube_c1()
    N = m.num_nodes
    g0 = np.random.default_rng(3).standard_normal((N, 3))
    m0 = (g0 / np.linalg.norm(g0, axis=1)[:, None]).ravel()
    Ms, A, Ku, alpha, gamma = 8e5, 1.3e-11, 5e5, 0.5, 2.211e5
    dem = O.DemagOracle(m.X, m.tets, Ms)
    T = 2e-12
    y, t, st = O.llg_bdf_relax(m.X, m.tets, m0, T, 1e-15, Ms, 1.3e-11, alpha, Ku=Ku, dem=dem, rtol=1e-7,
                               atol=1e-7, snes_rtol=1e-8, snes_atol=1e-10, ksp_rtol=1e-10)
    Kex = O.exchange_matrix(m.X, m.tets, dem.fem.V, A, Ms)
    f = lambda tt, yy: O.llg_rhs(yy, Kex @ yy + O.anisotropy_field(yy, (0, 0, 1), Ku, Ms) + dem.compute(yy),
                                 gamma, alpha, Ms)
    z = m0.copy()
    k1 = None
    nst = 400
    for _ in range(nst):
        z, _, k1 = O.dp5_step(f, 0.0, z, T / nst, k1)
    err = np.linalg.norm(y - z) / np.linalg.norm(z)
    # BDF2's global error at rtol 1e-7 over ~250 stiff steps is ~4e-5 here (the device
    # integrator against the device Dormand-Prince shows the same size)
    assert abs(t - T) <= 1e-9 * T and st["steps"] > 10 and err <= 1e-4, (st, err)
