"""Implicit (BDF Newton-Krylov) LLG steps on the device.

* smx_llg_pc_build: every node's 3x3 block equals inv(shift I - J_i) with
  J_i the oracle's exact local Jacobian (pinned by finite differences on the
  host) at kappa the signed mean diagonal of the exchange + uniaxial
  operator (exact form), and the reference's block built from it (precession
  and damping parts negated, stabilization kept, kappa the mean |diagonal|):
  1e-12;
* smx_multi_dot / smx_lincomb against numpy;
* GMRES on (shift I - J) x = b: the true residual (an independent
  application of the operator) meets the tolerance without a preconditioner
  and with the exact block, which cuts the iteration count (the reference's
  block and the exact one with |kappa| are run and reported);
* BDF integration of the 40 nm cube (exchange + anisotropy + FEM/BEM demag)
  agrees with the adaptive Dormand-Prince solution at t = 20 ps;
* LLG_GPU.relax (implicit) of a damped exchange + uniaxial box stops on the
  max |dm/dt| criterion in the uniform state along the easy axis;
* the device integrator against the oracle's restatement of the same BDF
  Newton-Krylov method (LU Poisson solves, numpy GMRES and blocks) on the
  40 nm cube with demag: same step count (+-2) and the same state at 2 ps
  (1e-6 relative).
"""
import math

import numpy as np
import pytest
import torch

import smx_oracle as O

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")


def _unit(n, seed):
    g = np.random.default_rng(seed).standard_normal((n, 3))
    return (g / np.linalg.norm(g, axis=1)[:, None]).ravel()


def _box():
    from smx import mesh as M
    return M.kuhn_box(8, 7, 5, 24.0, 21.0, 15.0, name="box")


def test_pc_blocks_match_oracle():
    from smx.implicit import JvContext
    from smx.llg import EffectiveField
    m = _box()
    N = m.num_nodes
    Ms, A, Ku, alpha, gamma = 8e5, 1.3e-11, 5e5, 0.2, 2.211e5
    hef = EffectiveField(m, Ms=Ms, Aex=A, Ku=Ku, n_ani_vec=(0.0, 0.6, 0.8), alpha=alpha, gamma=gamma, device=DEV)
    mv = _unit(N, 0) * 1.01
    hef.update_jac_state(torch.from_numpy(mv).to(DEV))
    ctx = JvContext(hef)
    shift = 3e12
    hm = hef.Hm_gpu.view(N, 3).cpu().numpy()
    fem = O.fem_operators(m.X, m.tets, Ms)
    Kex = O.exchange_matrix(m.X, m.tets, fem.V, A, Ms)
    n_ax = np.array([0.0, 0.6, 0.8])
    pref = 2 * Ku / (4e-7 * math.pi * Ms)
    dK = Kex.diagonal().reshape(N, 3) + pref * n_ax ** 2
    m3 = mv.reshape(N, 3)
    for exact in (False, True):
        ctx.update_pc_full_fast_gpu(shift, exact=exact, use_abs_kappa=not exact)
        inv = ctx.inv.view(N, 3, 3).cpu().numpy()
        kappa = dK.mean(axis=1) if exact else np.abs(dK).mean(axis=1)
        worst = 0.0
        for i in range(N):
            Jx = O.llg_local_jacobian(m3[i], hm[i], kappa[i], gamma, alpha, Ms, include_stab=False)
            Js = O.llg_local_jacobian(m3[i], hm[i], kappa[i], gamma, alpha, Ms) - Jx
            J = Js + Jx if exact else Js - Jx
            ref = np.linalg.inv((shift + 1e-14) * np.eye(3) - J)
            worst = max(worst, np.abs(inv[i] - ref).max() / np.abs(ref).max())
        print(f"[pc exact={exact}] worst block rel {worst:.2e}", flush=True)
        assert worst <= 1e-12
    x = torch.randn(3 * N, dtype=torch.float64, device=DEV)
    y = torch.empty_like(x)
    ctx.apply(x, y)
    ref = np.einsum("nij,nj->ni", inv, x.view(N, 3).cpu().numpy())
    assert np.linalg.norm(y.view(N, 3).cpu().numpy() - ref) <= 1e-14 * np.linalg.norm(ref)


@pytest.mark.parametrize("n,k", [(10007, 7), (5000, 31), (3, 1), (2048, 16)])
def test_multi_dot_and_lincomb(n, k):
    from smx import _capi
    L = _capi.lib()
    V = torch.randn(k, n, dtype=torch.float64, device=DEV)
    w = torch.randn(n, dtype=torch.float64, device=DEV)
    out = torch.zeros(k, dtype=torch.float64, device=DEV)
    work = torch.zeros(int(L.smx_multi_dot_ws_size(n, k)), dtype=torch.float64, device=DEV)
    _capi.check(L.smx_multi_dot(n, k, V.data_ptr(), n, w.data_ptr(), out.data_ptr(), work.data_ptr(),
                                _capi.stream_ptr()))
    ref = V.cpu().numpy() @ w.cpu().numpy()
    assert np.allclose(out.cpu().numpy(), ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    again = torch.zeros_like(out)
    _capi.check(L.smx_multi_dot(n, k, V.data_ptr(), n, w.data_ptr(), again.data_ptr(), work.data_ptr(),
                                _capi.stream_ptr()))
    assert torch.equal(again, out)                      # fixed reduction order
    c = torch.randn(k, dtype=torch.float64, device=DEV)
    y = torch.randn(n, dtype=torch.float64, device=DEV)
    y0 = y.cpu().numpy().copy()
    _capi.check(L.smx_lincomb(n, k, V.data_ptr(), n, c.data_ptr(), 0.5, y.data_ptr(), _capi.stream_ptr()))
    assert np.allclose(y.cpu().numpy(), 0.5 * y0 + c.cpu().numpy() @ V.cpu().numpy(), rtol=1e-12)


def test_gmres_newton_operator():
    from smx.implicit import GMRES, JvContext
    from smx.llg import EffectiveField
    m = _box()
    N = m.num_nodes
    hef = EffectiveField(m, Ms=8e5, Aex=1.3e-11, Ku=5e5, n_ani_vec=(0.0, 0.0, 1.0), alpha=0.1, device=DEV)
    hef.update_jac_state(torch.from_numpy(_unit(N, 1)).to(DEV))
    ctx = JvContext(hef)
    shift = 2e12                      # dt ~ 0.5 ps: beyond the explicit stability limit of the exchange
    b = torch.from_numpy(np.random.default_rng(2).standard_normal(3 * N)).to(DEV)
    out = {}
    for name, kw in (("none", None), ("exact_signed", dict(exact=True, use_abs_kappa=False)),
                     ("exact_abs", dict(exact=True, use_abs_kappa=True)),
                     ("reference", dict(exact=False, use_abs_kappa=True))):
        if kw is not None:
            ctx.update_pc_full_fast_gpu(shift, **kw)
        ctx.shift = shift
        g = GMRES(3 * N, DEV)
        x = torch.zeros_like(b)
        pc = ctx.apply if kw is not None else (lambda u, v: v.copy_(u))
        n_it, _ = g.solve(ctx.mult, pc, b, x, rtol=1e-8, max_it=400)
        r = torch.empty_like(b)
        ctx.mult(x, r)
        res = float(torch.linalg.norm(b - r) / torch.linalg.norm(b))
        print(f"[gmres pc={name}] its {n_it} true rel residual {res:.2e}", flush=True)
        out[name] = (n_it, res)
    assert out["none"][1] <= 2e-8
    assert out["exact_signed"][1] <= 2e-8
    assert out["exact_signed"][0] <= out["none"][0]


def test_bdf_matches_dormand_prince():
    from smx import mesh as M
    from smx.implicit import BDFIntegrator
    from smx.llg import DormandPrince, EffectiveField
    m = M.cube_c1()
    N = m.num_nodes
    kw = dict(Ms=8e5, Aex=1.3e-11, Ku=5e5, n_ani_vec=(0.0, 0.0, 1.0), alpha=0.5, device=DEV, use_demag=True,
              demag_kwargs=dict(ksp_rtol_u1=1e-12, ksp_rtol_u2=1e-12, ksp_atol=0.0))
    m0 = torch.from_numpy(_unit(N, 3)).to(DEV)
    t1 = 2e-11
    h1 = EffectiveField(m, **kw)
    y1 = m0.clone()
    dp = DormandPrince(h1, rtol=1e-9, atol=1e-9)
    s1 = dp.integrate(y1, 0.0, t1, 1e-15, exact_final_time="matchstep")
    h2 = EffectiveField(m, **kw)
    y2 = m0.clone()
    bdf = BDFIntegrator(h2, rtol=1e-7, atol=1e-7, snes_rtol=1e-6, snes_atol=1e-10, ksp_rtol=1e-8)
    s2 = bdf.integrate(y2, 0.0, t1, 1e-15, exact_final_time="matchstep")
    err = float(torch.linalg.norm(y2 - y1) / torch.linalg.norm(y1))
    print(f"[bdf] {s2} vs dp5 steps {s1['steps']}: rel diff {err:.2e}", flush=True)
    assert abs(s2["t"] - t1) <= 1e-9 * t1
    assert err <= 1e-4


def test_llg_relax_stops_on_dmdt():
    from smx.llg import LLG_GPU
    m = _box()
    N = m.num_nodes
    llg = LLG_GPU(m, Ms=8e5, Aex=1.3e-11, Ku=5e5, n_ani_vec=(0.0, 0.0, 1.0), alpha=1.0, device=DEV)
    g = np.random.default_rng(12).standard_normal((N, 3)) * 0.3 + np.array([0.0, 0.0, 1.0])
    m0 = (g / np.linalg.norm(g, axis=1)[:, None]).ravel()
    seen = []
    hef, st = llg.relax(m0, 0.0, 2e-9, 1e-14, stopping_dmdt=1.0, check_every_stop=5, return_stats=True,
                        monitor_fn=lambda step, t, y: seen.append(t))
    mz = hef.m_gpu.view(-1, 3)[:, 2].cpu().numpy()
    print(f"[relax] {st} min mz {mz.min():.12f}", flush=True)
    assert st["stopped_on_dmdt"] and st["maxdmdt_deg_ns"] < 1.0
    assert st["t"] < 2e-9 and len(seen) == st["steps"]
    assert mz.min() > 1.0 - 1e-4


def test_bdf_matches_oracle_integrator():
    from smx import mesh as M
    from smx.implicit import BDFIntegrator
    from smx.llg import EffectiveField
    m = M.cube_c1()
    N = m.num_nodes
    g0 = np.random.default_rng(3).standard_normal((N, 3))
    m0 = (g0 / np.linalg.norm(g0, axis=1)[:, None]).ravel()
    Ms, A, Ku, alpha = 8e5, 1.3e-11, 5e5, 0.5
    T = 2e-12
    tol = dict(rtol=1e-7, atol=1e-7, snes_rtol=1e-8, snes_atol=1e-10, ksp_rtol=1e-10)
    y_o, t_o, st_o = O.llg_bdf_relax(m.X, m.tets, m0, T, 1e-15, Ms, A, alpha, Ku=Ku, **tol)
    hef = EffectiveField(m, Ms=Ms, Aex=A, Ku=Ku, n_ani_vec=(0.0, 0.0, 1.0), alpha=alpha, device=DEV,
                         use_demag=True, demag_kwargs=dict(ksp_rtol_u1=1e-13, ksp_rtol_u2=1e-13, ksp_atol=0.0))
    y = torch.from_numpy(m0).to(DEV)
    st = BDFIntegrator(hef, **tol).integrate(y, 0.0, T, 1e-15, exact_final_time="matchstep")
    err = float(np.linalg.norm(y.cpu().numpy() - y_o) / np.linalg.norm(y_o))
    print(f"[bdf vs oracle] device {st} oracle {st_o}: rel diff {err:.2e}", flush=True)
    assert abs(st["t"] - T) <= 1e-9 * T and abs(t_o - T) <= 1e-9 * T
    assert abs(st["steps"] - st_o["steps"]) <= 2
    assert err <= 1e-6


def test_ifunction_and_shell_context_forms():
    """ifunction gives ydot - RHS; the (A, x, y) / (pc, x, y) forms of the
    shell context equal the two-argument ones."""
    from smx.implicit import JvContext
    from smx.llg import EffectiveField
    m = _box()
    N = m.num_nodes
    hef = EffectiveField(m, Ms=8e5, Aex=1.3e-11, Ku=5e5, n_ani_vec=(0.0, 0.0, 1.0), alpha=0.1, device=DEV)
    y = torch.from_numpy(_unit(N, 4)).to(DEV)
    ydot = torch.randn(3 * N, dtype=torch.float64, device=DEV)
    f = torch.empty_like(y)
    r = torch.empty_like(y)
    assert hef.ifunction(None, 1e-12, y, ydot, f) == 0
    hef.llg_rhs_vec(y, r)
    assert torch.equal(f, ydot - r)
    hef.update_jac_state(y)
    ctx = JvContext(hef)
    ctx.update_pc_full_fast_gpu(1e12, exact=True, use_abs_kappa=False)
    x = torch.randn_like(y)
    a, b = torch.empty_like(y), torch.empty_like(y)
    ctx.mult(x, a)
    ctx.mult(None, x, b)
    assert torch.equal(a, b)
    ctx.apply(x, a)
    ctx.apply(None, x, b)
    assert torch.equal(a, b)
