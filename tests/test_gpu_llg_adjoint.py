"""Adjoint through the LLG time steps (per-cell Ms): forward parity with the
effective-field path, the right-hand-side VJP against central differences,
the through-time gradient against central differences, and checkpointing
that changes nothing.

Central differences use the relative step h = 1e-6 of the configuration
and CG tolerances of 1e-12, so the difference quotients carry ~1e-8 relative
error; the bars below sit above that.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")
KW = dict(ksp_rtol=1e-12, demag_kwargs=dict(ksp_max_it_u1=20000, ksp_max_it_u2=20000))


def _cube():
    from smx import mesh as M
    return M.kuhn_box(6, 6, 6, 24.0, 24.0, 24.0, name="cube24_6")


def _random_m(n, seed):
    g = np.random.default_rng(seed).standard_normal((n, 3))
    return torch.from_numpy((g / np.linalg.norm(g, axis=1)[:, None]).ravel()).to(DEV)


def _model(mesh, alpha=0.02):
    from smx.llg_adjoint import LLGAdjointModel
    return LLGAdjointModel(mesh, A_ex=1.3e-11, alpha=alpha, device=DEV, **KW)


def test_uniform_ms_matches_effective_field():
    """Uniform Ms: the per-node model is the reference's exchange + demag RHS."""
    from smx.llg import EffectiveField
    mesh = _cube()
    model = _model(mesh, alpha=0.5)
    Ms = 8e5
    hef = EffectiveField(mesh, Ms=Ms, Aex=1.3e-11, alpha=0.5, use_demag=True, device=DEV,
                         demag_kwargs=dict(ksp_rtol_u1=1e-12, ksp_rtol_u2=1e-12, ksp_atol=0.0,
                                           ksp_max_it_u1=20000, ksp_max_it_u2=20000))
    m = _random_m(mesh.num_nodes, 1)
    r_ref = torch.empty_like(m)
    hef.llg_rhs_vec(m, r_ref)
    r = torch.empty_like(m)
    model.rhs(m, torch.full((model.T,), Ms, dtype=torch.float64, device=DEV), r)
    assert float(torch.linalg.norm(r - r_ref) / torch.linalg.norm(r_ref)) < 1e-10


def test_rhs_vjp_against_central_differences():
    mesh = _cube()
    model = _model(mesh)
    rng = np.random.default_rng(2)
    m = _random_m(mesh.num_nodes, 3)
    Ms = torch.from_numpy(rng.uniform(6e5, 1e6, model.T)).to(DEV)
    lam = torch.from_numpy(rng.standard_normal(3 * model.N)).to(DEV)
    gm, gMs = model.vjp(m, Ms, lam)
    gm, gMs = gm.clone(), gMs.clone()
    r = lambda mm, ms: model.rhs(mm, ms, torch.empty_like(m)).clone()
    for _ in range(2):
        v = torch.from_numpy(rng.standard_normal(3 * model.N)).to(DEV)
        eps = 1e-6
        fd = float(lam @ (r(m + eps * v, Ms) - r(m - eps * v, Ms))) / (2 * eps)
        an = float(gm @ v)
        assert abs(fd - an) <= 1e-6 * abs(an)
        w = torch.from_numpy(rng.standard_normal(model.T)).to(DEV) * Ms
        h = 1e-6
        fd = float(lam @ (r(m, Ms + h * w) - r(m, Ms - h * w))) / (2 * h)
        an = float(gMs @ w)
        assert abs(fd - an) <= 1e-6 * abs(an)


def test_gradient_through_time_against_central_differences():
    """dJ/dMs for J = <m_z> after 12 DP5 steps (dt 1e-13 s, alpha 0.02)."""
    from smx.llg_adjoint import integrate, mean_mz
    mesh = _cube()
    model = _model(mesh)
    rng = np.random.default_rng(4)
    m0 = _random_m(mesh.num_nodes, 5)
    Ms = torch.from_numpy(rng.uniform(6e5, 1e6, model.T)).to(DEV).requires_grad_(True)
    J = mean_mz(model, integrate(model, m0, Ms, 1e-13, 12, checkpoint_every=5))
    (g,) = torch.autograd.grad(J, Ms)
    with torch.no_grad():
        for _ in range(2):
            w = torch.from_numpy(rng.standard_normal(model.T)).to(DEV) * Ms
            h = 1e-6
            Jp = float(mean_mz(model, integrate(model, m0, Ms + h * w, 1e-13, 12)))
            Jm = float(mean_mz(model, integrate(model, m0, Ms - h * w, 1e-13, 12)))
            fd = (Jp - Jm) / (2 * h)
            an = float(g @ w)
            assert abs(fd - an) <= 1e-4 * abs(an), (fd, an)


def test_checkpointing_changes_nothing():
    from smx.llg_adjoint import integrate, mean_mz
    mesh = _cube()
    model = _model(mesh)
    m0 = _random_m(mesh.num_nodes, 7)
    Ms0 = torch.from_numpy(np.random.default_rng(6).uniform(6e5, 1e6, model.T)).to(DEV)
    grads = []
    for every in (4, 100):
        Ms = Ms0.clone().requires_grad_(True)
        J = mean_mz(model, integrate(model, m0, Ms, 1e-13, 9, checkpoint_every=every))
        (g,) = torch.autograd.grad(J, Ms)
        grads.append(g)
    assert torch.equal(grads[0], grads[1])
