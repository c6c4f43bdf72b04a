"""Cubic anisotropy, Zhang-Li and spin-orbit torques, and the energy
minimizer on the device, against the FP64 oracle.

* H_eff, LLG RHS and Jv with cubic anisotropy (uniform and per-node axes)
  against the oracle (exchange + cubic field, its JVP): 1e-12 relative;
* RHS with Zhang-Li STT (nodal current direction) and SOT (uniform and
  per-node polarization) against the oracle; Jv against central finite
  differences of the device RHS (h = 1e-6): 1e-8 relative;
* energies: cubic, uniaxial and Zeeman against their closed forms;
* minimizer kernels (projected direction, Cayley step, max norm, normalize)
  against the oracle; the LaBonte/BB minimizer drives a random state of an
  exchange + uniaxial box to the uniform minimum E = -Ku V.
"""
import math

import numpy as np
import pytest
import torch

import smx_oracle as O

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")
MU0 = 4e-7 * math.pi


def _mesh(jitter=0.0):
    from smx import mesh as M
    m = M.kuhn_box(8, 7, 5, 24.0, 21.0, 15.0, name="box")
    if jitter:
        X = m.X + np.random.default_rng(7).uniform(-jitter, jitter, m.X.shape)
        m = M.TetMesh(X=X, tets=m.tets, name="boxj")
    return m


def _unit(n, seed, bias=None):
    g = np.random.default_rng(seed).standard_normal((n, 3))
    if bias is not None:
        g = g * 0.3 + np.asarray(bias)
    return (g / np.linalg.norm(g, axis=1)[:, None]).ravel()


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("per_node", [False, True])
def test_cubic_field_rhs_jvp(per_node):
    from smx.llg import EffectiveField
    m = _mesh(0.4)
    N = m.num_nodes
    Ms, A, K1, alpha, gamma = 8e5, 1.3e-11, -4.5e4, 0.3, 2.211e5
    if per_node:
        rng = np.random.default_rng(8)
        u1 = rng.standard_normal((N, 3))
        u2 = rng.standard_normal((N, 3))
    else:
        u1, u2 = (1.0, 1.0, 0.0), (0.0, 0.3, 1.0)
    hef = EffectiveField(m, Ms=Ms, Aex=A, Kc1=K1, u1_cub=u1, u2_cub=u2, alpha=alpha, gamma=gamma, device=DEV)
    axes = O.cubic_axes_oracle(u1, u2)
    fem = O.fem_operators(m.X, m.tets, Ms)
    Kex = O.exchange_matrix(m.X, m.tets, fem.V, A, Ms)
    mv = _unit(N, 0)
    md = torch.from_numpy(mv).to(DEV)
    H = torch.empty_like(md)
    hef.compute_H_eff_vec(md, H)
    Href = Kex @ mv + O.cubic_field(mv, axes, K1, Ms)
    assert rel(H.cpu().numpy(), Href) <= 1e-12
    r = torch.empty_like(md)
    hef.llg_rhs_vec(md, r)
    assert rel(r.cpu().numpy(), O.llg_rhs(mv, Href, gamma, alpha, Ms)) <= 1e-12
    v = np.random.default_rng(1).standard_normal(3 * N)
    hef.update_jac_state(md)
    jv = torch.empty_like(md)
    hef.jac_vec_times_vec(torch.from_numpy(v).to(DEV), jv)
    Hv = Kex @ v + O.cubic_field_jvp(mv, v, axes, K1, Ms)
    ref = O.jac_vec(mv, Href, v, Hv, gamma, alpha, Ms)
    err = rel(jv.cpu().numpy(), ref)
    print(f"[cubic per_node={per_node}] Jv rel {err:.2e}", flush=True)
    assert err <= 1e-12
    E = hef.compute_Energy_terms(md)
    Eref = O.cubic_energy(mv, axes, K1, fem.V)
    assert abs(E["E_cubic"] - Eref) <= 1e-12 * abs(Eref)


@pytest.mark.parametrize("per_node_p", [False, True])
def test_stt_sot_rhs_and_jvp(per_node_p):
    from smx.llg import EffectiveField
    m = _mesh(0.3)
    N = m.num_nodes
    Ms, A, Ku, alpha, gamma = 8e5, 1.3e-11, 3e5, 0.1, 2.211e5
    rng = np.random.default_rng(9)
    Jdir = rng.standard_normal((N, 3))
    Jmag, P, beta = 5e11, 0.6, 0.04
    p = rng.standard_normal((N, 3)) if per_node_p else (0.2, 1.0, -0.3)
    H_dl, H_fl = 2e4, -7e3
    hef = EffectiveField(m, Ms=Ms, Aex=A, Ku=Ku, n_ani_vec=(0.0, 0.0, 1.0), alpha=alpha, gamma=gamma,
                         Jmagnitude=Jmag, Jdir_vec=Jdir.ravel(), P=P, beta=beta, H_dl=H_dl, H_fl=H_fl,
                         polarization_vec=p, device=DEV)
    fem = O.fem_operators(m.X, m.tets, Ms)
    Kex = O.exchange_matrix(m.X, m.tets, fem.V, A, Ms)
    KJ = O.zhang_li_matrix(m.X, m.tets, fem.V, Jdir)
    mv = _unit(N, 2)
    md = torch.from_numpy(mv).to(DEV)
    Href = Kex @ mv + O.anisotropy_field(mv, (0.0, 0.0, 1.0), Ku, Ms)
    Z = np.column_stack([KJ @ mv.reshape(-1, 3)[:, k] for k in range(3)]).ravel()
    ref = O.llg_rhs_torques(mv, Href, gamma, alpha, Ms, Z=Z, stt=(Jmag, P, beta), p=p, H_dl=H_dl, H_fl=H_fl)
    r = torch.empty_like(md)
    hef.llg_rhs_vec(md, r)
    err = rel(r.cpu().numpy(), ref)
    print(f"[stt/sot per_node_p={per_node_p}] rhs rel {err:.2e}", flush=True)
    assert err <= 1e-12
    # Jv against central differences of the device RHS
    v = np.random.default_rng(3).standard_normal(3 * N)
    h = 1e-6
    rp, rm = torch.empty_like(md), torch.empty_like(md)
    hef.llg_rhs_vec(torch.from_numpy(mv + h * v).to(DEV), rp)
    hef.llg_rhs_vec(torch.from_numpy(mv - h * v).to(DEV), rm)
    fd = ((rp - rm) / (2 * h)).cpu().numpy()
    hef.update_jac_state(md)
    jv = torch.empty_like(md)
    hef.jac_vec_times_vec(torch.from_numpy(v).to(DEV), jv)
    err = rel(jv.cpu().numpy(), fd)
    print(f"[stt/sot per_node_p={per_node_p}] Jv vs FD rel {err:.2e}", flush=True)
    assert err <= 1e-8


def test_energies_closed_forms():
    from smx.llg import EffectiveField
    m = _mesh()
    N = m.num_nodes
    Ms, Ku, K1 = 8e5, 5e5, 3e4
    H0 = (1e4, -2e4, 5e3)
    hef = EffectiveField(m, Ms=Ms, Aex=0.0, Ku=Ku, n_ani_vec=(0.0, 0.0, 1.0), Kc1=K1, u1_cub=(1, 0, 0),
                         u2_cub=(0, 1, 0), H0_static=H0, device=DEV)
    fem = O.fem_operators(m.X, m.tets, Ms)
    mv = _unit(N, 4)
    E = hef.compute_Energy_terms(torch.from_numpy(mv).to(DEV))
    m3 = mv.reshape(-1, 3)
    Vt = fem.V * 1e-27
    assert abs(E["E_ani"] - (-Ku * np.sum(Vt * m3[:, 2] ** 2))) <= 1e-12 * abs(E["E_ani"])
    S = (m3[:, 0] * m3[:, 1]) ** 2 + (m3[:, 1] * m3[:, 2]) ** 2 + (m3[:, 2] * m3[:, 0]) ** 2
    assert abs(E["E_cubic"] - (-K1 * np.sum(Vt * S))) <= 1e-12 * abs(E["E_cubic"])
    Ez = -MU0 * Ms * np.sum(Vt * (m3 @ np.asarray(H0)))
    assert abs(E["E_zeeman"] - Ez) <= 1e-12 * abs(Ez)


def test_minimizer_kernels():
    from smx import _capi
    from smx.device import reduce_workspace
    L = _capi.lib()
    n = 5000
    rng = np.random.default_rng(10)
    m, h, d = (rng.standard_normal(3 * n) for _ in range(3))
    md, hd, dd = (torch.from_numpy(x).to(DEV) for x in (m, h, d))
    out = torch.empty_like(md)
    st = _capi.stream_ptr()
    _capi.check(L.smx_projected_direction(n, md.data_ptr(), hd.data_ptr(), out.data_ptr(), st))
    torch.cuda.synchronize()
    assert rel(out.cpu().numpy(), O.projected_direction(m, h)) <= 1e-15
    _capi.check(L.smx_cayley_step(n, md.data_ptr(), dd.data_ptr(), 0.21, out.data_ptr(), st))
    torch.cuda.synchronize()
    assert rel(out.cpu().numpy(), O.cayley_step(m, d, 0.21)) <= 1e-14
    red = reduce_workspace(n + 64, DEV)
    mx = torch.zeros(1, dtype=torch.float64, device=DEV)
    _capi.check(L.smx_max_norm3(n, hd.data_ptr(), mx.data_ptr(), red.data_ptr(), st))
    assert float(mx.item()) == float(np.linalg.norm(h.reshape(-1, 3), axis=1).max())
    _capi.check(L.smx_normalize3(n, md.data_ptr(), st))
    torch.cuda.synchronize()
    m3 = m.reshape(-1, 3)
    assert rel(md.cpu().numpy(), (m3 / np.linalg.norm(m3, axis=1)[:, None]).ravel()) <= 1e-15


@pytest.mark.parametrize("variant", ["adaptive", "alternate"])
def test_minimizer_reaches_uniform_minimum(variant):
    from smx.llg import EffectiveField
    from smx.minimizer import LaBonteBBMinimizerGPU
    m = _mesh()
    N = m.num_nodes
    Ms, A, Ku = 8e5, 1.3e-11, 5e5
    hef = EffectiveField(m, Ms=Ms, Aex=A, Ku=Ku, n_ani_vec=(0.0, 0.0, 1.0), device=DEV)
    hef.set_m_from_cpu(_unit(N, 11, bias=(0.0, 0.0, 1.0)))
    res = LaBonteBBMinimizerGPU(hef).minimize(max_iter=3000, tol=1.0, bb_variant=variant)
    fem = O.fem_operators(m.X, m.tets, Ms)
    Emin = -Ku * fem.V.sum() * 1e-27
    mz = hef.m_gpu.view(-1, 3)[:, 2].cpu().numpy()
    print(f"[minimizer {variant}] {res['stop_reason']} its {res['iterations']} E {res['energy']:.6e} "
          f"vs {Emin:.6e}, max|d| {res['max_projected_field']:.2e}", flush=True)
    assert res["converged"]
    assert mz.min() > 1.0 - 1e-8
    assert abs(res["energy"] - Emin) <= 1e-8 * abs(Emin)


def test_minimizer_matches_oracle():
    """The device minimizer against the oracle's restatement
    (labonte_bb_minimize) on an exchange + tilted uniaxial + Zeeman box:
    same iteration count and stop reason, energies to 1e-9, states to 1e-6."""
    from smx.llg import EffectiveField
    from smx.minimizer import LaBonteBBMinimizerGPU
    m = _mesh()
    N = m.num_nodes
    Ms, A, Ku = 8e5, 1.3e-11, 5e5
    n_ax, H0 = (0.0, 0.6, 0.8), (2e4, 0.0, -1e4)
    hef = EffectiveField(m, Ms=Ms, Aex=A, Ku=Ku, n_ani_vec=n_ax, H0_static=H0, device=DEV)
    m0 = _unit(N, 13, bias=(0.0, 0.0, 1.0))
    hef.set_m_from_cpu(m0)
    kw = dict(max_iter=150, tol=1e-6)
    res = LaBonteBBMinimizerGPU(hef).minimize(**kw)
    fem = O.fem_operators(m.X, m.tets, Ms)
    K = O.exchange_matrix(m.X, m.tets, fem.V, A, Ms)
    Vt = np.repeat(fem.V, 3) * 1e-27
    h0 = np.tile(np.asarray(H0), N)
    hani = lambda mm: O.anisotropy_field(mm, n_ax, Ku, Ms)
    energy = lambda mm: float(MU0 * Ms * np.sum(Vt * mm * (-0.5 * (K @ mm) - 0.5 * hani(mm) - h0)))
    heff = lambda mm: K @ mm + hani(mm) + h0
    mo, info = O.labonte_bb_minimize(energy, heff, m0, **kw)
    mg = hef.m_gpu.cpu().numpy()
    dm = float(np.linalg.norm(mg - mo) / np.linalg.norm(mo))
    print(f"[minimizer vs oracle] device {res['iterations']} its {res['stop_reason']} E {res['energy']:.12e}; "
          f"oracle {info['iterations']} its {info['stop_reason']} E {info['energy']:.12e}; |dm| {dm:.2e}",
          flush=True)
    assert res["iterations"] == info["iterations"] and res["stop_reason"] == info["stop_reason"]
    assert abs(res["energy"] - info["energy"]) <= 1e-9 * abs(info["energy"])
    assert dm <= 1e-6
