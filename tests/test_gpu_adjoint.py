"""Adjoint gradient of the micromagnetic energy (exchange + anisotropy +
demag) with per-cell Ms and A.

* device energy and gradient vs the oracle's dense-FK-map gradient
  (oracle.energy_per_cell) on a small cube, <= 1e-8 relative;
* uniform Ms/A: energies equal the reference's lumped energies;
* central finite differences (h = 1e-6 relative) on random directions on a
  disk mesh, <= 1e-5 relative (cancellation-limited);
* torch.autograd.Function wiring.
"""
import numpy as np
import pytest

import smx_oracle as O

pytestmark = pytest.mark.gpu


def _unit(n, seed):
    g = np.random.default_rng(seed).standard_normal((n, 3))
    return (g / np.linalg.norm(g, axis=1)[:, None]).ravel()


@pytest.mark.parametrize("backend", ["dense", "hmatrix"])
def test_energy_gradient_vs_oracle(backend):
    import torch
    from smx import mesh as M
    from smx.adjoint import MicromagneticEnergy
    m = M.kuhn_box(6, 6, 6, 24.0, 24.0, 24.0)
    T = m.tets.shape[0]
    rng = np.random.default_rng(3)
    Ms = rng.uniform(6e5, 1e6, T)
    A = rng.uniform(1e-11, 1.5e-11, T)
    mv = _unit(m.num_nodes, 4)
    kw = {} if backend == "dense" else dict(boundary_backend="hmatrix", hmatrix_epsilon=1e-10, hmatrix_leaf_size=16)
    model = MicromagneticEnergy(m, Ku=5e5, n_ani_vec=(0.0, 0.0, 1.0), demag_kwargs=kw, ksp_rtol=1e-13)
    dev = torch.device("cuda")
    mt, Mt, At = (torch.from_numpy(a).to(dev) for a in (mv, Ms, A))
    E, (gm, gMs, gA) = model.value_and_grad(mt, Mt, At)
    dem = O.DemagOracle(m.X, m.tets, 1.0)
    Er, parts, gmr, gMsr, gAr = O.energy_per_cell(m.X, m.tets, mv, Ms, A, 5e5, (0.0, 0.0, 1.0), dem)
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
    print(f"[adjoint {backend}] E {E['E_total']:.6e} vs {Er:.6e}; grad m {rel(gm.cpu().numpy(), gmr):.2e} "
          f"Ms {rel(gMs.cpu().numpy(), gMsr):.2e} A {rel(gA.cpu().numpy(), gAr):.2e}", flush=True)
    tol = 1e-8 if backend == "dense" else 1e-6
    for k in ("E_ex", "E_an", "E_d"):
        assert abs(E[k] - parts[k]) <= tol * abs(Er), k
    assert rel(gm.cpu().numpy(), gmr) <= tol
    assert rel(gMs.cpu().numpy(), gMsr) <= tol
    assert rel(gA.cpu().numpy(), gAr) <= 1e-12


def test_uniform_energy_equals_lumped():
    import torch
    from smx import mesh as M
    from smx.adjoint import MicromagneticEnergy
    from smx.llg import EffectiveField
    m = M.cube_c1()
    T = m.tets.shape[0]
    dev = torch.device("cuda")
    mv = torch.from_numpy(_unit(m.num_nodes, 1)).to(dev)
    model = MicromagneticEnergy(m, Ku=5e5, n_ani_vec=(0.0, 0.0, 1.0))
    E = model.forward(mv, torch.full((T,), 8e5, dtype=torch.float64, device=dev),
                      torch.full((T,), 1.3e-11, dtype=torch.float64, device=dev))
    hef = EffectiveField(m, Ms=8e5, Aex=1.3e-11, Ku=5e5, n_ani_vec=(0.0, 0.0, 1.0), use_demag=True,
                         demag_kwargs=dict(ksp_rtol_u1=1e-13, ksp_rtol_u2=1e-13, ksp_atol=0.0))
    ref = hef.compute_Energy_terms(mv)
    print(f"[lumped] {E} vs {ref}", flush=True)
    assert abs(E["E_ex"] - ref["E_lin"]) <= 1e-10 * abs(ref["E_lin"])
    assert abs(E["E_an"] - ref["E_ani"]) <= 1e-12 * abs(ref["E_ani"])
    assert abs(E["E_d"] - ref["E_demag"]) <= 1e-8 * abs(ref["E_demag"])


def test_gradient_finite_differences_disk():
    """Central FD on 8 random directions in m, Ms and A (disk, 3 nm cells)."""
    import torch
    from smx import mesh as M
    from smx.adjoint import MicromagneticEnergy
    m = M.disk(diameter=60.0, height=9.0, h=3.0)
    T = m.tets.shape[0]
    dev = torch.device("cuda")
    rng = np.random.default_rng(7)
    Ms = torch.from_numpy(rng.uniform(6e5, 1e6, T)).to(dev)
    A = torch.from_numpy(rng.uniform(1e-11, 1.5e-11, T)).to(dev)
    mv = torch.from_numpy(_unit(m.num_nodes, 8)).to(dev)
    model = MicromagneticEnergy(m, Ku=5e5, n_ani_vec=(0.0, 0.0, 1.0), ksp_rtol=1e-13)
    E0, (gm, gMs, gA) = model.value_and_grad(mv, Ms, A)
    h = 1e-6
    worst = 0.0
    for k in range(8):
        dm = torch.from_numpy(rng.standard_normal(3 * m.num_nodes)).to(dev)
        dM = torch.from_numpy(rng.standard_normal(T)).to(dev) * 1e5
        dA = torch.from_numpy(rng.standard_normal(T)).to(dev) * 1e-12
        Ep = model.forward(mv + h * dm, Ms + h * dM, A + h * dA)["E_total"]
        Em = model.forward(mv - h * dm, Ms - h * dM, A - h * dA)["E_total"]
        fd = (Ep - Em) / (2 * h)
        an = float((gm * dm).sum() + (gMs * dM).sum() + (gA * dA).sum())
        worst = max(worst, abs(fd - an) / abs(an))
    print(f"[adjoint FD disk N={m.num_nodes}] worst rel {worst:.2e}", flush=True)
    assert worst <= 1e-5


def test_autograd_function():
    import torch
    from smx import mesh as M
    from smx.adjoint import MicromagneticEnergy, MicromagneticEnergyFunction
    m = M.kuhn_box(5, 5, 5, 20.0, 20.0, 20.0)
    T = m.tets.shape[0]
    dev = torch.device("cuda")
    model = MicromagneticEnergy(m, Ku=5e5, n_ani_vec=(0.0, 0.0, 1.0))
    mv = torch.from_numpy(_unit(m.num_nodes, 2)).to(dev).requires_grad_(True)
    Ms = torch.full((T,), 8e5, dtype=torch.float64, device=dev, requires_grad=True)
    A = torch.full((T,), 1.3e-11, dtype=torch.float64, device=dev, requires_grad=True)
    E = MicromagneticEnergyFunction.apply(mv, Ms, A, model)
    E.backward()
    _, (gm, gMs, gA) = model.value_and_grad(mv.detach(), Ms.detach(), A.detach())
    assert torch.allclose(mv.grad, gm) and torch.allclose(Ms.grad, gMs) and torch.allclose(A.grad, gA)


def test_hmatrix_transpose_matvec():
    import torch
    from smx import mesh as M
    from smx.boundary import build_boundary_geometry
    from smx.hmatrix import HMatrixBuildConfig, HMatrixGPUPackedFused, build_hmatrix_device
    m = M.cube_c1()
    geo = build_boundary_geometry(m.X, m.tets)
    bd = O.boundary_data(m.X, m.tets)
    B = O.dense_B_rows(bd)
    data = build_hmatrix_device(geo, HMatrixBuildConfig(epsilon=1e-8, leaf_size=32), torch.device("cuda"))
    H = HMatrixGPUPackedFused(data, use_fp32=False)
    x = np.random.default_rng(0).standard_normal(geo.Nb)
    y = H.rmatvec(torch.from_numpy(x).cuda()).cpu().numpy()
    e = np.linalg.norm(y - B.T @ x) / np.linalg.norm(B.T @ x)
    print(f"[hmatrix B^T] {e:.2e}", flush=True)
    assert e <= 1e-7
