"""The adjoint workloads at their stated sizes.

* Disk (200 nm diameter, 20 nm height, 3 nm mean edge; Ms_cell ~ U[6e5, 1e6],
  A_cell ~ U[1e-11, 1.5e-11], Ku 5e5 along z): the gradient of
  E = E_ex + E_an + E_d with respect to m, Ms_cell and A_cell against central
  differences (step h = 1e-6) on 64 random joint directions.  Bound: 1e-5
  relative per direction (CG at 1e-13; the difference quotient loses ~1e-9
  to cancellation and carries O(h^2) truncation).
* 150 nm cube (30^3 Kuhn cells, 162 000 cells): the gradient of the final
  <m_z> with respect to per-cell Ms after 500 fixed Dormand-Prince steps of
  1e-13 s (alpha 0.02, vortex start, checkpoint every 50 steps).  The FP64
  gradient against central differences (relative step 1e-6) on two random
  directions: 1e-6; the FP32-boundary forward's gradient against the FP64
  gradient: 1e-6 relative L2.
"""
import time

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")


def _random_m(n, seed):
    g = np.random.default_rng(seed).standard_normal((n, 3))
    return (g / np.linalg.norm(g, axis=1)[:, None]).ravel()


def _vortex_m(X, seed=0, center=(75.0, 75.0), rc=10.0, noise=0.05):
    rng = np.random.default_rng(seed)
    x = X[:, 0] - center[0]
    y = X[:, 1] - center[1]
    m = np.stack([-y, x, np.full_like(x, rc)], axis=1)
    m /= np.linalg.norm(m, axis=1)[:, None]
    m += rng.normal(0.0, noise, m.shape)
    m /= np.linalg.norm(m, axis=1)[:, None]
    return m.ravel()


def test_disk_energy_gradient_64_directions():
    from smx import mesh as M
    from smx.adjoint import MicromagneticEnergy
    mesh = M.disk(diameter=200.0, height=20.0, h=3.0)
    T = int(mesh.tets.shape[0])
    rng = np.random.default_rng(0)
    Ms = torch.from_numpy(rng.uniform(6e5, 1e6, T)).to(DEV)
    A = torch.from_numpy(rng.uniform(1e-11, 1.5e-11, T)).to(DEV)
    m = torch.from_numpy(_random_m(mesh.num_nodes, 1)).to(DEV)
    model = MicromagneticEnergy(mesh, Ku=5e5, n_ani_vec=(0.0, 0.0, 1.0), ksp_rtol=1e-13, device=DEV,
                                demag_kwargs=dict(B_dtype="float64"))
    t0 = time.perf_counter()
    E, (gm, gMs, gA) = model.value_and_grad(m, Ms, A)
    torch.cuda.synchronize()
    t_grad = time.perf_counter() - t0
    h = 1e-6
    errs = []
    for _ in range(64):
        dm = torch.from_numpy(rng.standard_normal(3 * mesh.num_nodes)).to(DEV)
        dM = torch.from_numpy(rng.standard_normal(T)).to(DEV) * 1e5
        dA = torch.from_numpy(rng.standard_normal(T)).to(DEV) * 1e-12
        Ep = model.forward(m + h * dm, Ms + h * dM, A + h * dA)["E_total"]
        Em = model.forward(m - h * dm, Ms - h * dM, A - h * dA)["E_total"]
        fd = (Ep - Em) / (2 * h)
        an = float((gm * dm).sum() + (gMs * dM).sum() + (gA * dA).sum())
        errs.append(abs(fd - an) / abs(an))
    print(f"[disk N={mesh.num_nodes} T={T} Nb={model.demag.Nb}] value_and_grad {t_grad * 1e3:.1f} ms; "
          f"64 directions: max {max(errs):.2e} median {np.median(errs):.2e}", flush=True)
    assert 30_000 < mesh.num_nodes < 50_000
    assert max(errs) <= 1e-5


def test_cube_adjoint_through_500_steps():
    from smx import mesh as M
    from smx.llg_adjoint import LLGAdjointModel, integrate, mean_mz
    mesh = M.cube_c2()
    m0 = torch.from_numpy(_vortex_m(mesh.X, seed=0)).to(DEV)
    dt, steps, every = 1e-13, 500, 50
    grads = {}
    for name, bdt in (("fp64", "float64"), ("fp32", "float32")):
        model = LLGAdjointModel(mesh, A_ex=1.3e-11, alpha=0.02, ksp_rtol=1e-10, device=DEV,
                                demag_kwargs=dict(B_dtype=bdt, ksp_max_it_u1=20000, ksp_max_it_u2=20000))
        Ms = torch.full((model.T,), 8e5, dtype=torch.float64, device=DEV, requires_grad=True)
        t0 = time.perf_counter()
        J = mean_mz(model, integrate(model, m0, Ms, dt, steps, every))
        (g,) = torch.autograd.grad(J, Ms)
        torch.cuda.synchronize()
        grads[name] = g
        print(f"[cube adjoint {name}] J={float(J.detach()):.12f}, forward + gradient {time.perf_counter() - t0:.1f}s, "
              f"{model.rhs_evals} RHS evaluations, {model.vjp_evals} VJPs", flush=True)
        if name == "fp64":
            rng = np.random.default_rng(11)
            errs = []
            with torch.no_grad():
                for _ in range(2):
                    w = torch.from_numpy(rng.standard_normal(model.T)).to(DEV) * 8e5
                    hh = 1e-6
                    Jp = float(mean_mz(model, integrate(model, m0, Ms.detach() + hh * w, dt, steps, every)))
                    Jm = float(mean_mz(model, integrate(model, m0, Ms.detach() - hh * w, dt, steps, every)))
                    fd = (Jp - Jm) / (2 * hh)
                    an = float(g @ w)
                    errs.append(abs(fd - an) / abs(an))
            print(f"[cube adjoint fp64] central differences: {errs}", flush=True)
            assert max(errs) <= 1e-6
        del model
        torch.cuda.empty_cache()
    e = float(torch.linalg.norm(grads["fp32"] - grads["fp64"]) / torch.linalg.norm(grads["fp64"]))
    print(f"[cube adjoint] FP32-boundary gradient vs FP64: {e:.2e}", flush=True)
    assert e <= 1e-6
