"""Speculative CG launches in the demag potential solves.

Once both iteration counts are known, each solve enqueues last count +
margin iterations without waiting and the evaluation checks both status
blocks once at its end.  Iterations past convergence return at once on the
device's done flag, so every result must equal the waiting path's bit for
bit; an iteration guess that falls short must be caught and repaired (the
solve runs again, waiting).  Both preconditioners (AMG and Jacobi), the
dense and the H-matrix boundary operator.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")


def _unit(n, seed):
    g = np.random.default_rng(seed).standard_normal((n, 3))
    return (g / np.linalg.norm(g, axis=1)[:, None]).ravel()


@pytest.mark.parametrize("pc,backend", [("gamg", "dense"), ("jacobi", "dense"), ("gamg", "hmatrix")])
def test_speculative_equals_waiting(pc, backend):
    from smx import mesh as M
    from smx.llg import EffectiveField
    m = M.cube_c1()
    N = m.num_nodes
    # the near-row mode fixed: its set-up timing choice changes the last bits;
    # a multi-level AMG (this mesh's 1331 rows would otherwise be one dense
    # coarse level, solved in one iteration, which no guess can fall short of)
    dk = dict(ksp_pc_type=pc, boundary_backend=backend, ksp_rtol_u1=1e-6, ksp_rtol_u2=1e-6, ksp_atol=1e-12,
              hmatrix_near_in_z=True, amg_max_coarse=256)
    kw = dict(Ms=8e5, Aex=1.3e-11, Ku=5e5, n_ani_vec=(0.0, 0.0, 1.0), alpha=0.5, use_demag=True, device=DEV)
    hs = EffectiveField(m, demag_kwargs=dict(cg_speculative=True, **dk), **kw)
    hw = EffectiveField(m, demag_kwargs=dict(cg_speculative=False, **dk), **kw)
    states = [torch.from_numpy(_unit(N, s)).to(DEV) for s in range(6)]
    rs, rw = torch.empty_like(states[0]), torch.empty_like(states[0])
    for k, x in enumerate(states):
        if k == 4:
            hs.demag_field.last_its_u1 = 1      # a guess that falls short: must be repaired
        hs.llg_rhs_vec(x, rs)
        hw.llg_rhs_vec(x, rw)
        torch.cuda.synchronize()
        assert torch.equal(rs, rw), k
        assert (hs.demag_field.last_its_u1, hs.demag_field.last_its_u2) == \
            (hw.demag_field.last_its_u1, hw.demag_field.last_its_u2), k
    d = hs.demag_field
    print(f"[speculative {pc}/{backend}] its {d.last_its_u1}/{d.last_its_u2}, redos {d.speculative_redos}",
          flush=True)
    assert d.speculative_redos >= 1
    # H_demag through compute_vec, and the energy path
    hd_s, hd_w = torch.empty_like(rs), torch.empty_like(rs)
    hs.demag_field.compute_vec(states[5], hd_s)
    hw.demag_field.compute_vec(states[5], hd_w)
    torch.cuda.synchronize()
    assert torch.equal(hd_s, hd_w)
    assert hs.compute_Energy_terms(states[2])["E_demag"] == hw.compute_Energy_terms(states[2])["E_demag"]


def test_overlapped_local_terms_equal_fused():
    """Local terms on the side stream + demag/RHS kernel against the single
    fused kernel, all terms on (exchange, DMI, uniaxial, cubic, Zhang-Li,
    SOT, static field): the two sum the field in another order, so equal to
    rounding (1e-14 relative); the RHS, H_eff and H_demag."""
    from smx import mesh as M
    from smx.llg import EffectiveField
    m = M.cube_c1()
    N = m.num_nodes
    kw = dict(Ms=8e5, Aex=1.3e-11, Ku=5e5, n_ani_vec=(0.0, 0.6, 0.8), D_int=2e-3, n0_int_vec=(0.0, 0.0, 1.0),
              Kc1=-4e4, u1_cub=(1.0, 0.0, 0.0), u2_cub=(0.0, 1.0, 0.0), Jmagnitude=5e11, Jdir_vec=(1.0, 0.0, 0.0),
              P=0.6, beta=0.04, H_dl=2e4, H_fl=-5e3, polarization_vec=(0.0, 1.0, 0.0), H0_static=(1e4, 0.0, 0.0),
              alpha=0.3, use_demag=True, device=DEV, demag_kwargs=dict(ksp_rtol_u1=1e-12, ksp_rtol_u2=1e-12))
    ho = EffectiveField(m, overlap_local_terms=True, **kw)
    hf = EffectiveField(m, overlap_local_terms=False, **kw)
    assert ho.overlap_local_terms and not hf.overlap_local_terms
    for seed in range(3):
        x = torch.from_numpy(_unit(N, 20 + seed)).to(DEV)
        ro, rf = torch.empty_like(x), torch.empty_like(x)
        ho.llg_rhs_vec(x, ro)
        hf.llg_rhs_vec(x, rf)
        torch.cuda.synchronize()
        for a, b in ((ro, rf), (ho.H_eff_gpu, hf.H_eff_gpu), (ho.H_demag_gpu, hf.H_demag_gpu)):
            assert float(torch.linalg.norm(a - b) / torch.linalg.norm(b)) <= 1e-14
        ha, hb = torch.empty_like(x), torch.empty_like(x)
        ho.compute_H_eff_vec(x, ha)
        hf.compute_H_eff_vec(x, hb)
        torch.cuda.synchronize()
        assert float(torch.linalg.norm(ha - hb) / torch.linalg.norm(hb)) <= 1e-14
