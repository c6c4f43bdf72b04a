"""Mixed-precision Poisson solves (smx_pcg_jacobi_mixed, smx_pcg_amg_mixed):
FP32 inner CG, FP64 iterative refinement, FP64 compensated inner products.

* Both solves of the demag field (u1: Neumann, mode 0; u2: Dirichlet lifting,
  mode 1), both preconditioners (AMG, Jacobi), on a 14 x 12 x 10 Kuhn box:
  the true FP64 residual of the mixed solution (recomputed on the host) meets
  the requested tolerance, for rtol 1e-6 (the FP32 solve to rtol / 2 and one
  FP64 check) and 1e-10 (the FP32 solve to 1e-4, then FP64 checks and correction
  solves to 1e-8 and 1e-10; the solution within 1e-8 of the FP64 solver's at
  1e-13).
* Without refinement the FP32 solve still converges in its recurrence, and
  its true residual is within 10x the tolerance at rtol 1e-6.
* The waiting call reports the iteration count of all its solves; the
  speculative launch (no host wait) gives the same bits when the guessed
  counts suffice.
* The demag field of the 40 nm cube with mixed solves: against the FP64
  solver at the workload's rtol 1e-6 (1e-6), and against the oracle's LU
  solves at rtol 1e-12 with two refinements (1e-9).
"""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

import smx_oracle as O

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")


def _box():
    from smx import mesh as M
    return M.kuhn_box(14, 12, 10, 42.0, 36.0, 30.0, name="box14")


def _demag(m, pc, mixed, refinements=1, rtol=1e-6, **kw):
    from smx.demag import DemagFieldLindholmGPU
    return DemagFieldLindholmGPU(m, None, None, 8e5, ksp_rtol_u1=rtol, ksp_rtol_u2=rtol, ksp_atol=0.0,
                                 ksp_pc_type=pc, cg_mixed=mixed, cg_refinements=refinements,
                                 cg_speculative=False, amg_fp32=mixed, ksp_max_it_u1=20000,
                                 ksp_max_it_u2=20000, **kw)


def _stiff(d):
    fem = d.ops.fem
    g = fem.graph
    r = np.repeat(np.arange(fem.N), np.diff(g.indptr))
    return sp.csr_matrix((fem.stiff, (r, g.indices)), shape=(fem.N, fem.N))


def _rhs(d, seed):
    g = np.random.default_rng(seed).standard_normal((d.N, 3))
    m = torch.from_numpy((g / np.linalg.norm(g, axis=1)[:, None]).ravel()).to(DEV)
    d.Div.spmv(m, d.b_u1, kind=1, partials=d.b_part)
    gb = torch.from_numpy(np.random.default_rng(seed + 1).standard_normal(d.Nb)).to(DEV)
    return d.b_u1.clone(), gb


def _residuals(d, A, b, x1, gb, x2):
    """True FP64 residuals: u1 (Neumann, mean-free b) and u2 (interior rows of
    A x = 0 with x = gb on the boundary)."""
    bh = b.cpu().numpy()
    bh = bh - bh.mean()
    r1 = np.linalg.norm(bh - A @ x1) / np.linalg.norm(bh)
    bn = d.ops.fem.bnodes
    gbh = np.zeros(d.N)
    gbh[d.opB.bverts] = gb.cpu().numpy()
    interior = np.setdiff1d(np.arange(d.N), bn)
    assert np.abs(x2[bn] - gbh[bn]).max() <= 1e-12 * np.abs(gbh).max()
    lifted = A[interior][:, bn] @ gbh[bn]
    # the solver's ||b|| holds the boundary rows (data) and the lifted interior rows
    bnorm = np.sqrt(np.linalg.norm(gbh[bn]) ** 2 + np.linalg.norm(lifted) ** 2)
    r2 = np.linalg.norm((A @ x2)[interior]) / bnorm
    return r1, r2


def _solve_both(d, b, gb):
    its1 = d.solve("u1", 0, d.u1, d.ws1, b=b, b_part=d.b_part)
    y = torch.zeros(d.Nb, dtype=torch.float64, device=DEV)
    y.copy_(gb)
    its2 = d.solve("u2", 1, d.u2, d.ws2, gb=y, project=0)
    torch.cuda.synchronize()
    return its1, its2, d.u1.cpu().numpy().copy(), d.u2.cpu().numpy().copy()


@pytest.mark.parametrize("pc", ["gamg", "jacobi"])
@pytest.mark.parametrize("rtol,refinements,rtol_inner", [(1e-6, 1, None), (1e-10, 2, 1e-4)])
def test_mixed_solves_meet_tolerance(pc, rtol, refinements, rtol_inner):
    m = _box()
    d = _demag(m, pc, True, refinements, rtol, cg_rtol_inner=rtol_inner)
    A = _stiff(d)
    b, gb = _rhs(d, 3)
    its1, its2, x1, x2 = _solve_both(d, b, gb)
    r1, r2 = _residuals(d, A, b, x1, gb, x2)
    split1, split2 = d._its_split["u1"], d._its_split["u2"]
    print(f"[mixed cg {pc} rtol={rtol:g} ref={refinements}] its u1 {its1} {split1} u2 {its2} {split2}; "
          f"true residuals {r1:.2e} {r2:.2e}", flush=True)
    assert sum(split1) == its1 and sum(split2) == its2
    if rtol_inner is not None:      # the FP32 solve stops at 1e-4: the corrections must run
        assert split1[0] > 0 and split2[0] > 0
    assert r1 <= 1.5 * rtol and r2 <= 1.5 * rtol
    if rtol < 1e-8:
        d64 = _demag(m, pc, False, rtol=1e-13)
        _, _, y1, y2 = _solve_both(d64, b, gb)
        e1 = np.linalg.norm((x1 - x1.mean()) - (y1 - y1.mean())) / np.linalg.norm(y1 - y1.mean())
        e2 = np.linalg.norm(x2 - y2) / np.linalg.norm(y2)
        print(f"  vs FP64 solver at 1e-13: u1 {e1:.2e} u2 {e2:.2e}", flush=True)
        assert e1 <= 1e-8 and e2 <= 1e-8


@pytest.mark.parametrize("pc", ["gamg", "jacobi"])
def test_single_fp32_solve_without_refinement(pc):
    m = _box()
    d = _demag(m, pc, True, refinements=0, rtol=1e-6)
    A = _stiff(d)
    b, gb = _rhs(d, 5)
    its1, its2, x1, x2 = _solve_both(d, b, gb)
    r1, r2 = _residuals(d, A, b, x1, gb, x2)
    print(f"[fp32 cg {pc}, no refinement] its {its1} {its2}; true residuals {r1:.2e} {r2:.2e}", flush=True)
    assert d._its_split["u1"][0] == 0 and r1 <= 1e-5 and r2 <= 1e-5


@pytest.mark.parametrize("pc", ["gamg", "jacobi"])
def test_speculative_equals_waiting(pc):
    m = _box()
    d = _demag(m, pc, True, refinements=1, rtol=1e-6)
    b, gb = _rhs(d, 7)
    _solve_both(d, b, gb)
    w1, w2 = d.u1.clone(), d.u2.clone()
    e1, e2 = d._its_split["u1"], d._its_split["u2"]
    d.u1.zero_()
    d.u2.zero_()
    y = gb.clone()
    assert d.solve("u1", 0, d.u1, d.ws1, b=b, b_part=d.b_part, expected=sum(e1), speculative=True) == -1
    assert d.solve("u2", 1, d.u2, d.ws2, gb=y, project=0, expected=sum(e2), speculative=True) == -1
    torch.cuda.synchronize()
    s1, s2 = d.ws1[:8].cpu().numpy(), d.ws2[:8].cpu().numpy()
    # the FP32 solves run to rtol / 2: the FP64 checks pass without a correction
    assert e1[0] == 0 and e2[0] == 0
    assert s1[0] == 1.0 and s2[0] == 1.0 and s1[5] == 0.0 and s2[5] == 0.0
    assert (int(s1[7]), int(s1[1])) == e1 and (int(s2[7]), int(s2[1])) == e2
    assert torch.equal(d.u1, w1) and torch.equal(d.u2, w2)


@pytest.mark.parametrize("pc", ["gamg", "jacobi"])
def test_demag_field_mixed_cg(pc):
    """40 nm cube: mixed solves against the FP64 solver at the workload's
    rtol 1e-6 (1e-6), and, at rtol 1e-12 with two refinements, against the
    oracle's LU solves (1e-9)."""
    from smx import mesh as M
    m = M.cube_c1()
    g = np.random.default_rng(11).standard_normal((m.num_nodes, 3))
    mv = (g / np.linalg.norm(g, axis=1)[:, None]).ravel()
    mt = torch.from_numpy(mv).to(DEV)

    def field(d):
        H = torch.empty(3 * m.num_nodes, dtype=torch.float64, device=DEV)
        d.compute_vec(mt, H)
        torch.cuda.synchronize()
        return H.cpu().numpy()

    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
    e_64 = rel(field(_demag(m, pc, True, 1, 1e-6)), field(_demag(m, pc, False, rtol=1e-6)))
    ref = O.DemagOracle(m.X, m.tets, 8e5).compute(mv)
    e_or = rel(field(_demag(m, pc, True, 2, 1e-12)), ref)
    print(f"[demag mixed cg c1 {pc}] vs FP64 solver (rtol 1e-6) {e_64:.2e}; vs oracle (rtol 1e-12, "
          f"2 refinements) {e_or:.2e}", flush=True)
    assert e_64 <= 1e-6 and e_or <= 1e-9
