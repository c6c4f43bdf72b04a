"""Dirichlet solves with trailing identity rows (smx_pcg_* nd > 0).

The u2 operator A_bc has identity rows on the boundary nodes and its interior
rows do not couple to the boundary columns.  With the interior rows first in
the solver order and the nd boundary rows after them, the solvers iterate the
interior block alone and write the boundary rows' data once.  Checked through
the C ABI on a Kuhn box (width-7 stencil) and a jittered one (width 15):

* mode 1 (lifting of boundary data gb) and mode 0 (right side b, whose
  boundary entries become x there) against a direct sparse solve of the full
  system (scipy splu), rtol 1e-12, bar 1e-9 relative, for the Jacobi and the
  AMG preconditioner, the FP64 and the mixed-precision forms;
* the boundary entries of x equal their data bit for bit;
* the same solve over the full system (nd = 0, identity rows iterated) reaches
  the same solution; the split solve's tolerance counts the boundary rows in
  ||b|| like the full system's, so both stop within an iteration or two.
"""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla
import torch

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")


def _problem(jitter):
    from smx import mesh as M
    from smx.fem import assemble_fem, compact_csr
    m = M.kuhn_box(14, 12, 10, 42.0, 36.0, 30.0, name="box14")
    if jitter:
        rng = np.random.default_rng(3)
        m = M.TetMesh(X=m.X + rng.uniform(-0.6, 0.6, m.X.shape), tets=m.tets, name="box14j")
    fem = assemble_fem(m.X, m.tets)
    g = fem.graph

    def csr(vals):
        ip, ix, vv = compact_csr(g.indptr, g.indices, vals)
        S = sp.csr_matrix((vv, ix, ip), shape=(fem.N, fem.N))
        S.sort_indices()
        return S

    return fem, csr(fem.stiff), csr(fem.stiffness_bc())


def _ell(S, rows):
    from smx.device import DeviceEll
    S = sp.csr_matrix(S)
    S.sort_indices()
    return DeviceEll.from_csr(S.indptr, S.indices, S.data, rows, DEV)


def _d(a, dt=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV).to(dt)


def _solve(pc, mode, split, jitter, mixed=False, rtol=1e-12):
    """One solve through the C ABI; returns (x in node order, iterations, the
    direct solution, the boundary mask)."""
    from smx import _capi
    from smx.amg import AmgHierarchy
    from smx.demag import solver_operator
    fem, Araw, Abc = _problem(jitter)
    N = fem.N
    isb = fem.is_boundary
    rng = np.random.default_rng(11)
    # a scrambled caller order, so perm is not the identity
    base = rng.permutation(N)
    order = np.concatenate([base[~isb[base]], base[isb[base]]]) if split else base
    n = int(np.count_nonzero(~isb)) if split else N
    nd = N - n
    Pbc = Abc[order][:, order].tocsr()
    Pbc.sort_indices()
    if split:
        assert Pbc[:n, n:].nnz == 0
    A_it = Pbc[:n, :n].tocsr()
    raw = Araw[order][:, order].tocsr()[:n]
    bpos = np.full(N, -1, dtype=np.int64)
    bnodes = np.nonzero(isb)[0]
    bpos[bnodes] = np.arange(bnodes.size)
    gb = rng.standard_normal(bnodes.size)
    b = rng.standard_normal(N)
    # direct solution of the full system, node order
    if mode == 1:
        rhs = -(Araw[:, bnodes] @ gb)
        rhs[isb] = gb
    else:
        rhs = b.copy()
    x_ref = spla.splu(sp.csc_matrix(Abc)).solve(rhs)

    L = _capi.lib()
    A = _ell(A_it, n)
    R = _ell(raw, n)
    dinv = _d(1.0 / A_it.diagonal())
    x = torch.zeros(N, dtype=torch.float64, device=DEV)
    ws = torch.zeros(int(max(L.smx_pcg_ws_size(N), L.smx_pcg_amg_ws_size(N))), dtype=torch.float64, device=DEV)
    perm = _d(order.astype(np.int32), torch.int32)
    bp = _d(bpos[order].astype(np.int32), torch.int32)
    gbd, bd = _d(gb), _d(b)
    its, rn, sums = C.c_int32(0), C.c_double(0.0), (C.c_double * 1)()
    keep = []
    if pc == "jacobi":
        Aj = solver_operator(A, dinv)
        mx = None
        if mixed:
            v32 = Aj.vals.float()
            keep.append(v32)
            mx = _capi.CGMixed(v32.data_ptr(), 0.5 * rtol, 1, 0)
        args = [n, nd, Aj.width, Aj.ld, _capi.ptr(Aj.cols), _capi.ptr(Aj.vals), _capi.ptr(dinv), mode,
                _capi.ptr(bd), None, 0, 0, R.width, R.ld, _capi.ptr(R.cols), _capi.ptr(R.vals), _capi.ptr(bp),
                _capi.ptr(gbd), _capi.ptr(x), 0, _capi.ptr(perm), rtol, 0.0, 20000, 32, 0, _capi.ptr(ws),
                C.byref(its), C.byref(rn), sums]
        rc = L.smx_pcg_jacobi_mixed(*args, C.byref(mx), _capi.stream_ptr()) if mixed else \
            L.smx_pcg_jacobi(*args, _capi.stream_ptr())
    else:
        # coarse levels up to 256 rows: with the default dense coarse size
        # the interior block of this box would be one dense level (an exact
        # solve in one iteration) and the full system not
        h = AmgHierarchy(A_it, DEV, fp32=mixed, max_coarse=256)
        assert h.sizes[0] == n and len(h.sizes) >= 2
        mx = None
        if mixed:
            v32 = A.vals.float()
            keep.append(v32)
            mx = _capi.CGMixed(v32.data_ptr(), 0.5 * rtol, 1, 0)
        args = [n, nd, A.width, A.ld, _capi.ptr(A.cols), _capi.ptr(A.vals), mode, _capi.ptr(bd), None, 0, 0,
                R.width, R.ld, _capi.ptr(R.cols), _capi.ptr(R.vals), _capi.ptr(bp), _capi.ptr(gbd), _capi.ptr(x),
                0, _capi.ptr(perm), C.addressof(h.c_levels), h.nlev, h.val_dtype, rtol, 0.0, 2000, 8, 0,
                _capi.ptr(ws), C.byref(its), C.byref(rn), sums]
        rc = L.smx_pcg_amg_mixed(*args, C.byref(mx), _capi.stream_ptr()) if mixed else \
            L.smx_pcg_amg(*args, _capi.stream_ptr())
    torch.cuda.synchronize()
    _capi.check(rc, "smx_pcg")
    return x.cpu().numpy(), its.value, x_ref, isb, (gb, b, bpos)


@pytest.mark.parametrize("jitter", [False, True])
@pytest.mark.parametrize("pc", ["jacobi", "amg"])
@pytest.mark.parametrize("mode", [1, 0])
def test_split_solve_matches_direct_solve(pc, mode, jitter):
    print(f"[dirichlet rows] pc={pc} mode={mode} jitter={jitter}", flush=True)
    x, its, x_ref, isb, (gb, b, bpos) = _solve(pc, mode, True, jitter)
    rel = np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref)
    assert rel <= 1e-9, rel
    data = gb[bpos[isb]] if mode == 1 else b[isb]
    assert np.array_equal(x[isb], data)           # written, not iterated
    xf, its_full, _, _, _ = _solve(pc, mode, False, jitter)
    assert np.linalg.norm(xf - x_ref) / np.linalg.norm(x_ref) <= 1e-9
    assert abs(its - its_full) <= 3, (its, its_full)


@pytest.mark.parametrize("pc", ["jacobi", "amg"])
def test_split_solve_mixed_precision(pc):
    print(f"[dirichlet rows] mixed pc={pc}", flush=True)
    x, its, x_ref, isb, (gb, b, bpos) = _solve(pc, 1, True, False, mixed=True, rtol=1e-10)
    rel = np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref)
    assert rel <= 1e-7, rel
    assert np.array_equal(x[isb], gb[bpos[isb]])
