"""Jacobi CG through the C ABI.

* smx_pcg_prepare against its definition (diagonal moved to slot 0, values
  a_ij * dinv_j): bit-exact, out of place and in place;
* an operator that is not in solver form is rejected before any iteration;
* CG solutions against a direct sparse solve (scipy splu) on the Kuhn stencil
  (width 7: the exact-width kernel) and on a jittered Kuhn mesh, whose P1
  stiffness keeps all 15 neighbours (width 15: the predicated kernel), with
  and without a solver row permutation.  rtol 1e-12 gives a solution error
  bounded by cond(A) * 1e-12; the bar is 1e-9 relative.
"""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla
import torch

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")


def _mesh(jitter):
    from smx import mesh as M
    m = M.kuhn_box(14, 12, 10, 42.0, 36.0, 30.0, name="box14")
    if jitter:
        rng = np.random.default_rng(3)
        X = m.X + rng.uniform(-0.6, 0.6, m.X.shape)   # 3 nm cells, +-20 %
        m = M.TetMesh(X=X, tets=m.tets, name="box14j")
    return m


def _csr(fem, vals, perm):
    from smx.fem import compact_csr
    g = fem.graph
    ip, ix, vv = compact_csr(g.indptr, g.indices, vals)
    S = sp.csr_matrix((vv, ix, ip), shape=(fem.N, fem.N))
    if perm is not None:
        S = S[perm][:, perm].tocsr()
    S.sort_indices()
    return S


def _ell(S):
    from smx.device import DeviceEll
    return DeviceEll.from_csr(S.indptr, S.indices, S.data, S.shape[0], DEV)


def _prepare(A, dinv, out=None):
    from smx import _capi
    from smx.device import DeviceEll
    if out is None:
        out = DeviceEll(nrows=A.nrows, width=A.width, ld=A.ld, cols=torch.empty_like(A.cols),
                        vals=torch.empty_like(A.vals), nnz=A.nnz)
    _capi.check(_capi.lib().smx_pcg_prepare(A.nrows, A.width, A.ld, _capi.ptr(A.cols), _capi.ptr(A.vals),
                                            _capi.ptr(dinv), _capi.ptr(out.cols), _capi.ptr(out.vals),
                                            _capi.stream_ptr()), "smx_pcg_prepare")
    return out


def _solve(A, dinv, b, perm, rtol=1e-12, project=0):
    """smx_pcg_jacobi mode 0; returns (rc, iterations, x in caller order)."""
    from smx import _capi
    L = _capi.lib()
    n = A.nrows
    x = torch.zeros(n, dtype=torch.float64, device=DEV)
    ws = torch.zeros(int(L.smx_pcg_ws_size(n)), dtype=torch.float64, device=DEV)
    bd = torch.from_numpy(b).to(DEV)
    bpart = torch.tensor([float(b.sum())], dtype=torch.float64, device=DEV)
    pd = None if perm is None else torch.from_numpy(perm.astype(np.int32)).to(DEV)
    its, rn, sums = C.c_int32(0), C.c_double(0.0), (C.c_double * 1)()
    rc = L.smx_pcg_jacobi(n, 0, A.width, A.ld, _capi.ptr(A.cols), _capi.ptr(A.vals), _capi.ptr(dinv), 0,
                          _capi.ptr(bd), _capi.ptr(bpart), 1, project, A.width, A.ld, _capi.ptr(A.cols),
                          _capi.ptr(A.vals), None, None, _capi.ptr(x), 0, _capi.ptr(pd), rtol, 0.0, 20000, 32,
                          0, _capi.ptr(ws), C.byref(its), C.byref(rn), sums, _capi.stream_ptr())
    torch.cuda.synchronize()
    return rc, its.value, x.cpu().numpy()


def _expected_solver_form(A, dinv):
    W, ld, n = A.width, A.ld, A.nrows
    cols = A.cols.cpu().numpy().reshape(W, ld).copy()
    vals = A.vals.cpu().numpy().reshape(W, ld).copy()
    di = dinv.cpu().numpy()
    for i in range(n):
        d = int(np.nonzero(cols[:, i] == i)[0][0])
        cols[[0, d], i] = cols[[d, 0], i]
        vals[[0, d], i] = vals[[d, 0], i]
    vals[:, :n] = vals[:, :n] * di[cols[:, :n]]
    return cols, vals


@pytest.mark.parametrize("jitter", [False, True])
def test_prepare_is_its_definition(jitter):
    from smx.fem import assemble_fem
    m = _mesh(jitter)
    fem = assemble_fem(m.X, m.tets)
    S = _csr(fem, fem.stiff, None)
    A = _ell(S)
    assert A.width == (15 if jitter else 7)
    dinv = torch.from_numpy(1.0 / S.diagonal()).to(DEV)
    ec, ev = _expected_solver_form(A, dinv)
    out = _prepare(A, dinv)
    torch.cuda.synchronize()
    n = A.nrows
    oc = out.cols.cpu().numpy().reshape(A.width, A.ld)
    ov = out.vals.cpu().numpy().reshape(A.width, A.ld)
    assert np.array_equal(oc[:, :n], ec[:, :n])
    assert np.array_equal(ov[:, :n], ev[:, :n])
    assert np.array_equal(oc[0, :n], np.arange(n))
    _prepare(A, dinv, out=A)                      # in place
    torch.cuda.synchronize()
    ic = A.cols.cpu().numpy().reshape(A.width, A.ld)
    iv = A.vals.cpu().numpy().reshape(A.width, A.ld)
    assert np.array_equal(ic[:, :n], oc[:, :n]) and np.array_equal(iv[:, :n], ov[:, :n])


def test_operator_not_in_solver_form_is_rejected():
    from smx import _capi
    from smx.fem import assemble_fem
    m = _mesh(False)
    fem = assemble_fem(m.X, m.tets)
    S = _csr(fem, fem.stiffness_bc(), None)
    A = _ell(S)                                   # plain A: sorted columns, diagonal not first
    dinv = torch.from_numpy(1.0 / S.diagonal()).to(DEV)
    b = np.random.default_rng(0).standard_normal(fem.N)
    rc, its, x = _solve(A, dinv, b, None)
    assert rc == _capi.SMX_ERR_VALUE
    assert its == 0
    assert "solver form" in _capi.lib().smx_last_error().decode()
    assert not np.any(x)


@pytest.mark.parametrize("jitter", [False, True])
@pytest.mark.parametrize("reorder", [False, True])
def test_dirichlet_solve_matches_direct(jitter, reorder):
    from smx.demag import OperatorSet
    from smx.fem import assemble_fem
    m = _mesh(jitter)
    fem = assemble_fem(m.X, m.tets)
    perm = OperatorSet(m, DEV, fem).solver_order() if reorder else None
    if reorder:
        assert not np.array_equal(perm, np.arange(fem.N))
    S0 = _csr(fem, fem.stiffness_bc(), None)      # caller order, for the direct solve
    S = _csr(fem, fem.stiffness_bc(), perm)
    dinv = torch.from_numpy(1.0 / S.diagonal()).to(DEV)
    A = _prepare(_ell(S), dinv)
    b = np.random.default_rng(1).standard_normal(fem.N)
    rc, its, x = _solve(A, dinv, b, perm)
    assert rc == 0 and its > 10
    x_ref = spla.splu(S0.tocsc()).solve(b)
    assert np.linalg.norm(x - x_ref) <= 1e-9 * np.linalg.norm(x_ref)


@pytest.mark.parametrize("jitter", [False, True])
def test_neumann_solve_matches_direct(jitter):
    """Singular Neumann operator, right side projected to mean zero: the
    solution is fixed up to a constant; compare with the mean removed."""
    from smx.fem import assemble_fem
    m = _mesh(jitter)
    fem = assemble_fem(m.X, m.tets)
    S = _csr(fem, fem.stiff, None)
    dinv = torch.from_numpy(1.0 / S.diagonal()).to(DEV)
    A = _prepare(_ell(S), dinv)
    b = np.random.default_rng(2).standard_normal(fem.N)
    rc, its, x = _solve(A, dinv, b, None, project=1)
    assert rc == 0
    bp = b - b.mean()
    Sp = S.tolil()
    Sp[0, :] = 0.0
    Sp[0, 0] = 1.0
    rhs = bp.copy()
    rhs[0] = 0.0
    x_ref = spla.splu(Sp.tocsc()).solve(rhs)      # one node pinned
    x_ref -= x_ref.mean()
    xc = x - x.mean()
    assert np.linalg.norm(xc - x_ref) <= 1e-9 * np.linalg.norm(x_ref)
