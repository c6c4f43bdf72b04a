"""AMG-preconditioned CG through the C ABI.

* smx_amg_apply (one V-cycle) against the host recursion of the same folded
  operators (smx.amg.apply_vcycle): FP64 values to 1e-13, FP32 values to the
  FP32 rounding of the operators (1e-6), on 1, 2 and 3+ level hierarchies;
* smx_pcg_amg solutions against a direct sparse solve (scipy splu) for the
  Neumann (u1) and Dirichlet (u2) operators, on the Kuhn stencil (width 7) and
  a jittered mesh (width 15), with and without a solver row permutation, FP64
  and FP32 hierarchies: rtol 1e-12, bar 1e-9 relative;
* iteration counts far below Jacobi CG on the same system;
* determinism (two solves bitwise equal) and the max_it / zero right-hand side
  edge cases.
"""
import copy
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla
import torch

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")


def _mesh(jitter, dims=(20, 16, 12)):
    from smx import mesh as M
    nx, ny, nz = dims
    m = M.kuhn_box(nx, ny, nz, 3.0 * nx, 3.0 * ny, 3.0 * nz, name="box")
    if jitter:
        rng = np.random.default_rng(3)
        X = m.X + rng.uniform(-0.6, 0.6, m.X.shape)
        m = M.TetMesh(X=X, tets=m.tets, name="boxj")
    return m


def _csr(fem, vals, perm=None):
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


def _hier(S, fp32=False, max_coarse=1024):
    from smx.amg import AmgHierarchy
    return AmgHierarchy(S, DEV, fp32=fp32, max_coarse=max_coarse)


def _apply(h, r):
    from smx import _capi
    L = _capi.lib()
    n = h.sizes[0]
    rd = torch.from_numpy(r).to(DEV)
    u = torch.zeros(n, dtype=torch.float64, device=DEV)
    ws = torch.zeros(int(L.smx_pcg_amg_ws_size(n)), dtype=torch.float64, device=DEV)
    _capi.check(L.smx_amg_apply(C.addressof(h.c_levels), h.nlev, h.val_dtype, rd.data_ptr(), u.data_ptr(),
                                ws.data_ptr(), _capi.stream_ptr()), "smx_amg_apply")
    torch.cuda.synchronize()
    return u.cpu().numpy()


@pytest.mark.parametrize("max_coarse,min_levels", [(4096, 1), (1024, 2), (20, 3)])
@pytest.mark.parametrize("fp32", [False, True])
def test_vcycle_matches_host(max_coarse, min_levels, fp32):
    from smx.amg import apply_vcycle
    from smx.fem import assemble_fem
    m = _mesh(False, (16, 14, 10))
    fem = assemble_fem(m.X, m.tets)
    S = _csr(fem, fem.stiff)
    h = _hier(S, fp32=fp32, max_coarse=max_coarse)
    assert h.nlev >= min_levels
    host = h.host
    if fp32:
        # the device reads FP32 values and accumulates in FP64: emulate exactly that
        host = copy.deepcopy(h.host)
        for L in host:
            for name in ("S", "Q", "R"):
                M = getattr(L, name)
                if M is not None:
                    M = M.copy()
                    M.data = M.data.astype(np.float32).astype(np.float64)
                    setattr(L, name, M)
    rng = np.random.default_rng(5)
    for _ in range(2):
        r = rng.standard_normal(fem.N)
        got = _apply(h, r)
        ref = apply_vcycle(host, r)
        err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        ref64 = apply_vcycle(h.host, r)
        err64 = np.linalg.norm(got - ref64) / np.linalg.norm(ref64)
        print(f"[amg apply] levels {h.describe()} fp32={fp32}: rel {err:.2e} (vs FP64 operators {err64:.2e})",
              flush=True)
        assert err <= 1e-13
        assert err64 <= (1e-6 if fp32 else 1e-13)


def _solve(A, h, b, perm, mode=0, rtol=1e-12, raw=None, bpos=None, gb=None, max_it=500, project=1):
    from smx import _capi
    L = _capi.lib()
    n = A.nrows
    x = torch.zeros(n, dtype=torch.float64, device=DEV)
    ws = torch.zeros(int(L.smx_pcg_amg_ws_size(n)), dtype=torch.float64, device=DEV)
    bd = None if b is None else torch.from_numpy(b).to(DEV)
    bpart = None if b is None else torch.tensor([float(b.sum())], dtype=torch.float64, device=DEV)
    pd = None if perm is None else torch.from_numpy(perm.astype(np.int32)).to(DEV)
    raw = raw or A
    its, rn, sums = C.c_int32(0), C.c_double(0.0), (C.c_double * 1)()
    rc = L.smx_pcg_amg(n, 0, A.width, A.ld, _capi.ptr(A.cols), _capi.ptr(A.vals), mode, _capi.ptr(bd),
                       _capi.ptr(bpart), 0 if b is None else 1, project, raw.width, raw.ld, _capi.ptr(raw.cols),
                       _capi.ptr(raw.vals), _capi.ptr(bpos), _capi.ptr(gb), _capi.ptr(x), 0, _capi.ptr(pd),
                       C.addressof(h.c_levels), h.nlev, h.val_dtype, rtol, 0.0, max_it, 8, 0,
                       _capi.ptr(ws), C.byref(its), C.byref(rn), sums, _capi.stream_ptr())
    torch.cuda.synchronize()
    return rc, its.value, x.cpu().numpy(), sums[0]


def _order(m, fem, reorder):
    from smx.demag import OperatorSet
    return OperatorSet(m, DEV, fem).solver_order() if reorder else None


@pytest.mark.parametrize("jitter", [False, True])
@pytest.mark.parametrize("reorder", [False, True])
@pytest.mark.parametrize("fp32", [False, True])
def test_neumann_solve_matches_direct(jitter, reorder, fp32):
    from smx.fem import assemble_fem
    m = _mesh(jitter)
    fem = assemble_fem(m.X, m.tets)
    perm = _order(m, fem, reorder)
    S0 = _csr(fem, fem.stiff)
    S = _csr(fem, fem.stiff, perm)
    A = _ell(S)
    assert A.width == (15 if jitter else 7)
    h = _hier(S, fp32=fp32, max_coarse=256)
    assert h.nlev >= 2
    b = np.random.default_rng(2).standard_normal(fem.N)
    rc, its, x, sumx = _solve(A, h, b, perm)
    print(f"[amg neumann] jitter={jitter} reorder={reorder} fp32={fp32} levels {h.describe()} its {its}",
          flush=True)
    assert rc == 0 and 0 < its <= 40
    assert abs(sumx - x.sum()) <= 1e-9 * np.abs(x).sum()
    bp = b - b.mean()
    Sp = S0.tolil()
    Sp[0, :] = 0.0
    Sp[0, 0] = 1.0
    rhs = bp.copy()
    rhs[0] = 0.0
    x_ref = spla.splu(Sp.tocsc()).solve(rhs)
    x_ref -= x_ref.mean()
    xc = x - x.mean()
    assert np.linalg.norm(xc - x_ref) <= 1e-9 * np.linalg.norm(x_ref)


@pytest.mark.parametrize("jitter", [False, True])
@pytest.mark.parametrize("reorder", [False, True])
def test_dirichlet_lifting_solve_matches_direct(jitter, reorder):
    """Mode 1: boundary data gb on the boundary rows, interior rows solve the
    raw stiffness with that data lifted to the right side."""
    from smx.fem import assemble_fem
    m = _mesh(jitter)
    fem = assemble_fem(m.X, m.tets)
    perm = _order(m, fem, reorder)
    sel = perm if perm is not None else np.arange(fem.N)
    Sraw = _csr(fem, fem.stiff, perm)
    Sbc = _csr(fem, fem.stiffness_bc(), perm)
    h = _hier(Sbc, max_coarse=256)
    bn = fem.bnodes
    g = np.random.default_rng(4).standard_normal(bn.size)
    pos_of = -np.ones(fem.N, dtype=np.int32)
    pos_of[bn] = np.arange(bn.size)
    bpos = torch.from_numpy(pos_of[sel].astype(np.int32)).to(DEV)
    gb = torch.from_numpy(g).to(DEV)
    rc, its, x, _ = _solve(_ell(Sbc), h, None, perm, mode=1, raw=_ell(Sraw), bpos=bpos, gb=gb, project=0)
    print(f"[amg dirichlet] jitter={jitter} reorder={reorder} levels {h.describe()} its {its}", flush=True)
    assert rc == 0 and 0 < its <= 40
    # reference: interior block solve with the boundary values lifted
    Sfull = _csr(fem, fem.stiff)
    inner = np.nonzero(~fem.is_boundary)[0]
    xb = np.zeros(fem.N)
    xb[bn] = g
    rhs = -(Sfull @ xb)[inner]
    xi = spla.splu(Sfull[inner][:, inner].tocsc()).solve(rhs)
    x_ref = xb.copy()
    x_ref[inner] = xi
    assert np.array_equal(x[bn], g)
    assert np.linalg.norm(x - x_ref) <= 1e-9 * np.linalg.norm(x_ref)


def test_far_fewer_iterations_than_jacobi_and_deterministic():
    from smx.fem import assemble_fem
    m = _mesh(False, (40, 32, 8))
    fem = assemble_fem(m.X, m.tets)
    S = _csr(fem, fem.stiff)
    A = _ell(S)
    h = _hier(S)
    b = np.random.default_rng(7).standard_normal(fem.N)
    rc1, its1, x1, _ = _solve(A, h, b, None, rtol=1e-6)
    rc2, its2, x2, _ = _solve(A, h, b, None, rtol=1e-6)
    assert rc1 == rc2 == 0 and its1 == its2
    assert np.array_equal(x1, x2)                    # fixed-order reductions
    bp = b - b.mean()
    d = S.diagonal()
    r = bp.copy()
    z = r / d
    p = z.copy()
    rz = r @ z
    its_j = 0
    x = np.zeros_like(bp)
    while np.linalg.norm(r) > 1e-6 * np.linalg.norm(bp) and its_j < 5000:
        q = S @ p
        a = rz / (p @ q)
        x += a * p
        r -= a * q
        z = r / d
        rz2 = r @ z
        p = z + rz2 / rz * p
        rz = rz2
        its_j += 1
    print(f"[amg vs jacobi] {h.describe()}: AMG {its1} its, Jacobi {its_j} its", flush=True)
    assert its1 * 5 < its_j


def test_edge_cases():
    from smx import _capi
    from smx.fem import assemble_fem
    m = _mesh(False, (10, 8, 6))
    fem = assemble_fem(m.X, m.tets)
    S = _csr(fem, fem.stiff)
    A = _ell(S)
    h = _hier(S, max_coarse=64)
    rc, its, x, _ = _solve(A, h, np.zeros(fem.N), None)   # zero right-hand side: converged at x0 = 0
    assert rc == 0 and its == 0 and not np.any(x)
    b = np.random.default_rng(0).standard_normal(fem.N)
    rc, its, x, _ = _solve(A, h, b, None, max_it=1)
    assert rc == _capi.SMX_ERR_NOT_CONVERGED and its == 1
    assert "did not converge" in _capi.lib().smx_last_error().decode()
