"""Multi-RHS dense boundary product on the matrix cores (smx_dense_bmatmul).

The kernel is checked against a float64 torch product of the same stored
operands: B in its storage precision widened to FP64, X rounded to the MFMA
input precision.  What remains is the accumulation error (FP32 accumulation
for BF16/FP32 storage, FP64 for FP64).  The error against the exact FP64
product is reported as the precision study of the dense boundary operator.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _geo(mesh):
    from smx.boundary import build_boundary_geometry
    return build_boundary_geometry(mesh.X, mesh.tets)


def _case(mesh, dtype, k, seed=0):
    from smx.demag import assemble_dense_B
    dev = torch.device("cuda:0")
    geo = _geo(mesh)
    B, ld = assemble_dense_B(geo, dev, dtype)
    n = geo.Nb
    X = torch.from_numpy(np.random.default_rng(seed).standard_normal((n, k))).to(dev)
    return B, ld, X, n


def _rel(a, b):
    return float(torch.linalg.norm(a - b) / torch.linalg.norm(b))


def _round_bf16(x):
    """Round float64 to the nearest bfloat16 (8 significant bits, ties to even), once."""
    m, e = np.frexp(np.asarray(x, dtype=np.float64))
    return np.ldexp(np.rint(m * 256.0), e - 8)


def _stored(X, dtype):
    """X as the kernel stages it: one rounding to the MFMA input precision."""
    if dtype == torch.float64:
        return X
    if dtype == torch.float32:
        return X.float().double()
    return torch.from_numpy(_round_bf16(X.cpu().numpy())).to(X.device)


def _acc_err(Y, Bs, Xs):
    """Accumulation error against the stored operands, relative to |B| |X|
    (the scale of the running sums; the Lindholm rows cancel strongly)."""
    return float((Y - Bs @ Xs).abs().max() / (Bs.abs() @ Xs.abs()).max())


# accumulation over a K slice of ~600 terms: K u ~ 4e-5 (FP32), ~7e-14 (FP64)
@pytest.mark.parametrize("dtype,tol", [(torch.float64, 2e-14), (torch.float32, 2e-5), (torch.bfloat16, 2e-5)])
def test_bmatmul_matches_stored_operands_c1(dtype, tol):
    from smx import mesh as M
    from smx.demag import dense_bmatmul
    B0, ld0, X, n = _case(M.cube_c1(), dtype, 70)     # 70 columns: a full and a partial launch
    # leading dimension 8 (mod 16): a BF16 lane chunk of 16 straddles the row end
    ld = (ld0 + 15) // 16 * 16 + 8
    B = torch.zeros((n, ld), dtype=dtype, device=X.device)
    B[:, :ld0] = B0
    Y = dense_bmatmul(B, ld, X)
    # the zero padding contributes exact zeros: same bits as the unpadded matrix
    assert torch.equal(dense_bmatmul(B0, ld0, X), Y)
    Bs = B[:, :n].double()
    assert _acc_err(Y, Bs, _stored(X, dtype)) < tol
    # deterministic: a second call is bitwise identical
    assert torch.equal(dense_bmatmul(B, ld, X), Y)


def test_bmatmul_precision_study_c2():
    """c2 (Nb = 5402), 64 right-hand sides: BF16/FP32-MFMA against the FP64 product."""
    from smx import mesh as M
    from smx.demag import assemble_dense_B, dense_bmatmul
    dev = torch.device("cuda:0")
    mesh = M.cube_c2()
    geo = _geo(mesh)
    n = geo.Nb
    X = torch.from_numpy(np.random.default_rng(3).standard_normal((n, 64))).to(dev)
    B64, ld = assemble_dense_B(geo, dev, torch.float64)
    Y64 = dense_bmatmul(B64, ld, X)
    ref = B64[:, :n] @ X
    assert _rel(Y64, ref) < 1e-13
    errs = {}
    for dt in (torch.float32, torch.bfloat16):
        Bl = B64.to(dt)
        Y = dense_bmatmul(Bl, ld, X)
        errs[str(dt)] = _rel(Y, ref)
        assert _acc_err(Y, Bl[:, :n].double(), _stored(X, dt)) < 2e-5
    print("c2 rel. error vs FP64 product:", errs)
    assert errs["torch.float32"] < 1e-6
    assert errs["torch.bfloat16"] < 1e-2


def test_bf16_storage_is_one_rounding():
    """The dense B stored in BF16 is the single round-to-nearest-even of FP64."""
    from smx import mesh as M
    from smx.demag import assemble_dense_B
    dev = torch.device("cuda:0")
    geo = _geo(M.cube_c1())
    B64, ld = assemble_dense_B(geo, dev, torch.float64)
    B16, _ = assemble_dense_B(geo, dev, torch.bfloat16)
    assert np.array_equal(B16.double().cpu().numpy(), _round_bf16(B64.cpu().numpy()))


def test_bmatmul_single_column_equals_matvec():
    """nrhs = 1 through the matrix cores agrees with the GEMV kernel."""
    from smx import _capi
    from smx import mesh as M
    from smx.demag import dense_bmatmul
    B, ld, X, n = _case(M.cube_c1(), torch.float64, 1, seed=5)
    y = torch.empty(n, dtype=torch.float64, device=X.device)
    L = _capi.lib()
    _capi.check(L.smx_dense_bmatvec(B.data_ptr(), _capi.SMX_F64, n, ld, X.data_ptr(), None, None, 0.0, None,
                                    y.data_ptr(), _capi.stream_ptr()))
    Y = dense_bmatmul(B, ld, X)
    assert _rel(Y[:, 0], y) < 1e-14
