"""smx_hmat_matvec_noalloc (the H-matrix matvec whose scratch is the
library's fixed device array) against smx_hmat_matvec (caller workspaces),
both called through the C ABI.

Two calls in a row on one GPU, the second with a smaller operator: the
150 nm cube (Nb = 5402, leaf 64), then the 40 nm cube (Nb = 602, leaf 32),
FP32 and FP64 factors, with the near rows in the leaf launch and in a launch
of their own.  The two entry points run the same kernels on the same data,
so the results must be bitwise equal; the scratch keeps the first call's
contents, which the second must not read.  A size beyond the fixed array is
rejected before any launch.
"""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")


def _operator(mesh, leaf):
    from smx.boundary import build_boundary_geometry
    from smx.hmatrix import HMatrixBuildConfig, build_hmatrix_device
    geo = build_boundary_geometry(mesh.X, mesh.tets)
    return geo, build_hmatrix_device(geo, HMatrixBuildConfig(epsilon=1e-6, leaf_size=leaf), DEV)


@pytest.fixture(scope="module")
def operators():
    from smx import mesh as M
    return [_operator(M.cube_c2(), 64), _operator(M.cube_c1(), 32)]


@pytest.mark.parametrize("fp32,near_in_z", [(True, False), (False, False), (True, True)])
def test_noalloc_matches_workspace_form(operators, fp32, near_in_z):
    from smx import _capi
    from smx.hmatrix import HMatrixGPUPackedFused
    L = _capi.lib()
    for k, (geo, data) in enumerate(operators):
        H = HMatrixGPUPackedFused(data, use_fp32=fp32, device=DEV, near_in_z=near_in_z)
        assert (H._H.ynear is not None) == near_in_z
        x = torch.from_numpy(np.random.default_rng(100 + k).standard_normal(geo.Nb)).to(DEV)
        y_ref = torch.full_like(x, np.nan)
        y_na = torch.full_like(x, np.nan)
        print(f"[noalloc] call {k}: Nb={geo.Nb} nz={H._nz} fp32={fp32} near_in_z={near_in_z}", flush=True)
        _capi.check(L.smx_hmat_matvec(C.byref(H._H), x.data_ptr(), None, 0.0, y_ref.data_ptr(),
                                      _capi.stream_ptr()), "smx_hmat_matvec")
        _capi.check(L.smx_hmat_matvec_noalloc(C.byref(H._H), H._nz, x.data_ptr(), None, 0.0, y_na.data_ptr(),
                                              _capi.stream_ptr()), "smx_hmat_matvec_noalloc")
        torch.cuda.synchronize()
        assert torch.isfinite(y_ref).all()
        assert torch.equal(y_na, y_ref), k
        # the Python form of the same entry point
        y_py = torch.full_like(x, np.nan)
        H.matvec_noalloc(x, y_py)
        torch.cuda.synchronize()
        assert torch.equal(y_py, y_ref), k


def test_noalloc_rejects_oversized_scratch(operators):
    from smx import _capi
    from smx.hmatrix import HMatrixGPUPackedFused
    geo, data = operators[1]
    H = HMatrixGPUPackedFused(data, use_fp32=True, device=DEV, near_in_z=False)
    x = torch.zeros(geo.Nb, dtype=torch.float64, device=DEV)
    y = torch.zeros_like(x)
    rc = _capi.lib().smx_hmat_matvec_noalloc(C.byref(H._H), (1 << 20), x.data_ptr(), None, 0.0, y.data_ptr(),
                                             _capi.stream_ptr())
    assert rc == _capi.SMX_ERR_SHAPE
    assert b"scratch" in _capi.lib().smx_last_error()
