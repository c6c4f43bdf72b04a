"""ctypes binding of libsmx_hip.so (declarations in include/smx_hip.h).

The library is loaded from this package directory only.  If it is missing the
import fails loudly: there is no CPU fallback for any operator of the path.
Status codes map onto the exception types the reference raises for the same
conditions (ValueError for shapes, TypeError for dtypes, RuntimeError for
device failures, MemoryError for workspaces).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import torch

LIB_PATH = Path(__file__).resolve().parent / "libsmx_hip.so"

SMX_OK, SMX_ERR_SHAPE, SMX_ERR_VALUE, SMX_ERR_NOT_CONVERGED, SMX_ERR_OOM, SMX_ERR_HIP, SMX_ERR_TYPE = range(7)
SMX_F64, SMX_F32, SMX_BF16 = 0, 1, 2
TRI_FIELDS = 34
LINDHOLM_TILE_COLS = 1024
PCG_DONE, PCG_ITERS, PCG_FAIL, PCG_SUMX, PCG_ITERS0 = 0, 1, 5, 6, 7
HMAT_Z_GROUP = 8
HMAT_NEAR_W_MAX = 8192
HMAT_KW_MAX = 16384
SVD_KMAX = 128
ACA_TILE_ROWS = 64

DTYPE_CODE = {torch.float64: SMX_F64, torch.float32: SMX_F32, torch.bfloat16: SMX_BF16}


class NotConvergedError(RuntimeError):
    """CG stopped at max_it or broke down (SMX_ERR_NOT_CONVERGED)."""


_EXC = {
    SMX_ERR_SHAPE: ValueError,
    SMX_ERR_VALUE: ValueError,
    SMX_ERR_NOT_CONVERGED: NotConvergedError,
    SMX_ERR_OOM: MemoryError,
    SMX_ERR_HIP: RuntimeError,
    SMX_ERR_TYPE: TypeError,
}

_lib = None

P = C.c_void_p
I32 = C.c_int32
I64 = C.c_int64
F64 = C.c_double


class FieldArgs(C.Structure):
    _fields_ = [
        ("n", I64),
        ("slice_ptr", P), ("cols", P),
        ("ex_vals", P), ("dmi_vals", P), ("grad_vals", P),
        ("u1", P), ("u2", P), ("u1_sum", P), ("inv_vol", P),
        ("m", P),
        ("ani_axis", P), ("ani_axis_u", F64 * 3), ("ani_pref", F64),
        ("h0", P), ("h0_u", F64 * 3),
        ("h_eff", P), ("h_demag", P), ("rhs", P),
        ("P", F64), ("alpha", F64), ("do_precess", F64), ("stab", F64),
        ("cub_pref", F64), ("cub_axes", P), ("cub_axes_u", F64 * 9),
        ("stt_vals", P), ("stt_pref", F64), ("stt_a", F64), ("stt_b", F64),
        ("sot_p", P), ("sot_p_u", F64 * 3), ("sot_fl", F64), ("sot_dl", F64),
        ("dmi_ncomp", I32), ("dmi_off", I32 * 9),
    ]


class CGMixed(C.Structure):
    """smx_cg_mixed: FP32 inner CG with FP64 iterative refinement."""
    _fields_ = [("vals32", P), ("rtol_inner", F64), ("refinements", I32), ("pad", I32)]


class NearTilesArgs(C.Structure):
    """smx_hmat_near_tiles."""
    _fields_ = [("ntile", I64), ("tile", P), ("row", P), ("col", P), ("val", P), ("iperm", P),
                ("val_dtype", I32), ("pad", I32)]


class HMatrixArgs(C.Structure):
    _fields_ = [
        ("nb", I64), ("gidx", P), ("perm", P), ("diag_p", P),
        ("near_ptr", P), ("near_idx", P), ("near_val", P), ("near_dtype", I32), ("factor_dtype", I32),
        ("nztask", I64), ("ztask", P), ("V", P),
        ("nleaf", I64), ("leaf_start", P), ("leaf_stop", P), ("leaf_u_off", P), ("leaf_kw", P),
        ("leaf_zg_off", P), ("leaf_zg", P), ("kw_max", I32), ("leaf_waves", I32),
        ("U", P), ("xp", P), ("z", P), ("ynear", P),
        ("near_task", P), ("n_near_task", I64), ("near_stage", P), ("near_lidx", P),
        ("near_w_max", I32), ("near_waves", I32), ("leaf_zr_off", P), ("leaf_zr", P),
    ]


byref = C.byref

_SIGS = {
    "smx_version": ([], C.c_int),
    "smx_last_error": ([], C.c_char_p),
    "smx_lindholm_fill": ([P, P, I64, P, I64, P, P, P, I64, I64, P, I64, I64, I32, P], C.c_int),
    "smx_lindholm_fill_batched": ([P, P, I64, P, P, P, P, P, P, P, I64, P, I32, P], C.c_int),
    "smx_dense_finish": ([P, I64, I64, P, P, I32, P], C.c_int),
    "smx_aca_batched": ([P, P, I64, F64, P, P, P, P, P, P], C.c_int),
    "smx_aca_large_ws_doubles": ([I64, I64], I64),
    "smx_aca_large": ([P, P, I64, P, P, P, I64, I32, F64, P, P, P, P, P, P, I64, P], C.c_int),
    "smx_svd_ws_doubles": ([I32, I32, I32], I64),
    "smx_svd_batched": ([P, P, I64, F64, P, P, P, P, P, P, P], C.c_int),
    "smx_rsvd_round": ([P, P, I64, P, F64, I32, P, P, P, P, P, P, P], C.c_int),
    "smx_hmat_matvec": ([C.POINTER(HMatrixArgs), P, P, F64, P, P], C.c_int),
    "smx_hmat_matmat_ws_size": ([C.POINTER(HMatrixArgs), I64], I64),
    "smx_hmat_matvec_noalloc": ([C.POINTER(HMatrixArgs), I64, P, P, F64, P, P], C.c_int),
    "smx_hmat_matmat": ([C.POINTER(HMatrixArgs), C.c_void_p, I64, P, I64, I32, P, I64, P, I64, P], C.c_int),
    "smx_dense_bmatvec": ([P, I32, I64, I64, P, P, P, F64, P, P, P], C.c_int),
    "smx_dense_bmatvec_t": ([P, I32, I64, I64, P, P, P, P], C.c_int),
    "smx_dense_bmatmul_ws_size": ([I64, I32], I64),
    "smx_dense_bmatmul": ([P, I32, I64, I64, P, I64, I32, P, I64, P, I64, P], C.c_int),
    "smx_sell_spmv": ([I64, P, P, P, I32, P, P, P, P], C.c_int),
    "smx_sell_spmv_blocks": ([I64], I64),
    "smx_reduce_ws_size": ([I64], I64),
    "smx_pcg_ws_size": ([I64], I64),
    "smx_pcg_prepare": ([I64, I32, I64, P, P, P, P, P, P], I32),
    "smx_pcg_jacobi": ([I64, I64, I32, I64, P, P, P, I32, P, P, I64, I32, I32, I64, P, P, P, P, P, I32, P, F64,
                        F64, I32, I32, I32, P, C.POINTER(I32), C.POINTER(F64), C.POINTER(F64), P],
                       C.c_int),
    "smx_pcg_jacobi_mixed": ([I64, I64, I32, I64, P, P, P, I32, P, P, I64, I32, I32, I64, P, P, P, P, P, I32, P,
                              F64, F64, I32, I32, I32, P, C.POINTER(I32), C.POINTER(F64), C.POINTER(F64),
                              C.POINTER(CGMixed), P], C.c_int),
    "smx_pcg_amg_ws_size": ([I64], I64),
    "smx_amg_apply": ([P, I32, I32, P, P, P, P], C.c_int),
    "smx_pcg_amg": ([I64, I64, I32, I64, P, P, I32, P, P, I64, I32, I32, I64, P, P, P, P, P, I32, P, P, I32, I32,
                     F64, F64, I32, I32, I32, P, C.POINTER(I32), C.POINTER(F64), C.POINTER(F64), P], C.c_int),
    "smx_pcg_amg_mixed": ([I64, I64, I32, I64, P, P, I32, P, P, I64, I32, I32, I64, P, P, P, P, P, I32, P, P, I32,
                           I32, F64, F64, I32, I32, I32, P, C.POINTER(I32), C.POINTER(F64), C.POINTER(F64),
                           C.POINTER(CGMixed), P], C.c_int),
    "smx_field_llg_rhs": ([C.POINTER(FieldArgs), P], C.c_int),
    "smx_llg_jvp": ([C.POINTER(FieldArgs), P, P, P, F64, P], C.c_int),
    "smx_dot3_weighted": ([I64, P, P, P, P, P, P], C.c_int),
    "smx_projected_direction": ([I64, P, P, P, P], C.c_int),
    "smx_cayley_step": ([I64, P, P, F64, P, P], C.c_int),
    "smx_max_norm3": ([I64, P, P, P, P], C.c_int),
    "smx_normalize3": ([I64, P, P], C.c_int),
    "smx_llg_pc_build": ([C.POINTER(FieldArgs), P, F64, F64, F64, I32, I32, I32, P, P], C.c_int),
    "smx_block3_apply": ([I64, P, P, P, P], C.c_int),
    "smx_multi_dot_ws_size": ([I64, I32], I64),
    "smx_multi_dot": ([I64, I32, P, I64, P, P, P, P], C.c_int),
    "smx_lincomb": ([I64, I32, P, I64, P, F64, P, P], C.c_int),
    "smx_rk_combine": ([I64, P, F64, I32, P, P, P, F64, F64, P, P, P, P], C.c_int),
    "smx_energy_blocks": ([I64], I64),
    "smx_energy_cell_forward": ([I64, P, P, P, P, P, P, P, P, P, P], C.c_int),
    "smx_energy_node_div": ([I64, P, P, P, P, P, P, P], C.c_int),
    "smx_energy_cell_backward": ([I64, P, P, P, P, P, P, P, P], C.c_int),
    "smx_energy_node_grad": ([I64, P, P, P, P, P, P, P, P, F64, P, P, P], C.c_int),
    "smx_vec_glue": ([I32, I64, P, P, P, F64, F64, P, P, F64, P, P], C.c_int),
    "smx_dot_partials": ([I64, P, P, P, P], C.c_int),
    "smx_node_rhs": ([I64, P, P, P, F64, F64, F64, P, P], C.c_int),
    "smx_node_rhs_vjp": ([I64, P, P, P, P, P, P, F64, P, P, F64, F64, F64, P, P, P, P, P, P], C.c_int),
    "smx_node_to_cell": ([I64, P, P, P, P, I32, P, P], C.c_int),
    "smx_cell_to_node": ([I64, P, P, P, P, P, P, P], C.c_int),
    "smx_row_scale": ([I64, P, P, P, P, P], C.c_int),
}


def exported_symbols():
    return list(_SIGS)


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(
                f"{LIB_PATH} is missing: build the HIP extension first "
                "(__graft_entry__.build() or python -m smx.build)")
        L = C.CDLL(str(LIB_PATH))
        for name, (args, res) in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def check(rc: int, what: str = ""):
    if rc != SMX_OK:
        msg = lib().smx_last_error().decode(errors="replace")
        raise _EXC.get(rc, RuntimeError)(f"{what}: {msg}" if what else msg)


def ptr(t) -> int | None:
    """Device pointer of a tensor (None for None)."""
    if t is None:
        return None
    return t.data_ptr()


def stream_ptr(stream=None) -> int:
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


def require_device(t: torch.Tensor, name: str, dtype=torch.float64, numel=None):
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise ValueError(f"{name} must be a device tensor")
    if dtype is not None and t.dtype != dtype:
        raise TypeError(f"{name} must be {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if numel is not None and t.numel() != numel:
        raise ValueError(f"{name} must have {numel} entries, got {t.numel()}")
    return t
