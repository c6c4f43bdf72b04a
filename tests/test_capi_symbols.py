"""The C-ABI library loads without a GPU and exports every function that
include/smx_hip.h declares (no compute calls here)."""
import ctypes
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    text = (ROOT / "include" / "smx_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(smx_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_path():
    names = declared_functions()
    for must in ("smx_lindholm_fill", "smx_lindholm_fill_batched", "smx_aca_batched", "smx_hmat_matvec",
                 "smx_dense_bmatvec", "smx_pcg_jacobi", "smx_field_llg_rhs", "smx_llg_jvp"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from smx import _capi
    lib = ctypes.CDLL(str(_capi.LIB_PATH))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding knows every declared entry point
    assert set(declared_functions()) <= set(_capi.exported_symbols())


def test_error_string_and_shape_check_without_gpu():
    from smx import _capi
    L = _capi.lib()
    assert L.smx_version() >= 100
    # argument validation runs on the host before any launch
    rc = L.smx_dense_bmatvec(None, 0, 10, 5, None, None, None, 0.0, None, None, None)
    assert rc == _capi.SMX_ERR_SHAPE
    assert b"bad shape" in L.smx_last_error()
    rc = L.smx_lindholm_fill(None, None, 4, None, 1, None, None, None, 1, 5000, None, 1, 1, 0, None)
    assert rc == _capi.SMX_ERR_SHAPE
