"""Device-resident operator containers (HBM layout of the path).

Everything the per-evaluation kernels read lives in device tensors allocated
once at set-up; PyTorch only provides the allocations and the stream.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import _capi
from .fem import SellMatrix


def default_device():
    if not torch.cuda.is_available():
        raise RuntimeError("smx needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def to_dev(a, device, dtype=None) -> torch.Tensor:
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.to(device).contiguous()


@dataclass
class DeviceSell:
    nrows: int
    k: int
    slice_ptr: torch.Tensor
    cols: torch.Tensor
    vals: torch.Tensor
    nnz: int

    @classmethod
    def upload(cls, m: SellMatrix, device, dtype=torch.float64):
        return cls(nrows=m.nrows, k=m.k,
                   slice_ptr=to_dev(m.slice_ptr.astype(np.int64), device),
                   cols=to_dev(m.cols.astype(np.int32), device),
                   vals=to_dev(m.vals.reshape(-1), device, dtype),
                   nnz=m.nnz)

    def with_vals(self, vals: np.ndarray, k: int, dtype=torch.float64, sell_positions=None):
        """Same pattern, other values (host values already in SELL order)."""
        return DeviceSell(nrows=self.nrows, k=k, slice_ptr=self.slice_ptr, cols=self.cols,
                          vals=to_dev(np.asarray(vals).reshape(-1), self.cols.device, dtype),
                          nnz=self.nnz)

    @property
    def padded_entries(self) -> int:
        return int(self.cols.numel())

    def nbytes(self) -> int:
        return int(self.slice_ptr.numel() * 8 + self.cols.numel() * 4 + self.vals.numel() * self.vals.element_size())

    def spmv(self, x: torch.Tensor, y: torch.Tensor, kind: int, partials=None, stream=None):
        """y = A x; with `partials`, block sums of y land in partials[0:blocks()]."""
        L = _capi.lib()
        _capi.check(L.smx_sell_spmv(self.nrows, _capi.ptr(self.slice_ptr), _capi.ptr(self.cols),
                                    _capi.ptr(self.vals), kind, _capi.ptr(x), _capi.ptr(y),
                                    _capi.ptr(partials), _capi.stream_ptr(stream)),
                    "smx_sell_spmv")

    def blocks(self) -> int:
        return int(_capi.lib().smx_sell_spmv_blocks(self.nrows))


@dataclass
class DeviceEll:
    """Fixed-width ELL matrix: entry j of row i at cols[j*ld + i]."""
    nrows: int
    width: int
    ld: int
    cols: torch.Tensor
    vals: torch.Tensor
    nnz: int

    @classmethod
    def from_csr(cls, indptr: np.ndarray, indices: np.ndarray, vals: np.ndarray, nrows: int, device,
                 dtype=torch.float64):
        lens = np.diff(indptr).astype(np.int64)
        W = max(int(lens.max()) if lens.size else 1, 1)
        ld = (nrows + 63) // 64 * 64
        cols = np.empty((W, ld), dtype=np.int32)
        cols[:] = np.minimum(np.arange(ld), max(nrows - 1, 0))[None, :]
        v = np.zeros((W, ld))
        r = np.repeat(np.arange(nrows, dtype=np.int64), lens)
        j = np.arange(indices.size, dtype=np.int64) - np.repeat(indptr[:-1], lens)
        cols[j, r] = indices
        v[j, r] = vals
        return cls(nrows=nrows, width=W, ld=ld, cols=to_dev(cols.ravel(), device),
                   vals=to_dev(v.ravel(), device, dtype), nnz=int(indices.size))

    def nbytes(self) -> int:
        return int(self.cols.numel() * 4 + self.vals.numel() * self.vals.element_size())


def sell_positions(indptr: np.ndarray, nrows: int):
    """SELL slot of every CSR entry (same rule as fem.to_sell)."""
    lens = np.diff(indptr).astype(np.int64)
    nslice = (nrows + 63) // 64
    lens_p = np.zeros(nslice * 64, dtype=np.int64)
    lens_p[:nrows] = lens
    widths = lens_p.reshape(nslice, 64).max(axis=1)
    slice_ptr = np.zeros(nslice + 1, dtype=np.int64)
    slice_ptr[1:] = np.cumsum(widths * 64)
    r = np.repeat(np.arange(nrows, dtype=np.int64), lens)
    j = np.arange(int(indptr[-1]), dtype=np.int64) - np.repeat(indptr[:-1], lens)
    s = r // 64
    return slice_ptr[s] + j * 64 + (r - 64 * s), int(slice_ptr[-1])


def sell_values(indptr: np.ndarray, vals: np.ndarray, nrows: int) -> np.ndarray:
    """Scatter CSR-ordered values (nnz, *k) into SELL order (entries, k)."""
    vals = np.asarray(vals, dtype=np.float64)
    k = 1 if vals.ndim == 1 else int(np.prod(vals.shape[1:]))
    pos, total = sell_positions(indptr, nrows)
    out = np.zeros((total, k))
    out[pos] = vals.reshape(-1, k)
    return out


def reduce_workspace(n: int, device) -> torch.Tensor:
    L = _capi.lib()
    return torch.zeros(int(L.smx_reduce_ws_size(int(n))), dtype=torch.float64, device=device)
