"""Lindholm block provider with the reference's protocol.

Counterpart of LindholmEntryProviderGPU (lindholm_hmatrix_gpu/
entry_provider_gpu.py:352-777): the object the H-matrix builder asks for
blocks of the double-layer operator B.  Here every entry is computed by the
batched Lindholm fill kernel on the device (smx_lindholm_fill_batched):
fill_block_device returns the block as a device tensor, fill_block and
fill_block_cpu return it as a host array (no entry is computed on the host).  The
provider wraps the boundary geometry that the demag class keeps as ``opB``.
"""
from __future__ import annotations

import numpy as np
import torch

from ..boundary import BoundaryGeometry
from .builder import HMATRIX_CONSTRUCTION_VERSION, DeviceBuildOps, geometry_hash


class LindholmEntryProviderGPU:
    supports_gpu = True
    resolved_build_backend = "gpu"

    def __init__(self, geo: BoundaryGeometry, device_id: int = 0, threads_per_block: int = 128,
                 kernel_mode: str = "auto", fast_math: bool = False, index_cache_max_bytes: int = 128 * 1024 ** 2):
        if int(device_id) < 0:
            raise ValueError("device_id must be non-negative.")
        if not 32 <= int(threads_per_block) <= 1024:
            raise ValueError("threads_per_block must be between 32 and 1024.")
        if str(kernel_mode).strip().lower() not in {"auto", "row", "atomic"}:
            raise ValueError("kernel_mode must be 'auto', 'row', or 'atomic'.")
        if int(index_cache_max_bytes) < 0:
            raise ValueError("index_cache_max_bytes must be non-negative.")
        self.geo = geo
        self.device_id = int(device_id)
        # accepted for the reference's signature; the fill kernel's launch shape is fixed
        self.threads_per_block = int(threads_per_block)
        self.kernel_mode = str(kernel_mode).strip().lower()
        self.fast_math = bool(fast_math)
        self.index_cache_max_bytes = int(index_cache_max_bytes)
        self._ops = None
        self._blocks = 0
        self._entries = 0

    @classmethod
    def from_opB(cls, opB: BoundaryGeometry, **kwargs) -> "LindholmEntryProviderGPU":
        return cls(opB, **kwargs)

    # ---- protocol attributes ----
    @property
    def size(self) -> int:
        return self.geo.Nb

    @property
    def Xb(self) -> np.ndarray:
        return self.geo.Xb

    @property
    def tri_lid(self) -> np.ndarray:
        return self.geo.tri_lid

    @property
    def incident_policy(self) -> int:
        return int(self.geo.incident_policy)

    def gpu_available(self) -> bool:
        return torch.cuda.is_available()

    @property
    def cache_signature(self) -> str:
        return self.cache_signature_for_backend("gpu")

    def cache_signature_for_backend(self, backend: str) -> str:
        return (f"{HMATRIX_CONSTRUCTION_VERSION}|lindholm|incident_policy={self.incident_policy}|"
                f"geometry={geometry_hash(self.geo.Xb, self.geo.tri_lid)}|backend={backend}")

    # ---- device lifetime ----
    def prepare_device(self):
        if self._ops is None:
            self._ops = DeviceBuildOps(self.geo, torch.device("cuda", self.device_id))
        return self._ops

    def release_construction_resources(self):
        self._ops = None

    def stats(self) -> dict:
        return {"provider_blocks": self._blocks, "provider_entries": self._entries,
                "lindholm_pair_evaluations": int(self._ops.evaluations) if self._ops is not None else 0}

    # ---- blocks ----
    def fill_block_device(self, rows, cols, *, row_key=None, col_key=None) -> torch.Tensor:
        """B[rows, cols] (boundary indices) as an (m, n) device tensor."""
        rows = np.asarray(rows, dtype=np.int64)
        cols = np.asarray(cols, dtype=np.int64)
        if rows.ndim != 1 or cols.ndim != 1:
            raise ValueError("rows and cols must be 1-D index arrays")
        nb = self.size
        if rows.size and (rows.min() < 0 or rows.max() >= nb) or cols.size and (cols.min() < 0 or cols.max() >= nb):
            raise ValueError("block indices out of range")
        ops = self.prepare_device()
        m, n = rows.size, cols.size
        out = ops.alloc(m * n)
        if m and n:
            ops.fill([(rows, cols)], out, np.zeros(1, dtype=np.int64))
        self._blocks += 1
        self._entries += m * n
        return out[:m * n].view(m, n)

    def fill_block(self, rows, cols) -> np.ndarray:
        return self.fill_block_device(rows, cols).cpu().numpy()

    def fill_block_cpu(self, rows, cols) -> np.ndarray:
        """The reference's host fill: the same entries, computed on the device."""
        return self.fill_block(rows, cols)
