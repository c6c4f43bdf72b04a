"""Host H-matrix storage with the reference's cache format.

``HMatrixCPUData`` / ``LowRankBlockCPU`` keep the fields and the ``.npz`` layout
of lindholm_hmatrix_gpu/storage.py:10-275 (keys perm, diag_jump, near_data,
near_indices, near_indptr, near_shape, far_meta (structured), far_U_flat,
far_V_flat, metadata_json), so caches written by either side load in the other.
Matrix-vector products run on the device only (backend.HMatrixGPUPackedFused).
"""
from __future__ import annotations

import json
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np
import scipy.sparse as sp

FAR_META_DTYPE = np.dtype([
    ("row_start", np.int64), ("row_stop", np.int64), ("col_start", np.int64),
    ("col_stop", np.int64), ("rank", np.int64), ("u_offset", np.int64), ("u_size", np.int64),
    ("v_offset", np.int64), ("v_size", np.int64), ("relative_residual", np.float64),
])


@dataclass
class LowRankBlockCPU:
    row_start: int
    row_stop: int
    col_start: int
    col_stop: int
    U: np.ndarray   # (m, r)
    V: np.ndarray   # (r, n)
    relative_residual: float = 0.0

    @property
    def rank(self) -> int:
        return int(self.U.shape[1])

    @property
    def storage_entries(self) -> int:
        return int(self.U.size + self.V.size)


@dataclass
class HMatrixCPUData:
    perm: np.ndarray
    diag_jump: np.ndarray
    near_csr: sp.csr_matrix
    far_blocks: list
    metadata: dict = field(default_factory=dict)

    @property
    def size(self) -> int:
        return int(self.perm.size)

    @property
    def inverse_perm(self) -> np.ndarray:
        inv = np.empty_like(self.perm)
        inv[self.perm] = np.arange(self.perm.size, dtype=self.perm.dtype)
        return inv

    @property
    def dense_theoretical_entries(self) -> int:
        return int(self.size * self.size)

    @property
    def far_storage_entries(self) -> int:
        return int(sum(b.storage_entries for b in self.far_blocks))

    @property
    def stored_entries(self) -> int:
        return int(self.near_csr.nnz + self.far_storage_entries + self.diag_jump.size)

    @property
    def compression_ratio_entries(self) -> float:
        return self.dense_theoretical_entries / max(self.stored_entries, 1)

    def save(self, path, *, compressed: bool = True):
        path = Path(path)
        path.parent.mkdir(parents=True, exist_ok=True)
        meta = np.empty(len(self.far_blocks), dtype=FAR_META_DTYPE)
        uo = vo = 0
        for i, b in enumerate(self.far_blocks):
            us, vs = b.U.size, b.V.size
            meta[i] = (b.row_start, b.row_stop, b.col_start, b.col_stop, b.rank, uo, us, vo, vs,
                       float(b.relative_residual))
            uo += us
            vo += vs
        U = (np.concatenate([np.ascontiguousarray(b.U, dtype=np.float64).ravel() for b in self.far_blocks])
             if self.far_blocks else np.empty(0))
        V = (np.concatenate([np.ascontiguousarray(b.V, dtype=np.float64).ravel() for b in self.far_blocks])
             if self.far_blocks else np.empty(0))
        writer = np.savez_compressed if compressed else np.savez
        with path.open("wb") as f:
            writer(f, perm=np.asarray(self.perm, dtype=np.int32),
                   diag_jump=np.asarray(self.diag_jump, dtype=np.float64),
                   near_data=np.asarray(self.near_csr.data, dtype=np.float64),
                   near_indices=np.asarray(self.near_csr.indices, dtype=np.int32),
                   near_indptr=np.asarray(self.near_csr.indptr, dtype=np.int32),
                   near_shape=np.asarray(self.near_csr.shape, dtype=np.int64),
                   far_meta=meta, far_U_flat=U, far_V_flat=V,
                   metadata_json=np.asarray(json.dumps(self.metadata, default=_json_default)))

    @classmethod
    def load(cls, path):
        ar = np.load(Path(path), allow_pickle=False)
        shape = tuple(np.asarray(ar["near_shape"], dtype=np.int64).tolist())
        near = sp.csr_matrix((np.asarray(ar["near_data"], dtype=np.float64),
                              np.asarray(ar["near_indices"], dtype=np.int32),
                              np.asarray(ar["near_indptr"], dtype=np.int32)), shape=shape)
        Uf = np.asarray(ar["far_U_flat"], dtype=np.float64)
        Vf = np.asarray(ar["far_V_flat"], dtype=np.float64)
        blocks = []
        for mt in ar["far_meta"]:
            m = int(mt["row_stop"] - mt["row_start"])
            n = int(mt["col_stop"] - mt["col_start"])
            r = int(mt["rank"])
            u0, v0 = int(mt["u_offset"]), int(mt["v_offset"])
            blocks.append(LowRankBlockCPU(
                int(mt["row_start"]), int(mt["row_stop"]), int(mt["col_start"]), int(mt["col_stop"]),
                np.ascontiguousarray(Uf[u0:u0 + int(mt["u_size"])].reshape(m, r)),
                np.ascontiguousarray(Vf[v0:v0 + int(mt["v_size"])].reshape(r, n)),
                float(mt["relative_residual"])))
        raw = str(np.asarray(ar["metadata_json"]).item())
        return cls(perm=np.asarray(ar["perm"], dtype=np.int32),
                   diag_jump=np.asarray(ar["diag_jump"], dtype=np.float64),
                   near_csr=near, far_blocks=blocks, metadata=json.loads(raw) if raw else {})


def _json_default(o):
    if isinstance(o, (np.integer,)):
        return int(o)
    if isinstance(o, (np.floating,)):
        return float(o)
    raise TypeError(type(o))
