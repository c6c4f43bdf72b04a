"""Packed H-matrix on the device and its matvec (HMatrixGPUPackedFused).

Same constructor / ``matvec(x, out=None)`` / ``benchmark`` / ``memory_report``
surface as lindholm_hmatrix_gpu/backend_fused.py:359-646.  Device layout:

* far factors of every block packed into two flat arrays (FP32 by default):
  U column-major (m x r: one contiguous m-vector per rank, so a wavefront of
  64 output rows reads 64 consecutive entries), V row-major (r x n);
* near field as CSR in permuted coordinates (FP64 values, int32 columns, the
  exact zeros dropped as storage.py:453-459 does);
* a leaf table: the coarsest row intervals on which every far block is either
  fully present or absent (cut to <= 64 rows), each with its list of
  (far block, row offset) pairs, so one workgroup produces each output row
  without atomics.
"""
from __future__ import annotations

import time

import numpy as np
import torch

from .. import _capi
from ..device import to_dev
from .storage import HMatrixCPUData

_U_F = {True: torch.float32, False: torch.float64}


def _leaf_table(nb, starts_b, stops_b, dev):
    """Coarsest intervals on which every block's output range is all-or-nothing,
    cut to <= 64 rows, with the (block, offset) pairs of each interval."""
    bounds = np.unique(np.concatenate([[0, nb], starts_b, stops_b]))
    starts, stops = [], []
    for a, b in zip(bounds[:-1], bounds[1:]):
        for s in range(int(a), int(b), 64):
            starts.append(s)
            stops.append(min(s + 64, int(b)))
    starts = np.array(starts, dtype=np.int64)
    stops = np.array(stops, dtype=np.int64)
    nleaf = starts.size
    nfar = np.asarray(starts_b).size
    if nfar:
        first = np.searchsorted(starts, starts_b)
        last = np.searchsorted(starts, stops_b)
        cnt = last - first
        blk = np.repeat(np.arange(nfar, dtype=np.int64), cnt)
        leaf = np.repeat(first, cnt) + (np.arange(cnt.sum()) - np.repeat(np.cumsum(cnt) - cnt, cnt))
        order = np.lexsort((blk, leaf))
        blk, leaf = blk[order], leaf[order]
        fe_row = starts[leaf] - np.asarray(starts_b)[blk]
        fe_ptr = np.zeros(nleaf + 1, dtype=np.int64)
        np.add.at(fe_ptr, leaf + 1, 1)
        fe_ptr = np.cumsum(fe_ptr)
    else:
        blk = np.zeros(0, np.int64)
        fe_row = np.zeros(0, np.int64)
        fe_ptr = np.zeros(nleaf + 1, dtype=np.int64)
    i32 = lambda a: to_dev(np.asarray(a, dtype=np.int32), dev)
    i64 = lambda a: to_dev(np.asarray(a, dtype=np.int64), dev)
    return dict(start=i32(starts), stop=i32(stops), fe_ptr=i64(fe_ptr), fe_block=i32(blk), fe_row=i32(fe_row))


class HMatrixGPUPackedFused:
    def __init__(self, data: HMatrixCPUData, threads_per_block: int = 128, use_fp32: bool = True,
                 device=None, near_fp32: bool = False, boundary_gather: np.ndarray | None = None):
        self.data = data
        self.device = device or torch.device("cuda", torch.cuda.current_device())
        dev = self.device
        self.use_fp32 = bool(use_fp32)
        self.threads_per_block = int(threads_per_block)
        nb = data.size
        self.nb = nb
        perm = np.asarray(data.perm, dtype=np.int64)
        self.perm = to_dev(perm.astype(np.int32), dev)
        self.gidx_plain = self.perm
        self.gidx_boundary = (to_dev(np.asarray(boundary_gather, dtype=np.int64)[perm].astype(np.int32), dev)
                              if boundary_gather is not None else None)
        self.diag_p = to_dev(np.asarray(data.diag_jump, dtype=np.float64)[perm], dev)
        near = data.near_csr.tocsr()
        self.near_nnz = int(near.nnz)
        self.near_ptr = to_dev(near.indptr.astype(np.int64), dev)
        self.near_idx = to_dev(near.indices.astype(np.int32), dev)
        nd = torch.float32 if near_fp32 else torch.float64
        self.near_val = to_dev(near.data, dev, nd)
        self.near_dtype = _capi.DTYPE_CODE[nd]
        fb = data.far_blocks
        nfar = len(fb)
        self.nfar = nfar
        r0 = np.array([b.row_start for b in fb], dtype=np.int64)
        r1 = np.array([b.row_stop for b in fb], dtype=np.int64)
        c0 = np.array([b.col_start for b in fb], dtype=np.int64)
        c1 = np.array([b.col_stop for b in fb], dtype=np.int64)
        rk = np.array([b.rank for b in fb], dtype=np.int64)
        m = r1 - r0
        n = c1 - c0
        u_off = np.concatenate([[0], np.cumsum(m * rk)[:-1]]).astype(np.int64) if nfar else np.zeros(0, np.int64)
        v_off = np.concatenate([[0], np.cumsum(n * rk)[:-1]]).astype(np.int64) if nfar else np.zeros(0, np.int64)
        z_off = np.concatenate([[0], np.cumsum(rk)[:-1]]).astype(np.int64) if nfar else np.zeros(0, np.int64)
        fdt = np.float32 if self.use_fp32 else np.float64
        Uflat = (np.concatenate([np.ascontiguousarray(b.U.T, dtype=fdt).ravel() for b in fb])
                 if nfar else np.zeros(1, fdt))
        Vflat = (np.concatenate([np.ascontiguousarray(b.V, dtype=fdt).ravel() for b in fb])
                 if nfar else np.zeros(1, fdt))
        self.far_storage_entries = int((m * rk).sum() + (n * rk).sum())
        self.U = to_dev(Uflat, dev)
        self.V = to_dev(Vflat, dev)
        self.factor_dtype = _capi.SMX_F32 if self.use_fp32 else _capi.SMX_F64
        i32 = lambda a: to_dev(np.asarray(a, dtype=np.int32), dev)
        i64 = lambda a: to_dev(np.asarray(a, dtype=np.int64), dev)
        self.far = dict(c0=i32(c0), n=i32(n), m=i32(m), rank=i32(rk), u_off=i64(u_off), v_off=i64(v_off),
                        z_off=i64(z_off))
        self._r = (r0, r1, c0, c1, rk, u_off, v_off, z_off)
        self.leaf = _leaf_table(nb, r0, r1, dev)
        self.nleaf = int(self.leaf["start"].numel())
        self.xp = torch.zeros(nb, dtype=torch.float64, device=dev)
        self.z = torch.zeros(max(int(rk.sum()), 1), dtype=torch.float64, device=dev)
        self._H = self._struct(self.gidx_plain)
        self._Hb = self._struct(self.gidx_boundary) if self.gidx_boundary is not None else None

    def _struct(self, gidx, transpose=False):
        """Kernel arguments; with transpose=True the same kernels apply B^T:
        z_b = U_b^T x[I_b] (U's rank vectors are contiguous, as V's are), the
        outputs J_b take V_b^T z_b, the near field is the transposed CSR and
        the leaf table runs over column intervals."""
        if transpose:
            return self._struct_t(gidx)
        H = _capi.HMatrixArgs()
        p = _capi.ptr
        H.nb = self.nb
        H.gidx = p(gidx)
        H.perm = p(self.perm)
        H.diag_p = p(self.diag_p)
        H.near_ptr = p(self.near_ptr) if self.near_nnz else None
        H.near_idx = p(self.near_idx) if self.near_nnz else None
        H.near_val = p(self.near_val) if self.near_nnz else None
        H.near_dtype = self.near_dtype
        H.factor_dtype = self.factor_dtype
        H.nfar = self.nfar
        f = self.far
        H.far_c0, H.far_n, H.far_m, H.far_rank = p(f["c0"]), p(f["n"]), p(f["m"]), p(f["rank"])
        H.far_u_off, H.far_v_off, H.far_z_off = p(f["u_off"]), p(f["v_off"]), p(f["z_off"])
        H.U, H.V = p(self.U), p(self.V)
        lf = self.leaf
        H.nleaf = self.nleaf
        H.leaf_start, H.leaf_stop = p(lf["start"]), p(lf["stop"])
        H.leaf_fe_ptr, H.leaf_fe_block, H.leaf_fe_row = p(lf["fe_ptr"]), p(lf["fe_block"]), p(lf["fe_row"])
        H.xp, H.z = p(self.xp), p(self.z)
        return H

    def _struct_t(self, gidx):
        if not hasattr(self, "_tt"):
            dev = self.device
            r0, r1, c0, c1, rk, u_off, v_off, z_off = self._r
            nt = self.data.near_csr.T.tocsr()
            nd = self.near_val.dtype
            i32 = lambda a: to_dev(np.asarray(a, dtype=np.int32), dev)
            i64 = lambda a: to_dev(np.asarray(a, dtype=np.int64), dev)
            self._tt = dict(near_ptr=to_dev(nt.indptr.astype(np.int64), dev),
                            near_idx=to_dev(nt.indices.astype(np.int32), dev),
                            near_val=to_dev(nt.data, dev, nd),
                            far=dict(c0=i32(r0), n=i32(r1 - r0), m=i32(c1 - c0), rank=i32(rk), u_off=i64(v_off),
                                     v_off=i64(u_off), z_off=i64(z_off)),
                            leaf=_leaf_table(self.nb, c0, c1, dev))
        tt = self._tt
        H = _capi.HMatrixArgs()
        p = _capi.ptr
        H.nb = self.nb
        H.gidx = p(gidx)
        H.perm = p(self.perm)
        H.diag_p = p(self.diag_p)
        H.near_ptr = p(tt["near_ptr"]) if self.near_nnz else None
        H.near_idx = p(tt["near_idx"]) if self.near_nnz else None
        H.near_val = p(tt["near_val"]) if self.near_nnz else None
        H.near_dtype = self.near_dtype
        H.factor_dtype = self.factor_dtype
        H.nfar = self.nfar
        f = tt["far"]
        H.far_c0, H.far_n, H.far_m, H.far_rank = p(f["c0"]), p(f["n"]), p(f["m"]), p(f["rank"])
        H.far_u_off, H.far_v_off, H.far_z_off = p(f["u_off"]), p(f["v_off"]), p(f["z_off"])
        H.U, H.V = p(self.V), p(self.U)
        lf = tt["leaf"]
        H.nleaf = int(lf["start"].numel())
        H.leaf_start, H.leaf_stop = p(lf["start"]), p(lf["stop"])
        H.leaf_fe_ptr, H.leaf_fe_block, H.leaf_fe_row = p(lf["fe_ptr"]), p(lf["fe_block"]), p(lf["fe_row"])
        H.xp, H.z = p(self.xp), p(self.z)
        return H

    # ------------------------------------------------------------------
    def rmatvec(self, x_gpu: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """y = B^T x (original boundary ordering): the adjoint boundary operator."""
        if x_gpu.shape != (self.nb,):
            raise ValueError(f"Expected x.shape == {(self.nb,)}, got {tuple(x_gpu.shape)}.")
        _capi.require_device(x_gpu, "x")
        if out is None:
            out = torch.empty_like(x_gpu)
        if not hasattr(self, "_HT"):
            self._HT = self._struct(self.gidx_plain, transpose=True)
        _capi.check(_capi.lib().smx_hmat_matvec(_capi.byref(self._HT), x_gpu.data_ptr(), None, 0.0,
                                                out.data_ptr(), _capi.stream_ptr()), "smx_hmat_matvec")
        return out

    def matvec(self, x_gpu: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """y = B x for x in the original boundary ordering (backend_fused.py:496-512)."""
        if x_gpu.shape != (self.nb,):
            raise ValueError(f"Expected x.shape == {(self.nb,)}, got {tuple(x_gpu.shape)}.")
        _capi.require_device(x_gpu, "x")
        if out is None:
            out = torch.empty_like(x_gpu)
        _capi.check(_capi.lib().smx_hmat_matvec(_capi.byref(self._H), x_gpu.data_ptr(), None, 0.0,
                                                out.data_ptr(), _capi.stream_ptr()), "smx_hmat_matvec")
        return out

    def matvec_numpy(self, x: np.ndarray) -> np.ndarray:
        xt = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64)).to(self.device)
        return self.matvec(xt).cpu().numpy()

    def matvec_from(self, src: torch.Tensor, y: torch.Tensor, gather=None, shift_ptr=None, shift_scale=0.0):
        """y = B x with x[k] = src[gather[k]] - shift (the demag boundary gather)."""
        H = self._Hb if gather is not None else self._H
        if gather is not None and H is None:
            raise ValueError("boundary_gather was not provided at construction")
        _capi.check(_capi.lib().smx_hmat_matvec(_capi.byref(H), src.data_ptr(), shift_ptr, float(shift_scale),
                                                y.data_ptr(), _capi.stream_ptr()), "smx_hmat_matvec")

    def algorithmic_bytes(self) -> int:
        """Bytes one matvec must move: factors, near CSR (values + columns +
        row pointers), diagonal, x gather / y scatter."""
        fb = 4 if self.use_fp32 else 8
        nv = self.near_val.element_size()
        return int(self.far_storage_entries * fb + self.near_nnz * (nv + 4) + (self.nb + 1) * 8
                   + self.nb * 8 + 4 * self.nb * 8)

    def benchmark(self, x=None, repeats=20, warmup=3):
        if x is None:
            x = torch.randn(self.nb, dtype=torch.float64, device=self.device)
        out = torch.empty_like(x)
        for _ in range(warmup):
            self.matvec(x, out)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(repeats):
            self.matvec(x, out)
        e1.record()
        e1.synchronize()
        total = e0.elapsed_time(e1) * 1e-3
        return {"timer": "hip_event", "repeats": repeats, "total_seconds": total,
                "average_seconds": total / repeats}

    def memory_report(self) -> dict:
        arrays = [self.U, self.V, self.near_ptr, self.near_idx, self.near_val, self.perm, self.diag_p,
                  self.xp, self.z, *self.far.values(), *self.leaf.values()]
        nbytes = int(sum(a.numel() * a.element_size() for a in arrays))
        return {"backend": "hmatrix", "factor_dtype": "float32" if self.use_fp32 else "float64",
                "far_blocks": self.nfar, "near_nnz": self.near_nnz,
                "far_storage_entries": self.far_storage_entries,
                "accounted_gpu_bytes": nbytes, "accounted_gpu_mib": nbytes / 1024 ** 2,
                "accounted_gpu_human": f"{nbytes / 1024 ** 2:.1f} MiB"}

    def print_memory_report(self, prefix="[HMatrixGPU]"):
        r = self.memory_report()
        print(f"{prefix} far={r['far_blocks']} near_nnz={r['near_nnz']} factors={r['factor_dtype']} "
              f"GPU={r['accounted_gpu_human']}", flush=True)
