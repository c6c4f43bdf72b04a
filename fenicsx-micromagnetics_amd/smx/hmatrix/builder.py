"""H-matrix construction on the device.

Partition rule (lindholm_hmatrix_gpu/builder.py:256-376, 630-760): for a
block (r, c)
  * admissible and its workspace does not fit (dense_bytes >
    max_temporary_block_bytes, or dense_bytes x the compressor's workspace
    multiplier above gpu_memory_fraction of the device memory available) ->
    subdivide; an unsplittable block is kept dense (gpu_fallback_to_cpu) or
    raises MemoryError;
  * admissible -> fill, compress (full-pivot ACA, truncated SVD or adaptive
    randomized SVD); accept when rank > 0, rel. residual <= eps (1 + 64 u)
    and (m+n) rank < 0.95 m n; else subdivide when the tolerance failed and a
    side can split; else keep it dense;
  * inadmissible leaf x leaf -> dense; otherwise subdivide the splittable side
    with the larger diameter (rows on ties).
A batch that runs out of device memory is halved and retried; a single block
that still does is subdivided (counted as an OOM subdivision) or, when it
cannot split, kept dense as above.  The reference fills such a dense block on
the CPU; here it is filled on the device in row slabs small enough to fit, so
no block entry is ever computed on the host.
Each block's fate depends only on the block itself, so the recursion is run
breadth first: every level's candidate blocks are filled by ONE batched
Lindholm launch and compressed by ONE batched ACA launch (a workgroup per
block), with a single host synchronization per batch.  The resulting set of
near/far blocks equals the depth-first recursion's; far blocks are emitted in
that recursion's order (a replay of the recorded decisions).

ACA stops at the first rank the storage rule can no longer accept (when no
max_rank is given): full-pivot ACA always reaches the tolerance eventually,
so such a block is kept dense exactly as the reference would keep it.
"""
from __future__ import annotations

import math
import time
from dataclasses import asdict, dataclass
from hashlib import sha256

import numpy as np
import scipy.sparse as sp
import torch

from .. import _capi
from ..boundary import BoundaryGeometry
from ..device import to_dev
from .storage import HMatrixCPUData, LowRankBlockCPU
from .tree import ClusterNode, admissible, build_cluster_tree

HMATRIX_CONSTRUCTION_VERSION = "smx-lindholm-hmatrix-v1"

FILL_TASK_DTYPE = np.dtype([("out_off", np.int64), ("ld_row", np.int64), ("ld_col", np.int64),
                            ("rows_off", np.int64), ("rel_off", np.int64), ("nrows", np.int32),
                            ("nrel", np.int32), ("ncols", np.int32), ("pad", np.int32)])
ACA_TASK_DTYPE = np.dtype([("a_off", np.int64), ("u_off", np.int64), ("v_off", np.int64),
                           ("m", np.int32), ("n", np.int32), ("rcap", np.int32), ("pad", np.int32)])
assert FILL_TASK_DTYPE.itemsize == 56 and ACA_TASK_DTYPE.itemsize == 40
# blocks of at least this many entries (16 MiB) are compressed with every ACA
# rank step spread over 64-row tiles (smx_aca_large)
ACA_LARGE_ENTRIES = 1 << 21
ACA_TILE_ROWS = _capi.ACA_TILE_ROWS


def aca_tile_table(m: np.ndarray):
    """Row tiles of smx_aca_large for blocks of m[b] rows: (tile_task,
    tile_row0, tile_first), block b owning tiles tile_first[b] ..
    tile_first[b+1]-1 in row order, ACA_TILE_ROWS rows each (the last one
    shorter)."""
    m = np.asarray(m, dtype=np.int64)
    if np.any(m <= 0):
        raise ValueError("every block needs at least one row")
    ntile = (m + ACA_TILE_ROWS - 1) // ACA_TILE_ROWS
    first = np.concatenate([[0], np.cumsum(ntile)]).astype(np.int32)
    tile_task = np.repeat(np.arange(m.size, dtype=np.int32), ntile)
    tile_row0 = (ACA_TILE_ROWS * (np.arange(int(ntile.sum())) - np.repeat(first[:-1], ntile))).astype(np.int32)
    return tile_task, tile_row0, first


@dataclass
class HMatrixBuildConfig:
    epsilon: float = 1e-6
    eta: float = 2.0
    leaf_size: int = 64
    compressor: str = "fullaca"
    max_rank: int | None = None
    max_temporary_block_bytes: int = 256 * 1024 * 1024
    low_rank_storage_factor: float = 0.95
    build_backend: str = "gpu"
    gpu_device_id: int = 0
    gpu_memory_fraction: float = 0.75
    gpu_fallback_to_cpu: bool = True
    gpu_aca_residual_check_interval: int = 1
    gpu_rsvd_initial_rank: int = 16
    gpu_rsvd_oversampling: int = 8
    gpu_rsvd_power_iterations: int = 1
    gpu_progress_interval_seconds: float = 5.0
    batch_bytes: int = 4 * 1024 ** 3

    # (field, predicate, what the value must satisfy); checked in this order
    _CHECKS = (
        ("epsilon", lambda v: v > 0.0, "> 0"),
        ("eta", lambda v: v > 0.0, "> 0"),
        ("leaf_size", lambda v: v >= 1, ">= 1"),
        ("max_rank", lambda v: v is None or int(v) >= 0, "None or >= 0"),
        ("max_temporary_block_bytes", lambda v: v >= 8, ">= 8 (one FP64 entry)"),
        ("low_rank_storage_factor", lambda v: 0.0 < v <= 1.0, "in (0, 1]"),
        ("gpu_device_id", lambda v: int(v) >= 0, ">= 0"),
        ("gpu_memory_fraction", lambda v: 0.0 < float(v) <= 1.0, "in (0, 1]"),
        ("gpu_aca_residual_check_interval", lambda v: v >= 1, ">= 1"),
        ("gpu_rsvd_initial_rank", lambda v: v >= 1, ">= 1"),
        ("gpu_rsvd_oversampling", lambda v: v >= 0, ">= 0"),
        ("gpu_rsvd_power_iterations", lambda v: v >= 0, ">= 0"),
        ("gpu_progress_interval_seconds", lambda v: v >= 0.0, ">= 0"),
    )
    _COMPRESSOR_ALIASES = {"fullaca": "fullaca", "aca": "fullaca", "svd": "svd", "rsvd": "rsvd",
                           "randomized_svd": "rsvd", "randomized-svd": "rsvd"}

    def __post_init__(self):
        for name, ok, rule in self._CHECKS:
            value = getattr(self, name)
            if not ok(value):
                raise ValueError(f"HMatrixBuildConfig.{name} = {value!r} is invalid: it must be {rule}")
        comp = self._COMPRESSOR_ALIASES.get(str(self.compressor).strip().lower())
        if comp is None:
            raise ValueError(f"HMatrixBuildConfig.compressor = {self.compressor!r} is not one of "
                             f"{sorted(set(self._COMPRESSOR_ALIASES))}")
        self.compressor = comp
        self.gpu_device_id = int(self.gpu_device_id)
        self.build_backend = str(self.build_backend).strip().lower()
        if self.build_backend not in {"gpu", "auto"}:
            raise ValueError(f"HMatrixBuildConfig.build_backend = {self.build_backend!r}: this library builds "
                             "the H-matrix on the device only ('gpu' or 'auto')")


def geometry_hash(Xb, tri_lid=None) -> str:
    d = sha256()
    d.update(np.ascontiguousarray(Xb, dtype=np.float64).view(np.uint8))
    if tri_lid is not None:
        d.update(np.ascontiguousarray(tri_lid, dtype=np.int32).view(np.uint8))
    return d.hexdigest()


class DeviceBuildOps:
    """Device side of the builder: batched Lindholm fill and batched ACA.

    The builder only talks to this interface (alloc / fill / aca / host /
    gather), so its partition logic can also be driven by a host stand-in."""

    def __init__(self, geo: BoundaryGeometry, device):
        self.geo = geo
        self.device = device
        self.Xb = to_dev(geo.Xb, device)
        self.geom = to_dev(geo.tri_table(), device)
        self.lid = to_dev(geo.tri_lid.astype(np.int32), device)
        self.inc = geo.incidence()
        self.evaluations = 0

    def relevant(self, cols: np.ndarray):
        """Relevant triangles of a column tile, cached per build by the tile's
        columns (a cluster's tiles recur as the columns of many blocks)."""
        cache = self.__dict__.setdefault("_rel_cache", {})
        key = (cols.size, hash(cols.tobytes()))
        hit = cache.get(key)
        if hit is not None and np.array_equal(hit[0], cols):
            return hit[1], hit[2]
        scratch = self.__dict__.setdefault("_col_of", np.full(self.geo.Nb, -1, dtype=np.int64))
        rel, rc = self.geo.relevant_triangles(cols, self.inc, scratch=scratch)
        cache[key] = (cols.copy(), rel, rc)
        return rel, rc

    def alloc(self, n: int):
        return torch.empty(max(int(n), 1), dtype=torch.float64, device=self.device)

    def available_bytes(self) -> int:
        """Free device memory plus what torch's caching allocator holds
        unused (builder.py:520-548 counts the pool's reusable bytes alike)."""
        free, _ = torch.cuda.mem_get_info(self.device)
        pool = torch.cuda.memory_reserved(self.device) - torch.cuda.memory_allocated(self.device)
        return int(free) + max(int(pool), 0)

    def release(self):
        torch.cuda.empty_cache()

    def host(self, t) -> np.ndarray:
        return t.cpu().numpy()

    def gather(self, t, ranges) -> np.ndarray:
        if not ranges:
            return np.zeros(0)
        return torch.cat([t[a:b] for a, b in ranges]).cpu().numpy()

    def fill_tasks(self, blocks, offsets):
        """Task table of a batch: one task per (block, column tile)."""
        T = _capi.LINDHOLM_TILE_COLS
        tasks = []
        rows_parts, rel_parts, rc_parts = [], [], []
        nrows_tot = nrel_tot = 0
        for b, (rows, cols) in enumerate(blocks):
            m, n = rows.size, cols.size
            rows_parts.append(rows)
            for t0 in range(0, n, T):
                tc = cols[t0:t0 + T]
                rel, rc = DeviceBuildOps.relevant(self, tc)
                tasks.append((int(offsets[b]) + t0, n, 1, nrows_tot, nrel_tot, m, rel.size, tc.size, 0))
                rel_parts.append(rel)
                rc_parts.append(rc.reshape(-1))
                nrel_tot += rel.size
                self.evaluations += m * rel.size
            nrows_tot += m
        task_arr = np.array(tasks, dtype=FILL_TASK_DTYPE)
        nwg = (task_arr["nrows"].astype(np.int64) + 3) // 4
        wg_task = np.repeat(np.arange(task_arr.size, dtype=np.int32), nwg)
        starts = np.cumsum(nwg) - nwg
        wg_row0 = (4 * (np.arange(int(nwg.sum()), dtype=np.int64) - np.repeat(starts, nwg))).astype(np.int32)
        rows = np.concatenate(rows_parts).astype(np.int32)
        rel = np.concatenate(rel_parts).astype(np.int32) if rel_parts else np.zeros(1, np.int32)
        rc = np.concatenate(rc_parts).astype(np.int32) if rc_parts else np.zeros(3, np.int32)
        return task_arr, wg_task, wg_row0, rows, rel, rc

    def fill(self, blocks, out, offsets):
        """blocks: list of (rows, cols) global boundary indices; block b is
        written row-major at out[offsets[b] : offsets[b] + m*n]."""
        task_arr, wg_task, wg_row0, rows, rel, rc = self.fill_tasks(blocks, offsets)
        dev = self.device
        keep = [to_dev(a, dev) for a in (rows, rel, rc, task_arr.view(np.uint8), wg_task, wg_row0)]
        L = _capi.lib()
        _capi.check(L.smx_lindholm_fill_batched(
            _capi.ptr(self.Xb), _capi.ptr(self.geom), self.geo.Nt, _capi.ptr(self.lid),
            *[_capi.ptr(k) for k in keep], int(wg_task.size), out.data_ptr(),
            int(self.geo.incident_policy), _capi.stream_ptr()), "smx_lindholm_fill_batched")
        torch.cuda.current_stream().synchronize()

    def aca(self, A, tasks: np.ndarray, U, V, rtol: float):
        """Full-pivot ACA of every task's block: blocks of at least
        ACA_LARGE_ENTRIES entries through smx_aca_large (each rank step
        spread over 64-row tiles, one workgroup each), the others through
        smx_aca_batched (one workgroup per block)."""
        dev = self.device
        nt = int(tasks.size)
        L = _capi.lib()
        ranks, rels, status = np.zeros(nt, np.int32), np.zeros(nt), np.zeros(nt, np.int32)
        mn = tasks["m"].astype(np.int64) * tasks["n"].astype(np.int64)
        for ids, large in ((np.flatnonzero(mn < ACA_LARGE_ENTRIES), False),
                           (np.flatnonzero(mn >= ACA_LARGE_ENTRIES), True)):
            if ids.size == 0:
                continue
            sub = np.ascontiguousarray(tasks[ids])
            nb = int(sub.size)
            tasks_d = to_dev(sub.view(np.uint8), dev)
            rank_d = torch.empty(nb, dtype=torch.int32, device=dev)
            res_d = torch.empty(nb, dtype=torch.float64, device=dev)
            st_d = torch.empty(nb, dtype=torch.int32, device=dev)
            if not large:
                _capi.check(L.smx_aca_batched(A.data_ptr(), tasks_d.data_ptr(), nb, float(rtol), U.data_ptr(),
                                              V.data_ptr(), rank_d.data_ptr(), res_d.data_ptr(), st_d.data_ptr(),
                                              _capi.stream_ptr()), "smx_aca_batched")
            else:
                tile_task, tile_row0, first = aca_tile_table(sub["m"])
                nws = int(L.smx_aca_large_ws_doubles(nb, int(tile_task.size)))
                work = torch.empty(nws, dtype=torch.float64, device=dev)
                keep = [to_dev(a, dev) for a in (tile_task, tile_row0, first)]
                _capi.check(L.smx_aca_large(A.data_ptr(), tasks_d.data_ptr(), nb, keep[0].data_ptr(),
                                            keep[1].data_ptr(), keep[2].data_ptr(), int(tile_task.size),
                                            int(sub["rcap"].max()), float(rtol), U.data_ptr(), V.data_ptr(),
                                            rank_d.data_ptr(), res_d.data_ptr(), st_d.data_ptr(), work.data_ptr(),
                                            nws, _capi.stream_ptr()), "smx_aca_large")
            ranks[ids], rels[ids], status[ids] = rank_d.cpu().numpy(), res_d.cpu().numpy(), st_d.cpu().numpy()
        return ranks, rels, status

    def compress(self, method, A, tasks, U, V, rtol, max_rank, seeds, cfg):
        if method == "fullaca":
            return self.aca(A, tasks, U, V, rtol)
        from .compressors import rsvd_batched, svd_batched
        if method == "svd":
            return svd_batched(A, tasks, U, V, rtol, max_rank)
        return rsvd_batched(A, tasks, U, V, rtol, seeds, max_rank, cfg.gpu_rsvd_initial_rank,
                            cfg.gpu_rsvd_oversampling, cfg.gpu_rsvd_power_iterations)


def is_device_oom(exc: BaseException) -> bool:
    """Allocation failures of the device (torch's OutOfMemoryError, the C
    ABI's SMX_ERR_OOM as MemoryError, or an allocator message), through the
    exception chain (builder.py:591-628)."""
    seen = set()
    cur = exc
    while cur is not None and id(cur) not in seen:
        seen.add(id(cur))
        if isinstance(cur, (MemoryError, torch.cuda.OutOfMemoryError)):
            return True
        msg = str(cur).lower()
        if any(t in msg for t in ("out of memory", "outofmemory", "memory allocation", "alloc failed")):
            return True
        cur = cur.__cause__ or cur.__context__
    return False


def _can_split(node: ClusterNode) -> bool:
    return not node.is_leaf


def _children(r: ClusterNode, c: ClusterNode):
    rs, cs = _can_split(r), _can_split(c)
    if rs and (not cs or r.diameter >= c.diameter):
        return "row", [(r.left, c), (r.right, c)]
    return "col", [(r, c.left), (r, c.right)]


def _max_ok_rank(m: int, n: int, factor: float) -> int:
    """Largest r with (m+n) r < factor m n."""
    lim = factor * m * n
    r = int(math.floor(lim / (m + n)))
    while r > 0 and (m + n) * r >= lim:
        r -= 1
    return max(r, 0)


def build_hmatrix_device(geo: BoundaryGeometry, config: HMatrixBuildConfig | None = None, device=None,
                         verbose: bool = False, ops=None) -> HMatrixCPUData:
    from collections import deque

    from .compressors import block_random_seed, workspace_multiplier
    cfg = config or HMatrixBuildConfig()
    t_start = time.perf_counter()
    tree = build_cluster_tree(geo.Xb, leaf_size=cfg.leaf_size)
    perm = tree.perm
    if ops is None:
        ops = DeviceBuildOps(geo, device or torch.device("cuda", torch.cuda.current_device()))
    margin = 1.0 + 64.0 * np.finfo(np.float64).eps
    mult = workspace_multiplier(cfg.compressor)
    stats = dict(dense_blocks=0, low_rank_blocks=0, rejected_low_rank_blocks=0,
                 forced_subdivisions_memory=0, compression_tolerance_subdivisions=0,
                 block_entry_evaluations=0, max_observed_rank=0, sum_observed_rank=0,
                 aca_batches=0, fill_batches=0, gpu_workspace_subdivisions=0, gpu_oom_subdivisions=0,
                 fallback_dense_blocks=0, oom_batch_splits=0, slab_filled_blocks=0, cpu_fallback_blocks=0)
    decision: dict = {}
    far_of: dict = {}
    near_rows, near_cols, near_vals = [], [], []
    last_report = [time.perf_counter()]

    def key(r, c):
        return (id(r), id(c))

    def idx(node):
        return perm[node.start:node.stop].astype(np.int64)

    def release():
        rel = getattr(ops, "release", None)
        if rel is not None:
            rel()

    def keep_nonzeros(host, offs, chunk, row_shift=None):
        nz = np.flatnonzero(host)
        if not nz.size:
            return
        b = np.searchsorted(offs, nz, side="right") - 1
        local = nz - offs[b]
        ncol = np.array([c.size for _, c in chunk], dtype=np.int64)[b]
        rs = np.array([r.start for r, _ in chunk], dtype=np.int64)[b]
        if row_shift is not None:
            rs = rs + row_shift
        cs = np.array([c.start for _, c in chunk], dtype=np.int64)[b]
        near_rows.append(rs + local // ncol)
        near_cols.append(cs + local % ncol)
        near_vals.append(host[nz])

    def fill_slabs(r, c):
        """One dense block filled in row slabs, halving the slab until it fits."""
        rows, cols = idx(r), idx(c)
        slab = max(1, r.size // 2)
        a = 0
        while a < r.size:
            b = min(a + slab, r.size)
            try:
                buf = ops.alloc((b - a) * c.size)
                ops.fill([(rows[a:b], cols)], buf, np.zeros(1, np.int64))
                host = ops.host(buf)[:(b - a) * c.size].copy()
            except Exception as exc:
                if not is_device_oom(exc) or slab == 1:
                    raise
                buf = None
                release()
                slab = max(1, slab // 2)
                continue
            del buf
            stats["fill_batches"] += 1
            keep_nonzeros(host, np.zeros(1, np.int64), [(r, c)], row_shift=a)
            a = b

    def add_near(pairs, budget):
        """Fill blocks densely and keep their nonzero entries (permuted coords)."""
        if not pairs:
            return
        work = deque(_batches(pairs, budget))
        while work:
            chunk = work.popleft()
            sizes = np.array([r.size * c.size for r, c in chunk], dtype=np.int64)
            offs = np.concatenate([[0], np.cumsum(sizes)[:-1]])
            try:
                buf = ops.alloc(int(sizes.sum()))
                ops.fill([(idx(r), idx(c)) for r, c in chunk], buf, offs)
                host = ops.host(buf)[:int(sizes.sum())].copy()
            except Exception as exc:
                if not is_device_oom(exc):
                    raise
                buf = None
                release()
                if len(chunk) > 1:
                    stats["oom_batch_splits"] += 1
                    h = len(chunk) // 2
                    work.appendleft(chunk[h:])
                    work.appendleft(chunk[:h])
                    continue
                if not cfg.gpu_fallback_to_cpu:
                    raise MemoryError(f"device OOM filling a dense {chunk[0][0].size}x{chunk[0][1].size} block") \
                        from exc
                fill_slabs(*chunk[0])
                stats["slab_filled_blocks"] += 1
                stats["block_entry_evaluations"] += int(sizes.sum())
                continue
            del buf
            stats["fill_batches"] += 1
            stats["block_entry_evaluations"] += int(sizes.sum())
            keep_nonzeros(host, offs, chunk)

    def compress_chunk(chunk):
        """Fill + compress one batch; returns per-block results and the host
        factors of the blocks that pass; commits nothing."""
        m_ = np.array([r.size for r, _ in chunk], dtype=np.int64)
        n_ = np.array([c.size for _, c in chunk], dtype=np.int64)
        sizes = m_ * n_
        offs = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
        rcap = np.empty(len(chunk), dtype=np.int64)
        storage_cap = np.zeros(len(chunk), dtype=bool)
        for i in range(len(chunk)):
            ok = _max_ok_rank(int(m_[i]), int(n_[i]), cfg.low_rank_storage_factor)
            full = int(min(m_[i], n_[i]))
            if cfg.max_rank is None:
                rcap[i] = min(ok + 1, full)
                storage_cap[i] = rcap[i] == ok + 1
            else:
                rcap[i] = min(max(int(cfg.max_rank), 0), full)
        uoff = np.concatenate([[0], np.cumsum(m_ * rcap)[:-1]]).astype(np.int64)
        voff = np.concatenate([[0], np.cumsum(n_ * rcap)[:-1]]).astype(np.int64)
        A = ops.alloc(int(sizes.sum()))
        U = ops.alloc(int((m_ * rcap).sum()))
        V = ops.alloc(int((n_ * rcap).sum()))
        ops.fill([(idx(r), idx(c)) for r, c in chunk], A, offs)
        tasks = np.zeros(len(chunk), dtype=ACA_TASK_DTYPE)
        tasks["a_off"], tasks["u_off"], tasks["v_off"] = offs, uoff, voff
        tasks["m"], tasks["n"], tasks["rcap"] = m_, n_, rcap
        seeds = np.array([block_random_seed(r.start, r.stop, c.start, c.stop) for r, c in chunk], dtype=np.int64)
        ranks, rels, status = ops.compress(cfg.compressor, A, tasks, U, V, float(cfg.epsilon), cfg.max_rank,
                                           seeds, cfg)
        fate = []
        for i, (r, c) in enumerate(chunk):
            rank, rel = int(ranks[i]), float(rels[i])
            if cfg.compressor == "fullaca" and status[i] == 2 and storage_cap[i]:
                fate.append("reject")
                continue
            meets = np.isfinite(rel) and rel <= cfg.epsilon * margin
            if rank > 0 and meets and (m_[i] + n_[i]) * rank < cfg.low_rank_storage_factor * sizes[i]:
                fate.append("far")
            elif not meets and (_can_split(r) or _can_split(c)):
                fate.append("split")
            elif status[i] == 3 and rcap[i] > 0:
                # the block is exactly zero: dense, with no entries to store
                fate.append("zero")
            else:
                # includes a rank cap of 0 (max_rank=0) on a nonzero block: the
                # compressor returns rank 0, so the block is kept dense with its
                # entries, as builder.py:343-371 keeps it as a dense rejection
                fate.append("reject")
        acc = [i for i, f in enumerate(fate) if f == "far"]
        Uh = Vh = None
        if acc:
            Uh = ops.gather(U, [(int(uoff[i]), int(uoff[i]) + int(m_[i]) * int(ranks[i])) for i in acc])
            Vh = ops.gather(V, [(int(voff[i]), int(voff[i]) + int(n_[i]) * int(ranks[i])) for i in acc])
        del A, U, V
        return fate, ranks, rels, m_, n_, Uh, Vh, int(sizes.sum())

    queue = [(tree.root, tree.root)]
    while queue:
        avail = float(ops.available_bytes()) * cfg.gpu_memory_fraction if hasattr(ops, "available_bytes") \
            else math.inf
        nxt, cand, dense_pairs, rejected = [], [], [], []

        def split(r, c):
            which, ch = _children(r, c)
            decision[key(r, c)] = ("split", which)
            nxt.extend(ch)

        def split_or_fallback(r, c, stat, sink):
            if _can_split(r) or _can_split(c):
                stats[stat] += 1
                split(r, c)
                return
            if not cfg.gpu_fallback_to_cpu:
                reason = "device OOM" if stat == "gpu_oom_subdivisions" else "device workspace limit"
                raise MemoryError(f"{reason} for an unsplittable H-matrix block {r.size}x{c.size}.")
            decision[key(r, c)] = ("dense",)
            sink.append((r, c))
            stats["dense_blocks"] += 1
            stats["fallback_dense_blocks"] += 1

        for r, c in queue:
            adm = admissible(r, c, cfg.eta)
            nbytes = r.size * c.size * 8
            if adm and (nbytes > cfg.max_temporary_block_bytes or nbytes * mult > avail):
                stats["forced_subdivisions_memory"] += 1
                split_or_fallback(r, c, "gpu_workspace_subdivisions", dense_pairs)
            elif adm:
                cand.append((r, c))
            elif r.is_leaf and c.is_leaf:
                decision[key(r, c)] = ("dense",)
                dense_pairs.append((r, c))
                stats["dense_blocks"] += 1
            else:
                split(r, c)
        fill_budget = int(max(min(cfg.batch_bytes, avail), 8))
        add_near(dense_pairs, fill_budget)
        work = deque(_batches(cand, int(max(min(cfg.batch_bytes // 2, avail / mult), 8))))
        while work:
            chunk = work.popleft()
            try:
                fate, ranks, rels, m_, n_, Uh, Vh, nent = compress_chunk(chunk)
            except Exception as exc:
                if not is_device_oom(exc):
                    raise
                release()
                if len(chunk) > 1:
                    stats["oom_batch_splits"] += 1
                    h = len(chunk) // 2
                    work.appendleft(chunk[h:])
                    work.appendleft(chunk[:h])
                else:
                    split_or_fallback(chunk[0][0], chunk[0][1], "gpu_oom_subdivisions", rejected)
                continue
            stats["fill_batches"] += 1
            stats["aca_batches"] += 1
            stats["block_entry_evaluations"] += nent
            pu = pv = 0
            for i, (r, c) in enumerate(chunk):
                f = fate[i]
                if f == "far":
                    decision[key(r, c)] = ("far",)
                    mm, nn, rk = int(m_[i]), int(n_[i]), int(ranks[i])
                    Ub = np.ascontiguousarray(Uh[pu:pu + mm * rk].reshape(rk, mm).T)
                    Vb = Vh[pv:pv + nn * rk].reshape(rk, nn).copy()
                    pu += mm * rk
                    pv += nn * rk
                    far_of[key(r, c)] = LowRankBlockCPU(r.start, r.stop, c.start, c.stop, Ub, Vb, float(rels[i]))
                    stats["low_rank_blocks"] += 1
                    stats["max_observed_rank"] = max(stats["max_observed_rank"], rk)
                    stats["sum_observed_rank"] += rk
                elif f == "split":
                    stats["compression_tolerance_subdivisions"] += 1
                    split(r, c)
                else:
                    decision[key(r, c)] = ("dense",)
                    stats["dense_blocks"] += 1
                    stats["rejected_low_rank_blocks"] += 1
                    if f == "reject":
                        rejected.append((r, c))
            if verbose and cfg.gpu_progress_interval_seconds > 0 and \
                    time.perf_counter() - last_report[0] >= cfg.gpu_progress_interval_seconds:
                last_report[0] = time.perf_counter()
                print(f"[HMatrix GPU build] far={stats['low_rank_blocks']}, near={stats['dense_blocks']}, "
                      f"entries={stats['block_entry_evaluations']:,}", flush=True)
        add_near(rejected, fill_budget)
        queue = nxt

    # far blocks in the order of the depth-first recursion; the block tree
    # (every visited block and its fate) is kept for inspection
    far_blocks = []
    visited = []
    fate_code = {"far": 0, "dense": 1}
    stack = [(tree.root, tree.root)]
    while stack:
        r, c = stack.pop()
        d = decision[key(r, c)]
        code = fate_code.get(d[0], 2 if d[0] == "split" and d[1] == "row" else 3)
        visited.append((r.start, r.stop, c.start, c.stop, code))
        if d[0] == "far":
            far_blocks.append(far_of[key(r, c)])
        elif d[0] == "split":
            ch = [(r.left, c), (r.right, c)] if d[1] == "row" else [(r, c.left), (r, c.right)]
            stack.append(ch[1])
            stack.append(ch[0])
    Nb = geo.Nb
    if near_rows:
        near = sp.coo_matrix((np.concatenate(near_vals), (np.concatenate(near_rows), np.concatenate(near_cols))),
                             shape=(Nb, Nb)).tocsr()
        near.sum_duplicates()
    else:
        near = sp.csr_matrix((Nb, Nb), dtype=np.float64)
    elapsed = time.perf_counter() - t_start
    nlr = stats["low_rank_blocks"]
    meta = dict(Nb=Nb, geometry_hash=geometry_hash(geo.Xb, geo.tri_lid), build_config=asdict(cfg),
                build_time_seconds=elapsed, average_far_rank=(stats["sum_observed_rank"] / nlr if nlr else 0.0),
                resolved_build_backend="gpu", construction_version=HMATRIX_CONSTRUCTION_VERSION,
                lindholm_pair_evaluations=int(getattr(ops, "evaluations", 0)), **stats)
    data = HMatrixCPUData(perm=perm.astype(np.int32), diag_jump=geo.diag_jump.astype(np.float64),
                          near_csr=near, far_blocks=far_blocks, metadata=meta)
    # rows (row_start, row_stop, col_start, col_stop, fate) in depth-first
    # order; fate 0 far, 1 dense, 2 split by rows, 3 split by columns
    data.block_tree = np.asarray(visited, dtype=np.int64).reshape(-1, 5)
    data.metadata["near_nnz"] = int(near.nnz)
    data.metadata["far_storage_entries"] = int(data.far_storage_entries)
    data.metadata["stored_entries_total"] = int(data.stored_entries)
    data.metadata["compression_ratio_entries"] = float(data.compression_ratio_entries)
    if verbose:
        print(f"[hmatrix] Nb={Nb} far={nlr} dense={stats['dense_blocks']} near_nnz={near.nnz} "
              f"far_entries={data.far_storage_entries} avg_rank={meta['average_far_rank']:.1f} "
              f"build {elapsed:.2f}s", flush=True)
    return data


def _batches(pairs, budget_bytes):
    out, cur, cur_b = [], [], 0
    for r, c in pairs:
        b = r.size * c.size * 8
        if cur and cur_b + b > budget_bytes:
            out.append(cur)
            cur, cur_b = [], 0
        cur.append((r, c))
        cur_b += b
    if cur:
        out.append(cur)
    return out


def build_or_load_hmatrix(cache_path, geo, *args, config: HMatrixBuildConfig | None = None, device=None,
                          force_rebuild=False, verbose=None, cache_compressed=None, diag_jump=None, provider=None):
    """Reuse a cache when its geometry hash, build configuration and construction
    version match (builder.py:814-927); otherwise build on the device and save.

    Two call forms: (cache_path, geo, config=None, device=None, ...) with the
    boundary geometry, or the reference's (cache_path, Xb, diag_jump,
    provider, config=None, force_rebuild=False, verbose=True,
    cache_compressed=True) with a LindholmEntryProviderGPU (its geometry must
    hold the same Xb; diag_jump is stored as given).  verbose and
    cache_compressed default to False in the first form and to True (the
    reference's defaults) in the second."""
    from pathlib import Path
    if not isinstance(geo, BoundaryGeometry):
        # the reference's positional form
        Xb = np.asarray(geo, dtype=np.float64)
        pos = list(args)
        names = ["diag_jump", "provider", "config", "force_rebuild", "verbose", "cache_compressed"]
        vals = dict(diag_jump=diag_jump, provider=provider, config=config, force_rebuild=force_rebuild,
                    verbose=True if verbose is None else verbose,
                    cache_compressed=True if cache_compressed is None else cache_compressed)
        for k, v in zip(names, pos):
            vals[k] = v
        prov = vals["provider"]
        if prov is None or not isinstance(getattr(prov, "geo", None), BoundaryGeometry):
            raise TypeError("provider must be a LindholmEntryProviderGPU (from_opB of the boundary geometry)")
        if prov.geo.Xb.shape != Xb.shape or not np.array_equal(prov.geo.Xb, Xb):
            raise ValueError("Xb differs from the provider's boundary geometry")
        dj = vals["diag_jump"]
        if dj is None or np.asarray(dj).shape != (Xb.shape[0],):
            raise ValueError("diag_jump must hold one value per boundary node")
        data = build_or_load_hmatrix(cache_path, prov.geo, config=vals["config"],
                                     device=device or torch.device("cuda", prov.device_id),
                                     force_rebuild=vals["force_rebuild"], verbose=vals["verbose"],
                                     cache_compressed=vals["cache_compressed"])
        dj = np.asarray(dj, dtype=np.float64)
        data.metadata["provider_signature"] = prov.cache_signature
        if not np.array_equal(data.diag_jump, dj):
            data.diag_jump = dj.copy()
            if cache_path is not None:
                data.save(cache_path, compressed=vals["cache_compressed"])
        return data
    if args:
        # (cache_path, geo, config, device, force_rebuild, verbose, cache_compressed) positionally
        names = ["config", "device", "force_rebuild", "verbose", "cache_compressed"]
        kw = dict(zip(names, args))
        config = kw.get("config", config)
        device = kw.get("device", device)
        force_rebuild = kw.get("force_rebuild", force_rebuild)
        verbose = kw.get("verbose", verbose)
        cache_compressed = kw.get("cache_compressed", cache_compressed)
    verbose = bool(verbose) if verbose is not None else False
    cache_compressed = bool(cache_compressed) if cache_compressed is not None else False
    cfg = config or HMatrixBuildConfig()
    if cache_path is not None and not force_rebuild and Path(cache_path).exists():
        try:
            data = HMatrixCPUData.load(cache_path)
            m = data.metadata
            if (m.get("geometry_hash") == geometry_hash(geo.Xb, geo.tri_lid)
                    and m.get("construction_version") == HMATRIX_CONSTRUCTION_VERSION
                    and m.get("build_config") == _json_roundtrip(asdict(cfg))):
                data.metadata["cache_loaded"] = True
                return data
        except Exception:
            pass
    data = build_hmatrix_device(geo, cfg, device, verbose)
    if cache_path is not None:
        data.save(cache_path, compressed=cache_compressed)
    return data


def _json_roundtrip(d):
    import json
    return json.loads(json.dumps(d))
