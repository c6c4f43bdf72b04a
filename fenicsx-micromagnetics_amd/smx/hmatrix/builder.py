"""H-matrix construction on the device.

Partition rule (lindholm_hmatrix_gpu/builder.py:256-376): for a block (r, c)
  * admissible and dense_bytes > max_temporary_block_bytes -> subdivide;
  * admissible -> fill, full-pivot ACA; accept when rank > 0, rel. residual
    <= eps (1 + 64 u) and (m+n) rank < 0.95 m n; else subdivide when the
    tolerance failed and a side can split; else keep it dense;
  * inadmissible leaf x leaf -> dense; otherwise subdivide the splittable side
    with the larger diameter (rows on ties).
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
    batch_bytes: int = 4 * 1024 ** 3

    def __post_init__(self):
        if self.epsilon <= 0.0:
            raise ValueError("epsilon must be positive.")
        if self.eta <= 0.0:
            raise ValueError("eta must be positive.")
        if self.leaf_size < 1:
            raise ValueError("leaf_size must be positive.")
        if self.max_temporary_block_bytes < 8:
            raise ValueError("max_temporary_block_bytes is too small.")
        if not (0.0 < self.low_rank_storage_factor <= 1.0):
            raise ValueError("low_rank_storage_factor must be in (0, 1].")
        comp = str(self.compressor).strip().lower()
        if comp == "aca":
            comp = "fullaca"
        if comp != "fullaca":
            raise NotImplementedError("only the full-pivot ACA compressor is built for the device path")
        self.compressor = comp
        self.build_backend = str(self.build_backend).strip().lower()
        if self.build_backend not in {"gpu", "auto"}:
            raise ValueError("build_backend must be 'gpu' (the device builds the H-matrix)")


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

    def alloc(self, n: int):
        return torch.empty(max(int(n), 1), dtype=torch.float64, device=self.device)

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
                rel, rc = self.geo.relevant_triangles(tc, self.inc)
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
        dev = self.device
        nt = int(tasks.size)
        tasks_d = to_dev(tasks.view(np.uint8), dev)
        rank_d = torch.empty(nt, dtype=torch.int32, device=dev)
        res_d = torch.empty(nt, dtype=torch.float64, device=dev)
        st_d = torch.empty(nt, dtype=torch.int32, device=dev)
        _capi.check(_capi.lib().smx_aca_batched(A.data_ptr(), tasks_d.data_ptr(), nt, float(rtol), U.data_ptr(),
                                                V.data_ptr(), rank_d.data_ptr(), res_d.data_ptr(),
                                                st_d.data_ptr(), _capi.stream_ptr()), "smx_aca_batched")
        return rank_d.cpu().numpy(), res_d.cpu().numpy(), st_d.cpu().numpy()


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
    cfg = config or HMatrixBuildConfig()
    t_start = time.perf_counter()
    tree = build_cluster_tree(geo.Xb, leaf_size=cfg.leaf_size)
    perm = tree.perm
    if ops is None:
        ops = DeviceBuildOps(geo, device or torch.device("cuda", torch.cuda.current_device()))
    margin = 1.0 + 64.0 * np.finfo(np.float64).eps
    stats = dict(dense_blocks=0, low_rank_blocks=0, rejected_low_rank_blocks=0,
                 forced_subdivisions_memory=0, compression_tolerance_subdivisions=0,
                 block_entry_evaluations=0, max_observed_rank=0, sum_observed_rank=0,
                 aca_batches=0, fill_batches=0)
    decision: dict = {}
    far_of: dict = {}
    near_rows, near_cols, near_vals = [], [], []

    def key(r, c):
        return (id(r), id(c))

    def idx(node):
        return perm[node.start:node.stop].astype(np.int64)

    def add_near(pairs):
        """Fill blocks densely and keep their nonzero entries (permuted coords)."""
        if not pairs:
            return
        for chunk in _batches(pairs, cfg.batch_bytes):
            sizes = np.array([r.size * c.size for r, c in chunk], dtype=np.int64)
            offs = np.concatenate([[0], np.cumsum(sizes)[:-1]])
            buf = ops.alloc(int(sizes.sum()))
            ops.fill([(idx(r), idx(c)) for r, c in chunk], buf, offs)
            stats["fill_batches"] += 1
            stats["block_entry_evaluations"] += int(sizes.sum())
            host = ops.host(buf)[:int(sizes.sum())]
            nz = np.flatnonzero(host)
            if nz.size:
                b = np.searchsorted(offs, nz, side="right") - 1
                local = nz - offs[b]
                ncol = np.array([c.size for _, c in chunk], dtype=np.int64)[b]
                rs = np.array([r.start for r, _ in chunk], dtype=np.int64)[b]
                cs = np.array([c.start for _, c in chunk], dtype=np.int64)[b]
                near_rows.append(rs + local // ncol)
                near_cols.append(cs + local % ncol)
                near_vals.append(host[nz])
            del buf

    queue = [(tree.root, tree.root)]
    while queue:
        nxt, cand, dense_pairs = [], [], []
        for r, c in queue:
            adm = admissible(r, c, cfg.eta)
            if adm and r.size * c.size * 8 > cfg.max_temporary_block_bytes:
                stats["forced_subdivisions_memory"] += 1
                if not _can_split(r) and not _can_split(c):
                    decision[key(r, c)] = ("dense",)
                    dense_pairs.append((r, c))
                    stats["dense_blocks"] += 1
                else:
                    which, ch = _children(r, c)
                    decision[key(r, c)] = ("split", which)
                    nxt += ch
            elif adm:
                cand.append((r, c))
            elif r.is_leaf and c.is_leaf:
                decision[key(r, c)] = ("dense",)
                dense_pairs.append((r, c))
                stats["dense_blocks"] += 1
            else:
                which, ch = _children(r, c)
                decision[key(r, c)] = ("split", which)
                nxt += ch
        add_near(dense_pairs)
        rejected = []
        for chunk in _batches(cand, cfg.batch_bytes // 2):
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
                    user = min(max(int(cfg.max_rank), 0), full)
                    rcap[i] = user
            uoff = np.concatenate([[0], np.cumsum(m_ * rcap)[:-1]]).astype(np.int64)
            voff = np.concatenate([[0], np.cumsum(n_ * rcap)[:-1]]).astype(np.int64)
            A = ops.alloc(int(sizes.sum()))
            U = ops.alloc(int((m_ * rcap).sum()))
            V = ops.alloc(int((n_ * rcap).sum()))
            ops.fill([(idx(r), idx(c)) for r, c in chunk], A, offs)
            stats["fill_batches"] += 1
            stats["block_entry_evaluations"] += int(sizes.sum())
            tasks = np.zeros(len(chunk), dtype=ACA_TASK_DTYPE)
            tasks["a_off"], tasks["u_off"], tasks["v_off"] = offs, uoff, voff
            tasks["m"], tasks["n"], tasks["rcap"] = m_, n_, rcap
            ranks, rels, status = ops.aca(A, tasks, U, V, float(cfg.epsilon))
            stats["aca_batches"] += 1
            acc_idx = []
            for i, (r, c) in enumerate(chunk):
                rank, rel = int(ranks[i]), float(rels[i])
                if status[i] == 2 and storage_cap[i]:
                    decision[key(r, c)] = ("dense",)
                    rejected.append((r, c))
                    stats["dense_blocks"] += 1
                    stats["rejected_low_rank_blocks"] += 1
                    continue
                meets = np.isfinite(rel) and rel <= cfg.epsilon * margin
                if rank > 0 and meets and (m_[i] + n_[i]) * rank < cfg.low_rank_storage_factor * sizes[i]:
                    acc_idx.append(i)
                    decision[key(r, c)] = ("far",)
                    continue
                if not meets and (_can_split(r) or _can_split(c)):
                    stats["compression_tolerance_subdivisions"] += 1
                    which, ch = _children(r, c)
                    decision[key(r, c)] = ("split", which)
                    nxt += ch
                    continue
                decision[key(r, c)] = ("dense",)
                stats["dense_blocks"] += 1
                stats["rejected_low_rank_blocks"] += 1
                if status[i] != 3:
                    rejected.append((r, c))
            if acc_idx:
                Uh = ops.gather(U, [(int(uoff[i]), int(uoff[i]) + int(m_[i]) * int(ranks[i])) for i in acc_idx])
                Vh = ops.gather(V, [(int(voff[i]), int(voff[i]) + int(n_[i]) * int(ranks[i])) for i in acc_idx])
                pu = pv = 0
                for i in acc_idx:
                    r, c = chunk[i]
                    mm, nn, rk = int(m_[i]), int(n_[i]), int(ranks[i])
                    Ub = np.ascontiguousarray(Uh[pu:pu + mm * rk].reshape(rk, mm).T)
                    Vb = Vh[pv:pv + nn * rk].reshape(rk, nn).copy()
                    pu += mm * rk
                    pv += nn * rk
                    far_of[key(r, c)] = LowRankBlockCPU(r.start, r.stop, c.start, c.stop, Ub, Vb, float(rels[i]))
                    stats["low_rank_blocks"] += 1
                    stats["max_observed_rank"] = max(stats["max_observed_rank"], rk)
                    stats["sum_observed_rank"] += rk
            del A, U, V
        add_near(rejected)
        queue = nxt

    # far blocks in the order of the depth-first recursion
    far_blocks = []
    stack = [(tree.root, tree.root)]
    while stack:
        r, c = stack.pop()
        d = decision[key(r, c)]
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


def build_or_load_hmatrix(cache_path, geo: BoundaryGeometry, config: HMatrixBuildConfig | None = None,
                          device=None, force_rebuild=False, verbose=False, cache_compressed=False):
    """Reuse a cache when its geometry hash, build configuration and construction
    version match (builder.py:814-927); otherwise build on the device and save."""
    from pathlib import Path
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
