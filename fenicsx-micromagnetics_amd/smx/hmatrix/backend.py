"""Packed H-matrix on the device and its matvec (HMatrixGPUPackedFused).

Same constructor / ``matvec(x, out=None)`` / ``benchmark`` / ``memory_report``
surface as lindholm_hmatrix_gpu/backend_fused.py:359-646.  Device layout
(every factor load of the kernels is 16 bytes and aligned):

* V of every far block row-major (r x n), each row padded to a multiple of 4
  entries, so ``z_b = V_b x`` streams whole 16-byte vectors; the z stage runs
  one wavefront per (block, group of 8 ranks) task, each task a 32-byte
  record with its offsets;
* leaves: the coarsest row intervals on which every far block is either fully
  present or absent (cut to <= 64 rows).  Each leaf owns one dense row-major
  matrix whose columns are the ranks of all far blocks over its rows (in block
  order), rows padded to a multiple of 4 entries.  The leaf's output row is a
  single dot product of that row with the concatenated z slices (gathered
  into LDS through a per-leaf index list), plus its near row, so one
  workgroup produces each output row without atomics;
* near field as CSR in permuted coordinates (FP64 values, int32 columns, the
  exact zeros dropped as storage.py:453-459 does), every row padded to an even
  entry count so values and columns load in aligned pairs.

The transposed operator (``rmatvec``) is the same layout built from the
swapped factors (U_b^T as the z-stage matrix, V_b^T as the leaf rows, column
leaves, near^T), packed on first use.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from .. import _capi
from ..device import to_dev
from .storage import HMatrixCPUData

_ZG = _capi.HMAT_Z_GROUP


# V row padding in entries by factor size (SMX_HMAT_V_ALIGN_F32 / _F64): FP32
# rows start on 128-byte cache lines (film z launch 710.7 -> 697.8 MB
# fetched, matvec 286.2 -> 284.3 us); FP64 rows keep 4 entries (128-byte rows
# made the FP64 matvec slower, 532.3 -> 536.4 us)
V_ROW_ENTRIES = {4: 32, 8: 4}
# near row padding in entries by value size: rows of values start on 128-byte
# lines (the pairs the kernels read need an even count)
NEAR_ROW_ENTRIES = {8: 16, 4: 32}


def _round_to(a, m):
    return (np.asarray(a, dtype=np.int64) + m - 1) // m * m


def _round4(a):
    return (np.asarray(a, dtype=np.int64) + 3) // 4 * 4


def leaf_intervals(nb, starts_b, stops_b):
    """Coarsest intervals on which every block's row range is all-or-nothing,
    cut to <= 64 rows."""
    bounds = np.unique(np.concatenate([[0, nb], np.asarray(starts_b, np.int64), np.asarray(stops_b, np.int64)]))
    starts, stops = [], []
    for a, b in zip(bounds[:-1], bounds[1:]):
        for s in range(int(a), int(b), 64):
            starts.append(s)
            stops.append(min(s + 64, int(b)))
    return np.array(starts, dtype=np.int64), np.array(stops, dtype=np.int64)


ZTASK_DTYPE = np.dtype([("v_off", np.int64), ("z_off", np.int64), ("c0", np.int32), ("n", np.int32),
                        ("kc", np.int32), ("pad", np.int32)])   # smx_hmat_ztask, 32 bytes


def _xcd_slots(work, per_wg):
    """Work items in position order (``work`` per item) cut into eight
    groups of equal total work, each largest first, laid out so that
    workgroup 8 i + x (``per_wg`` consecutive slots) takes group x's i-th
    items: the dispatcher sends workgroup b to XCD b % 8.  Returns the item
    of every slot, -1 for the padding of the shorter groups."""
    work = np.asarray(work, dtype=np.int64)
    before = np.cumsum(work) - work
    grp = np.minimum(before * 8 // max(int(work.sum()), 1), 7)
    lists = []
    for x in range(8):
        sel = np.flatnonzero(grp == x)
        lists.append(sel[np.argsort(-work[sel], kind="stable")])
    nwg = max((len(l) + per_wg - 1) // per_wg for l in lists)
    table = np.full((nwg, 8, per_wg), -1, dtype=np.int64)
    for x, l in enumerate(lists):
        col = np.full(nwg * per_wg, -1, dtype=np.int64)
        col[:len(l)] = l
        table[:, x, :] = col.reshape(nwg, per_wg)
    return table.reshape(-1)


def pack_operator(nb, row_start, row_stop, col_start, col_stop, rank, Ufac, Vfac, near_csr, fdt, near_dt,
                  xcd_order=True):
    """Host arrays of the packed layout for y = D x + Near x + sum_b U_b V_b x.

    Ufac[b] is (m_b x r_b), Vfac[b] is (r_b x n_b).  Returns a dict of numpy
    arrays (all offsets in entries) plus the unpadded storage counts.  Work
    items (z tasks, leaves) are stored largest first: independent items in any
    order give the same result, and largest-first lets the dispatcher balance
    the tail."""
    r0 = np.asarray(row_start, dtype=np.int64)
    r1 = np.asarray(row_stop, dtype=np.int64)
    c0 = np.asarray(col_start, dtype=np.int64)
    c1 = np.asarray(col_stop, dtype=np.int64)
    rk = np.asarray(rank, dtype=np.int64)
    nfar = rk.size
    m, n = r1 - r0, c1 - c0
    fb = np.dtype(fdt).itemsize
    nv = np.dtype(near_dt).itemsize
    P = {}
    # ---- z stage: V rows padded to 4 entries, tasks of <= _ZG rank rows ----
    ld = _round_to(n, V_ROW_ENTRIES[fb])
    v_off = np.zeros(nfar, dtype=np.int64)
    if nfar:
        v_off[1:] = np.cumsum(rk * ld)[:-1]
    V = np.zeros(max(int((rk * ld).sum()), 4), dtype=fdt)
    for b in range(nfar):
        if rk[b]:
            V[v_off[b]:v_off[b] + rk[b] * ld[b]].reshape(rk[b], ld[b])[:, :n[b]] = Vfac[b]
    # z entries grouped by row cluster (blocks sorted by row interval): the
    # blocks over one leaf then give a few contiguous z runs, one per row
    # cluster above it (film: 3.5 runs per leaf against 41 block runs)
    zkey = np.lexsort((np.arange(nfar), r1, r0))
    z_off = np.zeros(nfar, dtype=np.int64)
    if nfar:
        z_off[zkey] = np.cumsum(rk[zkey]) - rk[zkey]
    def tasks(group, xcd=False):
        """Tasks of <= group rank rows of one block (rows contiguous in V),
        largest first.  xcd: the tasks are cut by column position into eight
        groups of equal work, each largest first, and interleaved so that
        workgroup 8 i + x (of four tasks) runs group x's i-th four on XCD x:
        each XCD's L2 then holds the x slices of one eighth of the columns
        instead of all of them.  Short groups are padded with empty tasks
        (n = kc = 0)."""
        ng = (rk + group - 1) // group
        tb = np.repeat(np.arange(nfar, dtype=np.int64), ng)
        k0 = (np.arange(int(ng.sum()), dtype=np.int64) - np.repeat(np.cumsum(ng) - ng, ng)) * group
        kc = np.minimum(group, rk[tb] - k0)
        work = n[tb] * kc
        if xcd and tb.size:
            oc = np.lexsort((k0, tb, c0[tb]))             # by first column, block, rank row
            slots = _xcd_slots(work[oc], 4)
            order = np.where(slots >= 0, oc[np.maximum(slots, 0)], -1)
        else:
            order = np.argsort(-work, kind="stable")
        live = order >= 0
        ti = order[live]
        tb, k0, kc = tb[ti], k0[ti], kc[ti]
        zt = np.zeros(max(order.size, 1), dtype=ZTASK_DTYPE)   # empty tasks: n = kc = 0
        zt["v_off"][live] = v_off[tb] + k0 * ld[tb]
        zt["z_off"][live] = z_off[tb] + k0
        zt["c0"][live] = c0[tb]
        zt["n"][live] = n[tb]
        zt["kc"][live] = kc
        return zt, int(order.size)

    zt, nzt = tasks(_ZG, xcd=xcd_order)
    # tasks of the multi-RHS product: one 16-row MFMA tile of rank rows each
    # (smx_hmat_matmat's per-task limit; two or four tiles per task, sharing
    # the x loads, measured slower on the film: 558 -> 633 us for four), in
    # per-XCD column groups for either factor type (each L2 then holds the
    # staged X rows of one eighth of the columns: film 64-RHS product
    # 1377 -> 1339 us FP32, 2062 -> 1968-2007 us FP64)
    mm_rows = MM_TASK_ROWS
    ztm, nztm = tasks(mm_rows, xcd=True)
    P.update(V=V, ztask=zt.view(np.int64), nztask=nzt, ztask_mm=ztm.view(np.int64), nztask_mm=nztm,
             ztask_mm_rows=mm_rows)
    # ---- leaves ----
    starts, stops = leaf_intervals(nb, r0, r1)
    nleaf = starts.size
    first = np.searchsorted(starts, r0)
    last = np.searchsorted(starts, r1)
    cnt = last - first
    blk = np.repeat(np.arange(nfar, dtype=np.int64), cnt)
    leaf = np.repeat(first, cnt) + (np.arange(int(cnt.sum()), dtype=np.int64) - np.repeat(np.cumsum(cnt) - cnt, cnt))
    o = np.lexsort((blk, r1[blk], r0[blk], leaf))
    blk, leaf = blk[o], leaf[o]                  # (leaf, block) pairs, blocks in z order in a leaf
    fe_ptr = np.zeros(nleaf + 1, dtype=np.int64)
    np.add.at(fe_ptr, leaf + 1, 1)
    fe_ptr = np.cumsum(fe_ptr)
    rke = rk[blk]
    cum = np.cumsum(rke) - rke
    base = np.zeros(nleaf, dtype=np.int64)
    nonempty = np.diff(fe_ptr) > 0
    base[nonempty] = cum[fe_ptr[:-1][nonempty]]
    kpos = cum - base[leaf]                      # first column of each pair in its leaf matrix
    K = np.zeros(nleaf, dtype=np.int64)
    np.add.at(K, leaf, rke)
    kw = _round_to(K, V_ROW_ENTRIES[fb])        # U rows padded as V rows
    cl = stops - starts
    near = near_csr.tocsr()
    near.sort_indices()
    nptr0 = near.indptr.astype(np.int64)
    work = cl * kw * fb + (nptr0[stops] - nptr0[starts]) * (nv + 4)
    # storage slot -> leaf (-1: an empty padding slot); as the z tasks, the
    # leaves are cut by row position into eight groups of equal work, each
    # largest first, interleaved so that workgroup 8 i + x takes group x's
    # i-th leaf on XCD x (its L2 then holds the z entries of one eighth of
    # the rows' blocks)
    lo = _xcd_slots(work, 1) if (xcd_order and nleaf) else np.argsort(-work, kind="stable")
    live = lo >= 0
    lol = lo[live]
    u_off = np.zeros(nleaf, dtype=np.int64)      # per leaf
    sz = np.zeros(lo.size, dtype=np.int64)
    sz[live] = (cl * kw)[lol]
    u_off[lol] = (np.cumsum(sz) - sz)[live]
    U = np.zeros(max(int((cl * kw).sum()), 4), dtype=fdt)
    o_rows = starts[leaf] - r0[blk]
    for e in range(blk.size):
        b, L = int(blk[e]), int(leaf[e])
        c, w, orow = int(cl[L]), int(kw[L]), int(o_rows[e])
        U[u_off[L]:u_off[L] + c * w].reshape(c, w)[:, kpos[e]:kpos[e] + rke[e]] = Ufac[b][orow:orow + c]
    # z gather lists: column k of leaf L reads z[z_off[b] + j] for its (b, j)
    zg_leaf = np.repeat(z_off[blk], rke) + (np.arange(int(rke.sum()), dtype=np.int64)
                                            - np.repeat(np.cumsum(rke) - rke, rke))
    gptr = np.zeros(nleaf + 1, dtype=np.int64)
    gptr[1:] = np.cumsum(K)
    zg = np.concatenate([zg_leaf[gptr[L]:gptr[L + 1]] for L in lol]) if nleaf else np.zeros(0, np.int64)
    kslot = np.zeros(lo.size, dtype=np.int64)
    kslot[live] = K[lol]
    zg_off = np.zeros(lo.size + 1, dtype=np.int64)
    zg_off[1:] = np.cumsum(kslot)
    # the same lists as runs of consecutive z entries: (first z entry, first
    # column) per run, leaf_zr_off[L] .. leaf_zr_off[L+1] the runs of slot L
    kk = np.arange(zg.size, dtype=np.int64) - np.repeat(zg_off[:-1], kslot)
    rs = np.flatnonzero((kk == 0) | (np.diff(zg, prepend=-2) != 1))
    zr = np.stack([zg[rs], kk[rs]], axis=1).astype(np.int32).ravel()
    zr_off = np.zeros(lo.size + 1, dtype=np.int64)
    np.add.at(zr_off, np.repeat(np.arange(lo.size, dtype=np.int64), kslot)[rs] + 1, 1)
    zr_off = np.cumsum(zr_off)
    pick = lambda a: np.where(live, a[np.maximum(lo, 0)], 0) if nleaf else a[lo]
    P.update(leaf_start=pick(starts).astype(np.int32), leaf_stop=pick(stops).astype(np.int32),
             leaf_u_off=pick(u_off), leaf_kw=pick(kw).astype(np.int32), leaf_zg_off=zg_off,
             leaf_zg=(zg if zg.size else np.zeros(1, np.int64)).astype(np.int32), U=U,
             leaf_zr_off=zr_off, leaf_zr=(zr if zr.size else np.zeros(2, np.int32)),
             kw_max=int(kw.max()) if nleaf else 0)
    # ---- near rows padded to whole 128-byte lines of values ----
    cnt = np.diff(nptr0)
    pc = _round_to(cnt, NEAR_ROW_ENTRIES[nv])
    ptr = np.zeros(nb + 1, dtype=np.int64)
    ptr[1:] = np.cumsum(pc)
    idx = np.repeat(np.arange(nb, dtype=np.int64), pc)   # padding: a valid column with value 0
    val = np.zeros(max(int(ptr[-1]), 2), dtype=near_dt)
    rowid = np.repeat(np.arange(nb, dtype=np.int64), cnt)
    dest = ptr[rowid] + (np.arange(near.nnz, dtype=np.int64) - np.repeat(nptr0[:-1], cnt))
    idx[dest] = near.indices
    val[dest] = near.data
    if idx.size == 0:
        idx = np.zeros(2, dtype=np.int64)
    P.update(near_ptr=ptr, near_idx=idx.astype(np.int32), near_val=val)
    P.update(pack_near_tasks(ptr, idx, starts, stops, xcd_order))
    P["far_storage_entries"] = int((m * rk).sum() + (n * rk).sum())
    P["near_nnz"] = int(near.nnz)
    P["nz"] = int(rk.sum())
    if P["kw_max"] > _capi.HMAT_KW_MAX:
        raise ValueError(f"a leaf carries {P['kw_max']} far ranks; the matvec stages at most "
                         f"{_capi.HMAT_KW_MAX} in LDS")
    return P


MM_TASK_ROWS = 16   # rank rows per multi-RHS z task (SMX_HMAT_MM_ROWS)

NTASK_DTYPE = np.dtype([("s_off", np.int64), ("p0", np.int32), ("p1", np.int32), ("w", np.int32),
                        ("pad", np.int32)])   # smx_hmat_ntask, 24 bytes


def pack_near_tasks(ptr, idx, starts, stops, xcd_order=True, min_tasks=1024):
    """Tasks of the near launch (hm_near_lds_kernel): row ranges inside one
    leaf interval (at most rt rows, rt chosen so that there are about
    ``min_tasks`` tasks, between 4 and 64 rows), the distinct columns their
    entries touch (staged in LDS, in increasing order) and the slot of every
    entry's column in its task's list (uint16, stored as int16 bits).  Tasks
    in per-XCD row groups as the leaves (xcd_order) or largest first.  {} when
    a task would touch more than SMX_HMAT_NEAR_W_MAX columns."""
    ptr = np.asarray(ptr, dtype=np.int64)
    idx = np.asarray(idx, dtype=np.int64)
    nb = ptr.size - 1
    if nb == 0 or ptr[-1] == 0:
        return {}
    rt = int(min(64, max(4, 4 * -(-nb // (4 * min_tasks)))))
    starts = np.asarray(starts, dtype=np.int64)
    stops = np.asarray(stops, dtype=np.int64)
    nch = -(-(stops - starts) // rt)
    tl = np.repeat(np.arange(starts.size), nch)
    j = np.arange(int(nch.sum()), dtype=np.int64) - np.repeat(np.cumsum(nch) - nch, nch)
    p0 = starts[tl] + j * rt
    p1 = np.minimum(p0 + rt, stops[tl])
    ntask = p0.size
    e0, e1 = ptr[p0], ptr[p1]
    cnt = e1 - e0
    tid = np.repeat(np.arange(ntask, dtype=np.int64), cnt)     # tasks cover the entries in order
    uk, inv = np.unique(tid * nb + idx[:int(ptr[-1])], return_inverse=True)
    first = np.searchsorted(uk // nb, np.arange(ntask, dtype=np.int64))
    w = np.diff(np.append(first, uk.size))
    if int(w.max()) > _capi.HMAT_NEAR_W_MAX:
        return {}
    lidx = (inv.reshape(-1) - first[tid]).astype(np.uint16)
    order = _xcd_slots(cnt, 1) if xcd_order else np.argsort(-cnt, kind="stable")
    live = order >= 0
    tk = np.zeros(order.size, dtype=NTASK_DTYPE)              # empty tasks: p0 = p1 = w = 0
    o = order[live]
    tk["s_off"][live] = first[o]
    tk["p0"][live] = p0[o]
    tk["p1"][live] = p1[o]
    tk["w"][live] = w[o]
    return dict(near_task=tk.view(np.int64), n_near_task=int(order.size),
                near_stage=(uk % nb).astype(np.int32), near_lidx=lidx.view(np.int16),
                near_w_max=int(w.max()), near_stage_total=int(uk.size))


NTILE_DTYPE = np.dtype([("a_off", np.int64), ("col_off", np.int64), ("w", np.int32),
                        ("nrows", np.int32)])   # smx_hmat_ntile, 24 bytes
NTILE_ROWS = 16


def pack_near_tiles(nb, near_csr, leaf_starts, near_dt):
    """The near rows (permuted coordinates) as dense tiles of <= 16 rows for
    the multi-RHS product: tile t holds rows row[16 t + s] (s < nrows, -1
    beyond) and the union of their columns, col[col_off .. col_off + w), its
    values row-major at a_off with rows padded to a multiple of 4 entries
    (zeros where a row has no entry).

    Rows of one leaf interval (``leaf_starts``: the ascending starts of the
    intervals that partition 0..nb) share their near blocks, but their
    patterns still differ: on a flat face the coplanar entries are exact
    zeros and are not stored.  The rows of a leaf are therefore grouped by
    their column pattern (an order-free hash of the column set); each group
    gives full 16-row tiles, and the groups' remainders share tiles in
    pattern order.  Tiles are stored in the order of their first row (the
    device maps contiguous eighths of them to the eight XCDs, so each XCD's
    L2 serves the X rows of one part of the surface).  Every row is in
    exactly one tile, rows without near entries included."""
    near = near_csr.tocsr().copy()
    near.sum_duplicates()
    near.sort_indices()
    ptr = near.indptr.astype(np.int64)
    cols = near.indices.astype(np.int64)
    cnt = np.diff(ptr)
    key = np.zeros(nb, dtype=np.uint64)
    if cols.size:
        h = (cols.astype(np.uint64) + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15)
        h ^= h >> np.uint64(31)
        h *= np.uint64(0xBF58476D1CE4E5B9)
        live = cnt > 0
        key[live] = np.add.reduceat(h, ptr[:-1][live])
    leaf = np.searchsorted(np.asarray(leaf_starts, dtype=np.int64), np.arange(nb, dtype=np.int64),
                           side="right") - 1
    o = np.lexsort((np.arange(nb), key, leaf))          # rows by leaf, pattern, index
    lf, ky = leaf[o], key[o]
    newg = np.ones(nb, dtype=bool)
    newg[1:] = (lf[1:] != lf[:-1]) | (ky[1:] != ky[:-1])
    gid = np.cumsum(newg) - 1
    gstart = np.flatnonzero(newg)
    gsize = np.diff(np.append(gstart, nb))
    full = (np.arange(nb) - gstart[gid]) < (gsize[gid] // NTILE_ROWS) * NTILE_ROWS
    rows_full = o[full].reshape(-1, NTILE_ROWS)         # whole groups of 16 equal patterns
    rest, rest_leaf = o[~full], lf[~full]
    nr = rest.size
    newl = np.ones(nr, dtype=bool)
    newl[1:] = rest_leaf[1:] != rest_leaf[:-1]
    lstart = np.flatnonzero(newl)
    r_in_l = np.arange(nr) - lstart[np.cumsum(newl) - 1]
    chunk_new = newl | (r_in_l % NTILE_ROWS == 0)       # remainders of a leaf, 16 at a time
    cid = np.cumsum(chunk_new) - 1
    nrest_t = int(chunk_new.sum())
    ntile = rows_full.shape[0] + nrest_t
    tile_of = np.empty(nb, dtype=np.int64)
    slot_of = np.empty(nb, dtype=np.int64)
    tile_of[rows_full.ravel()] = np.repeat(np.arange(rows_full.shape[0]), NTILE_ROWS)
    slot_of[rows_full.ravel()] = np.tile(np.arange(NTILE_ROWS), rows_full.shape[0])
    cstart = np.flatnonzero(chunk_new)
    tile_of[rest] = rows_full.shape[0] + cid
    slot_of[rest] = np.arange(nr) - cstart[cid]
    nrows = np.bincount(tile_of, minlength=ntile).astype(np.int64)
    # column union per tile
    rowid = np.repeat(np.arange(nb, dtype=np.int64), cnt)
    te = tile_of[rowid]
    u, inv = np.unique(te * nb + cols, return_inverse=True)
    ut = u // nb
    w = np.bincount(ut, minlength=ntile).astype(np.int64)
    # storage order: by first row
    first = np.full(ntile, nb, dtype=np.int64)
    np.minimum.at(first, tile_of, np.arange(nb, dtype=np.int64))
    order = np.argsort(first, kind="stable")            # storage slot -> tile
    rank = np.empty(ntile, dtype=np.int64)
    rank[order] = np.arange(ntile)
    ld = _round4(w)
    a_sz = nrows * ld
    a_off = np.zeros(ntile, dtype=np.int64)
    a_off[order] = np.concatenate([[0], np.cumsum(a_sz[order])[:-1]]) if ntile else []
    col_off = np.zeros(ntile, dtype=np.int64)
    col_off[order] = np.concatenate([[0], np.cumsum(w[order])[:-1]]) if ntile else []
    # u is sorted by (tile, column): its tile runs start at the old-order offsets
    old_off = np.concatenate([[0], np.cumsum(w)[:-1]]) if ntile else np.zeros(0, np.int64)
    pos = inv - old_off[te]
    col = np.empty(max(int(w.sum()), 1), dtype=np.int32)
    col[:1] = 0
    col[col_off[ut] + (np.arange(u.size) - old_off[ut])] = (u % nb).astype(np.int32)
    val = np.zeros(max(int(a_sz.sum()), 4), dtype=near_dt)
    val[a_off[te] + slot_of[rowid] * ld[te] + pos] = near.data
    row = np.full((ntile, NTILE_ROWS), -1, dtype=np.int32)
    row[rank[tile_of], slot_of] = np.arange(nb, dtype=np.int32)
    tiles = np.zeros(ntile, dtype=NTILE_DTYPE)
    tiles["a_off"] = a_off[order]
    tiles["col_off"] = col_off[order]
    tiles["w"] = w[order]
    tiles["nrows"] = nrows[order]
    return dict(tile=tiles, row=row.ravel(), col=col, val=val, ntile=ntile,
                dense_entries=int(a_sz.sum()), mfma_entries=int((NTILE_ROWS * ld).sum()), nnz=int(near.nnz),
                w_max=int(w.max()) if ntile else 0, w_mean=float(w.mean()) if ntile else 0.0)


def near_from_tiles(nb, tiles, row, col, val):
    """The near matrix (permuted coordinates, scipy CSR) a tile pack encodes
    (host check of pack_near_tiles)."""
    import scipy.sparse as sp
    rs, cs, vs = [], [], []
    row = np.asarray(row).reshape(-1, NTILE_ROWS)
    for t in range(tiles.size):
        a0, c0, w, nr = (int(tiles["a_off"][t]), int(tiles["col_off"][t]), int(tiles["w"][t]),
                         int(tiles["nrows"][t]))
        ld = (w + 3) // 4 * 4
        A = np.asarray(val[a0:a0 + nr * ld]).reshape(nr, ld)[:, :w]
        rr, kk = np.nonzero(A)
        rs.append(row[t, rr])
        cs.append(np.asarray(col[c0:c0 + w])[kk])
        vs.append(A[rr, kk])
    cat = (lambda a: np.concatenate(a) if a else np.zeros(0))
    return sp.csr_matrix((cat(vs), (cat(rs).astype(np.int64), cat(cs).astype(np.int64))), shape=(nb, nb))


class HMatrixGPUPackedFused:
    def __init__(self, data: HMatrixCPUData, threads_per_block: int = 128, use_fp32: bool = True,
                 device=None, near_fp32: bool = False, boundary_gather: np.ndarray | None = None,
                 near_in_z="auto", xcd_order: bool | None = None, near_tasks: bool = True):
        self.data = data
        # near_tasks: the near launch as one workgroup per row range with its
        # x columns staged in LDS (pack_near_tasks); False: a wavefront per row
        self.near_tasks = bool(near_tasks)
        # the leaf launch stages its z entries by runs (leaf_zr) instead of
        # one gather index per entry (leaf_zg); the same values either way
        self.leaf_z_runs = True
        # xcd_order: z tasks and leaves in per-XCD position groups
        # (pack_operator); False: one list each, largest first.  None: on
        # (film matvec 310 -> 300 us with FP32 factors, 533 -> 530 us with
        # FP64; the 64-RHS product unchanged with FP64 factors, 1454 -> 1424
        # us with FP32)
        self.xcd_order = True if xcd_order is None else bool(xcd_order)
        # near_in_z: the near rows are summed by a launch of their own (one
        # wavefront per row) instead of inside the leaf launch.  It pays when
        # the leaves are too few to fill the device (the 150 nm cube has 84);
        # "auto" times both once (cache evicted) and keeps the faster.
        self._near_mode_request = near_in_z
        self.near_in_z = bool(near_in_z) if near_in_z != "auto" else True
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
        self._fdt = np.float32 if self.use_fp32 else np.float64
        self._ndt = np.float32 if near_fp32 else np.float64
        self.factor_dtype = _capi.SMX_F32 if self.use_fp32 else _capi.SMX_F64
        self.near_dtype = _capi.SMX_F32 if near_fp32 else _capi.SMX_F64
        fb = data.far_blocks
        self.nfar = len(fb)
        host = pack_operator(nb, [b.row_start for b in fb], [b.row_stop for b in fb],
                             [b.col_start for b in fb], [b.col_stop for b in fb], [b.rank for b in fb],
                             [b.U for b in fb], [b.V for b in fb], data.near_csr, self._fdt, self._ndt,
                             xcd_order=self.xcd_order)
        self.far_storage_entries = host["far_storage_entries"]
        self.near_nnz = host["near_nnz"]
        self.fwd = self._upload(host)
        self.nleaf = int(self.fwd["leaf_start"].numel())
        self.U, self.V = self.fwd["U"], self.fwd["V"]
        self.xp = torch.zeros(nb, dtype=torch.float64, device=dev)
        self.z = torch.zeros(max(host["nz"], 1), dtype=torch.float64, device=dev)
        self._nz = int(host["nz"])
        self.ynear = torch.zeros(max(nb, 1), dtype=torch.float64, device=dev)
        self._H = self._struct(self.fwd, self.gidx_plain)
        self._Hb = self._struct(self.fwd, self.gidx_boundary) if self.gidx_boundary is not None else None
        self.near_mode_seconds = {}
        if near_in_z == "auto" and self.near_nnz > 0:
            self._choose_near_mode()

    def _choose_near_mode(self, reps: int = 5):
        """Time the matvec with the near rows in the z launch and in the leaf
        launch (HIP events, after a warm-up, each timed call after a read
        sweep that evicts the 256 MB Infinity Cache, as inside an evaluation
        where the operator streams from HBM) and keep the faster."""
        x = torch.randn(self.nb, dtype=torch.float64, device=self.device)
        out = torch.empty_like(x)
        sweep = torch.ones(128 * 1024 * 1024, dtype=torch.float32, device=self.device)   # 512 MB
        for mode in (True, False):
            self.set_near_mode(mode)
            for _ in range(2):
                self.matvec(x, out)
            ts = []
            for _ in range(reps):
                sweep.sum()
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                self.matvec(x, out)
                e1.record()
                e1.synchronize()
                ts.append(e0.elapsed_time(e1) * 1e-3)
            self.near_mode_seconds[mode] = min(ts)
        del sweep
        self.set_near_mode(self.near_mode_seconds[True] < self.near_mode_seconds[False])

    def set_near_mode(self, near_in_z: bool):
        """Sum the near rows in a launch of their own (True) or inside the leaf
        launch (False) from the next matvec on."""
        self.near_in_z = bool(near_in_z)
        self._H = self._struct(self.fwd, self.gidx_plain)
        if self.gidx_boundary is not None:
            self._Hb = self._struct(self.fwd, self.gidx_boundary)

    def _upload(self, host):
        return {k: (to_dev(v, self.device) if isinstance(v, np.ndarray) else v) for k, v in host.items()}

    def _struct(self, pk, gidx):
        H = _capi.HMatrixArgs()
        p = _capi.ptr
        H.nb = self.nb
        H.gidx = p(gidx)
        H.perm = p(self.perm)
        H.diag_p = p(self.diag_p)
        has_near = pk["near_nnz"] > 0
        H.near_ptr = p(pk["near_ptr"]) if has_near else None
        H.near_idx = p(pk["near_idx"]) if has_near else None
        H.near_val = p(pk["near_val"]) if has_near else None
        H.near_dtype = self.near_dtype
        H.factor_dtype = self.factor_dtype
        H.nztask = int(pk["nztask"])
        H.ztask = p(pk["ztask"])
        H.V = p(pk["V"])
        H.nleaf = int(pk["leaf_start"].numel())
        H.leaf_start, H.leaf_stop = p(pk["leaf_start"]), p(pk["leaf_stop"])
        H.leaf_u_off, H.leaf_kw = p(pk["leaf_u_off"]), p(pk["leaf_kw"])
        H.leaf_zg_off, H.leaf_zg = p(pk["leaf_zg_off"]), p(pk["leaf_zg"])
        if self.leaf_z_runs and "leaf_zr" in pk:
            H.leaf_zr_off, H.leaf_zr = p(pk["leaf_zr_off"]), p(pk["leaf_zr"])
        H.kw_max = int(pk["kw_max"])
        H.U = p(pk["U"])
        H.xp, H.z = p(self.xp), p(self.z)
        H.ynear = p(self.ynear) if (self.near_in_z and has_near) else None
        if H.ynear and self.near_tasks and pk.get("near_task") is not None:
            H.near_task = p(pk["near_task"])
            H.n_near_task = int(pk["n_near_task"])
            H.near_stage, H.near_lidx = p(pk["near_stage"]), p(pk["near_lidx"])
            H.near_w_max = int(pk["near_w_max"])
            H.near_waves = int(getattr(self, "near_waves", 0))   # 0: the library's default (8)
        return H

    def _near_by_task(self) -> bool:
        return self.near_in_z and self.near_tasks and self.near_nnz > 0 and "near_task" in self.fwd

    def transposed_pack(self):
        """Packed layout of B^T (built on first use): z stage over U_b^T,
        leaf rows from V_b^T, leaves over column intervals, near^T."""
        if not hasattr(self, "_tpk"):
            fb = self.data.far_blocks
            host = pack_operator(self.nb, [b.col_start for b in fb], [b.col_stop for b in fb],
                                 [b.row_start for b in fb], [b.row_stop for b in fb], [b.rank for b in fb],
                                 [np.asarray(b.V).T for b in fb], [np.asarray(b.U).T for b in fb],
                                 self.data.near_csr.T.tocsr(), self._fdt, self._ndt, xcd_order=self.xcd_order)
            self._tpk = self._upload(host)
        return self._tpk

    # ------------------------------------------------------------------
    def rmatvec(self, x_gpu: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """y = B^T x (original boundary ordering): the adjoint boundary operator."""
        if x_gpu.shape != (self.nb,):
            raise ValueError(f"Expected x.shape == {(self.nb,)}, got {tuple(x_gpu.shape)}.")
        _capi.require_device(x_gpu, "x")
        if out is None:
            out = torch.empty_like(x_gpu)
        if not hasattr(self, "_HT"):
            self._HT = self._struct(self.transposed_pack(), self.gidx_plain)
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

    def matvec_noalloc(self, x_gpu: torch.Tensor, out: torch.Tensor) -> torch.Tensor:
        """y = B x through smx_hmat_matvec_noalloc: the same kernels with the
        x, z and near-sum scratch in the library's fixed device array instead
        of this object's workspaces (for callers that cannot allocate; out is
        the caller's)."""
        if x_gpu.shape != (self.nb,) or out.shape != (self.nb,):
            raise ValueError(f"Expected x.shape == out.shape == {(self.nb,)}.")
        _capi.require_device(x_gpu, "x")
        _capi.require_device(out, "out")
        _capi.check(_capi.lib().smx_hmat_matvec_noalloc(_capi.byref(self._H), self._nz, x_gpu.data_ptr(), None,
                                                        0.0, out.data_ptr(), _capi.stream_ptr()),
                    "smx_hmat_matvec_noalloc")
        return out

    def near_tiles(self):
        """The near rows as dense 16-row tiles on the device (pack_near_tiles),
        built on first use; returns the smx_hmat_near_tiles argument."""
        if getattr(self, "_NT", None) is None:
            fb = self.data.far_blocks
            starts, _ = leaf_intervals(self.nb, [b.row_start for b in fb], [b.row_stop for b in fb])
            host = pack_near_tiles(self.nb, self.data.near_csr, starts, self._ndt)
            self.near_tile_stats = {k: host[k] for k in ("ntile", "dense_entries", "mfma_entries", "nnz", "w_max", "w_mean")}
            dev = self._upload({k: host[k] for k in ("row", "col", "val")})
            dev["tile"] = to_dev(host["tile"].view(np.int64), self.device)
            perm = np.asarray(self.data.perm, dtype=np.int64)
            iperm = np.empty(self.nb, dtype=np.int32)
            iperm[perm] = np.arange(self.nb, dtype=np.int32)
            dev["iperm"] = to_dev(iperm, self.device)
            NT = _capi.NearTilesArgs()
            NT.ntile = int(host["ntile"])
            NT.tile, NT.row, NT.col, NT.val, NT.iperm = (_capi.ptr(dev[k]) for k in ("tile", "row", "col", "val",
                                                                                    "iperm"))
            NT.val_dtype = self.near_dtype
            self._nt_dev, self._NT = dev, NT
        return self._NT

    def matmat(self, X: torch.Tensor, out: torch.Tensor | None = None, near_tiles: bool = True) -> torch.Tensor:
        """Y = B X for up to 64 right-hand sides: X and Y are (k, nb) device
        tensors, row j = right-hand side j (boundary ordering).  The low-rank
        stages run as MFMA tiles over the columns (smx_hmat_matmat); the near
        field as FP64 MFMA tiles over dense 16-row near tiles (near_tiles=True,
        packed on the first call) or one wavefront per CSR row."""
        if X.dim() != 2 or X.shape[1] != self.nb or not 1 <= X.shape[0] <= 64:
            raise ValueError(f"Expected X of shape (k <= 64, {self.nb}), got {tuple(X.shape)}.")
        _capi.require_device(X, "X", numel=X.numel())
        X = X.contiguous()
        k = int(X.shape[0])
        if out is None:
            out = torch.empty_like(X)
        _capi.require_device(out, "out", numel=X.numel())
        nz = int(self._nz)
        if getattr(self, "_Hmm", None) is None:
            # the multi-RHS product walks its own task table (one MFMA row
            # tile of rank rows per task)
            Hm = self._struct(self.fwd, self.gidx_plain)
            Hm.ztask = _capi.ptr(self.fwd["ztask_mm"])
            Hm.nztask = int(self.fwd["nztask_mm"])
            Hm.ynear = None
            self._Hmm = Hm
        need = int(_capi.lib().smx_hmat_matmat_ws_size(C.byref(self._Hmm), nz))
        if getattr(self, "_mm_work", None) is None or self._mm_work.numel() < need:
            self._mm_work = torch.empty(max(need, 1), dtype=torch.uint8, device=self.device)
        NT = C.byref(self.near_tiles()) if near_tiles else None
        _capi.check(_capi.lib().smx_hmat_matmat(C.byref(self._Hmm), NT, nz, _capi.ptr(X), self.nb, k, _capi.ptr(out),
                                                self.nb, _capi.ptr(self._mm_work), self._mm_work.numel(),
                                                _capi.stream_ptr()), "smx_hmat_matmat")
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
        """Bytes one matvec must move, by the H-matrix matvec's per-unit
        figure: far_storage_entries x factor size + near_nnz x (value +
        4-byte column) + (nb + 1) x 4 row pointers + nb x 8 diagonal +
        4 nb x 8 for x and y with the permutation gather and scatter (padding
        excluded).  The forms differ in what they stream (stored_bytes): the
        near task form reads 2-byte column slots and per-task column lists
        instead of 4-byte columns."""
        fb = 4 if self.use_fp32 else 8
        nv = 4 if self.near_dtype == _capi.SMX_F32 else 8
        return int(self.far_storage_entries * fb + self.near_nnz * (nv + 4) + (self.nb + 1) * 4
                   + self.nb * 8 + 4 * self.nb * 8)

    def stored_bytes(self) -> int:
        """Bytes the kernels stream per matvec in this form, alignment padding
        included: factors, near values and column data (4-byte columns, or
        2-byte slots + task column lists + the staged x entries), int64 row
        pointers, the leaf z run table (or gather lists), task records,
        vectors."""
        pk = self.fwd
        size = lambda t: t.numel() * t.element_size()
        cols = (size(pk["near_lidx"]) + size(pk["near_stage"]) + 8 * int(pk["near_stage_total"])
                + size(pk["near_task"]) if self._near_by_task() else size(pk["near_idx"]))
        zlists = (size(pk["leaf_zr"]) + size(pk["leaf_zr_off"]) if self._H.leaf_zr
                  else size(pk["leaf_zg"]) + size(pk["leaf_zg_off"]))
        return int(size(pk["U"]) + size(pk["V"]) + size(pk["near_val"]) + cols + zlists + size(pk["ztask"])
                   + (self.nb + 1) * 8 + self.nb * 8 + 4 * self.nb * 8)

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
        arrays = [v for v in self.fwd.values() if isinstance(v, torch.Tensor)]
        arrays += [self.perm, self.diag_p, self.xp, self.z]
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
