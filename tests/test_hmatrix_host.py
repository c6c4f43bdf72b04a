"""H-matrix host logic on the CPU.

The builder is driven through a host stand-in of the device operations
(batched Lindholm fill, batched compression, memory queries).  The stand-in
checks every task of the tables the builder produces against the buffer sizes
the kernels would index, and evaluates the tasks with the oracle's Lindholm
weights and compressors.  The resulting partition must equal the oracle's
depth-first build, also under the device builder's memory rules (a workspace
limit, allocations that fail above a size) and with the SVD compressors.
"""
import math

import numpy as np
import pytest

import smx_oracle as O
from smx import mesh as M
from smx.boundary import build_boundary_geometry
from smx.hmatrix import HMatrixBuildConfig, HMatrixCPUData, build_cluster_tree, build_hmatrix_device
from smx.hmatrix.builder import DeviceBuildOps


class HostOps:
    def __init__(self, geo, bd, avail=None, alloc_limit=None):
        self.geo = geo
        self.bd = bd
        self.inc = geo.incidence()
        self.evaluations = 0
        self.avail = avail
        self.alloc_limit = alloc_limit
        self.failed_allocs = 0

    def available_bytes(self):
        return float("inf") if self.avail is None else self.avail

    def alloc(self, n):
        if self.alloc_limit is not None and 8 * int(n) > self.alloc_limit:
            self.failed_allocs += 1
            raise MemoryError("out of memory (host stand-in)")
        return np.full(max(int(n), 1), np.nan)

    def compress(self, method, A, tasks, U, V, rtol, max_rank, seeds, cfg):
        if method == "fullaca":
            return self.aca(A, tasks, U, V, rtol)
        n_t = tasks.size
        ranks, rels, status = np.zeros(n_t, np.int32), np.zeros(n_t), np.zeros(n_t, np.int32)
        for q, t in enumerate(tasks):
            m, n, rcap = int(t["m"]), int(t["n"]), int(t["rcap"])
            a0, u0, v0 = int(t["a_off"]), int(t["u_off"]), int(t["v_off"])
            assert a0 + m * n <= A.size and u0 + rcap * m <= U.size and v0 + rcap * n <= V.size
            R = A[a0:a0 + m * n].reshape(m, n)
            if not np.any(R) or (max_rank is not None and max_rank == 0):
                status[q] = 3
                continue
            if method == "svd":
                Uf, Vf, rel = O.truncated_svd(R, rtol, max_rank)
            else:
                Uf, Vf, rel = O.randomized_svd(R, rtol, max_rank, cfg.gpu_rsvd_initial_rank,
                                               cfg.gpu_rsvd_oversampling, cfg.gpu_rsvd_power_iterations,
                                               random_seed=int(seeds[q]))
            rk = Uf.shape[1]
            ranks[q], rels[q] = rk, rel
            if rk > rcap:
                status[q] = 2
                continue
            U[u0:u0 + rk * m] = Uf.T.ravel()
            V[v0:v0 + rk * n] = Vf.ravel()
        return ranks, rels, status

    def host(self, t):
        return t

    def gather(self, t, ranges):
        return np.concatenate([t[a:b] for a, b in ranges]) if ranges else np.zeros(0)

    def fill(self, blocks, out, offsets):
        tasks, wg_task, wg_row0, rows, rel, rc = DeviceBuildOps.fill_tasks(self, blocks, offsets)
        assert wg_task.size == wg_row0.size
        assert np.all(wg_row0 < tasks["nrows"][wg_task]) and np.all(wg_row0 % 4 == 0)
        for t in tasks:
            ro, nr, reo, nrel, nc = int(t["rows_off"]), int(t["nrows"]), int(t["rel_off"]), int(t["nrel"]), int(t["ncols"])
            assert 0 <= ro and ro + nr <= rows.size
            assert 0 <= reo and reo + nrel <= rel.size and 3 * (reo + nrel) <= rc.size
            assert 0 < nc <= 1024
            last = int(t["out_off"]) + (nr - 1) * int(t["ld_row"]) + (nc - 1) * int(t["ld_col"])
            assert 0 <= int(t["out_off"]) and last < out.size
            rr = rows[ro:ro + nr].astype(np.int64)
            tri = rel[reo:reo + nrel].astype(np.int64)
            cc = rc[3 * reo:3 * (reo + nrel)].reshape(-1, 3)
            assert cc.max() < nc and np.all(tri < self.bd.Nt)
            w = np.stack(O._weights_rows(self.bd, rr, tri), axis=2)
            lid = self.bd.tri_lid[tri].astype(np.int64)
            w[(lid[None] == rr[:, None, None]).any(axis=2)] = 0.0
            tile = np.zeros((nr, nc))
            for k in range(3):
                ok = cc[:, k] >= 0
                for i in range(nr):
                    np.add.at(tile[i], cc[ok, k], w[i, ok, k])
            for i in range(nr):
                base = int(t["out_off"]) + i * int(t["ld_row"])
                out[base:base + nc * int(t["ld_col"]):int(t["ld_col"])] = tile[i]

    def aca(self, A, tasks, U, V, rtol):
        n_t = tasks.size
        ranks = np.zeros(n_t, np.int32)
        rels = np.zeros(n_t)
        status = np.zeros(n_t, np.int32)
        for q, t in enumerate(tasks):
            m, n, rcap = int(t["m"]), int(t["n"]), int(t["rcap"])
            a0, u0, v0 = int(t["a_off"]), int(t["u_off"]), int(t["v_off"])
            assert a0 + m * n <= A.size and u0 + rcap * m <= U.size and v0 + rcap * n <= V.size
            R = A[a0:a0 + m * n].reshape(m, n)
            assert np.all(np.isfinite(R))
            nA = float(np.linalg.norm(R))
            rank, rel, st = 0, 1.0, 2
            if nA == 0.0 or rcap == 0:
                st, rel = 3, 0.0
            else:
                for k in range(rcap):
                    i, j = np.unravel_index(int(np.argmax(np.abs(R))), R.shape)
                    piv = R[i, j]
                    if abs(piv) <= np.finfo(float).eps * nA:
                        st = 1
                        break
                    u = R[:, j].copy()
                    v = R[i, :] / piv
                    U[u0 + k * m:u0 + (k + 1) * m] = u
                    V[v0 + k * n:v0 + (k + 1) * n] = v
                    R -= np.outer(u, v)
                    rel = float(np.linalg.norm(R) / nA)
                    rank = k + 1
                    if rel <= rtol:
                        st = 0
                        break
            ranks[q], rels[q], status[q] = rank, rel, st
        return ranks, rels, status


@pytest.fixture(scope="module")
def c1():
    m = M.cube_c1()
    return m, build_boundary_geometry(m.X, m.tets), O.boundary_data(m.X, m.tets)


def test_tree_bit_exact_vs_oracle(c1):
    _, geo, bd = c1
    for leaf in (16, 32, 64):
        t = build_cluster_tree(geo.Xb, leaf)
        perm_o, root_o = O.cluster_tree(bd.Xb, leaf)
        assert np.array_equal(t.perm, perm_o)
        a, b = [t.root], [root_o]
        while a:
            x, y = a.pop(), b.pop()
            assert (x.start, x.stop) == (y.start, y.stop)
            assert np.array_equal(x.bbox_min, y.bmin) and np.array_equal(x.bbox_max, y.bmax)
            assert x.is_leaf == y.is_leaf
            if not x.is_leaf:
                a += [x.left, x.right]
                b += [y.left, y.right]


@pytest.mark.parametrize("leaf,eps", [(32, 1e-6), (16, 1e-4)])
def test_partition_matches_oracle(c1, leaf, eps):
    _, geo, bd = c1
    cfg = HMatrixBuildConfig(epsilon=eps, leaf_size=leaf, batch_bytes=1 << 20)
    data = build_hmatrix_device(geo, cfg, ops=HostOps(geo, bd))
    ref = O.build_hmatrix(bd, eps=eps, eta=2.0, leaf_size=leaf)
    got = [(b.row_start, b.row_stop, b.col_start, b.col_stop, b.rank) for b in data.far_blocks]
    exp = [(r0, r1, c0, c1, U.shape[1]) for r0, r1, c0, c1, U, V, _ in ref.far]
    assert got == exp
    assert np.array_equal(data.near_csr.indptr, ref.near.indptr)
    assert np.array_equal(data.near_csr.indices, ref.near.indices)
    assert np.abs(data.near_csr.data - ref.near.data).max() < 1e-15
    for k in ("low_rank_blocks", "dense_blocks", "rejected_low_rank_blocks", "compression_tolerance_subdivisions",
              "forced_subdivisions_memory"):
        assert data.metadata[k] == ref.stats[k], k
    B = O.dense_B_rows(bd)
    rng = np.random.default_rng(1234)
    for _ in range(3):
        x = rng.standard_normal(bd.Nb)
        y = _matvec(data, x)
        assert np.linalg.norm(y - B @ x) / np.linalg.norm(B @ x) < 10 * eps


def test_forced_memory_subdivision(c1):
    _, geo, bd = c1
    cfg = HMatrixBuildConfig(epsilon=1e-6, leaf_size=32, max_temporary_block_bytes=1024)
    data = build_hmatrix_device(geo, cfg, ops=HostOps(geo, bd))
    ref = O.build_hmatrix(bd, eps=1e-6, leaf_size=32, max_temporary_block_bytes=1024)
    assert data.metadata["forced_subdivisions_memory"] == ref.stats["forced_subdivisions_memory"] > 0
    assert [(b.row_start, b.col_start) for b in data.far_blocks] == [(f[0], f[2]) for f in ref.far]


def test_cache_roundtrip(tmp_path, c1):
    _, geo, bd = c1
    data = build_hmatrix_device(geo, HMatrixBuildConfig(leaf_size=32), ops=HostOps(geo, bd))
    p = tmp_path / "h.npz"
    data.save(p, compressed=False)
    back = HMatrixCPUData.load(p)
    assert np.array_equal(back.perm, data.perm)
    assert (back.near_csr != data.near_csr).nnz == 0
    assert len(back.far_blocks) == len(data.far_blocks)
    for a, b in zip(back.far_blocks, data.far_blocks):
        assert np.array_equal(a.U, b.U) and np.array_equal(a.V, b.V)
    assert back.metadata["geometry_hash"] == data.metadata["geometry_hash"]
    # the archive has exactly the reference's key set
    ar = np.load(p, allow_pickle=False)
    assert set(ar.files) == {"perm", "diag_jump", "near_data", "near_indices", "near_indptr", "near_shape",
                             "far_meta", "far_U_flat", "far_V_flat", "metadata_json"}


def _matvec(data, x):
    """storage.py:284-299 semantics, evaluated here for the test only."""
    xp = x[data.perm]
    yp = data.diag_jump[data.perm] * xp + data.near_csr @ xp
    for b in data.far_blocks:
        yp[b.row_start:b.row_stop] += b.U @ (b.V @ xp[b.col_start:b.col_stop])
    y = np.empty_like(x)
    y[data.perm] = yp
    return y


def _emulate_packed(pk, nb, perm, diag_p, x):
    """Index-for-index host emulation of the three matvec kernels on the packed
    arrays of ``pack_operator`` (bounds, alignment and coverage checked)."""
    from smx.hmatrix.backend import ZTASK_DTYPE
    xp = x[perm]
    V = pk["V"].astype(np.float64)
    U = pk["U"].astype(np.float64)
    zG = 8
    z = np.full(max(pk["nz"], 1), np.nan)
    written = np.zeros(z.size, dtype=int)
    tasks = pk["ztask"].view(ZTASK_DTYPE)[:pk["nztask"]]
    for t in tasks:
        n, kc, c0, vo, zo = int(t["n"]), int(t["kc"]), int(t["c0"]), int(t["v_off"]), int(t["z_off"])
        if kc == 0:                      # padding of a short XCD group: no work
            assert n == 0
            continue
        assert 1 <= kc <= zG and c0 >= 0 and c0 + n <= nb
        al = {4: 32, 8: 4}[pk["V"].dtype.itemsize]  # row padding: 128 bytes FP32, 4 entries FP64
        ld = (n + al - 1) // al * al
        assert vo % al == 0 and vo + kc * ld <= V.size and zo + kc <= z.size
        Vt = V[vo:vo + kc * ld].reshape(kc, ld)
        assert np.all(Vt[:, n:] == 0)
        z[zo:zo + kc] = Vt[:, :n] @ xp[c0:c0 + n]
        written[zo:zo + kc] += 1
    assert np.all(written[:pk["nz"]] == 1)
    y = np.full(nb, np.nan)
    covered = np.zeros(nb, dtype=int)
    nptr, nidx, nval = pk["near_ptr"], pk["near_idx"], pk["near_val"].astype(np.float64)
    zg, zgo = pk["leaf_zg"], pk["leaf_zg_off"]
    for L in range(pk["leaf_start"].size):
        l0, l1 = int(pk["leaf_start"][L]), int(pk["leaf_stop"][L])
        c, kw, uo = l1 - l0, int(pk["leaf_kw"][L]), int(pk["leaf_u_off"][L])
        if c == 0:                       # padding of a short XCD group: no rows, no ranks
            assert kw == 0 and zgo[L + 1] == zgo[L]
            continue
        assert 0 < c <= 64 and kw % 4 == 0 and uo % 4 == 0 and uo + c * kw <= U.size
        assert kw <= pk["kw_max"]
        covered[l0:l1] += 1
        K = int(zgo[L + 1] - zgo[L])
        assert 0 <= kw - K < {4: 32, 8: 4}[pk["U"].dtype.itemsize]    # rows padded as V's
        idx = zg[zgo[L]:zgo[L + 1]]
        assert np.all((idx >= 0) & (idx < pk["nz"]))
        # the run form of the same list: contiguous z ranges, first columns increasing
        zr, zro = pk["leaf_zr"].reshape(-1, 2), pk["leaf_zr_off"]
        runs = zr[zro[L]:zro[L + 1]]
        ends = np.append(runs[1:, 1], K)
        assert runs.size == 0 or (runs[0, 1] == 0 and np.all(ends > runs[:, 1]))
        expanded = np.concatenate([np.arange(a, a + e - k) for (a, k), e in zip(runs, ends)] or [np.zeros(0, int)])
        assert np.array_equal(expanded, idx)
        zl = np.zeros(kw)
        zl[:K] = z[idx]
        UL = U[uo:uo + c * kw].reshape(c, kw)
        assert np.all(UL[:, K:] == 0)
        for i in range(c):
            p = l0 + i
            n0, n1 = int(nptr[p]), int(nptr[p + 1])
            assert n0 % 2 == 0 and n1 % 2 == 0 and n1 <= nidx.size
            assert np.all((nidx[n0:n1] >= 0) & (nidx[n0:n1] < nb))
            s = float(nval[n0:n1] @ xp[nidx[n0:n1]]) + float(UL[i] @ zl)
            y[perm[p]] = diag_p[p] * xp[p] + s
    assert np.all(covered == 1)
    return y


def _pack(data, transpose=False, fdt=np.float64, xcd_order=True):
    from smx.hmatrix.backend import pack_operator
    fb = data.far_blocks
    if transpose:
        return pack_operator(data.size, [b.col_start for b in fb], [b.col_stop for b in fb],
                             [b.row_start for b in fb], [b.row_stop for b in fb], [b.rank for b in fb],
                             [b.V.T for b in fb], [b.U.T for b in fb], data.near_csr.T.tocsr(), fdt, np.float64,
                             xcd_order=xcd_order)
    return pack_operator(data.size, [b.row_start for b in fb], [b.row_stop for b in fb],
                         [b.col_start for b in fb], [b.col_stop for b in fb], [b.rank for b in fb],
                         [b.U for b in fb], [b.V for b in fb], data.near_csr, fdt, np.float64, xcd_order=xcd_order)


@pytest.mark.parametrize("xcd_order", [True, False])
def test_packed_tables_and_matvec(c1, xcd_order):
    """Both work orders (per-XCD position groups with padding, or one
    largest-first list) hold every task and leaf once and give the matvec."""
    _, geo, bd = c1
    data = build_hmatrix_device(geo, HMatrixBuildConfig(leaf_size=32), ops=HostOps(geo, bd))
    pk = _pack(data, xcd_order=xcd_order)
    perm = np.asarray(data.perm)
    x = np.random.default_rng(0).standard_normal(geo.Nb)
    ref = _matvec(data, x)
    assert np.abs(_emulate_packed(pk, data.size, perm, data.diag_jump[perm], x) - ref).max() < 1e-12
    if xcd_order:
        # workgroup b runs on XCD b % 8: its slots must come from group b % 8,
        # whose first columns (z tasks) / rows (leaves) rise with the group
        from smx.hmatrix.backend import ZTASK_DTYPE
        t = pk["ztask"].view(ZTASK_DTYPE)[:pk["nztask"]]
        wg = np.arange(t.size) // 4
        live = t["kc"] > 0
        lo = [t["c0"][live & (wg % 8 == x)].min() for x in range(8) if np.any(live & (wg % 8 == x))]
        assert lo == sorted(lo)
        rows = pk["leaf_stop"] > pk["leaf_start"]
        first = [pk["leaf_start"][rows & (np.arange(rows.size) % 8 == x)].min() for x in range(8)
                 if np.any(rows & (np.arange(rows.size) % 8 == x))]
        assert first == sorted(first)
    assert pk["far_storage_entries"] == data.far_storage_entries
    assert pk["near_nnz"] == data.near_csr.nnz
    # FP32 factors: same tables, rounding only in the factors
    pk32 = _pack(data, fdt=np.float32)
    y32 = _emulate_packed(pk32, data.size, perm, data.diag_jump[perm], x)
    assert np.abs(y32 - ref).max() < 1e-5 * np.abs(ref).max()


def test_packed_transpose_tables(c1):
    """The transposed pack (factors swapped, column leaves, near^T) reproduces
    B^T x under the same index emulation."""
    _, geo, bd = c1
    data = build_hmatrix_device(geo, HMatrixBuildConfig(leaf_size=32), ops=HostOps(geo, bd))
    pk = _pack(data, transpose=True)
    perm = np.asarray(data.perm)
    x = np.random.default_rng(1).standard_normal(geo.Nb)
    xp = x[perm]
    yp = data.diag_jump[perm] * xp + data.near_csr.T @ xp
    for b in data.far_blocks:
        yp[b.col_start:b.col_stop] += b.V.T @ (b.U.T @ xp[b.row_start:b.row_stop])
    ref = np.empty_like(x)
    ref[perm] = yp
    assert np.abs(_emulate_packed(pk, data.size, perm, data.diag_jump[perm], x) - ref).max() < 1e-12


def test_pack_all_near(c1):
    """A leaf size of the whole boundary leaves no far block: no z tasks, empty
    leaf matrices, the matvec is diag + near only."""
    _, geo, bd = c1
    data = build_hmatrix_device(geo, HMatrixBuildConfig(leaf_size=4096), ops=HostOps(geo, bd))
    assert len(data.far_blocks) == 0
    pk = _pack(data)
    assert pk["nztask"] == 0 and pk["kw_max"] == 0 and pk["nz"] == 0
    perm = np.asarray(data.perm)
    x = np.random.default_rng(3).standard_normal(geo.Nb)
    assert np.abs(_emulate_packed(pk, data.size, perm, data.diag_jump[perm], x) - _matvec(data, x)).max() < 1e-12


def test_pack_fp32_near_values(c1):
    """FP32 near values: every row's entries and columns as with FP64 values,
    rounded once; rows padded (value 0, the row's own column) to 128-byte
    lines of values: 32 entries for FP32, 16 for FP64."""
    from smx.hmatrix.backend import pack_operator
    _, geo, bd = c1
    data = build_hmatrix_device(geo, HMatrixBuildConfig(leaf_size=32), ops=HostOps(geo, bd))
    fb = data.far_blocks
    pk = pack_operator(data.size, [b.row_start for b in fb], [b.row_stop for b in fb], [b.col_start for b in fb],
                       [b.col_stop for b in fb], [b.rank for b in fb], [b.U for b in fb], [b.V for b in fb],
                       data.near_csr, np.float64, np.float32)
    assert pk["near_val"].dtype == np.float32
    ref = _pack(data)
    cnt = np.diff(data.near_csr.indptr)
    for q, (P, R) in enumerate(((pk, 32), (ref, 16))):
        ptr = P["near_ptr"]
        assert np.all(ptr % R == 0) and np.all(np.diff(ptr) - cnt >= 0) and np.all(np.diff(ptr) - cnt < R)
    for row in range(data.size):
        a, b = int(pk["near_ptr"][row]), int(ref["near_ptr"][row])
        k = int(cnt[row])
        assert np.array_equal(pk["near_idx"][a:a + k], ref["near_idx"][b:b + k])
        assert np.array_equal(pk["near_val"][a:a + k], ref["near_val"][b:b + k].astype(np.float32))
        pad = slice(a + k, int(pk["near_ptr"][row + 1]))
        assert np.all(pk["near_val"][pad] == 0) and np.all(pk["near_idx"][pad] == row)


@pytest.mark.parametrize("compressor", ["svd", "rsvd"])
def test_svd_compressors_partition_matches_oracle(c1, compressor):
    _, geo, bd = c1
    eps = 1e-6
    cfg = HMatrixBuildConfig(epsilon=eps, leaf_size=32, compressor=compressor, batch_bytes=1 << 20)
    data = build_hmatrix_device(geo, cfg, ops=HostOps(geo, bd))
    if compressor == "svd":
        comp = "svd"
    else:
        comp = lambda A, rtol, mr, k: O.randomized_svd(A, rtol, mr, random_seed=O.block_random_seed(*k))
    ref = O.build_hmatrix(bd, eps=eps, leaf_size=32, compressor=comp)
    got = [(b.row_start, b.row_stop, b.col_start, b.col_stop, b.rank) for b in data.far_blocks]
    assert got == [(r0, r1, c0, c1, U.shape[1]) for r0, r1, c0, c1, U, V, _ in ref.far]
    _same_near(data.near_csr, ref.near)
    B = O.dense_B_rows(bd)
    x = np.random.default_rng(5).standard_normal(bd.Nb)
    assert np.linalg.norm(_matvec(data, x) - B @ x) / np.linalg.norm(B @ x) < 10 * eps


def test_config_compressor_names():
    assert HMatrixBuildConfig(compressor="randomized-svd").compressor == "rsvd"
    assert HMatrixBuildConfig(compressor="ACA").compressor == "fullaca"
    with pytest.raises(ValueError):
        HMatrixBuildConfig(compressor="qr")
    with pytest.raises(ValueError):
        HMatrixBuildConfig(gpu_memory_fraction=0.0)


def test_workspace_limit_subdivides_like_the_device_rule(c1):
    """A workspace limit: admissible blocks whose dense bytes x 3.5 exceed
    the available fraction split (counted), as in the oracle with the same
    predicate; the operator stays within tolerance."""
    _, geo, bd = c1
    avail = 100_000.0
    cfg = HMatrixBuildConfig(epsilon=1e-6, leaf_size=32, gpu_memory_fraction=0.5)
    data = build_hmatrix_device(geo, cfg, ops=HostOps(geo, bd, avail=avail))
    ref = O.build_hmatrix(bd, eps=1e-6, leaf_size=32, workspace_fits=lambda nb: nb * 3.5 <= avail * 0.5)
    assert data.metadata["gpu_workspace_subdivisions"] == ref.stats["gpu_workspace_subdivisions"] > 0
    assert data.metadata["forced_subdivisions_memory"] == ref.stats["forced_subdivisions_memory"]
    got = [(b.row_start, b.row_stop, b.col_start, b.col_stop, b.rank) for b in data.far_blocks]
    assert got == [(r0, r1, c0, c1, U.shape[1]) for r0, r1, c0, c1, U, V, _ in ref.far]
    assert max((f.row_stop - f.row_start) * (f.col_stop - f.col_start) for f in data.far_blocks) * 8 * 3.5 \
        <= avail * 0.5


@pytest.mark.parametrize("limit", [6 * 1024, 2 * 1024])
def test_oom_halves_batches_then_subdivides(c1, limit):
    """Allocations above 6 KiB (2 KiB) fail: batches are halved, single
    admissible blocks whose dense fill does not fit are subdivided (OOM
    count), leaf blocks that do not fit are filled in row slabs; the partition equals the
    oracle's with the same OOM predicate, and the operator stays within
    tolerance."""
    _, geo, bd = c1
    ops = HostOps(geo, bd, alloc_limit=limit)
    cfg = HMatrixBuildConfig(epsilon=1e-6, leaf_size=32)
    data = build_hmatrix_device(geo, cfg, ops=ops)
    assert ops.failed_allocs > 0
    ref = O.build_hmatrix(bd, eps=1e-6, leaf_size=32, oom=lambda m, n: m * n * 8 > limit)
    assert data.metadata["gpu_oom_subdivisions"] == ref.stats["gpu_oom_subdivisions"] > 0
    assert data.metadata["fallback_dense_blocks"] == ref.stats["fallback_dense_blocks"]
    assert data.metadata["oom_batch_splits"] > 0
    if limit < 19 * 19 * 8:             # smaller than the smallest leaf block
        assert data.metadata["slab_filled_blocks"] > 0 and data.metadata["fallback_dense_blocks"] > 0
    else:
        assert len(data.far_blocks) > 0
    assert data.metadata["cpu_fallback_blocks"] == 0
    got = [(b.row_start, b.row_stop, b.col_start, b.col_stop, b.rank) for b in data.far_blocks]
    assert got == [(r0, r1, c0, c1, U.shape[1]) for r0, r1, c0, c1, U, V, _ in ref.far]
    _same_near(data.near_csr, ref.near)
    B = O.dense_B_rows(bd)
    x = np.random.default_rng(6).standard_normal(bd.Nb)
    assert np.linalg.norm(_matvec(data, x) - B @ x) / np.linalg.norm(B @ x) < 1e-5


def test_unsplittable_without_fallback_raises(c1):
    _, geo, bd = c1
    cfg = HMatrixBuildConfig(epsilon=1e-6, leaf_size=32, max_temporary_block_bytes=1024, gpu_fallback_to_cpu=False)
    with pytest.raises(MemoryError):
        build_hmatrix_device(geo, cfg, ops=HostOps(geo, bd))


def _same_near(a, b):
    """Same sparsity pattern; values equal to the last bits (the stand-in sums
    a tile's triangle contributions in another order than the oracle)."""
    assert np.array_equal(a.indptr, b.indptr) and np.array_equal(a.indices, b.indices)
    assert np.abs(a.data - b.data).max() < 1e-15


def test_sixteen_row_task_table(c1):
    """The multi-RHS product's 16-row z tasks: in bounds of V and z, every z
    entry produced once, and the same z as the matvec's 8-row tasks."""
    from smx.hmatrix.backend import ZTASK_DTYPE, pack_operator
    _, geo, bd = c1
    d = build_hmatrix_device(geo, HMatrixBuildConfig(leaf_size=16, epsilon=1e-4), ops=HostOps(geo, bd))
    fb = d.far_blocks
    pk = pack_operator(d.size, [b.row_start for b in fb], [b.row_stop for b in fb], [b.col_start for b in fb],
                       [b.col_stop for b in fb], [b.rank for b in fb], [b.U for b in fb], [b.V for b in fb],
                       d.near_csr, np.float32, np.float64)
    V = pk["V"].astype(np.float64)
    xp = np.random.default_rng(0).standard_normal(d.size)[np.asarray(d.perm)]

    def zvec(tasks, cap):
        z = np.full(pk["nz"], np.nan)
        hits = np.zeros(pk["nz"], dtype=int)
        for t in tasks:
            n, kc, vo, zo, c0 = int(t["n"]), int(t["kc"]), int(t["v_off"]), int(t["z_off"]), int(t["c0"])
            if kc == 0:                  # padding of a short XCD group
                assert n == 0
                continue
            al = {4: 32, 8: 4}[pk["V"].dtype.itemsize]  # row padding by factor type
            ld = (n + al - 1) // al * al
            assert 1 <= kc <= cap and vo % al == 0 and vo + kc * ld <= V.size and zo + kc <= pk["nz"]
            assert 0 <= c0 and c0 + n <= d.size
            z[zo:zo + kc] = V[vo:vo + kc * ld].reshape(kc, ld)[:, :n] @ xp[c0:c0 + n]
            hits[zo:zo + kc] += 1
        assert np.all(hits == 1)
        return z

    z8 = zvec(pk["ztask"].view(ZTASK_DTYPE)[:pk["nztask"]], 8)
    assert pk["ztask_mm_rows"] == 16                # one MFMA row tile per task
    zmm = zvec(pk["ztask_mm"].view(ZTASK_DTYPE)[:pk["nztask_mm"]], 16)
    live8 = int(np.count_nonzero(pk["ztask"].view(ZTASK_DTYPE)[:pk["nztask"]]["kc"]))
    assert pk["nztask_mm"] < live8
    assert np.abs(z8 - zmm).max() <= 1e-12


@pytest.mark.parametrize("leaf", [16, 32])
def test_near_tiles_reproduce_near_csr(c1, leaf):
    """The multi-RHS product's dense near tiles: every row in exactly one tile,
    every index in bounds, the near matrix recovered entry for entry, and the
    tile products equal to the CSR product."""
    from smx.hmatrix.backend import NTILE_ROWS, leaf_intervals, near_from_tiles, pack_near_tiles
    _, geo, bd = c1
    d = build_hmatrix_device(geo, HMatrixBuildConfig(leaf_size=leaf), ops=HostOps(geo, bd))
    fb = d.far_blocks
    starts, _ = leaf_intervals(d.size, [b.row_start for b in fb], [b.row_stop for b in fb])
    pk = pack_near_tiles(d.size, d.near_csr, starts, np.float64)
    t, row, col, val = pk["tile"], pk["row"], pk["col"], pk["val"]
    R = row.reshape(-1, NTILE_ROWS)
    assert R.shape[0] == t.size == pk["ntile"]
    assert np.all((t["nrows"] >= 1) & (t["nrows"] <= NTILE_ROWS))
    for q in range(t.size):
        nr = int(t["nrows"][q])
        assert np.all(R[q, :nr] >= 0) and np.all(R[q, nr:] == -1)
        ld = (int(t["w"][q]) + 3) // 4 * 4
        assert t["a_off"][q] + nr * ld <= val.size and t["col_off"][q] + t["w"][q] <= col.size
        assert t["a_off"][q] % 4 == 0
    got = np.sort(R[R >= 0])
    assert np.array_equal(got, np.arange(d.size))
    assert np.all(np.diff(R.min(axis=1, where=R >= 0, initial=d.size)) > 0)   # stored in row order
    ref = d.near_csr.tocsr()
    back = near_from_tiles(d.size, t, row, col, val)
    assert (back != ref).nnz == 0
    assert pk["nnz"] == ref.nnz and pk["dense_entries"] >= ref.nnz
    X = np.random.default_rng(2).standard_normal((d.size, 5))
    Y = np.zeros_like(X)
    for q in range(t.size):
        nr, w, a0, c0 = int(t["nrows"][q]), int(t["w"][q]), int(t["a_off"][q]), int(t["col_off"][q])
        ld = (w + 3) // 4 * 4
        Y[R[q, :nr]] = val[a0:a0 + nr * ld].reshape(nr, ld)[:, :w] @ X[col[c0:c0 + w]]
    assert np.abs(Y - ref @ X).max() <= 1e-12 * max(1.0, np.abs(ref @ X).max())


@pytest.mark.parametrize("compressor", ["fullaca", "svd"])
def test_max_rank_zero_keeps_blocks_dense(c1, compressor):
    """max_rank = 0: every admissible block compresses to rank 0 and must stay
    in the operator as a dense rejection (builder.py:343-371), so the
    H-matrix equals the dense B exactly up to rounding of the near sums."""
    _, geo, bd = c1
    cfg = HMatrixBuildConfig(epsilon=1e-6, leaf_size=32, max_rank=0, compressor=compressor, batch_bytes=1 << 20)
    data = build_hmatrix_device(geo, cfg, ops=HostOps(geo, bd))
    ref = O.build_hmatrix(bd, eps=1e-6, leaf_size=32, max_rank=0, compressor=compressor)
    assert not data.far_blocks and not ref.far
    assert data.metadata["rejected_low_rank_blocks"] == ref.stats["rejected_low_rank_blocks"] > 0
    assert np.array_equal(data.near_csr.indices, ref.near.indices)
    B = O.dense_B_rows(bd)
    rng = np.random.default_rng(1234)
    x = rng.standard_normal(bd.Nb)
    assert np.linalg.norm(_matvec(data, x) - B @ x) <= 1e-13 * np.linalg.norm(B @ x)


def test_build_config_rejects_bad_values():
    for kw in (dict(epsilon=0.0), dict(eta=-1.0), dict(leaf_size=0), dict(max_rank=-1),
               dict(low_rank_storage_factor=1.5), dict(compressor="lu"), dict(build_backend="cpu")):
        with pytest.raises(ValueError):
            HMatrixBuildConfig(**kw)
    assert HMatrixBuildConfig(compressor="ACA").compressor == "fullaca"
    assert HMatrixBuildConfig(compressor="randomized-svd").compressor == "rsvd"
    assert HMatrixBuildConfig(max_rank=0).max_rank == 0


@pytest.mark.parametrize("xcd_order", [True, False])
@pytest.mark.parametrize("min_tasks", [1024, 8])
def test_near_task_tables(c1, xcd_order, min_tasks):
    """Near tasks of the LDS near launch: every row in one task, each task
    inside one leaf interval, its staged columns sorted and distinct, and
    every entry's 2-byte slot naming its column; a host replay of the
    launch (staged x, rows by slot) gives the near product."""
    from smx.hmatrix.backend import NTASK_DTYPE, leaf_intervals, pack_near_tasks
    _, geo, bd = c1
    data = build_hmatrix_device(geo, HMatrixBuildConfig(leaf_size=32), ops=HostOps(geo, bd))
    pk = _pack(data, xcd_order=xcd_order)
    fb = data.far_blocks
    starts, stops = leaf_intervals(data.size, [b.row_start for b in fb], [b.row_stop for b in fb])
    nt = pack_near_tasks(pk["near_ptr"], pk["near_idx"], starts, stops, xcd_order, min_tasks=min_tasks)
    ptr, idx, val = pk["near_ptr"], pk["near_idx"].astype(np.int64), pk["near_val"].astype(np.float64)
    tasks = nt["near_task"].view(NTASK_DTYPE)[:nt["n_near_task"]]
    stage, lidx = nt["near_stage"], nt["near_lidx"].view(np.uint16).astype(np.int64)
    x = np.random.default_rng(3).standard_normal(data.size)
    y = np.full(data.size, np.nan)
    seen = np.zeros(data.size, dtype=int)
    for t in tasks:
        p0, p1, w, so = int(t["p0"]), int(t["p1"]), int(t["w"]), int(t["s_off"])
        if p1 == p0:
            assert w == 0
            continue
        L = np.searchsorted(starts, p0, side="right") - 1
        assert starts[L] <= p0 < p1 <= stops[L] and p1 - p0 <= 64
        cols = stage[so:so + w]
        assert w <= nt["near_w_max"] and np.all(np.diff(cols) > 0)
        xs = x[cols]
        for p in range(p0, p1):
            e = np.arange(ptr[p], ptr[p + 1])
            assert np.array_equal(cols[lidx[e]], idx[e])
            y[p] = float(val[e] @ xs[lidx[e]])
            seen[p] += 1
    assert np.all(seen == 1)
    ref = data.near_csr @ x
    assert np.abs(y - ref).max() <= 1e-12 * max(np.abs(ref).max(), 1.0)
