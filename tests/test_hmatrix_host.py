"""H-matrix host logic on the CPU.

The builder is driven through a host stand-in of the two device operations
(batched Lindholm fill, batched full-pivot ACA).  The stand-in checks every
task of the tables the builder produces against the buffer sizes the kernels
would index, and evaluates the tasks with the oracle's Lindholm weights.  The
resulting partition must equal the oracle's depth-first build.
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
    def __init__(self, geo, bd):
        self.geo = geo
        self.bd = bd
        self.inc = geo.incidence()
        self.evaluations = 0

    def alloc(self, n):
        return np.full(max(int(n), 1), np.nan)

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


def _emulate_packed_matvec(P, x):
    """Index-for-index host emulation of the three matvec kernels on the packed
    tables (bounds checked), FP64 factors assumed."""
    t = lambda a: a.cpu().numpy()
    nb = P.nb
    perm, dp = t(P.perm), t(P.diag_p)
    U, V = t(P.U).astype(np.float64), t(P.V).astype(np.float64)
    f = {k: t(v) for k, v in P.far.items()}
    lf = {k: t(v) for k, v in P.leaf.items()}
    xp = x[perm]
    z = np.zeros(max(int(f["rank"].sum()), 1))
    for b in range(P.nfar):
        c0, n, r = int(f["c0"][b]), int(f["n"][b]), int(f["rank"][b])
        vo, zo = int(f["v_off"][b]), int(f["z_off"][b])
        assert c0 + n <= nb and vo + r * n <= V.size and zo + r <= z.size
        z[zo:zo + r] = V[vo:vo + r * n].reshape(r, n) @ xp[c0:c0 + n]
    y = np.full(nb, np.nan)
    nptr, nidx, nval = t(P.near_ptr), t(P.near_idx), t(P.near_val)
    covered = np.zeros(nb, dtype=int)
    for L in range(P.nleaf):
        l0, l1 = int(lf["start"][L]), int(lf["stop"][L])
        assert 0 < l1 - l0 <= 64
        covered[l0:l1] += 1
        acc = np.zeros(l1 - l0)
        for e in range(int(lf["fe_ptr"][L]), int(lf["fe_ptr"][L + 1])):
            b = int(lf["fe_block"][e])
            o = int(lf["fe_row"][e])
            m, r, uo, zo = int(f["m"][b]), int(f["rank"][b]), int(f["u_off"][b]), int(f["z_off"][b])
            assert 0 <= o and o + (l1 - l0) <= m and uo + r * m <= U.size
            Ub = U[uo:uo + r * m].reshape(r, m).T
            acc += Ub[o:o + (l1 - l0)] @ z[zo:zo + r]
        for i in range(l0, l1):
            s = sum(nval[e] * xp[nidx[e]] for e in range(nptr[i], nptr[i + 1]))
            y[perm[i]] = dp[i] * xp[i] + s + acc[i - l0]
    assert np.all(covered == 1)
    return y


def test_packed_tables_and_matvec(c1):
    import torch
    from smx.hmatrix import HMatrixGPUPackedFused
    _, geo, bd = c1
    data = build_hmatrix_device(geo, HMatrixBuildConfig(leaf_size=32), ops=HostOps(geo, bd))
    P = HMatrixGPUPackedFused(data, use_fp32=False, device=torch.device("cpu"), boundary_gather=geo.bverts)
    x = np.random.default_rng(0).standard_normal(geo.Nb)
    assert np.abs(_emulate_packed_matvec(P, x) - _matvec(data, x)).max() < 1e-12
    assert P.far_storage_entries == data.far_storage_entries


def test_packed_transpose_tables(c1):
    """The transposed view (factors swapped, column leaves, near^T) reproduces
    B^T x under the same index emulation."""
    import torch
    from smx.hmatrix import HMatrixGPUPackedFused
    _, geo, bd = c1
    data = build_hmatrix_device(geo, HMatrixBuildConfig(leaf_size=32), ops=HostOps(geo, bd))
    P = HMatrixGPUPackedFused(data, use_fp32=False, device=torch.device("cpu"))
    P._struct(P.gidx_plain, transpose=True)
    tt = P._tt

    class View:
        pass
    v = View()
    v.nb, v.perm, v.diag_p, v.nfar = P.nb, P.perm, P.diag_p, P.nfar
    v.U, v.V = P.V, P.U
    v.far, v.leaf = tt["far"], tt["leaf"]
    v.nleaf = int(tt["leaf"]["start"].numel())
    v.near_ptr, v.near_idx, v.near_val = tt["near_ptr"], tt["near_idx"], tt["near_val"]
    x = np.random.default_rng(1).standard_normal(geo.Nb)
    xp = x[data.perm]
    yp = data.diag_jump[data.perm] * xp + data.near_csr.T @ xp
    for b in data.far_blocks:
        yp[b.col_start:b.col_stop] += b.V.T @ (b.U.T @ xp[b.row_start:b.row_stop])
    ref = np.empty_like(x)
    ref[data.perm] = yp
    assert np.abs(_emulate_packed_matvec(v, x) - ref).max() < 1e-12
