"""Pins of the oracle's discrete H-matrix rules, each by a formulation that a
plausible slip in the restatement would fail.

* Cluster tree (tree.py:44-99).  The rule "split on the first axis of largest
  bbox extent, stable mergesort of that coordinate, left = first n//2" is
  equivalent to a closed form: the points entering a node at depth d are in
  lexicographic order of (x_{a_{d-1}}, ..., x_{a_0}, original index), a_k the
  split axes of its ancestors (each stable sort keeps the previous order
  among ties).  The test checks every internal node against np.lexsort on that
  key chain, on Kuhn-grid boundaries where coordinate ties straddle the cut
  and where bbox extents tie between axes.  Planted mistakes in a copy of the
  oracle (unstable sort, last-argmax axis, ceil(n/2) cut) must fail it.
* Full-pivot ACA (compressors.py:35-99).  From the returned factors alone:
  R_k = A - U_k V_k; pivot k is the first row-major argmax |R_k|, so
  V[k, j_k] = 1 and U[:, k] = R_k[:, j_k]; the rank is the first k with
  ||R_k||_F <= eps ||A||_F, or the step whose pivot is <= u ||A||_F.
  Exact-arithmetic cases with tied maxima pin "first occurrence".
* Acceptance (builder.py:331-376): blocks whose compressor reports values just
  on either side of each threshold: rank > 0, rel <= eps (1 + 64 u),
  (m + n) r < 0.95 m n.  The expected fate is written from the reference's
  rule, and the product's builder (driven by the host stand-in) must make the
  same decisions as the oracle.
"""
import math

import numpy as np
import pytest

import smx_oracle as O
from smx import mesh as M

U64 = np.finfo(np.float64).eps


def _surface_points(nx, ny, nz, lx, ly, lz):
    """Boundary nodes of a Kuhn box in global node order (= bverts order)."""
    m = M.kuhn_box(nx, ny, nz, lx, ly, lz)
    X = m.X
    lo, hi = X.min(axis=0), X.max(axis=0)
    on = np.any((X == lo) | (X == hi), axis=1)
    return X[on]


def _check_tree(points, perm, root, leaf):
    """Every node against the lexsort closed form; returns counts of the
    nodes where the rule had teeth (ties across the cut, tied extents)."""
    points = np.asarray(points)
    n = points.shape[0]
    assert np.array_equal(np.sort(perm), np.arange(n))
    tie_cuts = tie_axes = 0
    stack = [(root, [])]
    while stack:
        node, axes = stack.pop()
        ids = perm[node.start:node.stop]
        P = points[ids]
        assert np.array_equal(P.min(axis=0), node.bmin) and np.array_equal(P.max(axis=0), node.bmax)
        if node.is_leaf:
            assert 1 <= node.size <= leaf
            continue
        assert node.size > leaf
        ext = P.max(axis=0) - P.min(axis=0)
        a = min(i for i in range(3) if ext[i] == ext.max())
        tie_axes += int(np.sum(ext == ext.max()) > 1)
        # incoming order: keys for np.lexsort, the last key is the primary one
        inc = ids[np.lexsort([ids] + [points[ids, ax] for ax in axes])]
        ranked = inc[np.lexsort((np.arange(inc.size), points[inc, a]))]
        cut = node.size // 2
        left = node.left
        assert left.start == node.start and left.stop == node.start + cut
        assert set(perm[left.start:left.stop].tolist()) == set(ranked[:cut].tolist())
        if points[ranked[cut - 1], a] == points[ranked[cut], a]:
            tie_cuts += 1
        stack.append((node.right, axes + [a]))
        stack.append((node.left, axes + [a]))
    return tie_cuts, tie_axes


GRIDS = [((10, 10, 10, 40.0, 40.0, 40.0), 16), ((10, 10, 10, 40.0, 40.0, 40.0), 64),
         ((30, 12, 3, 75.0, 30.0, 7.5), 32), ((40, 40, 2, 100.0, 100.0, 5.0), 64)]


@pytest.mark.parametrize("grid,leaf", GRIDS)
def test_tree_closed_form(grid, leaf):
    pts = _surface_points(*grid)
    perm, root = O.cluster_tree(pts, leaf)
    tie_cuts, tie_axes = _check_tree(pts, perm, root, leaf)
    # the grids exercise the tie rules, so a slip in them cannot pass unseen
    assert tie_cuts > 0
    if grid[0] == grid[1] == grid[2]:
        assert tie_axes > 0


def test_tree_closed_form_film_boundary():
    """The 500 x 500 x 10 nm film's 83 202 boundary nodes, leaf 64."""
    pts = _surface_points(200, 200, 4, 500.0, 500.0, 10.0)
    assert pts.shape[0] == 83202
    perm, root = O.cluster_tree(pts, 64)
    tie_cuts, tie_axes = _check_tree(pts, perm, root, 64)
    assert tie_cuts > 100 and tie_axes > 0


def _mutant_tree(points, leaf, kind="mergesort", axis_rule="first", cut_rule="floor"):
    """A copy of the bisection with one planted mistake (CPU only)."""
    order = []

    def rec(idx):
        lp = points[idx]
        bmin, bmax = lp.min(axis=0), lp.max(axis=0)
        if idx.size <= leaf:
            s = len(order)
            order.extend(idx.tolist())
            return O.Cluster(s, len(order), bmin, bmax)
        ext = bmax - bmin
        ax = int(np.argmax(ext)) if axis_rule == "first" else int(2 - np.argmax(ext[::-1]))
        srt = idx[np.argsort(lp[:, ax], kind=kind)]
        mid = srt.size // 2 if cut_rule == "floor" else (srt.size + 1) // 2
        L, R = rec(srt[:mid]), rec(srt[mid:])
        return O.Cluster(L.start, R.stop, bmin, bmax, L, R)

    root = rec(np.arange(points.shape[0], dtype=np.int32))
    return np.asarray(order, dtype=np.int32), root


@pytest.mark.parametrize("mutation", [dict(kind="quicksort"), dict(axis_rule="last"), dict(cut_rule="ceil")])
def test_tree_check_catches_planted_mistakes(mutation):
    pts = _surface_points(10, 10, 10, 40.0, 40.0, 40.0)
    # 601 points (one removed) so that some cuts have odd sizes
    pts = pts[:-1]
    perm, root = _mutant_tree(pts, 16, **mutation)
    with pytest.raises(AssertionError):
        _check_tree(pts, perm, root, 16)
    perm_o, root_o = O.cluster_tree(pts, 16)
    _check_tree(pts, perm_o, root_o, 16)


# ---------------------------------------------------------------- ACA
def _check_aca(A, U, V, rel, eps, max_rank=None):
    """The pivot and stop rules from the factors; returns steps checked."""
    m, n = A.shape
    r = U.shape[1]
    normA = np.linalg.norm(A)
    checked = 0
    for k in range(r):
        Rk = A - U[:, :k] @ V[:k]
        flat = np.abs(Rk).ravel()
        top = np.partition(flat, -2)[-2:] if flat.size > 1 else flat
        i, j = divmod(int(np.argmax(flat)), n)
        if flat.size > 1 and top[1] - top[0] <= 1e-9 * top[1]:
            continue          # near tie up to the rounding of R_k: not decidable from the factors
        assert V[k, j] == pytest.approx(1.0, abs=1e-12)
        assert np.allclose(U[:, k], Rk[:, j], rtol=0, atol=1e-10 * normA)
        assert abs(U[i, k]) == pytest.approx(flat.max(), rel=1e-9)
        if k > 0:
            assert np.linalg.norm(Rk) > eps * normA * (1 - 1e-9)
        checked += 1
    Rr = A - U @ V
    stop_tol = np.linalg.norm(Rr) <= eps * normA * (1 + 1e-9)
    stop_piv = np.abs(Rr).max() <= U64 * normA * (1 + 1e-6)
    stop_cap = max_rank is not None and r == max_rank
    assert stop_tol or stop_piv or stop_cap or r == min(m, n)
    assert rel == pytest.approx(np.linalg.norm(Rr) / normA, rel=1e-6, abs=1e-15)
    return checked


def _lindholm_blocks():
    """Admissible blocks of the 40 nm cube's B (real operator entries)."""
    m = M.cube_c1()
    bd = O.boundary_data(m.X, m.tets)
    perm, root = O.cluster_tree(bd.Xb, 32)
    leaves = []
    st = [root]
    while st:
        c = st.pop()
        if c.is_leaf:
            leaves.append(c)
        else:
            st += [c.left, c.right]
    out = []
    for a in leaves:
        for b in leaves:
            if O.admissible(a, b, 2.0) and len(out) < 12:
                out.append(O.block_B(bd, perm[a.start:a.stop], perm[b.start:b.stop]))
    return out


def test_aca_pivot_and_stop_rule_on_lindholm_blocks():
    total = 0
    for A in _lindholm_blocks():
        for eps in (1e-2, 1e-6, 1e-10):
            U, V, rel = O.fullaca_dense(A, eps)
            total += _check_aca(A, U, V, rel, eps)
    assert total > 50


def test_aca_pivot_first_occurrence_on_ties():
    # |5| appears at (0, 2) and (1, 0): row-major first is (0, 2)
    A = np.array([[1.0, 2.0, 5.0], [-5.0, 1.0, 1.0], [2.0, 3.0, 1.0]])
    U, V, rel = O.fullaca_dense(A, 1e-30, max_rank=1)
    assert np.array_equal(U[:, 0], A[:, 2]) and np.array_equal(V[0], A[0] / 5.0)
    # equal magnitudes with opposite signs, the negative one first
    A = np.array([[0.0, -4.0], [4.0, 1.0]])
    U, V, _ = O.fullaca_dense(A, 1e-30, max_rank=1)
    assert np.array_equal(U[:, 0], A[:, 1]) and np.array_equal(V[0], A[0] / -4.0)


def test_aca_rank_on_exact_low_rank_and_caps():
    u = np.array([[1.0, 2.0], [3.0, -1.0], [0.0, 4.0], [2.0, 2.0], [1.0, 0.0]])
    v = np.array([[1.0, 0.0, 2.0, -1.0], [0.0, 3.0, 1.0, 1.0]])
    A = u @ v                      # rank 2
    U, V, rel = O.fullaca_dense(A, 1e-30)
    assert U.shape[1] == 2 and rel <= 1e-15
    U, V, rel = O.fullaca_dense(A, 1e-30, max_rank=1)
    assert U.shape[1] == 1 and rel > 0.01
    # tolerance met after the first step: the rank is 1, not 2
    A1 = np.outer([1.0, 2.0, 3.0], [1.0, 1.0, 2.0]) + 1e-9 * np.outer([1.0, -1.0, 0.5], [0.3, 1.0, -2.0])
    U, V, rel = O.fullaca_dense(A1, 1e-6)
    assert U.shape[1] == 1 and rel <= 1e-6
    # zero block and max_rank 0: rank 0, residual reported as 0
    for Z, mr in ((np.zeros((3, 4)), None), (A, 0)):
        U, V, rel = O.fullaca_dense(Z, 1e-6, max_rank=mr)
        assert U.shape == (Z.shape[0], 0) and V.shape == (0, Z.shape[1]) and rel == 0.0


def test_aca_check_catches_a_late_stop():
    """One step past the tolerance (the planted mistake) fails the stop rule."""
    A = _lindholm_blocks()[0]
    eps = 1e-6
    U, V, rel = O.fullaca_dense(A, eps)
    r = U.shape[1]
    U2, V2, rel2 = O.fullaca_dense(A, eps * 1e-3, max_rank=r + 1)
    assert U2.shape[1] == r + 1
    with pytest.raises(AssertionError):
        # the extra step's residual check: R_r already met eps
        _check_aca(A, U2, V2, rel2, eps)


# ---------------------------------------------------------------- acceptance
MARGIN = 1.0 + 64.0 * U64


def _scenario(key):
    """Deterministic choice of the compressor's reported outcome per block."""
    r0, r1, c0, c1 = key
    return (r0 * 7 + c0 * 13 + (r1 - r0) * 3 + (c1 - c0)) % 6


def _crafted(A_shape, key, eps):
    m, n = A_shape
    sc = _scenario(key)
    ok_rank = int(math.ceil(0.95 * m * n / (m + n))) - 1     # largest r with (m+n) r < 0.95 m n
    while (m + n) * (ok_rank + 1) < 0.95 * m * n:
        ok_rank += 1
    while ok_rank > 0 and not (m + n) * ok_rank < 0.95 * m * n:
        ok_rank -= 1
    if sc == 0:
        return 1, eps * MARGIN, "far"                           # exactly on the tolerance: accepted
    if sc == 1:
        return 1, np.nextafter(eps * MARGIN, np.inf), "fail"     # just above: fails the tolerance
    if sc == 2:
        return 0, 0.0, "dense"                                  # rank 0: kept dense
    if sc == 3:
        return max(ok_rank, 1), eps * 0.5, ("far" if ok_rank >= 1 else "dense")   # storage just below 0.95 m n
    if sc == 4:
        return ok_rank + 1, eps * 0.5, "dense"                  # storage at/above 0.95 m n: dense rejection
    return 2, eps * (1.0 + 32.0 * U64), "far"                   # inside the rounding margin: accepted


def test_acceptance_thresholds_oracle_and_builder():
    from smx.boundary import build_boundary_geometry
    from smx.hmatrix import HMatrixBuildConfig, build_hmatrix_device
    from test_hmatrix_host import HostOps

    m = M.cube_c1()
    bd = O.boundary_data(m.X, m.tets)
    geo = build_boundary_geometry(m.X, m.tets)
    eps = 1e-6
    seen = {}

    def comp_oracle(A, rtol, max_rank, key):
        rk, rel, fate = _crafted(A.shape, key, eps)
        seen[key] = (fate, A.shape)
        return np.zeros((A.shape[0], rk)), np.zeros((rk, A.shape[1])), rel

    ref = O.build_hmatrix(bd, eps=eps, eta=2.0, leaf_size=32, compressor=comp_oracle)
    far = {(f[0], f[1], f[2], f[3]) for f in ref.far}
    n_fail = 0
    for key, (fate, shape) in seen.items():
        if fate == "far":
            assert key in far, (key, fate)
        else:
            assert key not in far, (key, fate)
        n_fail += fate == "fail"
    assert {f for f, _ in seen.values()} >= {"far", "fail", "dense"} and n_fail > 0
    assert ref.stats["compression_tolerance_subdivisions"] > 0 and ref.stats["rejected_low_rank_blocks"] > 0

    class CraftedOps(HostOps):
        """Host stand-in whose compressor reports the crafted outcome of each
        block; the block of each task comes from the builder's per-block seed
        call, which precedes every compression batch."""
        keys = []

        def compress(self, method, A, tasks, U, V, rtol, max_rank, seeds, cfg):
            nt = tasks.size
            keys, self.keys[:] = list(self.keys[:nt]), self.keys[nt:]
            assert len(keys) == nt
            ranks, rels, status = np.zeros(nt, np.int32), np.zeros(nt), np.zeros(nt, np.int32)
            for q, t in enumerate(tasks):
                mm, nn = int(t["m"]), int(t["n"])
                assert (keys[q][1] - keys[q][0], keys[q][3] - keys[q][2]) == (mm, nn)
                rk, rel, _ = _crafted((mm, nn), keys[q], eps)
                ranks[q], rels[q] = rk, rel
                # status 3 is reserved for an all-zero block; these blocks are not
                status[q] = 2 if rk > int(t["rcap"]) else 0
                if status[q] == 0:
                    u0, v0 = int(t["u_off"]), int(t["v_off"])
                    U[u0:u0 + rk * mm] = 0.0
                    V[v0:v0 + rk * nn] = 0.0
            return ranks, rels, status

    ops = CraftedOps(geo, bd)
    from smx.hmatrix import compressors as Bmod
    orig = Bmod.block_random_seed

    def seed_and_record(r0, r1, c0, c1):
        ops.keys.append((r0, r1, c0, c1))
        return orig(r0, r1, c0, c1)

    Bmod.block_random_seed = seed_and_record
    try:
        # compressor "svd": the crafted outcomes are taken as reported (the
        # full-pivot ACA path adds the kernel's own rank-cap status)
        cfg = HMatrixBuildConfig(epsilon=eps, leaf_size=32, compressor="svd", batch_bytes=1 << 14)
        data = build_hmatrix_device(geo, cfg, ops=ops)
    finally:
        Bmod.block_random_seed = orig
    got = [(b.row_start, b.row_stop, b.col_start, b.col_stop, b.rank) for b in data.far_blocks]
    exp = [(f[0], f[1], f[2], f[3], f[4].shape[1]) for f in ref.far]
    assert got == exp
    for k in ("low_rank_blocks", "dense_blocks", "rejected_low_rank_blocks", "compression_tolerance_subdivisions"):
        assert data.metadata[k] == ref.stats[k], k
    assert np.array_equal(data.near_csr.indptr, ref.near.indptr)
    assert np.array_equal(data.near_csr.indices, ref.near.indices)
