"""The 500 x 500 x 10 nm film (N = 202 005, Nb = 83 202): the ACA H-matrix
workload, on the device, against the oracle.

* Tree and partition: the device build's permutation equals the oracle's
  cluster tree; its block tree (every visited block, depth-first) equals the
  oracle's geometric walk (eta admissibility, the 256 MiB forced split, the
  row/column subdivision rule, dense inadmissible leaf pairs), given the
  compressor's outcome on each admissible block.
* Those outcomes: a seeded sample of admissible blocks is filled by the
  oracle's Lindholm weights and compressed by its full-pivot ACA; the rank and
  the accept / split / dense decision must match (0 acceptance flips), and the
  device factors must meet eps on the oracle's block.
* Matvec: the validate_against_dense protocol (integration.py:180-205; seed
  1234, 3 trials, x ~ N(0, 1)) on a seeded sample of 64 rows of the oracle's
  dense FP64 B (83 202 columns each): relative error <= 10 eps for the FP64
  factors, <= 10 eps + 1e-6 for the FP32 factors; B 1 = -1 through the
  H-matrix.
* Demag (FP64 factors): the uniform-m patch test (u1 = Ms z + const, so the
  u1 part of H is -Ms m) and constant-shift invariance of u1 + u2.
* Mixed (FP32 factors, FP32 inner CG with one FP64 refinement restart, FP32
  AMG values) against FP64 on the film material: H_demag and the LLG RHS
  within 1e-6 relative L2.
"""
import math

import numpy as np
import pytest

import smx_oracle as O

pytestmark = pytest.mark.gpu

EPS = 1e-6
MARGIN = 1.0 + 64.0 * np.finfo(np.float64).eps


@pytest.fixture(scope="module")
def film():
    import time
    import torch
    from smx import mesh as M
    from smx.boundary import build_boundary_geometry
    from smx.demag import OperatorSet
    from smx.hmatrix import HMatrixBuildConfig, build_hmatrix_device
    dev = torch.device("cuda")
    m = M.film_c3()
    t0 = time.perf_counter()
    geo = build_boundary_geometry(m.X, m.tets)
    data = build_hmatrix_device(geo, HMatrixBuildConfig(epsilon=EPS, eta=2.0, leaf_size=64), dev)
    t1 = time.perf_counter()
    rows = np.sort(np.random.default_rng(1234).choice(geo.Nb, 64, replace=False))
    bd = O.boundary_data_vec(m.X, m.tets, omega_rows=rows)
    Brows = O.dense_B_rows(bd, rows, chunk=8)
    ops = OperatorSet(m, dev)
    md = data.metadata
    print(f"[film] N={m.num_nodes} Nb={geo.Nb} build {t1 - t0:.1f}s far={md['low_rank_blocks']} "
          f"dense={md['dense_blocks']} rejected={md['rejected_low_rank_blocks']} "
          f"tol_splits={md['compression_tolerance_subdivisions']} near_nnz={md['near_nnz']}; "
          f"oracle rows {time.perf_counter() - t1:.1f}s", flush=True)
    assert geo.Nb == 83202 and m.num_nodes == 202005
    return dict(mesh=m, geo=geo, data=data, rows=rows, bd=bd, Brows=Brows, ops=ops, dev=dev)


def _oracle_walk(bd, fate_of, eta=2.0, leaf=64, max_bytes=256 * 1024 ** 2):
    """Depth-first block walk of builder.py:256-376 over the oracle's tree;
    admissible blocks take the compressor outcome fate_of[(r0, r1, c0, c1)]
    (0 far, 1 dense, 'split'), whose split direction is the oracle's."""
    perm, root = O.cluster_tree(bd.Xb, leaf)
    out = []

    def split_code(r, c):
        rs, cs = not r.is_leaf, not c.is_leaf
        return 2 if rs and (not cs or r.diameter >= c.diameter) else 3

    def children(r, c, code):
        return [(r.left, c), (r.right, c)] if code == 2 else [(r, c.left), (r, c.right)]

    stack = [(root, root)]
    while stack:
        r, c = stack.pop()
        k = (r.start, r.stop, c.start, c.stop)
        adm = O.admissible(r, c, eta)
        if adm and r.size * c.size * 8 > max_bytes:
            code = split_code(r, c)
        elif adm:
            f = fate_of.get(k)
            assert f is not None, f"admissible block {k} missing from the device build"
            code = split_code(r, c) if f == "split" else f
            if f == "split":
                assert not (r.is_leaf and c.is_leaf)
        elif r.is_leaf and c.is_leaf:
            code = 1
        else:
            code = split_code(r, c)
        out.append((*k, code))
        if code >= 2:
            ch = children(r, c, code)
            stack.append(ch[1])
            stack.append(ch[0])
    return perm, np.asarray(out, dtype=np.int64)


def test_film_tree_and_partition(film):
    data, bd = film["data"], film["bd"]
    bt = data.block_tree
    fate = {}
    for r0, r1, c0, c1, code in bt.tolist():
        fate[(r0, r1, c0, c1)] = "split" if code >= 2 else code
    perm_o, walk = _oracle_walk(bd, fate)
    assert np.array_equal(data.perm, perm_o)
    assert np.array_equal(film["geo"].bverts, bd.bverts)
    assert walk.shape == bt.shape and np.array_equal(walk, bt)
    n_far = int((bt[:, 4] == 0).sum())
    assert n_far == len(data.far_blocks) == data.metadata["low_rank_blocks"]
    print(f"[film partition] {bt.shape[0]} blocks visited, {n_far} far, "
          f"{int((bt[:, 4] == 1).sum())} dense: identical to the oracle's walk", flush=True)


def test_film_aca_sample_no_flips(film):
    """Oracle fill + full-pivot ACA on a seeded sample of the admissible
    blocks (up to 2.5e5 entries each): same decision and rank as the device
    build; the device factors meet eps on the oracle's block."""
    data, bd = film["data"], film["bd"]
    bt = data.block_tree
    far = {(b.row_start, b.row_stop, b.col_start, b.col_stop): b for b in data.far_blocks}
    _, root = O.cluster_tree(bd.Xb, 64)
    nodes = {}
    st = [root]
    while st:
        c = st.pop()
        nodes[(c.start, c.stop)] = c
        if not c.is_leaf:
            st += [c.left, c.right]
    adm = []
    for r0, r1, c0, c1, code in bt.tolist():
        r, c = nodes[(r0, r1)], nodes[(c0, c1)]
        if O.admissible(r, c, 2.0) and (r1 - r0) * (c1 - c0) * 8 <= 256 * 1024 ** 2:
            adm.append((r0, r1, c0, c1, code))
    rng = np.random.default_rng(7)
    small = [a for a in adm if (a[1] - a[0]) * (a[3] - a[2]) <= 250_000]
    pick = [small[i] for i in rng.choice(len(small), min(24, len(small)), replace=False)]
    pick += [a for a in adm if a[4] != 0][:8]           # every non-far outcome, up to 8
    flips = rank_diff = 0
    for r0, r1, c0, c1, code in pick:
        A = O.block_B(bd, data.perm[r0:r1], data.perm[c0:c1])
        U, V, rel = O.fullaca_dense(A, EPS)
        mm, nn = A.shape
        k = U.shape[1]
        meets = rel <= EPS * MARGIN
        if k > 0 and meets and (mm + nn) * k < 0.95 * mm * nn:
            exp = 0
        elif not meets and not (nodes[(r0, r1)].is_leaf and nodes[(c0, c1)].is_leaf):
            exp = 2
        else:
            exp = 1
        got = 0 if code == 0 else (1 if code == 1 else 2)
        flips += exp != got
        if code == 0:
            b = far[(r0, r1, c0, c1)]
            rank_diff += b.rank != k
            assert np.linalg.norm(b.U @ b.V - A) <= EPS * (1 + 1e-9) * np.linalg.norm(A)
    print(f"[film ACA sample] {len(pick)} blocks: {flips} acceptance flips, {rank_diff} rank differences",
          flush=True)
    assert flips == 0 and rank_diff == 0


@pytest.mark.parametrize("fp32", [False, True])
def test_film_matvec_vs_oracle_rows(film, fp32):
    import torch
    from smx.hmatrix import HMatrixGPUPackedFused
    data, rows, Brows = film["data"], film["rows"], film["Brows"]
    H = HMatrixGPUPackedFused(data, use_fp32=fp32, device=film["dev"])
    rng = np.random.default_rng(1234)
    worst = 0.0
    for _ in range(3):
        x = rng.standard_normal(film["geo"].Nb)
        y = H.matvec(torch.from_numpy(x).to(film["dev"])).cpu().numpy()
        ref = Brows @ x
        worst = max(worst, float(np.linalg.norm(y[rows] - ref) / np.linalg.norm(ref)))
    one = H.matvec(torch.ones(film["geo"].Nb, dtype=torch.float64, device=film["dev"])).cpu().numpy()
    e_one = float(np.linalg.norm(one + 1.0) / math.sqrt(one.size))
    e_one_oracle = float(np.abs(Brows.sum(axis=1) + 1.0).max())
    print(f"[film matvec fp32={fp32}] 64 oracle rows: rel err {worst:.3e}; ||B 1 + 1||/||1|| {e_one:.3e} "
          f"(oracle rows {e_one_oracle:.1e})", flush=True)
    assert e_one_oracle <= 1e-12
    assert worst <= 10 * EPS + (1e-6 if fp32 else 0.0)
    assert e_one <= 10 * EPS + (1e-6 if fp32 else 0.0)
    del H
    torch.cuda.empty_cache()


def _demag(film, fp32, **kw):
    from smx.demag import DemagFieldLindholmGPU
    args = dict(ksp_rtol_u1=1e-6, ksp_rtol_u2=1e-6, ksp_atol=1e-12)
    args.update(kw)
    return DemagFieldLindholmGPU(film["mesh"], None, None, 8e5, boundary_backend="hmatrix",
                                 hmatrix_prebuilt=film["data"], hmatrix_use_fp32=fp32, geometry=film["geo"],
                                 operators=film["ops"], **args)


def test_film_demag_patch_and_shift(film):
    """FP64 factors, CG to 1e-12: uniform m = z gives u1 = Ms z + const (the
    P1 patch test, so the u1 part of H is -Ms m exactly); adding a constant
    to u1 before the boundary matvec leaves H unchanged (B 1 = -1 makes u2
    absorb it) to the H-matrix accuracy."""
    import torch
    m = film["mesh"]
    d = _demag(film, False, ksp_rtol_u1=1e-12, ksp_rtol_u2=1e-12, ksp_atol=0.0, ksp_max_it_u1=5000,
               ksp_max_it_u2=20000, cg_speculative=False)
    mv = torch.from_numpy(np.tile([0.0, 0.0, 1.0], m.num_nodes)).to(film["dev"])
    H = torch.empty_like(mv)
    d.compute_vec(mv, H)
    torch.cuda.synchronize()
    u1 = d.u1.cpu().numpy()
    z = m.X[:, 2]
    exp = 8e5 * (z - z.mean())
    e_patch = float(np.linalg.norm(u1 - u1.mean() - exp) / np.linalg.norm(exp))
    Hz = H.cpu().numpy().reshape(-1, 3)
    vol = d.ops.fem.volume
    avg = float((Hz[:, 2] * vol).sum() / vol.sum() / 8e5)
    # constant shift: g = B (u1 + c) then u2, against the unshifted u1 + u2
    rng = np.random.default_rng(3)
    g = rng.standard_normal((m.num_nodes, 3))
    mr = torch.from_numpy((g / np.linalg.norm(g, axis=1)[:, None]).ravel()).to(film["dev"])
    d.compute_vec(mr, H)
    torch.cuda.synchronize()
    phi0 = (d.u1 - d.u1.mean() + d.u2).clone()
    c = 10.0 * float(torch.sqrt(torch.mean(d.u1 * d.u1)))
    u1s = d.u1 - d.u1.mean() + c
    y = torch.empty(d.Nb, dtype=torch.float64, device=film["dev"])
    d.boundary_matvec(u1s, y, gather=d.bverts_gpu)
    u2s = torch.zeros_like(d.u2)
    d.solve("u2", 1, u2s, d.ws2, gb=y, project=0)
    torch.cuda.synchronize()
    phis = u1s + u2s
    e_shift = float(torch.linalg.norm(phis - phi0) / torch.linalg.norm(phi0))
    print(f"[film demag fp64] patch u1 {e_patch:.2e}; <Hz>/Ms {avg:.4f}; shift c = {c:.3e}: "
          f"|phi(c) - phi(0)|/|phi| {e_shift:.2e}", flush=True)
    assert e_patch <= 1e-9
    assert -1.0 < avg < -0.9          # thin film: demag factor close to 1
    assert e_shift <= 1e-3            # without B 1 = -1 the shift would move phi by O(1)


def test_film_mixed_vs_fp64(film):
    """The film workload's two modes with the same solver settings (rtol 1e-6,
    zero initial guess): H_demag and the LLG RHS within 1e-6."""
    import torch
    from smx.llg import EffectiveField
    m = film["mesh"]
    mat = dict(Ms=8e5, Aex=1.3e-11, alpha=0.5, Ku=5e5, n_ani_vec=(0.0, 0.0, 1.0), D_int=3e-3,
               n0_int_vec=(0.0, 0.0, 1.0))
    hk = dict(boundary_backend="hmatrix", geometry=film["geo"], hmatrix_prebuilt=film["data"],
              ksp_rtol_u1=1e-6, ksp_rtol_u2=1e-6, ksp_atol=1e-12)
    h64 = EffectiveField(m, use_demag=True, operators=film["ops"], demag_kwargs=dict(hmatrix_use_fp32=False, **hk),
                         **mat)
    hmx = EffectiveField(m, use_demag=True, operators=film["ops"],
                         demag_kwargs=dict(hmatrix_use_fp32=True, cg_mixed=True, **hk), **mat)
    g = np.random.default_rng(0).standard_normal((m.num_nodes, 3))
    mv = torch.from_numpy((g / np.linalg.norm(g, axis=1)[:, None]).ravel()).to(film["dev"])
    r64, rmx = torch.empty_like(mv), torch.empty_like(mv)
    errs = []
    for _ in range(2):
        h64.llg_rhs_vec(mv, r64)
        hmx.llg_rhs_vec(mv, rmx)
        torch.cuda.synchronize()
        Hd64 = torch.empty_like(mv)
        Hdmx = torch.empty_like(mv)
        h64.demag_field.compute_vec(mv, Hd64)
        hmx.demag_field.compute_vec(mv, Hdmx)
        torch.cuda.synchronize()
        errs.append((float(torch.linalg.norm(Hdmx - Hd64) / torch.linalg.norm(Hd64)),
                     float(torch.linalg.norm(rmx - r64) / torch.linalg.norm(r64))))
    print(f"[film mixed vs fp64] (H_demag, RHS) rel err {errs}; its u1/u2 fp64 "
          f"{h64.demag_field.last_its_u1}/{h64.demag_field.last_its_u2} mixed "
          f"{hmx.demag_field.last_its_u1}/{hmx.demag_field.last_its_u2}", flush=True)
    assert all(eh <= 1e-6 and er <= 1e-6 for eh, er in errs)


def test_film_fill_deterministic(film):
    """A 1024 x 1024 film block whose relevant triangles share vertices
    (lanes of one wavefront add into the same columns): two fills are
    bitwise equal, and equal to the oracle's block within 1e-13 of its max."""
    import torch
    from smx import _capi
    from smx.device import to_dev
    geo, bd, data, dev = film["geo"], film["bd"], film["data"], film["dev"]
    rows = data.perm[:1024].astype(np.int64)
    cols = data.perm[512:1536].astype(np.int64)
    ref = O.block_B(bd, rows, cols)
    rel, rel_cols = geo.relevant_triangles(cols)
    L = _capi.lib()
    Xb_d, rows_d = to_dev(geo.Xb, dev), to_dev(rows.astype(np.int32), dev)
    geom_d, lid_d = to_dev(geo.tri_table(), dev), to_dev(geo.tri_lid, dev)
    rel_d, rc_d = to_dev(rel, dev), to_dev(rel_cols.reshape(-1), dev)
    outs = []
    for _ in range(2):
        out = torch.full((rows.size, cols.size), 7.0, dtype=torch.float64, device=dev)
        _capi.check(L.smx_lindholm_fill(_capi.ptr(Xb_d), _capi.ptr(rows_d), rows.size, _capi.ptr(geom_d), geo.Nt,
                                        _capi.ptr(lid_d), _capi.ptr(rel_d), _capi.ptr(rc_d), rel.size, cols.size,
                                        out.data_ptr(), cols.size, 1, 0, _capi.stream_ptr()))
        outs.append(out)
    torch.cuda.synchronize()
    shared = np.bincount(geo.tri_lid[rel].ravel(), minlength=geo.Nb)[cols].max()
    err = float(np.abs(outs[0].cpu().numpy() - ref).max() / np.abs(ref).max())
    print(f"[film fill] {rel.size} relevant triangles (a column shared by up to {shared}); "
          f"two fills equal: {torch.equal(outs[0], outs[1])}; vs oracle {err:.2e}", flush=True)
    assert shared > 1 and torch.equal(outs[0], outs[1]) and err <= 1e-13


def test_film_rebuild_bitwise(film):
    """A second device build of the film H-matrix: the same block tree and
    bitwise-equal factors and near field (0 acceptance flips across builds)."""
    from smx.hmatrix import HMatrixBuildConfig, build_hmatrix_device
    a = film["data"]
    b = build_hmatrix_device(film["geo"], HMatrixBuildConfig(epsilon=EPS, eta=2.0, leaf_size=64), film["dev"])
    assert np.array_equal(a.block_tree, b.block_tree)
    assert len(a.far_blocks) == len(b.far_blocks)
    same = all(np.array_equal(x.U, y.U) and np.array_equal(x.V, y.V) for x, y in zip(a.far_blocks, b.far_blocks))
    assert same
    assert np.array_equal(a.near_csr.indptr, b.near_csr.indptr)
    assert np.array_equal(a.near_csr.indices, b.near_csr.indices)
    assert np.array_equal(a.near_csr.data, b.near_csr.data)
    print(f"[film rebuild] {len(b.far_blocks)} far blocks bitwise equal, near field bitwise equal", flush=True)
