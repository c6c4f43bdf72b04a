"""H-matrix on the device: batched fill + ACA build and the packed matvec.

Partition: compared with the oracle's depth-first build (ACA acceptance flips
are counted and must be zero on these meshes).  Matvec: the
validate_against_dense protocol (seed 1234, 3 trials) against the dense FP64
Lindholm B, bounded by 10 eps (FP64 factors) and 10 eps + 1e-6 (FP32 factors).
"""
import numpy as np
import pytest

import smx_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c1():
    from smx import mesh as M
    from smx.boundary import build_boundary_geometry
    m = M.cube_c1()
    bd = O.boundary_data(m.X, m.tets)
    return m, build_boundary_geometry(m.X, m.tets), bd, O.dense_B_rows(bd)


@pytest.mark.parametrize("leaf,eps", [(32, 1e-6), (16, 1e-4)])
def test_device_build_matches_oracle(c1, leaf, eps):
    import torch
    from smx.hmatrix import HMatrixBuildConfig, build_hmatrix_device
    m, geo, bd, B = c1
    data = build_hmatrix_device(geo, HMatrixBuildConfig(epsilon=eps, leaf_size=leaf), torch.device("cuda"))
    ref = O.build_hmatrix(bd, eps=eps, eta=2.0, leaf_size=leaf)
    got = {(b.row_start, b.col_start): b for b in data.far_blocks}
    exp = {(f[0], f[2]): f for f in ref.far}
    flips = set(got) ^ set(exp)
    print(f"[hmatrix build leaf={leaf} eps={eps}] far {len(got)} vs oracle {len(exp)}, flips {len(flips)}, "
          f"build {data.metadata['build_time_seconds']:.2f}s", flush=True)
    assert not flips
    rank_diff = sum(got[k].rank != exp[k][4].shape[1] for k in got)
    assert rank_diff == 0
    assert np.array_equal(data.near_csr.indices, ref.near.indices)
    assert np.abs(data.near_csr.data - ref.near.data).max() <= 1e-13 * np.abs(ref.near.data).max()
    for k, b in got.items():
        e = exp[k]
        blk = B[np.ix_(data.perm[b.row_start:b.row_stop], data.perm[b.col_start:b.col_stop])]
        assert np.linalg.norm(b.U @ b.V - blk) <= eps * (1 + 1e-9) * np.linalg.norm(blk) + 1e-300


@pytest.mark.parametrize("fp32", [False, True])
def test_device_matvec_validate_against_dense(c1, fp32):
    import torch
    from smx.hmatrix import HMatrixBuildConfig, HMatrixGPUPackedFused, build_hmatrix_device
    m, geo, bd, B = c1
    eps = 1e-6
    data = build_hmatrix_device(geo, HMatrixBuildConfig(epsilon=eps, leaf_size=32), torch.device("cuda"))
    H = HMatrixGPUPackedFused(data, use_fp32=fp32, boundary_gather=geo.bverts)
    rng = np.random.default_rng(1234)
    worst = 0.0
    for _ in range(3):
        x = rng.standard_normal(geo.Nb)
        y = H.matvec_numpy(x)
        worst = max(worst, np.linalg.norm(y - B @ x) / np.linalg.norm(B @ x))
    print(f"[hmatrix matvec fp32={fp32}] rel err vs dense {worst:.3e}", flush=True)
    assert worst <= 10 * eps + (1e-6 if fp32 else 0.0)
    # boundary gather + shift path used by the demag field
    N = m.num_nodes
    u = torch.from_numpy(rng.standard_normal(N)).cuda()
    s = torch.tensor([u.sum().item()], dtype=torch.float64, device="cuda")
    y = torch.empty(geo.Nb, dtype=torch.float64, device="cuda")
    H.matvec_from(u, y, gather=True, shift_ptr=s.data_ptr(), shift_scale=1.0 / N)
    torch.cuda.synchronize()
    xb = u.cpu().numpy()[geo.bverts] - u.cpu().numpy().mean()
    assert np.linalg.norm(y.cpu().numpy() - B @ xb) / np.linalg.norm(B @ xb) <= 10 * eps + 1e-6


def test_demag_hmatrix_vs_dense(c1):
    import torch
    from smx.demag import DemagFieldLindholmGPU
    m, geo, bd, B = c1
    kw = dict(ksp_rtol_u1=1e-13, ksp_rtol_u2=1e-13, ksp_atol=0.0, ksp_max_it_u1=5000, ksp_max_it_u2=5000)
    dd = DemagFieldLindholmGPU(m, None, None, 8e5, **kw)
    dh = DemagFieldLindholmGPU(m, None, None, 8e5, boundary_backend="hmatrix", hmatrix_epsilon=1e-6,
                               hmatrix_leaf_size=32, **kw)
    rng = np.random.default_rng(0)
    g = rng.standard_normal((m.num_nodes, 3))
    mv = torch.from_numpy((g / np.linalg.norm(g, axis=1)[:, None]).ravel()).cuda()
    Hd = torch.empty_like(mv)
    Hh = torch.empty_like(mv)
    dd.compute_vec(mv, Hd)
    dh.compute_vec(mv, Hh)
    torch.cuda.synchronize()
    e = float(torch.linalg.norm(Hh - Hd) / torch.linalg.norm(Hd))
    print(f"[demag hmatrix vs dense c1] {e:.3e}", flush=True)
    assert e <= 1e-5
    rep = dh.boundary_memory_report()
    assert rep["backend"] == "hmatrix" and rep["far_blocks"] > 0


def test_c2_build_and_matvec():
    """150 nm cube (Nb = 5402): leaf 64, eta 2, eps 1e-6; dense FP64 B from the
    device as the reference operator."""
    import time
    import torch
    from smx import mesh as M
    from smx.boundary import build_boundary_geometry
    from smx.demag import assemble_dense_B
    from smx.hmatrix import HMatrixBuildConfig, HMatrixGPUPackedFused, build_hmatrix_device
    m = M.cube_c2()
    geo = build_boundary_geometry(m.X, m.tets)
    dev = torch.device("cuda")
    t0 = time.perf_counter()
    data = build_hmatrix_device(geo, HMatrixBuildConfig(epsilon=1e-6, leaf_size=64), dev)
    t_build = time.perf_counter() - t0
    Bd, ld = assemble_dense_B(geo, dev)
    Bh = Bd[:, :geo.Nb]
    md = data.metadata
    print(f"[c2 hmatrix] build {t_build:.2f}s far={md['low_rank_blocks']} dense={md['dense_blocks']} "
          f"near_nnz={md['near_nnz']} far_entries={md['far_storage_entries']} "
          f"avg_rank={md['average_far_rank']:.1f} compression={md['compression_ratio_entries']:.2f}", flush=True)
    for fp32 in (False, True):
        H = HMatrixGPUPackedFused(data, use_fp32=fp32)
        rng = np.random.default_rng(1234)
        worst = 0.0
        for _ in range(3):
            x = torch.from_numpy(rng.standard_normal(geo.Nb)).cuda()
            y = H.matvec(x)
            ref = Bh @ x
            worst = max(worst, float(torch.linalg.norm(y - ref) / torch.linalg.norm(ref)))
        bm = H.benchmark(repeats=20)
        print(f"[c2 hmatrix fp32={fp32}] rel err {worst:.3e}, {bm['average_seconds']*1e6:.1f} us/matvec, "
              f"{H.algorithmic_bytes()/bm['average_seconds']/1e9:.0f} GB/s", flush=True)
        assert worst <= 1e-5


@pytest.mark.parametrize("fp32", [False, True])
def test_near_rows_in_z_launch(c1, fp32):
    """Near rows summed by their own launch against the leaf kernel's own
    near loop: equal to rounding (the two sum near and far parts of a row in
    separate wave totals), forward and transposed."""
    import torch
    from smx.hmatrix import HMatrixBuildConfig, HMatrixGPUPackedFused, build_hmatrix_device
    m, geo, bd, B = c1
    data = build_hmatrix_device(geo, HMatrixBuildConfig(epsilon=1e-6, leaf_size=32), torch.device("cuda"))
    Ha = HMatrixGPUPackedFused(data, use_fp32=fp32, near_in_z=True)
    Hb = HMatrixGPUPackedFused(data, use_fp32=fp32, near_in_z=False)
    x = torch.from_numpy(np.random.default_rng(9).standard_normal(geo.Nb)).cuda()
    for fa, fb in ((Ha.matvec, Hb.matvec), (Ha.rmatvec, Hb.rmatvec)):
        ya, yb = fa(x), fb(x)
        assert float(torch.linalg.norm(ya - yb) / torch.linalg.norm(yb)) <= 1e-15
    assert np.linalg.norm(Ha.matvec_numpy(x.cpu().numpy()) - B @ x.cpu().numpy()) <= \
        (10 * 1e-6 + (1e-6 if fp32 else 0.0)) * np.linalg.norm(B @ x.cpu().numpy())


@pytest.mark.parametrize("fp32", [False, True])
def test_near_task_forms_agree(c1, fp32):
    """The near launch's forms give the same bits: a wavefront per row, and
    one workgroup per row task with 4, 8 or 16 wavefronts (plain, boundary
    gather with a shift, transposed); other task widths are rejected."""
    import torch
    from smx.hmatrix import HMatrixBuildConfig, HMatrixGPUPackedFused, build_hmatrix_device
    m, geo, bd, B = c1
    data = build_hmatrix_device(geo, HMatrixBuildConfig(epsilon=1e-6, leaf_size=32), torch.device("cuda"))
    kw = dict(use_fp32=fp32, near_in_z=True, boundary_gather=geo.bverts)
    forms = {"row": HMatrixGPUPackedFused(data, near_tasks=False, **kw),
             "task": HMatrixGPUPackedFused(data, **kw)}
    assert forms["row"]._H.near_task is None and forms["task"]._H.near_task is not None
    rng = np.random.default_rng(11)
    x = torch.from_numpy(rng.standard_normal(geo.Nb)).cuda()
    u = torch.from_numpy(rng.standard_normal(m.num_nodes)).cuda()
    shift = torch.tensor([0.37], dtype=torch.float64, device="cuda")
    gather = torch.from_numpy(np.asarray(geo.bverts)).cuda()

    def outputs(H):
        yb = torch.empty_like(x)
        H.matvec_from(u, yb, gather=gather, shift_ptr=shift.data_ptr(), shift_scale=1.5)
        return H.matvec(x).clone(), yb, H.rmatvec(x).clone()

    ref = outputs(forms["row"])
    print(f"[near forms fp32={fp32}] Nb={geo.Nb} tasks={forms['task'].fwd['n_near_task']}", flush=True)
    for name, H in forms.items():
        for nw in ((4, 8, 16, 0) if name != "row" else (0,)):
            for S in (H._H, H._Hb):
                S.near_waves = nw
            if hasattr(H, "_HT"):
                H._HT.near_waves = nw
            got = outputs(H)
            for what, a, b in zip(("matvec", "gather+shift", "rmatvec"), got, ref):
                assert torch.equal(a, b), (name, nw, what, float((a - b).abs().max()))
    H = forms["task"]
    H._H.near_waves = 12
    with pytest.raises(ValueError):
        H.matvec(x)
    H._H.near_waves = 0
    yb = ref[1].cpu().numpy()
    xb = u.cpu().numpy()[np.asarray(geo.bverts)] - 0.37 * 1.5
    assert np.linalg.norm(yb - B @ xb) <= (10 * 1e-6 + (1e-6 if fp32 else 0.0)) * np.linalg.norm(B @ xb)


@pytest.mark.parametrize("fp32", [False, True])
def test_leaf_workgroup_shapes_agree(c1, fp32):
    """4, 8 and 16 wavefronts per leaf workgroup give the same bits (each
    row is one wavefront's dot product either way); other values are
    rejected before any launch."""
    import torch
    from smx import _capi
    from smx.hmatrix import HMatrixBuildConfig, HMatrixGPUPackedFused, build_hmatrix_device
    m, geo, bd, B = c1
    data = build_hmatrix_device(geo, HMatrixBuildConfig(epsilon=1e-6, leaf_size=32), torch.device("cuda"))
    H = HMatrixGPUPackedFused(data, use_fp32=fp32, near_in_z=False)
    x = torch.from_numpy(np.random.default_rng(10).standard_normal(geo.Nb)).cuda()
    ref = H.matvec(x).clone()
    for nw in (4, 8, 16, 0):
        H._H.leaf_waves = nw
        assert torch.equal(H.matvec(x), ref), nw
    H._H.leaf_waves = 12
    with pytest.raises(ValueError):
        H.matvec(x)
    H._H.leaf_waves = 0


@pytest.mark.parametrize("fp32,near_in_z", [(False, False), (True, False), (True, True)])
def test_leaf_z_runs_equal_gather(c1, fp32, near_in_z):
    """The leaf launch's z staging by contiguous runs (leaf_zr) against the
    per-entry gather (leaf_zg): the same bits, forward and transposed; one of
    the two run pointers alone is rejected."""
    import torch
    from smx.hmatrix import HMatrixBuildConfig, HMatrixGPUPackedFused, build_hmatrix_device
    m, geo, bd, B = c1
    data = build_hmatrix_device(geo, HMatrixBuildConfig(epsilon=1e-6, leaf_size=32), torch.device("cuda"))
    H = HMatrixGPUPackedFused(data, use_fp32=fp32, near_in_z=near_in_z)
    x = torch.from_numpy(np.random.default_rng(12).standard_normal(geo.Nb)).cuda()
    assert H._H.leaf_zr is not None
    runs = H.fwd["leaf_zr"].numel() // 2
    print(f"[leaf z runs fp32={fp32} near_in_z={near_in_z}] Nb={geo.Nb} runs={runs} "
          f"entries={H.fwd['leaf_zg'].numel()}", flush=True)
    ref = (H.matvec(x).clone(), H.rmatvec(x).clone())
    for S in (H._H, H._HT):
        S.leaf_zr, S.leaf_zr_off = None, None
    got = (H.matvec(x).clone(), H.rmatvec(x).clone())
    for a, b in zip(got, ref):
        assert torch.equal(a, b)
    H._H.leaf_zr_off = H.fwd["leaf_zr_off"].data_ptr()
    with pytest.raises(ValueError):
        H.matvec(x)
    assert np.linalg.norm(ref[0].cpu().numpy() - B @ x.cpu().numpy()) <= \
        (10 * 1e-6 + (1e-6 if fp32 else 0.0)) * np.linalg.norm(B @ x.cpu().numpy())


@pytest.mark.parametrize("fp32,k,near_fp32", [(False, 64, False), (False, 7, False), (True, 64, False),
                                               (True, 1, False), (True, 64, True)])
def test_matmat_against_matvecs(fp32, k, near_fp32):
    """Multi-RHS product (MFMA low-rank stages) against k separate matvecs on
    the 150 nm cube, with the near field from the dense near tiles (FP64 MFMA)
    and from the CSR rows: FP64 factors 1e-12 (the MFMA tiles sum in another
    order), FP32 factors 1e-6 (X and Z rounded to FP32 once) relative per
    column; the two near forms agree to 1e-13 of the column norm."""
    import torch
    from smx import mesh as M
    from smx.boundary import build_boundary_geometry
    from smx.hmatrix import HMatrixBuildConfig, HMatrixGPUPackedFused, build_hmatrix_device
    m = M.cube_c2()
    geo = build_boundary_geometry(m.X, m.tets)
    data = build_hmatrix_device(geo, HMatrixBuildConfig(epsilon=1e-6, leaf_size=64), torch.device("cuda"))
    H = HMatrixGPUPackedFused(data, use_fp32=fp32, near_in_z=False, near_fp32=near_fp32)
    X = torch.from_numpy(np.random.default_rng(k).standard_normal((k, geo.Nb))).cuda()
    print(f"[matmat fp32={fp32} k={k} near_fp32={near_fp32}] Nb={geo.Nb}", flush=True)
    Y = H.matmat(X)
    Yc = H.matmat(X, near_tiles=False)
    torch.cuda.synchronize()
    st = H.near_tile_stats
    assert st["nnz"] == H.near_nnz and st["dense_entries"] >= st["nnz"]
    ref = torch.stack([H.matvec(X[j]) for j in range(k)])
    nrm = torch.linalg.norm(ref, dim=1)
    err = (torch.linalg.norm(Y - ref, dim=1) / nrm).max().item()
    errc = (torch.linalg.norm(Yc - ref, dim=1) / nrm).max().item()
    dforms = (torch.linalg.norm(Y - Yc, dim=1) / nrm).max().item()
    print(f"  max column rel err tiles {err:.2e} csr {errc:.2e} between {dforms:.2e}; "
          f"{st['ntile']} tiles, {st['dense_entries']} dense entries for {st['nnz']} near", flush=True)
    assert err <= (1e-6 if fp32 else 1e-12) and errc <= (1e-6 if fp32 else 1e-12)
    assert dforms <= 1e-13
