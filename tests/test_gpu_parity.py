"""Device parity: HIP kernels through the C ABI vs the FP64 oracle.

Tolerances: Lindholm assembly 1e-13 (max abs error / max |B|); FP64 demag field
and LLG RHS 1e-10 relative L2 (CG run to ||r|| <= 1e-14 ||b||); FP32/BF16
boundary matvec errors are measured against the FP64 product and bounded by
the storage rounding (2^-24 and 2^-8 relative per entry).
"""
import math

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import smx_oracle as O

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    return torch


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def c1():
    from smx import mesh
    m = mesh.cube_c1()
    bd = O.boundary_data(m.X, m.tets)
    B = O.dense_B_rows(bd)
    return m, bd, B


def random_unit(n, seed=0):
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((n, 3))
    return (g / np.linalg.norm(g, axis=1)[:, None]).ravel()


def test_dense_B_c1(c1):
    torch = _torch()
    from smx.boundary import build_boundary_geometry
    from smx.demag import assemble_dense_B
    m, bd, B = c1
    geo = build_boundary_geometry(m.X, m.tets)
    Bd, ld = assemble_dense_B(geo, torch.device("cuda"), torch.float64)
    Bh = Bd.cpu().numpy()[:, :geo.Nb]
    err = np.abs(Bh - B).max() / np.abs(B).max()
    print(f"[dense_B_c1] Nb={geo.Nb} max err {err:.3e}", flush=True)
    assert err <= 1e-13
    assert np.abs(Bh.sum(axis=1) + 1.0).max() <= 1e-12


def test_block_fill_matches_oracle(c1):
    torch = _torch()
    from smx import _capi
    from smx.boundary import build_boundary_geometry
    from smx.device import to_dev
    m, bd, _ = c1
    geo = build_boundary_geometry(m.X, m.tets)
    rng = np.random.default_rng(3)
    rows = np.sort(rng.choice(geo.Nb, 77, replace=False))
    cols = np.sort(rng.choice(geo.Nb, 131, replace=False))
    ref = O.block_B(bd, rows, cols)
    dev = torch.device("cuda")
    rel, rel_cols = geo.relevant_triangles(cols)
    out = torch.full((rows.size, cols.size), 7.0, dtype=torch.float64, device=dev)
    L = _capi.lib()
    # keep every device argument alive until the kernel has finished
    Xb_d, rows_d = to_dev(geo.Xb, dev), to_dev(rows.astype(np.int32), dev)
    geom_d, lid_d = to_dev(geo.tri_table(), dev), to_dev(geo.tri_lid, dev)
    rel_d, rc_d = to_dev(rel, dev), to_dev(rel_cols.reshape(-1), dev)
    _capi.check(L.smx_lindholm_fill(_capi.ptr(Xb_d), _capi.ptr(rows_d), rows.size, _capi.ptr(geom_d), geo.Nt,
                                    _capi.ptr(lid_d), _capi.ptr(rel_d), _capi.ptr(rc_d), rel.size, cols.size,
                                    out.data_ptr(), cols.size, 1, 0, _capi.stream_ptr()))
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    err = np.abs(got - ref).max() / max(np.abs(ref).max(), 1e-300)
    print(f"[block_fill] err {err:.3e}", flush=True)
    assert err <= 1e-13


@pytest.mark.parametrize("dtype", ["float64", "float32", "bfloat16"])
def test_dense_matvec_dtypes(c1, dtype):
    torch = _torch()
    from smx import _capi
    m, bd, B = c1
    n = B.shape[0]
    ld = (n + 7) // 8 * 8
    tdt = getattr(torch, dtype)
    Bp = np.zeros((n, ld))
    Bp[:, :n] = B
    Bt = torch.from_numpy(Bp).to(tdt).cuda()
    x = np.random.default_rng(1).standard_normal(n)
    xt = torch.from_numpy(x).cuda()
    y = torch.zeros(n, dtype=torch.float64, device="cuda")
    L = _capi.lib()
    _capi.check(L.smx_dense_bmatvec(Bt.data_ptr(), _capi.DTYPE_CODE[tdt], n, ld, xt.data_ptr(), None, None,
                                    0.0, None, y.data_ptr(), _capi.stream_ptr()))
    torch.cuda.synchronize()
    # exact product with the stored (rounded) matrix
    Bs = Bt.to(torch.float64).cpu().numpy()[:, :n]
    exact = Bs @ x
    e_exact = rel_l2(y.cpu().numpy(), exact)
    e_ref = rel_l2(y.cpu().numpy(), B @ x)
    print(f"[matvec {dtype}] vs stored {e_exact:.2e} vs fp64 {e_ref:.2e}", flush=True)
    assert e_exact <= 1e-13
    bound = {"float64": 1e-13, "float32": 1e-6, "bfloat16": 2e-2}[dtype]
    assert e_ref <= bound
    # transposed product
    w = torch.zeros(((n + 255) // 256) * n, dtype=torch.float64, device="cuda")
    yt = torch.zeros(n, dtype=torch.float64, device="cuda")
    _capi.check(L.smx_dense_bmatvec_t(Bt.data_ptr(), _capi.DTYPE_CODE[tdt], n, ld, xt.data_ptr(), yt.data_ptr(),
                                      w.data_ptr(), _capi.stream_ptr()))
    torch.cuda.synchronize()
    assert rel_l2(yt.cpu().numpy(), Bs.T @ x) <= 1e-13


def test_sell_spmv_kinds(c1):
    torch = _torch()
    from smx.demag import OperatorSet
    m, _, _ = c1
    ops = OperatorSet(m, torch.device("cuda"))
    f = ops.fem
    g = f.graph
    r = np.repeat(np.arange(f.N), np.diff(g.indptr))
    rng = np.random.default_rng(5)
    x1 = rng.standard_normal(f.N)
    x3 = rng.standard_normal(3 * f.N)
    dev = torch.device("cuda")
    A = sp.csr_matrix((f.stiff, (r, g.indices)), shape=(f.N, f.N))
    y = torch.zeros(f.N, dtype=torch.float64, device=dev)
    As = ops.graph_values(f.stiff, 1)
    s = torch.zeros(As.blocks(), dtype=torch.float64, device=dev)
    As.spmv(torch.from_numpy(x1).cuda(), y, 0, partials=s)
    torch.cuda.synchronize()
    assert rel_l2(y.cpu().numpy(), A @ x1) <= 1e-14
    assert abs(s.sum().item() - (A @ x1).sum()) <= 1e-10 * np.abs(A @ x1).sum()
    Div = sp.csr_matrix((f.div.reshape(-1), (np.repeat(r, 3), (3 * g.indices[:, None] + np.arange(3)).ravel())),
                        shape=(f.N, 3 * f.N))
    ops.graph_values(f.div, 3).spmv(torch.from_numpy(x3).cuda(), y, 1, partials=s)
    torch.cuda.synchronize()
    assert rel_l2(y.cpu().numpy(), Div @ x3) <= 1e-13
    G = sp.csr_matrix((f.grad.reshape(-1), ((3 * r[:, None] + np.arange(3)).ravel(), np.repeat(g.indices, 3))),
                      shape=(3 * f.N, f.N))
    y3 = torch.zeros(3 * f.N, dtype=torch.float64, device=dev)
    ops.graph_values(f.grad, 3).spmv(torch.from_numpy(x1).cuda(), y3, 2)
    torch.cuda.synchronize()
    assert rel_l2(y3.cpu().numpy(), G @ x1) <= 1e-13
    blocks = rng.standard_normal((g.nnz, 3, 3))
    kr = np.broadcast_to(3 * r[:, None, None] + np.arange(3)[None, :, None], blocks.shape).ravel()
    kc = np.broadcast_to(3 * g.indices[:, None, None] + np.arange(3)[None, None, :], blocks.shape).ravel()
    K = sp.csr_matrix((blocks.reshape(-1), (kr, kc)), shape=(3 * f.N, 3 * f.N))
    ops.graph_values(blocks.reshape(-1, 9), 9).spmv(torch.from_numpy(x3).cuda(), y3, 3)
    torch.cuda.synchronize()
    assert rel_l2(y3.cpu().numpy(), K @ x3) <= 1e-13


def _demag_pair(m, bd, B, **kw):
    torch = _torch()
    from smx.demag import DemagFieldLindholmGPU
    Ms = 8e5
    d = DemagFieldLindholmGPU(m, None, None, Ms, ksp_rtol_u1=1e-14, ksp_rtol_u2=1e-14, ksp_atol=0.0,
                              ksp_max_it_u1=5000, ksp_max_it_u2=5000, **kw)
    ref = O.DemagOracle(m.X, m.tets, Ms, B=B, bd=bd)
    return d, ref


@pytest.mark.parametrize("pc", ["gamg", "jacobi", "auto"])
def test_demag_field_c1_fp64(c1, pc):
    torch = _torch()
    m, bd, B = c1
    d, ref = _demag_pair(m, bd, B, ksp_pc_type=pc)
    mv = random_unit(m.num_nodes, 0)
    H = torch.zeros(3 * m.num_nodes, dtype=torch.float64, device="cuda")
    d.compute_vec(torch.from_numpy(mv).cuda(), H)
    torch.cuda.synchronize()
    Hr, parts = ref.compute(mv, return_parts=True)
    e = rel_l2(H.cpu().numpy(), Hr)
    u1 = d.u1.cpu().numpy()
    e_u1 = rel_l2(u1 - u1.mean(), parts["u1"])
    e_u2 = rel_l2(d.u2.cpu().numpy(), parts["u2"])
    print(f"[demag c1 {pc}] rel L2 H {e:.3e} u1 {e_u1:.3e} u2 {e_u2:.3e} its {d.last_its_u1}/{d.last_its_u2}", flush=True)
    assert e_u1 <= 1e-10 and e_u2 <= 1e-10
    assert e <= 1e-10
    E = d.Energy_lumped_gpu(torch.from_numpy(mv).cuda(), H)
    Er = ref.energy(mv, Hr)
    assert abs(E - Er) <= 1e-10 * abs(Er)


def test_demag_uniform_patch(c1):
    """Uniform m: u1 = Ms m.x (+c) exactly, so the u1 part of H is -Ms m."""
    torch = _torch()
    m, bd, B = c1
    d, ref = _demag_pair(m, bd, B)
    mv = np.tile([0.0, 0.0, 1.0], m.num_nodes)
    H = torch.zeros(3 * m.num_nodes, dtype=torch.float64, device="cuda")
    d.compute_vec(torch.from_numpy(mv).cuda(), H)
    torch.cuda.synchronize()
    u1 = d.u1.cpu().numpy()
    z = m.X[:, 2]
    expect = 8e5 * (z - z.mean())
    assert rel_l2(u1 - u1.mean(), expect) <= 1e-11
    Hz = H.cpu().numpy().reshape(-1, 3)
    vol = d.ops.fem.volume
    avg = (Hz[:, 2] * vol).sum() / vol.sum() / 8e5
    print(f"[uniform cube] <Hz>/Ms = {avg:.4f}", flush=True)
    assert -0.40 < avg < -0.28


@pytest.mark.parametrize("dtype", ["float32", "bfloat16"])
def test_demag_lowprec_boundary_error(c1, dtype):
    torch = _torch()
    m, bd, B = c1
    d64, ref = _demag_pair(m, bd, B)
    dlo, _ = _demag_pair(m, bd, B, B_dtype=dtype)
    mv = torch.from_numpy(random_unit(m.num_nodes, 2)).cuda()
    H64 = torch.zeros_like(mv)
    Hlo = torch.zeros_like(mv)
    d64.compute_vec(mv, H64)
    dlo.compute_vec(mv, Hlo)
    torch.cuda.synchronize()
    e = rel_l2(Hlo.cpu().numpy(), H64.cpu().numpy())
    print(f"[demag {dtype} boundary] rel err vs fp64 {e:.3e}", flush=True)
    assert e <= (1e-6 if dtype == "float32" else 1e-2)


def _hef_pair(m, bd, B, **extra):
    torch = _torch()
    from smx.llg import EffectiveField
    kw = dict(Ms=8e5, Aex=1.3e-11, Ku=5e5, n_ani_vec=(0.0, 0.0, 1.0), D_int=3e-3,
              n0_int_vec=(0.0, 0.0, 1.0), alpha=0.5)
    kw.update(extra)
    hef = EffectiveField(m, use_demag=True,
                         demag_kwargs=dict(ksp_rtol_u1=1e-14, ksp_rtol_u2=1e-14, ksp_atol=0.0,
                                           ksp_max_it_u1=5000, ksp_max_it_u2=5000), **kw)
    dem = O.DemagOracle(m.X, m.tets, kw["Ms"], B=B, bd=bd)
    fem = O.fem_operators(m.X, m.tets, 1.0)
    Kex = O.exchange_matrix(m.X, m.tets, fem.V, kw["Aex"], kw["Ms"])
    Kd = O.dmi_interfacial_matrix(m.X, m.tets, fem.V, kw["D_int"], kw["Ms"], kw["n0_int_vec"])
    if kw.get("D_bulk", 0.0):
        Kd = Kd + O.dmi_bulk_matrix(m.X, m.tets, fem.V, kw["D_bulk"], kw["Ms"])

    def field(mv, demag=True):
        H = Kex @ mv + Kd @ mv + O.anisotropy_field(mv, kw["n_ani_vec"], kw["Ku"], kw["Ms"])
        if demag:
            H = H + dem.compute(mv)
        return H

    return hef, field, kw


def test_llg_rhs_c1(c1):
    torch = _torch()
    m, bd, B = c1
    hef, field, kw = _hef_pair(m, bd, B)
    mv = random_unit(m.num_nodes, 4)
    out = torch.zeros(3 * m.num_nodes, dtype=torch.float64, device="cuda")
    hef.llg_rhs_vec(torch.from_numpy(mv).cuda(), out)
    torch.cuda.synchronize()
    Hr = field(mv)
    rr = O.llg_rhs(mv, Hr, 2.211e5, kw["alpha"], kw["Ms"])
    eH = rel_l2(hef.H_eff_gpu.cpu().numpy(), Hr)
    er = rel_l2(out.cpu().numpy(), rr)
    print(f"[llg rhs c1] H {eH:.3e} rhs {er:.3e}", flush=True)
    assert eH <= 1e-10 and er <= 1e-10


@pytest.mark.parametrize("D_int", [0.0, 3e-3])
def test_llg_rhs_bulk_dmi_c1(c1, D_int):
    """Bulk DMI (D = 3e-3 J/m^2) alone and with interfacial DMI, through the
    fused field kernel, against the oracle's dmi_bulk_matrix (DMI_Bulk_GPU.py:
    43-86, 248-264): H and RHS within 1e-10 relative L2; the JVP too."""
    torch = _torch()
    m, bd, B = c1
    hef, field, kw = _hef_pair(m, bd, B, D_bulk=3e-3, D_int=D_int)
    mv = random_unit(m.num_nodes, 14)
    out = torch.zeros(3 * m.num_nodes, dtype=torch.float64, device="cuda")
    mt = torch.from_numpy(mv).cuda()
    hef.llg_rhs_vec(mt, out)
    torch.cuda.synchronize()
    Hr = field(mv)
    rr = O.llg_rhs(mv, Hr, 2.211e5, kw["alpha"], kw["Ms"])
    eH = rel_l2(hef.H_eff_gpu.cpu().numpy(), Hr)
    er = rel_l2(out.cpu().numpy(), rr)
    v = np.random.default_rng(15).standard_normal(3 * m.num_nodes)
    hef.update_jac_state(mt)
    jv = torch.zeros_like(mt)
    hef.jac_vec_times_vec(torch.from_numpy(v).cuda(), jv)
    torch.cuda.synchronize()
    ej = rel_l2(jv.cpu().numpy(), O.jac_vec(mv, field(mv, demag=False), v, field(v, demag=False), 2.211e5,
                                            kw["alpha"], kw["Ms"]))
    print(f"[llg rhs bulk DMI c1, D_int={D_int}] H {eH:.3e} rhs {er:.3e} jvp {ej:.3e}", flush=True)
    assert eH <= 1e-10 and er <= 1e-10 and ej <= 1e-10


def test_llg_jvp_c1(c1):
    torch = _torch()
    m, bd, B = c1
    hef, field, kw = _hef_pair(m, bd, B)
    rng = np.random.default_rng(7)
    mv = random_unit(m.num_nodes, 5)
    v = rng.standard_normal(3 * m.num_nodes)
    mt = torch.from_numpy(mv).cuda()
    hef.update_jac_state(mt)
    out = torch.zeros_like(mt)
    hef.jac_vec_times_vec(torch.from_numpy(v).cuda(), out)
    torch.cuda.synchronize()
    Hm = field(mv, demag=False)
    Hv = field(v, demag=False)
    ref = O.jac_vec(mv, Hm, v, Hv, 2.211e5, kw["alpha"], kw["Ms"])
    e = rel_l2(out.cpu().numpy(), ref)
    print(f"[jvp c1] {e:.3e}", flush=True)
    assert e <= 1e-10


def test_dp5_fixed_steps_c1(c1):
    """The 40 nm cube workload: exchange + FEM/BEM demag, alpha 0.5, seeded
    random m, 100 fixed Dormand-Prince steps of 1e-13 s (601 RHS
    evaluations) on the device against the oracle's (LU Poisson solves).
    Bound: 1e-9 relative L2 on the final state (each evaluation agrees to
    ~1e-12; the bound leaves room for their accumulation over 100 steps)."""
    torch = _torch()
    from smx.llg import DormandPrince
    m, bd, B = c1
    hef, field, kw = _hef_pair(m, bd, B, Ku=0.0, D_int=0.0)
    mv = random_unit(m.num_nodes, 6)
    y = torch.from_numpy(mv.copy()).cuda()
    dp = DormandPrince(hef)
    dt = 1e-13
    nsteps = 100
    f = lambda t, yy: O.llg_rhs(yy, field(yy), 2.211e5, kw["alpha"], kw["Ms"])
    yr = mv.copy()
    k1 = None
    worst = 0.0
    for chunk in range(4):
        dp.run_steps(y, chunk * 25 * dt, dt, 25)
        torch.cuda.synchronize()
        for _ in range(25):
            yr, _, k1 = O.dp5_step(f, 0.0, yr, dt, k1)
        worst = max(worst, rel_l2(y.cpu().numpy(), yr))
    d = rel_l2(y.cpu().numpy(), mv)
    print(f"[dp5 c1] {nsteps} steps: worst rel err {worst:.3e} (change of m {d:.3e}, "
          f"{dp.rhs_evals} device RHS evaluations)", flush=True)
    assert worst <= 1e-9 and d > 1e-2


def test_demag_against_golden_fixture():
    """Device FP64 demag field vs the committed oracle fixture (40 nm cube)."""
    from pathlib import Path
    torch = _torch()
    from smx import mesh
    from smx.demag import DemagFieldLindholmGPU
    G = np.load(Path(__file__).parent / "golden" / "cube40_oracle.npz", allow_pickle=False)
    m = mesh.cube_c1()
    d = DemagFieldLindholmGPU(m, None, None, 8e5, ksp_rtol_u1=1e-14, ksp_rtol_u2=1e-14, ksp_atol=0.0,
                              ksp_max_it_u1=5000, ksp_max_it_u2=5000)
    H = torch.empty(3 * m.num_nodes, dtype=torch.float64, device="cuda")
    d.compute_vec(torch.from_numpy(G["m"]).cuda(), H)
    torch.cuda.synchronize()
    e = rel_l2(H.cpu().numpy(), G["H_demag"])
    print(f"[golden demag] {e:.3e}", flush=True)
    assert e <= 1e-10


@pytest.mark.parametrize("dmi", [dict(D_int=3e-3), dict(D_int=0.0, D_bulk=3e-3), dict(D_int=3e-3, D_bulk=2e-3)])
def test_dmi_compact_layout_matches_full_block(c1, dmi):
    """The fused field kernel and the JVP read the DMI blocks in the compact
    layout (only the components nonzero somewhere: interfacial DMI along z
    has none in the xx, xy, yx, yy places, bulk DMI none on the diagonal);
    the same product with all nine components stored gives the same bits
    (H, RHS, JVP)."""
    torch = _torch()
    from smx.fem import dmi_bulk_values, dmi_interfacial_values
    m, bd, B = c1
    hef, field, kw = _hef_pair(m, bd, B, **dmi)
    ncomp = int(np.count_nonzero(hef.dmi_off >= 0))
    mt = torch.from_numpy(random_unit(m.num_nodes, 21)).cuda()
    v = torch.from_numpy(np.random.default_rng(22).standard_normal(3 * m.num_nodes)).cuda()

    def run():
        out = torch.zeros_like(mt)
        hef.llg_rhs_vec(mt, out)
        hef.update_jac_state(mt)
        jv = torch.zeros_like(mt)
        hef.jac_vec_times_vec(v, jv)
        torch.cuda.synchronize()
        return hef.H_eff_gpu.clone(), out, jv

    compact = run()
    fem = hef.ops.fem
    full = 0.0
    if dmi.get("D_int"):
        full = full + dmi_interfacial_values(m.X, m.tets, fem, dmi["D_int"], kw["Ms"], (0.0, 0.0, 1.0))
    if dmi.get("D_bulk"):
        full = full + dmi_bulk_values(m.X, m.tets, fem, dmi["D_bulk"], kw["Ms"])
    hef.dmi = hef.ops.graph_values(full.reshape(-1, 9), 9)
    hef.dmi_off = np.arange(9, dtype=np.int32)
    nine = run()
    print(f"[dmi layout {dmi}] {ncomp} of 9 components stored", flush=True)
    assert ncomp <= (5 if not dmi.get("D_bulk") else 7 if dmi.get("D_int") else 6)
    for a, b in zip(compact, nine):
        assert torch.equal(a, b)
