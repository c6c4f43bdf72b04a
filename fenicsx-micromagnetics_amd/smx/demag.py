"""Fredkin-Koehler (FEM/BEM) demagnetizing field with the Lindholm double layer.

Drop-in counterpart of ``DemagFieldLindholmGPU``
(gpu/fields/Demag_Lindholm_GPU.py:82-724): same constructor keywords, same
``compute_vec(m_vec, out_vec)`` / ``Energy_lumped_gpu`` / memory-report surface,
with torch device tensors in place of PETSc CUDA Vecs.  One evaluation is

    b   = Div m,  b -= mean(b)                   (Div SpMV + fixed-order sum)
    u1  : A u1 = b  (Jacobi CG, Neumann)          (one kernel per iteration)
    g|G = B (u1|G - mean(u1))                    (dense GEMV or H-matrix)
    u2  : A_bc u2 = lift(g)                        (Jacobi CG, Dirichlet)
    H   = G (u1 - mean(u1) + u2) / V                (node-graph gradient)

exactly the order of _solve_u1_gpu / _apply_B_gpu / _solve_u2_gpu / compute_vec
(:569-670).  The mean removal is the nullspace projection of the single
connected component (ns_u1.remove).
"""
from __future__ import annotations

import ctypes as C
import math
import time

import numpy as np
import scipy.sparse as sp
from scipy.sparse.csgraph import connected_components, reverse_cuthill_mckee
import torch

from . import _capi
from .boundary import BoundaryGeometry, build_boundary_geometry
from .device import DeviceEll, DeviceSell, default_device, reduce_workspace, to_dev
from .fem import FEMSystem, assemble_fem, compact_csr, to_sell
from .mesh import TetMesh

MU0 = 4.0 * math.pi * 1e-7

_B_DTYPES = {
    "float64": torch.float64, "fp64": torch.float64, "f64": torch.float64,
    "float32": torch.float32, "fp32": torch.float32, "f32": torch.float32,
    "bfloat16": torch.bfloat16, "bf16": torch.bfloat16,
}


def _torch_dtype(B_dtype):
    if isinstance(B_dtype, torch.dtype):
        return B_dtype
    if isinstance(B_dtype, str):
        key = B_dtype.lower()
    else:
        key = np.dtype(B_dtype).name
    if key not in _B_DTYPES:
        raise TypeError(f"unsupported B_dtype {B_dtype!r}")
    return _B_DTYPES[key]


def solver_operator(A: DeviceEll, dinv: torch.Tensor) -> DeviceEll:
    """The CG operator in solver form (smx_pcg_prepare): the diagonal of every
    row in slot 0 and every value a_ij * dinv_j, i.e. A D^-1 (a new copy)."""
    out = DeviceEll(nrows=A.nrows, width=A.width, ld=A.ld, cols=A.cols.clone(), vals=A.vals.clone(),
                    nnz=A.nnz)                    # rows n..ld-1 (never read) keep A's padding
    _capi.check(_capi.lib().smx_pcg_prepare(A.nrows, A.width, A.ld, _capi.ptr(A.cols), _capi.ptr(A.vals),
                                            _capi.ptr(dinv), _capi.ptr(out.cols), _capi.ptr(out.vals),
                                            _capi.stream_ptr()), "smx_pcg_prepare")
    return out


class OperatorSet:
    """Device copies of the FEM operators one mesh needs (shared between the
    demag field and the effective field)."""

    def __init__(self, mesh: TetMesh, device=None, fem: FEMSystem | None = None):
        self.device = device or default_device()
        self.mesh = mesh
        self.fem = fem or assemble_fem(mesh.X, mesh.tets)
        f = self.fem
        g = f.graph
        self.N = f.N
        self.graph_sell = to_sell(g.indptr, g.indices, np.zeros(g.nnz), f.N)
        self.graph = DeviceSell.upload(self.graph_sell, self.device)
        self._sell_pos = None
        self.inv_vol = to_dev(1.0 / f.volume, self.device)
        self.vol = to_dev(f.volume, self.device)

    def graph_values(self, vals: np.ndarray, k: int) -> DeviceSell:
        """Upload graph-ordered values (nnz, *k) in the node-graph SELL layout."""
        from .device import sell_values
        sv = sell_values(self.fem.graph.indptr, vals, self.N)
        return self.graph.with_vals(sv, k)

    def compact_host(self, vals: np.ndarray, perm: np.ndarray | None = None) -> sp.csr_matrix:
        """Scalar operator with (near-)zero entries dropped (host CSR); with
        ``perm`` (solver row i = node perm[i]) it is P A P^T."""
        g = self.fem.graph
        ip, ix, vv = compact_csr(g.indptr, g.indices, vals)
        S = sp.csr_matrix((vv, ix, ip), shape=(self.N, self.N))
        if perm is not None:
            S = S[perm][:, perm].tocsr()
        S.sort_indices()
        return S

    def compact(self, vals: np.ndarray, perm: np.ndarray | None = None) -> DeviceEll:
        """``compact_host`` as ELL for the CG kernels."""
        S = self.compact_host(vals, perm)
        return DeviceEll.from_csr(S.indptr, S.indices, S.data, self.N, self.device)

    def solver_order(self) -> np.ndarray:
        """Row order for the CG solves: the node numbering or its reverse
        Cuthill-McKee order, whichever has fewer matrix entries crossing the
        eight contiguous row ranges the CG kernel gives the eight XCDs (those
        entries are fetched from another XCD's L2 every iteration)."""
        if not hasattr(self, "_order"):
            g = self.fem.graph
            N = self.N
            r = np.repeat(np.arange(N), np.diff(g.indptr))
            adj = sp.csr_matrix((np.ones(g.nnz, dtype=np.int8), (r, g.indices)), shape=(N, N))
            rcm = np.asarray(reverse_cuthill_mckee(adj, symmetric_mode=True), dtype=np.int64)
            pos = np.empty(N, dtype=np.int64)
            pos[rcm] = np.arange(N)
            chunk = (N + 7) // 8
            cross_id = int(np.count_nonzero(r // chunk != g.indices // chunk))
            cross_rcm = int(np.count_nonzero(pos[r] // chunk != pos[g.indices] // chunk))
            self._order = rcm if cross_rcm < cross_id else np.arange(N, dtype=np.int64)
        return self._order


class DemagFieldLindholmGPU:
    """FEM/BEM demag field on one MI355X (see module docstring)."""

    def __init__(self, mesh, V=None, V1=None, Ms=8e5, VolN=None, serial_mesh_path=None,
                 cache_dir=None, cache_prefix="Bcache_lindholm_ownedrows", cache_tag=None,
                 map_tol=1e-10, ksp_rtol_u1=1e-6, ksp_rtol_u2=1e-6, incident_policy="skip",
                 volume_scale=1e27, B_dtype=np.float64, boundary_backend="dense",
                 dense_gpu_max_bytes=512 * 1024 ** 2, hmatrix_cache_path=None,
                 hmatrix_cache_dir="./hmatrix_cache", hmatrix_epsilon=1e-5, hmatrix_eta=2.0,
                 hmatrix_leaf_size=64, hmatrix_compressor="fullaca", hmatrix_max_rank=None,
                 hmatrix_max_temporary_block_bytes=256 * 1024 ** 2, hmatrix_force_rebuild=False,
                 hmatrix_cache_compressed=False, hmatrix_build_backend="gpu",
                 hmatrix_build_device_id=0, hmatrix_build_threads_per_block=128,
                 hmatrix_gpu_entry_kernel="auto", hmatrix_gpu_fast_math=False,
                 hmatrix_gpu_index_cache_max_bytes=128 * 1024 ** 2,
                 hmatrix_gpu_memory_fraction=0.75, hmatrix_gpu_fallback_to_cpu=True,
                 hmatrix_gpu_aca_residual_check_interval=1, hmatrix_gpu_rsvd_initial_rank=16,
                 hmatrix_gpu_rsvd_oversampling=8, hmatrix_gpu_rsvd_power_iterations=1,
                 hmatrix_gpu_progress_interval_seconds=5.0, hmatrix_threads_per_block=128,
                 hmatrix_use_fp32=True, hmatrix_direct_gpu_pack=False,
                 hmatrix_release_build_memory=True, hmatrix_verbose=False,
                 *, ksp_atol=1e-12, ksp_max_it_u1=2000, ksp_max_it_u2=12000,
                 warm_start=False, device=None, operators: OperatorSet | None = None,
                 geometry: BoundaryGeometry | None = None, hmatrix_prebuilt=None,
                 hmatrix_near_fp32=False, hmatrix_near_in_z="auto", cg_reorder=True, cg_speculative=True,
                 ksp_pc_type="gamg",
                 ksp_pc_type_u2=None,
                 amg_fp32=None, amg_max_coarse=None, amg_theta=None, cg_mixed=False, cg_refinements=1,
                 cg_rtol_inner=None, verbose=False):
        """Keywords as the reference's constructor (Demag_Lindholm_GPU.py:101-155),
        plus, keyword-only: the solver settings the reference hard-codes
        (ksp_atol, max_it), ``ksp_pc_type`` / ``ksp_pc_type_u2`` in {"gamg",
        "jacobi", "auto"} for the u1 / u2 preconditioner (u2 defaults to u1's;
        "auto" times both on a fixed right side at set-up and keeps the faster),
        ``amg_fp32`` (FP32 AMG operator values; default: FP32 when the boundary
        operator is stored below FP64), ``cg_mixed`` (the mixed-precision
        Poisson solves: FP32 inner CG, FP64 iterative refinement; off unless
        asked for): the FP32 solve to ``cg_rtol_inner`` (default rtol / 2),
        then up to ``cg_refinements`` FP64 residual checks, each followed
        by a correction solve when it fails; and prebuilt geometry / operators."""
        if not isinstance(mesh, TetMesh):
            raise TypeError("mesh must be a smx.mesh.TetMesh")
        self.mesh = mesh
        self.Ms = float(Ms)
        self.mu0 = MU0
        self.volume_scale = float(volume_scale)
        self.ops = operators or OperatorSet(mesh, device)
        self.device = self.ops.device
        dev = self.device
        fem = self.ops.fem
        self.N = fem.N
        self.local_dofs = self.N
        self.local_size = 3 * self.N
        self.last_its_u1 = 1
        self.last_its_u2 = 1
        self.ksp_rtol_u1 = float(ksp_rtol_u1)
        self.ksp_rtol_u2 = float(ksp_rtol_u2)
        self.ksp_atol = float(ksp_atol)
        self.ksp_max_it_u1 = int(ksp_max_it_u1)
        self.ksp_max_it_u2 = int(ksp_max_it_u2)
        self.warm_start = bool(warm_start)
        self.verbose = bool(verbose)
        self.converged = True
        # speculative solves: once both iteration counts are known, each solve
        # launches last count + margin iterations without a host wait and the
        # status of both is checked once, after the kernel that consumes them
        self.cg_speculative = bool(cg_speculative)
        self._its_known = False
        self._pending_m = None
        self.speculative_redos = 0
        self._status_host = torch.zeros(16, dtype=torch.float64).pin_memory()

        if VolN is not None:
            vol = np.asarray(VolN, dtype=np.float64).reshape(-1, 3)[:, 0]
            self.vol_scalar_gpu = to_dev(vol, dev)
        else:
            self.vol_scalar_gpu = self.ops.vol
        self.inv_vol_scalar_gpu = 1.0 / self.vol_scalar_gpu

        # single connected component: nullspace = constants (ns_u1)
        g = fem.graph
        r = np.repeat(np.arange(self.N), np.diff(g.indptr))
        adj = sp.csr_matrix((np.ones(g.nnz, dtype=np.int8), (r, g.indices)), shape=(self.N, self.N))
        ncomp, _ = connected_components(adj, directed=False)
        if ncomp != 1:
            raise NotImplementedError(
                f"mesh has {ncomp} connected components; the device path projects one constant nullspace")

        # ---------------- FEM operators ----------------
        self.Div = self.ops.graph_values(self.Ms * fem.div, 3)
        self.G = self.ops.graph_values(fem.grad, 3)
        # CG operators in the solver's row order (P A P^T); vectors stay in node order
        perm = self.ops.solver_order() if cg_reorder else None
        self.cg_perm = to_dev(perm.astype(np.int32), dev) if perm is not None else None
        ident = np.arange(self.N)
        sel = perm if perm is not None else ident
        # u2 in an order of its own: the interior rows first (in the order of
        # u1), the Dirichlet rows of A_bc after them.  The interior rows of
        # A_bc do not couple to the Dirichlet columns, so the solver iterates
        # the interior block alone and writes the Dirichlet rows' data once
        # (smx_pcg_* trailing identity rows): the same iterates as the full
        # system, without streaming its identity rows every iteration.
        isb = fem.is_boundary[sel]
        self.n_int = int(np.count_nonzero(~isb))
        split = 0 < self.n_int < self.N
        sel2 = np.concatenate([sel[~isb], sel[isb]]) if split else sel
        n2 = self.n_int if split else self.N
        self.cg_perm2 = to_dev(sel2.astype(np.int32), dev)
        d1 = self._diag(fem.stiff)
        d2 = self._diag(fem.stiffness_bc())
        self.dinv_u1 = to_dev(1.0 / d1[sel], dev)
        self.dinv_u2 = to_dev(1.0 / d2[sel2][:n2], dev)
        pc1 = str(ksp_pc_type).strip().lower()
        pc2 = pc1 if ksp_pc_type_u2 is None else str(ksp_pc_type_u2).strip().lower()
        if pc1 not in ("gamg", "jacobi", "auto") or pc2 not in ("gamg", "jacobi", "auto"):
            raise ValueError("ksp_pc_type / ksp_pc_type_u2 must be 'gamg', 'jacobi' or 'auto'")
        A1h = self.ops.compact_host(fem.stiff, perm)
        A2f = self.ops.compact_host(fem.stiffness_bc(), sel2)
        if split and A2f[:n2, n2:].nnz:
            raise RuntimeError("interior rows of the Dirichlet operator couple to Dirichlet columns")
        A2h = A2f[:n2, :n2].tocsr()
        # plain P A P^T: the operator of the AMG-CG kernels (u1: also the
        # lifting operator); the Jacobi CG kernels take the solver form A D^-1
        # with the diagonal in slot 0.  u2 lifts b = -A_raw g with the raw
        # stiffness rows of its iterated rows, in its own order.
        self.A_raw = DeviceEll.from_csr(A1h.indptr, A1h.indices, A1h.data, self.N, dev)
        R2 = self.ops.compact_host(fem.stiff, sel2)[:n2].tocsr()
        raw2 = DeviceEll.from_csr(R2.indptr, R2.indices, R2.data, n2, dev)
        A_bc = DeviceEll.from_csr(A2h.indptr, A2h.indices, A2h.data, n2, dev)
        self._cg = {}
        for key, plain, dinv, pc, n, raw, cperm in (
                ("u1", self.A_raw, self.dinv_u1, pc1, self.N, self.A_raw, self.cg_perm),
                ("u2", A_bc, self.dinv_u2, pc2, n2, raw2, self.cg_perm2)):
            self._cg[key] = {"plain": plain, "dinv": dinv, "request": pc, "pc": pc,
                             "jacobi": solver_operator(plain, dinv) if pc in ("jacobi", "auto") else None,
                             "amg": None, "n": n, "nd": self.N - n, "raw": raw, "perm": cperm}
        self._sel = (sel, sel2)
        from .amg import MAX_COARSE, THETA
        self._amg_host = (A1h, A2h, int(MAX_COARSE if amg_max_coarse is None else amg_max_coarse),
                          float(THETA if amg_theta is None else amg_theta))

        # ---------------- boundary operator ----------------
        self.opB = geometry or build_boundary_geometry(mesh.X, mesh.tets, incident_policy)
        geo = self.opB
        self.Nb = geo.Nb
        self.bverts_gpu = to_dev(geo.bverts.astype(np.int32), dev)
        self.bpos_gpu = to_dev(geo.bpos.astype(np.int32), dev)
        # boundary positions in each solve's row order
        for key, order in zip(("u1", "u2"), self._sel):
            bp = geo.bpos[order]
            c = self._cg[key]
            if key == "u2" and c["nd"] and not (np.all(bp[:c["n"]] < 0) and np.all(bp[c["n"]:] >= 0)):
                raise RuntimeError("the boundary geometry and the Dirichlet rows of the FEM system differ")
            c["bpos"] = to_dev(bp.astype(np.int32), dev)
        del self._sel
        self.diag_jump_gpu = to_dev(geo.diag_jump, dev)
        self.B_dtype = _torch_dtype(B_dtype)

        requested = str(boundary_backend).strip().lower()
        if requested not in {"dense", "hmatrix", "auto"}:
            raise ValueError("boundary_backend must be 'dense', 'hmatrix', or 'auto'.")
        dense_bytes = self.Nb * self.Nb * torch.tensor([], dtype=self.B_dtype).element_size()
        if requested == "auto":
            requested = "dense" if dense_bytes <= int(dense_gpu_max_bytes) else "hmatrix"
        self.boundary_backend = requested
        self.B_dense_gpu = None
        self.B_hmatrix_gpu = None
        self.B_hmatrix_cpu = None
        self.hmatrix_cache_path = None
        t0 = time.perf_counter()
        if requested == "dense":
            self.B_dense_gpu, self.B_ld = assemble_dense_B(geo, dev, self.B_dtype)
        elif hmatrix_prebuilt is not None:
            from .hmatrix import HMatrixGPUPackedFused
            self.B_hmatrix_cpu = hmatrix_prebuilt
            self.B_hmatrix_gpu = HMatrixGPUPackedFused(hmatrix_prebuilt, use_fp32=bool(hmatrix_use_fp32),
                                                       device=dev, near_fp32=bool(hmatrix_near_fp32),
                                                       near_in_z=hmatrix_near_in_z,
                                                       boundary_gather=geo.bverts)
        else:
            from .hmatrix import build_device_hmatrix
            self.B_hmatrix_cpu, self.B_hmatrix_gpu = build_device_hmatrix(
                geo, dev, epsilon=float(hmatrix_epsilon), eta=float(hmatrix_eta),
                leaf_size=int(hmatrix_leaf_size), max_rank=hmatrix_max_rank,
                max_temporary_block_bytes=int(hmatrix_max_temporary_block_bytes),
                use_fp32=bool(hmatrix_use_fp32), cache_path=hmatrix_cache_path,
                force_rebuild=bool(hmatrix_force_rebuild), verbose=bool(hmatrix_verbose),
                near_fp32=bool(hmatrix_near_fp32), near_in_z=hmatrix_near_in_z, compressor=str(hmatrix_compressor),
                gpu_device_id=int(hmatrix_build_device_id), gpu_memory_fraction=float(hmatrix_gpu_memory_fraction),
                gpu_fallback_to_cpu=bool(hmatrix_gpu_fallback_to_cpu),
                gpu_aca_residual_check_interval=int(hmatrix_gpu_aca_residual_check_interval),
                gpu_rsvd_initial_rank=int(hmatrix_gpu_rsvd_initial_rank),
                gpu_rsvd_oversampling=int(hmatrix_gpu_rsvd_oversampling),
                gpu_rsvd_power_iterations=int(hmatrix_gpu_rsvd_power_iterations),
                gpu_progress_interval_seconds=float(hmatrix_gpu_progress_interval_seconds))
        torch.cuda.synchronize(dev)
        self.setup_boundary_seconds = time.perf_counter() - t0

        # ---------------- AMG hierarchies (PETSc GAMG counterpart) ----------------
        if amg_fp32 is None:
            # follow the boundary operator: a preconditioner in FP32 when the
            # boundary operator is stored below FP64 (the outer CG stays FP64)
            if self.boundary_backend == "dense":
                amg_fp32 = self.B_dtype != torch.float64
            else:
                amg_fp32 = bool(hmatrix_use_fp32)
        self.amg_fp32 = bool(amg_fp32)
        # mixed-precision Poisson solves (FP32 inner CG, FP64 refinement)
        self.cg_mixed = bool(cg_mixed)
        self.cg_refinements = int(cg_refinements)
        self.cg_rtol_inner = None if cg_rtol_inner is None else float(cg_rtol_inner)
        if self.cg_mixed:
            for c in self._cg.values():
                c["plain32"] = c["plain"].vals.float()
                c["jacobi32"] = c["jacobi"].vals.float() if c["jacobi"] is not None else None
        # iteration counts of the last solves: (earlier inner solves, last solve)
        self._its_split = {"u1": (0, 1), "u2": (0, 1)}
        A1h, A2h, mc, th = self._amg_host
        del self._amg_host
        t1 = time.perf_counter()
        for key, Ah in (("u1", A1h), ("u2", A2h)):
            if self._cg[key]["request"] in ("gamg", "auto"):
                from .amg import AmgHierarchy
                self._cg[key]["amg"] = AmgHierarchy(Ah, dev, fp32=self.amg_fp32, max_coarse=mc, theta=th)
        self.setup_amg_seconds = time.perf_counter() - t1

        # ---------------- per-evaluation buffers ----------------
        N = self.N
        z = lambda n: torch.zeros(n, dtype=torch.float64, device=dev)
        L = _capi.lib()
        self.b_u1 = z(N)
        self.u1 = z(N)
        self.u2 = z(N)
        self.xb = z(self.Nb)
        self.y_rows = z(self.Nb)
        self.red = reduce_workspace(3 * N + 64, dev)
        self.b_part = z(self.Div.blocks())
        nws = max(int(L.smx_pcg_ws_size(N)), int(L.smx_pcg_amg_ws_size(N)))
        self.ws1 = z(nws)
        self.ws2 = z(nws)
        self._have_u = False
        self._energy_out = z(1)
        if "auto" in (pc1, pc2):
            self._choose_preconditioners()

    # ------------------------------------------------------------------
    @property
    def pc_type(self) -> str:
        """Preconditioner of the u1 solve ("gamg" or "jacobi")."""
        return self._cg["u1"]["pc"]

    @property
    def pc_type_u2(self) -> str:
        return self._cg["u2"]["pc"]

    @property
    def amg_u1(self):
        return self._cg["u1"]["amg"]

    @property
    def amg_u2(self):
        return self._cg["u2"]["amg"]

    @property
    def A_u1(self) -> DeviceEll:
        """The operator the u1 solve iterates on (solver form for Jacobi)."""
        c = self._cg["u1"]
        return c["jacobi"] if c["pc"] == "jacobi" else c["plain"]

    @property
    def A_u2(self) -> DeviceEll:
        c = self._cg["u2"]
        return c["jacobi"] if c["pc"] == "jacobi" else c["plain"]

    def _choose_preconditioners(self, reps: int = 3):
        """For every solve asked to be "auto": solve a fixed right side (Div m of
        a seeded random unit m for u1, seeded random boundary data for u2) with
        each preconditioner, time it with device events, keep the faster.  Both
        solve to the same tolerance; the choice only moves the time."""
        rng = np.random.default_rng(0)
        g = rng.standard_normal((self.N, 3))
        m = to_dev((g / np.linalg.norm(g, axis=1)[:, None]).ravel(), self.device)
        self.Div.spmv(m, self.b_u1, kind=1, partials=self.b_part)
        gb = to_dev(rng.standard_normal(self.Nb), self.device)
        self.choice_seconds = {}
        for key in ("u1", "u2"):
            c = self._cg[key]
            if c["request"] != "auto":
                continue
            best = None
            for pc in ("jacobi", "gamg"):
                c["pc"] = pc
                run = (lambda: self.solve("u1", 0, self.u1, self.ws1, b=self.b_u1, b_part=self.b_part,
                                          expected=its)) if key == "u1" else \
                      (lambda: self.solve("u2", 1, self.u2, self.ws2, gb=gb, project=0, expected=its))
                its = 0
                its = run()
                ts = []
                for _ in range(reps):
                    e0 = torch.cuda.Event(enable_timing=True)
                    e1 = torch.cuda.Event(enable_timing=True)
                    e0.record()
                    run()
                    e1.record()
                    e1.synchronize()
                    ts.append(e0.elapsed_time(e1) * 1e-3)
                self.choice_seconds[(key, pc)] = min(ts)
                if best is None or min(ts) < best[0]:
                    best = (min(ts), pc)
            c["pc"] = best[1]
        self.u1.zero_()
        self.u2.zero_()
        self.converged = True

    # ------------------------------------------------------------------
    def _diag(self, vals):
        g = self.ops.fem.graph
        r = np.repeat(np.arange(self.N), np.diff(g.indptr))
        d = np.zeros(self.N)
        m = r == g.indices
        d[r[m]] = vals[m]
        return d

    @property
    def u1_sum_ptr(self) -> int:
        """Device address of sum(u1) left by the u1 solve."""
        return self.ws1.data_ptr() + 8 * _capi.PCG_SUMX

    def solve(self, op: str, mode: int, x: torch.Tensor, ws: torch.Tensor, *, b=None, b_part=None,
              project: int = 1, gb=None, rtol=None, max_it=None, expected: int = 0, warm: bool = False,
              speculative: bool = False) -> int:
        """One preconditioned CG solve on operator ``op`` ("u1": the Neumann
        stiffness; "u2": the stiffness with Dirichlet identity rows) with that
        operator's preconditioner.  mode 0: A x = b - mean(b)*project (b and
        x in node order); mode 1: the Dirichlet lifting of boundary data gb.
        Leaves sum(x) in ws[PCG_SUMX].  Returns the iteration count, or -1
        for a speculative launch (expected + margin iterations enqueued, no
        wait; the status block ws[0:8] tells later whether it converged)."""
        c = self._cg[op]
        spec = bool(speculative) and expected > 0
        if rtol is None:
            rtol = self.ksp_rtol_u1 if op == "u1" else self.ksp_rtol_u2
        if max_it is None:
            max_it = self.ksp_max_it_u1 if op == "u1" else self.ksp_max_it_u2
        L = _capi.lib()
        its = C.c_int32(0)
        rn = C.c_double(0.0)
        sums = (C.c_double * 1)()
        raw = c["raw"]
        n, nd = c["n"], c["nd"]
        nbp = 0 if b_part is None else b_part.numel()
        amg_pc = c["pc"] == "gamg"
        mixed = None
        if self.cg_mixed:
            v32 = c["plain32"] if amg_pc else c["jacobi32"]
            # default: the FP32 solve runs to rtol / 2, so that the FP64 residual
            # (which drifts from the FP32 recurrence by ~1e-7 of ||b||) meets rtol
            # and the check does not ask for a correction solve
            rin = 0.5 * float(rtol) if self.cg_rtol_inner is None else self.cg_rtol_inner
            mixed = _capi.CGMixed(v32.data_ptr(), min(max(rin, 1e-300), 0.5), self.cg_refinements, 0)
        if amg_pc:
            A, amg = c["plain"], c["amg"]
            args = [n, nd, A.width, A.ld, _capi.ptr(A.cols), _capi.ptr(A.vals), mode,
                    _capi.ptr(b), _capi.ptr(b_part), nbp, int(project), raw.width, raw.ld,
                    _capi.ptr(raw.cols), _capi.ptr(raw.vals), _capi.ptr(c["bpos"]),
                    _capi.ptr(gb), _capi.ptr(x), int(warm), _capi.ptr(c["perm"]),
                    C.addressof(amg.c_levels), amg.nlev, amg.val_dtype, float(rtol),
                    float(self.ksp_atol), int(max_it), -1 if spec else 8, int(expected),
                    _capi.ptr(ws), C.byref(its), C.byref(rn), sums]
            if mixed is None:
                rc, what = L.smx_pcg_amg(*args, _capi.stream_ptr()), "smx_pcg_amg"
            else:
                rc, what = L.smx_pcg_amg_mixed(*args, C.byref(mixed), _capi.stream_ptr()), "smx_pcg_amg_mixed"
        else:
            A = c["jacobi"]
            args = [n, nd, A.width, A.ld, _capi.ptr(A.cols), _capi.ptr(A.vals),
                    _capi.ptr(c["dinv"]), mode, _capi.ptr(b), _capi.ptr(b_part), nbp,
                    int(project), raw.width, raw.ld, _capi.ptr(raw.cols), _capi.ptr(raw.vals),
                    _capi.ptr(c["bpos"]), _capi.ptr(gb), _capi.ptr(x), int(warm),
                    _capi.ptr(c["perm"]), float(rtol), float(self.ksp_atol), int(max_it),
                    -2 if spec else 16, int(expected), _capi.ptr(ws), C.byref(its), C.byref(rn), sums]
            if mixed is None:
                rc, what = L.smx_pcg_jacobi(*args, _capi.stream_ptr()), "smx_pcg_jacobi"
            else:
                rc, what = L.smx_pcg_jacobi_mixed(*args, C.byref(mixed), _capi.stream_ptr()), "smx_pcg_jacobi_mixed"
        if rc == _capi.SMX_ERR_NOT_CONVERGED:
            # The reference never checks the KSP converged reason; keep going.
            self.converged = False
        else:
            _capi.check(rc, what)
        if its.value >= 0 and mixed is not None:
            # the waiting form synchronized the stream: split the count by solve
            st = ws[:8].cpu().numpy()
            self._its_split[op] = (int(st[_capi.PCG_ITERS0]), int(st[_capi.PCG_ITERS]))
        elif its.value >= 0:
            self._its_split[op] = (0, int(its.value))
        return int(its.value)

    def boundary_matvec(self, src: torch.Tensor, y: torch.Tensor, gather=None, shift_ptr=None,
                        shift_scale=0.0):
        """y = B x with x = src[gather] - (*shift)*shift_scale (dense or H-matrix)."""
        L = _capi.lib()
        if self.boundary_backend == "dense":
            _capi.check(L.smx_dense_bmatvec(_capi.ptr(self.B_dense_gpu),
                                            _capi.DTYPE_CODE[self.B_dense_gpu.dtype], self.Nb,
                                            self.B_ld, _capi.ptr(src), _capi.ptr(gather),
                                            shift_ptr, float(shift_scale), _capi.ptr(self.xb),
                                            _capi.ptr(y), _capi.stream_ptr()), "smx_dense_bmatvec")
        else:
            self.B_hmatrix_gpu.matvec_from(src, y, gather=gather, shift_ptr=shift_ptr,
                                           shift_scale=shift_scale)

    def solve_potential(self, m_vec: torch.Tensor, speculative=None):
        """u1, u2 for magnetization m (3N); leaves sum(u1) at u1_sum_ptr.
        With speculative solves (the default once both iteration counts are
        known) nothing waits for the device: the caller enqueues what reads
        u1/u2 and then calls finish_potential, which checks both solves and,
        if one had not converged within its launches, solves again waiting
        and returns True (the caller then recomputes what it enqueued)."""
        _capi.require_device(m_vec, "m_vec", numel=self.local_size)
        if self._pending_m is not None:
            self.finish_potential()
        if speculative is None:
            speculative = self.cg_speculative and self._its_known
        warm = self.warm_start and self._have_u
        self.Div.spmv(m_vec, self.b_u1, kind=1, partials=self.b_part)
        its1 = self.solve("u1", 0, self.u1, self.ws1, b=self.b_u1, b_part=self.b_part,
                          expected=self.last_its_u1, warm=warm, speculative=speculative)
        self.boundary_matvec(self.u1, self.y_rows, gather=self.bverts_gpu,
                             shift_ptr=self.u1_sum_ptr, shift_scale=1.0 / self.N)
        its2 = self.solve("u2", 1, self.u2, self.ws2, gb=self.y_rows, project=0,
                          expected=self.last_its_u2, warm=warm, speculative=speculative)
        self._have_u = True
        if its1 < 0 or its2 < 0:
            self._status_host[:8].copy_(self.ws1[:8], non_blocking=True)
            self._status_host[8:].copy_(self.ws2[:8], non_blocking=True)
            self._pending_m = m_vec
        else:
            self.last_its_u1, self.last_its_u2 = its1, its2
            self._its_known = True

    def finish_potential(self) -> bool:
        """Check the speculative solves of the last solve_potential (one wait
        for the stream); returns True when they had to be solved again."""
        m_vec = self._pending_m
        if m_vec is None:
            return False
        self._pending_m = None
        torch.cuda.current_stream().synchronize()
        st = self._status_host.numpy()
        redo = False
        for off, name in ((0, "u1"), (8, "u2")):
            done, its, fail = st[off + _capi.PCG_DONE], st[off + _capi.PCG_ITERS], st[off + _capi.PCG_FAIL]
            if fail == 3.0:
                raise ValueError(f"{name} CG operator is not in solver form")
            if done == 0.0:
                redo = True
                continue
            if fail != 0.0:
                self.converged = False      # the reference never checks KSP convergence: keep going
            its0 = st[off + _capi.PCG_ITERS0] if self.cg_mixed else 0.0
            self._its_split[name] = (int(its0), int(its))
            setattr(self, "last_its_" + name, int(its0 + its))
        if redo:
            self.speculative_redos += 1
            self.solve_potential(m_vec, speculative=False)
        return redo

    def field_args(self, m_vec, out_h_demag=None) -> _capi.FieldArgs:
        a = _capi.FieldArgs()
        a.n = self.N
        a.slice_ptr = _capi.ptr(self.G.slice_ptr)
        a.cols = _capi.ptr(self.G.cols)
        a.grad_vals = _capi.ptr(self.G.vals)
        a.u1 = _capi.ptr(self.u1)
        a.u2 = _capi.ptr(self.u2)
        a.u1_sum = self.u1_sum_ptr
        a.inv_vol = _capi.ptr(self.inv_vol_scalar_gpu)
        a.m = _capi.ptr(m_vec)
        a.h_demag = _capi.ptr(out_h_demag)
        return a

    def compute_vec(self, m_vec: torch.Tensor, out_vec: torch.Tensor):
        """H_demag(m) into out_vec (Demag_Lindholm_GPU.py:643-670)."""
        _capi.require_device(out_vec, "out_vec", numel=self.local_size)
        self.solve_potential(m_vec)
        a = self.field_args(m_vec, out_vec)
        _capi.check(_capi.lib().smx_field_llg_rhs(C.byref(a), _capi.stream_ptr()), "smx_field_llg_rhs")
        if self.finish_potential():
            _capi.check(_capi.lib().smx_field_llg_rhs(C.byref(a), _capi.stream_ptr()), "smx_field_llg_rhs")

    def Energy_lumped_gpu(self, m_vec: torch.Tensor, h_vec: torch.Tensor) -> float:
        """E = -1/2 mu0 Ms sum_i V_i m_i.H_i / volume_scale (:706-724)."""
        L = _capi.lib()
        _capi.check(L.smx_dot3_weighted(self.N, _capi.ptr(m_vec), _capi.ptr(h_vec),
                                        _capi.ptr(self.vol_scalar_gpu), _capi.ptr(self._energy_out),
                                        _capi.ptr(self.red), _capi.stream_ptr()), "smx_dot3_weighted")
        return -0.5 * self.mu0 * self.Ms * float(self._energy_out.item()) / self.volume_scale

    def boundary_memory_report(self) -> dict:
        if self.boundary_backend == "dense":
            nbytes = int(self.B_dense_gpu.numel() * self.B_dense_gpu.element_size())
            return {"backend": "dense", "Nb": self.Nb, "dtype": str(self.B_dense_gpu.dtype),
                    "accounted_gpu_bytes": nbytes, "accounted_gpu_mib": nbytes / 1024 ** 2}
        rep = self.B_hmatrix_gpu.memory_report()
        rep["Nb"] = self.Nb
        rep["cache_path"] = self.hmatrix_cache_path
        return rep

    def print_boundary_memory_report(self):
        rep = self.boundary_memory_report()
        print(f"[Demag Lindholm GPU] backend={rep['backend']}, Nb={rep['Nb']}, "
              f"GPU memory={rep.get('accounted_gpu_mib', 0.0):.3f} MiB", flush=True)


def assemble_dense_B(geo: BoundaryGeometry, device, dtype=torch.float64):
    """Dense Lindholm B on the device (diag = omega/4pi - 1), row-major with a
    leading dimension padded to 8 entries (16-byte rows for every dtype)."""
    L = _capi.lib()
    Nb, Nt = geo.Nb, geo.Nt
    ld = (Nb + 7) // 8 * 8
    B64 = torch.zeros((Nb, ld), dtype=torch.float64, device=device)
    Xb = to_dev(geo.Xb, device)
    geom = to_dev(geo.tri_table(), device)
    lid = to_dev(geo.tri_lid.astype(np.int32), device)
    rows = torch.arange(Nb, dtype=torch.int32, device=device)
    inc = geo.incidence()
    st = _capi.stream_ptr()
    T = _capi.LINDHOLM_TILE_COLS
    keep = []
    for c0 in range(0, Nb, T):
        cols = np.arange(c0, min(c0 + T, Nb))
        rel, rel_cols = geo.relevant_triangles(cols, inc)
        rel_d = to_dev(rel, device)
        rc_d = to_dev(rel_cols.reshape(-1), device)
        keep.append((rel_d, rc_d))
        _capi.check(L.smx_lindholm_fill(_capi.ptr(Xb), _capi.ptr(rows), Nb, _capi.ptr(geom), Nt,
                                        _capi.ptr(lid), _capi.ptr(rel_d), _capi.ptr(rc_d),
                                        int(rel.size), int(cols.size), B64.data_ptr() + 8 * c0,
                                        ld, 1, int(geo.incident_policy), st), "smx_lindholm_fill")
    diag = to_dev(geo.diag_jump, device)
    if dtype == torch.float64:
        _capi.check(L.smx_dense_finish(B64.data_ptr(), Nb, ld, _capi.ptr(diag), None, 0, st),
                    "smx_dense_finish")
        torch.cuda.synchronize(device)
        return B64, ld
    lo = torch.zeros((Nb, ld), dtype=dtype, device=device)
    _capi.check(L.smx_dense_finish(B64.data_ptr(), Nb, ld, _capi.ptr(diag), lo.data_ptr(),
                                   _capi.DTYPE_CODE[dtype], st), "smx_dense_finish")
    torch.cuda.synchronize(device)
    del B64
    return lo, ld


def dense_bmatmul(B: torch.Tensor, ld: int, X: torch.Tensor, out: torch.Tensor | None = None,
                  work: torch.Tensor | None = None) -> torch.Tensor:
    """Y = B X for a block of right-hand sides on the matrix cores.

    B: dense boundary matrix from ``assemble_dense_B`` (FP64, FP32 or BF16,
    row-major with leading dimension ``ld``).  X: (n, k) FP64 on the device,
    any k (processed 64 columns per launch).  Returns Y (n, k) FP64.  The
    storage precision of B sets the MFMA kind (BF16 / FP32 / FP64)."""
    n = int(B.shape[0])
    if X.dim() != 2 or X.shape[0] != n or X.dtype != torch.float64:
        raise ValueError(f"X must be ({n}, k) float64, got {tuple(X.shape)} {X.dtype}")
    _capi.require_device(X, "X")
    L = _capi.lib()
    code = _capi.DTYPE_CODE[B.dtype]
    k = int(X.shape[1])
    Xc = X.t().contiguous()                      # column j contiguous: X[:, j] at j*n
    Yc = torch.empty((k, n), dtype=torch.float64, device=X.device)
    need = int(L.smx_dense_bmatmul_ws_size(n, code))
    if work is None or work.numel() < need:
        work = torch.empty(max(need, 1), dtype=torch.uint8, device=X.device)
    st = _capi.stream_ptr()
    for j0 in range(0, k, 64):
        nr = min(64, k - j0)
        _capi.check(L.smx_dense_bmatmul(B.data_ptr(), code, n, ld, Xc.data_ptr() + 8 * j0 * n, n, nr,
                                        Yc.data_ptr() + 8 * j0 * n, n, work.data_ptr(), work.numel(), st),
                    "smx_dense_bmatmul")
    Y = Yc.t()
    if out is not None:
        out.copy_(Y)
        return out
    return Y.contiguous()
