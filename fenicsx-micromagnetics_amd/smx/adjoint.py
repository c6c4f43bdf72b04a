"""Micromagnetic energy with per-cell Ms and A, and its hand-written adjoint.

    E(m, Ms, A) = E_ex + E_an + E_d
    E_ex = 1e-9  sum_c A_c  m_c^T K^c m_c                 (consistent exchange, J)
    E_an = -1e-27 Ku sum_i V_i (m_i . n_i)^2              (Anisotropy_GPU.py:250-275)
    E_d  = 1/2 mu0 1e-27 b^T phi,  b = Div_Ms m,  phi = L b

L is the Fredkin-Koehler map b -> phi = u1 + u2 of the demag field (u1 Neumann
solve with the constant projected out, boundary operator B, harmonic lifting
E of the boundary data):  L = (I + E B Gamma) K  with K = Pi A^+ Pi.  For a
uniform Ms this E_d is exactly the reference's lumped demag energy
-1/2 mu0 Ms sum_i V_i m_i . H_i (Demag_Lindholm_GPU.py:706-724), because the
lumped gradient G equals -Div^T / Ms; E_ex equals the lumped exchange energy.

Gradient: dE_d/db = 1/2 mu0 1e-27 (phi + L^T b) with
    L^T b = K (b + Gamma^T B^T E^T b),
    E^T y = y_Gamma - (A_raw z)_Gamma,  A_bc z = y_interior (interior Dirichlet solve),
so one backward costs an interior Dirichlet CG solve, one transposed boundary
matvec (dense B^T or the H-matrix with its factors swapped) and one Neumann CG
solve; CG iterates are never stored (each solve restarts from zero).  Then
dE/dm = Div_Ms^T dE/db + exchange/anisotropy terms and the per-cell
derivatives follow from the same cell kernels.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from . import _capi
from .demag import DemagFieldLindholmGPU, OperatorSet
from .device import default_device, to_dev
from .fem import tet_gradients
from .mesh import TetMesh

MU0 = 4.0 * math.pi * 1e-7


class MicromagneticEnergy:
    def __init__(self, mesh: TetMesh, Ku=0.0, n_ani_vec=None, use_demag=True, demag_kwargs=None,
                 device=None, ksp_rtol=1e-12, operators: OperatorSet | None = None):
        self.mesh = mesh
        self.device = device or default_device()
        dev = self.device
        self.ops = operators or OperatorSet(mesh, dev)
        fem = self.ops.fem
        self.N = fem.N
        tets = np.asarray(mesh.tets, dtype=np.int64)
        self.T = int(tets.shape[0])
        vol, g = tet_gradients(np.asarray(mesh.X, dtype=np.float64), tets)
        cgv = (vol / 4.0)[:, None, None] * g                       # (T,4,3)
        ck = vol[:, None, None] * np.einsum("tak,tbk->tab", g, g)  # (T,4,4)
        self.cells = to_dev(tets.astype(np.int32).ravel(), dev)
        self.cgv = to_dev(cgv.ravel(), dev)
        self.ck = to_dev(ck.reshape(-1), dev)
        # node -> (cell, local vertex) incidence, ordered by (node, cell, local)
        flat = tets.ravel()
        order = np.argsort(flat, kind="stable")
        ptr = np.zeros(self.N + 1, dtype=np.int64)
        np.add.at(ptr, flat + 1, 1)
        self.inc_ptr = to_dev(np.cumsum(ptr), dev)
        self.inc = to_dev(order.astype(np.int32), dev)
        self.Ku = float(Ku)
        self.axis_gpu = None
        self.axis_u = (C.c_double * 3)(0.0, 0.0, 0.0)
        if self.Ku != 0.0:
            na = np.asarray(n_ani_vec, dtype=np.float64).reshape(-1, 3)
            na = na / np.maximum(np.linalg.norm(na, axis=1), 1e-300)[:, None]
            if na.shape[0] == 1:
                self.axis_u = (C.c_double * 3)(*na[0])
            else:
                self.axis_gpu = to_dev(na.ravel(), dev)
        self.demag = None
        if use_demag:
            kw = dict(ksp_rtol_u1=ksp_rtol, ksp_rtol_u2=ksp_rtol, ksp_atol=0.0, ksp_max_it_u1=20000,
                      ksp_max_it_u2=20000)
            kw.update(demag_kwargs or {})
            self.demag = DemagFieldLindholmGPU(mesh, None, None, 1.0, operators=self.ops, **kw)
            self.stiff_graph = self.ops.graph_values(fem.stiff, 1)
        z = lambda n: torch.zeros(n, dtype=torch.float64, device=dev)
        nbT = int(_capi.lib().smx_energy_blocks(self.T))
        nbN = int(_capi.lib().smx_energy_blocks(self.N))
        self.s = z(3 * self.T)
        self.f = z(12 * self.T)
        self.e = z(self.T)
        self.tc = z(3 * self.T)
        self.pT = z(nbT)
        self.pN = z(nbN)
        self.pN2 = z(nbN)
        self.b = z(self.N)
        self.phi = z(self.N)
        self.y = z(self.N)
        self.zv = z(self.N)
        self.Az = z(self.N)
        self.w = z(self.N)
        self.v = z(self.N)
        self.gb = z(self.N)
        self.ones = torch.ones(self.N, dtype=torch.float64, device=dev)
        self.grad_m = z(3 * self.N)
        self.grad_Ms = z(self.T)
        if self.demag is not None:
            self.q = z(self.demag.Nb)
            self.t = z(self.demag.Nb)
            self.ws3 = z(max(int(_capi.lib().smx_pcg_ws_size(self.N)), int(_capi.lib().smx_pcg_amg_ws_size(self.N))))
            self.tmp_t = z(((self.demag.Nb + 255) // 256) * self.demag.Nb) \
                if self.demag.boundary_backend == "dense" else None

    # ------------------------------------------------------------------
    def _L(self):
        return _capi.lib()

    def _check(self, rc, what):
        _capi.check(rc, what)

    def _solve(self, op, mode, x, ws, b=None, b_part=None, project=1, gb=None):
        """A solve of the demag field's operator ``op`` with its preconditioner,
        at the u1 tolerance (non-convergence is not an error, as there)."""
        d = self.demag
        return d.solve(op, mode, x, ws, b=b, b_part=b_part, project=project, gb=gb,
                       rtol=d.ksp_rtol_u1, max_it=d.ksp_max_it_u1)

    def _sum(self, partials) -> float:
        return float(np.sum(partials.cpu().numpy()))

    def _glue(self, op, n, x, y=None, idx=None, a=1.0, b=1.0, sx=None, sy=None, out=None):
        self._check(self._L().smx_vec_glue(op, n, _capi.ptr(x), _capi.ptr(y), _capi.ptr(idx), float(a), float(b),
                                           sx, sy, float(self.N), _capi.ptr(out), _capi.stream_ptr()),
                    "smx_vec_glue")

    def _check_inputs(self, m, Ms_cell, A_cell):
        _capi.require_device(m, "m", numel=3 * self.N)
        _capi.require_device(Ms_cell, "Ms_cell", numel=self.T)
        _capi.require_device(A_cell, "A_cell", numel=self.T)

    # ------------------------------------------------------------------
    def forward(self, m, Ms_cell, A_cell) -> dict:
        """Energies (J); keeps b and phi for the backward pass."""
        self._check_inputs(m, Ms_cell, A_cell)
        L = self._L()
        st = _capi.stream_ptr()
        self._check(L.smx_energy_cell_forward(self.T, _capi.ptr(self.cells), _capi.ptr(self.ck),
                                              _capi.ptr(Ms_cell), _capi.ptr(A_cell), _capi.ptr(m),
                                              _capi.ptr(self.s), _capi.ptr(self.f), _capi.ptr(self.e),
                                              _capi.ptr(self.pT), st), "smx_energy_cell_forward")
        out = {"E_ex": self._sum(self.pT)}
        self._check(L.smx_energy_node_grad(self.N, _capi.ptr(self.inc_ptr), _capi.ptr(self.inc), None, None,
                                           _capi.ptr(m), _capi.ptr(self.ops.vol), _capi.ptr(self.axis_gpu),
                                           self.axis_u, self.Ku, _capi.ptr(self.grad_m), _capi.ptr(self.pN), st),
                    "smx_energy_node_grad")
        out["E_an"] = self._sum(self.pN)
        if self.demag is not None:
            d = self.demag
            self._check(L.smx_energy_node_div(self.N, _capi.ptr(self.inc_ptr), _capi.ptr(self.inc),
                                              _capi.ptr(self.cgv), _capi.ptr(self.s), _capi.ptr(self.b),
                                              _capi.ptr(self.pN), st), "smx_energy_node_div")
            self.its_forward = [self._solve("u1", 0, d.u1, d.ws1, b=self.b, b_part=self.pN)]
            d.boundary_matvec(d.u1, d.y_rows, gather=d.bverts_gpu, shift_ptr=d.u1_sum_ptr,
                              shift_scale=1.0 / self.N)
            self.its_forward.append(self._solve("u2", 1, d.u2, d.ws2, gb=d.y_rows, project=0))
            self._glue(3, self.N, d.u1, d.u2, sx=d.u1_sum_ptr, out=self.phi)
            self._check(L.smx_dot_partials(self.N, _capi.ptr(self.b), _capi.ptr(self.phi), _capi.ptr(self.pN2), st),
                        "smx_dot_partials")
            out["E_d"] = 0.5 * MU0 * 1e-27 * self._sum(self.pN2)
        out["E_total"] = out["E_ex"] + out["E_an"] + out.get("E_d", 0.0)
        return out

    def lt_apply(self, y: torch.Tensor, out: torch.Tensor):
        """out = L^T y (L: b -> phi of the demag potential): an interior Dirichlet
        solve for E^T, the transposed boundary operator, and a Neumann solve.
        Leaves the iteration counts of the two solves in self.its_backward."""
        L = self._L()
        st = _capi.stream_ptr()
        d = self.demag
        N, Nb = self.N, d.Nb
        # E^T y: interior Dirichlet solve, then the boundary residual
        self._glue(0, N, y, idx=d.bpos_gpu, out=self.y)
        its = [self._solve("u2", 0, self.zv, self.ws3, b=self.y, project=0)]
        self.stiff_graph.spmv(self.zv, self.Az, 0)
        self._glue(1, Nb, y, self.Az, idx=d.bverts_gpu, out=self.q)
        # B^T
        if d.boundary_backend == "dense":
            self._check(L.smx_dense_bmatvec_t(_capi.ptr(d.B_dense_gpu), _capi.DTYPE_CODE[d.B_dense_gpu.dtype],
                                              Nb, d.B_ld, _capi.ptr(self.q), _capi.ptr(self.t),
                                              _capi.ptr(self.tmp_t), st), "smx_dense_bmatvec_t")
        else:
            d.B_hmatrix_gpu.rmatvec(self.q, self.t)
        # K (y + Gamma^T t), K = Pi A^+ Pi
        self._glue(2, N, y, self.t, idx=d.bpos_gpu, out=self.w)
        self._check(L.smx_dot_partials(N, _capi.ptr(self.w), _capi.ptr(self.ones), _capi.ptr(self.pN2), st),
                    "smx_dot_partials")
        its.append(self._solve("u1", 0, self.v, self.ws3, b=self.w, b_part=self.pN2, project=1))
        self._glue(3, N, self.v, self.v, a=0.0, b=1.0, sy=self.ws3.data_ptr() + 8 * _capi.PCG_SUMX, out=out)
        self.its_backward = its
        return out

    def backward(self, m, Ms_cell, A_cell):
        """(dE/dm, dE/dMs_cell, dE/dA_cell) at the state of the last forward()."""
        L = self._L()
        st = _capi.stream_ptr()
        tptr = None
        if self.demag is not None:
            d = self.demag
            # dE_d/db = c (phi + L^T b)
            self.lt_apply(self.b, self.gb)
            c = 0.5 * MU0 * 1e-27
            self._glue(3, self.N, self.phi, self.gb, a=c, b=c, out=self.gb)
            self._check(L.smx_energy_cell_backward(self.T, _capi.ptr(self.cells), _capi.ptr(self.cgv),
                                                   _capi.ptr(Ms_cell), _capi.ptr(self.gb), _capi.ptr(m),
                                                   _capi.ptr(self.tc), _capi.ptr(self.grad_Ms), st),
                        "smx_energy_cell_backward")
            tptr = _capi.ptr(self.tc)
        else:
            self.grad_Ms.zero_()
        self._check(L.smx_energy_node_grad(self.N, _capi.ptr(self.inc_ptr), _capi.ptr(self.inc), _capi.ptr(self.f),
                                           tptr, _capi.ptr(m), _capi.ptr(self.ops.vol), _capi.ptr(self.axis_gpu),
                                           self.axis_u, self.Ku, _capi.ptr(self.grad_m), None, st),
                    "smx_energy_node_grad")
        return self.grad_m.clone(), self.grad_Ms.clone(), self.e.clone()

    def value_and_grad(self, m, Ms_cell, A_cell):
        E = self.forward(m, Ms_cell, A_cell)
        return E, self.backward(m, Ms_cell, A_cell)


class MicromagneticEnergyFunction(torch.autograd.Function):
    """E = MicromagneticEnergyFunction.apply(m, Ms_cell, A_cell, model): total
    energy (J).  backward() re-runs the forward solves at the saved inputs (the
    potentials are recomputed, not kept across calls) and then the adjoint."""

    @staticmethod
    def forward(ctx, m, Ms_cell, A_cell, model: MicromagneticEnergy):
        m = m.detach().contiguous()
        Ms_cell = Ms_cell.detach().contiguous()
        A_cell = A_cell.detach().contiguous()
        E = model.forward(m, Ms_cell, A_cell)
        ctx.model = model
        ctx.save_for_backward(m, Ms_cell, A_cell)
        return torch.tensor(E["E_total"], dtype=torch.float64, device=m.device)

    @staticmethod
    def backward(ctx, grad_out):
        m, Ms_cell, A_cell = ctx.saved_tensors
        model = ctx.model
        model.forward(m, Ms_cell, A_cell)
        gm, gMs, gA = model.backward(m, Ms_cell, A_cell)
        return grad_out * gm, grad_out * gMs, grad_out * gA, None
