"""Adjoint through the LLG time integration (gradients w.r.t. per-cell Ms).

Forward model (exchange + Fredkin-Koehler demag + uniform H0, m on nodes, Ms
per cell), the reference's LLG right-hand side (llg_module_GPU.py:100-122,
1261-1285) with a nodal saturation magnetization:

    Ms_i = sum_{c ni i} |c|/4 Ms_c / V_i
    H    = cex (K m) + Hd(m; Ms) + H0,   cex_i = -2A/(mu0 Ms_i V_i) 1e18
    Hd   = V^-1 G phi,  phi = L b,  b = Div_Ms m   (b_a = int Ms m . grad l_a)
    r    = P (p m x H + alpha m x (m x H)) + S_i (1 - |m|^2) m,  S_i = s0 Ms_i

(uniform Ms reproduces EffectiveField.llg_rhs_vec exactly).  The Dormand-
Prince 5(4) steps (fixed dt, FSAL) run as torch tensor arithmetic around
``LLGRHSFunction``, whose backward is the hand-written vector-Jacobian
product:

    dr/dm^T lam = gm + K (cex gH) + Div_Ms^T L^T G^T (V^-1 gH)
    dr/dMs^T lam = d(y^T Div_Ms m)/dMs |_{y = L^T G^T V^-1 gH}
                   + nodal terms (1/Ms_i in cex, Ms_i in S) mapped to cells

with gH = dr/dH^T lam (node_rhs_vjp kernel), G^T = -Div_1 (the FEM gradient
is minus the transposed divergence), and L^T from the energy adjoint (an
interior Dirichlet solve, the transposed boundary operator, a Neumann
solve).  The potentials are recomputed in the backward (CG iterates are not
stored), and ``integrate`` checkpoints the state every ``checkpoint`` steps
(torch.utils.checkpoint), recomputing the stages in between.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch
from torch.utils.checkpoint import checkpoint

from . import _capi
from .adjoint import MU0, MicromagneticEnergy
from .device import sell_values, to_dev
from .fem import tet_gradients
from .mesh import TetMesh

# Dormand-Prince (1980) 5(4) tableau
_DP_A = [[], [1 / 5], [3 / 40, 9 / 40], [44 / 45, -56 / 15, 32 / 9],
         [19372 / 6561, -25360 / 2187, 64448 / 6561, -212 / 729],
         [9017 / 3168, -355 / 33, 46732 / 5247, 49 / 176, -5103 / 18656]]
_DP_B = [35 / 384, 0.0, 500 / 1113, 125 / 192, -2187 / 6784, 11 / 84]


class LLGAdjointModel:
    def __init__(self, mesh: TetMesh, A_ex=1.3e-11, alpha=0.02, gamma=2.211e5, do_precess=1, H0=(0.0, 0.0, 0.0),
                 use_demag=True, demag_kwargs=None, ksp_rtol=1e-10, device=None):
        self.energy = MicromagneticEnergy(mesh, Ku=0.0, use_demag=use_demag, demag_kwargs=demag_kwargs,
                                          device=device, ksp_rtol=ksp_rtol)
        e = self.energy
        self.device = e.device
        dev = self.device
        self.ops = e.ops
        fem = self.ops.fem
        self.N, self.T = e.N, e.T
        self.A_ex = float(A_ex)
        self.alpha = float(alpha)
        self.gamma = float(gamma)
        self.do_precess = float(do_precess)
        self.P = -self.gamma / (1.0 + self.alpha ** 2)
        self.s0 = 0.5 * self.gamma / (1.0 + self.alpha ** 2)
        self.h0 = (C.c_double * 3)(*[float(v) for v in H0])
        vol, _ = tet_gradients(np.asarray(mesh.X, dtype=np.float64), np.asarray(mesh.tets, dtype=np.int64))
        self.qvol = to_dev(vol / 4.0, dev)
        g = fem.graph
        self.kex = self.ops.graph_values(fem.stiff, 1)                 # stiffness in SELL slots
        rows = np.repeat(np.arange(self.N, dtype=np.float64), np.diff(g.indptr)) + 1.0
        self.slot_row = to_dev((sell_values(g.indptr, rows, self.N) - 1.0).astype(np.int32), dev)
        self.ex_vals = torch.zeros_like(self.kex.vals)
        z = lambda n: torch.zeros(n, dtype=torch.float64, device=dev)
        self.Ms_node = z(self.N)
        self.cex = z(self.N)
        self.S = z(self.N)
        self.h = z(3 * self.N)
        self.hd = z(3 * self.N)
        self.gm = z(3 * self.N)
        self.gk = z(3 * self.N)
        self.gd = z(3 * self.N)
        self.kg = z(3 * self.N)
        self.gnode = z(self.N)
        self.yv = z(self.N)
        self.lt = z(self.N)
        self.gMs = z(self.T)
        self.zero_A = z(self.T)
        self._Ms_key = None
        self.rhs_evals = 0
        self.vjp_evals = 0

    # ------------------------------------------------------------------
    def _L(self):
        return _capi.lib()

    def set_Ms(self, Ms_cell: torch.Tensor):
        """Nodal Ms, exchange prefactors and stabilization for this Ms field."""
        # the cache holds a reference to the tensor, so no other tensor can be
        # given its storage while the key is alive (in-place updates bump _version)
        key = (Ms_cell.data_ptr(), Ms_cell._version)
        if self._Ms_key is not None and key == self._Ms_key[0]:
            return
        _capi.require_device(Ms_cell, "Ms_cell", numel=self.T)
        e, L, st = self.energy, self._L(), _capi.stream_ptr()
        _capi.check(L.smx_cell_to_node(self.N, _capi.ptr(e.inc_ptr), _capi.ptr(e.inc), _capi.ptr(self.qvol),
                                       _capi.ptr(self.ops.inv_vol), _capi.ptr(Ms_cell), _capi.ptr(self.Ms_node), st),
                    "smx_cell_to_node")
        # cex = -2A 1e18 / (mu0 Ms V);  S = s0 Ms  (device arithmetic on N entries)
        torch.mul(self.Ms_node, self.ops.vol, out=self.cex)
        torch.reciprocal(self.cex, out=self.cex)
        self.cex.mul_(-2.0 * self.A_ex * 1e18 / MU0)
        torch.mul(self.Ms_node, self.s0, out=self.S)
        _capi.check(L.smx_row_scale(self.kex.vals.numel(), _capi.ptr(self.kex.vals), _capi.ptr(self.slot_row),
                                    _capi.ptr(self.cex), _capi.ptr(self.ex_vals), st), "smx_row_scale")
        self._Ms_key = (key, Ms_cell)

    def _field(self, m, Ms_cell, ex_vals, out_h, out_hd=None, demag=True):
        """out_h = ex_vals-exchange(m) [+ Hd + H0]: the fused field kernel."""
        e, d = self.energy, self.energy.demag
        a = _capi.FieldArgs()
        a.n = self.N
        a.slice_ptr = _capi.ptr(self.kex.slice_ptr)
        a.cols = _capi.ptr(self.kex.cols)
        a.ex_vals = _capi.ptr(ex_vals)
        a.m = _capi.ptr(m)
        if demag and d is not None:
            a.grad_vals = _capi.ptr(d.G.vals)
            a.u1, a.u2 = _capi.ptr(d.u1), _capi.ptr(d.u2)
            a.u1_sum = d.u1_sum_ptr
            a.inv_vol = _capi.ptr(self.ops.inv_vol)
            a.h_demag = _capi.ptr(out_hd)
        if demag:
            a.h0_u[0], a.h0_u[1], a.h0_u[2] = self.h0[0], self.h0[1], self.h0[2]
        a.h_eff = _capi.ptr(out_h)
        _capi.check(self._L().smx_field_llg_rhs(C.byref(a), _capi.stream_ptr()), "smx_field_llg_rhs")

    def _potentials(self, m, Ms_cell):
        """b = Div_Ms m, u1, u2 (the demag potentials at m)."""
        e = self.energy
        if e.demag is None:
            return
        L, st = self._L(), _capi.stream_ptr()
        d = e.demag
        _capi.check(L.smx_energy_cell_forward(self.T, _capi.ptr(e.cells), _capi.ptr(e.ck), _capi.ptr(Ms_cell),
                                              _capi.ptr(self.zero_A), _capi.ptr(m), _capi.ptr(e.s), _capi.ptr(e.f),
                                              _capi.ptr(e.e), _capi.ptr(e.pT), st), "smx_energy_cell_forward")
        _capi.check(L.smx_energy_node_div(self.N, _capi.ptr(e.inc_ptr), _capi.ptr(e.inc), _capi.ptr(e.cgv),
                                          _capi.ptr(e.s), _capi.ptr(e.b), _capi.ptr(e.pN), st), "smx_energy_node_div")
        e._solve("u1", 0, d.u1, d.ws1, b=e.b, b_part=e.pN)
        d.boundary_matvec(d.u1, d.y_rows, gather=d.bverts_gpu, shift_ptr=d.u1_sum_ptr, shift_scale=1.0 / self.N)
        e._solve("u2", 1, d.u2, d.ws2, gb=d.y_rows, project=0)

    # ------------------------------------------------------------------
    def rhs(self, m: torch.Tensor, Ms_cell: torch.Tensor, out: torch.Tensor) -> torch.Tensor:
        """out = r(m; Ms_cell); leaves H in self.h and Hd in self.hd."""
        _capi.require_device(m, "m", numel=3 * self.N)
        self.set_Ms(Ms_cell)
        self._potentials(m, Ms_cell)
        self._field(m, Ms_cell, self.ex_vals, self.h, self.hd)
        _capi.check(self._L().smx_node_rhs(self.N, _capi.ptr(m), _capi.ptr(self.h), _capi.ptr(self.S), self.P,
                                           self.alpha, self.do_precess, _capi.ptr(out), _capi.stream_ptr()),
                    "smx_node_rhs")
        self.rhs_evals += 1
        return out

    def vjp(self, m: torch.Tensor, Ms_cell: torch.Tensor, lam: torch.Tensor):
        """(dr/dm^T lam, dr/dMs_cell^T lam) at (m, Ms_cell); recomputes the potentials."""
        e = self.energy
        L, st = self._L(), _capi.stream_ptr()
        r = torch.empty_like(m)
        self.rhs(m, Ms_cell, r)
        lam = lam.contiguous()
        _capi.check(L.smx_node_rhs_vjp(self.N, _capi.ptr(m), _capi.ptr(self.h), _capi.ptr(self.hd),
                                       C.cast(self.h0, C.c_void_p), _capi.ptr(self.Ms_node), _capi.ptr(self.S),
                                       self.s0, _capi.ptr(self.cex), _capi.ptr(self.ops.inv_vol), self.P, self.alpha,
                                       self.do_precess, _capi.ptr(lam), _capi.ptr(self.gm), _capi.ptr(self.gk),
                                       _capi.ptr(self.gd), _capi.ptr(self.gnode), st), "smx_node_rhs_vjp")
        # exchange transpose: K (cex gH) with the unscaled stiffness
        self._field(self.gk, Ms_cell, self.kex.vals, self.kg, demag=False)
        gm = self.gm + self.kg
        # nodal Ms terms -> cells
        _capi.check(L.smx_node_to_cell(self.T, _capi.ptr(e.cells), _capi.ptr(self.qvol), _capi.ptr(self.ops.inv_vol),
                                       _capi.ptr(self.gnode), 0, _capi.ptr(self.gMs), st), "smx_node_to_cell")
        gMs = self.gMs.clone()
        if e.demag is not None:
            d = e.demag
            # y = G^T (gH / V) = -Div_1 gd;  lt = L^T y
            d.Div.spmv(self.gd, self.yv, kind=1)
            self.yv.neg_()
            e.lt_apply(self.yv, self.lt)
            # Div_Ms^T lt (m part) and d(lt^T Div_Ms m)/dMs (cell part)
            _capi.check(L.smx_energy_cell_backward(self.T, _capi.ptr(e.cells), _capi.ptr(e.cgv), _capi.ptr(Ms_cell),
                                                   _capi.ptr(self.lt), _capi.ptr(m), _capi.ptr(e.tc),
                                                   _capi.ptr(e.grad_Ms), st), "smx_energy_cell_backward")
            _capi.check(L.smx_energy_node_grad(self.N, _capi.ptr(e.inc_ptr), _capi.ptr(e.inc), None, _capi.ptr(e.tc),
                                               _capi.ptr(m), _capi.ptr(self.ops.vol), None, e.axis_u, 0.0,
                                               _capi.ptr(e.grad_m), None, st), "smx_energy_node_grad")
            gm += e.grad_m
            gMs += e.grad_Ms
        self.vjp_evals += 1
        return gm, gMs


class LLGRHSFunction(torch.autograd.Function):
    """r = LLGRHSFunction.apply(m, Ms_cell, model)."""

    @staticmethod
    def forward(ctx, m, Ms_cell, model: LLGAdjointModel):
        m = m.detach().contiguous()
        Ms_cell = Ms_cell.detach().contiguous()
        out = torch.empty_like(m)
        model.rhs(m, Ms_cell, out)
        ctx.model = model
        ctx.save_for_backward(m, Ms_cell)
        return out

    @staticmethod
    def backward(ctx, lam):
        m, Ms_cell = ctx.saved_tensors
        gm, gMs = ctx.model.vjp(m, Ms_cell, lam)
        return gm, gMs, None


def dp5_step(f, m, k1, dt):
    """One fixed-step Dormand-Prince 5(4) step with FSAL: returns (m_new, k7)."""
    ks = [k1]
    for s in range(1, 6):
        y = m
        for j, a in enumerate(_DP_A[s]):
            if a != 0.0:
                y = y + (dt * a) * ks[j]
        ks.append(f(y))
    m_new = m
    for j, b in enumerate(_DP_B):
        if b != 0.0:
            m_new = m_new + (dt * b) * ks[j]
    return m_new, f(m_new)


def integrate(model: LLGAdjointModel, m0: torch.Tensor, Ms_cell: torch.Tensor, dt: float, steps: int,
              checkpoint_every: int = 50):
    """m after `steps` fixed DP5 steps; differentiable w.r.t. m0 and Ms_cell.
    The state (m, k1) is checkpointed every `checkpoint_every` steps and the
    stages in between are recomputed in the backward pass."""
    f = lambda y: LLGRHSFunction.apply(y, Ms_cell, model)

    def segment(m, k1, n):
        for _ in range(n):
            m, k1 = dp5_step(f, m, k1, dt)
        return m, k1

    m = m0
    k1 = f(m)
    done = 0
    while done < steps:
        n = min(checkpoint_every, steps - done)
        if torch.is_grad_enabled() and (m.requires_grad or Ms_cell.requires_grad):
            m, k1 = checkpoint(segment, m, k1, n, use_reentrant=False)
        else:
            m, k1 = segment(m, k1, n)
        done += n
    return m


def mean_mz(model: LLGAdjointModel, m: torch.Tensor) -> torch.Tensor:
    """J = sum_i V_i m_z,i / sum_i V_i."""
    vol = model.ops.vol
    return (vol * m.view(-1, 3)[:, 2]).sum() / vol.sum()
