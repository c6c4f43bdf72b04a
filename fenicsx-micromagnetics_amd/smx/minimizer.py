"""Energy minimization by LaBonte-type steepest descent on the unit sphere.

Counterpart of LaBonteBBMinimizerGPU / EnergyMinimizerGPU
(gpu/solvers/Minimizer_GPU.py:171-846, 860-1225): the descent direction is
the projected effective field d = H - (m.H) m, the update is the
norm-preserving Cayley step m <- (I - C)^-1 (I + C) m with C v = c x v and
c = (step/2) m x d, the step length follows Barzilai-Borwein rules (BB1, BB2,
alternating, or the adaptive rule of Loris, Bertero, De Mol, Zanella and
Zanni), and a trial is accepted against the largest energy of a sliding
window (nonmonotone acceptance) with backtracking.  The per-node work runs
in HIP kernels (smx_projected_direction, smx_cayley_step, smx_max_norm3,
smx_dot3_weighted); the energies come from EffectiveField.compute_Energy_terms.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _capi
from .device import reduce_workspace
from .llg import EffectiveField


class LaBonteBBMinimizerGPU:
    def __init__(self, effective_field: EffectiveField):
        self.hef = effective_field
        n = effective_field.local_size
        dev = effective_field.device
        z = lambda: torch.zeros(n, dtype=torch.float64, device=dev)
        self.m_old, self.m_trial, self.H, self.d, self.d_old = z(), z(), z(), z(), z()
        self.s_vec, self.y_vec = z(), z()
        self.red = reduce_workspace(n + 64, dev)
        self._out = torch.zeros(1, dtype=torch.float64, device=dev)
        self.N = effective_field.N
        self.last_energy = None
        self.last_max_grad = None
        self.last_iterations = 0
        self._bb_reset(0.5, 2)

    # ---------------- device helpers ----------------
    def _L(self):
        return _capi.lib()

    def _max_norm3(self, v) -> float:
        _capi.check(self._L().smx_max_norm3(self.N, _capi.ptr(v), _capi.ptr(self._out), _capi.ptr(self.red),
                                            _capi.stream_ptr()), "smx_max_norm3")
        return float(self._out.item())

    def _dot(self, a, b) -> float:
        _capi.check(self._L().smx_dot3_weighted(self.N, _capi.ptr(a), _capi.ptr(b), None, _capi.ptr(self._out),
                                                _capi.ptr(self.red), _capi.stream_ptr()), "smx_dot3_weighted")
        return float(self._out.item())

    def _direction(self, m):
        """d = H_eff(m) - (m.H_eff) m (Minimizer_GPU.py:306-383)."""
        self.hef.compute_H_eff_vec(m, self.H)
        _capi.check(self._L().smx_projected_direction(self.N, _capi.ptr(m), _capi.ptr(self.H), _capi.ptr(self.d),
                                                      _capi.stream_ptr()), "smx_projected_direction")

    def _cayley(self, m, d, step, out):
        _capi.check(self._L().smx_cayley_step(self.N, _capi.ptr(m), _capi.ptr(d), float(step), _capi.ptr(out),
                                              _capi.stream_ptr()), "smx_cayley_step")

    def _energy(self, m) -> float:
        return float(self.hef.compute_Energy_terms(m)["E_total"])

    def _normalize(self, m):
        _capi.check(self._L().smx_normalize3(self.N, _capi.ptr(m), _capi.stream_ptr()), "smx_normalize3")

    def _sub(self, x, y, out):
        """out = x - y (smx_vec_glue op 3 with a = 1, b = -1)."""
        _capi.check(self._L().smx_vec_glue(3, 3 * self.N, _capi.ptr(x), _capi.ptr(y), None, 1.0, -1.0, None, None,
                                           1.0, _capi.ptr(out), _capi.stream_ptr()), "smx_vec_glue")

    # ---------------- Barzilai-Borwein step ----------------
    def _bb_reset(self, tau0, malpha, tau_min=1e-12, tau_max=1e12):
        self.bb_tau = min(max(float(tau0), tau_min), tau_max)
        self.bb_Malpha = max(int(malpha), 0)
        self.bb_tau_min, self.bb_tau_max = float(tau_min), float(tau_max)
        self.bb_alpha2_history = []
        self.bb_last_rule = "init"

    def _bb_step(self, alpha_old, alpha_min, alpha_max, variant, iteration):
        """BB step from s = m_n - m_{n-1} and y = g_n - g_{n-1}, g = -d
        (Minimizer_GPU.py:418-535)."""
        sty = self._dot(self.s_vec, self.y_vec)
        sts = self._dot(self.s_vec, self.s_vec)
        yty = self._dot(self.y_vec, self.y_vec)
        clamp = lambda a: min(max(float(a), alpha_min), alpha_max)
        if sts < 1e-300 or yty < 1e-300:
            self.bb_last_rule = "reuse_old_degenerate"
            return clamp(alpha_old)
        key = str(variant).lower().replace("-", "_")
        if key in ("adaptive", "gpss", "loris", "exl", "loris_zanni"):
            if sty <= 1e-300:
                self.bb_last_rule = "adaptive_alpha_max_nonpositive_sty"
                return float(alpha_max)
            a1, a2 = clamp(sts / sty), clamp(sty / yty)
            self.bb_alpha2_history.append(a2)
            self.bb_alpha2_history = self.bb_alpha2_history[-(self.bb_Malpha + 1):]
            if a2 / max(a1, 1e-300) <= self.bb_tau:
                self.bb_tau = max(self.bb_tau * 0.9, self.bb_tau_min)
                self.bb_last_rule = "adaptive_recent_min_bb2"
                return min(self.bb_alpha2_history)
            self.bb_tau = min(self.bb_tau * 1.1, self.bb_tau_max)
            self.bb_last_rule = "adaptive_bb1"
            return a1
        if sty <= 1e-300:
            self.bb_last_rule = "nonpositive_sty"
            return clamp(alpha_old)
        if key == "bb1" or (key == "alternate" and iteration % 2 == 0):
            self.bb_last_rule = "bb1"
            return clamp(sts / sty)
        if key in ("bb2", "alternate"):
            self.bb_last_rule = "bb2"
            return clamp(sty / yty)
        raise ValueError(f"unknown bb_variant {variant!r}")

    # ---------------- driver ----------------
    def minimize(self, max_iter=5000, tol=1.0, alpha0=5e-11, alpha_min=1e-13, alpha_max=1e-8,
                 bb_variant="adaptive", bb_adaptive_tau0=0.5, bb_adaptive_Malpha=2,
                 nonmonotone_window=10, max_backtracking=10, shrink=0.5, energy_accept_rtol=1e-10,
                 energy_accept_atol=1e-24, energy_stagnation_rtol=1e-10, energy_stagnation_atol=1e-24,
                 stagnation_window=30, max_rejected_at_alpha_min=30, alpha_restart=5e-11,
                 max_alpha_restarts=3, restart_residual_factor=50.0, stagnation_residual_stop=50.0,
                 print_every=0):
        """Minimize from hef.m_gpu in place; stops when max |d| < tol (A/m),
        on energy stagnation at a low residual, or on repeated rejection at
        alpha_min (Minimizer_GPU.py:537-846).  Returns a summary dict."""
        self._bb_reset(bb_adaptive_tau0, bb_adaptive_Malpha)
        m = self.hef.m_gpu
        self._normalize(m)
        E = self._energy(m)
        hist = [E]
        alpha = min(max(float(alpha0), alpha_min), alpha_max)
        rejected = restarts = 0
        stop = "max_iter"
        self._direction(m)
        it = 0
        for it in range(max_iter):
            self.last_iterations = it
            max_g = self._max_norm3(self.d)
            self.last_max_grad = max_g
            if print_every and it % print_every == 0:
                print(f"[LaBonte-BB] it={it:06d} E={E:.12e} max|projH|={max_g:.6e} alpha={alpha:.6e}", flush=True)
            if max_g < tol:
                stop = "projected_field_tol"
                break
            if len(hist) > stagnation_window:
                E_old, E_new = hist[-stagnation_window], hist[-1]
                if abs(E_new - E_old) <= energy_stagnation_atol + energy_stagnation_rtol * max(abs(E_old), abs(E_new), 1e-300):
                    if max_g <= restart_residual_factor * tol:
                        stop = "energy_stagnation_low_residual"
                        break
                    if restarts < max_alpha_restarts:
                        alpha = min(max(alpha_restart, alpha_min), alpha_max)
                        restarts += 1
                        rejected = 0
                        hist = [E]
                        self._direction(m)
                        continue
                    if max_g < stagnation_residual_stop:
                        stop = "energy_stagnation_high_residual"
                        break
            self.m_old.copy_(m)
            self.d_old.copy_(self.d)
            E_ref = max(hist[-nonmonotone_window:])
            accepted = False
            a_try = alpha
            for _ in range(max_backtracking + 1):
                self._cayley(self.m_old, self.d, a_try, self.m_trial)
                E_trial = self._energy(self.m_trial)
                if E_trial <= E_ref + energy_accept_atol + energy_accept_rtol * max(abs(E_ref), abs(E_trial), 1e-300):
                    m.copy_(self.m_trial)
                    E = E_trial
                    hist.append(E)
                    accepted = True
                    rejected = 0
                    break
                a_try = max(a_try * shrink, alpha_min)
            if not accepted:
                m.copy_(self.m_old)
                rejected += 1
                alpha = alpha_min
                if rejected >= max_rejected_at_alpha_min:
                    stop = "repeated_rejection_at_alpha_min"
                    break
                self._direction(m)
                continue
            self._direction(m)
            # s = m - m_old; y = g - g_old = -(d - d_old)
            self._sub(m, self.m_old, self.s_vec)
            self._sub(self.d_old, self.d, self.y_vec)
            alpha = min(max(self._bb_step(a_try, alpha_min, alpha_max, bb_variant, it), alpha_min), alpha_max)
        self._normalize(m)
        self._direction(m)
        self.last_energy = self._energy(m)
        self.last_max_grad = self._max_norm3(self.d)
        return {"iterations": int(self.last_iterations), "energy": float(self.last_energy),
                "max_projected_field": float(self.last_max_grad), "alpha": float(alpha),
                "bb_variant": str(bb_variant), "bb_last_rule": self.bb_last_rule, "alpha_restarts": int(restarts),
                "rejected_at_alpha_min": int(rejected), "stop_reason": stop,
                "converged": stop in ("projected_field_tol", "energy_stagnation_low_residual")}


class EnergyMinimizerGPU:
    """Term builder + minimizer (Minimizer_GPU.py:860-1225)."""

    def __init__(self, mesh, Ms, gamma=2.211e5, alpha=0.5, device=None):
        self.mesh = mesh
        self._kw = dict(Ms=Ms, Aex=0.0, gamma=gamma, alpha=alpha)
        self._demag = None
        self.device = device
        self.hef = None

    def add_exchange(self, Aex):
        self._kw["Aex"] = float(Aex)

    def add_demag(self, method="lindholm", **kwargs):
        self._demag = (method, kwargs)

    def add_anisotropy(self, Ku, n_vec):
        self._kw.update(Ku=Ku, n_ani_vec=n_vec)

    add_uniaxial_anisotropy = add_anisotropy

    def add_dmi_bulk(self, D_bulk):
        self._kw["D_bulk"] = D_bulk

    def add_dmi_interfacial(self, D_int, n0_vec):
        self._kw.update(D_int=D_int, n0_int_vec=n0_vec)

    def add_external_field(self, H0_vec=None, H_time_func=None):
        self._kw.update(H0_static=H0_vec, H_time_func=H_time_func)

    def set_uniform_field(self, Hx, Hy, Hz):
        self._kw["H0_static"] = (Hx, Hy, Hz)
        if self.hef is not None:
            self.hef.set_uniform_field(Hx, Hy, Hz)

    def add_cubic_anisotropy(self, Kc1, u1_vec, u2_vec):
        self._kw.update(Kc1=Kc1, u1_cub=u1_vec, u2_cub=u2_vec)

    def _build_effective_field(self):
        kw = dict(self._kw)
        if self._demag is not None:
            kw.update(use_demag=True, demag_method=self._demag[0], demag_kwargs=self._demag[1])
        self.hef = EffectiveField(self.mesh, device=self.device, **kw)
        return self.hef

    def minimize(self, m0_array=None, **kwargs):
        if self.hef is None:
            self._build_effective_field()
        if m0_array is not None:
            self.hef.set_m_from_cpu(np.asarray(m0_array, dtype=np.float64))
        self.minimizer = LaBonteBBMinimizerGPU(self.hef)
        return self.minimizer.minimize(**kwargs)
