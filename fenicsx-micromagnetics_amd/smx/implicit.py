"""Implicit LLG time stepping: variable-step BDF with Newton-Krylov solves.

Counterpart of the reference's implicit relax (gpu/solvers/llg_module_GPU.py:
2052-2381: PETSc TS ``bdf`` with the ``basic`` step adaptor, SNES
``newtonls`` with a backtracking line search, KSP ``gmres`` with the
matrix-free Jacobian and the 3x3 block preconditioner of ``JvContext``,
:1412-1544).  Everything per-node runs in HIP kernels: the RHS (fused field
kernel + demag), the Jacobian-vector product (smx_llg_jvp), the block
preconditioner (smx_llg_pc_build / smx_block3_apply), the Arnoldi work of the
GMRES (smx_multi_dot, smx_lincomb, smx_vec_glue, smx_dot3_weighted) and the
error norm (smx_rk_combine).  The host runs the Newton, GMRES (Givens
rotations on the Hessenberg matrix) and step-size logic, as PETSc does.

Time integration (BDF of order 1 on the first step, then 2, variable step):
  BDF2  dy/dt(t_{n+1}) ~ [a0 y_{n+1} + a1 y_n + a2 y_{n-1}] / h,
        w = h/h_prev, a0 = (1+2w)/(1+w), a1 = -(1+w), a2 = w^2/(1+w);
  Newton on F(y) = a0/h y + (a1 y_n + a2 y_{n-1})/h - f(y) = 0 with the
  Jacobian shift I - J (shift = a0/h, J excludes demag as the reference's);
  local error (Milne): C (y - y_pred), y_pred the polynomial extrapolation of
  the last three (BDF2; C = 2/11) or two (BDF1; C = 1/2) accepted states,
  weighted RMS norm with atol + rtol max(|y_n|, |y_{n+1}|), step factor
  0.9 err^(-1/(q+1)) clipped to [0.1, 3], x0.25 after a failed solve; a
  step that would leave less than 1e-3 of itself before t1 is stretched to
  land on t1 (no sliver step, whose huge shift the Newton tolerance cannot
  resolve).
"""
from __future__ import annotations

import ctypes as C
import math
import time

import numpy as np
import torch

from . import _capi
from .device import reduce_workspace
from .llg import EffectiveField
from .stepcontrol import BasicAdaptor, StepRejectedError, check_final_time_mode


class JvContext:
    """Matrix-free Newton operator y = shift x - J x and its 3x3 block
    preconditioner (llg_module_GPU.py:1412-1544)."""

    def __init__(self, hef: EffectiveField, eps_reg=1e-14, det_eps=1e-30):
        self.hef = hef
        self.shift = 0.0
        self.eps_reg = float(eps_reg)
        self.det_eps = float(det_eps)
        self.calls = 0
        self.callsPre = 0
        n = hef.local_size
        dev = hef.device
        self.inv = torch.zeros(3 * n, dtype=torch.float64, device=dev)    # 9 per node
        self.Jv = torch.zeros(n, dtype=torch.float64, device=dev)
        self._pc_ready = False

    def update_pc_full_fast_gpu(self, shift, include_stab=True, use_abs_kappa=True, exact=False):
        """Rebuild the block inverses at the Jacobian state (hef.M_state_gpu /
        hef.Hm_gpu, set by hef.update_jac_state).  exact=False: the
        reference's block; exact=True: the RHS's own local derivative."""
        self.shift = float(shift)
        h = self.hef
        a = h._args(h.M_state_gpu, with_demag=False, with_h0=False)
        _capi.check(_capi.lib().smx_llg_pc_build(C.byref(a), _capi.ptr(h.Hm_gpu), self.shift, self.eps_reg,
                                                 self.det_eps, int(bool(include_stab)), int(bool(use_abs_kappa)),
                                                 int(bool(exact)), _capi.ptr(self.inv), _capi.stream_ptr()),
                    "smx_llg_pc_build")
        self._pc_ready = True

    def mult(self, *args):
        """y = shift x - J x; called as mult(x, y) or, like the reference's
        matrix-free shell context, mult(A, x, y) (llg_module_GPU.py:1510-1517)."""
        x, y = args[-2], args[-1]
        self.calls += 1
        self.hef.jac_vec_times_vec(x, self.Jv)
        _glue_axpby(self.hef.local_size, self.shift, x, -1.0, self.Jv, y)

    def apply(self, *args):
        """y = P^-1 x with the 3x3 blocks; apply(x, y) or apply(pc, x, y)
        (llg_module_GPU.py:1519-1544)."""
        x, y = args[-2], args[-1]
        self.callsPre += 1
        if not self._pc_ready:
            y.copy_(x)
            return
        _capi.check(_capi.lib().smx_block3_apply(self.hef.N, _capi.ptr(self.inv), _capi.ptr(x), _capi.ptr(y),
                                                 _capi.stream_ptr()), "smx_block3_apply")


def _glue_axpby(n, a, x, b, y, out):
    """out = a x + b y (y may be None)."""
    _capi.check(_capi.lib().smx_vec_glue(3, int(n), _capi.ptr(x), _capi.ptr(y), None, float(a), float(b), None, None,
                                         1.0, _capi.ptr(out), _capi.stream_ptr()), "smx_vec_glue")


class GMRES:
    """Right-preconditioned restarted GMRES with classical Gram-Schmidt applied
    twice (one multi-dot and one linear-combination launch per pass), the
    Krylov vectors kept on the device (PETSc KSP gmres, restart 30)."""

    def __init__(self, n: int, device, restart: int = 30):
        self.n = int(n)
        self.restart = int(restart)
        z = lambda *s: torch.zeros(*s, dtype=torch.float64, device=device)
        self.V = z(self.restart + 1, self.n)
        self.Z = z(self.restart, self.n)
        self.r = z(self.n)
        self.t = z(self.n)
        self.h_dev = z(self.restart + 1)
        self.y_dev = z(self.restart)
        self.work = z(int(_capi.lib().smx_multi_dot_ws_size(self.n, self.restart + 1)))
        self.red = reduce_workspace(self.n + 64, device)
        self._s = z(1)
        self.iterations = 0

    def _dot(self, a, b) -> float:
        _capi.check(_capi.lib().smx_dot3_weighted(self.n // 3, _capi.ptr(a), _capi.ptr(b), None, _capi.ptr(self._s),
                                                  _capi.ptr(self.red), _capi.stream_ptr()), "smx_dot3_weighted")
        return float(self._s.item())

    def _orth(self, j: int, w) -> np.ndarray:
        """w <- w - V[:j+1]^T (V[:j+1] w), twice; returns the coefficients."""
        L = _capi.lib()
        st = _capi.stream_ptr()
        tot = np.zeros(j + 1)
        for _ in range(2):
            _capi.check(L.smx_multi_dot(self.n, j + 1, self.V.data_ptr(), self.n, _capi.ptr(w),
                                        self.h_dev.data_ptr(), self.work.data_ptr(), st), "smx_multi_dot")
            h = self.h_dev[: j + 1].cpu().numpy()
            tot += h
            self.y_dev[: j + 1].copy_(torch.from_numpy(-h))
            _capi.check(L.smx_lincomb(self.n, j + 1, self.V.data_ptr(), self.n, self.y_dev.data_ptr(), 1.0,
                                      _capi.ptr(w), st), "smx_lincomb")
        return tot

    def solve(self, mult, precond, b, x, rtol, atol=0.0, max_it=200) -> tuple[int, float]:
        """x <- approximate solution of A x = b, ||b - A x|| <= max(rtol ||b||, atol)."""
        L = _capi.lib()
        st = _capi.stream_ptr()
        bnorm = math.sqrt(max(self._dot(b, b), 0.0))
        tol = max(rtol * bnorm, atol)
        mult(x, self.t)
        _glue_axpby(self.n, 1.0, b, -1.0, self.t, self.r)
        beta = math.sqrt(max(self._dot(self.r, self.r), 0.0))
        total = 0
        m = self.restart
        while beta > tol and total < max_it:
            _glue_axpby(self.n, 1.0 / beta, self.r, 0.0, None, self.V[0])
            H = np.zeros((m + 1, m))
            cs, sn = np.zeros(m), np.zeros(m)
            g = np.zeros(m + 1)
            g[0] = beta
            k = 0
            for j in range(m):
                precond(self.V[j], self.Z[j])
                mult(self.Z[j], self.V[j + 1])
                H[: j + 1, j] = self._orth(j, self.V[j + 1])
                hn = math.sqrt(max(self._dot(self.V[j + 1], self.V[j + 1]), 0.0))
                H[j + 1, j] = hn
                if hn > 0.0:
                    _glue_axpby(self.n, 1.0 / hn, self.V[j + 1], 0.0, None, self.V[j + 1])
                for i in range(j):                       # previous rotations
                    a, c = H[i, j], H[i + 1, j]
                    H[i, j], H[i + 1, j] = cs[i] * a + sn[i] * c, -sn[i] * a + cs[i] * c
                den = math.hypot(H[j, j], H[j + 1, j])
                cs[j], sn[j] = (1.0, 0.0) if den == 0.0 else (H[j, j] / den, H[j + 1, j] / den)
                H[j, j] = cs[j] * H[j, j] + sn[j] * H[j + 1, j]
                H[j + 1, j] = 0.0
                g[j + 1] = -sn[j] * g[j]
                g[j] = cs[j] * g[j]
                k = j + 1
                total += 1
                if abs(g[j + 1]) <= tol or total >= max_it or hn == 0.0:
                    break
            y = np.zeros(k)
            for i in range(k - 1, -1, -1):
                y[i] = (g[i] - H[i, i + 1:k] @ y[i + 1:k]) / H[i, i]
            self.y_dev[:k].copy_(torch.from_numpy(y))
            _capi.check(L.smx_lincomb(self.n, k, self.Z.data_ptr(), self.n, self.y_dev.data_ptr(), 1.0, _capi.ptr(x),
                                      st), "smx_lincomb")
            mult(x, self.t)
            _glue_axpby(self.n, 1.0, b, -1.0, self.t, self.r)
            beta = math.sqrt(max(self._dot(self.r, self.r), 0.0))
        self.iterations = total
        return total, beta


class BDFIntegrator:
    """Variable-step BDF(1, 2) Newton-Krylov integrator (see the module notes)."""

    def __init__(self, hef: EffectiveField, rtol=1e-6, atol=1e-6, snes_rtol=1e-2, snes_atol=1e-4,
                 snes_max_it=8, ksp_rtol=1e-4, ksp_max_it=200, dt_min=1e-17, dt_max=1e-10, safety=0.9,
                 clip=(0.1, 3.0), scale_solve_failed=0.25, use_pc=True, pc_form="exact", reject_safety=0.1,
                 max_reject=10):
        self.hef = hef
        self.rtol, self.atol = float(rtol), float(atol)
        self.snes_rtol, self.snes_atol, self.snes_max_it = float(snes_rtol), float(snes_atol), int(snes_max_it)
        self.ksp_rtol, self.ksp_max_it = float(ksp_rtol), int(ksp_max_it)
        self.dt_min, self.dt_max = float(dt_min), float(dt_max)
        self.safety, self.clip, self.scale_failed = safety, clip, scale_solve_failed
        self.adaptor = BasicAdaptor(safety=float(safety), reject_safety=float(reject_safety), clip=tuple(clip),
                                    dt_min=self.dt_min, dt_max=self.dt_max, max_reject=int(max_reject))
        self.use_pc = bool(use_pc)
        if pc_form not in ("exact", "reference"):
            raise ValueError(f"pc_form must be 'exact' or 'reference', got {pc_form!r}")
        # "exact": the RHS's local derivative with the signed mean diagonal
        # (GMRES on the Newton operator of an 8x7x5 exchange + uniaxial box at
        # shift 2e12: 36 iterations against 86 without a preconditioner);
        # "reference": the reference's block (|kappa|, the precession and
        # damping parts negated), which stalls on the same system.
        self.pc_kw = dict(exact=True, use_abs_kappa=False) if pc_form == "exact" else \
            dict(exact=False, use_abs_kappa=True)
        n = hef.local_size
        dev = hef.device
        z = lambda: torch.zeros(n, dtype=torch.float64, device=dev)
        self.n = n
        self.hist = [z(), z(), z()]          # y_n, y_{n-1}, y_{n-2}
        self.times = [0.0, 0.0, 0.0]
        self.nhist = 0
        self.y, self.ypred, self.c_hist, self.F, self.f, self.delta, self.ytry = (z() for _ in range(7))
        self.ctx = JvContext(hef)
        self.gmres = GMRES(n, dev)
        self.red = reduce_workspace(n, dev)
        self.err = torch.zeros(1, dtype=torch.float64, device=dev)
        self._s = torch.zeros(1, dtype=torch.float64, device=dev)
        self.u_prev = z()
        self.last_max_dmdt_deg_ns = float("nan")
        self.stats = {"steps": 0, "rejected": 0, "failed_solves": 0, "newton": 0, "gmres": 0, "rhs_evals": 0}

    # ---------------- device helpers ----------------
    def _norm(self, v) -> float:
        _capi.check(_capi.lib().smx_dot3_weighted(self.n // 3, _capi.ptr(v), _capi.ptr(v), None, _capi.ptr(self._s),
                                                  _capi.ptr(self.red), _capi.stream_ptr()), "smx_dot3_weighted")
        return math.sqrt(max(float(self._s.item()), 0.0))

    def _comb(self, coeffs, vecs, out):
        """out = sum c_i v_i (two or three vectors)."""
        _glue_axpby(self.n, coeffs[0], vecs[0], coeffs[1], vecs[1], out)
        if len(vecs) > 2:
            _glue_axpby(self.n, 1.0, out, coeffs[2], vecs[2], out)

    def _rhs(self, t, y, out):
        self.hef.current_time = t
        self.hef.llg_rhs_vec(y, out)
        self.stats["rhs_evals"] += 1

    def _residual(self, t, y, shift, out) -> float:
        """out = shift y - c_hist - f(y); returns ||out||."""
        self._rhs(t, y, self.f)
        _glue_axpby(self.n, shift, y, -1.0, self.f, out)
        _glue_axpby(self.n, 1.0, out, -1.0, self.c_hist, out)
        return self._norm(out)

    def _error_norm(self, coef) -> float:
        """WRMS of coef (y - ypred) with atol + rtol max(|y_n|, |y|)."""
        L = _capi.lib()
        yn = self.hist[0]
        _glue_axpby(self.n, 1.0, self.y, -1.0, yn, self.delta)       # k0 = y - y_n
        _glue_axpby(self.n, 1.0, yn, -1.0, self.ypred, self.ytry)    # k1 = y_n - ypred
        ptrs = (C.c_void_p * 7)(self.delta.data_ptr(), self.ytry.data_ptr(), None, None, None, None, None)
        c = (C.c_double * 7)(1.0, 0.0, 0, 0, 0, 0, 0)
        e = (C.c_double * 7)(coef, coef, 0, 0, 0, 0, 0)
        _capi.check(L.smx_rk_combine(self.n, yn.data_ptr(), 1.0, 2, C.cast(ptrs, C.c_void_p), C.cast(c, C.c_void_p),
                                     C.cast(e, C.c_void_p), self.atol, self.rtol, self.F.data_ptr(),
                                     self.err.data_ptr(), self.red.data_ptr(), _capi.stream_ptr()), "smx_rk_combine")
        return float(self.err.item())

    # ---------------- one step ----------------
    def _newton(self, t1, shift) -> bool:
        F0 = self._residual(t1, self.y, shift, self.F)
        Fn = F0
        tol = max(self.snes_atol, self.snes_rtol * F0)
        for it in range(self.snes_max_it):
            if Fn <= tol and it > 0:
                return True
            self.stats["newton"] += 1
            self.hef.update_jac_state(self.y)
            self.ctx.shift = shift
            if self.use_pc:
                self.ctx.update_pc_full_fast_gpu(shift, include_stab=True, **self.pc_kw)
            # (shift I - J) delta = -F
            _glue_axpby(self.n, -1.0, self.F, 0.0, None, self.f)
            self.delta.zero_()
            its, _ = self.gmres.solve(self.ctx.mult, self.ctx.apply if self.use_pc else (lambda x, y: y.copy_(x)),
                                      self.f, self.delta, self.ksp_rtol, 0.0, self.ksp_max_it)
            self.stats["gmres"] += its
            # backtracking line search on ||F||
            lam = 1.0
            for _ in range(5):
                _glue_axpby(self.n, 1.0, self.y, lam, self.delta, self.ytry)
                Ft = self._residual(t1, self.ytry, shift, self.F)
                if Ft <= (1.0 - 1e-4 * lam) * Fn or lam < 0.1:
                    break
                lam *= 0.5
            self.y.copy_(self.ytry)
            Fn = Ft
            if Fn <= tol:
                return True
        return Fn <= tol

    def _max_dmdt_deg_ns(self, t, t_prev) -> float:
        """max_i |y_i - u_prev_i| / (t - t_prev) in degrees per ns
        (llg_module_GPU.py:330-378); u_prev <- y."""
        _glue_axpby(self.n, 1.0, self.hist[0], -1.0, self.u_prev, self.delta)
        _capi.check(_capi.lib().smx_max_norm3(self.n // 3, _capi.ptr(self.delta), _capi.ptr(self._s),
                                              _capi.ptr(self.red), _capi.stream_ptr()), "smx_max_norm3")
        self.u_prev.copy_(self.hist[0])
        return float(self._s.item()) / (t - t_prev) * (180.0 / math.pi) * 1e-9

    def integrate(self, y, t0, t1, dt, max_steps=10 ** 7, stopping_dmdt=0.0, check_every=5, monitor_fn=None,
                  exact_final_time="stepover"):
        """Advance y (device, in place) from t0 until t >= t1.  stopping_dmdt > 0:
        stop early once max |dm/dt| over the last check_every accepted steps
        falls below it (deg/ns).  monitor_fn(step, t, y) after each accepted
        step.  exact_final_time "stepover" (the reference's) lets the last
        step pass t1; "matchstep" lands on t1 (a step that would leave less
        than 1e-3 of itself before t1 is stretched to reach it)."""
        match = check_final_time_mode(exact_final_time) == "matchstep"
        failed = False
        reject_run = 0
        t = float(t0)
        self.hist[0].copy_(y)
        self.times[0] = t
        self.nhist = 1
        dt = min(max(float(dt), self.dt_min), self.dt_max)
        steps = 0
        self.u_prev.copy_(y)
        t_check = t
        self.stats["stopped_on_dmdt"] = False
        while t < t1 * (1 - 1e-15) and steps < max_steps:
            rem = t1 - t
            if match and (dt >= rem or rem - dt < 1e-3 * dt):   # land on t1: no sliver step before it
                dt = rem
            q = 2 if self.nhist >= 2 else 1
            yn = self.hist[0]
            if q == 2:
                w = dt / (self.times[0] - self.times[1])
                a0, a1, a2 = (1 + 2 * w) / (1 + w), -(1 + w), w * w / (1 + w)
                shift = a0 / dt
                self._comb((-a1 / dt, -a2 / dt), (yn, self.hist[1]), self.c_hist)
            else:
                shift = 1.0 / dt
                _glue_axpby(self.n, 1.0 / dt, yn, 0.0, None, self.c_hist)
            # predictor: polynomial extrapolation of the accepted states
            tn = [self.times[i] for i in range(min(self.nhist, 3))]
            tq = t + dt
            lag = [math.prod((tq - tn[j]) / (tn[i] - tn[j]) for j in range(len(tn)) if j != i) for i in range(len(tn))]
            if len(tn) == 1:
                self.ypred.copy_(yn)
            else:
                self._comb(tuple(lag), tuple(self.hist[: len(tn)]), self.ypred)
            self.y.copy_(self.ypred)
            ok = self._newton(t + dt, shift)
            if not ok:
                self.stats["failed_solves"] += 1
                dt = max(dt * self.scale_failed, self.dt_min)
                if dt <= self.dt_min * (1 + 1e-12) and self.stats["failed_solves"] > 50:
                    raise RuntimeError("BDF: nonlinear solves keep failing at dt_min")
                continue
            if len(tn) >= q + 1:
                err = self._error_norm(2.0 / 11.0 if q == 2 else 0.5)
            else:
                err = 0.0                                  # no estimate yet: accept
            accept, dt_next = self.adaptor.choose(err, dt, q + 1, failed)
            if accept:
                failed = False
                reject_run = 0
                # rotate history
                self.hist = [self.hist[2], self.hist[0], self.hist[1]]
                self.hist[0].copy_(self.y)
                self.times = [t + dt, self.times[0], self.times[1]]
                self.nhist = min(self.nhist + 1, 3)
                t += dt
                steps += 1
                self.stats["steps"] += 1
                dt = dt_next
                if monitor_fn is not None:
                    monitor_fn(steps, t, self.hist[0])
                if stopping_dmdt > 0.0 and steps % max(int(check_every), 1) == 0:
                    self.last_max_dmdt_deg_ns = self._max_dmdt_deg_ns(t, t_check)
                    t_check = t
                    if self.last_max_dmdt_deg_ns < stopping_dmdt:
                        self.stats["stopped_on_dmdt"] = True
                        break
            else:
                self.stats["rejected"] += 1
                reject_run = (reject_run + 1) if failed else 1
                if reject_run > self.adaptor.max_reject:
                    raise StepRejectedError(f"BDF: {reject_run} consecutive rejections at t = {t:.6e} "
                                            f"(dt = {dt:.3e}, error norm {err:.3e})")
                failed = True
                dt = dt_next
        y.copy_(self.hist[0])
        return {"t": t, "dt": dt, "maxdmdt_deg_ns": self.last_max_dmdt_deg_ns, **self.stats}


def relax_implicit(hef: EffectiveField, t_final, dt_init=1e-14, ts_rtol=1e-6, ts_atol=1e-6, snes_rtol=1e-2,
                   snes_atol=1e-4, ksp_rtol=1e-4, m0_array=None, use_python_pc=True, max_steps=5_000_000,
                   t0=0.0, stopping_dmdt=0.0, check_every_stop=5, monitor_fn=None, pc_form="exact",
                   exact_final_time="stepover"):
    """Implicit relax of hef.m_gpu from t0 to t_final (llg_module_GPU.py:
    2052-2381); the output-directory and snapshot machinery of the reference
    is left to monitor_fn(step, t, m)."""
    if m0_array is not None:
        hef.set_m_from_cpu(m0_array)
    integ = BDFIntegrator(hef, rtol=ts_rtol, atol=ts_atol, snes_rtol=snes_rtol, snes_atol=snes_atol,
                          ksp_rtol=ksp_rtol, use_pc=use_python_pc, pc_form=pc_form)
    t0w = time.perf_counter()
    out = integ.integrate(hef.m_gpu, float(t0), float(t_final), float(dt_init), max_steps=max_steps,
                          stopping_dmdt=stopping_dmdt, check_every=check_every_stop, monitor_fn=monitor_fn,
                          exact_final_time=exact_final_time)
    out["elapsed_s"] = time.perf_counter() - t0w
    return out
