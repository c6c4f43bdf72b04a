"""Adaptive step-size decision shared by the explicit and implicit integrators.

The reference drives both integrators through PETSc TS with the "basic"
adaptor and these options (llg_module_GPU.py:1764-1769, 2078-2084): clip
[0.1, 3], safety 0.9, reject safety 0.1, dt in [dt_min, dt_max]; PETSc's own
defaults add at most 10 rejections of one step (``ts_max_reject``).  The rule
(PETSc TSAdaptChoose_Basic, v3.25):

- the step is accepted when the weighted error norm is <= 1, and also when it
  is not but the step is already at dt_min (within sqrt(machine eps));
- the next step is dt * clip(s * err^(-1/order), clip) with s = safety, or
  safety * reject_safety when this attempt failed right after another failed
  attempt of the same step; then dt is kept within [dt_min, dt_max];
- err = 0 asks for the largest growth (clip[1]).

A non-finite error norm (a diverged stage, e.g. a Poisson solve that did not
converge) raises instead of shrinking dt forever.  Both integrators land the
last step on t_final only with ``exact_final_time="matchstep"``; the default
``"stepover"`` lets the last step pass t_final, as the reference's
ExactFinalTime.STEPOVER does (llg_module_GPU.py:1752, 2106).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

_SQRT_EPS = math.sqrt(2.220446049250313e-16)

FINAL_TIME_MODES = ("stepover", "matchstep")


class StepRejectedError(RuntimeError):
    """More consecutive rejections of one step than max_reject allows."""


@dataclass
class BasicAdaptor:
    safety: float = 0.9
    reject_safety: float = 0.1
    clip: tuple = (0.1, 3.0)
    dt_min: float = 1e-18
    dt_max: float = 1e-10
    max_reject: int = 10

    def choose(self, err: float, dt: float, order: float, prev_attempt_failed: bool):
        """(accept, dt_next) for an attempt of size dt with error norm err."""
        err = float(err)
        if not math.isfinite(err):
            raise FloatingPointError(f"step error norm is {err} at dt = {dt:.3e}: a stage did not produce "
                                     "finite values")
        s = self.safety
        if err > 1.0:
            if prev_attempt_failed:
                s *= self.reject_safety
            accept = dt < (1.0 + _SQRT_EPS) * self.dt_min
        else:
            accept = True
        fac = s * err ** (-1.0 / float(order)) if err > 0.0 else math.inf
        fac = min(max(fac, self.clip[0]), self.clip[1])
        return accept, min(max(dt * fac, self.dt_min), self.dt_max)


def check_final_time_mode(mode: str) -> str:
    mode = str(mode).strip().lower()
    if mode not in FINAL_TIME_MODES:
        raise ValueError(f"exact_final_time must be one of {FINAL_TIME_MODES}, got {mode!r}")
    return mode
