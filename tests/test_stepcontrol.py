"""Step-size decisions of the integrators (host logic, CPU).

The expected values are worked out by hand from PETSc's TSAdaptChoose_Basic
with the reference's options (safety 0.9, reject safety 0.1, clip [0.1, 3],
dt_min 1e-18, dt_max 1e-10; llg_module_GPU.py:1764-1769)."""
import math

import pytest

from smx.stepcontrol import BasicAdaptor, check_final_time_mode


def test_accept_and_growth():
    a = BasicAdaptor()
    acc, dt = a.choose(0.5 ** 5, 1e-13, 5, False)          # err^(-1/5) = 2 -> factor 1.8
    assert acc and dt == pytest.approx(1.8e-13, rel=1e-14)
    acc, dt = a.choose(0.0, 1e-13, 5, False)               # no error: the largest growth
    assert acc and dt == pytest.approx(3e-13, rel=1e-15)
    acc, dt = a.choose(1e-30, 5e-11, 5, False)             # growth capped by dt_max
    assert acc and dt == 1e-10
    acc, dt = a.choose(1.0, 1e-13, 5, False)               # err == 1 is accepted
    assert acc and dt == pytest.approx(0.9e-13, rel=1e-14)


def test_rejection_and_reject_safety():
    a = BasicAdaptor()
    acc, dt = a.choose(32.0, 1e-13, 5, False)              # 0.9 * 32^(-1/5) = 0.45
    assert not acc and dt == pytest.approx(0.45e-13, rel=1e-14)
    # a second failed attempt of the same step: safety 0.9 * 0.1 -> factor 0.045, clipped to 0.1
    acc, dt = a.choose(32.0, 1e-13, 5, True)
    assert not acc and dt == pytest.approx(0.1e-13, rel=1e-14)
    acc, dt = a.choose(1.0 + 1e-9, 1e-13, 5, True)         # 0.09 * ~1 -> 0.1 after the clip
    assert not acc and dt == pytest.approx(0.1e-13, rel=1e-12)
    acc, dt = a.choose(1e6, 1e-13, 2, False)               # BDF1 order 2: 0.9e-3 -> clip 0.1
    assert not acc and dt == pytest.approx(1e-14, rel=1e-14)


def test_step_at_dt_min_is_accepted():
    a = BasicAdaptor(dt_min=1e-18)
    acc, dt = a.choose(50.0, 1e-18, 5, True)
    assert acc and dt == 1e-18
    acc, dt = a.choose(50.0, 1e-18 * (1 + 1e-9), 5, False)
    assert acc
    acc, _ = a.choose(50.0, 1.001e-18, 5, False)           # above dt_min by more than sqrt(eps): rejected
    assert not acc


@pytest.mark.parametrize("bad", [math.nan, math.inf, -math.inf])
def test_non_finite_error_raises(bad):
    with pytest.raises(FloatingPointError):
        BasicAdaptor().choose(bad, 1e-13, 5, False)


def test_final_time_modes():
    assert check_final_time_mode("StepOver") == "stepover"
    assert check_final_time_mode("matchstep") == "matchstep"
    with pytest.raises(ValueError):
        check_final_time_mode("interpolate")


def test_oracle_bdf_stepover_and_dt_min():
    """The oracle's BDF: "stepover" passes t1 with the last step, "matchstep"
    lands on it; a problem whose error exceeds 1 at every dt is accepted at
    dt_min instead of rejecting forever."""
    import numpy as np
    import smx_oracle as O
    A = np.array([[-1.0, 0.0], [0.0, -2.0]])
    f = lambda t, y: A @ y
    ops = lambda y, shift: ((lambda v: shift * v - A @ v), (lambda v: v / (shift + 1.5)))
    y0 = np.array([1.0, 1.0])
    kw = dict(rtol=1e-6, atol=1e-6, snes_rtol=1e-12, snes_atol=1e-14, ksp_rtol=1e-14, dt_max=0.7, dt_min=1e-6)
    _, t_m, _ = O.bdf_integrate(f, ops, y0, 0.0, 1.0, 0.3, exact_final_time="matchstep", **kw)
    _, t_s, _ = O.bdf_integrate(f, ops, y0, 0.0, 1.0, 0.3, exact_final_time="stepover", **kw)
    assert t_m == pytest.approx(1.0, abs=1e-15) and t_s >= 1.0
    # tolerances so tight that every step fails: the steps at dt_min are accepted
    tight = dict(kw, rtol=1e-30, atol=1e-30, dt_min=1e-3)
    _, t, st = O.bdf_integrate(f, ops, y0, 0.0, 0.01, 1e-3, **tight)
    assert t == pytest.approx(0.01, rel=1e-12) and st["rejected"] >= 1 and st["steps"] >= 4
