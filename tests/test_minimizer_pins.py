"""Pins of the oracle's LaBonte / Barzilai-Borwein minimizer.

* bb_step_rule on hand-computed s, y: BB1 = s.s/s.y, BB2 = s.y/y.y, the
  alternating rule by iteration parity, the adaptive rule's choice and its
  tau update, the degenerate and non-positive curvature branches;
* a macrospin in a field along its easy axis relaxes to m = n with the
  closed-form energy -mu0 Ms H - Ku (unit volume);
* an exchange + uniaxial box relaxes to the uniform state, E = -Ku V.
"""
import math

import numpy as np
import pytest

import smx_oracle as O

MU0 = 4e-7 * math.pi


def _state(tau=0.5, malpha=2):
    return dict(tau=tau, Malpha=malpha, tau_min=1e-12, tau_max=1e12, hist=[])


def test_bb_rules_by_hand():
    amin, amax = 1e-3, 1e3
    # s = (1, 1), y = (1, 0): s.s = 2, s.y = 1, y.y = 1 -> BB1 = 2, BB2 = 1
    st = _state()
    a, rule = O.bb_step_rule(2.0, 1.0, 1.0, 0.7, amin, amax, "adaptive", 0, st)
    assert (a, rule) == (1.0, "adaptive_recent_min_bb2") and st["tau"] == pytest.approx(0.45)
    # s = (1, 0), y = (2, 0): BB1 = BB2 = 0.5, ratio 1 > 0.45 -> BB1, tau * 1.1
    a, rule = O.bb_step_rule(1.0, 2.0, 4.0, 0.7, amin, amax, "adaptive", 1, st)
    assert (a, rule) == (0.5, "adaptive_bb1") and st["tau"] == pytest.approx(0.495)
    assert st["hist"] == [1.0, 0.5]
    assert O.bb_step_rule(2.0, 1.0, 1.0, 0.7, amin, amax, "alternate", 4, _state()) == (2.0, "bb1")
    assert O.bb_step_rule(2.0, 1.0, 1.0, 0.7, amin, amax, "alternate", 5, _state()) == (1.0, "bb2")
    assert O.bb_step_rule(2.0, 1.0, 1.0, 0.7, amin, amax, "bb2", 4, _state()) == (1.0, "bb2")
    # clamping, degenerate and non-positive curvature
    assert O.bb_step_rule(2.0, 1e-6, 1.0, 0.7, amin, amax, "bb1", 0, _state()) == (amax, "bb1")
    assert O.bb_step_rule(0.0, 0.0, 1.0, 0.7, amin, amax, "bb1", 0, _state()) == (0.7, "reuse_old_degenerate")
    assert O.bb_step_rule(1.0, -2.0, 4.0, 0.7, amin, amax, "adaptive", 0, _state()) == \
        (amax, "adaptive_alpha_max_nonpositive_sty")
    assert O.bb_step_rule(1.0, -2.0, 4.0, 0.7, amin, amax, "bb1", 0, _state()) == (0.7, "nonpositive_sty")


@pytest.mark.parametrize("variant", ["adaptive", "alternate", "bb1"])
def test_macrospin_relaxes_to_easy_axis(variant):
    Ms, Ku, H = 8e5, 5e5, 3e4
    n = np.array([0.0, 0.6, 0.8])
    pref = 2 * Ku / (MU0 * Ms)
    energy = lambda m: float(-MU0 * Ms * H * (m @ n) - Ku * (m @ n) ** 2)
    heff = lambda m: H * n + pref * (m @ n) * n
    # start inside the convex basin (s.y > 0 along the path: the BB rules apply);
    # the relative energy tests are off, so the run must reach the field tolerance
    m0 = n + np.array([0.3, 0.0, 0.0])
    # alpha_max above 1/|H_eff| ~ 1e-6 lets the BB step reach the macrospin's natural step
    m, info = O.labonte_bb_minimize(energy, heff, m0, max_iter=2000, tol=1e-3, bb_variant=variant, alpha_max=1e-5,
                                    energy_stagnation_rtol=0.0, energy_stagnation_atol=0.0, energy_accept_rtol=0.0)
    assert info["converged"] and info["iterations"] < 200, info
    assert np.abs(m - n).max() < 1e-8
    assert info["energy"] == pytest.approx(-MU0 * Ms * H - Ku, rel=1e-12)


def test_box_relaxes_to_uniform_state():
    from smx import mesh as M
    m = M.kuhn_box(6, 5, 4, 18.0, 15.0, 12.0, name="b")
    N = m.num_nodes
    Ms, A, Ku = 8e5, 1.3e-11, 5e5
    fem = O.fem_operators(m.X, m.tets, Ms)
    K = O.exchange_matrix(m.X, m.tets, fem.V, A, Ms)
    Vt = np.repeat(fem.V, 3) * 1e-27
    n = (0.0, 0.0, 1.0)
    hani = lambda mm: O.anisotropy_field(mm, n, Ku, Ms)
    energy = lambda mm: float(-0.5 * MU0 * Ms * np.sum(Vt * mm * (K @ mm + hani(mm))))
    heff = lambda mm: K @ mm + hani(mm)
    g = np.random.default_rng(11).standard_normal((N, 3)) * 0.3 + np.array([0.0, 0.0, 1.0])
    mm, info = O.labonte_bb_minimize(energy, heff, (g / np.linalg.norm(g, axis=1)[:, None]).ravel(),
                                     max_iter=3000, tol=1.0)
    assert info["converged"]
    assert mm.reshape(-1, 3)[:, 2].min() > 1 - 1e-8
    assert info["energy"] == pytest.approx(-Ku * fem.V.sum() * 1e-27, rel=1e-8)
