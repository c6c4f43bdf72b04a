"""Host pins of the implicit-step pieces.

* the oracle's local 3x3 Jacobian of the LLG RHS equals the central
  finite-difference Jacobian of llg_rhs(m', H + kappa (m' - m)) at m' = m
  (the field's self-derivative taken as kappa I), with and without the
  norm stabilization;
* at kappa = 0 it is the exact derivative of the RHS for a fixed field.
"""
import numpy as np
import pytest

import smx_oracle as O


@pytest.mark.parametrize("stab", [True, False])
def test_local_jacobian_is_the_fd_jacobian(stab):
    rng = np.random.default_rng(0)
    gamma, alpha, Ms = 2.211e5, 0.3, 8e5
    for _ in range(5):
        m = rng.standard_normal(3)
        m /= np.linalg.norm(m) * 1.02       # off the unit sphere: the stabilization term is live
        h = rng.standard_normal(3) * 1e5
        kappa = rng.uniform(-3e5, 3e5)
        J = O.llg_local_jacobian(m, h, kappa, gamma, alpha, Ms, include_stab=stab)
        S = Ms * gamma / (1 + alpha ** 2) * 0.5

        def r(mp):
            out = O.llg_rhs(mp, h + kappa * (mp - m), gamma, alpha, Ms)
            if not stab:
                out = out - S * (1.0 - mp @ mp) * mp
            return out

        eps = 1e-6
        Jfd = np.column_stack([(r(m + eps * e) - r(m - eps * e)) / (2 * eps) for e in np.eye(3)])
        assert np.abs(J - Jfd).max() <= 1e-7 * np.abs(Jfd).max()
