"""Pins of the oracle's SVD compressors.

* truncated_svd on matrices of prescribed singular values: the rank is the
  smallest r whose tail energy sqrt(sum_{i>r} s_i^2) is within rtol of
  ||s||, computed here from the prescribed s (no SVD); the reported residual
  equals ||A - U V||_F / ||A||_F; max_rank caps the rank;
* randomized_svd: exact recovery of rank-r input (residual at rounding
  level); the reported residual equals ||A - U V||_F / ||A||_F (for an
  orthonormal Q, ||A - Q B_r||^2 = ||A||^2 - sum_{i<=r} s_i(B)^2; both sides
  squared agree to the rounding of ||A||^2); a too-small
  initial rank grows until the tolerance is met, and the rank found is at
  least the optimal (SVD) rank and at most the sketch width reached;
* the per-block seed mixing on hand-computed values.  (The product applies
  the rank rule inside its compressor kernels; the GPU tests compare its
  ranks with these.)
"""
import math

import numpy as np
import pytest

import smx_oracle as O
from smx.hmatrix.compressors import block_random_seed


def _prescribed(m, n, s, seed):
    rng = np.random.default_rng(seed)
    Q1, _ = np.linalg.qr(rng.standard_normal((m, m)))
    Q2, _ = np.linalg.qr(rng.standard_normal((n, n)))
    k = s.size
    return (Q1[:, :k] * s) @ Q2[:, :k].T


def _expected_rank(s, rtol, max_rank):
    tot = float(np.sum(s * s))
    for r in range(0, min(max_rank, s.size) + 1):
        if math.sqrt(max(float(np.sum(s[r:] ** 2)), 0.0)) <= rtol * math.sqrt(tot):
            return max(r, 1)
    return min(max_rank, s.size)


@pytest.mark.parametrize("decay,rtol", [(0.37, 1e-6), (0.61, 1e-4), (0.83, 1e-8)])
def test_truncated_svd_rank_and_residual(decay, rtol):
    m, n = 70, 50
    s = decay ** np.arange(n) * (1.0 + 0.013 * np.arange(n) % 0.1)   # no ties at the threshold
    s = np.sort(s)[::-1]
    A = _prescribed(m, n, s, 1)
    U, V, rel = O.truncated_svd(A, rtol)
    r = U.shape[1]
    assert r == _expected_rank(s, rtol, n)
    true = np.linalg.norm(A - U @ V) / np.linalg.norm(A)
    assert true <= rtol * (1 + 1e-6)
    assert abs(rel ** 2 - true ** 2) <= 64 * n * np.finfo(float).eps     # cancellation in ||A||^2 - captured
    # capped
    U2, V2, rel2 = O.truncated_svd(A, rtol, max_rank=max(r - 3, 1))
    assert U2.shape[1] == max(r - 3, 1) and rel2 > rtol


def test_randomized_svd_exact_low_rank():
    s = np.array([5.0, 3.0, 2.0, 1.0, 0.5, 0.25, 0.1])
    A = _prescribed(90, 60, s, 2)
    U, V, rel = O.randomized_svd(A, 1e-10, initial_rank=16)
    assert U.shape[1] == 7
    assert np.linalg.norm(A - U @ V) / np.linalg.norm(A) <= 1e-13
    assert np.allclose(np.linalg.svd(U @ V, compute_uv=False)[:7], s, rtol=1e-12)


@pytest.mark.parametrize("initial", [1, 2, 16])
def test_randomized_svd_adaptive_growth(initial):
    n = 80
    s = 0.7 ** np.arange(n)
    A = _prescribed(100, n, s, 3)
    rtol = 1e-7
    widths = []

    def omega(nn, kk):
        widths.append(kk)
        return np.random.default_rng(len(widths)).standard_normal((nn, kk))

    U, V, rel = O.randomized_svd(A, rtol, initial_rank=initial, oversampling=4, omega=omega)
    r = U.shape[1]
    true = np.linalg.norm(A - U @ V) / np.linalg.norm(A)
    assert true <= rtol * (1 + 1e-3)
    assert abs(rel ** 2 - true ** 2) <= 64 * n * np.finfo(float).eps
    assert _expected_rank(s, rtol, n) <= r <= widths[-1]
    # the target doubles (target + 1 at least) until a rank meets rtol
    t = initial
    for w in widths:
        assert w == min(n, t + 4)
        t = min(n, max(t + 1, 2 * t))


def test_zero_block_and_seed():
    A = np.zeros((5, 7))
    for f in (O.truncated_svd, O.randomized_svd):
        U, V, rel = f(A, 1e-6)
        assert U.shape == (5, 0) and V.shape == (0, 7) and rel == 0.0
    # 73856093 ^ 2*19349663 ^ 3*83492791 ^ 4*2654435761, masked to 31 bits
    v = (1 * 73856093) ^ (2 * 19349663) ^ (3 * 83492791) ^ (4 * 2654435761)
    assert block_random_seed(1, 2, 3, 4) == O.block_random_seed(1, 2, 3, 4) == v & 0x7FFFFFFF
    assert block_random_seed(0, 0, 0, 0) == 0
