"""Row order of the CG solves (host logic): a permutation, and the node order or
its reverse Cuthill-McKee order, whichever leaves fewer matrix entries across
the eight contiguous row ranges the CG kernel gives the eight XCDs."""
import numpy as np
import scipy.sparse as sp
import torch

from smx import mesh as M
from smx.demag import OperatorSet


def _crossing(ops, order):
    g = ops.fem.graph
    N = ops.N
    pos = np.empty(N, dtype=np.int64)
    pos[order] = np.arange(N)
    r = np.repeat(np.arange(N), np.diff(g.indptr))
    chunk = (N + 7) // 8
    return int(np.count_nonzero(pos[r] // chunk != pos[g.indices] // chunk))


def test_film_like_box_takes_rcm():
    mesh = M.kuhn_box(40, 40, 2, 100.0, 100.0, 5.0)     # x-fastest numbering, thin in z
    ops = OperatorSet(mesh, torch.device("cpu"))
    order = ops.solver_order()
    assert np.array_equal(np.sort(order), np.arange(ops.N))
    assert not np.array_equal(order, np.arange(ops.N))
    assert _crossing(ops, order) < _crossing(ops, np.arange(ops.N))


def test_cube_keeps_the_better_order():
    mesh = M.kuhn_box(8, 8, 8, 32.0, 32.0, 32.0)
    ops = OperatorSet(mesh, torch.device("cpu"))
    order = ops.solver_order()
    assert np.array_equal(np.sort(order), np.arange(ops.N))
    assert _crossing(ops, order) <= _crossing(ops, np.arange(ops.N))


def test_permuted_operator_is_p_a_pt():
    mesh = M.kuhn_box(6, 5, 2, 12.0, 10.0, 4.0)
    ops = OperatorSet(mesh, torch.device("cpu"))
    fem = ops.fem
    perm = ops.solver_order()
    E = ops.compact(fem.stiff, perm)
    # densify the ELL and compare with P A P^T of the CSR stiffness
    W, ld = E.width, E.ld
    cols = E.cols.numpy().reshape(W, ld)[:, :ops.N]
    vals = E.vals.numpy().reshape(W, ld)[:, :ops.N]
    D = np.zeros((ops.N, ops.N))
    for j in range(W):
        np.add.at(D, (np.arange(ops.N), cols[j]), vals[j])
    g = fem.graph
    r = np.repeat(np.arange(ops.N), np.diff(g.indptr))
    A = sp.csr_matrix((fem.stiff, (r, g.indices)), shape=(ops.N, ops.N)).toarray()
    A[np.abs(A) <= 1e-14 * np.abs(A).max(axis=1, keepdims=True)] = 0.0
    assert np.abs(D - A[np.ix_(perm, perm)]).max() == 0.0
