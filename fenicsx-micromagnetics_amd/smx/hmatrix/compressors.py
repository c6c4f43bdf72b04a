"""Truncated-SVD and adaptive randomized-SVD compression of batches of
admissible blocks on the device.

Counterparts of truncated_svd_gpu / randomized_svd_gpu
(lindholm_hmatrix_gpu/compressors.py:373-492).  The blocks of one builder
batch sit row-major in one device buffer (the batched Lindholm fill's
output); every block of a batch is compressed by one launch of the library's
own kernels, one workgroup per block (`smx_svd_batched`: pivoted Householder
QR + one-sided Jacobi SVD; `smx_rsvd_round`: sketch products, Jacobi
orthonormalization and the projected Jacobi SVD), whatever the block shapes.
The rank rule is the reference's: the first rank r whose captured singular
energy leaves sqrt(||A||_F^2 - sum_{i<=r} s_i^2) <= rtol ||A||_F, else the
largest usable rank.  Factors go to the builder's U/V buffers in the layout
of the ACA kernel: column k of U (scaled by s_k) at u_off + k m, row k of V
at v_off + k n, for ranks up to the task's rcap.

The randomized test matrices come from a device generator seeded per block
(builder.py:170-183 seed mixing), so a block's factors do not depend on the
batch it is compressed in; they are not the reference's CuPy stream.  Both
kernels resolve ranks up to SMX_SVD_KMAX (128): a block that needs more QR
steps to reach the truncated SVD's floor, or a sketch wider than 128, comes
back unresolved (status 4, infinite residual) and the builder subdivides it.
"""
from __future__ import annotations

import numpy as np
import torch

from .. import _capi

STATUS_STORED, STATUS_OVER_CAP, STATUS_ZERO, STATUS_FAILED, STATUS_WIDER = 0, 2, 3, 4, 7

SVD_TASK_DTYPE = np.dtype([("a_off", np.int64), ("u_off", np.int64), ("v_off", np.int64), ("w_off", np.int64),
                           ("om_off", np.int64), ("m", np.int32), ("n", np.int32), ("rcap", np.int32),
                           ("mr", np.int32), ("s", np.int32), ("pad", np.int32)])
assert SVD_TASK_DTYPE.itemsize == 64


def block_random_seed(r0: int, r1: int, c0: int, c1: int) -> int:
    """builder.py:170-183."""
    v = int(r0) * 73856093 ^ int(r1) * 19349663 ^ int(c0) * 83492791 ^ int(c1) * 2654435761
    return int(v & 0x7FFFFFFF)


def _max_rank(m, n, max_rank):
    return min(m, n) if max_rank is None else min(max(int(max_rank), 0), m, n)


def _task_table(tasks, ids, max_rank, widths):
    """Device task table for the blocks `ids` of an ACA-layout task array;
    widths[q] is the sketch width (0: truncated SVD).  Returns the table and
    the workspace size in doubles."""
    L = _capi.lib()
    st = np.zeros(len(ids), dtype=SVD_TASK_DTYPE)
    for f in ("a_off", "u_off", "v_off", "m", "n", "rcap"):
        st[f] = tasks[f][ids]
    st["mr"] = [_max_rank(int(m), int(n), max_rank) for m, n in zip(st["m"], st["n"])]
    st["s"] = widths
    ws = np.array([L.smx_svd_ws_doubles(int(m), int(n), int(s)) for m, n, s in zip(st["m"], st["n"], widths)],
                  dtype=np.int64)
    if np.any(ws < 0):
        raise ValueError("invalid block shape or sketch width for the SVD compressors")
    st["w_off"] = np.concatenate([[0], np.cumsum(ws)[:-1]])
    return st, int(ws.sum())


def _outputs(nt, dev):
    return (torch.empty(nt, dtype=torch.int32, device=dev), torch.empty(nt, dtype=torch.float64, device=dev),
            torch.empty(nt, dtype=torch.int32, device=dev))


def svd_batched(A, tasks, U, V, rtol: float, max_rank=None):
    """Truncated SVD of every task's block (compressors.py:373-414) in one
    launch.  Zero blocks (and max_rank 0) give status 3 with rank 0; a block
    that the QR does not resolve within 128 steps gives status 4 with an
    infinite residual, so the builder subdivides it or keeps it dense.  The
    blocks in A are overwritten."""
    nt = int(tasks.size)
    ranks, rels, status = np.zeros(nt, np.int32), np.zeros(nt), np.zeros(nt, np.int32)
    if nt == 0:
        return ranks, rels, status
    dev = A.device
    st, nws = _task_table(tasks, np.arange(nt), max_rank, np.zeros(nt, np.int32))
    st_d = torch.from_numpy(st.view(np.uint8)).to(dev)
    work = torch.empty(max(nws, 1), dtype=torch.float64, device=dev)
    rk, rl, ss = _outputs(nt, dev)
    _capi.check(_capi.lib().smx_svd_batched(A.data_ptr(), st_d.data_ptr(), nt, float(rtol), U.data_ptr(),
                                            V.data_ptr(), work.data_ptr(), rk.data_ptr(), rl.data_ptr(),
                                            ss.data_ptr(), _capi.stream_ptr()), "smx_svd_batched")
    return rk.cpu().numpy(), rl.cpu().numpy(), ss.cpu().numpy()


def rsvd_batched(A, tasks, U, V, rtol: float, seeds, max_rank=None, initial_rank=16, oversampling=8,
                 power_iterations=1):
    """Adaptive randomized SVD of every task's block (compressors.py:417-492):
    each round is one launch over all blocks still open, each with its own
    sketch width min(mr, target + oversampling); a block whose sketch finds
    no rank meeting rtol (status 7) doubles its target and returns in the next
    round.  Each block's test matrices come from its own generator, seeded
    with seeds[i]."""
    nt = int(tasks.size)
    ranks, rels, status = np.zeros(nt, np.int32), np.zeros(nt), np.zeros(nt, np.int32)
    dev = A.device
    initial_rank = max(1, int(initial_rank))
    oversampling = max(0, int(oversampling))
    power_iterations = max(0, int(power_iterations))
    gens, target, mr = {}, {}, {}
    active = []
    for i in range(nt):
        mr[i] = _max_rank(int(tasks["m"][i]), int(tasks["n"][i]), max_rank)
        if mr[i] == 0:
            status[i] = STATUS_ZERO
            continue
        g = torch.Generator(device=dev)
        g.manual_seed(int(seeds[i]))
        gens[i] = g
        target[i] = min(initial_rank, mr[i])
        active.append(i)
    L = _capi.lib()
    while active:
        ids = np.array(active, dtype=np.int64)
        widths = np.array([min(mr[i], target[i] + oversampling, _capi.SVD_KMAX) for i in active], np.int32)
        oms = [torch.randn(int(tasks["n"][i]), int(k), generator=gens[i], device=dev, dtype=torch.float64)
               .reshape(-1) for i, k in zip(active, widths)]
        st, nws = _task_table(tasks, ids, max_rank, widths)
        st["om_off"] = np.concatenate([[0], np.cumsum([o.numel() for o in oms])[:-1]])
        omega = torch.cat(oms)
        st_d = torch.from_numpy(st.view(np.uint8)).to(dev)
        work = torch.empty(max(nws, 1), dtype=torch.float64, device=dev)
        rk, rl, ss = _outputs(len(active), dev)
        _capi.check(L.smx_rsvd_round(A.data_ptr(), st_d.data_ptr(), len(active), omega.data_ptr(), float(rtol),
                                     power_iterations, U.data_ptr(), V.data_ptr(), work.data_ptr(), rk.data_ptr(),
                                     rl.data_ptr(), ss.data_ptr(), _capi.stream_ptr()), "smx_rsvd_round")
        rk, rl, ss = rk.cpu().numpy(), rl.cpu().numpy(), ss.cpu().numpy()
        nxt = []
        for q, i in enumerate(active):
            if ss[q] == STATUS_WIDER:
                if widths[q] >= _capi.SVD_KMAX:
                    ranks[i], rels[i], status[i] = 0, np.inf, STATUS_FAILED
                else:
                    target[i] = min(mr[i], max(target[i] + 1, 2 * target[i]))
                    nxt.append(i)
                continue
            ranks[i], rels[i], status[i] = rk[q], rl[q], ss[q]
        del omega, oms, work
        active = nxt
    return ranks, rels, status


def workspace_multiplier(method: str) -> float:
    """Device workspace per byte of dense block (compressors.py:575-598)."""
    return {"fullaca": 3.5, "svd": 4.0, "rsvd": 2.0}.get(method, 3.0)
