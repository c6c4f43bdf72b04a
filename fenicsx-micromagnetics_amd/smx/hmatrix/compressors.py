"""Truncated-SVD and adaptive randomized-SVD compression of batches of
admissible blocks on the device.

Counterparts of truncated_svd_gpu / randomized_svd_gpu
(lindholm_hmatrix_gpu/compressors.py:373-492).  The blocks of one builder
batch sit row-major in one device buffer (the batched Lindholm fill's
output); blocks of equal shape are compressed together by batched library
calls on the device: rocBLAS GEMMs for the sketches and projections, rocSOLVER
for the QR factorizations and the SVDs (the reference calls cuBLAS/cuSOLVER
the same way).  The rank rule is the reference's: the first rank r whose
captured singular energy leaves sqrt(||A||_F^2 - sum_{i<=r} s_i^2) <=
rtol ||A||_F, else the largest usable rank.  Factors go to the builder's U/V
buffers in the layout of the ACA kernel: column k of U (scaled by s_k) at
u_off + k m, row k of V at v_off + k n, for ranks up to the task's rcap.

The randomized test matrices come from a device generator seeded per block
(builder.py:170-183 seed mixing), so a block's factors do not depend on the
batch it is compressed in; they are not the reference's CuPy stream.
"""
from __future__ import annotations

import math

import numpy as np
import torch

STATUS_STORED, STATUS_OVER_CAP, STATUS_ZERO, STATUS_FAILED = 0, 2, 3, 4


def block_random_seed(r0: int, r1: int, c0: int, c1: int) -> int:
    """builder.py:170-183."""
    v = int(r0) * 73856093 ^ int(r1) * 19349663 ^ int(c0) * 83492791 ^ int(c1) * 2654435761
    return int(v & 0x7FFFFFFF)


def _rank_rule(s: np.ndarray, norm_sq: float, rtol: float, max_rank: int):
    usable = min(int(max_rank), int(s.size))
    captured = np.cumsum(s[:usable] * s[:usable])
    rels = np.sqrt(np.maximum(norm_sq - captured, 0.0)) / math.sqrt(norm_sq)
    ok = np.flatnonzero(rels <= rtol)
    rank = int(ok[0] + 1) if ok.size else usable
    rel = float(rels[rank - 1]) if rank else 1.0
    return rank, rel, bool(ok.size)


def _groups(tasks, idx):
    out = {}
    for i in idx:
        out.setdefault((int(tasks["m"][i]), int(tasks["n"][i])), []).append(int(i))
    return out


def _stack(A, tasks, ids, m, n):
    return torch.stack([A[int(tasks["a_off"][i]):int(tasks["a_off"][i]) + m * n].view(m, n) for i in ids])


def _store(U, V, tasks, i, m, n, Uf, Vf, rank):
    """Uf: (m, rank) scaled left factor; Vf: (rank, n)."""
    if rank == 0:
        return
    u0, v0 = int(tasks["u_off"][i]), int(tasks["v_off"][i])
    U[u0:u0 + rank * m].view(rank, m).copy_(Uf.transpose(0, 1))
    V[v0:v0 + rank * n].view(rank, n).copy_(Vf)


def _max_rank(m, n, max_rank):
    return min(m, n) if max_rank is None else min(max(int(max_rank), 0), m, n)


def _svd(X):
    """Batched SVD as (U, s, Vh, s on the host) per element (one host copy of
    the singular values per batch); where the library's iteration does not
    converge for some element, the batch is redone element by element and a
    failing element comes back as None."""
    try:
        Us, s, Vh = torch.linalg.svd(X, full_matrices=False)
        sh = s.cpu().numpy()
        return [(Us[b], s[b], Vh[b], sh[b]) for b in range(X.shape[0])]
    except RuntimeError:
        out = []
        for b in range(X.shape[0]):
            try:
                u, sv, vh = torch.linalg.svd(X[b], full_matrices=False)
                out.append((u, sv, vh, sv.cpu().numpy()))
            except RuntimeError:
                out.append(None)
        return out


def svd_batched(A, tasks, U, V, rtol: float, max_rank=None):
    """Truncated SVD of every task's block (compressors.py:373-414).  Zero
    blocks never reach the SVD (the library reports non-convergence on an
    exactly zero matrix); a block whose SVD does not converge is returned
    with an infinite residual (status 4), so the builder subdivides it or
    keeps it dense."""
    nt = int(tasks.size)
    ranks = np.zeros(nt, np.int32)
    rels = np.zeros(nt)
    status = np.zeros(nt, np.int32)
    for (m, n), ids in _groups(tasks, range(nt)).items():
        X = _stack(A, tasks, ids, m, n)
        norm_sq = (X * X).sum(dim=(1, 2)).cpu().numpy()
        mr = _max_rank(m, n, max_rank)
        live = [b for b in range(len(ids)) if norm_sq[b] > 0.0 and mr > 0]
        for b in sorted(set(range(len(ids))) - set(live)):
            ranks[ids[b]], rels[ids[b]], status[ids[b]] = 0, 0.0, STATUS_ZERO
        if not live:
            continue
        res = _svd(X[live] if len(live) < len(ids) else X)
        for q, b in enumerate(live):
            i = ids[b]
            if res[q] is None:
                ranks[i], rels[i], status[i] = 0, math.inf, STATUS_FAILED
                continue
            Ub_, sb_, Vhb_, sh_ = res[q]
            rank, rel, _ = _rank_rule(sh_, float(norm_sq[b]), rtol, mr)
            ranks[i], rels[i] = rank, rel
            if rank > int(tasks["rcap"][i]):
                status[i] = STATUS_OVER_CAP
                continue
            _store(U, V, tasks, i, m, n, Ub_[:, :rank] * sb_[:rank][None, :], Vhb_[:rank, :], rank)
        del X, res
    return ranks, rels, status


def rsvd_batched(A, tasks, U, V, rtol: float, seeds, max_rank=None, initial_rank=16, oversampling=8,
                 power_iterations=1):
    """Adaptive randomized SVD of every task's block (compressors.py:417-492):
    per round, blocks of equal shape and sketch width go through one batched
    sketch / power iteration / QR / projected SVD; a block whose sketch finds
    no rank meeting rtol doubles its target and returns in the next round."""
    nt = int(tasks.size)
    ranks = np.zeros(nt, np.int32)
    rels = np.zeros(nt)
    status = np.zeros(nt, np.int32)
    dev = A.device
    initial_rank = max(1, int(initial_rank))
    oversampling = max(0, int(oversampling))
    power_iterations = max(0, int(power_iterations))
    gens, target, mr, norm_sq = {}, {}, {}, {}
    active = []
    for (m, n), ids in _groups(tasks, range(nt)).items():
        X = _stack(A, tasks, ids, m, n)
        ns = (X * X).sum(dim=(1, 2)).cpu().numpy()
        del X
        for b, i in enumerate(ids):
            mr[i] = _max_rank(m, n, max_rank)
            norm_sq[i] = float(ns[b])
            if norm_sq[i] == 0.0 or mr[i] == 0:
                status[i] = STATUS_ZERO
                continue
            g = torch.Generator(device=dev)
            g.manual_seed(int(seeds[i]))
            gens[i] = g
            target[i] = min(initial_rank, mr[i])
            active.append(i)
    while active:
        keyed = {}
        for i in active:
            m, n = int(tasks["m"][i]), int(tasks["n"][i])
            k = min(mr[i], target[i] + oversampling)
            keyed.setdefault((m, n, k), []).append(i)
        nxt = []
        for (m, n, k), ids in keyed.items():
            X = _stack(A, tasks, ids, m, n)
            Om = torch.stack([torch.randn(n, k, generator=gens[i], device=dev, dtype=torch.float64) for i in ids])
            Y = torch.matmul(X, Om)
            Xt = X.transpose(1, 2).contiguous()
            for _ in range(power_iterations):
                Q, _ = torch.linalg.qr(Y, mode="reduced")
                Y = torch.matmul(X, torch.matmul(Xt, Q))
            Q, _ = torch.linalg.qr(Y, mode="reduced")
            Bm = torch.matmul(Q.transpose(1, 2).contiguous(), X)
            res = _svd(Bm)
            for b, i in enumerate(ids):
                if res[b] is None:
                    ranks[i], rels[i], status[i] = 0, math.inf, STATUS_FAILED
                    continue
                Ub_, sb_, Vhb_, sh_ = res[b]
                rank, rel, ok = _rank_rule(sh_, norm_sq[i], rtol, mr[i])
                if not ok and k < mr[i]:
                    target[i] = min(mr[i], max(target[i] + 1, 2 * target[i]))
                    nxt.append(i)
                    continue
                ranks[i], rels[i] = rank, rel
                if rank > int(tasks["rcap"][i]):
                    status[i] = STATUS_OVER_CAP
                    continue
                Uf = torch.matmul(Q[b], Ub_[:, :rank]) * sb_[:rank][None, :]
                _store(U, V, tasks, i, m, n, Uf, Vhb_[:rank, :], rank)
            del X, Xt, Om, Y, Q, Bm, res
        active = nxt
    return ranks, rels, status


def workspace_multiplier(method: str) -> float:
    """Device workspace per byte of dense block (compressors.py:575-598)."""
    return {"fullaca": 3.5, "svd": 4.0, "rsvd": 2.0}.get(method, 3.0)
