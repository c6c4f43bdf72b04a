"""Smoothed-aggregation AMG hierarchy for the two Poisson solves.

The reference preconditions both FEM solves with PETSc's GAMG
(Demag_Lindholm_GPU.py:216-226 for u1, 271-281 for u2: ``ksp.getPC().
setType("gamg")`` with PETSc defaults).  GAMG is smoothed aggregation: a
strength graph, aggregates, a tentative prolongator from the near-nullspace
(the constants for a scalar Laplacian), one damped-Jacobi smoothing step of the
prolongator with w = 4/(3 rho(D^-1 A)), Galerkin coarse operators, and a
V-cycle with Jacobi-type smoothing.  This module builds that hierarchy on the
host (once per mesh) and folds each level of the V-cycle into three sparse
operators the device solver applies without any other work:

    S = 2w D^-1 - w^2 D^-1 A D^-1      pre-smooth + residual + post-smooth
    Q = (I - w D^-1 A) P               prolongation, then post-smoothing
    R = Q^T                            pre-smoothing, residual, restriction

so that M^-1 b = S b + Q M_c^-1 (R b) is one damped-Jacobi pre- and
post-smoothing sweep per level around the coarse correction (symmetric, so CG
can use it).  The coarsest level (at most ``max_coarse`` unknowns) is solved
with a dense pseudo-inverse, which also covers the singular Neumann operator of
u1 (its nullspace, the constants, is preserved on every level because the
tentative prolongator is built from them).

The aggregation follows the classical three-pass scheme of Vanek, Mandel and
Brezina ("Algebraic multigrid by smoothed aggregation for second and fourth
order elliptic problems", Computing 56, 1996): roots whose strong
neighbourhood is still free form aggregates with it, remaining nodes join a
neighbouring aggregate, and leftovers form aggregates of their own.  Rows with
no strong off-diagonal (the Dirichlet identity rows of u2) stay unaggregated:
the smoother solves them exactly up to the damping.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp

from . import _capi


def strength_graph(A: sp.csr_matrix, theta: float = 0.0) -> sp.csr_matrix:
    """Symmetric strength of connection: |a_ij| > theta sqrt(|a_ii a_jj|), i != j."""
    A = sp.csr_matrix(A)
    n = A.shape[0]
    d = np.abs(A.diagonal())
    r = np.repeat(np.arange(n), np.diff(A.indptr))
    c = A.indices
    v = A.data
    keep = (r != c) & (v != 0.0) & (np.abs(v) > theta * np.sqrt(d[r] * d[c]))
    S = sp.csr_matrix((np.ones(int(keep.sum()), dtype=np.int8), (r[keep], c[keep])), shape=(n, n))
    S.sort_indices()
    return S


def aggregate(S: sp.csr_matrix) -> tuple[np.ndarray, int]:
    """Aggregate id of every node (-1: unaggregated, isolated) and the count.

    Pass 1 visits nodes in order: a free node whose strong neighbours are all
    free becomes a root and takes them.  Pass 2 puts each remaining node with
    an aggregated neighbour into the first such neighbour's aggregate (the
    aggregate ids of pass 1).  Pass 3 turns each node still free into a new
    aggregate together with its free neighbours.  Aggregates are numbered in
    order of creation, so coarse unknowns follow the fine row order."""
    n = S.shape[0]
    ip = S.indptr.tolist()
    ix = S.indices.tolist()
    agg = [-1] * n
    na = 0
    for i in range(n):
        if agg[i] >= 0:
            continue
        a, b = ip[i], ip[i + 1]
        if a == b:
            continue
        nb = ix[a:b]
        if all(agg[j] < 0 for j in nb):
            agg[i] = na
            for j in nb:
                agg[j] = na
            na += 1
    first = list(agg)
    for i in range(n):
        if first[i] >= 0:
            continue
        for j in ix[ip[i]:ip[i + 1]]:
            if first[j] >= 0:
                agg[i] = first[j]
                break
    for i in range(n):
        if agg[i] >= 0:
            continue
        a, b = ip[i], ip[i + 1]
        if a == b:
            continue
        agg[i] = na
        for j in ix[a:b]:
            if agg[j] < 0:
                agg[j] = na
        na += 1
    return np.asarray(agg, dtype=np.int64), na


def spectral_radius_dinv_a(A: sp.csr_matrix, iters: int = 30, seed: int = 0) -> float:
    """rho(D^-1 A) by power iteration on the symmetric D^-1/2 A D^-1/2
    (deterministic start vector)."""
    d = A.diagonal()
    s = 1.0 / np.sqrt(np.abs(d))
    B = sp.diags(s) @ A @ sp.diags(s)
    v = np.random.default_rng(seed).standard_normal(A.shape[0])
    v /= np.linalg.norm(v)
    lam = 0.0
    for _ in range(iters):
        w = B @ v
        lam = float(v @ w)
        nw = np.linalg.norm(w)
        if nw == 0.0:
            return 1.0
        v = w / nw
    return abs(lam)


def tentative_prolongator(agg: np.ndarray, na: int, B: np.ndarray):
    """T (n x na) and the coarse near-nullspace: column j of T is B restricted
    to aggregate j, normalized (the one-vector case of the per-aggregate QR of
    smoothed aggregation), so T B_c = B on every aggregated node."""
    n = agg.size
    rows = np.nonzero(agg >= 0)[0]
    norm = np.sqrt(np.bincount(agg[rows], weights=B[rows] ** 2, minlength=na))
    T = sp.csr_matrix((B[rows] / norm[agg[rows]], (rows, agg[rows])), shape=(n, na))
    return T, norm


@dataclass
class HostLevel:
    n: int
    S: sp.csr_matrix
    Q: sp.csr_matrix | None = None
    R: sp.csr_matrix | None = None
    omega: float = 0.0
    rho: float = 0.0


# Smoother of the V-cycle: damped Jacobi with weight SMOOTH_WEIGHT / rho(D^-1 A)
# (the prolongator keeps the smoothed-aggregation weight 4/3 / rho), one sweep
# before and one after the coarse correction (COARSE_SWEEPS on the levels
# below the finest).  Measured with host PCG to rtol 1e-6 from zero in the
# solver row order: weight 4/3 -> 1.6 takes the film's u1 from 11 to 10
# iterations, the 150 nm cube's u1 from 11 to 10 and its u2 from 10 to 9,
# with the operators' widths unchanged.  Two sweeps on the coarse levels take
# the film's u1 to 9 iterations on the device, but the folded level-1
# smoother grows from 24 to 229 entries per row and an iteration from 47 to
# 61.5 us: slower, so one sweep.  The V-cycle stays SPD while the weight is
# below 2 / rho; the power-iteration rho is within 5 % of the true one on
# these operators, so the effective weight is at most 1.68.
SMOOTH_WEIGHT = 1.6
COARSE_SWEEPS = 1

# Strength threshold of the aggregation and the largest level solved by the
# dense coarse inverse.  theta 0.02 keeps the film's hierarchy at three
# levels with a 1174-row coarse level (1600 >= 1174) and takes its u1 from 10
# to 9 AMG-CG iterations (the 150 nm cube: 10 -> 9); theta 0 with 1024 rows
# gives 202005 -> 25339 -> 856.  A larger theta adds a level on the film
# (two more launches per iteration) for the same 9 iterations.
THETA = 0.02
MAX_COARSE = 1600


def build_levels(A: sp.csr_matrix, max_coarse: int = MAX_COARSE, theta: float = THETA,
                 max_levels: int = 12, smooth_weight: float = SMOOTH_WEIGHT,
                 coarse_sweeps: int = COARSE_SWEEPS) -> list[HostLevel]:
    """Folded V-cycle operators of every level (host, FP64).  With nu damped
    Jacobi sweeps from a zero guess the pre-smoother is one symmetric operator
    Sm = sum_{j<nu} E^j w D^-1 (E = I - w D^-1 A), the post-smoother the same
    Sm applied to the residual, so S = 2 Sm - Sm A Sm and Q = (I - Sm A) P."""
    A = sp.csr_matrix(A, dtype=np.float64)
    A.sort_indices()
    levels: list[HostLevel] = []
    B = np.ones(A.shape[0])                 # near-nullspace: the constants
    while A.shape[0] > max_coarse and len(levels) + 1 < max_levels:
        n = A.shape[0]
        agg, na = aggregate(strength_graph(A, theta))
        if na < 8 or na >= n:
            # a coarse level of a few aggregates is mostly nullspace: solve
            # this level densely instead
            break
        T, B = tentative_prolongator(agg, na, B)
        rho = spectral_radius_dinv_a(A)
        w = float(smooth_weight) / rho
        Dinv = sp.diags(1.0 / A.diagonal())
        DA = (Dinv @ A).tocsr()
        P = (T - (4.0 / (3.0 * rho)) * (DA @ T)).tocsr()
        nu = 1 if not levels else max(1, int(coarse_sweeps))
        Sm = (w * Dinv).tocsr()                  # sum_{j<nu} E^j w D^-1
        Ej = sp.identity(n, format="csr")
        for _ in range(nu - 1):
            Ej = (Ej - w * (DA @ Ej)).tocsr()
            Sm = (Sm + Ej @ (w * Dinv)).tocsr()
        SmA = (Sm @ A).tocsr()
        S = (2.0 * Sm - SmA @ Sm).tocsr()
        S = (S + S.T) * 0.5                      # exact symmetry
        Q = (P - SmA @ P).tocsr()
        for M in (S, Q, P):
            M.sum_duplicates()
            M.sort_indices()
        R = Q.T.tocsr()
        R.sort_indices()
        levels.append(HostLevel(n=n, S=S, Q=Q, R=R, omega=w, rho=rho))
        A = (P.T @ A @ P).tocsr()
        A = (A + A.T) * 0.5          # exact symmetry of the Galerkin product
        A.sum_duplicates()
        A.sort_indices()
    # pseudo-inverse from the symmetric eigendecomposition: eigenvalues below
    # 1e-12 of the largest are the nullspace (u1's constants) and are dropped
    lam, V = np.linalg.eigh(A.toarray())
    keep = lam > 1e-12 * max(abs(lam).max(), 1e-300)
    Z = (V[:, keep] / lam[keep]) @ V[:, keep].T
    Z = 0.5 * (Z + Z.T)
    # the dense coarse inverse as full CSR rows (explicit zeros kept)
    nc = A.shape[0]
    Zc = sp.csr_matrix((Z.ravel(), np.tile(np.arange(nc), nc), np.arange(nc + 1) * nc), shape=(nc, nc))
    levels.append(HostLevel(n=nc, S=Zc))
    return levels


def apply_vcycle(levels: list[HostLevel], b: np.ndarray, l: int = 0) -> np.ndarray:
    """M^-1 b on the host (the recursion the device solver runs)."""
    L = levels[l]
    if L.Q is None:
        return L.S @ b
    return L.S @ b + L.Q @ apply_vcycle(levels, L.R @ b, l + 1)


def lanes_per_row(*ops, per_lane: int = 2) -> int:
    """Lanes per CSR row for the device launch: about ``per_lane`` entries per
    lane (a power of two, at most a wavefront)."""
    avg = max((M.nnz / max(M.shape[0], 1)) for M in ops if M is not None)
    t = 1
    while t < 64 and per_lane * t < avg:
        t *= 2
    return t


SLICE_THREADS = 256


def sell_layout(M: sp.csr_matrix, tpr: int, dense: bool = False):
    """Sliced-ELL arrays of the device operator (smx_amg_op): 256/tpr rows per
    slice, each slice as wide as its longest row; slot j of local row i at
    off[s] + j*rps + i (tpr == 1) or off[s] + i*W_s + j.  Padding slots: value
    0, column 0.  ``dense``: full rows, no column array (slot j is column j).
    Returns (off int32, col int32 or None, val float64)."""
    M = sp.csr_matrix(M)
    M.sort_indices()
    rps = SLICE_THREADS // tpr
    n = M.shape[0]
    nsl = (n + rps - 1) // rps
    lens = np.diff(M.indptr).astype(np.int64)
    if dense and (M.nnz != n * M.shape[1] or not np.array_equal(M.indices, np.tile(np.arange(M.shape[1]), n))):
        raise ValueError("dense operator must store every column of every row in order")
    lp = np.zeros(nsl * rps, dtype=np.int64)
    lp[:n] = lens
    W = lp.reshape(nsl, rps).max(axis=1) if nsl else np.zeros(0, dtype=np.int64)
    off = np.zeros(nsl + 1, dtype=np.int64)
    off[1:] = np.cumsum(W * rps)
    if off[-1] >= 2 ** 31:
        raise ValueError("operator too large for int32 slot offsets")
    r = np.repeat(np.arange(n, dtype=np.int64), lens)
    j = np.arange(M.nnz, dtype=np.int64) - np.repeat(M.indptr[:-1].astype(np.int64), lens)
    sl, i = r // rps, r % rps
    pos = off[sl] + (j * rps + i if tpr == 1 else i * W[sl] + j)
    val = np.zeros(int(off[-1]))
    val[pos] = M.data
    col = None
    if not dense:
        col = np.zeros(int(off[-1]), dtype=np.int32)
        col[pos] = M.indices
    return off.astype(np.int32), col, val


def sell_matvec(off, col, val, rows: int, tpr: int, v: np.ndarray) -> np.ndarray:
    """Host restatement of the device slice loop (layout check)."""
    rps = SLICE_THREADS // tpr
    y = np.zeros(rows)
    for s in range(len(off) - 1):
        W = (int(off[s + 1]) - int(off[s])) // rps
        for i in range(rps):
            row = s * rps + i
            if row >= rows:
                break
            jj = np.arange(W)
            pos = off[s] + (jj * rps + i if tpr == 1 else i * W + jj)
            c = jj if col is None else col[pos]
            y[row] = float(val[pos] @ v[c])
    return y


class AmgOp(C.Structure):
    _fields_ = [("slice_off", C.c_void_p), ("col", C.c_void_p), ("val", C.c_void_p), ("rows", C.c_int64)]


class AmgLevel(C.Structure):
    _fields_ = [("n", C.c_int64), ("S", AmgOp), ("Q", AmgOp), ("R", AmgOp),
                ("s_tpr", C.c_int32), ("r_tpr", C.c_int32),
                ("b", C.c_void_p), ("x", C.c_void_p)]


class AmgHierarchy:
    """Device copy of the folded V-cycle (smx_pcg_amg's ``levels``)."""

    def __init__(self, A: sp.csr_matrix, device, fp32: bool = False, max_coarse: int = MAX_COARSE,
                 theta: float = THETA, host_levels: list[HostLevel] | None = None):
        import torch
        self.device = device
        self.fp32 = bool(fp32)
        self.host = host_levels if host_levels is not None else build_levels(A, max_coarse, theta)
        vdt = torch.float32 if self.fp32 else torch.float64
        self.val_dtype = _capi.SMX_F32 if self.fp32 else _capi.SMX_F64
        self._keep = []
        nlev = len(self.host)
        self.c_levels = (AmgLevel * nlev)()
        self.sizes = [L.n for L in self.host]
        for l, L in enumerate(self.host):
            c = self.c_levels[l]
            c.n = L.n
            # the fine restriction runs inside the CG update launch, whose every
            # workgroup first re-reduces the partial sums: fewer, fuller
            # workgroups there.  The fine S and Q (a few entries per row) run
            # one row per lane.
            c.s_tpr = 1 if l == 0 and L.Q is not None else lanes_per_row(L.S, L.Q)
            c.r_tpr = lanes_per_row(L.R, per_lane=8 if l == 0 else 2) if L.Q is not None else 1
            c.S = self._op(L.S, vdt, c.s_tpr, dense=L.Q is None)
            if L.Q is not None:
                c.Q = self._op(L.Q, vdt, c.s_tpr)
                c.R = self._op(L.R, vdt, c.r_tpr)
            if l > 0:
                b = torch.zeros(L.n, dtype=torch.float64, device=device)
                x = torch.zeros(L.n, dtype=torch.float64, device=device)
                self._keep += [b, x]
                c.b = b.data_ptr()
                c.x = x.data_ptr()
        self.nbytes = sum(t.numel() * t.element_size() for t in self._keep)

    def _op(self, M: sp.csr_matrix, vdt, tpr: int, dense: bool = False) -> AmgOp:
        """Upload one operator in the sliced-ELL layout for ``tpr`` lanes per
        row; ``dense``: full rows without column indices (the coarse inverse)."""
        import torch
        off, col, val = sell_layout(M, tpr, dense)
        offd = torch.from_numpy(off).to(self.device)
        vald = torch.from_numpy(val).to(self.device).to(vdt)
        self._keep += [offd, vald]
        cp = None
        if col is not None:
            cold = torch.from_numpy(col).to(self.device)
            self._keep.append(cold)
            cp = cold.data_ptr()
        return AmgOp(offd.data_ptr(), cp, vald.data_ptr(), M.shape[0])

    @property
    def nlev(self) -> int:
        return len(self.host)

    def describe(self) -> str:
        parts = []
        for L in self.host:
            nnz = L.S.nnz + (L.Q.nnz + L.R.nnz if L.Q is not None else 0)
            parts.append(f"{L.n}({nnz / max(L.n, 1):.1f}/row)")
        return " -> ".join(parts)
