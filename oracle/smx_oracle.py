"""FP64 CPU oracle for the Fredkin-Koehler demag + LLG RHS hot path.

TEST INFRASTRUCTURE ONLY.  This module is imported by ``tests/``, by
``__graft_entry__.smoke()`` (as the checker) and by the ``cpu_baseline`` leg of
``bench.py``.  The product path (the ``smx`` package and ``libsmx_hip.so``) never
imports, links or calls it.

It restates, in plain numpy/scipy float64, the reference's algorithm for every
function on the path.  Each function names the reference file:line it follows
(paths relative to the reference's ``src/spinfemx``).  Where the reference calls
DOLFINx/UFL/PETSc, the oracle evaluates the same weak forms with exact P1
element integrals and uses direct sparse LU (``scipy.sparse.linalg.splu``), as
the reference's CPU path does below 60 000 dofs
(cpu/fields/Demag/Demag_Lindholm.py:669-680).

How each part is pinned (see tests/test_oracle_*.py):
  solid_angle         closed forms (octant = pi/2, face = 2 pi ...), sign
  lindholm_weights    sum_k w_k = -Omega/4pi; brute-force Gauss quadrature of
                      (1/4pi) int phi_k (x-y).n/|x-y|^3 dS on far targets;
                      scale invariance
  boundary_data       counts of the benchmark meshes; omega_sum in {pi/2, pi, 2pi}
                      on boxes; outward normals (divergence theorem)
  dense_B             B.1 = -1 per row
  fem_operators       A.1 = 0; interior Kuhn rows = h * 7-point stencil;
                      G(linear) = -grad * V; Div = -Ms G^T; sum V = volume
  demag_field         patch test (uniform m: H = -Ms m for the u1-part);
                      uniformly magnetized cube -> <H> ~ -Ms/3 (convergence)
  exchange/DMI/aniso  K 1 = 0; 7-point Laplacian; continuum formulas on linear m
  llg_rhs / jac_vec   m.r = S(1-|m|^2)|m|^2; central finite differences
  tree / partition    permutation + interval invariants, leaf sizes,
                      admissibility by brute force
  fullaca_dense       ||A - UV||_F <= eps ||A||_F ; exact rank on rank-r input
  truncated_svd       matrices of prescribed singular values: the rank is the
                      first r with sqrt(sum_{i>r} s_i^2) <= rtol ||s||, the
                      reported residual equals ||A - UV||_F / ||A||_F
  randomized_svd      exact recovery of rank-r input; reported residual equals
                      ||A - UV||_F / ||A||_F (the projection identity);
                      adaptive growth from a too-small initial rank
  build_hmatrix       workspace / OOM hooks: every far block fits, matvec
                      against dense B
  hmatrix matvec      validate_against_dense protocol (seed 1234, 3 trials)
  cubic anisotropy    field = -(1/mu0 Ms) dE/dm (central FD of the energy);
                      closed forms at m along an axis / a face diagonal
  zhang_li_matrix     linear m: Z = (J.grad) m exactly at every node; the
                      per-node J rule against the element closed form
  torques             macrospin: damping-like SOT turns m toward p; STT/SOT
                      torques orthogonal to m; JVPs by central FD
  cayley_step         Rodrigues rotation by 2 atan|c| about c; |y| = |m|
  gmres_right         solves against numpy's dense solve; restarts and the
                      right preconditioner change the iterates, not the answer
  labonte_bb_minimize macrospin in a field along the easy axis -> m = n;
                      exchange + uniaxial box -> E = -Ku V; BB rules on
                      hand-computed s, y
  bdf_integrate       fixed-step BDF2 error ratio ~4 on halving dt (second
                      order) on y' = A y against expm; the adaptive run meets
                      its tolerance there and on the damped macrospin's
                      closed-form solution
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

MU0 = 4.0 * math.pi * 1e-7
TINY = 1e-300


# --------------------------------------------------------------------------
# Lindholm kernel (gpu/fields/Lindholm_kernel.py:10-125)
# --------------------------------------------------------------------------
def _dot(a, b):
    return a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1] + a[..., 2] * b[..., 2]


def _det3(a, b, c):
    return (a[..., 0] * (b[..., 1] * c[..., 2] - b[..., 2] * c[..., 1])
            - a[..., 1] * (b[..., 0] * c[..., 2] - b[..., 2] * c[..., 0])
            + a[..., 2] * (b[..., 0] * c[..., 1] - b[..., 1] * c[..., 0]))


def solid_angle(x0, p0, p1, p2):
    """Van Oosterom-Strackee solid angle (Lindholm_kernel.py:30-47)."""
    r0 = p0 - x0
    r1 = p1 - x0
    r2 = p2 - x0
    n0 = np.sqrt(_dot(r0, r0)) + TINY
    n1 = np.sqrt(_dot(r1, r1)) + TINY
    n2 = np.sqrt(_dot(r2, r2)) + TINY
    det = _det3(r0, r1, r2)
    denom = n0 * n1 * n2 + _dot(r0, r1) * n2 + _dot(r1, r2) * n0 + _dot(r2, r0) * n1
    return 2.0 * np.arctan2(det, denom)


def _p_log1p(ri, rj, s):
    """Lindholm_kernel.py:51-61: log((ri+rj+s)/(ri+rj-s)) via log1p."""
    denom = np.maximum(ri + rj - s, TINY)
    return np.log1p((2.0 * s) / denom)


def lindholm_weights(x0, p0, p1, p2, n_unit, area, s_edge, eta, gamma):
    """(w0, w1, w2) of the linear double layer (Lindholm_kernel.py:65-125).

    Broadcasting: x0 (...,3); triangle arrays p* (...,3), n_unit (...,3),
    area (...), s_edge (...,3), eta (...,3,3), gamma (...,3,3).
    """
    rho0 = p0 - x0
    rho1 = p1 - x0
    rho2 = p2 - x0
    r0 = np.sqrt(_dot(rho0, rho0)) + TINY
    r1 = np.sqrt(_dot(rho1, rho1)) + TINY
    r2 = np.sqrt(_dot(rho2, rho2)) + TINY
    h = _dot(n_unit, rho0)
    e0 = _dot(eta[..., 0, :], rho0)
    e1 = _dot(eta[..., 1, :], rho1)
    e2 = _dot(eta[..., 2, :], rho2)
    s0 = s_edge[..., 0]
    s1 = s_edge[..., 1]
    s2 = s_edge[..., 2]
    P0 = _p_log1p(r0, r1, s0)
    P1 = _p_log1p(r1, r2, s1)
    P2 = _p_log1p(r2, r0, s2)
    omega = solid_angle(x0, p0, p1, p2)
    g = gamma
    gP0 = g[..., 0, 0] * P0 + g[..., 0, 1] * P1 + g[..., 0, 2] * P2
    gP1 = g[..., 1, 0] * P0 + g[..., 1, 1] * P1 + g[..., 1, 2] * P2
    gP2 = g[..., 2, 0] * P0 + g[..., 2, 1] * P1 + g[..., 2, 2] * P2
    c = 1.0 / (8.0 * math.pi * (area + TINY))
    # opposite-edge convention (Lindholm_kernel.py:118-123)
    w0 = s1 * c * (e1 * omega - h * gP0)
    w1 = s2 * c * (e2 * omega - h * gP1)
    w2 = s0 * c * (e0 * omega - h * gP2)
    return w0, w1, w2


# --------------------------------------------------------------------------
# Boundary data (gpu/fields/Demag_Lindholm.py:177-285)
# --------------------------------------------------------------------------
_TET_FACES = ((1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2))


def exterior_facets(tets):
    """Exterior facets as (sorted vertex triple, owning cell), ordered
    lexicographically by the sorted triple (the facet numbering used throughout)."""
    T = tets.shape[0]
    faces = []
    owner = []
    for f in _TET_FACES:
        faces.append(np.sort(tets[:, f], axis=1))
        owner.append(np.arange(T))
    faces = np.concatenate(faces).astype(np.int64)
    owner = np.concatenate(owner)
    uniq, inv, counts = np.unique(faces, axis=0, return_inverse=True, return_counts=True)
    ext = np.nonzero(counts[inv.ravel()] == 1)[0]
    order = np.lexsort((faces[ext, 2], faces[ext, 1], faces[ext, 0]))
    ext = ext[order]
    return faces[ext], owner[ext]


@dataclass
class BoundaryData:
    bverts: np.ndarray    # (Nb,) mesh node id of each boundary index
    Xb: np.ndarray
    tri_lid: np.ndarray   # (Nt,3) int32
    tri_pts: np.ndarray   # (Nt,3,3)
    ntri: np.ndarray      # (Nt,3)
    area: np.ndarray
    s_edge: np.ndarray    # (Nt,3)
    eta: np.ndarray       # (Nt,3,3)
    gamma: np.ndarray     # (Nt,3,3)
    omega_sum: np.ndarray  # (Nb,)

    @property
    def Nb(self):
        return int(self.Xb.shape[0])

    @property
    def Nt(self):
        return int(self.tri_lid.shape[0])

    @property
    def diag_jump(self):
        return self.omega_sum / (4.0 * math.pi) - 1.0


def boundary_data(X, tets) -> BoundaryData:
    """Restates build_boundary_data_serial (Demag_Lindholm.py:177-285) one facet
    at a time: outward orientation by the facet-centroid vs cell-centroid test
    (:207-214), bverts = unique(tris), edge/eta/gamma precomputation (:232-263)
    and omega_sum = sum over tets of |Omega| at each boundary vertex (:265-283)."""
    X = np.asarray(X, dtype=np.float64)
    faces, owner = exterior_facets(np.asarray(tets))
    tris = []
    normals = []
    for (v0, v1, v2), cell in zip(faces.tolist(), owner.tolist()):
        p0, p1, p2 = X[v0], X[v1], X[v2]
        n = np.cross(p1 - p0, p2 - p0)
        nn = np.linalg.norm(n)
        if nn == 0.0:
            continue
        n = n / nn
        cc = X[tets[cell]].mean(axis=0)
        fc = (p0 + p1 + p2) / 3.0
        if np.dot(fc - cc, n) < 0.0:
            v1, v2 = v2, v1
            p1, p2 = X[v1], X[v2]
            n = np.cross(p1 - p0, p2 - p0)
            n /= (np.linalg.norm(n) + 1e-30)
        tris.append((v0, v1, v2))
        normals.append(n)
    tris = np.asarray(tris, dtype=np.int64)
    ntri = np.asarray(normals, dtype=np.float64)
    bverts = np.unique(tris.ravel())
    Nb = bverts.size
    bmap = -np.ones(X.shape[0], dtype=np.int64)
    bmap[bverts] = np.arange(Nb)
    tri_lid = bmap[tris].astype(np.int32)
    tri_pts = X[tris].astype(np.float64)
    cross = np.cross(tri_pts[:, 1] - tri_pts[:, 0], tri_pts[:, 2] - tri_pts[:, 0])
    area = 0.5 * np.linalg.norm(cross, axis=1)
    e = [tri_pts[:, 1] - tri_pts[:, 0], tri_pts[:, 2] - tri_pts[:, 1], tri_pts[:, 0] - tri_pts[:, 2]]
    s = [np.linalg.norm(ei, axis=1) + 1e-30 for ei in e]
    s_edge = np.stack(s, axis=1)
    xi = np.stack([e[i] / s[i][:, None] for i in range(3)], axis=1)
    eta = np.stack([np.cross(ntri, xi[:, i]) for i in range(3)], axis=1)
    gamma = np.zeros((xi.shape[0], 3, 3))
    for i in range(3):
        for j in range(3):
            gamma[:, i, j] = np.einsum("ij,ij->i", xi[:, (i + 1) % 3], xi[:, j])
    omega_sum = np.zeros(Nb)
    for c in range(tets.shape[0]):
        vs = tets[c]
        pts = X[vs]
        for j in range(4):
            lid = bmap[int(vs[j])]
            if lid < 0:
                continue
            omega_sum[lid] += abs(float(solid_angle(pts[j], pts[(j + 1) % 4],
                                                    pts[(j + 2) % 4], pts[(j + 3) % 4])))
    return BoundaryData(bverts=bverts.astype(np.int64), Xb=X[bverts].copy(), tri_lid=tri_lid,
                        tri_pts=tri_pts, ntri=ntri, area=area, s_edge=s_edge, eta=eta,
                        gamma=gamma, omega_sum=omega_sum)


def boundary_data_vec(X, tets, omega_rows=None) -> BoundaryData:
    """boundary_data with the facet and tet loops vectorized, for meshes where
    the per-facet loop is too slow (the 83k-node film).  Same operations in
    the same order per facet, so the fields equal boundary_data's (checked
    against it in the tests).  omega_rows: boundary indices whose omega_sum is
    wanted (the sum runs over the tets touching them, in tet order); the other
    entries are NaN.  None computes all of them."""
    X = np.asarray(X, dtype=np.float64)
    tets = np.asarray(tets, dtype=np.int64)
    faces, owner = exterior_facets(tets)
    p0, p1, p2 = X[faces[:, 0]], X[faces[:, 1]], X[faces[:, 2]]
    n = np.cross(p1 - p0, p2 - p0)
    nn = np.linalg.norm(n, axis=1)
    keep = nn != 0.0
    faces, owner, p0, p1, p2, n, nn = faces[keep], owner[keep], p0[keep], p1[keep], p2[keep], n[keep], nn[keep]
    n = n / nn[:, None]
    cc = X[tets[owner]].mean(axis=1)
    fc = (p0 + p1 + p2) / 3.0
    flip = np.einsum("ij,ij->i", fc - cc, n) < 0.0
    tris = faces.copy()
    tris[flip, 1], tris[flip, 2] = faces[flip, 2], faces[flip, 1]
    nf = np.cross(X[tris[flip, 1]] - p0[flip], X[tris[flip, 2]] - p0[flip])
    n[flip] = nf / (np.linalg.norm(nf, axis=1) + 1e-30)[:, None]
    bverts = np.unique(tris.ravel())
    Nb = bverts.size
    bmap = -np.ones(X.shape[0], dtype=np.int64)
    bmap[bverts] = np.arange(Nb)
    tri_lid = bmap[tris].astype(np.int32)
    tri_pts = X[tris].astype(np.float64)
    cross = np.cross(tri_pts[:, 1] - tri_pts[:, 0], tri_pts[:, 2] - tri_pts[:, 0])
    area = 0.5 * np.linalg.norm(cross, axis=1)
    e = [tri_pts[:, 1] - tri_pts[:, 0], tri_pts[:, 2] - tri_pts[:, 1], tri_pts[:, 0] - tri_pts[:, 2]]
    sl = [np.linalg.norm(ei, axis=1) + 1e-30 for ei in e]
    xi = np.stack([e[i] / sl[i][:, None] for i in range(3)], axis=1)
    eta = np.stack([np.cross(n, xi[:, i]) for i in range(3)], axis=1)
    gamma = np.zeros((xi.shape[0], 3, 3))
    for i in range(3):
        for j in range(3):
            gamma[:, i, j] = np.einsum("ij,ij->i", xi[:, (i + 1) % 3], xi[:, j])
    want = np.ones(Nb, dtype=bool) if omega_rows is None else np.isin(np.arange(Nb), np.asarray(omega_rows))
    omega_sum = np.where(want, 0.0, np.nan)
    lid4 = bmap[tets]                                      # (T, 4)
    sel = np.nonzero((lid4 >= 0) & want[np.maximum(lid4, 0)])
    c, j = sel                                             # tet-major order
    pts = X[tets[c]]                                       # (k, 4, 3)
    ar = np.arange(c.size)
    om = np.abs(solid_angle(pts[ar, j], pts[ar, (j + 1) % 4], pts[ar, (j + 2) % 4], pts[ar, (j + 3) % 4]))
    np.add.at(omega_sum, lid4[c, j], om)
    return BoundaryData(bverts=bverts.astype(np.int64), Xb=X[bverts].copy(), tri_lid=tri_lid,
                        tri_pts=tri_pts, ntri=n, area=area, s_edge=np.stack(sl, axis=1), eta=eta,
                        gamma=gamma, omega_sum=omega_sum)


# --------------------------------------------------------------------------
# Dense B and block fill (gpu/fields/Demag_Lindholm.py:51-116;
# lindholm_hmatrix_gpu/entry_provider.py:41-117, 187-223)
# --------------------------------------------------------------------------
def _weights_rows(bd: BoundaryData, rows, tri):
    x0 = bd.Xb[rows][:, None, :]
    return lindholm_weights(x0, bd.tri_pts[tri, 0][None], bd.tri_pts[tri, 1][None],
                            bd.tri_pts[tri, 2][None], bd.ntri[tri][None], bd.area[tri][None],
                            bd.s_edge[tri][None], bd.eta[tri][None], bd.gamma[tri][None])


def dense_B_rows(bd: BoundaryData, rows=None, incident_policy=0, chunk=32):
    """Rows of the dense Lindholm B (assemble_B_rows_owned, Demag_Lindholm.py:51-116):
    diagonal omega_sum/4pi - 1, off-diagonal sum of (w0,w1,w2) over triangles not
    incident to the target (incident_policy 0 = "skip").  Accumulation order per
    row follows the reference loop (triangle-major, vertex a then b then c)."""
    Nb, Nt = bd.Nb, bd.Nt
    rows = np.arange(Nb) if rows is None else np.asarray(rows, dtype=np.int64)
    out = np.zeros((rows.size, Nb))
    tri = np.arange(Nt)
    lid = bd.tri_lid.astype(np.int64)
    for c0 in range(0, rows.size, chunk):
        rr = rows[c0:c0 + chunk]
        w0, w1, w2 = _weights_rows(bd, rr, tri)
        w = np.stack([w0, w1, w2], axis=2)  # (R, Nt, 3)
        if incident_policy == 0:
            inc = (lid[None, :, :] == rr[:, None, None]).any(axis=2)
            w[inc] = 0.0
        R = rr.size
        idx = (np.arange(R)[:, None, None] * Nb + lid[None, :, :]).ravel()
        blk = np.bincount(idx, weights=w.ravel(), minlength=R * Nb).reshape(R, Nb)
        out[c0:c0 + R] = blk
    out[np.arange(rows.size), rows] = bd.omega_sum[rows] / (4.0 * math.pi) - 1.0
    return out


def relevant_triangles(bd: BoundaryData, cols):
    """_relevant_triangles (entry_provider.py:187-223): triangles incident to at
    least one block column (sorted, unique) and the local column of each of their
    vertices (-1 if not in the block)."""
    cols = np.asarray(cols, dtype=np.int64)
    in_block = np.zeros(bd.Nb, dtype=bool)
    in_block[cols] = True
    rel = np.nonzero(in_block[bd.tri_lid].any(axis=1))[0]
    col_of = -np.ones(bd.Nb, dtype=np.int64)
    col_of[cols] = np.arange(cols.size)
    return rel.astype(np.int32), col_of[bd.tri_lid[rel]].astype(np.int32)


def block_B(bd: BoundaryData, rows, cols, incident_policy=0):
    """Off-diagonal block B[rows, cols] with the diagonal jump excluded
    (LindholmEntryProviderCPU.fill_block, entry_provider.py:225-247)."""
    rows = np.asarray(rows, dtype=np.int64)
    cols = np.asarray(cols, dtype=np.int64)
    rel, rel_cols = relevant_triangles(bd, cols)
    out = np.zeros((rows.size, cols.size))
    if rel.size == 0:
        return out
    lid = bd.tri_lid[rel].astype(np.int64)
    for c0 in range(0, rows.size, 64):
        rr = rows[c0:c0 + 64]
        w0, w1, w2 = _weights_rows(bd, rr, rel)
        w = np.stack([w0, w1, w2], axis=2)
        if incident_policy == 0:
            w[(lid[None] == rr[:, None, None]).any(axis=2)] = 0.0
        valid = rel_cols >= 0
        R = rr.size
        idx = (np.arange(R)[:, None, None] * cols.size + np.where(valid, rel_cols, 0)[None])
        wv = np.where(valid[None], w, 0.0)
        out[c0:c0 + R] = np.bincount(idx.ravel(), weights=wv.ravel(),
                                     minlength=R * cols.size).reshape(R, cols.size)
    return out


# --------------------------------------------------------------------------
# P1 FEM operators (UFL forms of gpu/fields/Demag_Lindholm_GPU.py:188-261, 554-561;
# Exchange_GPU.py:53-68; DMI_Interfacial_GPU.py:134-168; DMI_Bulk_GPU.py:248-264;
# nodal volume llg_module_GPU.py:469-487)
# --------------------------------------------------------------------------
def element_geometry(X, tets):
    """Per tet: volume and gradients of the 4 barycentric functions (T,4,3)."""
    P = X[tets]  # (T,4,3)
    J = np.stack([P[:, 1] - P[:, 0], P[:, 2] - P[:, 0], P[:, 3] - P[:, 0]], axis=2)  # columns
    det = np.linalg.det(J)
    vol = np.abs(det) / 6.0
    Jinv = np.linalg.inv(J)  # rows = grad lambda_1..3
    grads = np.empty((tets.shape[0], 4, 3))
    grads[:, 1:] = Jinv
    grads[:, 0] = -Jinv.sum(axis=1)
    return vol, grads


def _coo(rows, cols, vals, shape):
    M = sp.coo_matrix((vals.ravel(), (rows.ravel(), cols.ravel())), shape=shape).tocsr()
    M.sum_duplicates()
    return M


@dataclass
class FEMOperators:
    N: int
    A: sp.csr_matrix          # int grad v . grad u           (N x N)
    Div: sp.csr_matrix        # Ms int grad v . m             (N x 3N)
    G: sp.csr_matrix          # int v . (-grad phi)           (3N x N)
    V: np.ndarray             # lumped nodal volume int phi_i (N)
    bnodes: np.ndarray        # boundary node ids (sorted)
    A_bc: sp.csr_matrix       # A with Dirichlet rows/cols -> identity
    extra: dict = field(default_factory=dict)


def fem_operators(X, tets, Ms) -> FEMOperators:
    X = np.asarray(X, dtype=np.float64)
    tets = np.asarray(tets, dtype=np.int64)
    N = X.shape[0]
    vol, gr = element_geometry(X, tets)
    T = tets.shape[0]
    a_idx = np.repeat(tets[:, :, None], 4, axis=2)
    b_idx = np.repeat(tets[:, None, :], 4, axis=1)
    Ke = vol[:, None, None] * np.einsum("tak,tbk->tab", gr, gr)
    A = _coo(a_idx, b_idx, Ke, (N, N))
    # Div[a, 3b+c] = Ms * int d_c lambda_a * lambda_b = Ms * vol/4 * d_c lambda_a
    rows = np.repeat(tets[:, :, None, None], 4, axis=2).repeat(3, axis=3)
    cols = 3 * tets[:, None, :, None] + np.arange(3)[None, None, None, :]
    cols = np.broadcast_to(cols, (T, 4, 4, 3))
    vals = Ms * (vol / 4.0)[:, None, None, None] * np.broadcast_to(gr[:, :, None, :], (T, 4, 4, 3))
    Div = _coo(rows, cols, vals, (N, 3 * N))
    # G[3a+c, b] = - int lambda_a d_c lambda_b = -vol/4 * d_c lambda_b
    grows = 3 * tets[:, :, None, None] + np.arange(3)[None, None, None, :]
    grows = np.broadcast_to(grows, (T, 4, 4, 3))
    gcols = np.broadcast_to(tets[:, None, :, None], (T, 4, 4, 3))
    gvals = -(vol / 4.0)[:, None, None, None] * np.broadcast_to(gr[:, None, :, :], (T, 4, 4, 3))
    G = _coo(grows, gcols, gvals, (3 * N, N))
    V = np.bincount(tets.ravel(), weights=np.repeat(vol / 4.0, 4), minlength=N)
    faces, _ = exterior_facets(tets)
    bnodes = np.unique(faces.ravel())
    A_bc = dirichlet_identity(A, bnodes)
    return FEMOperators(N=N, A=A, Div=Div, G=G, V=V, bnodes=bnodes, A_bc=A_bc)


def dirichlet_identity(A, bnodes):
    """DOLFINx assemble_matrix(a, bcs): zero BC rows and columns, unit diagonal."""
    N = A.shape[0]
    keep = np.ones(N)
    keep[bnodes] = 0.0
    D = sp.diags(keep)
    Abc = (D @ A @ D).tocsr()
    I = sp.diags(1.0 - keep)
    Abc = (Abc + I).tocsr()
    Abc.sum_duplicates()
    return Abc


def exchange_matrix(X, tets, V, A_ex, Ms):
    """H_ex = diag(-2A/(mu0 Ms)/V_i/1e-18) * (int grad u : grad v) m (Exchange_GPU.py:53-68)."""
    fem = fem_operators(X, tets, 1.0)
    K = sp.kron(fem.A, sp.identity(3), format="csr")
    pref = np.repeat(-2.0 * A_ex / (MU0 * Ms) / V / 1e-18, 3)
    return (sp.diags(pref) @ K).tocsr()


def _tri_mass3(area):
    """int_T lambda_a lambda_b lambda_e dA for a,b,e in 0..2 (exact)."""
    M = np.empty((area.size, 3, 3, 3))
    for a in range(3):
        for b in range(3):
            for e in range(3):
                mult = np.bincount([a, b, e], minlength=3)
                num = math.prod(math.factorial(int(k)) for k in mult)
                M[:, a, b, e] = 2.0 * area * num / math.factorial(5)
    return M


def _boundary_tris_outward(X, tets):
    faces, owner = exterior_facets(tets)
    P = X[faces]
    n = np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0])
    nn = np.linalg.norm(n, axis=1)
    n = n / nn[:, None]
    cc = X[tets[owner]].mean(axis=1)
    fc = P.mean(axis=1)
    flip = np.einsum("ij,ij->i", fc - cc, n) < 0
    n[flip] *= -1.0
    return faces, n, 0.5 * nn


_EPS = np.zeros((3, 3, 3))
_EPS[0, 1, 2] = _EPS[1, 2, 0] = _EPS[2, 0, 1] = 1.0
_EPS[0, 2, 1] = _EPS[2, 1, 0] = _EPS[1, 0, 2] = -1.0


def dmi_interfacial_matrix(X, tets, V, D, Ms, n_nodes):
    """Interfacial DMI (DMI_Interfacial_GPU.py:134-168):
    K = int (n div u - grad(n.u)) . v dx + 1/2 int ((n x nu) x u) . v ds with the
    P1 interface normal n, scaled by -2D/(mu0 Ms)/1e-9/V_i."""
    X = np.asarray(X, dtype=np.float64)
    tets = np.asarray(tets, dtype=np.int64)
    N = X.shape[0]
    n_nodes = np.broadcast_to(np.asarray(n_nodes, dtype=np.float64).reshape(-1, 3), (N, 3))
    vol, gr = element_geometry(X, tets)
    T = tets.shape[0]
    ne = n_nodes[tets]  # (T,4,3)
    mass = (np.ones((4, 4)) + np.eye(4)) / 20.0  # int lambda_a lambda_b / vol
    # int lambda_a n_c = vol * sum_e mass[a,e] n_{e,c}
    ln = vol[:, None, None] * np.einsum("ae,tec->tac", mass, ne)          # (T,4,3) [a,c]
    dn = np.einsum("tek,tec->tkc", gr, ne)                                # d_k n_c  (T,3,3)
    # entry[a,b,c,d] = int lambda_a [ d_d lam_b n_c - lam_b d_c n_d - n_d d_c lam_b ]
    t1 = np.einsum("tac,tbd->tabcd", ln, gr)
    t2 = np.einsum("ab,t,tcd->tabcd", mass, vol, dn)
    t3 = np.einsum("tad,tbc->tabcd", ln, gr)
    Ke = t1 - t2 - t3
    faces, nu, farea = _boundary_tris_outward(X, tets)
    M3 = _tri_mass3(farea)  # (F,3,3,3)
    nf = n_nodes[faces]     # (F,3,3)
    # ((n x nu) x e_d)_c = n_c? use vector identity (a x b) x e_d with a = n x nu
    # coefficient: int lam_a lam_b ((n x nu) x e_d)_c = sum_e M3[a,b,e] ((n_e x nu) x e_d)_c
    a_vec = np.cross(nf, nu[:, None, :])  # (F,3(e),3)
    # (a x e_d)_c = eps_{c i j} a_i delta_jd = eps_{c i d} a_i
    axd = np.einsum("cid,fei->fedc", _EPS, a_vec)  # (F,e,d,c)
    Kb = 0.5 * np.einsum("fabe,fedc->fabcd", M3, axd)
    rows = 3 * tets[:, :, None, None, None] + np.arange(3)[None, None, None, :, None]
    cols = 3 * tets[:, None, :, None, None] + np.arange(3)[None, None, None, None, :]
    rows = np.broadcast_to(rows, Ke.shape)
    cols = np.broadcast_to(cols, Ke.shape)
    brows = np.broadcast_to(3 * faces[:, :, None, None, None] + np.arange(3)[None, None, None, :, None], Kb.shape)
    bcols = np.broadcast_to(3 * faces[:, None, :, None, None] + np.arange(3)[None, None, None, None, :], Kb.shape)
    K = _coo(np.concatenate([rows.ravel(), brows.ravel()]),
             np.concatenate([cols.ravel(), bcols.ravel()]),
             np.concatenate([Ke.ravel(), Kb.ravel()]), (3 * N, 3 * N))
    pref = np.repeat(-2.0 * D / (MU0 * Ms) / 1e-9 / V, 3)
    return (sp.diags(pref) @ K).tocsr()


def dmi_bulk_matrix(X, tets, V, D, Ms):
    """Bulk DMI (DMI_Bulk_GPU.py:248-264): K = -int curl(u).v dx + 1/2 int (nu x u).v ds,
    scaled by 2D/(mu0 Ms)*1e9/V_i."""
    X = np.asarray(X, dtype=np.float64)
    tets = np.asarray(tets, dtype=np.int64)
    N = X.shape[0]
    vol, gr = element_geometry(X, tets)
    # curl(lam_b e_d)_c = eps_{c i d} d_i lam_b ; int lam_a = vol/4
    Ke = -np.einsum("t,cid,tbi->tbcd", vol / 4.0, _EPS, gr)  # independent of a
    Ke = np.broadcast_to(Ke[:, None], (tets.shape[0], 4, 4, 3, 3))
    faces, nu, farea = _boundary_tris_outward(X, tets)
    mass = (np.ones((3, 3)) + np.eye(3)) / 12.0
    # (nu x e_d)_c = eps_{c i d} nu_i
    nxd = np.einsum("cid,fi->fcd", _EPS, nu)
    Kb = 0.5 * np.einsum("f,ab,fcd->fabcd", farea, mass, nxd)
    rows = np.broadcast_to(3 * tets[:, :, None, None, None] + np.arange(3)[None, None, None, :, None], Ke.shape)
    cols = np.broadcast_to(3 * tets[:, None, :, None, None] + np.arange(3)[None, None, None, None, :], Ke.shape)
    brows = np.broadcast_to(3 * faces[:, :, None, None, None] + np.arange(3)[None, None, None, :, None], Kb.shape)
    bcols = np.broadcast_to(3 * faces[:, None, :, None, None] + np.arange(3)[None, None, None, None, :], Kb.shape)
    K = _coo(np.concatenate([rows.ravel(), brows.ravel()]),
             np.concatenate([cols.ravel(), bcols.ravel()]),
             np.concatenate([np.ascontiguousarray(Ke).ravel(), Kb.ravel()]), (3 * N, 3 * N))
    pref = np.repeat(2.0 * D / (MU0 * Ms) * 1e9 / V, 3)
    return (sp.diags(pref) @ K).tocsr()


def anisotropy_field(m, n_ani, Ku, Ms):
    """H = 2Ku/(mu0 Ms) (m.n) n with n normalized per node (Anisotropy_GPU.py:25-38, 128-136)."""
    n = np.asarray(n_ani, dtype=np.float64).reshape(-1, 3)
    n = n / np.maximum(np.linalg.norm(n, axis=1), 1e-300)[:, None]
    m3 = m.reshape(-1, 3)
    n = np.broadcast_to(n, m3.shape)
    pref = 2.0 * Ku / (MU0 * Ms)
    return (pref * np.einsum("ij,ij->i", m3, n)[:, None] * n).ravel()


# --------------------------------------------------------------------------
# Demag field: u1 solve -> B u1|G -> u2 solve -> H = G(u1+u2)/V
# (gpu/fields/Demag_Lindholm_GPU.py:569-670; cpu/fields/Demag/Demag_Lindholm.py:785-839)
# --------------------------------------------------------------------------
class DemagOracle:
    def __init__(self, X, tets, Ms, B=None, bd: BoundaryData | None = None):
        self.X = np.asarray(X, dtype=np.float64)
        self.tets = np.asarray(tets)
        self.Ms = float(Ms)
        self.fem = fem_operators(self.X, self.tets, self.Ms)
        self.bd = bd if bd is not None else boundary_data(self.X, self.tets)
        self.B = dense_B_rows(self.bd) if B is None else B
        N = self.fem.N
        # u1: singular Neumann system; pin node 0 for the factorization and
        # project the constant afterwards (the nullspace removal of :572-582).
        A = self.fem.A.tolil()
        A[0, :] = 0.0
        A[0, 0] = 1.0
        self._lu1 = spla.splu(A.tocsc())
        self._lu2 = spla.splu(self.fem.A_bc.tocsc())
        self.N = N

    def solve_u1(self, m):
        b = self.fem.Div @ m
        b = b - b.mean()
        bb = b.copy()
        bb[0] = 0.0
        u1 = self._lu1.solve(bb)
        return u1 - u1.mean()

    def boundary_rhs(self, u1):
        xb = u1[self.bd.bverts]
        y = self.B @ xb
        g = np.zeros(self.N)
        g[self.bd.bverts] = y
        return g

    def solve_u2(self, g):
        b = -(self.fem.A @ g)
        bn = self.fem.bnodes
        b[bn] = g[bn]
        return self._lu2.solve(b)

    def compute(self, m, return_parts=False):
        m = np.asarray(m, dtype=np.float64)
        u1 = self.solve_u1(m)
        g = self.boundary_rhs(u1)
        u2 = self.solve_u2(g)
        H = (self.fem.G @ (u1 + u2)) / np.repeat(self.fem.V, 3)
        if return_parts:
            return H, dict(u1=u1, u2=u2, g=g)
        return H

    def energy(self, m, H):
        """Energy_lumped_gpu (Demag_Lindholm_GPU.py:706-724)."""
        dens = np.einsum("ij,ij->i", m.reshape(-1, 3), H.reshape(-1, 3))
        return -0.5 * MU0 * self.Ms * float(np.sum(dens * self.fem.V)) / 1e27


# --------------------------------------------------------------------------
# LLG RHS and Jacobian-vector product (gpu/solvers/llg_module_GPU.py:100-164)
# --------------------------------------------------------------------------
def llg_rhs(m, H, gamma, alpha, Ms, do_precess=1.0):
    """r = P (p m x H + alpha m x (m x H)) + S (1-|m|^2) m with P = -gamma/(1+alpha^2),
    S = Ms gamma/(1+alpha^2)/2 (llg_module_GPU.py:100-122, 430-431)."""
    m3 = m.reshape(-1, 3)
    h3 = H.reshape(-1, 3)
    P = -gamma / (1.0 + alpha * alpha)
    S = Ms * gamma / (1.0 + alpha * alpha) * 0.5
    mxh = np.cross(m3, h3)
    mxmxh = np.cross(m3, mxh)
    n2 = np.einsum("ij,ij->i", m3, m3)
    r = P * (do_precess * mxh + alpha * mxmxh) + (S * (1.0 - n2))[:, None] * m3
    return r.ravel()


def jac_vec(m, H, v, Hv, gamma, alpha, Ms, do_precess=1.0):
    """Jv of the LLG RHS for fixed-field derivative Hv (llg_module_GPU.py:125-164)."""
    m3, h3, v3, hv3 = (a.reshape(-1, 3) for a in (m, H, v, Hv))
    S = Ms * gamma / (1.0 + alpha * alpha) * 0.5
    coef = -gamma / (1.0 + alpha * alpha)
    dot = lambda a, b: np.einsum("ij,ij->i", a, b)[:, None]
    mdHm, mdHv, vdHm, mdv, mdmm = dot(m3, h3), dot(m3, hv3), dot(v3, h3), dot(m3, v3), dot(m3, m3)
    damp = v3 * mdHm - 2.0 * h3 * mdv + m3 * (vdHm + mdHv) - hv3 * mdmm
    prec = np.cross(v3, h3) + np.cross(m3, hv3)
    out = coef * (do_precess * prec + alpha * damp) + S * (1.0 - mdmm) * v3 - 2.0 * S * m3 * mdv
    return out.ravel()


# --------------------------------------------------------------------------
# Dormand-Prince 5(4) (the PETSc TS "rk"/"5dp" scheme of llg_module_GPU.py:1724-1787)
# --------------------------------------------------------------------------
DP_A = [
    [],
    [1 / 5],
    [3 / 40, 9 / 40],
    [44 / 45, -56 / 15, 32 / 9],
    [19372 / 6561, -25360 / 2187, 64448 / 6561, -212 / 729],
    [9017 / 3168, -355 / 33, 46732 / 5247, 49 / 176, -5103 / 18656],
    [35 / 384, 0.0, 500 / 1113, 125 / 192, -2187 / 6784, 11 / 84],
]
DP_C = [0.0, 1 / 5, 3 / 10, 4 / 5, 8 / 9, 1.0, 1.0]
DP_B5 = [35 / 384, 0.0, 500 / 1113, 125 / 192, -2187 / 6784, 11 / 84, 0.0]
DP_B4 = [5179 / 57600, 0.0, 7571 / 16695, 393 / 640, -92097 / 339200, 187 / 2100, 1 / 40]


def dp5_step(f, t, y, dt, k1=None):
    """One Dormand-Prince step; returns (y5, err_vec, k7) with FSAL k7 = f(t+dt, y5)."""
    k = [k1 if k1 is not None else f(t, y)]
    for s in range(1, 7):
        ys = y + dt * sum(DP_A[s][j] * k[j] for j in range(s) if DP_A[s][j] != 0.0)
        k.append(f(t + DP_C[s] * dt, ys))
    y5 = y + dt * sum(DP_B5[j] * k[j] for j in range(7) if DP_B5[j] != 0.0)
    err = dt * sum((DP_B5[j] - DP_B4[j]) * k[j] for j in range(7))
    return y5, err, k[6]


# --------------------------------------------------------------------------
# Cluster tree, admissibility, block partition, full-pivot ACA, H-matvec
# (gpu/fields/lindholm_hmatrix_gpu/tree.py:10-118, builder.py:256-376,
#  compressors.py:35-99, storage.py:284-299, 435-499)
# --------------------------------------------------------------------------
@dataclass
class Cluster:
    start: int
    stop: int
    bmin: np.ndarray
    bmax: np.ndarray
    left: "Cluster | None" = None
    right: "Cluster | None" = None

    @property
    def size(self):
        return self.stop - self.start

    @property
    def is_leaf(self):
        return self.left is None and self.right is None

    @property
    def diameter(self):
        return float(np.linalg.norm(self.bmax - self.bmin))


def cluster_tree(points, leaf_size=64):
    """Recursive bisection (tree.py:44-99): split axis = first argmax of the
    bbox extent, stable mergesort of that coordinate, left = first n//2."""
    points = np.ascontiguousarray(points, dtype=np.float64)
    order = []

    def rec(idx):
        lp = points[idx]
        bmin, bmax = lp.min(axis=0), lp.max(axis=0)
        if idx.size <= leaf_size:
            s = len(order)
            order.extend(idx.tolist())
            return Cluster(s, len(order), bmin, bmax)
        ax = int(np.argmax(bmax - bmin))
        srt = idx[np.argsort(lp[:, ax], kind="mergesort")]
        mid = srt.size // 2
        if mid == 0 or mid == srt.size:
            mid = max(1, srt.size // 2)
        L = rec(srt[:mid])
        R = rec(srt[mid:])
        return Cluster(L.start, R.stop, bmin, bmax, L, R)

    root = rec(np.arange(points.shape[0], dtype=np.int32))
    return np.asarray(order, dtype=np.int32), root


def admissible(a: Cluster, b: Cluster, eta):
    """tree.py:101-118."""
    gap = np.maximum(np.maximum(a.bmin - b.bmax, b.bmin - a.bmax), 0.0)
    d = float(np.linalg.norm(gap))
    if d <= 0.0:
        return False
    return max(a.diameter, b.diameter) <= float(eta) * d


def fullaca_dense(A, rtol=1e-6, max_rank=None):
    """Full-pivot ACA on a dense block (compressors.py:35-99)."""
    A = np.asarray(A, dtype=np.float64)
    m, n = A.shape
    max_rank = min(m, n) if max_rank is None else min(max(int(max_rank), 0), m, n)
    normA = float(np.linalg.norm(A))
    if normA == 0.0 or max_rank == 0:
        return np.zeros((m, 0)), np.zeros((0, n)), 0.0
    R = A.copy()
    U, V = [], []
    rel = 1.0
    for _ in range(max_rank):
        i, j = np.unravel_index(int(np.argmax(np.abs(R))), R.shape)
        piv = R[i, j]
        if abs(piv) <= np.finfo(np.float64).eps * normA:
            break
        u = R[:, j].copy()
        v = R[i, :].copy() / piv
        U.append(u)
        V.append(v)
        R -= np.outer(u, v)
        rel = float(np.linalg.norm(R) / normA)
        if rel <= rtol:
            break
    if not U:
        return np.zeros((m, 0)), np.zeros((0, n)), rel
    return np.column_stack(U), np.vstack(V), rel


def _rank_from_singular_values(s, norm_sq, rtol, max_rank):
    """First rank whose captured energy leaves a relative residual <= rtol,
    else the largest usable rank (compressors.py:395-401, 466-474)."""
    usable = min(int(max_rank), int(s.size))
    captured = np.cumsum(s[:usable] ** 2)
    rels = np.sqrt(np.maximum(norm_sq - captured, 0.0)) / math.sqrt(norm_sq)
    ok = np.flatnonzero(rels <= rtol)
    rank = int(ok[0] + 1) if ok.size else usable
    rel = float(rels[rank - 1]) if rank else 1.0
    return rank, rel, bool(ok.size)


def truncated_svd(A, rtol=1e-6, max_rank=None):
    """Truncated SVD compressor, device form (compressors.py:373-414): the
    residual from ||A||_F^2 minus the captured singular energy."""
    A = np.asarray(A, dtype=np.float64)
    m, n = A.shape
    max_rank = min(m, n) if max_rank is None else min(max(int(max_rank), 0), m, n)
    norm_sq = float(np.sum(A * A))
    if norm_sq == 0.0 or max_rank == 0:
        return np.zeros((m, 0)), np.zeros((0, n)), 0.0
    U, s, Vh = np.linalg.svd(A, full_matrices=False)
    rank, rel, _ = _rank_from_singular_values(s, norm_sq, rtol, max_rank)
    return U[:, :rank] * s[:rank][None, :], Vh[:rank].copy(), rel


def randomized_svd(A, rtol=1e-6, max_rank=None, initial_rank=16, oversampling=8, power_iterations=1,
                   random_seed=12345, omega=None):
    """Adaptive randomized SVD (compressors.py:148-235, 417-492): sketch
    Y = A Omega with target + oversampling columns, power iterations
    Y <- A (A^T qr(Y)), Q = qr(Y), SVD of Q^T A; the target doubles until a
    rank meets rtol or the sketch reaches max_rank.  omega(n, k) supplies the
    test matrices (default: numpy's generator seeded with random_seed)."""
    A = np.asarray(A, dtype=np.float64)
    m, n = A.shape
    max_rank = min(m, n) if max_rank is None else min(max(int(max_rank), 0), m, n)
    initial_rank = max(1, int(initial_rank))
    oversampling = max(0, int(oversampling))
    norm_sq = float(np.sum(A * A))
    if norm_sq == 0.0 or max_rank == 0:
        return np.zeros((m, 0)), np.zeros((0, n)), 0.0
    if omega is None:
        rng = np.random.default_rng(int(random_seed))
        omega = lambda nn, kk: rng.standard_normal((nn, kk))
    target = min(initial_rank, max_rank)
    while True:
        k = min(max_rank, target + oversampling)
        Y = A @ omega(n, k)
        for _ in range(max(0, int(power_iterations))):
            Q, _ = np.linalg.qr(Y)
            Y = A @ (A.T @ Q)
        Q, _ = np.linalg.qr(Y)
        Ub, s, Vh = np.linalg.svd(Q.T @ A, full_matrices=False)
        rank, rel, ok = _rank_from_singular_values(s, norm_sq, rtol, max_rank)
        if ok or k >= max_rank:
            break
        target = min(max_rank, max(target + 1, 2 * target))
    if rank == 0:
        return np.zeros((m, 0)), np.zeros((0, n)), rel
    return (Q @ Ub[:, :rank]) * s[:rank][None, :], Vh[:rank].copy(), rel


def block_random_seed(r0, r1, c0, c1):
    """Per-block seed of the randomized compressor (builder.py:170-183)."""
    v = int(r0) * 73856093 ^ int(r1) * 19349663 ^ int(c0) * 83492791 ^ int(c1) * 2654435761
    return int(v & 0x7FFFFFFF)


@dataclass
class HMatrixOracle:
    perm: np.ndarray
    diag_jump: np.ndarray
    near: sp.csr_matrix
    far: list          # (r0, r1, c0, c1, U, V, rel)
    stats: dict

    def matvec(self, x):
        """storage.py:284-299."""
        xp = x[self.perm]
        yp = self.diag_jump[self.perm] * xp + self.near @ xp
        for r0, r1, c0, c1, U, V, _ in self.far:
            yp[r0:r1] += U @ (V @ xp[c0:c1])
        y = np.empty_like(x)
        y[self.perm] = yp
        return y


def build_hmatrix(bd: BoundaryData, eps=1e-6, eta=2.0, leaf_size=64,
                  max_temporary_block_bytes=256 * 1024 ** 2, storage_factor=0.95,
                  max_rank=None, fill=None, compressor="fullaca", workspace_fits=None, oom=None,
                  fallback_dense=True):
    """HMatrixBuilder._build_block / _subdivide (builder.py:256-376), and the
    device builder's memory rules (builder.py:520-548, 630-760) when
    ``workspace_fits(dense_bytes)`` / ``oom(m, n)`` are given: an admissible
    block whose workspace does not fit, or whose fill/compression runs out of
    memory, is subdivided (counted as workspace / OOM subdivisions), and kept
    dense when it cannot split (``fallback_dense``; else MemoryError).
    ``fill(rows, cols)`` defaults to block_B; ``compressor`` is "fullaca",
    "svd" or a callable (A, rtol, max_rank, (r0, r1, c0, c1)) -> (U, V, rel)."""
    fill = fill or (lambda r, c: block_B(bd, r, c))
    if compressor == "fullaca":
        compress = lambda A, key: fullaca_dense(A, eps, max_rank)
    elif compressor == "svd":
        compress = lambda A, key: truncated_svd(A, eps, max_rank)
    else:
        compress = lambda A, key: compressor(A, eps, max_rank, key)
    perm, root = cluster_tree(bd.Xb, leaf_size)
    nr, nc, nd = [], [], []
    far = []
    stats = dict(dense_blocks=0, low_rank_blocks=0, rejected_low_rank_blocks=0,
                 forced_subdivisions_memory=0, compression_tolerance_subdivisions=0,
                 gpu_workspace_subdivisions=0, gpu_oom_subdivisions=0, fallback_dense_blocks=0)
    margin = 1.0 + 64.0 * np.finfo(np.float64).eps

    def add_dense(r, c, A=None):
        if A is None:
            A = fill(perm[r.start:r.stop], perm[c.start:c.stop])
        i, j = np.nonzero(A)
        if i.size:
            nr.append(i + r.start)
            nc.append(j + c.start)
            nd.append(A[i, j])
        stats["dense_blocks"] += 1

    def subdivide(r, c):
        rs, cs = not r.is_leaf, not c.is_leaf
        if not rs and not cs:
            add_dense(r, c)
            return
        if rs and (not cs or r.diameter >= c.diameter):
            build(r.left, c)
            build(r.right, c)
        else:
            build(r, c.left)
            build(r, c.right)

    def split_or_fallback(r, c, key):
        if not r.is_leaf or not c.is_leaf:
            stats[key] += 1
            subdivide(r, c)
            return
        if not fallback_dense:
            raise MemoryError(f"unsplittable block {r.size}x{c.size}")
        add_dense(r, c)
        stats["fallback_dense_blocks"] += 1

    def build(r, c):
        adm = admissible(r, c, eta)
        nbytes = r.size * c.size * 8
        if workspace_fits is not None and adm and (nbytes > max_temporary_block_bytes
                                                   or not workspace_fits(nbytes)):
            stats["forced_subdivisions_memory"] += 1
            split_or_fallback(r, c, "gpu_workspace_subdivisions")
            return
        if adm and nbytes > max_temporary_block_bytes:
            stats["forced_subdivisions_memory"] += 1
            subdivide(r, c)
            return
        if adm and oom is not None and oom(r.size, c.size):
            split_or_fallback(r, c, "gpu_oom_subdivisions")
            return
        if adm:
            A = fill(perm[r.start:r.stop], perm[c.start:c.stop])
            U, V, rel = compress(A, (r.start, r.stop, c.start, c.stop))
            rank = U.shape[1]
            ok = np.isfinite(rel) and rel <= eps * margin
            if rank > 0 and ok and U.size + V.size < storage_factor * A.size:
                far.append((r.start, r.stop, c.start, c.stop, U, V, rel))
                stats["low_rank_blocks"] += 1
                return
            if not ok and (not r.is_leaf or not c.is_leaf):
                stats["compression_tolerance_subdivisions"] += 1
                subdivide(r, c)
                return
            add_dense(r, c, A)
            stats["rejected_low_rank_blocks"] += 1
            return
        if r.is_leaf and c.is_leaf:
            add_dense(r, c)
            return
        subdivide(r, c)

    build(root, root)
    Nb = bd.Nb
    if nr:
        near = sp.coo_matrix((np.concatenate(nd), (np.concatenate(nr), np.concatenate(nc))),
                             shape=(Nb, Nb)).tocsr()
        near.sum_duplicates()
    else:
        near = sp.csr_matrix((Nb, Nb))
    return HMatrixOracle(perm=perm, diag_jump=bd.diag_jump, near=near, far=far, stats=stats)


# --------------------------------------------------------------------------
# Energy with per-cell Ms and A, and its gradient from the dense FK map
# (the product's adjoint kernels are checked against this; the reference has
# no adjoint).  E_d = 1/2 mu0 1e-27 b^T L b with b = Div_Ms m; for uniform Ms this
# is the lumped energy of Demag_Lindholm_GPU.py:706-724 (G = -Div^T/Ms).
# --------------------------------------------------------------------------
def div_per_cell(X, tets, Ms_cell):
    """Div_Ms[a, 3b+c] = sum_{cells} Ms_cell * vol/4 * d_c lambda_a (N x 3N)."""
    X = np.asarray(X, dtype=np.float64)
    tets = np.asarray(tets, dtype=np.int64)
    N = X.shape[0]
    vol, gr = element_geometry(X, tets)
    T = tets.shape[0]
    rows = np.broadcast_to(tets[:, :, None, None], (T, 4, 4, 3))
    cols = np.broadcast_to(3 * tets[:, None, :, None] + np.arange(3)[None, None, None, :], (T, 4, 4, 3))
    vals = (np.asarray(Ms_cell) * vol / 4.0)[:, None, None, None] * np.broadcast_to(gr[:, :, None, :], (T, 4, 4, 3))
    return _coo(rows, cols, vals, (N, 3 * N))


def fk_linear_map(dem: "DemagOracle"):
    """Dense L: b -> phi = u1 + u2 of the FK scheme (columns by LU solves)."""
    N = dem.N
    Bm = np.eye(N)
    Bm -= Bm.mean(axis=0, keepdims=True)  # projected right-hand sides
    Bm[0, :] = 0.0
    U1 = dem._lu1.solve(Bm)
    U1 -= U1.mean(axis=0, keepdims=True)
    G = np.zeros((N, N))
    G[dem.bd.bverts, :] = dem.B @ U1[dem.bd.bverts, :]
    Rhs = -(dem.fem.A @ G)
    bn = dem.fem.bnodes
    Rhs[bn, :] = G[bn, :]
    U2 = dem._lu2.solve(Rhs)
    return U1 + U2


def energy_per_cell(X, tets, m, Ms_cell, A_cell, Ku=0.0, n_ani=None, dem=None, L=None):
    """(E_total, parts, dE/dm, dE/dMs_cell, dE/dA_cell) in J, FP64, dense L."""
    X = np.asarray(X, dtype=np.float64)
    tets = np.asarray(tets, dtype=np.int64)
    N = X.shape[0]
    vol, gr = element_geometry(X, tets)
    Ke = vol[:, None, None] * np.einsum("tak,tbk->tab", gr, gr)
    mc = m.reshape(-1, 3)[tets]                                 # (T,4,3)
    quad = np.einsum("tab,tac,tbc->t", Ke, mc, mc)
    parts = {"E_ex": 1e-9 * float(np.sum(A_cell * quad))}
    gA = 1e-9 * quad
    fm = 2e-9 * np.asarray(A_cell)[:, None, None] * np.einsum("tab,tbc->tac", Ke, mc)
    gm = np.zeros((N, 3))
    np.add.at(gm, tets.ravel(), fm.reshape(-1, 3))
    V = np.bincount(tets.ravel(), weights=np.repeat(vol / 4.0, 4), minlength=N)
    if Ku:
        n = np.asarray(n_ani, dtype=np.float64).reshape(-1, 3)
        n = n / np.linalg.norm(n, axis=1)[:, None]
        n = np.broadcast_to(n, (N, 3))
        mn = np.einsum("ij,ij->i", m.reshape(-1, 3), n)
        parts["E_an"] = -1e-27 * Ku * float(np.sum(V * mn * mn))
        gm += (-2e-27 * Ku * V * mn)[:, None] * n
    else:
        parts["E_an"] = 0.0
    gMs = np.zeros(tets.shape[0])
    if dem is not None:
        D = div_per_cell(X, tets, Ms_cell)
        b = D @ m
        if L is None:
            L = fk_linear_map(dem)
        c = 0.5 * MU0 * 1e-27
        parts["E_d"] = c * float(b @ (L @ b))
        gb = c * (L @ b + L.T @ b)
        gm += (D.T @ gb).reshape(-1, 3)
        # dE/dMs_c = gb^T (D^c m) = sum_a gb_a vol/4 grad(lambda_a) . sum_b m_b
        gMs = np.einsum("ta,tak,tk->t", gb[tets], (vol / 4.0)[:, None, None] * gr, mc.sum(axis=1))
    else:
        parts["E_d"] = 0.0
    E = parts["E_ex"] + parts["E_an"] + parts["E_d"]
    return E, parts, gm.ravel(), gMs, gA


# --------------------------------------------------------------------------
# Cubic anisotropy (gpu/fields/Cubic_Anisotropy_GPU.py:33-53, 179-190, 205-276,
# 278-380, 501-531), Zhang-Li STT (gpu/solvers/llg_stt_module_GPU.py:101-139,
# 393-435, 566-581), spin-orbit torque (gpu/solvers/llg_SOT_module_GPU.py:
# 170-205, 813-817) and the minimizer's Cayley step / projected direction
# (gpu/solvers/Minimizer_GPU.py:46-88)
# --------------------------------------------------------------------------
def cubic_axes_oracle(u1, u2):
    """c1 = u1/|u1|, c2 = u2 - (u2.c1) c1 normalized, c3 = c1 x c2 (:179-190)."""
    a = np.asarray(u1, dtype=np.float64).reshape(-1, 3)
    b = np.asarray(u2, dtype=np.float64).reshape(-1, 3)
    a = a / np.linalg.norm(a, axis=1)[:, None]
    b = b - np.einsum("ij,ij->i", b, a)[:, None] * a
    b = b / np.linalg.norm(b, axis=1)[:, None]
    c = np.cross(a, b)
    c = c / np.linalg.norm(c, axis=1)[:, None]
    return a, b, c


def _cubic_proj(m, axes):
    m3 = np.asarray(m, dtype=np.float64).reshape(-1, 3)
    c1, c2, c3 = (np.broadcast_to(c, m3.shape) for c in axes)
    dot = lambda x, y: np.einsum("ij,ij->i", x, y)
    return m3, (c1, c2, c3), (dot(m3, c1), dot(m3, c2), dot(m3, c3))


def cubic_field(m, axes, K1, Ms):
    """H = 2 K1/(mu0 Ms) (s1 c1 + s2 c2 + s3 c3), s1 = a1 (a2^2 + a3^2) cyclic,
    a_i = m.c_i (:33-53, 205-276)."""
    m3, (c1, c2, c3), (a1, a2, a3) = _cubic_proj(m, axes)
    pref = 2.0 * K1 / (MU0 * Ms)
    s1, s2, s3 = a1 * (a2 ** 2 + a3 ** 2), a2 * (a3 ** 2 + a1 ** 2), a3 * (a1 ** 2 + a2 ** 2)
    return (pref * (s1[:, None] * c1 + s2[:, None] * c2 + s3[:, None] * c3)).ravel()


def cubic_field_jvp(m, v, axes, K1, Ms):
    """dH/dm v of cubic_field (:278-380)."""
    m3, (c1, c2, c3), (a1, a2, a3) = _cubic_proj(m, axes)
    v3 = np.asarray(v, dtype=np.float64).reshape(-1, 3)
    d1, d2, d3 = (np.einsum("ij,ij->i", v3, c) for c in (c1, c2, c3))
    pref = 2.0 * K1 / (MU0 * Ms)
    s1 = d1 * (a2 ** 2 + a3 ** 2) + 2.0 * a1 * (a2 * d2 + a3 * d3)
    s2 = d2 * (a3 ** 2 + a1 ** 2) + 2.0 * a2 * (a3 * d3 + a1 * d1)
    s3 = d3 * (a1 ** 2 + a2 ** 2) + 2.0 * a3 * (a1 * d1 + a2 * d2)
    return (pref * (s1[:, None] * c1 + s2[:, None] * c2 + s3[:, None] * c3)).ravel()


def cubic_energy(m, axes, K1, V):
    """E = -K1 sum_i V_i (a1^2 a2^2 + a2^2 a3^2 + a3^2 a1^2) * 1e-27 (:501-531)."""
    _, _, (a1, a2, a3) = _cubic_proj(m, axes)
    S = a1 ** 2 * a2 ** 2 + a2 ** 2 * a3 ** 2 + a3 ** 2 * a1 ** 2
    return float(-K1 * np.sum(V * S) * 1e-27)


def zhang_li_matrix(X, tets, V, Jdir):
    """K_J / V: K_J[a, b] = int (J . grad phi_b) phi_a dx (UFL form
    inner(dot(grad(u), J), v) dx, rows scaled by 1/VolN; :393-435), J one
    vector or one per node (P1).  The product of the linear J with the linear
    phi_a is integrated with the 4-point degree-2 rule on each tet."""
    X = np.asarray(X, dtype=np.float64)
    tets = np.asarray(tets, dtype=np.int64)
    N = X.shape[0]
    vol, g = element_geometry(X, tets)
    J = np.asarray(Jdir, dtype=np.float64).reshape(-1, 3)
    Jn = np.broadcast_to(J, (N, 3)) if J.shape[0] == 1 else J
    # degree-2 rule: barycentric points (a, b, b, b) with a = 0.5854101966249685,
    # b = 0.1381966011250105, equal weights |T|/4
    qa, qb = 0.5854101966249685, 0.1381966011250105
    lam = np.full((4, 4), qb) + np.eye(4) * (qa - qb)      # lam[q, c] at point q
    Jq = np.einsum("qc,tck->tqk", lam, Jn[tets])          # J at the points
    # int phi_a (J . g_b) = sum_q |T|/4 lam[q, a] (J_q . g_b)
    elem = np.einsum("t,qa,tqk,tbk->tab", vol / 4.0, lam, Jq, g)
    rows = np.repeat(tets[:, :, None], 4, axis=2)
    cols = np.repeat(tets[:, None, :], 4, axis=1)
    K = _coo(rows, cols, elem, (N, N))
    return sp.diags(1.0 / V) @ K


def llg_rhs_torques(m, H, gamma, alpha, Ms, Z=None, stt=None, p=None, H_dl=0.0, H_fl=0.0, do_precess=1.0):
    """LLG RHS plus the Zhang-Li torque - pref (a m x Z + b m x (m x Z)) with
    stt = (J, P, beta): pref = J P muB/(e Ms (1+beta^2))/(1+alpha^2)/1e-9,
    a = beta - alpha, b = 1 + alpha beta (llg_stt_module_GPU.py:101-139,
    566-581), plus the spin-orbit torque c_fl m x p + c_dl m x (m x p),
    c_fl = -gamma (H_fl - alpha H_dl)/(1+alpha^2), c_dl = -gamma (H_dl +
    alpha H_fl)/(1+alpha^2) (llg_SOT_module_GPU.py:170-205, 813-817)."""
    r = llg_rhs(m, H, gamma, alpha, Ms, do_precess).reshape(-1, 3)
    m3 = np.asarray(m, dtype=np.float64).reshape(-1, 3)
    if stt is not None:
        J, P, beta = stt
        pref = J * P * 9.27400915e-24 / (1.6021766e-19 * Ms * (1.0 + beta ** 2)) / (1.0 + alpha ** 2) / 1e-9
        z3 = np.asarray(Z, dtype=np.float64).reshape(-1, 3)
        mz = np.cross(m3, z3)
        r -= pref * ((beta - alpha) * mz + (1.0 + alpha * beta) * np.cross(m3, mz))
    if p is not None and (H_dl or H_fl):
        p3 = np.broadcast_to(np.asarray(p, dtype=np.float64).reshape(-1, 3), m3.shape)
        p3 = p3 / np.linalg.norm(p3, axis=1)[:, None]
        den = 1.0 + alpha * alpha
        c_fl = -gamma * (H_fl - alpha * H_dl) / den
        c_dl = -gamma * (H_dl + alpha * H_fl) / den
        mp = np.cross(m3, p3)
        r += c_fl * mp + c_dl * np.cross(m3, mp)
    return r.ravel()


def projected_direction(m, h):
    """d = h - (m.h) m (Minimizer_GPU.py:75-88)."""
    m3, h3 = np.asarray(m).reshape(-1, 3), np.asarray(h).reshape(-1, 3)
    return (h3 - np.einsum("ij,ij->i", m3, h3)[:, None] * m3).ravel()


def cayley_step(m, d, step):
    """(I - C)^-1 (I + C) m, C v = c x v, c = (step/2) m x d, solved as a 3x3
    linear system per node (Minimizer_GPU.py:46-72 states the closed form)."""
    m3, d3 = np.asarray(m).reshape(-1, 3), np.asarray(d).reshape(-1, 3)
    c = 0.5 * step * np.cross(m3, d3)
    out = np.empty_like(m3)
    for i in range(m3.shape[0]):
        Cm = np.array([[0.0, -c[i, 2], c[i, 1]], [c[i, 2], 0.0, -c[i, 0]], [-c[i, 1], c[i, 0], 0.0]])
        out[i] = np.linalg.solve(np.eye(3) - Cm, (np.eye(3) + Cm) @ m3[i])
    return out.ravel()


# --------------------------------------------------------------------------
# Implicit steps: local 3x3 block of the Newton operator
# (gpu/solvers/llg_module_GPU.py:165-290, 1412-1505)
# --------------------------------------------------------------------------
def gmres_right(mult, pc, b, x, rtol, atol=0.0, max_it=200, restart=30):
    """Right-preconditioned restarted GMRES with classical Gram-Schmidt applied
    twice and Givens rotations (the KSP gmres of the implicit relax,
    llg_module_GPU.py:2112-2140, restart 30).  x is updated in place;
    returns (iterations, final true residual norm)."""
    bnorm = float(np.linalg.norm(b))
    tol = max(rtol * bnorm, atol)
    r = b - mult(x)
    beta = float(np.linalg.norm(r))
    total = 0
    while beta > tol and total < max_it:
        V = [r / beta]
        Z = []
        H = np.zeros((restart + 1, restart))
        cs, sn = np.zeros(restart), np.zeros(restart)
        g = np.zeros(restart + 1)
        g[0] = beta
        k = 0
        for j in range(restart):
            Z.append(pc(V[j]))
            w = mult(Z[j])
            h = np.zeros(j + 1)
            for _ in range(2):
                c = np.array([v @ w for v in V])
                w = w - sum(ci * v for ci, v in zip(c, V))
                h += c
            H[:j + 1, j] = h
            hn = float(np.linalg.norm(w))
            H[j + 1, j] = hn
            V.append(w / hn if hn > 0.0 else w)
            for i in range(j):
                a, c2 = H[i, j], H[i + 1, j]
                H[i, j], H[i + 1, j] = cs[i] * a + sn[i] * c2, -sn[i] * a + cs[i] * c2
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
        x += sum(yi * zi for yi, zi in zip(y, Z))
        r = b - mult(x)
        beta = float(np.linalg.norm(r))
    return total, beta


def bdf_integrate(f, newton_ops, y0, t0, t1, dt0, rtol=1e-6, atol=1e-6, snes_rtol=1e-2, snes_atol=1e-4,
                  snes_max_it=8, ksp_rtol=1e-4, ksp_max_it=200, dt_min=1e-17, dt_max=1e-10, fixed_dt=False,
                  max_steps=10 ** 7, budget_s=None, exact_final_time="matchstep", max_reject=10):
    """Variable-step BDF(1, 2) with Newton-Krylov solves: the implicit relax of
    llg_module_GPU.py:2052-2381 (PETSc TS bdf with the basic adaptor, SNES
    newtonls with a backtracking line search, KSP gmres).  Restated from the
    published method (BDF2 with variable step ratio w = h/h_prev:
    a0 = (1+2w)/(1+w), a1 = -(1+w), a2 = w^2/(1+w); Milne error estimate
    C (y - y_pred) with C = 2/11 (BDF2) or 1/2 (BDF1) against a polynomial
    predictor through the last accepted states, weighted RMS norm with
    atol + rtol max(|y_n|, |y|), step factor 0.9 err^(-1/(q+1)) clipped to
    [0.1, 3], x0.25 after a failed solve; a step that would leave less than
    1e-3 of itself before t1 is stretched to land on t1).

    f(t, y) is the right-hand side; newton_ops(y, shift) returns (mult, pc),
    the operator x -> shift x - J(y) x and its preconditioner.  fixed_dt
    accepts every step at dt (no error control).  budget_s stops after that
    many seconds of wall time.  Step decisions follow PETSc's TSAdaptChoose_Basic
    (v3.25, the adaptor of llg_module_GPU.py:2078-2084): a rejected attempt
    that follows another rejected attempt of the same step scales the safety
    by reject_safety 0.1; a step at dt_min (within sqrt(eps)) is accepted
    whatever its error; more than max_reject consecutive rejections, or a
    non-finite error, raise.  exact_final_time "matchstep" lands on t1 (the
    stretch rule above), "stepover" (the reference's ExactFinalTime) lets the
    last step pass it.  Returns (y, t, stats)."""
    import time as _time
    y = np.array(y0, dtype=np.float64)
    hist, times = [y.copy()], [float(t0)]
    t = float(t0)
    dt = min(max(float(dt0), dt_min), dt_max)
    st = dict(steps=0, rejected=0, failed_solves=0, newton=0, gmres=0, rhs_evals=0)
    start = _time.perf_counter()

    def rhs(tt, yy):
        st["rhs_evals"] += 1
        return f(tt, yy)

    prev_failed, run = False, 0
    while t < t1 * (1 - 1e-15) and st["steps"] < max_steps:
        if budget_s is not None and _time.perf_counter() - start >= budget_s:
            break
        rem = t1 - t
        if exact_final_time == "matchstep" and (dt >= rem or rem - dt < 1e-3 * dt):
            dt = rem                               # land on t1: no sliver step before it
        q = 2 if len(hist) >= 2 else 1
        if q == 2:
            w = dt / (times[0] - times[1])
            a0, a1, a2 = (1 + 2 * w) / (1 + w), -(1 + w), w * w / (1 + w)
            shift = a0 / dt
            c_hist = (-a1 / dt) * hist[0] + (-a2 / dt) * hist[1]
        else:
            shift = 1.0 / dt
            c_hist = hist[0] / dt
        tn = times[:3]
        tq = t + dt
        lag = [math.prod((tq - tn[j]) / (tn[i] - tn[j]) for j in range(len(tn)) if j != i)
               for i in range(len(tn))]
        ypred = hist[0].copy() if len(tn) == 1 else sum(c * h for c, h in zip(lag, hist))
        yv = ypred.copy()
        F = shift * yv - rhs(tq, yv) - c_hist
        Fn = float(np.linalg.norm(F))
        tol = max(snes_atol, snes_rtol * Fn)
        ok = False
        for it in range(snes_max_it):
            if Fn <= tol and it > 0:
                ok = True
                break
            st["newton"] += 1
            mult, pc = newton_ops(yv, shift)
            d = np.zeros_like(yv)
            its, _ = gmres_right(mult, pc, -F, d, ksp_rtol, 0.0, ksp_max_it)
            st["gmres"] += its
            lam = 1.0
            for _ in range(5):
                yt = yv + lam * d
                Ft = shift * yt - rhs(tq, yt) - c_hist
                Ftn = float(np.linalg.norm(Ft))
                if Ftn <= (1.0 - 1e-4 * lam) * Fn or lam < 0.1:
                    break
                lam *= 0.5
            yv, F, Fn = yt, Ft, Ftn
            if Fn <= tol:
                ok = True
                break
        if not ok:
            st["failed_solves"] += 1
            dt = max(dt * 0.25, dt_min)
            if dt <= dt_min * (1 + 1e-12) and st["failed_solves"] > 50:
                raise RuntimeError("BDF: nonlinear solves keep failing at dt_min")
            continue
        if fixed_dt or len(tn) < q + 1:
            err = 0.0
        else:
            coef = 2.0 / 11.0 if q == 2 else 0.5
            sc = atol + rtol * np.maximum(np.abs(hist[0]), np.abs(yv))
            err = math.sqrt(float(np.mean((coef * (yv - ypred) / sc) ** 2)))
        if not math.isfinite(err):
            raise FloatingPointError("BDF: non-finite error norm")
        at_min = dt < (1.0 + math.sqrt(np.finfo(np.float64).eps)) * dt_min
        if err <= 1.0 or at_min:
            hist = [yv.copy()] + hist[:2]
            times = [tq] + times[:2]
            t = tq
            st["steps"] += 1
            prev_failed, run = False, 0
            if not fixed_dt:
                fac = 0.9 * err ** (-1.0 / (q + 1)) if err > 0 else 3.0
                dt = min(max(dt * min(max(fac, 0.1), 3.0), dt_min), dt_max)
        else:
            st["rejected"] += 1
            run = run + 1 if prev_failed else 1
            if run > max_reject:
                raise RuntimeError("BDF: too many consecutive rejections")
            safety = 0.9 * (0.1 if prev_failed else 1.0)
            prev_failed = True
            fac = safety * err ** (-1.0 / (q + 1))
            dt = min(max(dt * min(max(fac, 0.1), 3.0), dt_min), dt_max)
    return hist[0], t, st


def llg_local_jacobian(m_i, h_i, kappa, gamma, alpha, Ms, do_precess=1.0, include_stab=True):
    """d r_i / d m_i of the LLG RHS r = -c1 m x H - c2 m x (m x H) + S (1-|m|^2) m
    at one node, with the field's self-derivative dH/dm = kappa I.  Written
    from the product rule: d(m x H) = dm x H + kappa m x dm;
    d(m (m.H) - H |m|^2) = dm (m.H) + m (dm.H + kappa m.dm) - kappa dm |m|^2
    - 2 H (m.dm)."""
    m = np.asarray(m_i, dtype=np.float64)
    h = np.asarray(h_i, dtype=np.float64)
    den = 1.0 + alpha * alpha
    c1, c2 = gamma * do_precess / den, gamma * alpha / den
    S = Ms * gamma / den * 0.5
    J = np.zeros((3, 3))
    for c in range(3):
        e = np.zeros(3)
        e[c] = 1.0
        dprec = np.cross(e, h) + kappa * np.cross(m, e)
        ddamp = e * (m @ h) + m * (e @ h + kappa * (m @ e)) - kappa * e * (m @ m) - 2.0 * h * (m @ e)
        col = -c1 * dprec - c2 * ddamp
        if include_stab:
            col = col + S * ((1.0 - m @ m) * e - 2.0 * m * (m @ e))
        J[:, c] = col
    return J


def llg_local_jacobians(m, h, kappa, gamma, alpha, Ms, do_precess=1.0, include_stab=True):
    """llg_local_jacobian for every node at once: m, h (3N), kappa (N) ->
    (N, 3, 3), the same product-rule columns written with einsum."""
    m3 = np.asarray(m, dtype=np.float64).reshape(-1, 3)
    h3 = np.asarray(h, dtype=np.float64).reshape(-1, 3)
    kap = np.broadcast_to(np.asarray(kappa, dtype=np.float64), (m3.shape[0],))
    den = 1.0 + alpha * alpha
    c1, c2 = gamma * do_precess / den, gamma * alpha / den
    S = Ms * gamma / den * 0.5
    E = np.eye(3)
    mh = np.einsum("ni,ni->n", m3, h3)
    mm = np.einsum("ni,ni->n", m3, m3)
    # column c: e_c x h + kappa m x e_c
    dprec = np.cross(E[None, :, :], h3[:, None, :]) + kap[:, None, None] * np.cross(m3[:, None, :], E[None, :, :])
    # column c: e_c (m.h) + m (h_c + kappa m_c) - kappa e_c |m|^2 - 2 h m_c
    ddamp = (E[None] * mh[:, None, None] + m3[:, None, :] * (h3 + kap[:, None] * m3)[:, :, None]
             - kap[:, None, None] * E[None] * mm[:, None, None] - 2.0 * h3[:, None, :] * m3[:, :, None])
    cols = -c1 * dprec - c2 * ddamp                           # (N, c, i): column c, component i
    if include_stab:
        cols = cols + S * ((1.0 - mm)[:, None, None] * E[None] - 2.0 * m3[:, None, :] * m3[:, :, None])
    return np.transpose(cols, (0, 2, 1))


def llg_bdf_relax(X, tets, m0, t1, dt0, Ms, A_ex, alpha, Ku=0.0, n_ani=(0.0, 0.0, 1.0), gamma=2.211e5, B=None,
                  dem=None, **bdf_kw):
    """The implicit relax of an exchange + uniaxial + FEM/BEM demag magnet
    (llg_module_GPU.py:2052-2381) on the oracle's operators: RHS with the
    demag field from DemagOracle (LU solves), Newton operator shift I - J
    with J the JVP without demag (llg_module_GPU.py:1293-1323) at the field
    without demag, block preconditioner inv(shift I - J_i) with J_i the exact
    local derivative at kappa_i = mean of the linear operator's diagonal.
    Returns (m, t, stats)."""
    X = np.asarray(X, dtype=np.float64)
    dem = dem if dem is not None else DemagOracle(X, tets, Ms, B=B)
    fem = dem.fem
    Kex = exchange_matrix(X, tets, fem.V, A_ex, Ms).tocsr()
    nax = np.asarray(n_ani, dtype=np.float64)
    pref = 2.0 * Ku / (MU0 * Ms)
    kappa = (Kex.diagonal().reshape(-1, 3) + pref * nax[None, :] ** 2).mean(axis=1)

    def h_lin(v):
        out = Kex @ v
        if Ku:
            out = out + anisotropy_field(v, nax, Ku, Ms)
        return out

    def f(t, y):
        return llg_rhs(y, h_lin(y) + dem.compute(y), gamma, alpha, Ms)

    def newton_ops(y, shift):
        Hm = h_lin(y)
        J = llg_local_jacobians(y, Hm, kappa, gamma, alpha, Ms)
        inv = np.linalg.inv((shift + 1e-14) * np.eye(3)[None] - J)
        mult = lambda v: shift * v - jac_vec(y, Hm, v, h_lin(v), gamma, alpha, Ms)
        pc = lambda v: np.einsum("nij,nj->ni", inv, v.reshape(-1, 3)).ravel()
        return mult, pc

    return bdf_integrate(f, newton_ops, m0, 0.0, t1, dt0, **bdf_kw)


def bb_step_rule(sts, sty, yty, alpha_old, alpha_min, alpha_max, variant, iteration, state):
    """Barzilai-Borwein step from s = m_n - m_{n-1}, y = g_n - g_{n-1}
    (Minimizer_GPU.py:418-535): BB1 = s.s/s.y, BB2 = s.y/y.y (Barzilai and
    Borwein 1988), alternating, or the adaptive rule of Loris, Bertero, De
    Mol, Zanella and Zanni (2009): take the smallest BB2 of the last
    Malpha + 1 when BB2/BB1 <= tau (then tau *= 0.9), else BB1 (tau *= 1.1).
    state holds tau, Malpha, tau_min, tau_max and the BB2 history; returns
    (alpha, rule)."""
    clamp = lambda a: min(max(float(a), alpha_min), alpha_max)
    if sts < 1e-300 or yty < 1e-300:
        return clamp(alpha_old), "reuse_old_degenerate"
    key = str(variant).lower().replace("-", "_")
    if key in ("adaptive", "gpss", "loris", "exl", "loris_zanni"):
        if sty <= 1e-300:
            return float(alpha_max), "adaptive_alpha_max_nonpositive_sty"
        a1, a2 = clamp(sts / sty), clamp(sty / yty)
        state["hist"] = (state["hist"] + [a2])[-(state["Malpha"] + 1):]
        if a2 / max(a1, 1e-300) <= state["tau"]:
            state["tau"] = max(state["tau"] * 0.9, state["tau_min"])
            return min(state["hist"]), "adaptive_recent_min_bb2"
        state["tau"] = min(state["tau"] * 1.1, state["tau_max"])
        return a1, "adaptive_bb1"
    if sty <= 1e-300:
        return clamp(alpha_old), "nonpositive_sty"
    if key == "bb1" or (key == "alternate" and iteration % 2 == 0):
        return clamp(sts / sty), "bb1"
    if key in ("bb2", "alternate"):
        return clamp(sty / yty), "bb2"
    raise ValueError(f"unknown bb_variant {variant!r}")


def labonte_bb_minimize(energy, heff, m0, max_iter=5000, tol=1.0, alpha0=5e-11, alpha_min=1e-13, alpha_max=1e-8,
                        bb_variant="adaptive", bb_adaptive_tau0=0.5, bb_adaptive_Malpha=2, nonmonotone_window=10,
                        max_backtracking=10, shrink=0.5, energy_accept_rtol=1e-10, energy_accept_atol=1e-24,
                        energy_stagnation_rtol=1e-10, energy_stagnation_atol=1e-24, stagnation_window=30,
                        max_rejected_at_alpha_min=30, alpha_restart=5e-11, max_alpha_restarts=3,
                        restart_residual_factor=50.0, stagnation_residual_stop=50.0):
    """LaBonte steepest descent on the unit sphere with Barzilai-Borwein steps
    (LaBonteBBMinimizerGPU.minimize, Minimizer_GPU.py:537-846): direction
    d = H - (m.H) m, Cayley update m <- (I - C)^-1 (I + C) m with
    c = (step/2) m x d, nonmonotone acceptance against the largest of the
    last nonmonotone_window energies with backtracking, restarts of the step
    on energy stagnation, stop on max_i |d_i| < tol.  energy(m) -> float,
    heff(m) -> H.  Returns (m, info)."""
    def direction(mm):
        return projected_direction(mm, heff(mm))

    def max_norm3(v):
        return float(np.linalg.norm(np.asarray(v).reshape(-1, 3), axis=1).max())

    state = dict(tau=min(max(float(bb_adaptive_tau0), 1e-12), 1e12), Malpha=max(int(bb_adaptive_Malpha), 0),
                 tau_min=1e-12, tau_max=1e12, hist=[])
    rule = "init"
    m3 = np.asarray(m0, dtype=np.float64).reshape(-1, 3)
    m = (m3 / np.linalg.norm(m3, axis=1)[:, None]).ravel()
    E = energy(m)
    hist = [E]
    alpha = min(max(float(alpha0), alpha_min), alpha_max)
    rejected = restarts = 0
    stop = "max_iter"
    d = direction(m)
    it = 0
    for it in range(max_iter):
        max_g = max_norm3(d)
        if max_g < tol:
            stop = "projected_field_tol"
            break
        if len(hist) > stagnation_window:
            E_old, E_new = hist[-stagnation_window], hist[-1]
            if abs(E_new - E_old) <= energy_stagnation_atol + energy_stagnation_rtol * max(abs(E_old), abs(E_new),
                                                                                            1e-300):
                if max_g <= restart_residual_factor * tol:
                    stop = "energy_stagnation_low_residual"
                    break
                if restarts < max_alpha_restarts:
                    alpha = min(max(alpha_restart, alpha_min), alpha_max)
                    restarts += 1
                    rejected = 0
                    hist = [E]
                    d = direction(m)
                    continue
                if max_g < stagnation_residual_stop:
                    stop = "energy_stagnation_high_residual"
                    break
        m_old, d_old = m.copy(), d.copy()
        E_ref = max(hist[-nonmonotone_window:])
        accepted = False
        a_try = alpha
        for _ in range(max_backtracking + 1):
            m_trial = cayley_step(m_old, d, a_try)
            E_trial = energy(m_trial)
            if E_trial <= E_ref + energy_accept_atol + energy_accept_rtol * max(abs(E_ref), abs(E_trial), 1e-300):
                m, E = m_trial, E_trial
                hist.append(E)
                accepted = True
                rejected = 0
                break
            a_try = max(a_try * shrink, alpha_min)
        if not accepted:
            m = m_old
            rejected += 1
            alpha = alpha_min
            if rejected >= max_rejected_at_alpha_min:
                stop = "repeated_rejection_at_alpha_min"
                break
            d = direction(m)
            continue
        d = direction(m)
        s, y = m - m_old, d_old - d
        alpha, rule = bb_step_rule(float(s @ s), float(s @ y), float(y @ y), a_try, alpha_min, alpha_max, bb_variant,
                                   it, state)
        alpha = min(max(alpha, alpha_min), alpha_max)
    m3 = m.reshape(-1, 3)
    m = (m3 / np.linalg.norm(m3, axis=1)[:, None]).ravel()
    d = direction(m)
    info = dict(iterations=int(it), energy=float(energy(m)), max_projected_field=max_norm3(d), alpha=float(alpha),
                bb_last_rule=rule, alpha_restarts=int(restarts), rejected_at_alpha_min=int(rejected),
                stop_reason=stop, converged=stop in ("projected_field_tol", "energy_stagnation_low_residual"))
    return m, info
