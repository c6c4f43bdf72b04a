"""Boundary geometry of the Lindholm double-layer operator (host setup).

Restates build_boundary_data_serial (gpu/fields/Demag_Lindholm.py:177-285) with
vectorized numpy for meshes of any size:

* exterior facets = tet faces that occur once, numbered lexicographically by
  their sorted vertex triple (vertex order inside a facet: ascending id);
* outward orientation by the facet-centroid vs owning-cell-centroid test
  (:207-214), re-normalizing the flipped normal with the +1e-30 guard;
* bverts = unique(tris) (ascending mesh id), tri_lid, tri_pts, area,
  s_edge = |e_i| + 1e-30, xi_i = e_i/s_i, eta_i = n x xi_i,
  gamma_ij = xi_{i+1} . xi_j (:219-263);
* omega_sum[v] = sum over tets of |Omega| at the boundary vertex v (:265-283),
  accumulated in the same cell-major order as the reference loop.

The triangle table handed to the device kernels is the structure-of-arrays
layout of include/smx_hip.h (SMX_TRI_FIELDS doubles per triangle).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

TINY = 1e-300
_FACES = ((1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2))


def _solid_angle(x, p0, p1, p2):
    r0, r1, r2 = p0 - x, p1 - x, p2 - x
    n0 = np.sqrt(np.sum(r0 * r0, axis=-1)) + TINY
    n1 = np.sqrt(np.sum(r1 * r1, axis=-1)) + TINY
    n2 = np.sqrt(np.sum(r2 * r2, axis=-1)) + TINY
    det = np.einsum("...i,...i->...", r0, np.cross(r1, r2))
    den = (n0 * n1 * n2 + np.sum(r0 * r1, axis=-1) * n2 + np.sum(r1 * r2, axis=-1) * n0
           + np.sum(r2 * r0, axis=-1) * n1)
    return 2.0 * np.arctan2(det, den)


def exterior_facets(tets: np.ndarray, num_nodes: int):
    """(F,3) sorted vertex triples of the exterior facets and their owning cell."""
    tets = np.asarray(tets, dtype=np.int64)
    T = tets.shape[0]
    tri = np.concatenate([np.sort(tets[:, f], axis=1) for f in _FACES])
    owner = np.tile(np.arange(T, dtype=np.int64), 4)
    n = np.int64(num_nodes)
    key = (tri[:, 0] * n + tri[:, 1]) * n + tri[:, 2]
    order = np.argsort(key, kind="stable")
    ks = key[order]
    first = np.ones(ks.size, dtype=bool)
    first[1:] = ks[1:] != ks[:-1]
    last = np.ones(ks.size, dtype=bool)
    last[:-1] = ks[1:] != ks[:-1]
    single = order[first & last]  # sorted by key == lexicographic
    return tri[single], owner[single]


@dataclass
class BoundaryGeometry:
    bverts: np.ndarray     # (Nb,) int64 mesh node of each boundary index
    bpos: np.ndarray       # (N,) int32 boundary index of each node or -1
    Xb: np.ndarray
    tri_lid: np.ndarray    # (Nt,3) int32
    tri_pts: np.ndarray
    ntri: np.ndarray
    area: np.ndarray
    s_edge: np.ndarray
    eta: np.ndarray
    gamma: np.ndarray
    omega_sum: np.ndarray
    incident_policy: int = 0

    @property
    def Nb(self) -> int:
        return int(self.Xb.shape[0])

    @property
    def Nt(self) -> int:
        return int(self.tri_lid.shape[0])

    @property
    def diag_jump(self) -> np.ndarray:
        return self.omega_sum / (4.0 * math.pi) - 1.0

    def tri_table(self) -> np.ndarray:
        """(SMX_TRI_FIELDS, Nt) float64 structure-of-arrays triangle table."""
        Nt = self.Nt
        f = np.empty((34, Nt), dtype=np.float64)
        f[0:9] = self.tri_pts.reshape(Nt, 9).T
        f[9:12] = self.ntri.T
        f[12:15] = self.s_edge.T
        f[15:24] = self.eta.reshape(Nt, 9).T
        f[24:33] = self.gamma.reshape(Nt, 9).T
        f[33] = 1.0 / (8.0 * math.pi * (self.area + TINY))
        return np.ascontiguousarray(f)

    def incidence(self):
        """CSR node -> incident triangles (build_node_incident_triangle_lists,
        lindholm_hmatrix_gpu/entry_provider.py:13-38)."""
        lid = self.tri_lid.astype(np.int64).ravel()
        tri = np.repeat(np.arange(self.Nt, dtype=np.int64), 3)
        order = np.argsort(lid, kind="stable")
        ptr = np.zeros(self.Nb + 1, dtype=np.int64)
        np.add.at(ptr, lid + 1, 1)
        return np.cumsum(ptr), tri[order]

    def relevant_triangles(self, cols: np.ndarray, inc=None, scratch=None):
        """Triangles touching at least one of ``cols`` (ascending) and the local
        column of each vertex or -1 (entry_provider.py:187-223).  ``scratch``:
        an Nb-array of -1 to use as the column map (left as -1 on return)."""
        cols = np.asarray(cols, dtype=np.int64)
        ptr, tri = inc if inc is not None else self.incidence()
        counts = ptr[cols + 1] - ptr[cols]
        if counts.sum() == 0:
            return np.empty(0, np.int32), np.empty((0, 3), np.int32)
        idx = np.repeat(ptr[cols], counts) + (np.arange(counts.sum()) - np.repeat(np.cumsum(counts) - counts, counts))
        rel = np.unique(tri[idx])
        col_of = scratch if scratch is not None else np.full(self.Nb, -1, dtype=np.int64)
        col_of[cols] = np.arange(cols.size)
        out = col_of[self.tri_lid[rel]].astype(np.int32)
        if scratch is not None:
            col_of[cols] = -1
        return rel.astype(np.int32), out


def build_boundary_geometry(X: np.ndarray, tets: np.ndarray, incident_policy: str = "skip") -> BoundaryGeometry:
    X = np.asarray(X, dtype=np.float64)
    tets = np.asarray(tets, dtype=np.int64)
    N = X.shape[0]
    faces, owner = exterior_facets(tets, N)
    v0, v1, v2 = faces[:, 0].copy(), faces[:, 1].copy(), faces[:, 2].copy()
    p0, p1, p2 = X[v0], X[v1], X[v2]
    n = np.cross(p1 - p0, p2 - p0)
    nn = np.linalg.norm(n, axis=1)
    keep = nn != 0.0
    v0, v1, v2, p0, p1, p2, n, nn, owner = (a[keep] for a in (v0, v1, v2, p0, p1, p2, n, nn, owner))
    n = n / nn[:, None]
    cc = X[tets[owner]].mean(axis=1)
    fc = (p0 + p1 + p2) / 3.0
    flip = np.einsum("ij,ij->i", fc - cc, n) < 0.0
    v1f = np.where(flip, v2, v1)
    v2f = np.where(flip, v1, v2)
    if flip.any():
        q1, q2 = X[v1f[flip]], X[v2f[flip]]
        nf = np.cross(q1 - p0[flip], q2 - p0[flip])
        n[flip] = nf / (np.linalg.norm(nf, axis=1) + 1e-30)[:, None]
    tris = np.stack([v0, v1f, v2f], axis=1)
    bverts = np.unique(tris.ravel())
    Nb = bverts.size
    bpos = np.full(N, -1, dtype=np.int64)
    bpos[bverts] = np.arange(Nb)
    tri_lid = bpos[tris].astype(np.int32)
    tri_pts = X[tris]
    cr = np.cross(tri_pts[:, 1] - tri_pts[:, 0], tri_pts[:, 2] - tri_pts[:, 0])
    area = 0.5 * np.linalg.norm(cr, axis=1)
    e = (tri_pts[:, 1] - tri_pts[:, 0], tri_pts[:, 2] - tri_pts[:, 1], tri_pts[:, 0] - tri_pts[:, 2])
    s = [np.linalg.norm(ei, axis=1) + 1e-30 for ei in e]
    xi = np.stack([e[i] / s[i][:, None] for i in range(3)], axis=1)
    eta = np.cross(n[:, None, :], xi)
    gamma = np.einsum("tik,tjk->tij", xi[:, [1, 2, 0]], xi)
    # omega_sum in the reference's cell-major order
    P = X[tets]                              # (T,4,3)
    js = np.arange(4)
    x = P[:, js]                             # apex
    t1 = P[:, (js + 1) % 4]
    t2 = P[:, (js + 2) % 4]
    t3 = P[:, (js + 3) % 4]
    lid = bpos[tets]                         # (T,4)
    mask = lid >= 0
    om = np.abs(_solid_angle(x[mask], t1[mask], t2[mask], t3[mask]))
    omega_sum = np.zeros(Nb)
    np.add.at(omega_sum, lid[mask], om)
    pol = 0 if str(incident_policy).lower().startswith("skip") else 1
    return BoundaryGeometry(bverts=bverts.astype(np.int64), bpos=bpos.astype(np.int32),
                            Xb=X[bverts].copy(), tri_lid=tri_lid, tri_pts=tri_pts, ntri=n,
                            area=area, s_edge=np.stack(s, axis=1), eta=eta, gamma=gamma,
                            omega_sum=omega_sum, incident_policy=pol)
