"""P1 finite-element operators on the node graph, packed for the device.

The reference assembles these UFL forms with DOLFINx (host) and converts them
to AIJCUSPARSE: A_u1 = int grad v.grad u, Div = Ms int grad v.m, A_u2 (raw and
with Dirichlet identity rows), G = int v.(-grad phi)
(gpu/fields/Demag_Lindholm_GPU.py:188-261, 554-561), the exchange stiffness
(Exchange_GPU.py:53-68), the DMI operators (DMI_Interfacial_GPU.py:134-168,
DMI_Bulk_GPU.py:248-264) and the lumped nodal volume
(gpu/solvers/llg_module_GPU.py:469-487).

Here every operator is a value array over ONE node graph (the vertex adjacency
of the tetrahedra, diagonal included, columns ascending), so the per-node
kernels read a single index structure for exchange, DMI and the demag gradient.
Values per graph entry: A 1, Div 3, G 3, exchange 1, DMI 9 (3x3 block).
Element integrals are exact for P1 (UFL's estimated quadrature degree is exact
for these forms).  Device storage is SELL-64 (include/smx_hip.h).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .boundary import exterior_facets

MU0 = 4.0 * math.pi * 1e-7

_LEVI = np.zeros((3, 3, 3))
for _i, _j, _k in ((0, 1, 2), (1, 2, 0), (2, 0, 1)):
    _LEVI[_i, _j, _k] = 1.0
    _LEVI[_i, _k, _j] = -1.0


def tet_gradients(X: np.ndarray, tets: np.ndarray):
    """Volumes and barycentric gradients via the cofactor identity
    inv(J) rows = (c2 x c3, c3 x c1, c1 x c2) / det(J), c_k = x_k - x_0."""
    P = X[tets]
    c1, c2, c3 = P[:, 1] - P[:, 0], P[:, 2] - P[:, 0], P[:, 3] - P[:, 0]
    r1, r2, r3 = np.cross(c2, c3), np.cross(c3, c1), np.cross(c1, c2)
    det = np.einsum("ij,ij->i", c1, r1)
    g = np.empty((tets.shape[0], 4, 3))
    g[:, 1] = r1 / det[:, None]
    g[:, 2] = r2 / det[:, None]
    g[:, 3] = r3 / det[:, None]
    g[:, 0] = -(g[:, 1] + g[:, 2] + g[:, 3])
    return np.abs(det) / 6.0, g


@dataclass
class NodeGraph:
    n: int
    indptr: np.ndarray   # (n+1,) int64
    indices: np.ndarray  # (nnz,) int64
    pos: np.ndarray      # (T,4,4) graph position of element pair (a,b)

    @property
    def nnz(self) -> int:
        return int(self.indices.size)

    def scatter(self, elem_vals: np.ndarray, pos=None) -> np.ndarray:
        """Sum element contributions (T,4,4,*k) into graph values (nnz,*k)."""
        pos = self.pos if pos is None else pos
        k = elem_vals.shape[3:]
        flat_pos = pos.ravel()
        ev = elem_vals.reshape(flat_pos.size, -1)
        out = np.empty((self.nnz, ev.shape[1]))
        for q in range(ev.shape[1]):
            out[:, q] = np.bincount(flat_pos, weights=ev[:, q], minlength=self.nnz)
        return out.reshape((self.nnz,) + k)

    def pair_positions(self, rows: np.ndarray, cols: np.ndarray) -> np.ndarray:
        key = rows.astype(np.int64) * self.n + cols.astype(np.int64)
        gkey = np.repeat(np.arange(self.n, dtype=np.int64), np.diff(self.indptr)) * self.n + self.indices
        p = np.searchsorted(gkey, key)
        if np.any(gkey[np.minimum(p, gkey.size - 1)] != key):
            raise ValueError("pair not in the node graph")
        return p


def node_graph(tets: np.ndarray, n: int) -> NodeGraph:
    tets = np.asarray(tets, dtype=np.int64)
    a = np.broadcast_to(tets[:, :, None], (tets.shape[0], 4, 4))
    b = np.broadcast_to(tets[:, None, :], (tets.shape[0], 4, 4))
    key = (a * n + b).ravel()
    uniq, inv = np.unique(key, return_inverse=True)
    rows = uniq // n
    cols = uniq - rows * n
    indptr = np.zeros(n + 1, dtype=np.int64)
    np.add.at(indptr, rows + 1, 1)
    return NodeGraph(n=n, indptr=np.cumsum(indptr), indices=cols, pos=inv.reshape(tets.shape[0], 4, 4))


@dataclass
class SellMatrix:
    """Host SELL-64 matrix; `vals` has k components per entry."""
    nrows: int
    k: int
    slice_ptr: np.ndarray  # int64 (nslice+1,)
    cols: np.ndarray       # int32
    vals: np.ndarray       # float64 (entries, k)
    nnz: int               # stored (unpadded) entries

    @property
    def padded_entries(self) -> int:
        return int(self.cols.size)


def to_sell(indptr: np.ndarray, indices: np.ndarray, vals: np.ndarray, nrows: int) -> SellMatrix:
    vals = np.asarray(vals, dtype=np.float64)
    k = 1 if vals.ndim == 1 else int(np.prod(vals.shape[1:]))
    vals = vals.reshape(-1, k)
    lens = np.diff(indptr).astype(np.int64)
    nslice = (nrows + 63) // 64
    lens_p = np.zeros(nslice * 64, dtype=np.int64)
    lens_p[:nrows] = lens
    widths = lens_p.reshape(nslice, 64).max(axis=1)
    slice_ptr = np.zeros(nslice + 1, dtype=np.int64)
    slice_ptr[1:] = np.cumsum(widths * 64)
    total = int(slice_ptr[-1])
    rows_all = np.arange(nslice * 64, dtype=np.int64)
    fill_col = np.where(rows_all < nrows, rows_all, 0)
    # padding: every slot defaults to (own row, 0.0)
    cols = np.empty(total, dtype=np.int32)
    out = np.zeros((total, k))
    if total:
        # slot (s, j, lane) -> row 64 s + lane
        lane = np.tile(np.arange(64), total // 64)
        srow = np.repeat(np.arange(nslice, dtype=np.int64), widths * 64) * 64 + lane
        cols[:] = fill_col[srow]
        r = np.repeat(np.arange(nrows, dtype=np.int64), lens)
        j = np.arange(indices.size, dtype=np.int64) - np.repeat(indptr[:-1], lens)
        s = r // 64
        dest = slice_ptr[s] + j * 64 + (r - 64 * s)
        cols[dest] = indices
        out[dest] = vals
    return SellMatrix(nrows=nrows, k=k, slice_ptr=slice_ptr, cols=cols, vals=out, nnz=int(indices.size))


def compact_csr(indptr, indices, vals, rel_drop=1e-14):
    """Drop entries with |a_ij| <= rel_drop * max_j |a_ij| (scalar values)."""
    vals = np.asarray(vals)
    lens = np.diff(indptr)
    r = np.repeat(np.arange(lens.size), lens)
    rowmax = np.zeros(lens.size)
    np.maximum.at(rowmax, r, np.abs(vals))
    keep = np.abs(vals) > rel_drop * rowmax[r]
    new_ptr = np.zeros_like(indptr)
    np.add.at(new_ptr, r[keep] + 1, 1)
    return np.cumsum(new_ptr), indices[keep], vals[keep]


@dataclass
class FEMSystem:
    """Host operators on the node graph of one mesh."""
    N: int
    graph: NodeGraph
    volume: np.ndarray        # lumped nodal volume V_i (nm^3)
    stiff: np.ndarray         # A (nnz,)
    div: np.ndarray           # Div / Ms (nnz,3): int d_c lambda_a lambda_b
    grad: np.ndarray          # G (nnz,3)
    bnodes: np.ndarray        # boundary nodes
    is_boundary: np.ndarray   # (N,) bool
    elem_vol: np.ndarray
    elem_grad: np.ndarray

    def stiffness_bc(self) -> np.ndarray:
        """A with Dirichlet rows/cols replaced by identity (assemble_matrix with bcs)."""
        g = self.graph
        r = np.repeat(np.arange(self.N), np.diff(g.indptr))
        c = g.indices
        vals = self.stiff.copy()
        bmask = self.is_boundary[r] | self.is_boundary[c]
        vals[bmask] = 0.0
        vals[(r == c) & self.is_boundary[r]] = 1.0
        return vals


def assemble_fem(X: np.ndarray, tets: np.ndarray) -> FEMSystem:
    X = np.asarray(X, dtype=np.float64)
    tets = np.asarray(tets, dtype=np.int64)
    N = X.shape[0]
    vol, g = tet_gradients(X, tets)
    graph = node_graph(tets, N)
    stiff = graph.scatter((vol[:, None, None] * np.einsum("tak,tbk->tab", g, g))[..., None])[:, 0]
    q = vol / 4.0
    div = graph.scatter(q[:, None, None, None] * np.broadcast_to(g[:, :, None, :], (tets.shape[0], 4, 4, 3)))
    grad = graph.scatter(-q[:, None, None, None] * np.broadcast_to(g[:, None, :, :], (tets.shape[0], 4, 4, 3)))
    volume = np.bincount(tets.ravel(), weights=np.repeat(q, 4), minlength=N)
    faces, _ = exterior_facets(tets, N)
    bnodes = np.unique(faces.ravel())
    isb = np.zeros(N, dtype=bool)
    isb[bnodes] = True
    return FEMSystem(N=N, graph=graph, volume=volume, stiff=stiff, div=div, grad=grad,
                     bnodes=bnodes, is_boundary=isb, elem_vol=vol, elem_grad=g)


def exchange_values(fem: FEMSystem, A_ex: float, Ms: float) -> np.ndarray:
    """Exchange K_ab = -2A/(mu0 Ms)/V_a/1e-18 * stiff_ab (same for x, y, z)."""
    r = np.repeat(np.arange(fem.N), np.diff(fem.graph.indptr))
    pref = -2.0 * A_ex / (MU0 * Ms) / fem.volume / 1e-18
    return pref[r] * fem.stiff


def _boundary_faces(X, tets, N):
    faces, owner = exterior_facets(tets, N)
    P = X[faces]
    nrm = np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0])
    two_area = np.linalg.norm(nrm, axis=1)
    nu = nrm / two_area[:, None]
    out = np.einsum("ij,ij->i", P.mean(axis=1) - X[tets[owner]].mean(axis=1), nu) < 0.0
    nu[out] = -nu[out]
    return faces, nu, 0.5 * two_area


def _face_positions(graph: NodeGraph, faces):
    a = np.broadcast_to(faces[:, :, None], (faces.shape[0], 3, 3)).ravel()
    b = np.broadcast_to(faces[:, None, :], (faces.shape[0], 3, 3)).ravel()
    return graph.pair_positions(a, b).reshape(faces.shape[0], 3, 3)


def dmi_interfacial_values(X, tets, fem: FEMSystem, D: float, Ms: float, n0) -> np.ndarray:
    """Interfacial DMI blocks (nnz,3,3): int (n div u - grad(n.u)).v dx
    + 1/2 int ((n x nu) x u).v ds with P1 interface normal n, scaled by
    -2D/(mu0 Ms)/1e-9/V_a."""
    X = np.asarray(X, dtype=np.float64)
    tets = np.asarray(tets, dtype=np.int64)
    N = fem.N
    nn = np.broadcast_to(np.asarray(n0, dtype=np.float64).reshape(-1, 3), (N, 3))
    vol, g = fem.elem_vol, fem.elem_grad
    ne = nn[tets]                                       # (T,4,3) nodal normals
    M = (np.ones((4, 4)) + np.eye(4)) / 20.0            # P1 mass / vol
    Ln = vol[:, None, None] * np.einsum("ae,tec->tac", M, ne)   # int lam_a n_c
    Gn = np.einsum("tek,tec->tkc", g, ne)               # d_k n_c (const per tet)
    blk = (np.einsum("tac,tbd->tabcd", Ln, g)
           - np.einsum("ab,t,tcd->tabcd", M, vol, Gn)
           - np.einsum("tad,tbc->tabcd", Ln, g))
    vals = fem.graph.scatter(blk)
    faces, nu, farea = _boundary_faces(X, tets, N)
    if faces.size:
        # int_T lam_a lam_b lam_e dA = 2|T| p!q!r!/5!, p,q,r = multiplicities
        M3 = np.empty((3, 3, 3))
        for i in range(3):
            for j in range(3):
                for k in range(3):
                    cnt = np.bincount([i, j, k], minlength=3)
                    M3[i, j, k] = 2.0 * math.prod(math.factorial(int(x)) for x in cnt) / 120.0
        a = np.cross(nn[faces], nu[:, None, :])         # (F,3,3) n_e x nu
        axd = np.einsum("cid,fei->fedc", _LEVI, a)     # ((n_e x nu) x e_d)_c
        fb = 0.5 * np.einsum("f,abe,fedc->fabcd", farea, M3, axd)
        pos = _face_positions(fem.graph, faces)
        flat = pos.ravel()
        fbv = fb.reshape(flat.size, 9)
        add = np.empty((fem.graph.nnz, 9))
        for q in range(9):
            add[:, q] = np.bincount(flat, weights=fbv[:, q], minlength=fem.graph.nnz)
        vals = vals + add.reshape(-1, 3, 3)
    r = np.repeat(np.arange(N), np.diff(fem.graph.indptr))
    pref = -2.0 * D / (MU0 * Ms) / 1e-9 / fem.volume
    return pref[r][:, None, None] * vals


def dmi_bulk_values(X, tets, fem: FEMSystem, D: float, Ms: float) -> np.ndarray:
    """Bulk DMI blocks: -int curl(u).v dx + 1/2 int (nu x u).v ds, scaled by
    2D/(mu0 Ms)*1e9/V_a."""
    X = np.asarray(X, dtype=np.float64)
    tets = np.asarray(tets, dtype=np.int64)
    N = fem.N
    vol, g = fem.elem_vol, fem.elem_grad
    cb = -np.einsum("t,cid,tbi->tbcd", vol / 4.0, _LEVI, g)
    blk = np.broadcast_to(cb[:, None], (tets.shape[0], 4, 4, 3, 3))
    vals = fem.graph.scatter(blk)
    faces, nu, farea = _boundary_faces(X, tets, N)
    if faces.size:
        M2 = (np.ones((3, 3)) + np.eye(3)) / 12.0
        nxd = np.einsum("cid,fi->fcd", _LEVI, nu)
        fb = 0.5 * np.einsum("f,ab,fcd->fabcd", farea, M2, nxd)
        pos = _face_positions(fem.graph, faces).ravel()
        fbv = fb.reshape(pos.size, 9)
        add = np.empty((fem.graph.nnz, 9))
        for q in range(9):
            add[:, q] = np.bincount(pos, weights=fbv[:, q], minlength=fem.graph.nnz)
        vals = vals + add.reshape(-1, 3, 3)
    r = np.repeat(np.arange(N), np.diff(fem.graph.indptr))
    pref = 2.0 * D / (MU0 * Ms) * 1e9 / fem.volume
    return pref[r][:, None, None] * vals


def zhang_li_values(fem: FEMSystem, tets: np.ndarray, Jdir) -> np.ndarray:
    """Zhang-Li operator K_J[a, b] = int (J . grad phi_b) phi_a dx with its rows
    scaled by 1/V_a (llg_stt_module_GPU.py:393-435: the UFL form
    inner(dot(grad(u), J), v) dx, then diagonalScale(1/VolN, None)), on the node
    graph.  J is one 3-vector or one per node (P1 interpolated); for P1 J the
    integral is exact: int phi_c phi_a = |T| (1 + delta_ca) / 20, so
    K_ab = |T|/20 (sum_c J_c + J_a) . grad phi_b."""
    tets = np.asarray(tets, dtype=np.int64)
    J = np.asarray(Jdir, dtype=np.float64).reshape(-1, 3)
    if J.shape[0] == 1:
        Jt = np.broadcast_to(J[0], (tets.shape[0], 4, 3))
    elif J.shape[0] == fem.N:
        Jt = J[tets]
    else:
        raise ValueError("Jdir_vec must hold one vector or one vector per node")
    w = (fem.elem_vol / 20.0)[:, None, None] * (Jt.sum(axis=1)[:, None, :] + Jt)      # (T, a, 3)
    elem = np.einsum("tak,tbk->tab", w, fem.elem_grad)                                  # (T, a, b)
    vals = fem.graph.scatter(elem[..., None])[:, 0]
    r = np.repeat(np.arange(fem.N), np.diff(fem.graph.indptr))
    return vals / fem.volume[r]
