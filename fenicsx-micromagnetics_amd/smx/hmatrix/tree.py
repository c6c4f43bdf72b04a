"""Geometric cluster tree and eta-admissibility (host).

Restates lindholm_hmatrix_gpu/tree.py:10-118 so that the permutation and the
block tree are bit-identical to the reference's: recursive bisection where the
split axis is the FIRST argmax of the bounding-box extent, points ordered by a
stable mergesort of that coordinate and the left child taking the first n//2;
leaves hold <= leaf_size points; a block (a, b) is admissible iff the bbox
distance is > 0 and max(diam a, diam b) <= eta * dist (diam = 2-norm of the
bbox diagonal).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass(eq=False)
class ClusterNode:
    start: int
    stop: int
    bbox_min: np.ndarray
    bbox_max: np.ndarray
    left: "ClusterNode | None" = None
    right: "ClusterNode | None" = None
    _diam: float | None = field(default=None, repr=False)

    @property
    def size(self) -> int:
        return self.stop - self.start

    @property
    def is_leaf(self) -> bool:
        return self.left is None and self.right is None

    @property
    def diameter(self) -> float:
        if self._diam is None:
            self._diam = float(np.linalg.norm(self.bbox_max - self.bbox_min))
        return self._diam


@dataclass
class ClusterTree:
    points: np.ndarray
    perm: np.ndarray
    root: ClusterNode
    leaf_size: int

    def leaves(self):
        out = []
        stack = [self.root]
        while stack:
            n = stack.pop()
            if n.is_leaf:
                out.append(n)
            else:
                stack.append(n.right)
                stack.append(n.left)
        return out


def build_cluster_tree(points: np.ndarray, leaf_size: int = 64) -> ClusterTree:
    pts = np.ascontiguousarray(points, dtype=np.float64)
    if pts.ndim != 2 or pts.shape[1] != 3:
        raise ValueError("points must have shape (N, 3).")
    n = int(pts.shape[0])
    if n == 0:
        raise ValueError("At least one point is required.")
    if leaf_size < 1:
        raise ValueError("leaf_size must be positive.")
    order: list[np.ndarray] = []
    count = [0]

    def split(idx: np.ndarray) -> ClusterNode:
        sub = pts[idx]
        lo = sub.min(axis=0)
        hi = sub.max(axis=0)
        if idx.size <= leaf_size:
            s = count[0]
            order.append(idx)
            count[0] += idx.size
            return ClusterNode(s, count[0], lo, hi)
        axis = int(np.argmax(hi - lo))
        ranked = idx[np.argsort(sub[:, axis], kind="mergesort")]
        half = ranked.size // 2
        if half == 0 or half == ranked.size:
            half = max(1, ranked.size // 2)
        a = split(ranked[:half])
        b = split(ranked[half:])
        return ClusterNode(a.start, b.stop, lo, hi, a, b)

    root = split(np.arange(n, dtype=np.int32))
    perm = np.concatenate(order).astype(np.int32)
    if perm.size != n or np.unique(perm).size != n:
        raise RuntimeError("Invalid cluster-tree permutation.")
    return ClusterTree(points=pts, perm=perm, root=root, leaf_size=int(leaf_size))


def bbox_distance(a: ClusterNode, b: ClusterNode) -> float:
    gap = np.maximum(np.maximum(a.bbox_min - b.bbox_max, b.bbox_min - a.bbox_max), 0.0)
    return float(np.linalg.norm(gap))


def admissible(a: ClusterNode, b: ClusterNode, eta: float) -> bool:
    d = bbox_distance(a, b)
    if d <= 0.0:
        return False
    return max(a.diameter, b.diameter) <= float(eta) * d
