"""Geometric cluster tree and eta-admissibility (host).

Follows lindholm_hmatrix_gpu/tree.py:10-118, whose rules fix the permutation
and the block tree bit for bit:

- a cluster with more than ``leaf_size`` points is bisected along the FIRST
  axis of largest bounding-box extent (``np.argmax`` of max - min);
- its points are ranked by a stable mergesort of that coordinate (ties keep
  the incoming order) and the lower ``n // 2`` go to the left child;
- the permutation lists the leaves left to right;
- a block (a, b) is admissible iff the bounding boxes are a positive distance
  apart and max(diam a, diam b) <= eta * dist, with diam the 2-norm of the
  bounding-box diagonal.

The construction is an explicit work stack rather than recursion: nodes are
created when popped, children are pushed right first so the left subtree is
always finished first, and each leaf's points are appended to the permutation
in that order.  Since ``leaf_size >= 1``, a split cluster holds at least two
points and both halves are non-empty.
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
    """Bisection tree over ``points`` (N x 3) and the leaf-order permutation."""
    pts = np.ascontiguousarray(points, dtype=np.float64)
    if pts.ndim != 2 or pts.shape[1] != 3:
        raise ValueError(f"cluster tree: expected an (N, 3) array of points, got shape {pts.shape}")
    n = int(pts.shape[0])
    if n < 1:
        raise ValueError("cluster tree: the point set is empty")
    leaf_size = int(leaf_size)
    if leaf_size < 1:
        raise ValueError(f"cluster tree: leaf_size must be at least 1, got {leaf_size}")

    perm = np.empty(n, dtype=np.int32)
    filled = 0
    root_holder: list[ClusterNode] = []
    # Work items: (point ids in the incoming order, parent node, is_left_child).
    # A node's interval is known only once its subtree is done, so the stop of
    # every internal node is patched on the way back (the "close" items).
    work: list[tuple] = [("open", np.arange(n, dtype=np.int32), None, False)]
    while work:
        item = work.pop()
        if item[0] == "close":
            node = item[1]
            node.stop = node.right.stop
            continue
        _, ids, parent, is_left = item
        sub = pts[ids]
        lo = sub.min(axis=0)
        hi = sub.max(axis=0)
        node = ClusterNode(filled, filled, lo, hi)
        if parent is None:
            root_holder.append(node)
        elif is_left:
            parent.left = node
        else:
            parent.right = node
        if ids.size <= leaf_size:
            perm[filled:filled + ids.size] = ids
            filled += ids.size
            node.stop = filled
            continue
        axis = int(np.argmax(hi - lo))
        ranked = ids[np.argsort(sub[:, axis], kind="mergesort")]
        cut = ranked.size // 2
        # the left subtree must be completed before the right one starts, and
        # this node closes after both
        work.append(("close", node))
        work.append(("open", ranked[cut:], node, False))
        work.append(("open", ranked[:cut], node, True))

    if filled != n or np.bincount(perm, minlength=n).max() != 1:
        raise RuntimeError("cluster tree: the leaf order is not a permutation of the points")
    return ClusterTree(points=pts, perm=perm, root=root_holder[0], leaf_size=leaf_size)


def bbox_distance(a: ClusterNode, b: ClusterNode) -> float:
    """Euclidean gap between the two axis-aligned bounding boxes (0 if they touch)."""
    gap = np.maximum(np.maximum(a.bbox_min - b.bbox_max, b.bbox_min - a.bbox_max), 0.0)
    return float(np.linalg.norm(gap))


def admissible(a: ClusterNode, b: ClusterNode, eta: float) -> bool:
    d = bbox_distance(a, b)
    if d <= 0.0:
        return False
    return max(a.diameter, b.diameter) <= float(eta) * d
