"""Tetrahedral mesh generators for the demag / LLG hot path.

The reference reads DOLFINx meshes from XDMF (gpu/io_utils/mesh.py:8-31).  DOLFINx
is not used here: meshes are generated as plain arrays:

* ``X``    (N, 3) float64 vertex coordinates in nm,
* ``tets`` (T, 4) int32 vertex indices per tetrahedron.

Vertex numbering of the structured box is lexicographic,
``id = i + (nx+1) * (j + (ny+1) * k)``.  Every hexahedron is split into the six
Kuhn tetrahedra that share its main diagonal, which gives a conforming mesh and
the node counts quoted for the benchmark meshes (10^3 hexes -> 1331 nodes and 602
boundary nodes; 30^3 -> 29791/5402; 200x200x4 -> 202005/83202).
"""
from __future__ import annotations

from dataclasses import dataclass
from itertools import permutations

import numpy as np


@dataclass
class TetMesh:
    X: np.ndarray      # (N, 3) float64, nm
    tets: np.ndarray   # (T, 4) int32
    name: str = "mesh"

    @property
    def num_nodes(self) -> int:
        return int(self.X.shape[0])

    @property
    def num_cells(self) -> int:
        return int(self.tets.shape[0])


def kuhn_box(nx: int, ny: int, nz: int, lx: float, ly: float, lz: float,
             origin=(0.0, 0.0, 0.0), name: str | None = None) -> TetMesh:
    """Structured box [0,lx]x[0,ly]x[0,lz] (nm) with nx*ny*nz hexes, 6 tets each."""
    nx, ny, nz = int(nx), int(ny), int(nz)
    xs = origin[0] + np.linspace(0.0, lx, nx + 1)
    ys = origin[1] + np.linspace(0.0, ly, ny + 1)
    zs = origin[2] + np.linspace(0.0, lz, nz + 1)
    zz, yy, xx = np.meshgrid(zs, ys, xs, indexing="ij")
    X = np.stack([xx.ravel(), yy.ravel(), zz.ravel()], axis=1).astype(np.float64)

    sx, sy, sz = 1, nx + 1, (nx + 1) * (ny + 1)
    step = (sx, sy, sz)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    base = (i * sx + j * sy + k * sz).ravel().astype(np.int64)

    cells = []
    for perm in permutations((0, 1, 2)):
        a, b, c = perm
        v0 = base
        v1 = v0 + step[a]
        v2 = v1 + step[b]
        v3 = v2 + step[c]
        cells.append(np.stack([v0, v1, v2, v3], axis=1))
    # hex-major ordering: the six tets of one hex are consecutive
    tets = np.stack(cells, axis=1).reshape(-1, 4).astype(np.int32)
    return TetMesh(X=X, tets=tets, name=name or f"box{nx}x{ny}x{nz}")


def cube_c1() -> TetMesh:
    """Cube 40 nm per side, 10^3 hexes (1331 nodes, 602 boundary nodes)."""
    return kuhn_box(10, 10, 10, 40.0, 40.0, 40.0, name="cube40_10")


def cube_c2() -> TetMesh:
    """Cube 150 nm per side, 30^3 hexes (29791 nodes, 5402 boundary nodes)."""
    return kuhn_box(30, 30, 30, 150.0, 150.0, 150.0, name="cube150_30")


def film_c3() -> TetMesh:
    """Film 500x500x10 nm at 2.5 nm spacing (201x201x5 = 202005 nodes)."""
    return kuhn_box(200, 200, 4, 500.0, 500.0, 10.0, name="film500_200x200x4")


def disk(diameter: float = 200.0, height: float = 20.0, h: float = 3.0,
         name: str | None = None) -> TetMesh:
    """Cylinder (disk) mesh built from a Kuhn box by radial mapping.

    A square [-1,1]^2 cross-section grid is mapped onto the unit disk with the
    elliptical (Fong) square-to-disk map, so the mesh is conforming and its
    boundary is a polygonal approximation of the cylinder.  ``h`` is the target
    mean edge in nm.
    """
    r = 0.5 * float(diameter)
    n = max(2, int(round(2.0 * r / h)))
    nzc = max(1, int(round(float(height) / h)))
    box = kuhn_box(n, n, nzc, 2.0, 2.0, float(height), origin=(-1.0, -1.0, 0.0))
    u = box.X[:, 0]
    v = box.X[:, 1]
    x = u * np.sqrt(np.maximum(0.0, 1.0 - 0.5 * v * v))
    y = v * np.sqrt(np.maximum(0.0, 1.0 - 0.5 * u * u))
    X = np.stack([r * x, r * y, box.X[:, 2]], axis=1)
    return TetMesh(X=X, tets=box.tets, name=name or f"disk{diameter:g}x{height:g}_h{h:g}")
