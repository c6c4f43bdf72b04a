"""Writes tests/golden/cube40_oracle.npz: FP64 oracle outputs on the 40 nm cube
(10^3 hexes), seeded inputs.  The reference ships no golden vectors for this
path and cannot run here (its DOLFINx/PETSc/numba/CuPy stack is absent), so
these fixtures are regression vectors of the oracle, which is itself pinned by
tests/test_oracle_pins.py.  Run:  python tests/golden/make_golden.py
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT / "oracle"))
sys.path.insert(0, str(ROOT / "fenicsx-micromagnetics_amd"))
import smx_oracle as O  # noqa: E402
from smx import mesh as M  # noqa: E402


def main():
    m = M.cube_c1()
    bd = O.boundary_data(m.X, m.tets)
    B = O.dense_B_rows(bd)
    rng = np.random.default_rng(20240601)
    g = rng.standard_normal((m.num_nodes, 3))
    mv = (g / np.linalg.norm(g, axis=1)[:, None]).ravel()
    dem = O.DemagOracle(m.X, m.tets, 8e5, B=B, bd=bd)
    H, parts = dem.compute(mv, return_parts=True)
    rows = np.arange(0, bd.Nb, 37)
    perm, _ = O.cluster_tree(bd.Xb, leaf_size=32)
    fem = O.fem_operators(m.X, m.tets, 1.0)
    Kex = O.exchange_matrix(m.X, m.tets, fem.V, 1.3e-11, 8e5)
    rhs = O.llg_rhs(mv, Kex @ mv + H, 2.211e5, 0.5, 8e5)
    out = Path(__file__).with_name("cube40_oracle.npz")
    np.savez_compressed(out, m=mv, H_demag=H, u1=parts["u1"], u2=parts["u2"], B_rows=B[rows], rows=rows,
                        omega_sum=bd.omega_sum, tri_lid=bd.tri_lid, perm_leaf32=perm, rhs_exchange_demag=rhs)
    print("wrote", out, out.stat().st_size, "bytes")


if __name__ == "__main__":
    main()
