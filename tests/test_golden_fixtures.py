"""Regression of the oracle and the host set-up against the committed fixtures
(tests/golden/cube40_oracle.npz, written by tests/golden/make_golden.py)."""
from pathlib import Path

import numpy as np

import smx_oracle as O
from smx import mesh as M
from smx.boundary import build_boundary_geometry
from smx.hmatrix import build_cluster_tree

G = np.load(Path(__file__).parent / "golden" / "cube40_oracle.npz", allow_pickle=False)


def test_boundary_and_tree_fixtures():
    m = M.cube_c1()
    geo = build_boundary_geometry(m.X, m.tets)
    assert np.array_equal(geo.tri_lid, G["tri_lid"])
    assert np.array_equal(geo.omega_sum, G["omega_sum"])
    assert np.array_equal(build_cluster_tree(geo.Xb, 32).perm, G["perm_leaf32"])


def test_oracle_fixtures():
    m = M.cube_c1()
    bd = O.boundary_data(m.X, m.tets)
    rows = G["rows"]
    assert np.abs(O.dense_B_rows(bd, rows) - G["B_rows"]).max() < 1e-15
    dem = O.DemagOracle(m.X, m.tets, 8e5, bd=bd)
    H = dem.compute(G["m"])
    assert np.linalg.norm(H - G["H_demag"]) / np.linalg.norm(G["H_demag"]) < 1e-12
