"""Host-side tables of the device compressors (CPU, no GPU calls).

* aca_tile_table: every row of every block is covered by exactly one tile,
  tiles of a block are consecutive and in row order, each at most
  ACA_TILE_ROWS rows; the workspace size of smx_aca_large follows.
* The SVD compressors' task table: offsets and shapes copied from the ACA
  task layout, usable ranks min(m, n) or the cap, workspace offsets from
  smx_svd_ws_doubles (non-overlapping, in task order), and invalid sizes
  rejected before any launch.
"""
import numpy as np
import pytest


def test_aca_tile_table_covers_every_row_once():
    from smx.hmatrix.builder import ACA_TILE_ROWS, aca_tile_table
    m = np.array([1, 63, 64, 65, 5200, 129])
    tt, r0, first = aca_tile_table(m)
    assert first[0] == 0 and first[-1] == tt.size and np.all(np.diff(first) >= 1)
    for b, mb in enumerate(m):
        tiles = np.arange(first[b], first[b + 1])
        assert np.all(tt[tiles] == b)
        rows = np.concatenate([np.arange(r, min(r + ACA_TILE_ROWS, mb)) for r in r0[tiles]])
        assert np.array_equal(rows, np.arange(mb))          # each row once, in order
        assert np.all(np.diff(r0[tiles]) == ACA_TILE_ROWS)
    with pytest.raises(ValueError):
        aca_tile_table(np.array([3, 0]))


def test_aca_large_workspace_size():
    from smx import _capi
    L = _capi.lib()
    assert L.smx_aca_large_ws_doubles(9, 738) == 9 * 8 + 3 * 738 + 2
    assert L.smx_aca_large_ws_doubles(-1, 3) == -1


def test_svd_task_table_and_workspace():
    from smx import _capi
    from smx.hmatrix.builder import ACA_TASK_DTYPE
    from smx.hmatrix.compressors import SVD_TASK_DTYPE, _task_table
    L = _capi.lib()
    shapes = [(60, 40), (5, 300), (200, 150), (1, 1)]
    t = np.zeros(len(shapes), dtype=ACA_TASK_DTYPE)
    t["m"] = [a for a, _ in shapes]
    t["n"] = [b for _, b in shapes]
    t["rcap"] = [40, 5, 150, 1]
    t["a_off"] = np.concatenate([[0], np.cumsum(t["m"].astype(np.int64) * t["n"])[:-1]])
    t["u_off"] = 7
    t["v_off"] = 11
    for widths, max_rank in ((np.zeros(4, np.int32), None), (np.array([8, 5, 24, 1], np.int32), 20)):
        st, nws = _task_table(t, np.arange(len(shapes)), max_rank, widths)
        assert st.dtype == SVD_TASK_DTYPE and st.itemsize == 64
        for f in ("a_off", "u_off", "v_off", "m", "n", "rcap"):
            assert np.array_equal(st[f], t[f])
        mn = np.minimum(t["m"], t["n"])
        assert np.array_equal(st["mr"], mn if max_rank is None else np.minimum(mn, max_rank))
        ws = [L.smx_svd_ws_doubles(int(a), int(b), int(s)) for (a, b), s in zip(shapes, widths)]
        assert np.array_equal(st["w_off"], np.concatenate([[0], np.cumsum(ws)[:-1]])) and nws == sum(ws)
    # the truncated SVD's workspace: 2m + k + k m + k^2 with k = min(m, n, 128)
    assert L.smx_svd_ws_doubles(400, 300, 0) == 2 * 400 + 128 + 128 * 400 + 128 * 128
    assert L.smx_svd_ws_doubles(60, 40, 0) == 2 * 60 + 40 + 40 * 60 + 40 * 40
    # randomized: s (m + n) + s^2, s at most 128
    assert L.smx_svd_ws_doubles(60, 40, 8) == 8 * 100 + 64
    assert L.smx_svd_ws_doubles(60, 40, 129) == -1
    with pytest.raises(ValueError):
        _task_table(t, np.arange(1), None, np.array([200], np.int32))
