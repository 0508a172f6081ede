"""Pins of the oracle's grid exports used by the GPU grid tests: Sparse keys and for_each_near
(grid.rs:204-226, :371-399), against brute force over the cell keys of tiny point sets."""
import numpy as np


def _keys(pts, cell):
    with np.errstate(invalid="ignore", over="ignore"):
        q = np.floor(pts.astype(np.float32) / np.float32(cell))
    k = np.clip(q, -2147483648.0, 2147483647.0)
    return k.astype(np.int64)


def _finite(pts):
    return np.isfinite(pts).all(1)


def test_sparse_keys_are_sorted_occupied_cells(orc):
    rng = np.random.default_rng(2)
    pts = (rng.random((300, 2)) * 0.02).astype(np.float32)
    pts[5] = (50.0, -70.0)
    pts[9, 0] = np.nan
    keys = orc.grid_keys(pts, 1.0e-3)
    fin = _finite(pts)
    k = _keys(pts[fin], 1.0e-3)
    want = sorted(set(map(tuple, k.tolist())))
    assert [tuple(r) for r in keys.tolist()] == want
    assert not orc.grid_build(pts, 1.0e-3)["dense"]


def test_near_is_the_3x3_block_in_visiting_order(orc):
    rng = np.random.default_rng(4)
    cell = 1.0e-3
    pts = (rng.random((400, 2)) * 0.012).astype(np.float32)
    pts[::17, 1] = np.inf
    q = (rng.random((60, 2)) * 0.014 - 0.001).astype(np.float32)
    q[3] = (np.nan, 0.0)
    off, out = orc.grid_near(pts, cell, q)
    fin = _finite(pts)
    kp = _keys(pts, cell)
    for a in range(len(q)):
        got = out[off[a]:off[a + 1]].tolist()
        if not np.isfinite(q[a]).all():
            assert got == []
            continue
        kq = _keys(q[a:a + 1], cell)[0]
        want = []
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                sel = fin & (kp[:, 0] == kq[0] + dx) & (kp[:, 1] == kq[1] + dy)
                want += np.nonzero(sel)[0].tolist()
        assert got == want
