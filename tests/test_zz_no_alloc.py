"""mg_grid_build_noalloc (the grid build that keeps its scratch in a fixed device array and
allocates nothing) against mg_grid_build on the same points: layout, cell_start, indices and
Sparse keys bit-exact, on calls in a row (the later ones smaller), through the C ABI."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gm():
    from milldynamics3_amd import grid
    L = grid._lib()
    L.mg_grid_build_noalloc.argtypes = [C.c_void_p, C.c_uint32, C.c_float, C.POINTER(C.c_int64), C.c_void_p,
                                        C.c_uint64, C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64, C.c_void_p,
                                        C.c_char_p, C.c_size_t]
    return grid


def _noalloc(gm, pts, cell, cs_cap=None):
    L = gm._lib()
    n = len(pts)
    dpts = gm.DeviceArray.from_numpy(pts)
    cs_cap = 2 * n + 513 if cs_cap is None else cs_cap
    idx_cap, keys_cap = max(n, 1), 2 * max(n, 1)
    dcs, didx, dk = gm.DeviceArray(4 * cs_cap), gm.DeviceArray(4 * idx_cap), gm.DeviceArray(4 * keys_cap)
    meta = np.zeros(8, dtype=np.int64)
    err = C.create_string_buffer(256)
    rc = L.mg_grid_build_noalloc(dpts.ptr, n, cell, meta.ctypes.data_as(C.POINTER(C.c_int64)), dcs.ptr, cs_cap,
                                 didx.ptr, idx_cap, dk.ptr, keys_cap, None, err, 256)
    if rc != 0:
        return rc, err.value.decode(), None
    ncs, m = int(meta[5]), int(meta[6])
    nk = 0 if meta[0] else ncs - 1
    return rc, meta, {"cell_start": dcs.to_numpy(np.uint32, ncs), "indices": didx.to_numpy(np.uint32, m),
                      "keys": dk.to_numpy(np.int32, 2 * nk).reshape(-1, 2)}


def _handle(gm, pts, cell):
    g = gm.UniformGrid(max(len(pts), 1))
    lay = g.build(pts, cell).layout()
    g.close()
    return lay


def test_noalloc_matches_handle_build_on_calls_in_a_row(gm):
    rng = np.random.default_rng(21)
    big = (rng.random((300000, 2)) * 0.4).astype(np.float32)           # dense, 3 radix passes
    small = (rng.random((5000, 2)) * 0.03).astype(np.float32)           # sparse (outlier), non-finite
    small[17] = (40.0, -25.0)
    small[::9, 1] = np.nan
    one_cell = np.array([[0.2, 0.2], [0.2001, 0.2], [0.2, 0.2002]], dtype=np.float32)
    for pts, cell, label in ((big, 1.0e-3, "dense 300000"), (small, 1.0e-3, "sparse 5000"),
                             (one_cell, 1.0, "one cell"), (np.zeros((0, 2), np.float32), 1.0, "empty")):
        print(f"noalloc grid {label}", flush=True)
        rc, meta, out = _noalloc(gm, pts, cell)
        assert rc == 0, meta
        ref = _handle(gm, pts, cell)
        assert bool(meta[0]) == ref["dense"] and int(meta[6]) == ref["n_indices"], label
        if ref["dense"]:
            assert (int(meta[1]), int(meta[2]), int(meta[3]), int(meta[4])) == \
                (ref["min_cx"], ref["min_cy"], ref["width"], ref["height"])
        assert np.array_equal(out["cell_start"], ref["cell_start"]), label
        assert np.array_equal(out["indices"], ref["indices"]), label
        assert np.array_equal(out["keys"], ref["keys"]), label


def test_noalloc_rejects_oversize_and_short_outputs(gm):
    pts = (np.random.default_rng(1).random((1000, 2)) * 0.05).astype(np.float32)
    rc, msg, _ = _noalloc(gm, pts, 1.0e-3, cs_cap=10)
    assert rc == -3 and "too small" in msg
    L = gm._lib()
    dpts = gm.DeviceArray(64)
    meta = np.zeros(8, dtype=np.int64)
    err = C.create_string_buffer(256)
    rc = L.mg_grid_build_noalloc(dpts.ptr, (1 << 20) + 1, 1.0, meta.ctypes.data_as(C.POINTER(C.c_int64)), dpts.ptr,
                                 16, dpts.ptr, 16, dpts.ptr, 16, None, err, 256)
    assert rc == -4 and b"at most 1048576" in err.value
    # a good call after the rejected ones still matches
    rc, meta, out = _noalloc(gm, pts, 1.0e-3)
    assert rc == 0
    assert np.array_equal(out["indices"], _handle(gm, pts, 1.0e-3)["indices"])
