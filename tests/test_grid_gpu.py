"""GPU parity of the reference-layout uniform grid (grid.rs) and neighbour lists (pbf.rs:262-288).

Bar: bit-exact against the oracle's restatement (oracle/mg_oracle.cpp Grid, pinned in
tests/test_oracle_pins.py and tests/test_grid_oracle.py): layout (Dense/Sparse, bounding box),
cell_start, indices, Sparse keys, candidate pairs in emission order, neighbour lists entry for
entry, near-query candidates. Point sets: the drum's own initial states (the realtime preset and the
2000-disc drum at resolution 100, fluid and balls), a 1,048,576-point jittered lattice, sparse
bounding boxes, non-finite points, keys at the i32 saturation boundary, one-cell and empty sets.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REALTIME = {"simulation": {"max_balls": 300, "resolution": 25}}
C2 = {"mill": {"diameter_m": 0.1397, "speed_mode": "percent_critical", "speed_value": 75.0},
      "media": {"density_kg_m3": 7800.0}, "simulation": {"resolution": 100, "max_balls": 2000}}


@pytest.fixture(scope="module")
def grid_mod():
    from milldynamics3_amd import grid
    return grid


def _check_layout(g, orc, pts, cell, label):
    lay = g.layout()
    ref = orc.grid_build(pts, cell)
    assert lay["dense"] == ref["dense"], label
    if ref["dense"]:
        assert (lay["min_cx"], lay["min_cy"], lay["width"], lay["height"]) == \
            (ref["min_cx"], ref["min_cy"], ref["width"], ref["height"]), label
    assert np.array_equal(lay["cell_start"], ref["cell_start"]), f"{label}: cell_start differs"
    assert np.array_equal(lay["indices"], ref["indices"]), f"{label}: indices differ"
    if not ref["dense"]:
        assert np.array_equal(lay["keys"], orc.grid_keys(pts, cell)), f"{label}: sparse keys differ"
    return lay


def _check_all(grid_mod, orc, pts, cell, label, h=None, queries=None):
    pts = np.ascontiguousarray(np.asarray(pts, dtype=np.float32).reshape(-1, 2))
    print(f"grid parity {label}: {len(pts)} points, cell {cell}", flush=True)
    g = grid_mod.UniformGrid(max(len(pts), 1))
    g.build(pts, cell)
    lay = _check_layout(g, orc, pts, cell, label)
    pairs = g.candidate_pairs()
    assert np.array_equal(pairs, orc.grid_pairs(pts, cell)), f"{label}: candidate pairs differ"
    if h is not None:
        off, nb = g.neighbor_lists(h)
        roff, rnb = orc.neighbor_lists(pts, h)  # the oracle builds its grid with cell h
        assert np.array_equal(off, roff), f"{label}: neighbour offsets differ"
        assert np.array_equal(nb, rnb), f"{label}: neighbour entries differ"
    if queries is not None:
        qoff, qout = g.near(queries)
        roff, rout = orc.grid_near(pts, cell, queries)
        assert np.array_equal(qoff, roff) and np.array_equal(qout, rout), f"{label}: near candidates differ"
    g.close()
    return lay


@pytest.mark.parametrize("params,label", [(REALTIME, "realtime"), (C2, "2000-disc drum")])
def test_drum_initial_states(grid_mod, orc, params, label):
    o = orc.OracleSim(params, "ref")
    st = o.state()
    fx = st["fluid_x"].astype(np.float32)
    bx = st["ball_x"].astype(np.float32)
    h, r = float(np.float32(o.h)), float(np.float32(o.ball_radius))
    lay = _check_all(grid_mod, orc, fx, h, f"{label} fluid", h=h, queries=bx)
    assert lay["dense"]
    # the overlap-push ball grid (pbf.rs:1214-1222) and a DEM broadphase cell (dem.rs:626-633)
    _check_all(grid_mod, orc, bx, max(2 * r, r + 0.05 * h), f"{label} balls", queries=fx)
    _check_all(grid_mod, orc, bx, 2 * r * 1.05, f"{label} balls (DEM cell)")
    _check_all(grid_mod, orc, bx, 2 * r, f"{label} balls (contact lists)", h=2 * r)


def test_million_point_lattice(grid_mod, orc):
    x, _ = orc.c3_seed_state({"side": 1024})
    h = float(np.float32(2.0e-3))
    lay = _check_all(grid_mod, orc, x, h, "1M lattice", h=h)
    assert lay["dense"] and lay["n_indices"] == 1024 * 1024


def test_sparse_layout_with_far_outliers(grid_mod, orc):
    rng = np.random.default_rng(7)
    pts = rng.random((20000, 2)).astype(np.float32) * np.float32(0.05)
    pts[123] = (3.0e3, -2.0e3)
    pts[9876] = (-1.0e3, 5.0e2)
    q = np.vstack([pts[:500], [[3.0e3, -2.0e3], [1e9, 1e9]]]).astype(np.float32)
    lay = _check_all(grid_mod, orc, pts, 1.0e-3, "sparse", h=1.0e-3, queries=q)
    assert not lay["dense"]


def test_non_finite_points_are_skipped(grid_mod, orc):
    rng = np.random.default_rng(11)
    pts = rng.random((5000, 2)).astype(np.float32) * np.float32(0.03)
    pts[::7, 0] = np.nan
    pts[3::11, 1] = np.inf
    pts[5::13] = -np.inf
    q = np.vstack([pts[:300], [[np.nan, 0.0]]]).astype(np.float32)
    lay = _check_all(grid_mod, orc, pts, 1.2e-3, "non-finite", h=1.2e-3, queries=q)
    assert lay["n_indices"] < len(pts)


def test_saturation_boundary_keys(grid_mod, orc):
    # keys at i32::MAX / MIN: saturating offsets fold onto the same cell, where the reference emits
    # some pairs once per folding offset; the GPU walk must reproduce that order and multiplicity
    big = np.float32(3.0e30)
    pts = np.array([[big, big], [big, big], [big, 0.0], [big, 0.001], [-big, -big], [-big, -big],
                    [0.0, big], [0.0005, 0.0005], [0.0004, 0.0006], [big, -big]], dtype=np.float32)
    q = pts.copy()
    _check_all(grid_mod, orc, pts, 1.0e-3, "saturation", h=1.0e-3, queries=q)


def test_one_cell_and_empty_sets(grid_mod, orc):
    pts = np.array([[0.1, 0.1], [0.1001, 0.1002], [0.1003, 0.1], [0.1002, 0.1004]], dtype=np.float32)
    lay = _check_all(grid_mod, orc, pts, 1.0, "one cell", h=1.0, queries=pts)
    assert lay["dense"] and lay["n_cell_start"] == 2
    lay = _check_all(grid_mod, orc, np.full((6, 2), np.nan, dtype=np.float32), 1.0, "all non-finite", h=1.0,
                     queries=pts)
    assert not lay["dense"] and lay["n_cell_start"] == 1
    g = grid_mod.UniformGrid(4)
    g.build(np.zeros((0, 2), dtype=np.float32), 1.0)
    assert g.layout()["n_cell_start"] == 1 and len(g.candidate_pairs()) == 0


def test_rebuild_on_one_handle(grid_mod, orc):
    rng = np.random.default_rng(3)
    g = grid_mod.UniformGrid(30000)
    for n, cell in ((30000, 1.0e-3), (1000, 2.0e-3), (17, 5.0e-4)):
        pts = (rng.random((n, 2)) * 0.04).astype(np.float32)
        g.build(pts, cell)
        _check_layout(g, orc, pts, cell, f"rebuild n={n}")
        off, nb = g.neighbor_lists(cell)
        roff, rnb = orc.neighbor_lists(pts, cell)
        assert np.array_equal(off, roff) and np.array_equal(nb, rnb)


def test_bad_cell_size_is_rejected(grid_mod):
    g = grid_mod.UniformGrid(8)
    for bad in (0.0, -1.0, float("nan"), float("inf")):
        with pytest.raises(ValueError):
            g.build(np.zeros((3, 2), dtype=np.float32), bad)
    with pytest.raises(ValueError):
        g.build(np.zeros((9, 2), dtype=np.float32), 1.0)
