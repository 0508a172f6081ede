"""Pins of the oracle's metrics::mixing_index restatement (metrics.rs:506-576) on fields whose Lacey
index is known: fully segregated (0), mixed within every cell (1), and the reference's None cases."""
import numpy as np


def _cell_centres(radius, per_side=16):
    cell = 2 * radius / per_side
    c = -radius + (np.arange(per_side) + 0.5) * cell
    gx, gy = np.meshgrid(c, c)
    return np.stack([gx.ravel(), gy.ravel()], 1)


def test_segregated_and_mixed_fields(orc):
    R = 0.0315
    cen = _cell_centres(R)
    pts = np.repeat(cen, 4, axis=0).astype(np.float32)
    left = (pts[:, 0] < 0).astype(np.float32)
    assert orc.mixing_index_of(pts, left, R) == 0.0  # every cell pure: S^2 = S0^2
    alt = np.tile([0.0, 1.0, 0.0, 1.0], len(cen)).astype(np.float32)
    assert orc.mixing_index_of(pts, alt, R) == 1.0  # every cell at the mean: S^2 = 0


def test_none_cases(orc):
    R = 0.0315
    pts = np.array([[0.001, 0.001], [0.0011, 0.0012]], dtype=np.float32)
    assert orc.mixing_index_of(pts, [0.0, 1.0], R) is None            # one occupied cell
    cen = _cell_centres(R).astype(np.float32)
    assert orc.mixing_index_of(cen, np.ones(len(cen)), R) is None      # no dye variance
    assert orc.mixing_index_of(np.zeros((0, 2)), np.zeros(0), R) is None
    assert orc.mixing_index_of(cen, np.ones(len(cen)), 0.0) is None


def test_count_weighting(orc):
    # two occupied cells: a bulk one with 40 particles at mean 0.5 and a stray one with 1 particle
    # of dye 1; the weighted variance keeps the stray cell's weight at 1/41
    R = 0.0315
    a = np.tile([[-0.02, -0.02]], (40, 1))
    b = np.array([[0.02, 0.02]])
    pts = np.vstack([a, b]).astype(np.float32)
    dye = np.concatenate([np.tile([0.0, 1.0], 20), [1.0]]).astype(np.float32)
    p = dye.mean(dtype=np.float64)
    ov = (40 * (0.5 - p) ** 2 + 1 * (1.0 - p) ** 2) / 41
    s0 = p * (1 - p)
    sr = s0 / ((40 * 40 + 1) / 41)
    assert abs(orc.mixing_index_of(pts, dye, R) - min(max((s0 - ov) / (s0 - sr), 0), 1)) < 1e-5
