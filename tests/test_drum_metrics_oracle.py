"""Pins of the oracle's metrics::compute restatement (metrics.rs:195-710).

Hand-placed drum states whose metrics are known in closed form, the reference's own known-answer
cases (metrics.rs:783-850: a ball on a lifter tip counts as wall-adjacent; metrics.rs:1093-1135:
empty populations give None everywhere), and the cases where a plausible mistake would show: the
wrap-around gap of the largest-gap rule, the toe/shoulder swap for a clockwise drum, the
centrifuged (no-gap) charge, the seam-crossing pool, the one-sided compression error.
"""
import math

import numpy as np
import pytest

R = 0.5


def _m(orc, drum, balls=(), fluid=(), angle=0.0, br=0.01, v=None, w=None, dens=None, h=0.01, rest=1000.0,
       mass=1.0, inertia=1.0, dye=None):
    bx = np.asarray(balls, dtype=np.float32).reshape(-1, 2)
    fx = np.asarray(fluid, dtype=np.float32).reshape(-1, 2)
    nb, nf = len(bx), len(fx)
    bv = np.zeros((nb, 2), np.float32) if v is None else np.asarray(v, np.float32)
    bw = np.zeros(nb, np.float32) if w is None else np.asarray(w, np.float32)
    de = np.full(nf, rest, np.float32) if dens is None else np.asarray(dens, np.float32)
    dy = np.zeros(nf, np.float32) if dye is None else np.asarray(dye, np.float32)
    return orc.drum_metrics(drum, angle, bx, bv, bw, br, mass, inertia, fx, dy, de, h, rest)


def _ring(radius, angles):
    a = np.asarray(angles, dtype=np.float32)
    return np.stack([np.cos(a) * radius, np.sin(a) * radius], 1).astype(np.float32)


def test_empty_populations_are_none(orc):
    # metrics.rs:1093-1135
    m = _m(orc, orc.drum_array(R, 0.0))
    for k in ("has_toe_shoulder", "has_centroid", "has_pool_extent", "has_free_surface", "has_pool_depth",
              "has_mixing_index", "has_density_errors"):
        assert m[k] == 0.0, k
    assert m["total_kinetic_energy_j"] == 0.0 and m["max_ball_overlap_fraction"] == 0.0
    assert m["max_ball_wall_overlap_fraction"] == 0.0 and m["fluid_particle_count"] == 0


def test_ball_on_a_lifter_tip_is_wall_adjacent(orc):
    # metrics.rs:783-850: five balls hugging the plain circle plus one on lifter 0's tip
    lifters = {"count": 8, "height_m": 0.05, "base_width_m": 0.05, "top_width_m": 0.02, "phase_deg": 0.0}
    drum = orc.drum_array(R, 0.5, lifters)
    five = _ring(R - 0.005, [1.0 + i * 0.05 for i in range(5)])
    tip = np.array([[0.44, 0.0]], dtype=np.float32)
    assert orc.sdf_world(drum, tip[0], 0.0)[0] < 2.5 * 0.01
    assert _m(orc, drum, five)["has_toe_shoulder"] == 0.0  # five wall balls are too few
    m = _m(orc, drum, np.vstack([five, tip]))
    assert m["has_toe_shoulder"] == 1.0
    # sorted angles [0, 1.0, ..., 1.2]: the wrap gap 0 + 2 pi - 1.2 is the largest, so the cluster
    # runs from 0 (toe, omega >= 0) to the ball at 1.2 (shoulder)
    assert m["toe_angle_rad"] == 0.0
    assert m["shoulder_angle_rad"] == pytest.approx(1.2, abs=1e-6)
    # a clockwise drum swaps them
    m2 = _m(orc, orc.drum_array(R, -0.5, lifters), np.vstack([five, tip]))
    assert (m2["toe_angle_rad"], m2["shoulder_angle_rad"]) == (m["shoulder_angle_rad"], m["toe_angle_rad"])


def test_toe_shoulder_picks_the_first_largest_inner_gap(orc):
    # wall balls at 0.5..1.0 and 4.0..4.5 rad: inner gaps 3.0 and (4.5 -> 0.5 wrap) 2.28; the cluster
    # after the 3.0 gap starts at 4.0 and ends at 1.0 (wrapping)
    ang = [0.5, 0.6, 0.7, 1.0, 4.0, 4.2, 4.5]
    m = _m(orc, orc.drum_array(R, 1.0), _ring(R - 0.01, ang))
    assert m["has_toe_shoulder"] == 1.0
    assert m["toe_angle_rad"] == pytest.approx(4.0, abs=1e-6)
    assert m["shoulder_angle_rad"] == pytest.approx(1.0, abs=1e-6)
    # a ring with no gap above 0.3 pi reads as centrifuged: None
    ring = _ring(R - 0.01, np.arange(24) * (2 * math.pi / 24))
    assert _m(orc, orc.drum_array(R, 1.0), ring)["has_toe_shoulder"] == 0.0
    # balls beyond 2.5 radii of the wall are not wall-adjacent
    assert _m(orc, orc.drum_array(R, 1.0), _ring(R - 0.03, ang))["has_toe_shoulder"] == 0.0


def test_centroid_and_kinetic_energy(orc):
    balls = [[0.1, 0.2], [-0.1, 0.2], [0.1, -0.4], [-0.1, -0.4]]
    v = [[3.0, 4.0], [0.0, 0.0], [1.0, 0.0], [0.0, -2.0]]
    w = [2.0, 0.0, -1.0, 0.5]
    m = _m(orc, orc.drum_array(R, 0.0), balls, v=v, w=w, mass=0.25, inertia=0.5)
    assert m["has_centroid"] == 1.0
    assert m["centroid_x"] == 0.0 and m["centroid_y"] == pytest.approx(-0.1, abs=1e-7)
    ke = sum(0.5 * 0.25 * (a * a + b * b) + 0.5 * 0.5 * c * c for (a, b), c in zip(v, w))
    assert m["total_kinetic_energy_j"] == pytest.approx(ke, rel=1e-6)


def test_ball_and_wall_overlaps(orc):
    r = 0.01
    drum = orc.drum_array(R, 0.0)
    m = _m(orc, drum, [[0.0, 0.0], [1.5 * r, 0.0], [0.2, 0.2]], br=r)
    assert m["max_ball_overlap_fraction"] == pytest.approx(0.5, abs=1e-5)  # (2r - 1.5r) / r
    m = _m(orc, drum, [[0.0, 0.0], [0.0, 0.0]], br=r)
    assert m["max_ball_overlap_fraction"] == 2.0  # coincident centres read 200 %
    m = _m(orc, drum, [[0.0, 0.0], [2.5 * r, 0.0]], br=r)
    assert m["max_ball_overlap_fraction"] == 0.0
    # a ball whose centre sits r/2 from the wall penetrates it by half its radius
    m = _m(orc, drum, [[0.0, -(R - 0.5 * r)], [0.1, 0.0]], br=r)
    assert m["max_ball_wall_overlap_fraction"] == pytest.approx(0.5, abs=1e-4)


def test_pool_extent_across_the_seam_and_depth(orc):
    h = 0.01
    drum = orc.drum_array(R, 0.0)
    wall = _ring(R - 0.5 * h, [6.0, 6.1, 0.1, 0.2])
    m = _m(orc, drum, fluid=wall, h=h)
    # sorted [0.1, 0.2, 6.0, 6.1]: the largest gap is 0.2 -> 6.0, so the pool runs 6.0 -> 0.2
    assert m["has_pool_extent"] == 1.0
    assert m["pool_angle_min_rad"] == pytest.approx(6.0, abs=1e-6)
    assert m["pool_angle_max_rad"] == pytest.approx(0.2, abs=1e-6)
    one = _m(orc, drum, fluid=wall[:1], h=h)
    assert one["pool_angle_min_rad"] == one["pool_angle_max_rad"] == pytest.approx(6.0, abs=1e-6)
    # depth: nearest fluid to the centre within 5 degrees of straight down
    band = [[0.0, -(R - 0.02)], [0.01, -(R - 0.03)], [0.2, -0.2]]
    d = _m(orc, drum, fluid=band, h=h)
    assert d["has_pool_depth"] == 1.0
    dist = math.hypot(0.01, R - 0.03)
    assert d["pool_depth_m"] == pytest.approx(R - dist, abs=1e-6)
    assert _m(orc, drum, fluid=[[0.2, -0.2], [0.0, 0.3]], h=h)["has_pool_depth"] == 0.0


def test_free_surface_line_of_points_on_a_line(orc):
    k, b = 0.5, -0.1
    xs = np.linspace(-0.2, 0.2, 41)
    pts = np.stack([xs, k * xs + b], 1).astype(np.float32)
    m = _m(orc, orc.drum_array(R, 0.0), fluid=pts, h=0.01)
    assert m["has_free_surface"] == 1.0
    ang = math.atan(k)
    assert m["free_surface_angle_rad"] == pytest.approx(ang, abs=1e-5)
    assert m["free_surface_offset_m"] == pytest.approx(b * math.cos(ang), abs=1e-6)
    # fewer than three interior points: None
    assert _m(orc, orc.drum_array(R, 0.0), fluid=pts[:2], h=0.01)["has_free_surface"] == 0.0


def test_density_and_compression_errors(orc):
    dens = [900.0, 1000.0, 1100.0, 1200.0]
    m = _m(orc, orc.drum_array(R, 0.0), fluid=[[0.0, 0.0]] * 4, dens=dens, rest=1000.0)
    assert m["has_density_errors"] == 1.0
    assert m["max_fluid_density_error_fraction"] == pytest.approx(0.2, abs=1e-6)
    assert m["mean_fluid_density_error_fraction"] == pytest.approx(0.1, abs=1e-6)
    # one-sided: the deficient particle (900) counts as 0 compression
    assert m["max_fluid_compression_error_fraction"] == pytest.approx(0.2, abs=1e-6)
    assert m["mean_fluid_compression_error_fraction"] == pytest.approx(0.075, abs=1e-6)
    assert _m(orc, orc.drum_array(R, 0.0), fluid=[[0.0, 0.0]], rest=0.0)["has_density_errors"] == 0.0


def test_mixing_index_is_the_standalone_one(orc):
    rng = np.random.default_rng(2)
    pts = (rng.random((500, 2)) - 0.5).astype(np.float32) * 0.6
    dye = (pts[:, 0] < 0).astype(np.float32)
    m = _m(orc, orc.drum_array(R, 0.0), fluid=pts, dye=dye)
    assert m["has_mixing_index"] == 1.0
    assert np.float32(m["mixing_index"]) == np.float32(orc.mixing_index_of(pts, dye, R))


def test_non_finite_wall_ball_panics_like_the_reference(orc):
    # sdf_world of a non-finite point reads 0 (on the wall): its NaN angle reaches the sort
    balls = np.vstack([_ring(R - 0.005, [1.0, 1.1, 1.2, 1.3, 1.4, 1.5]), [[np.nan, 0.0]]]).astype(np.float32)
    with pytest.raises(ValueError):
        _m(orc, orc.drum_array(R, 1.0), balls)
