"""GPU parity of metrics::compute (metrics.rs:195-254) through mg_drum_metrics_batch.

Bar: bit-exact against the oracle's restatement (oracle/mg_oracle.cpp drum_metrics, pinned in
tests/test_drum_metrics_oracle.py): every field equal as an f32 (None where the oracle has None),
on the oracle's hand-placed cases, on drum states the oracle has stepped (the realtime preset, the
2000-disc drum, a lifter drum turning clockwise), on a batch mixing all of them with empty and
one-population drums, and on a drum whose wall lists run to thousands of entries (several sort
tiles). Fluid densities are synthetic (the metrics read them as given).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("toe_angle_rad", "shoulder_angle_rad", "centroid_x", "centroid_y", "pool_angle_min_rad",
        "pool_angle_max_rad", "free_surface_angle_rad", "free_surface_offset_m", "pool_depth_m", "mixing_index",
        "total_kinetic_energy_j", "max_ball_overlap_fraction", "max_ball_wall_overlap_fraction",
        "max_fluid_density_error_fraction", "mean_fluid_density_error_fraction",
        "max_fluid_compression_error_fraction", "mean_fluid_compression_error_fraction", "drum_omega_rad_s",
        "fluid_particle_count")
HAS = {"toe_angle_rad": "has_toe_shoulder", "shoulder_angle_rad": "has_toe_shoulder",
       "centroid_x": "has_centroid", "centroid_y": "has_centroid", "pool_angle_min_rad": "has_pool_extent",
       "pool_angle_max_rad": "has_pool_extent", "free_surface_angle_rad": "has_free_surface",
       "free_surface_offset_m": "has_free_surface", "pool_depth_m": "has_pool_depth",
       "mixing_index": "has_mixing_index", "max_fluid_density_error_fraction": "has_density_errors",
       "mean_fluid_density_error_fraction": "has_density_errors",
       "max_fluid_compression_error_fraction": "has_density_errors",
       "mean_fluid_compression_error_fraction": "has_density_errors"}

LIFTERS = {"count": 8, "height_m": 0.005, "base_width_m": 0.006, "top_width_m": 0.003, "phase_deg": 10.0}


@pytest.fixture(scope="module")
def mm():
    from milldynamics3_amd import metrics
    return metrics


def _oracle(orc, d):
    drum = orc.drum_array(d["radius_m"], d["omega"], d.get("lifters"))
    return orc.drum_metrics(drum, d["drum_angle"], d["ball_x"], d["ball_v"], d["ball_omega"], d["ball_radius"],
                            d["ball_mass"], d["ball_inertia"], d["fluid_x"], d["dye"], d["density"], d["h"],
                            d["rest_density"])


def _check(gpu, ref, label):
    for k in KEYS:
        flag = HAS.get(k)
        if flag is not None and ref[flag] == 0.0:
            assert gpu[k] is None, f"{label}: {k} should be None"
            continue
        g, r = np.float32(gpu[k]), np.float32(ref[k])
        assert g.view(np.uint32) == r.view(np.uint32) or (g == r == 0.0), f"{label}: {k} gpu {g!r} oracle {r!r}"


def _drum(radius, omega, angle, bx, bv, bw, br, bm, bi, fx, dye, dens, h, rest=1000.0, lifters=None):
    return {"radius_m": radius, "omega": omega, "drum_angle": angle, "lifters": lifters, "ball_radius": br,
            "ball_mass": bm, "ball_inertia": bi, "h": h, "rest_density": rest,
            "ball_x": np.asarray(bx, np.float32).reshape(-1, 2), "ball_v": np.asarray(bv, np.float32).reshape(-1, 2),
            "ball_omega": np.asarray(bw, np.float32).reshape(-1), "fluid_x": np.asarray(fx, np.float32).reshape(-1, 2),
            "dye": np.asarray(dye, np.float32).reshape(-1), "density": np.asarray(dens, np.float32).reshape(-1)}


def _stepped(orc, params, substeps, seed=0):
    o = orc.OracleSim(params, "ref")
    o.step_fixed(substeps)
    st, sc = o.state(), o.scalars()
    p = orc._merge(orc._DEFAULTS, params)
    rng = np.random.default_rng(seed)
    nf = len(st["fluid_x"])
    dens = (1000.0 * (1.0 + 0.03 * rng.standard_normal(nf))).astype(np.float32)
    li = p["lifters"] if p["lifters"]["count"] else None
    return _drum(p["mill"]["diameter_m"] / 2, o.omega, sc["drum_angle"], st["ball_x"], st["ball_v"],
                 st["ball_omega"], o.ball_radius, o.ball_mass, o.ball_inertia, st["fluid_x"], st["dye"], dens, o.h,
                 1000.0, li)


def _cases(orc):
    R = 0.5
    ring = lambda r, a: np.stack([np.cos(np.float32(a)) * r, np.sin(np.float32(a)) * r], 1).astype(np.float32)
    five = ring(R - 0.005, [1.0 + i * 0.05 for i in range(5)])
    tip = np.array([[0.44, 0.0]], np.float32)
    lifters = {"count": 8, "height_m": 0.05, "base_width_m": 0.05, "top_width_m": 0.02, "phase_deg": 0.0}
    z = lambda n: np.zeros((n, 2), np.float32)
    balls6 = np.vstack([five, tip])
    wall = ring(R - 0.005, [6.0, 6.1, 0.1, 0.2])
    line = np.stack([np.linspace(-0.2, 0.2, 41), 0.5 * np.linspace(-0.2, 0.2, 41) - 0.1], 1)
    fl = np.vstack([wall, line, [[0.0, -(R - 0.02)], [0.01, -(R - 0.03)]]]).astype(np.float32)
    nfl = len(fl)
    return [
        ("lifter tip", _drum(R, 0.5, 0.0, balls6, z(6), np.zeros(6), 0.01, 0.03, 1e-6, z(0), [], [], 0.01,
                             lifters=lifters)),
        ("lifter tip clockwise", _drum(R, -0.5, 0.0, balls6, z(6), np.zeros(6), 0.01, 0.03, 1e-6, z(0), [], [],
                                       0.01, lifters=lifters)),
        ("lifters turned away", _drum(R, 0.5, 0.3, balls6, z(6), np.zeros(6), 0.01, 0.03, 1e-6, z(0), [], [],
                                      0.01, lifters=lifters)),
        ("pool, line, depth", _drum(R, 1.0, 0.0, [[0.0, 0.0], [0.015, 0.0]], [[3.0, 4.0], [1.0, 0.0]], [2.0, -1.0],
                                    0.01, 0.25, 0.5, fl, (np.arange(nfl) % 2).astype(np.float32),
                                    np.linspace(900, 1200, nfl), 0.01)),
        ("empty", _drum(R, 0.0, 0.0, z(0), z(0), [], 0.01, 1.0, 1.0, z(0), [], [], 0.01)),
    ]


def test_hand_placed_cases(mm, orc):
    cases = _cases(orc)
    got = mm.drum_metrics_batch([d for _, d in cases])
    for (label, d), g in zip(cases, got):
        print(f"drum metrics: {label}", flush=True)
        _check(g, _oracle(orc, d), label)


@pytest.mark.parametrize("label,params,substeps", [
    ("realtime", {"simulation": {"max_balls": 300, "resolution": 25}}, 60),
    ("2000-disc drum", {"mill": {"diameter_m": 0.1397, "speed_mode": "percent_critical", "speed_value": 75.0},
                        "media": {"density_kg_m3": 7800.0}, "simulation": {"resolution": 100, "max_balls": 2000}}, 12),
    ("lifters clockwise", {"mill": {"direction": "clockwise"}, "lifters": LIFTERS,
                           "simulation": {"max_balls": 400, "resolution": 30}}, 40),
])
def test_stepped_drum_states(mm, orc, label, params, substeps):
    d = _stepped(orc, params, substeps)
    print(f"drum metrics: {label}, {len(d['ball_x'])} balls, {len(d['fluid_x'])} fluid", flush=True)
    _check(mm.drum_metrics_batch([d])[0], _oracle(orc, d), label)


def test_batch_equals_single_drums(mm, orc):
    drums = [d for _, d in _cases(orc)]
    drums += [_stepped(orc, {"simulation": {"max_balls": 200, "resolution": 20, "seed": s}}, 20, seed=s)
              for s in (1, 2, 3)]
    d = dict(drums[-1])  # balls only, then fluid only
    drums.append(dict(d, fluid_x=np.zeros((0, 2), np.float32), dye=np.zeros(0, np.float32),
                      density=np.zeros(0, np.float32)))
    drums.append(dict(d, ball_x=np.zeros((0, 2), np.float32), ball_v=np.zeros((0, 2), np.float32),
                      ball_omega=np.zeros(0, np.float32)))
    print(f"drum metrics batch: {len(drums)} drums", flush=True)
    got = mm.drum_metrics_batch(drums)
    for k, (g, d) in enumerate(zip(got, drums)):
        _check(g, _oracle(orc, d), f"batch drum {k}")


def test_long_wall_lists(mm, orc):
    # 6000 balls and 40000 fluid particles near a lifter wall: toe and pool lists of thousands of
    # angles (several sort tiles), thousands of free-surface points (many sum chunks)
    rng = np.random.default_rng(9)
    R = 0.3
    r = np.float32(R) * np.sqrt(rng.uniform(0.85, 1.0, 6000)).astype(np.float32)
    a = rng.uniform(0.0, 5.0, 6000).astype(np.float32)
    bx = np.stack([r * np.cos(a), r * np.sin(a)], 1).astype(np.float32)
    rf = np.float32(R) * np.sqrt(rng.uniform(0.0, 1.0, 40000)).astype(np.float32)
    af = rng.uniform(3.3, 6.1, 40000).astype(np.float32)
    fx = np.stack([rf * np.cos(af), rf * np.sin(af)], 1).astype(np.float32)
    d = _drum(R, 2.0, 0.7, bx, rng.standard_normal((6000, 2)), rng.standard_normal(6000), 0.002, 1e-4, 1e-10,
              fx, (fx[:, 0] > 0).astype(np.float32), 1000.0 + 30.0 * rng.standard_normal(40000), 0.003,
              lifters={"count": 12, "height_m": 0.02, "base_width_m": 0.02, "top_width_m": 0.01, "phase_deg": 5.0})
    ref = _oracle(orc, d)
    assert ref["has_toe_shoulder"] and ref["has_pool_extent"] and ref["has_free_surface"]
    _check(mm.drum_metrics_batch([d])[0], ref, "long lists")


def test_non_finite_wall_ball_is_an_error(mm, orc):
    R = 0.5
    a = np.float32([1.0, 1.1, 1.2, 1.3, 1.4, 1.5])
    bx = np.vstack([np.stack([np.cos(a) * (R - 0.005), np.sin(a) * (R - 0.005)], 1), [[np.nan, 0.0]]])
    d = _drum(R, 1.0, 0.0, bx, np.zeros((7, 2)), np.zeros(7), 0.01, 1.0, 1.0, np.zeros((0, 2)), [], [], 0.01)
    with pytest.raises(ValueError):
        _oracle(orc, d)
    with pytest.raises(ValueError, match="drum 0"):
        mm.drum_metrics_batch([d])
