"""GPU parity of DemState::step_recorded (dem.rs:530-1066) through mg_dem_step.

Bar: bit-exact against the oracle's standalone DEM step (oracle/mg_oracle.cpp Sim::dem_step, pinned
by the reference's DEM unit tests in tests/test_dem_oracle.py): after every step, every ball's
x, v, theta and omega equal as f32, the DemStepStats (wall work, dissipated and coupling energy,
collisions, impact histogram, largest displacement, candidate pairs) and the impact records
(kind, balls, speeds, restitution) equal. Populations, stepped together in one batch: the
realtime preset's 407 balls, the 2000-disc drum at 75 % critical speed, a clockwise lifter drum
with static friction above kinetic (the exp blend) and coupling impulses, balls fired through
the wall (the swept time-of-impact clamp), and a dozen balls spread over a large drum (the Sparse
grid layout); a cluster of heavily overlapping balls (the projection's restart with every
candidate pair).
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LIFTERS = {"count": 6, "height_m": 0.005, "base_width_m": 0.006, "top_width_m": 0.003, "phase_deg": 15.0}


@pytest.fixture(scope="module")
def dm():
    from milldynamics3_amd import dem
    return dem


def _pop(orc, params, initial=None, ext_seed=None, state_edit=None, mode="ref"):
    o = orc.OracleSim(dict(params, slurry={"enabled": False}), mode, initial_balls=initial)
    if state_edit is not None:
        st = o.state()
        state_edit(st)
        o.set_state(st)
    st, pr = o.state(), o.dem_params()
    p = orc._merge(orc._DEFAULTS, params)
    pop = {"radius_m": pr["radius_m"], "omega": pr["omega"], "drum_angle": 0.0,
           "lifters": {"count": int(pr["lifter_count"]), "height_m": pr["lifter_height_m"],
                       "base_width_m": pr["lifter_base_width_m"], "top_width_m": pr["lifter_top_width_m"],
                       "phase_deg": pr["lifter_phase_deg"]},
           "ball_radius": pr["ball_radius"], "ball_mass": pr["ball_mass"], "ball_inertia": pr["ball_inertia"],
           "iterations": p["simulation"]["dem_iterations"], "dt": o.fixed_sub_dt,
           "x": st["ball_x"].astype(np.float32), "v": st["ball_v"].astype(np.float32),
           "theta": st["ball_theta"].astype(np.float32), "omega_b": st["ball_omega"].astype(np.float32),
           "energy_f64": mode != "ref"}
    for k in ("restitution_ball_ball", "restitution_ball_wall", "friction_ball_ball", "friction_ball_wall",
              "friction_ball_ball_static", "friction_ball_wall_static", "friction_velocity_scale_m_s",
              "rolling_friction"):
        pop[k] = pr[k]
    rng = np.random.default_rng(ext_seed) if ext_seed is not None else None
    return o, pop, rng


def _bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


def _check(o, pop, stats, label):
    st = o.state()
    for key, ok in (("x", "ball_x"), ("v", "ball_v"), ("theta", "ball_theta"), ("omega_b", "ball_omega")):
        g, r = _bits(pop[key]), _bits(st[ok])
        bad = np.flatnonzero((g != r).reshape(len(g), -1).any(1))
        assert bad.size == 0, f"{label}: {key} differs at balls {bad[:8].tolist()}"
    return st


def _compare_stats(g, r, imp, label):
    for k in ("wall_work_j", "dissipated_energy_j", "coupling_work_j", "max_substep_displacement_over_diameter"):
        assert _bits(g[k]) == _bits(r[k]), f"{label}: {k} gpu {g[k]!r} oracle {r[k]!r}"
    assert g["collision_count"] == r["collision_count"] and g["n_pairs"] == r["n_pairs"], label
    assert g["impact_energy_histogram"] == [int(r[f"impact_energy_histogram_{b}"]) for b in range(12)], label
    gi = g["impacts"]
    assert len(gi) == len(imp), f"{label}: {len(gi)} records vs {len(imp)}"
    if len(imp):
        part = np.where(imp[:, 2] < 0, 0xFFFFFFFF, imp[:, 2]).astype(np.uint64)
        assert (gi["kind"] == imp[:, 0]).all() and (gi["ball"] == imp[:, 1]).all(), label
        assert (gi["partner"].astype(np.uint64) == part).all(), label
        for c, name in ((3, "normal_speed"), (4, "slip_speed"), (5, "restitution")):
            assert (_bits(gi[name]) == _bits(imp[:, c])).all(), f"{label}: record {name} differs"


def _run(dm, cases, steps, pairs_per_ball=24):
    """Steps every case `steps` times on both sides; returns the GPU stats of the first step."""
    batch = dm.DemBatch(len(cases), max(max(len(p["x"]) for _, _, p, _ in cases), 1), pairs_per_ball=pairs_per_ball,
                        max_impacts=8192)
    first = None
    for s in range(steps):
        refs = []
        for label, o, p, rng in cases:
            ext_i = ext_a = None
            if rng is not None:
                n = len(p["x"])
                scale = p["ball_mass"] * 0.05
                ext_i = (rng.standard_normal((n, 2)) * scale).astype(np.float32)
                ext_a = (rng.standard_normal(n) * p["ball_inertia"] * 5.0).astype(np.float32)
                p["ext_impulse"], p["ext_angular"] = ext_i, ext_a
            r = o.dem_step(p["drum_angle"], p["iterations"], p["dt"], ext_i, ext_a)
            refs.append((r, o.impacts()))
        print(f"dem step {s}: " + ", ".join(f"{label} {len(p['x'])}" for label, _, p, _ in cases), flush=True)
        got = batch.step([p for _, _, p, _ in cases])
        first = first or got
        for (label, o, p, _), g, (r, imp) in zip(cases, got, refs):
            _check(o, p, g, f"{label} step {s}")
            _compare_stats(g, r, imp, f"{label} step {s}")
            a = np.float32(p["drum_angle"]) + np.float32(p["omega"]) * np.float32(p["dt"])
            p["drum_angle"] = float(np.fmod(a, np.float32(2 * math.pi)))
    batch.close()
    return first


def test_populations_in_one_batch_match_oracle(dm, orc):
    cases = []
    o, p, rng = _pop(orc, {"simulation": {"max_balls": 300, "resolution": 25}})
    cases.append(("realtime", o, p, rng))
    o, p, rng = _pop(orc, {"mill": {"diameter_m": 0.1397, "speed_mode": "percent_critical", "speed_value": 75.0},
                           "media": {"density_kg_m3": 7800.0}, "simulation": {"max_balls": 2000}})
    cases.append(("2000 balls", o, p, rng))
    o, p, rng = _pop(orc, {"mill": {"direction": "clockwise", "speed_value": 180.0}, "lifters": LIFTERS,
                           "media": {"friction_ball_ball_static": 0.4, "friction_ball_wall_static": 0.6},
                           "simulation": {"max_balls": 400}}, ext_seed=3)
    cases.append(("lifters, static friction, impulses", o, p, rng))
    first = _run(dm, cases, 12)
    print("projection restarts at the first step: " + ", ".join(
        f"{label} {g['projection_restarts']}" for (label, _, _, _), g in zip(cases, first)), flush=True)


def test_balls_through_the_wall_use_the_swept_clamp(dm, orc):
    # balls next to the wall fired outward fast enough to end a step deep in it
    def fire(st):
        x = st["ball_x"]
        d = np.linalg.norm(x, axis=1)
        out = x / np.maximum(d, 1e-9)[:, None]
        far = np.argsort(-d)[:20]
        st["ball_v"][far] = out[far] * 60.0
    o, p, rng = _pop(orc, {"lifters": LIFTERS, "mill": {"speed_value": 300.0}, "simulation": {"max_balls": 200}},
                     state_edit=fire)
    # the fast balls widen the grid cell (2 r + max speed * dt + margin): most balls share a few
    # cells, so the candidate pairs number a couple of hundred per ball
    first = _run(dm, [("fired", o, p, rng)], 6, pairs_per_ball=320)
    assert first[0]["n_swept"] > 0  # the clamp ran


def test_too_many_pairs_is_reported(dm, orc):
    def fire(st):
        st["ball_v"][:20] = 60.0
    o, p, _ = _pop(orc, {"simulation": {"max_balls": 200}}, state_edit=fire)
    batch = dm.DemBatch(1, len(p["x"]), pairs_per_ball=8)
    with pytest.raises(OverflowError):
        batch.step([p])
    batch.close()


def test_sparse_layout_of_a_few_spread_balls(dm, orc):
    rng = np.random.default_rng(8)
    ang = rng.uniform(0, 2 * math.pi, 12)
    rad = rng.uniform(0.0, 0.9, 12)
    xy = np.stack([rad * np.cos(ang), rad * np.sin(ang)], 1).astype(np.float32)
    xy[3] = xy[2] + np.float32(0.0015)  # one touching pair
    o, p, _ = _pop(orc, {"mill": {"diameter_m": 2.0, "speed_value": 5.0}, "media": {"ball_diameter_m": 0.002},
                         "simulation": {"coarse_graining_mode": "off"}}, initial=xy)
    _run(dm, [("sparse", o, p, None)], 8)


def test_overlapping_cluster_restarts_the_projection(dm, orc):
    """The projection walks only the pairs closer than 2.5 r at the start and checks that every
    other candidate pair's start distance less both balls' travel stays above 2 r. A hexagonal
    cluster at spacing 1.4 r (every neighbour overlapping by 0.6 r, each correction moving a ball
    0.2 r, six neighbours, two iterations) travels far more than the 0.5 r margin, so the check fails
    and the projection restarts (with a wider set of pairs, or every candidate pair); the results
    stay the reference's."""
    r = 0.001
    pts = []
    for a in range(-4, 5):
        for b in range(-4, 5):
            if abs(a + b) <= 4:
                pts.append(((a + 0.5 * b) * 1.4 * r, b * 1.4 * r * math.sqrt(3.0) / 2.0 - 0.012))
    xy = np.asarray(pts, np.float32)
    o, p, _ = _pop(orc, {"media": {"ball_diameter_m": 2 * r}, "simulation": {"coarse_graining_mode": "off"}},
                   initial=xy)
    first = _run(dm, [("cluster", o, p, None)], 4)
    assert first[0]["projection_restarts"] >= 1


def test_f64_energy_sums_match_the_mixed_policy(dm, orc):
    """energy_f64 = 1: the energies and wall work are the same f32 terms summed in f64 in ball order
    (the oracle's "mixed" policy, which the drum engine runs); the state stays bit-exact and so do the
    f32 roundings of the f64 sums (wall work, dissipated energy)."""
    cases = []
    o, p, rng = _pop(orc, {"simulation": {"max_balls": 300, "resolution": 25}}, mode="mixed")
    cases.append(("realtime", o, p, rng))
    o, p, rng = _pop(orc, {"mill": {"direction": "clockwise", "speed_value": 180.0}, "lifters": LIFTERS,
                           "simulation": {"max_balls": 400}}, ext_seed=5, mode="mixed")
    cases.append(("lifters, impulses", o, p, rng))
    _run(dm, cases, 8)
