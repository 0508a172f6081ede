"""Pins of the oracle's DemState::step_recorded restatement (dem.rs:530-1066) run on its own: the
reference's DEM unit tests (dem.rs:1148-1560) ported onto the oracle's standalone DEM step —
rest on a flat wall and on a lifter tip, the e^2 rebound height, x-momentum of a head-on
collision, rolling without slip on a turning wall, a recorded head-on impact's speeds — plus the
sanitize rule (a non-finite ball restarts at the origin at rest)."""
import math

import numpy as np


def _sim(orc, drum_radius, r, xy, media=None, omega=0.0, lifters=None):
    params = {"mill": {"diameter_m": 2.0 * drum_radius, "speed_mode": "rpm",
                       "speed_value": omega * 60.0 / (2.0 * math.pi)},
              "media": dict({"ball_diameter_m": 2.0 * r, "density_kg_m3": 7800.0}, **(media or {})),
              "slurry": {"enabled": False},
              "lifters": lifters or {"count": 0},
              "simulation": {"coarse_graining_mode": "off"}}
    return orc.OracleSim(params, "ref", initial_balls=np.asarray(xy, np.float32))


def _set_v(o, v):
    st = o.state()
    st["ball_v"] = np.asarray(v, np.float64).reshape(-1, 2)
    o.set_state(st)


def test_ball_at_rest_on_flat_wall_stays_put(orc):
    # dem.rs:1148-1184
    R, r = 10.0, 0.05
    o = _sim(orc, R, r, [[0.0, -(R - r)]])
    for _ in range(240):
        o.dem_step(0.0, 4, 1.0 / 240.0)
    st = o.state()
    x, v = st["ball_x"][0], st["ball_v"][0]
    d = R - math.hypot(*x)
    assert r * 0.9 <= d <= r * 1.02 and abs(x[0]) < 0.02 and math.hypot(*v) < 0.05


def test_ball_rests_on_a_lifter_tip(orc):
    # dem.rs:1186-1234
    R, r = 10.0, 0.01
    lifters = {"count": 1, "height_m": 0.05, "base_width_m": 0.1, "top_width_m": 0.06, "phase_deg": -90.0}
    tip = R - 0.05
    o = _sim(orc, R, r, [[0.0, -(tip - r)]], lifters=lifters)
    for _ in range(240):
        o.dem_step(0.0, 4, 1.0 / 240.0)
    st = o.state()
    x, v = st["ball_x"][0], st["ball_v"][0]
    d = tip - math.hypot(*x)
    assert r * 0.9 <= d <= r * 1.02 and abs(x[0]) < 0.005 and math.hypot(*v) < 0.05


def test_rebound_height_is_e_squared(orc):
    # dem.rs:1236-1287
    R, r, e, h0 = 10.0, 0.05, 0.8, 0.5
    o = _sim(orc, R, r, [[0.0, -(R - r) + h0]],
             media={"restitution_ball_wall": e, "friction_ball_wall": 0.0, "friction_ball_wall_static": 0.0,
                    "rolling_friction": 0.0})
    dt = 1.0 / 480.0
    top, bounced, after = -1e9, False, 0
    for _ in range(20000):
        o.dem_step(0.0, 4, dt)
        st = o.state()
        height = R - math.hypot(*st["ball_x"][0]) - r
        vy = st["ball_v"][0][1]
        if not bounced:
            bounced = vy > 0.0 and height < h0 * 0.5
        else:
            top = max(top, height)
            if vy <= 0.0:
                after += 1
                if after > 5:
                    break
    assert bounced and abs(top - e * e * h0) / (e * e * h0) < 0.25


def test_head_on_collision_conserves_x_momentum(orc):
    # dem.rs:1289-1332
    r = 0.05
    media = {"restitution_ball_ball": 0.9, "friction_ball_ball": 0.0, "friction_ball_ball_static": 0.0,
             "rolling_friction": 0.0}
    o = _sim(orc, 1000.0, r, [[-0.5, 0.0], [0.5, 0.0]], media=media)
    _set_v(o, [[2.0, 0.0], [-2.0, 0.0]])
    px0 = o.ball_mass * (2.0 - 2.0)
    for _ in range(600):
        o.dem_step(0.0, 4, 1.0 / 480.0)
    st = o.state()
    px1 = o.ball_mass * (st["ball_v"][0][0] + st["ball_v"][1][0])
    assert abs(px1 - px0) / max(abs(px0), 1e-9) < 0.05  # the symmetric pair keeps px = 0 exactly
    assert st["ball_x"][1][0] > st["ball_x"][0][0] + 2.0 * r * 0.9
    # they bounced apart at 0.9 of the approach speed
    assert abs(st["ball_v"][1][0] - 0.9 * 2.0) < 1e-3 and st["ball_v"][0][0] < 0.0


def test_rolling_without_slip_on_a_turning_wall(orc):
    # dem.rs:1334-1380
    R, r, omega = 10.0, 0.05, 0.5
    o = _sim(orc, R, r, [[0.0, -(R - r)]], omega=omega)
    w = o.omega
    for _ in range(5 * 480):
        o.dem_step(0.0, 4, 1.0 / 480.0)
    st = o.state()
    x, v, om = st["ball_x"][0], st["ball_v"][0], st["ball_omega"][0]
    d, n = orc.sdf_world(orc.drum_array(R, w), x, 0.0)
    t = np.array([-n[1], n[0]])
    c = x - r * n
    vw = w * np.array([-c[1], c[0]])
    slip = float((v - vw) @ t - r * om)
    assert abs(slip) < 0.05 * max(float(np.linalg.norm(vw)), 1e-6)


def test_recorded_head_on_impact(orc):
    # dem.rs:1519-1559
    r = 0.05
    media = {"restitution_ball_ball": 0.6, "friction_ball_ball": 0.0, "friction_ball_ball_static": 0.0,
             "rolling_friction": 0.0}
    o = _sim(orc, 1000.0, r, [[-0.99 * r, 0.0], [0.99 * r, 0.0]], media=media)
    _set_v(o, [[1.0, 0.0], [-1.0, 0.0]])
    s = o.dem_step(0.0, 4, 1.0 / 480.0)
    imp = o.impacts()
    assert s["collision_count"] == 1 and len(imp) == 1
    kind, ball, partner, vn, slip, e = imp[0]
    assert kind == 0 and abs(vn - 2.0) < 1e-4 and slip < 1e-4 and abs(e - 0.6) < 1e-6


def test_non_finite_ball_restarts_at_rest_at_the_origin(orc):
    o = _sim(orc, 0.5, 0.01, [[0.0, -0.3], [0.1, -0.3]])
    st = o.state()
    st["ball_x"][1] = (np.nan, 0.0)
    o.set_state(st)
    o.dem_step(0.0, 2, 1e-4)
    x, v = o.state()["ball_x"][1], o.state()["ball_v"][1]
    # sanitized to (0, 0) at rest, then one predict step under gravity
    assert x[0] == 0.0 and abs(x[1] - (-9.81 * 1e-4) * 1e-4) < 1e-9 and v[0] == 0.0
