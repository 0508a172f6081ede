"""Served sweeps on the GPU (mg_server_*): continuous batching of independent drum runs.

Bars:
- A run served beside other runs (admitted while others run, waiting in the queue for a slot,
  stepping in batches whose membership changes every few ticks) ends with exactly the results of
  the same drum stepped alone with Simulation.step_fixed for the same number of substeps: power,
  torque, dissipated power and collision rate equal bit for bit, and each per-revolution record
  equals the lone drum's values at the substep where its clock first reaches that revolution
  (mixing index and the revolution's impact-energy histogram included).
- Admission: a budget that holds one drum's footprint runs one drum at a time, in arrival order;
  a drum too large for the server or invalid Params are refused at submission.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FAST = {"mill": {"speed_value": 200.0}, "simulation": {"max_balls": 300, "resolution": 25}}
LIFT = {"mill": {"direction": "clockwise", "speed_value": 240.0},
        "lifters": {"count": 6, "height_m": 0.005, "base_width_m": 0.006, "top_width_m": 0.003, "phase_deg": 15.0},
        "simulation": {"max_balls": 250, "resolution": 30}}
WIDE = {"mill": {"diameter_m": 0.08, "speed_mode": "percent_critical", "speed_value": 140.0},
        "simulation": {"max_balls": 400, "resolution": 30}}


@pytest.fixture(scope="module")
def srv():
    from milldynamics3_amd import server
    return server


def _alone(params, substeps, rev_period):
    """The drum stepped alone: final scalars and the values at each revolution boundary."""
    from milldynamics3_amd import Simulation
    s = Simulation(params)
    revs, nxt = [], rev_period
    for _ in range(substeps):
        s.step_fixed(1)
        if s.sim_time() >= nxt:
            mix = s.mixing_index()
            revs.append({"power": s.power_draw_w(), "torque": s.torque_nm(), "coll": s.collision_rate_per_s(),
                         "diss": s.dissipated_power_w(), "mix": mix, "sim_time": s.sim_time()})
            nxt += rev_period
    out = {"power": s.power_draw_w(), "torque": s.torque_nm(), "coll": s.collision_rate_per_s(),
           "diss": s.dissipated_power_w(), "revs": revs}
    s.close()
    return out


def _f32(x):
    return float(np.float32(x))


def test_served_runs_equal_drums_stepped_alone(srv):
    s = srv.Server(max_drums=2, max_balls=2048)
    plan = [("fast", FAST, 1.2), ("lifters", LIFT, 0.6), ("wide", WIDE, 1.1)]
    ids = []
    for label, p, revs in plan:
        ids.append(s.submit(p, revs))
        print(f"submitted {label}", flush=True)
    st = s.poll(1)
    assert (st["running"], st["queued"]) == (2, 1), st  # two slots: the third waits
    ticks = 1
    while st["running"] or st["queued"]:
        st = s.poll(16)
        ticks += 16
        assert ticks < 20000
    print(f"served: {st}", flush=True)
    for (label, p, revs), rid in zip(plan, ids):
        r = s.result(rid)
        assert r["state"] == "done" and r["substeps_done"] == r["substeps_total"], label
        assert r["arrival_s"] <= r["start_s"] <= r["done_s"], label
        from milldynamics3_amd import Simulation
        probe = Simulation(p)
        dt = probe.fixed_sub_dt()
        probe.close()
        period = 2.0 * math.pi / _omega(p)
        assert r["substeps_total"] == max(1, math.ceil(revs * period / dt)), label
        a = _alone(p, r["substeps_total"], period)
        print(f"{label}: {r['substeps_total']} substeps, {len(a['revs'])} revolution records", flush=True)
        for k, key in (("power", "power_draw_w"), ("torque", "torque_nm"), ("coll", "collision_rate_per_s"),
                       ("diss", "dissipated_power_w")):
            assert _f32(r[key]) == _f32(a[k]), f"{label}: {key} served {r[key]!r} alone {a[k]!r}"
        stream = s.stream(rid)
        assert len(stream) == len(a["revs"]) == r["revolutions_done"], label
        for q, b in zip(stream, a["revs"]):
            assert q["sim_time_s"] == b["sim_time"], label
            assert _f32(q["power_draw_w"]) == _f32(b["power"]) and _f32(q["torque_nm"]) == _f32(b["torque"]), label
            assert _f32(q["collision_rate_per_s"]) == _f32(b["coll"]), label
            assert _f32(q["dissipated_power_w"]) == _f32(b["diss"]), label
            if b["mix"] is None:
                assert q["mixing_index"] is None, label
            else:
                assert _f32(q["mixing_index"]) == _f32(b["mix"]), label
            assert len(q["impact_energy_histogram"]) == 12 and min(q["impact_energy_histogram"]) >= 0, label
        s.release(rid)
    s.close()


def _omega(p):
    """|omega| of the Params (rpm 2 pi / 60, critical speed 42.3 / sqrt(D) rpm), in f32 as the library."""
    from milldynamics3_amd.sim import DEFAULT_PARAMS
    mill = dict(DEFAULT_PARAMS["mill"], **p.get("mill", {}))
    d = np.float32(mill["diameter_m"])
    v = np.float32(mill["speed_value"])
    rpm = v if mill["speed_mode"] == "rpm" else np.float32(np.float32(42.3) / np.sqrt(d)) * v / np.float32(100.0)
    return float(np.float32(np.float32(rpm) * np.float32(2.0) * np.float32(math.pi) / np.float32(60.0)))


def test_impact_histograms_sum_to_the_runs_counts(srv):
    """Each revolution's histogram counts the impacts of its own substeps only: their total over
    the records equals the lone drum's histogram summed over the same substeps."""
    from milldynamics3_amd import Simulation
    s = srv.Server(max_drums=4, max_balls=2048)
    rid = s.submit(FAST, 2.05)
    while s.poll(32)["running"] or s.poll(0)["queued"]:
        pass
    r, stream = s.result(rid), s.stream(rid)
    assert len(stream) == 2
    g = Simulation(FAST)
    period = 2.0 * math.pi / _omega(FAST)
    hist, nxt, got = np.zeros(12, np.int64), period, []
    for _ in range(r["substeps_total"]):
        g.step_fixed(1)
        hist += np.asarray(g.stats()["impact_energy_histogram"], np.int64)
        if g.sim_time() >= nxt:
            got.append(hist.copy())
            hist[:] = 0
            nxt += period
    for q, h in zip(stream, got):
        assert q["impact_energy_histogram"] == h.tolist()
    s.close()


def test_admission_by_budget_and_refusals(srv):
    probe = srv.Server(max_drums=4, max_balls=2048)
    rid = probe.submit(FAST, 0.01)
    fp = probe.result(rid)["footprint_bytes"]
    probe.close()
    # a budget for the slots' fixed workspaces and one drum: two runs go one after the other
    base = srv.Server(max_drums=2, max_balls=512)
    used0 = base.poll(0)["used_bytes"]
    base.close()
    s = srv.Server(max_drums=2, max_balls=512, hbm_budget_bytes=used0 + fp + fp // 2)
    a, b = s.submit(FAST, 0.02), s.submit(FAST, 0.02)
    st = s.poll(1)
    assert (st["running"], st["queued"]) == (1, 1), st
    assert s.result(a)["state"] == "running" and s.result(b)["state"] == "queued"
    while st["running"] or st["queued"]:
        st = s.poll(8)
    ra, rb = s.result(a), s.result(b)
    assert ra["done_s"] <= rb["start_s"]
    assert ra["power_draw_w"] == rb["power_draw_w"]  # the same drum, alone both times
    with pytest.raises(MemoryError):
        s.submit({"mill": {"diameter_m": 0.09}, "simulation": {"resolution": 25}}, 1.0)  # 830 balls > 512
    with pytest.raises(ValueError):
        s.submit({"mill": {"diameter_m": -1.0}}, 1.0)
    with pytest.raises(ValueError):
        s.submit({"mill": {"speed_value": 0.0}}, 1.0)
    with pytest.raises(ValueError):
        s.submit(FAST, 0.0)
    s.close()
