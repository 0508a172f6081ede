"""Drum substeps on the GPU through the C ABI (Simulation::step_fixed, lib.rs:438-565).

A substep is the fluid operator (FluidParticles::step_coupled), the DEM operator with the fluid's
coupling impulses (DemState::step_recorded), the accumulators (power, torque, dissipation EMAs,
energy budget, StressAccum::update with the step's impact records) and the drum clock. The checker
is the oracle's Simulation::advance ("mixed" policy: f32 state, f64 sums and CG dot products).

Bars:
- Initial state (dem.rs:355-409, pbf.rs:1044-1119, rng.rs): bit for bit.
- With the viscosity solve held at zero iterations on both sides, every substep is the reference's
  f32 arithmetic: after each of 12 substeps, every ball's x, v, theta, omega and every fluid
  particle's x and v are compared bit for bit, and the impact records; the EMAs (power, torque,
  dissipated power), which add f32 wall work formed from f64 sums, within 1e-6 relative; the
  stress accumulator within 1e-9 relative; the drum clock exactly.
- With the solve (the reference's tolerance), one substep at a time from the same state (the
  oracle is set to the GPU's state before each): ball and fluid positions within 1e-5 of the drum
  radius, velocities within 1e-5 of v_ref = |omega| R + sqrt(g R), angular velocities within 1e-5 of
  v_ref / r, CG and density iteration counts equal.
- A batch of different drums equals the same drums stepped one by one, bit for bit.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REALTIME = {"simulation": {"max_balls": 300, "resolution": 25}}
LIFTERS = {"mill": {"direction": "clockwise", "speed_value": 150.0},
           "lifters": {"count": 6, "height_m": 0.005, "base_width_m": 0.006, "top_width_m": 0.003,
                       "phase_deg": 15.0},
           "media": {"friction_ball_ball_static": 0.4, "friction_ball_wall_static": 0.6},
           "slurry": {"rheology": "bingham", "yield_stress_pa": 0.5, "viscosity_pa_s": 0.01},
           "simulation": {"max_balls": 250, "resolution": 35}}


@pytest.fixture(scope="module")
def mg():
    import milldynamics3_amd as m
    return m


def _bits(a):
    return np.ascontiguousarray(np.asarray(a, np.float32)).view(np.uint32)


def _same(g, r, what):
    g, r = _bits(g), _bits(r)
    bad = np.flatnonzero((g != r).reshape(len(g), -1).any(1)) if g.size else np.zeros(0, int)
    assert bad.size == 0, f"{what} differs at {bad[:8].tolist()} of {len(g)}"


def test_initial_state_matches_oracle(mg, orc):
    g = mg.Simulation(REALTIME)
    o = orc.OracleSim(REALTIME, "mixed")
    st = o.state()
    gb, gf = g.balls(), g.fluid()
    assert (g.n_balls, g.n_fluid) == (407, 794)
    assert np.array_equal(gb.x.astype(np.float64), st["ball_x"])
    assert np.array_equal(gf.x.astype(np.float64), st["fluid_x"])
    assert np.array_equal(gb.v.astype(np.float64), st["ball_v"])
    assert mg.drum_engine_enabled()


@pytest.mark.parametrize("label,params", [("realtime", REALTIME), ("lifters, bingham", LIFTERS)])
def test_substeps_without_solve_are_bit_exact(mg, orc, label, params):
    g = mg.Simulation(params)
    o = orc.OracleSim(params, "mixed")
    g.force_iterations(3, 0)
    o.force_iterations(3, 0)
    for s in range(12):
        print(f"{label}: substep {s}", flush=True)
        g.step_fixed(1)
        o.step_fixed(1)
        st, gb, gf = o.state(), g.balls(), g.fluid()
        for key, got in (("ball_x", gb.x), ("ball_v", gb.v), ("ball_theta", gb.theta), ("ball_omega", gb.omega),
                         ("fluid_x", gf.x), ("fluid_v", gf.v)):
            _same(got, st[key], f"{label} substep {s}: {key}")
        gi, oi = g.impacts(), o.impacts()
        assert gi.shape == oi.shape, f"{label} substep {s}: {len(gi)} vs {len(oi)} impact records"
        assert np.array_equal(gi, oi), f"{label} substep {s}: impact records differ"
        gs, os_ = g.stats(), o.scalars()
        assert gs["drum_angle"] == os_["drum_angle"] and gs["sim_time"] == os_["sim_time"]
        for k in ("power_draw_w", "torque_nm", "dissipated_power_w"):
            assert abs(gs[k] - os_[k]) <= 1e-6 * max(abs(os_[k]), 1e-12), f"{label} substep {s}: {k}"
        for k in ("density_iterations", "n_groups", "n_links", "links_in_range", "n_ball_pairs", "collision_count"):
            assert gs[k] == os_[k], f"{label} substep {s}: {k} gpu {gs[k]} oracle {os_[k]}"
        gst, ost = g.stress_state(), o.stress()
        assert np.allclose(gst, ost, rtol=1e-9, atol=1e-12 * max(np.abs(ost).max(), 1.0)), f"{label} substep {s}"


@pytest.mark.parametrize("label,params", [("realtime", REALTIME), ("lifters, bingham", LIFTERS)])
def test_one_substep_matches_oracle(mg, orc, label, params):
    g = mg.Simulation(params)
    o = orc.OracleSim(params, "mixed")
    R = o.dem_params()["radius_m"]
    vref = abs(o.omega) * R + math.sqrt(9.81 * R)
    r = o.ball_radius
    for s in range(6):
        gb, gf = g.balls(), g.fluid()
        guess, has_guess = g.viscous_guess()
        o.set_state({"ball_x": gb.x, "ball_v": gb.v, "ball_theta": gb.theta, "ball_omega": gb.omega,
                     "fluid_x": gf.x, "fluid_v": gf.v, "guess": guess, "has_guess": has_guess})
        o.set_scalars(g.drum_angle(), g.sim_time())
        print(f"{label}: substep {s}", flush=True)
        g.step_fixed(1)
        o.step_fixed(1)
        st, gb, gf, gs, os_ = o.state(), g.balls(), g.fluid(), g.stats(), o.scalars()
        assert gs["viscosity_iterations"] == os_["viscosity_iterations"], f"{label} substep {s}: CG iterations"
        assert gs["density_iterations"] == os_["density_iterations"], f"{label} substep {s}: density passes"
        for key, got, scale in (("ball_x", gb.x, R), ("fluid_x", gf.x, R), ("ball_v", gb.v, vref),
                                ("fluid_v", gf.v, vref), ("ball_omega", gb.omega, vref / r)):
            err = np.abs(got - st[key]).max() / scale if len(got) else 0.0
            assert err <= 1e-5, f"{label} substep {s}: {key} off by {err:.2e} of its scale"


def test_batch_equals_single_drums(mg):
    plist = [REALTIME, LIFTERS, {"simulation": {"max_balls": 150, "resolution": 30}}]
    b = mg.Batch(plist)
    singles = [mg.Simulation(p) for p in plist]
    b.step(4)
    b.sync()
    for s in singles:
        s.step_fixed(4)
    for k, (d, s) in enumerate(zip(b.drums, singles)):
        db, sb, df, sf = d.balls(), s.balls(), d.fluid(), s.fluid()
        for what, x, y in (("ball_x", db.x, sb.x), ("ball_v", db.v, sb.v), ("ball_omega", db.omega, sb.omega),
                           ("fluid_x", df.x, sf.x), ("fluid_v", df.v, sf.v)):
            _same(x, y, f"drum {k}: {what}")
        assert d.power_draw_w() == s.power_draw_w(), f"drum {k}: power"
        assert np.array_equal(d.stress_state(), s.stress_state()), f"drum {k}: stress"
    b.close()
