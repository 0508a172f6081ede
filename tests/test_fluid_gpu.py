"""GPU parity of FluidParticles::step_coupled (pbf.rs:1174-2365) through mg_fluid_step.

The checker is the oracle's fluid step (oracle/mg_oracle.cpp Sim::fluid_step, "mixed" policy: f32
state, f64 energies, wall work and CG dot products), run on the same drum state with the same balls.

Bars:
- Before the viscosity solve everything is the reference's f32 arithmetic, so with the solve held
  at zero iterations on both sides (density passes forced to pbf_iterations, CG iterations to 0)
  fluid x, v, the warm start, every ball's impulse and angular impulse, the density pass count and
  the group / link / in-range counts are compared bit for bit. The f64 sums (wall impulse, angular
  impulse, kinetic-energy deltas) are formed per tile on the GPU and sequentially by the oracle:
  within 1e-9 of their scale; the f32 wall bins within 1e-5 relative; wall work (f32 of an f64 sum)
  within 1e-6 relative.
- With the solve (adaptive, the reference's tolerance 1e-3 and caps 50 / 200): positions are
  bit-exact (the solve moves only velocities); iteration counts equal; velocities within 1e-5 of the
  drum's largest speed and impulses within 1e-5 of the largest impulse (the f64 dot products are
  summed in another order, so the iterates agree to rounding).
Cases, stepped together in one batch: the realtime preset (407 balls, 794 fluid particles), the
2000-disc drum (12,697 fluid particles), a clockwise lifter drum with a Bingham slurry, an inviscid
slurry (no solve) and a drum without balls.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LIFTERS = {"count": 6, "height_m": 0.005, "base_width_m": 0.006, "top_width_m": 0.003, "phase_deg": 15.0}
C2 = {"mill": {"diameter_m": 0.1397, "speed_mode": "percent_critical", "speed_value": 75.0},
      "media": {"density_kg_m3": 7800.0}, "simulation": {"resolution": 100, "max_balls": 2000}}
CASES = [
    ("realtime", {"simulation": {"max_balls": 300, "resolution": 25}}, 30),
    ("2000-disc drum", C2, 4),
    ("lifters, bingham", {"mill": {"direction": "clockwise", "speed_value": 180.0}, "lifters": LIFTERS,
                          "slurry": {"rheology": "bingham", "yield_stress_pa": 0.5, "viscosity_pa_s": 0.01},
                          "simulation": {"max_balls": 250, "resolution": 40}}, 20),
    ("inviscid", {"slurry": {"viscosity_pa_s": 0.0}, "simulation": {"max_balls": 200, "resolution": 30}}, 12),
    ("no balls", {"media": {"fill_fraction": 0.0}, "simulation": {"resolution": 30}}, 10),
]


@pytest.fixture(scope="module")
def fl():
    from milldynamics3_amd import fluid
    return fluid


def _bits(a):
    return np.ascontiguousarray(np.asarray(a, np.float32)).view(np.uint32)


def _case(orc, params, substeps):
    o = orc.OracleSim(params, "mixed")
    o.step_fixed(substeps)
    st, fp, dp = o.state(), o.fluid_params(), o.dem_params()
    drum = {"radius_m": dp["radius_m"], "omega": dp["omega"], "drum_angle": o.scalars()["drum_angle"],
            "lifters": {"count": int(dp["lifter_count"]), "height_m": dp["lifter_height_m"],
                        "base_width_m": dp["lifter_base_width_m"], "top_width_m": dp["lifter_top_width_m"],
                        "phase_deg": dp["lifter_phase_deg"]},
            "rheology": int(fp["rheology"]), "iterations": int(fp["pbf_iterations"]), "dt": o.fixed_sub_dt,
            "x": st["fluid_x"].astype(np.float32), "v": st["fluid_v"].astype(np.float32),
            "guess": st["guess"].astype(np.float32), "has_guess": st["has_guess"],
            "ball_x": st["ball_x"].astype(np.float32), "ball_v": st["ball_v"].astype(np.float32),
            "ball_omega": st["ball_omega"].astype(np.float32)}
    for k in ("h", "rest_density", "particle_mass", "ball_radius", "ball_mass", "ball_inertia", "ball_true_radius",
              "viscosity_pa_s", "yield_stress_pa", "wall_no_slip", "ball_no_slip", "wettability",
              "surface_tension_n_m"):
        drum[k] = fp[k]
    return o, drum


def _assert_bits(g, r, what):
    g, r = _bits(g), _bits(r)
    bad = np.flatnonzero((g != r).reshape(len(g), -1).any(1)) if g.size else np.zeros(0, int)
    assert bad.size == 0, f"{what} differs at {bad[:8].tolist()} of {len(g)}"


def _close(g, r, scale, tol, what):
    assert abs(g - r) <= tol * max(abs(scale), 1e-30), f"{what}: gpu {g!r} oracle {r!r}"


def _compare_sums(label, g, r):
    ke = max(abs(r[k]) for k in ("ke_delta_noslip_j", "ke_delta_drag_j", "ke_delta_buoyancy_j",
                                 "ke_delta_viscosity_j", "ke_delta_clamp_j"))
    for k in ("ke_delta_noslip_j", "ke_delta_drag_j", "ke_delta_buoyancy_j", "ke_delta_viscosity_j",
              "ke_delta_clamp_j"):
        _close(g[k], r[k], ke, 1e-9, f"{label}: {k}")
    wi = max(abs(r["wall_impulse_x"]), abs(r["wall_impulse_y"]))
    _close(g["wall_impulse_x"], r["wall_impulse_x"], wi, 1e-9, f"{label}: wall_impulse_x")
    _close(g["wall_impulse_y"], r["wall_impulse_y"], wi, 1e-9, f"{label}: wall_impulse_y")
    _close(g["wall_angular_impulse"], r["wall_angular_impulse"], abs(r["wall_angular_impulse"]), 1e-9,
           f"{label}: wall_angular_impulse")
    _close(g["wall_work_j"], r["wall_work_j"], r["wall_work_j"], 1e-6, f"{label}: wall_work_j")
    rb = np.asarray(r["wall_angle_bins"], np.float64)
    assert np.allclose(g["wall_angle_bins"], rb, rtol=1e-5, atol=1e-5 * max(np.abs(rb).max(), 1e-30)), label
    for k in ("n_groups", "n_links", "links_in_range", "density_iterations", "coupling_clamp_hits"):
        assert g[k] == int(r[k]), f"{label}: {k} gpu {g[k]} oracle {r[k]}"


def _run(orc, fl, force):
    built = [(label,) + _case(orc, p, k) for label, p, k in CASES]
    batch = fl.FluidBatch(len(built))
    if force:
        batch.force_iterations(3, 0)
    refs = []
    for label, o, d in built:
        if force:
            o.force_iterations(3, 0)
        imp, ang, st = o.fluid_step(d["drum_angle"], d["iterations"], d["dt"])
        refs.append((imp, ang, st, o.state()))
    print("fluid step: " + ", ".join(f"{label} {len(d['x'])}/{len(d['ball_x'])}" for label, _, d in built),
          flush=True)
    got = batch.step([d for _, _, d in built])
    batch.close()
    return built, refs, got


def test_stages_before_the_solve_are_bit_exact(orc, fl):
    built, refs, got = _run(orc, fl, force=True)
    for (label, o, d), (imp, ang, st, ost), g in zip(built, refs, got):
        assert g["error_bits"] == 0, label
        _assert_bits(d["x"], ost["fluid_x"], f"{label}: fluid x")
        _assert_bits(d["v"], ost["fluid_v"], f"{label}: fluid v")
        _assert_bits(g["impulse"], imp, f"{label}: ball impulse")
        _assert_bits(g["angular"], ang, f"{label}: ball angular impulse")
        if ost["has_guess"]:
            assert d["has_guess"], label
            _assert_bits(d["guess"], ost["guess"], f"{label}: warm start")
        assert g["viscosity_iterations"] == 0
        _compare_sums(label, g, st)


def test_step_with_the_solve_matches(orc, fl):
    built, refs, got = _run(orc, fl, force=False)
    for (label, o, d), (imp, ang, st, ost), g in zip(built, refs, got):
        assert g["error_bits"] == 0, label
        _assert_bits(d["x"], ost["fluid_x"], f"{label}: fluid x")
        assert g["viscosity_iterations"] == int(st["viscosity_iterations"]), \
            f"{label}: CG iterations gpu {g['viscosity_iterations']} oracle {st['viscosity_iterations']}"
        assert g["density_iterations"] == int(st["density_iterations"]), label
        vs = max(np.abs(ost["fluid_v"]).max(), 1e-30)
        dv = np.abs(d["v"] - ost["fluid_v"]).max()
        assert dv <= 1e-5 * vs, f"{label}: fluid v max diff {dv:.3e} of {vs:.3e}"
        if len(imp):
            isc = max(np.abs(imp).max(), 1e-30)
            di = np.abs(g["impulse"] - imp).max()
            assert di <= 1e-5 * isc, f"{label}: impulse max diff {di:.3e} of {isc:.3e}"
            asc = max(np.abs(ang).max(), 1e-30)
            da = np.abs(g["angular"] - ang).max()
            assert da <= 1e-5 * asc, f"{label}: angular impulse max diff {da:.3e} of {asc:.3e}"
        if ost["has_guess"]:
            gsc = max(np.abs(ost["guess"]).max(), 1e-30)
            assert np.abs(d["guess"] - ost["guess"]).max() <= 1e-5 * max(gsc, vs), label
        for k in ("n_groups", "n_links", "links_in_range"):
            assert g[k] == int(st[k]), f"{label}: {k}"
        _close(g["mean_shear_rate_per_s"], st["mean_shear_rate_per_s"], st["mean_shear_rate_per_s"], 1e-4,
               f"{label}: mean shear rate")
