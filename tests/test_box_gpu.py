"""GPU parity of the periodic-box PBF path (HIP kernels via the C ABI) against the oracle.

Bar: neighbour lists, cell-sorted traversal order, positions and densities bit-exact against the
oracle's mixed policy (f32 state, f64 CG dots); velocities after the CG within 1e-6 of max |v|
(the GPU's FP64 dot products are tree-ordered, the oracle's sequential, so an alpha or beta can
differ by one f32 ulp); CG iteration counts equal. At 1M particles: the same comparison against
the oracle, plus conservation of momentum and the CG's stopping rule.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

V_TOL = 1e-6


@pytest.fixture(scope="module")
def box_mod():
    from milldynamics3_amd import box
    return box


def _compare(b, o, label):
    x, v = b.state()
    xo, vo = o.state()
    off, nb = b.neighbors()
    ooff, onb = o.neighbors()
    assert np.array_equal(off, ooff), f"{label}: neighbour offsets differ"
    assert np.array_equal(nb, onb), f"{label}: neighbour entries differ"
    assert np.array_equal(x.astype(np.float64), xo), f"{label}: positions differ (max {np.abs(x - xo).max()})"
    assert np.array_equal(b.density().astype(np.float64), o.density()), f"{label}: densities differ"
    vmax = max(np.abs(vo).max(), 1e-30)
    assert np.abs(v - vo).max() / vmax <= V_TOL, f"{label}: velocities differ"
    assert b.stats()["cg_iterations"] == o.stats()["cg_iterations"], f"{label}: CG iterations differ"


def test_initial_state_matches_oracle_seed(box_mod, orc):
    cfg = {"side": 64}
    b = box_mod.Box(cfg)
    x, v = b.state()
    xs, vs = orc.c3_seed_state(cfg)
    assert np.array_equal(x, xs) and np.array_equal(v, vs)
    g = box_mod.box_geometry(cfg)
    og = orc.c3_box(cfg)
    assert g == og


# Up to side 64 every wave of 64 particles touches a periodic seam, so every pair loop takes the
# minimum image; from side 128 on most waves lie 2h inside the box and take plain differences.
@pytest.mark.parametrize("side,mu,steps", [(64, 1.0e-3, 3), (40, 1.0, 2), (48, 0.0, 2), (200, 1.0e-3, 2),
                                             (8, 1.0e-3, 3), (10, 1.0, 2), (128, 1.0, 2)])
def test_steps_match_oracle(box_mod, orc, side, mu, steps):
    cfg = {"side": side, "mu": mu}
    print(f"box parity side={side} mu={mu}", flush=True)
    b = box_mod.Box(cfg)
    o = orc.C3Oracle(cfg, "mixed")
    for k in range(steps):
        b.step()
        o.step()
        _compare(b, o, f"side={side} mu={mu} step={k}")
    if mu == 0.0:
        assert b.stats()["cg_iterations"] == 0


# Density pass counts other than 4: with no pass the pair table is built by its own kernel (no pass A
# to fuse with), with one pass the table-building pass A is the only one.
@pytest.mark.parametrize("passes,side", [(0, 64), (1, 64), (0, 136), (2, 136)])
def test_density_pass_counts_match_oracle(box_mod, orc, passes, side):
    cfg = {"side": side, "mu": 1.0e-3, "passes": passes}
    print(f"box parity passes={passes} side={side}", flush=True)
    b = box_mod.Box(cfg)
    o = orc.C3Oracle(cfg, "mixed")
    for k in range(2):
        b.step()
        o.step()
        _compare(b, o, f"passes={passes} side={side} step={k}")


def test_deterministic_across_runs(box_mod):
    cfg = {"side": 96, "mu": 1.0}
    a = box_mod.Box(cfg)
    c = box_mod.Box(cfg)
    a.step(2)
    c.step(2)
    xa, va = a.state()
    xc, vc = c.state()
    assert np.array_equal(xa, xc) and np.array_equal(va, vc)


def test_full_size_million_particles(box_mod, orc):
    cfg = {"side": 1024}
    print("box full size 1024^2", flush=True)
    b = box_mod.Box(cfg)
    x0, v0 = b.state()
    b.step()
    o = orc.C3Oracle(cfg, "mixed")
    o.step()
    _compare(b, o, "side=1024")
    x1, v1 = b.state()
    st = b.stats()
    assert st["residual"] <= 1e-6 * st["b_norm"] * (1 + 1e-9)
    # momentum: f32 positions quantise v = (x - x0)/dt at ulp(L)/dt per particle
    ulp = float(np.spacing(np.float32(1.024))) * 1440.0
    assert np.abs(v1.sum(0) - v0.sum(0)).max() < 8 * ulp * np.sqrt(b.n)


def test_four_million_particles(box_mod, orc):
    # the 2048^2 box of the throughput runs: 1,048,577 cell starts (1025 scan chunks) and the
    # interior fast path on ~96 % of waves; one step bit-exact against the oracle
    cfg = {"side": 2048}
    print("box full size 2048^2", flush=True)
    b = box_mod.Box(cfg)
    b.step()
    o = orc.C3Oracle(cfg, "mixed")
    o.step()
    _compare(b, o, "side=2048")


def test_full_size_windows_match_oracle(box_mod, orc):
    # five steps in one call at 1M particles (one window, as the throughput runs time them)
    cfg = {"side": 1024}
    print("box full size 1024^2, 5 steps in one call", flush=True)
    b = box_mod.Box(cfg)
    b.step(5)
    o = orc.C3Oracle(cfg, "mixed")
    o.step(5)
    x, v = b.state()
    xo, vo = o.state()
    assert np.array_equal(x.astype(np.float64), xo)
    assert np.array_equal(b.density().astype(np.float64), o.density())
    assert np.abs(v - vo).max() / max(np.abs(vo).max(), 1e-30) <= V_TOL
    assert b.stats()["cg_iterations"] == o.stats()["cg_iterations"]


def test_capped_cg_matches_oracle(box_mod, orc):
    # the CG stops at max_cg without converging; both sides must stop at the same iterate
    cfg = {"side": 48, "mu": 1.0, "max_cg": 3}
    b = box_mod.Box(cfg)
    o = orc.C3Oracle(cfg, "mixed")
    for k in range(2):
        b.step()
        o.step()
        _compare(b, o, f"max_cg=3 step={k}")
    assert b.stats()["cg_iterations"] == 3


@pytest.mark.parametrize("side,mu,steps", [(64, 1.0e-3, 20), (64, 1.0, 5), (48, 0.0, 17), (24, 3.0, 37)])
def test_multi_step_windows_match_oracle(box_mod, orc, side, mu, steps):
    # K steps in one call run as windows with checkpoints; a step whose CG outruns the predicted
    # iteration count stops the rest of its window and is re-run from its checkpoint, and the next
    # window is shorter (mu = 1 forces that on the first window; at mu = 3 the counts vary by one
    # from step to step over 37 steps)
    cfg = {"side": side, "mu": mu}
    b = box_mod.Box(cfg)
    o = orc.C3Oracle(cfg, "mixed")
    b.step(steps)
    o.step(steps)
    x, v = b.state()
    xo, vo = o.state()
    off, nb = b.neighbors()
    ooff, onb = o.neighbors()
    assert np.array_equal(off, ooff) and np.array_equal(nb, onb)
    assert np.array_equal(x.astype(np.float64), xo)
    assert np.array_equal(b.density().astype(np.float64), o.density())
    assert np.abs(v - vo).max() / max(np.abs(vo).max(), 1e-30) <= V_TOL
    assert b.stats()["cg_iterations"] == o.stats()["cg_iterations"]


def test_window_rerun_after_low_prediction_matches_oracle(box_mod, orc):
    # at mu = 3 the CG needs 31, 32, 31, 32, 32, ... iterations; after two steps a prediction of 31
    # makes the second step of the next window overrun, so the rest of the window is stopped and the
    # overrun step is re-run from its checkpoint: results must not change
    cfg = {"side": 24, "mu": 3.0}
    b = box_mod.Box(cfg)
    o = orc.C3Oracle(cfg, "mixed")
    b.step(2)
    o.step(2)
    b.set_cg_prediction(31)
    b.step(6)
    o.step(6)
    x, v = b.state()
    xo, vo = o.state()
    off, nb = b.neighbors()
    ooff, onb = o.neighbors()
    assert np.array_equal(off, ooff) and np.array_equal(nb, onb)
    assert np.array_equal(x.astype(np.float64), xo)
    assert np.abs(v - vo).max() / max(np.abs(vo).max(), 1e-30) <= V_TOL
    assert b.stats()["cg_iterations"] == o.stats()["cg_iterations"]


# ---- FP64 state (fp64 = 1) against the oracle's f64 policy -------------------------------------
# Bar: the initial state, neighbour lists and, within a step, positions and densities equal the
# oracle's f64 policy; the CG's FP64 dot products are summed per workgroup in a fixed tree while the
# oracle sums sequentially, and in f64 the CG scalars are not rounded to f32, so velocities (and,
# from the next step on, positions) agree to within X64_TOL / V64_TOL relative, not bit for bit.
X64_TOL = 1e-12
V64_TOL = 1e-10


def _compare64(b, o, label, xtol=X64_TOL):
    x, v = b.state()
    xo, vo = o.state()
    assert x.dtype == np.float64 and v.dtype == np.float64
    off, nb = b.neighbors()
    ooff, onb = o.neighbors()
    assert np.array_equal(off, ooff) and np.array_equal(nb, onb), f"{label}: neighbour lists differ"
    L = o.cfg["side"] * o.cfg["dx"]
    assert np.abs(x - xo).max() <= xtol * L, f"{label}: positions differ (max {np.abs(x - xo).max()})"
    d, do = b.density(), o.density()
    assert np.abs(d - do).max() <= xtol * do.max(), f"{label}: densities differ"
    vmax = max(np.abs(vo).max(), 1e-30)
    assert np.abs(v - vo).max() / vmax <= V64_TOL, f"{label}: velocities differ"
    assert b.stats()["cg_iterations"] == o.stats()["cg_iterations"], f"{label}: CG iterations differ"


@pytest.mark.parametrize("side,mu", [(32, 1.0e-3), (24, 1.0), (40, 0.0)])
def test_fp64_steps_match_oracle(box_mod, orc, side, mu):
    cfg = {"side": side, "mu": mu, "fp64": 1}
    print(f"box fp64 side={side} mu={mu}", flush=True)
    b = box_mod.Box(cfg)
    o = orc.C3Oracle(cfg, "fp64")
    x, v = b.state()
    xo, vo = o.state()
    assert np.array_equal(x, xo) and np.array_equal(v, vo)  # same f32 seed, promoted
    for k in range(3):
        b.step()
        o.step()
        _compare64(b, o, f"fp64 side={side} mu={mu} step={k}")
    if mu == 1.0e-3:  # first step: everything before the CG is bit-exact
        b2 = box_mod.Box(cfg)
        o2 = orc.C3Oracle(cfg, "fp64")
        b2.step()
        o2.step()
        assert np.array_equal(b2.state()[0], o2.state()[0])
        assert np.array_equal(b2.density(), o2.density())


def test_fp64_interior_path_matches_oracle(box_mod, orc):
    # side 128: most waves take the plain-difference path; mu = 1 runs ~20 CG iterations per step
    cfg = {"side": 128, "mu": 1.0, "fp64": 1}
    b = box_mod.Box(cfg)
    o = orc.C3Oracle(cfg, "fp64")
    for k in range(2):
        b.step()
        o.step()
        _compare64(b, o, f"fp64 side=128 mu=1 step={k}")


def test_fp64_multi_step_windows_match_oracle(box_mod, orc):
    cfg = {"side": 24, "mu": 3.0, "fp64": 1}
    b = box_mod.Box(cfg)
    o = orc.C3Oracle(cfg, "fp64")
    b.step(19)
    o.step(19)
    _compare64(b, o, "fp64 windows")


def test_fp64_full_size_million_particles(box_mod, orc):
    cfg = {"side": 1024, "fp64": 1}
    print("box fp64 full size 1024^2", flush=True)
    b = box_mod.Box(cfg)
    b.step()
    o = orc.C3Oracle(cfg, "fp64")
    o.step()
    _compare64(b, o, "fp64 side=1024")


def test_pair_arith_matches_ieee():
    # the f32 and f64 square root and quotient of the pair loops (core sequences without range
    # scaling, a shared reciprocal) against sqrt and '/' on 2^28 random operands of their range
    import ctypes as C
    from milldynamics3_amd import sim
    L = sim.lib()
    L.mg_box_check_arith.argtypes = [C.c_uint64, C.c_uint64, C.POINTER(C.c_uint64)]
    out = (C.c_uint64 * 4)()
    assert L.mg_box_check_arith(1 << 28, 20261203, out) == 0
    assert list(out) == [0, 0, 0, 0], f"mismatches (f32 sqrt, f32 quotient, f64 sqrt, f64 quotient): {list(out)}"


@pytest.mark.parametrize("fp64,yield_stress", [(0, 0.0), (1, 0.0), (0, 1.0), (1, 1.0)])
def test_dense_cluster_rescans_match_oracle(box_mod, orc, fp64, yield_stress):
    # 300 particles packed at 0.35 dx: up to ~114 pairs each, so the pair table overflows (more than
    # 32 pairs) and every kernel takes its rescan path for them, the CG's operator included (with a
    # yield stress: its effective viscosities too)
    side, dx = 32, 1.0e-3
    cfg = {"side": side, "mu": 1.0, "fp64": fp64, "yield_stress": yield_stress}
    print(f"box dense cluster fp64={fp64} yield_stress={yield_stress}", flush=True)
    b = box_mod.Box(cfg)
    o = orc.C3Oracle(cfg, "fp64" if fp64 else "mixed")
    x, v = o.state()
    k = np.arange(300)
    x[:300, 0] = side * dx / 2 + (k % 17 - 8) * 0.35 * dx
    x[:300, 1] = side * dx / 2 + (k // 17 - 9) * 0.35 * dx
    x = x.astype(np.float32).astype(np.float64)
    v = v.astype(np.float32).astype(np.float64)
    b.set_state(x, v)
    o.set_state(x, v)
    b.step()
    o.step()
    assert np.diff(o.neighbors()[0]).max() > 32
    if fp64:
        _compare64(b, o, "dense cluster fp64")
    else:
        _compare(b, o, "dense cluster")
    b.step()
    o.step()
    if fp64:  # the blob bursts (|v| ~ 10 m/s, 1000x the lattice): reduction-order differences in the CG move x more
        _compare64(b, o, "dense cluster fp64 step 2", xtol=1e-10)
    else:
        _compare(b, o, "dense cluster step 2")


def test_replay_timing_leaves_state_unchanged(box_mod, orc):
    """mg_box_replay_ms relaunches the last density passes A and B: the state, the densities and
    the next step are those of a box that never replayed (and both equal the oracle's)."""
    cfg = {"side": 128, "mu": 1.0e-3}
    print("box replay timing side=128", flush=True)
    b = box_mod.Box(cfg)
    o = orc.C3Oracle(cfg, "mixed")
    b.step(2)
    o.step()
    o.step()
    x0, v0 = b.state()
    assert b.replay_ms(1, 5) > 0.0
    assert b.replay_ms(2, 5) > 0.0
    x1, v1 = b.state()
    assert np.array_equal(x0, x1) and np.array_equal(v0, v1)
    with pytest.raises(RuntimeError):
        b.replay_ms(5, 1)  # CG classes are not relaunchable
    b.step()
    o.step()
    _compare(b, o, "after replay")


# Bingham boxes (pbf.rs:821-868, :2147-2151): shear rates and effective viscosities on the GPU, in
# the table's pair order. The f32 exp of the regularisation is the device's expf (the oracle's is
# the host's), which may differ by an ulp; the effective viscosities only enter the CG, whose
# velocities are held to the same 1e-6 bar.
@pytest.mark.parametrize("side,mu,yield_stress,steps", [(64, 1.0e-3, 1.0, 3), (128, 1.0e-3, 1.0, 2),
                                                        (40, 1.0e-2, 0.5, 2)])
def test_bingham_steps_match_oracle(box_mod, orc, side, mu, yield_stress, steps):
    cfg = {"side": side, "mu": mu, "yield_stress": yield_stress}
    print(f"box bingham side={side} mu={mu} yield_stress={yield_stress}", flush=True)
    b = box_mod.Box(cfg)
    o = orc.C3Oracle(cfg, "mixed")
    newt = orc.C3Oracle({"side": side, "mu": mu}, "mixed")
    for k in range(steps):
        b.step()
        o.step()
        newt.step()
        _compare(b, o, f"bingham side={side} step={k}")
    assert o.stats()["cg_iterations"] > newt.stats()["cg_iterations"]  # the yield stress is active


def test_fp64_bingham_matches_oracle(box_mod, orc):
    cfg = {"side": 32, "mu": 1.0e-3, "yield_stress": 1.0, "fp64": 1}
    print("box fp64 bingham side=32", flush=True)
    b = box_mod.Box(cfg)
    o = orc.C3Oracle(cfg, "fp64")
    for k in range(3):
        b.step()
        o.step()
        _compare64(b, o, f"fp64 bingham step={k}")


# Mean shear rate of the last step (pbf.rs:800-818): per-particle rates in the reference's pair order,
# FP64 sums. Without viscosity the velocities are bit-exact, so the f32 means agree to one ulp (the
# sums differ only in order); after the CG the velocities carry the CG's 1e-6 bar.
@pytest.mark.parametrize("cfg", [{"side": 48, "mu": 0.0}, {"side": 64, "mu": 1.0e-3},
                                 {"side": 128, "mu": 1.0e-3, "yield_stress": 1.0},
                                 {"side": 32, "mu": 1.0e-3, "fp64": 1}])
def test_mean_shear_rate_matches_oracle(box_mod, orc, cfg):
    print(f"box mean shear rate {cfg}", flush=True)
    b = box_mod.Box(cfg)
    o = orc.C3Oracle(cfg, "fp64" if cfg.get("fp64") else "mixed")
    assert b.mean_shear_rate() == 0.0
    for _ in range(2):
        b.step()
        o.step()
    m, mo = b.mean_shear_rate(), o.mean_shear_rate()
    assert mo > 0.0
    if cfg.get("fp64"):
        assert abs(m - mo) <= 1e-9 * mo
    elif cfg["mu"] == 0.0:
        assert abs(m - mo) <= float(np.spacing(np.float32(mo)))
    else:
        assert abs(m - mo) <= 1e-5 * mo
    x0, v0 = b.state()
    assert b.mean_shear_rate() == m and np.array_equal(b.state()[0], x0)  # a read: no state change


# Bingham boxes through multi-step windows (the path bench.py times for its Bingham line): the
# shear-rate / effective-viscosity kernel and the Bingham CG set-up inside windows, and a window
# re-run forced by a low iteration prediction (a Bingham step at water viscosity needs ~6 CG
# iterations, so a prediction of 1 makes the window's first step overrun). Windows must give the
# results of running every step alone on the GPU, bit for bit. Against the oracle the device's expf
# (one ulp from the host's at most) perturbs mu_eff, and over 16 steps the PBF flow grows that to a
# few f32 ulps of position and ~2e-4 of max |v| (measured on MI355X), so the oracle bar here is
# 1e-3 of a spacing and 1e-3 of max |v| and of the mean shear rate.
@pytest.mark.parametrize("fp64", [0, 1])
def test_bingham_windows_match_single_steps(box_mod, orc, fp64):
    cfg = {"side": 64, "mu": 1.0e-3, "yield_stress": 1.0, "fp64": fp64}
    print(f"box bingham windows fp64={fp64}", flush=True)
    b = box_mod.Box(cfg)
    s = box_mod.Box(cfg)
    o = orc.C3Oracle(cfg, "fp64" if fp64 else "mixed")
    b.step(2)
    b.set_cg_prediction(1)
    b.step(14)
    for _ in range(16):
        s.step()
    o.step(16)
    x, v = b.state()
    xs, vs = s.state()
    assert np.array_equal(x, xs) and np.array_equal(v, vs)
    assert b.stats()["cg_iterations"] == s.stats()["cg_iterations"]
    assert b.mean_shear_rate() == s.mean_shear_rate()
    xo, vo = o.state()
    L = box_mod.box_geometry(cfg)["L"]
    dx = L / cfg["side"]
    d = x.astype(np.float64) - xo
    d -= L * np.round(d / L)
    assert np.abs(d).max() <= 1e-3 * dx
    assert np.abs(v - vo).max() / max(np.abs(vo).max(), 1e-30) <= 1e-3
    m, mo = b.mean_shear_rate(), o.mean_shear_rate()
    assert mo > 0.0 and abs(m - mo) <= 1e-3 * mo


# ---- options the drum step has: adaptive density passes with the reference's CG constants,
# gravity in the predict, the step-0 reset of non-finite particles, the speed clamp and the f64
# kinetic energies (pbf.rs:924-926, :987-996, :1200-1235, :1261-1333) ---------------------------
OPTION_CASES = [
    {"side": 64, "passes": 3, "adaptive": 1, "density_ratio": 1.005, "tol": 1.0e-3, "max_cg": 50},
    {"side": 128, "passes": 1, "adaptive": 1, "density_ratio": 1.005, "mu": 1.0, "tol": 1.0e-3, "max_cg": 50},
    {"side": 64, "passes": 3, "adaptive": 1},
    {"side": 48, "passes": 3, "adaptive": 1, "density_ratio": 0.9},
    {"side": 10, "passes": 1, "adaptive": 1, "density_ratio": 1.005},
    {"side": 64, "gravity": -9.81},
    {"side": 64, "mu": 1.0, "speed_clamp": 0.015, "energy_stats": 1},
    {"side": 128, "tol": 1.0e-3, "max_cg": 50, "energy_stats": 1},
]


@pytest.mark.parametrize("cfg", OPTION_CASES)
def test_step_options_match_oracle(box_mod, orc, cfg):
    print(f"box options {cfg}", flush=True)
    b = box_mod.Box(cfg)
    o = orc.C3Oracle(cfg, "mixed")
    for k in range(3):
        b.step()
        o.step()
        _compare(b, o, f"{cfg} step={k}")
        so = o.step_stats()
        assert b.density_passes() == so["density_passes"]
        if cfg.get("energy_stats"):
            e = b.energy()
            assert e["ke_before_cg"] == pytest.approx(so["ke_before_cg"], rel=1e-12)
            assert e["ke_after_cg"] == pytest.approx(so["ke_after_cg"], rel=1e-5)
            assert e["ke_after_clamp"] == pytest.approx(so["ke_after_clamp"], rel=1e-5)


@pytest.mark.parametrize("cfg", [{"side": 64, "passes": 3, "adaptive": 1, "density_ratio": 1.005, "mu": 1.0},
                                 {"side": 48, "mu": 1.0e-3, "gravity": -9.81, "speed_clamp": 0.02}])
def test_step_options_through_windows(box_mod, orc, cfg):
    b = box_mod.Box(cfg)
    o = orc.C3Oracle(cfg, "mixed")
    b.step(12)
    o.step(12)
    x, v = b.state()
    xo, vo = o.state()
    assert np.array_equal(x.astype(np.float64), xo)
    assert np.abs(v - vo).max() / max(np.abs(vo).max(), 1e-30) <= V_TOL
    assert b.density_passes() == o.step_stats()["density_passes"]


def test_non_finite_particles_restart_at_the_origin(box_mod, orc):
    cfg = {"side": 64, "mu": 1.0e-3}
    b = box_mod.Box(cfg)
    o = orc.C3Oracle(cfg, "mixed")
    x, v = o.state()
    x[5] = (np.nan, 0.01)
    v[77] = (np.inf, 0.0)
    x[1000] = (-np.inf, np.nan)
    b.set_state(x.astype(np.float32), v.astype(np.float32))
    o.set_state(x, v)
    for k in range(2):
        b.step()
        o.step()
        _compare(b, o, f"non-finite step={k}")
    assert np.isfinite(b.state()[0]).all()


def test_fp64_adaptive_passes_match_oracle(box_mod, orc):
    cfg = {"side": 32, "fp64": 1, "adaptive": 1, "passes": 3, "density_ratio": 1.005, "mu": 1.0}
    b = box_mod.Box(cfg)
    o = orc.C3Oracle(cfg, "fp64")
    for k in range(3):
        b.step()
        o.step()
        _compare64(b, o, f"fp64 adaptive step={k}")
        assert b.density_passes() == o.step_stats()["density_passes"]


def test_million_particles_adaptive_reference_constants(box_mod, orc):
    cfg = {"side": 1024, "passes": 3, "adaptive": 1, "density_ratio": 1.005, "tol": 1.0e-3, "max_cg": 50}
    print("box 1M adaptive", flush=True)
    b = box_mod.Box(cfg)
    o = orc.C3Oracle(cfg, "mixed")
    b.step()
    o.step()
    _compare(b, o, "1M adaptive")
    assert b.density_passes() == o.step_stats()["density_passes"]
