"""Pins of the periodic-box oracle (oracle/c3_oracle.cpp) without a GPU.

The periodic box has no reference trace, so its oracle is pinned by what fixes it independently:
the reference grid's own neighbour order on inputs that never wrap, brute force on small inputs,
the lattice closed form of the mass, conservation laws of the density and viscosity operators, the
CG's own stopping rule, and the FP64 policy.
"""
import numpy as np
import pytest


def _brute_neighbors(pts, L, h):
    n = len(pts)
    d = pts[:, None, :].astype(np.float32) - pts[None, :, :].astype(np.float32)
    half = np.float32(0.5) * np.float32(L)
    d = np.where(d > half, d - np.float32(L), np.where(d < -half, d + np.float32(L), d)).astype(np.float32)
    r2 = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]).astype(np.float32)
    m = r2 <= np.float32(h) * np.float32(h)
    np.fill_diagonal(m, False)
    return [set(np.nonzero(m[i])[0].tolist()) for i in range(n)]


def test_neighbors_match_brute_force_with_wrap(orc):
    cfg = {"side": 24}
    box = orc.c3_box(cfg)
    rng = np.random.default_rng(5)
    pts = (rng.random((600, 2)) * box["L"]).astype(np.float32)
    pts[:40] = np.float32(box["L"]) * np.float32(0.999)  # pile some points at the wrap seam
    pts[40:80, 0] = np.float32(0.0)
    off, nb = orc.c3_neighbors_of(cfg, pts)
    brute = _brute_neighbors(pts, box["L"], box["h"])
    for i in range(len(pts)):
        row = nb[off[i]:off[i + 1]].tolist()
        assert len(row) == len(set(row))
        assert set(row) == brute[i]


def test_neighbor_order_is_reference_order_without_wrap(orc):
    # h = 0.25 and L = 8 are exact, so the box cells are the reference grid's cells floor(x/h);
    # points far from the seam never wrap, and the reference emission order (grid.rs:291-362 via
    # pbf.rs:262-288, restated by orc_neighbor_lists) must come out entry for entry.
    cfg = {"side": 64, "dx": 0.125}
    box = orc.c3_box(cfg)
    assert box["h"] == 0.25 and box["L"] == 8.0 and box["cw"] == 0.25
    rng = np.random.default_rng(11)
    pts = (2.0 + rng.random((900, 2)) * 4.0).astype(np.float32)
    off, nb = orc.c3_neighbors_of(cfg, pts)
    roff, rnb = orc.neighbor_lists(pts, box["h"])
    assert np.array_equal(off, roff)
    assert np.array_equal(nb, rnb)


def test_mass_from_lattice_closed_form(orc):
    box = orc.c3_box({"side": 32})
    dx = 1.0e-3
    h = 2 * dx
    # Poly6 4/(pi h^8) (h^2 - r^2)^3 over the 9 lattice sites with r < h (r = 0, dx, sqrt(2) dx)
    w = 4.0 / (np.pi * h ** 8)
    s = w * (h ** 6 + 4 * (h * h - dx * dx) ** 3 + 4 * (h * h - 2 * dx * dx) ** 3)
    assert box["mass"] == pytest.approx(1.02 * 1000.0 / s, rel=1e-5)
    assert box["ncx"] == 16


def test_seed_lattice_jitter_and_velocity_statistics(orc):
    cfg = {"side": 128}
    x, v = orc.c3_seed_state(cfg)
    x2, v2 = orc.c3_seed_state(cfg)
    assert np.array_equal(x, x2) and np.array_equal(v, v2)
    dx = 1.0e-3
    ix = np.tile(np.arange(128), 128)
    iy = np.repeat(np.arange(128), 128)
    jx = x[:, 0] - (ix + 0.5) * dx
    jy = x[:, 1] - (iy + 0.5) * dx
    assert np.abs(jx).max() <= 0.1 * dx * (1 + 1e-4)
    assert np.abs(jy).max() <= 0.1 * dx * (1 + 1e-4)
    assert abs(v.mean()) < 3e-4
    assert v.std() == pytest.approx(0.01, rel=0.03)


def test_unjittered_density_is_two_percent_over_rest(orc):
    # The first pass's rho on the exact lattice is 1.02 rho0 for every particle; with zero
    # velocity and jitter-free positions (set_state), the first lambda sees C = 0.02 everywhere,
    # so one pass moves nothing (every dp cancels by symmetry).
    cfg = {"side": 32, "passes": 1, "mu": 0.0}
    s = orc.C3Oracle(cfg, "fp64")
    dx = float(np.float32(1.0e-3))  # the box is built from the f32 spacing
    ix = np.tile(np.arange(32), 32)
    iy = np.repeat(np.arange(32), 32)
    x = np.stack([(ix + 0.5) * dx, (iy + 0.5) * dx], 1)
    s.set_state(x, np.zeros_like(x))
    s.step()
    xn, vn = s.state()
    assert np.abs(xn - x).max() < 1e-15
    assert np.allclose(s.density(), 1020.0, rtol=1e-6)


@pytest.mark.parametrize("mu", [1.0e-3, 1.0])
def test_step_conserves_momentum_and_cg_meets_tolerance(orc, mu):
    cfg = {"side": 48, "mu": mu}
    s = orc.C3Oracle(cfg, "fp64")
    x0, v0 = s.state()
    s.step()
    x1, v1 = s.state()
    st = s.stats()
    # density passes and the viscous operator are antisymmetric in (i, j): total momentum stays
    assert np.abs(v1.sum(0) - v0.sum(0)).max() < 1e-9 * np.abs(v0).sum()
    assert st["residual"] <= 1e-6 * st["b_norm"] * (1 + 1e-12)
    assert st["cg_iterations"] >= 1
    if mu >= 1.0:
        assert st["cg_iterations"] > 5


def test_viscosity_dissipates_energy(orc):
    cfg = {"side": 48, "passes": 4}
    a = orc.C3Oracle(dict(cfg, mu=0.0), "fp64")
    b = orc.C3Oracle(dict(cfg, mu=1.0), "fp64")
    a.step()
    b.step()
    _, va = a.state()
    _, vb = b.state()
    assert (vb ** 2).sum() < (va ** 2).sum()


def test_mixed_policy_tracks_fp64(orc):
    cfg = {"side": 48, "mu": 1.0}
    m = orc.C3Oracle(cfg, "mixed")
    d = orc.C3Oracle(cfg, "fp64")
    m.step()
    d.step()
    xm, vm = m.state()
    xd, vd = d.state()
    L = orc.c3_box(cfg)["L"]
    assert np.abs(xm - xd).max() / L < 1e-6
    assert np.abs(vm - vd).max() / np.abs(vd).max() < 1e-3


# ---- Bingham rheology (pbf.rs:821-868, :2147-2151) ----------------------------------------------
def _lattice(side, dx=1.0e-3):
    ix, iy = np.meshgrid(np.arange(side), np.arange(side))
    return np.stack([(ix.ravel() + 0.5) * dx, (iy.ravel() + 0.5) * dx], 1).astype(np.float32)


def _interior(pts, c, dx, margin=3):
    return (np.abs(pts[:, 0] - c) < c - margin * dx) & (np.abs(pts[:, 1] - c) < c - margin * dx)


def test_shear_rate_of_linear_fields_on_a_lattice(orc):
    """Constant velocity: exactly 0. Rigid rotation: 0 up to rounding (D is the symmetric part).
    Simple shear v = (g y, 0): rate = g S with the lattice sum S = sum_j V (dy_j)^2 c(r_j) / r_j,
    c = 30 / (pi h^5) (h - r)^2 (the spiky gradient of a linear field), evaluated here by brute
    force in f64; the same along x; doubling the velocities doubles the rates exactly; doubling the
    densities halves them."""
    side, dx = 32, 1.0e-3
    cfg = {"side": side, "dx": dx}
    box = orc.c3_box(cfg)
    h, m = box["h"], box["mass"]
    pts = _lattice(side, dx)
    n = len(pts)
    c = side * dx / 2
    inner = _interior(pts, c, dx)
    rho = np.full(n, 1020.0, np.float32)
    const = np.tile(np.array([[0.3, -0.2]], np.float32), (n, 1))
    assert not orc.c3_shear_rates_of(cfg, pts, const, rho, m).any()
    w = 5.0
    rot = np.stack([-w * (pts[:, 1] - c), w * (pts[:, 0] - c)], 1)
    assert orc.c3_shear_rates_of(cfg, pts, rot, rho, m)[inner].max() / w < 1e-5
    # brute-force lattice sum over the offsets inside the kernel radius
    S = 0.0
    for a in range(-3, 4):
        for b in range(-3, 4):
            r = dx * np.hypot(a, b)
            if 0 < r < h:
                S += (m / 1020.0) * (b * dx) ** 2 * 30.0 / (np.pi * h ** 5) * (h - r) ** 2 / r
    g = 10.0
    shear = np.stack([g * (pts[:, 1] - c), np.zeros(n)], 1)
    r1 = orc.c3_shear_rates_of(cfg, pts, shear, rho, m)
    assert np.allclose(r1[inner] / g, S, rtol=1e-5, atol=0)
    shear_x = np.stack([np.zeros(n), g * (pts[:, 0] - c)], 1)
    assert np.allclose(orc.c3_shear_rates_of(cfg, pts, shear_x, rho, m)[inner] / g, S, rtol=1e-5, atol=0)
    r2 = orc.c3_shear_rates_of(cfg, pts, 2.0 * shear.astype(np.float32), rho, m)
    assert np.array_equal(r2, 2.0 * r1)
    r3 = orc.c3_shear_rates_of(cfg, pts, shear, 2.0 * rho, m)
    assert np.allclose(r3, 0.5 * r1, rtol=1e-6, atol=0)


def test_bingham_viscosity_limits(orc):
    """g -> 0 (and non-finite g): the cap (1 + 100) mu; g >> g0: mu + tau_y / g; continuous across
    the 1e-4 g0 switch to its limit (to f32 cancellation); decreasing in g."""
    mu, tau = 1.0e-3, 2.0
    g0 = tau / (100.0 * mu)
    out = orc.c3_bingham_of(np.array([0.0, np.nan, np.inf], np.float32), mu, tau)
    assert np.allclose(out, 101.0 * mu, rtol=1e-6, atol=0)
    big = np.array([1e3 * g0, 1e5 * g0], np.float32)
    assert np.allclose(orc.c3_bingham_of(big, mu, tau), mu + tau / big.astype(np.float64), rtol=1e-6, atol=0)
    lo, hi = orc.c3_bingham_of(np.array([0.999e-4 * g0, 1.001e-4 * g0], np.float32), mu, tau)
    # 1 - exp(-x) at x = 1e-4 cancels in f32 (ulp(1) / x = 6e-4 relative), as in the reference
    assert abs(hi - lo) / lo < 1e-3
    g = np.logspace(-1, 4, 200).astype(np.float32) * np.float32(g0)  # above the f32 cancellation range
    vals = orc.c3_bingham_of(g, mu, tau).astype(np.float64)
    assert np.all(np.diff(vals) <= 1e-9 * vals[:-1])


def test_bingham_step_is_more_viscous_and_tends_to_newtonian(orc):
    """A yield stress only raises the effective viscosity: more CG iterations and less kinetic
    energy than the Newtonian step; a vanishing yield stress gives the Newtonian step back."""
    base = {"side": 24, "mu": 1.0e-3}
    newt = orc.C3Oracle(base, "mixed")
    bing = orc.C3Oracle(dict(base, yield_stress=1.0), "mixed")
    tiny = orc.C3Oracle(dict(base, yield_stress=1.0e-9), "mixed")
    for s in (newt, bing, tiny):
        s.step()
    _, vn = newt.state()
    _, vb = bing.state()
    _, vt = tiny.state()
    assert bing.stats()["cg_iterations"] > newt.stats()["cg_iterations"]
    assert (vb * vb).sum() < (vn * vn).sum()
    assert np.abs(vt - vn).max() <= 1e-5 * np.abs(vn).max()


def test_mean_shear_rate_is_the_mean_of_the_rates(orc):
    """The oracle's per-step mean shear rate is the f32 rounding of the f64 mean of the shear rates
    that the rate function (pinned above) gives for the step's final positions, velocities and
    densities over the step's own lists (pbf.rs:2303-2304)."""
    cfg = {"side": 24, "mu": 1.0e-3}
    o = orc.C3Oracle(cfg, "mixed")
    assert o.mean_shear_rate() == 0.0
    for _ in range(2):
        o.step()
    x, v = o.state()
    box = orc.c3_box(cfg)
    rates = orc.c3_shear_rates_of(cfg, x, v, o.density(), box["mass"], lists=o.neighbors())
    assert np.all(np.isfinite(rates))
    m = o.mean_shear_rate()
    assert m > 0.0
    assert m == float(np.float32(np.sum(rates.astype(np.float64)) / rates.size))


# ---- options the drum step has: adaptive passes, gravity, non-finite reset, speed clamp, energies --
def test_adaptive_passes_follow_the_reference_rule(orc):
    # The count lies in [base, base + 6] (pbf.rs:1261-1333), and the state equals a run with exactly
    # that many passes (the adaptive loop only decides where to stop).
    for side, base in ((48, 3), (64, 1)):
        a = orc.C3Oracle({"side": side, "passes": base, "adaptive": 1, "mu": 1.0e-3}, "fp64")
        a.step()
        k = a.step_stats()["density_passes"]
        assert base <= k <= base + 6
        e = orc.C3Oracle({"side": side, "passes": k, "mu": 1.0e-3}, "fp64")
        e.step()
        assert np.array_equal(a.state()[0], e.state()[0]) and np.array_equal(a.state()[1], e.state()[1])


def test_adaptive_passes_stop_at_the_tolerance_or_the_cap(orc):
    # The workload's lattice is 2 % over-dense on average, so some constraint always exceeds 0.02
    # and the count reaches the cap, base + 6; an under-dense lattice (ratio 0.9) stops right at the
    # base; a ratio of 1.005 stops in between, and in fewer passes as the flow relaxes.
    def counts(ratio, base, steps):
        s = orc.C3Oracle({"side": 48, "passes": base, "adaptive": 1, "mu": 1.0e-3, "density_ratio": ratio}, "fp64")
        out = []
        for _ in range(steps):
            s.step()
            out.append(s.step_stats()["density_passes"])
        return out
    assert counts(1.02, 3, 2) == [9, 9]
    assert counts(0.9, 3, 2) == [3, 3]
    mid = counts(1.005, 1, 3)
    assert all(1 < k < 7 for k in mid) and mid[-1] < mid[0]


def test_gravity_adds_g_dt_to_every_velocity(orc):
    g, dt = -9.81, 1.0 / 1440.0
    s = orc.C3Oracle({"side": 32, "passes": 0, "mu": 0.0, "gravity": g}, "fp64")
    _, v0 = s.state()
    s.step()
    _, v1 = s.state()
    dv = v1 - v0
    assert np.allclose(dv[:, 1], np.float32(g) * np.float32(dt), rtol=1e-6, atol=1e-12)
    assert np.abs(dv[:, 0]).max() < 1e-12


def test_non_finite_particles_restart_at_the_origin(orc):
    s = orc.C3Oracle({"side": 32, "passes": 0, "mu": 0.0}, "fp64")
    x, v = s.state()
    x[3] = (np.nan, 0.01)
    v[7] = (np.inf, 0.0)
    x[11] = (-np.inf, np.nan)
    s.set_state(x, v)
    s.step()
    x1, v1 = s.state()
    for i in (3, 7, 11):
        assert np.all(x1[i] == 0.0) and np.all(v1[i] == 0.0)
    assert np.isfinite(x1).all() and np.isfinite(v1).all()


def test_speed_clamp_and_kinetic_energies(orc):
    vmax = 0.015
    cfg = {"side": 32, "mu": 1.0e-3, "speed_clamp": vmax}
    s = orc.C3Oracle(cfg, "fp64")
    u = orc.C3Oracle(dict(cfg, speed_clamp=0.0), "fp64")
    s.step()
    u.step()
    _, vs = s.state()
    _, vu = u.state()
    speed_u = np.linalg.norm(vu, axis=1)
    assert np.linalg.norm(vs, axis=1).max() <= vmax * (1 + 1e-12)
    below = speed_u <= vmax
    assert below.any() and (~below).any()
    assert np.array_equal(vs[below], vu[below])
    assert np.allclose(vs[~below], vu[~below] * (vmax / speed_u[~below])[:, None], rtol=1e-12)
    st = s.step_stats()
    m = orc.c3_box(cfg)["mass"]
    assert st["ke_after_clamp"] == pytest.approx(0.5 * m * (vs ** 2).sum(), rel=1e-12)
    assert st["ke_after_cg"] == pytest.approx(0.5 * m * (vu ** 2).sum(), rel=1e-12)
    assert st["ke_after_clamp"] < st["ke_after_cg"] < st["ke_before_cg"]
