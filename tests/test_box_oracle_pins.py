"""Closed-form and brute-force pins of the periodic-box oracle's density pass, Morris operator and CG.

The box oracle's density pass (pbf.rs:1257-1334), viscous operator (pbf.rs:312-377) and implicit
viscosity CG (pbf.rs:423-480) are checked on hand-placed particles (the box is far larger than the
clusters, so no pair wraps) against:
  * closed forms of symmetric clusters, with the reference's constants written out (Poly6
    4/(pi h^8), spiky -30/(pi h^5), EPSILON_RELAX 200, CENTRAL_VISCOSITY_FACTOR 4, eta^2 = 0.01 h^2);
  * the reference's own pin of the central-force operator (pbf.rs:2428-2473): no force for a rigid
    rotation of a hexagonal patch, dissipation for a shear;
  * an FP64 dense assembly of (I + dt L) over all pairs (no grid, no lists) solved by
    numpy.linalg.solve, with the residual b - A v recomputed outside the oracle.
Each closed-form test also evaluates a plausible mistake (a dropped m/rho0 in grad C, a dropped
|sum g|^2, a factor 2 instead of 4, a missing eta^2, the plain Morris form) and asserts that the
configuration separates it from the oracle's value, so the test fails if the oracle carried it.
"""
import numpy as np
import pytest

CFG = {"side": 256, "dx": 1.0e-3, "rest": 1000.0}
CENTER = np.array([0.128, 0.128])
# the oracle's FP64 policy evaluates the reference's f32 literals as f32 values promoted (0.01 is
# 0.0099999998); the closed forms use the same constant
ETA = float(np.float32(0.01))


def _consts(orc, cfg):
    box = orc.c3_box(cfg)
    h = box["h"]
    return box, h, 4.0 / (np.pi * h ** 8), 30.0 / (np.pi * h ** 5)


def _w(r, h, k6):
    return k6 * (h * h - r * r) ** 3 if r < h else 0.0


def _q(r, h, ks):  # |spiky gradient| at distance r
    return ks * (h - r) ** 2 if 1e-9 < r < h else 0.0


def test_density_pass_closed_form_star(orc):
    """Square of four particles around a fifth, plus a sparse far pair (C clamps to 0)."""
    box, h, k6, ks = _consts(orc, CFG)
    m, rho0 = box["mass"], CFG["rest"]
    a = 0.15 * h  # half side of the square
    star = np.array([[0, 0], [a, a], [-a, a], [a, -a], [-a, -a]])
    pair = np.array([[0.0, 0.0], [0.5 * h, 0.0]]) + [20 * h, 0]
    pts = np.vstack([star, pair]) + CENTER
    dens, lam, dp = orc.c3_pass_of(CFG, pts, "fp64")

    r1, r2, r3 = a * np.sqrt(2), 2 * a, 2 * a * np.sqrt(2)  # corner-centre, corner-side, corner-diagonal
    q1, q2, q3 = _q(r1, h, ks), _q(r2, h, ks), _q(r3, h, ks)
    w0 = _w(0.0, h, k6)
    rho_c = m * (w0 + 4 * _w(r1, h, k6))
    rho_k = m * (w0 + _w(r1, h, k6) + 2 * _w(r2, h, k6) + _w(r3, h, k6))
    C_c, C_k = rho_c / rho0 - 1, rho_k / rho0 - 1
    assert C_c > 0 and C_k > 0
    f = m / rho0
    sum_sq_c, sq_sum_c = f * f * 4 * q1 * q1, 0.0
    sum_sq_k = f * f * (q1 * q1 + 2 * q2 * q2 + q3 * q3)
    sq_sum_k = f * f * 2 * (q1 / np.sqrt(2) + q2 + q3 / np.sqrt(2)) ** 2
    lam_c = -C_c / (sum_sq_c + sq_sum_c + 200.0)
    lam_k = -C_k / (sum_sq_k + sq_sum_k + 200.0)
    # corner (a, a): every spiky gradient points from it toward its neighbour, so with lambda < 0
    # its displacement points outward along (1, 1)
    dp_k = -f * ((lam_k + lam_c) * q1 / np.sqrt(2) + 2 * lam_k * (q2 + q3 / np.sqrt(2)))
    assert dp_k > 0

    assert dens[0] == pytest.approx(rho_c, rel=1e-12)
    assert np.allclose(dens[1:5], rho_k, rtol=1e-12, atol=0)
    assert lam[0] == pytest.approx(lam_c, rel=1e-10)
    assert np.allclose(lam[1:5], lam_k, rtol=1e-10, atol=0)
    assert np.abs(dp[0]).max() < 1e-12 * dp_k
    signs = np.sign(star[1:5])
    assert np.allclose(dp[1:5], dp_k * signs, rtol=1e-10, atol=0)
    # the sparse pair: C = max(rho/rho0 - 1, 0) = 0, so lambda and dp vanish exactly
    assert dens[5] < rho0 and lam[5] == 0.0 and lam[6] == 0.0
    assert np.all(dp[5:] == 0.0)
    # plausible mistakes this configuration separates from the oracle's values
    no_mass_factor = -C_k / ((sum_sq_k + sq_sum_k) / (f * f) + 200.0)  # grad C without m/rho0
    no_sq_sum = -C_k / (sum_sq_k + 200.0)                            # |sum g|^2 dropped
    no_relax = -C_k / (sum_sq_k + sq_sum_k)                           # EPSILON_RELAX dropped
    for wrong in (no_mass_factor, no_sq_sum, no_relax):
        assert abs(wrong - lam_k) > 1e-6 * abs(lam_k)
    one_sided_dp = -f * 2 * lam_k * (q1 / np.sqrt(2) + q2 + q3 / np.sqrt(2))  # lambda_i + lambda_i
    assert abs(one_sided_dp - dp_k) > 1e-3 * dp_k


def test_density_pass_f32_matches_fp64_star(orc):
    a = 0.15 * 2.0e-3
    pts = np.array([[0, 0], [a, a], [-a, a], [a, -a], [-a, -a]]) + CENTER
    d64, l64, p64 = orc.c3_pass_of(CFG, pts, "fp64")
    d32, l32, p32 = orc.c3_pass_of(CFG, pts.astype(np.float32), "ref")
    # f32 positions near 0.128 carry ~1e-8 m of rounding on 3e-4 m separations
    assert np.allclose(d32, d64, rtol=1e-4)
    assert np.allclose(l32, l64, rtol=1e-3)
    assert np.allclose(p32, p64, rtol=1e-3, atol=1e-4 * np.abs(p64).max())


def _two_particle(orc, cfg, r, e, b0, b1):
    box, h, k6, ks = _consts(orc, cfg)
    m, mu, dt = box["mass"], cfg["mu"], np.float64(np.float32(cfg["dt"]))
    rho = m * (_w(0.0, h, k6) + _w(r, h, k6))
    return m, rho, h, ks, mu, dt


def test_two_particle_viscosity_closed_form(orc):
    """(I + dt L) v = b for one pair: the relative normal velocity shrinks by 1 + 8 c dt, the
    tangential part and the centre-of-mass velocity stay (c = m (mu_i + mu_j)/(rho_i rho_j)
    |d . grad W| / (r^2 + 0.01 h^2), L's factor 4 applied to both rows)."""
    cfg = dict(CFG, mu=1.0, dt=1.0 / 1440.0, tol=1e-13, max_cg=50)
    h = orc.c3_box(cfg)["h"]
    r = 0.5 * h
    e = np.array([np.cos(np.pi / 6), np.sin(np.pi / 6)])
    pts = np.array([CENTER, CENTER + r * e])
    b = np.array([[0.01, 0.02], [-0.03, 0.005]])
    m, rho, h, ks, mu, dt = _two_particle(orc, cfg, r, e, b[0], b[1])
    d = pts[0] - pts[1]
    rr = np.linalg.norm(d)
    c = m * (2 * mu) / (rho * rho) * rr * _q(rr, h, ks) / (rr * rr + ETA * h * h)
    U = b[0] - b[1]
    un = U @ e
    S = b[0] + b[1]
    Unew = (U - un * e) + un / (1 + 8 * c * dt) * e
    expect = np.array([(S + Unew) / 2, (S - Unew) / 2])
    v, st = orc.c3_viscous_of(cfg, pts, b, True, "fp64")
    assert st["cg_iterations"] <= 2
    assert np.allclose(v, expect, rtol=1e-10, atol=1e-14)
    # the operator alone: (L b)_0 = 4 c e (e . (b0 - b1)) = -(L b)_1
    Lb = orc.c3_viscous_of(cfg, pts, b, False, "fp64")
    assert np.allclose(Lb[0], 4 * c * un * e, rtol=1e-11, atol=0)
    assert np.allclose(Lb[1], -Lb[0], rtol=1e-14, atol=0)
    # the configuration separates a factor 2 (instead of 4) and a missing eta^2 from the oracle
    assert 8 * c * dt > 1.0
    c_no_eta = m * (2 * mu) / (rho * rho) * _q(rr, h, ks) / rr
    for wrong in (un / (1 + 4 * c * dt), un / (1 + 8 * c_no_eta * dt)):
        assert abs(wrong - Unew @ e) > 1e-3 * abs(Unew @ e)


def _hex_patch(dx):
    pts = []
    for j in range(-4, 5):
        for i in range(-4, 5):
            pts.append([dx * (i + 0.5 * (j & 1)), dx * 0.8660254 * j])
    return np.array(pts)


@pytest.mark.parametrize("mode", ["ref", "fp64"])
def test_viscous_operator_vanishes_for_rigid_rotation(orc, mode):
    """The reference's pin (pbf.rs:2428-2473): a hexagonal patch (one-sided sums on its rim) feels no
    force from a rigid rotation under the central-force operator, and a shear is dissipative."""
    dx = 0.02
    cfg = dict(CFG, side=64, dx=dx, mu=1.0)
    c0 = np.array([0.64, 0.64])
    rel = _hex_patch(dx)
    pts = rel + c0
    if mode == "ref":
        pts = pts.astype(np.float32).astype(np.float64)
        rel = pts - c0
    omega = 3.0
    v = omega * np.stack([-rel[:, 1], rel[:, 0]], 1)
    out = orc.c3_viscous_of(cfg, pts, v, False, mode)
    _, off, nb, w = orc.c3_weights_of(cfg, pts)
    scale = w.max() * omega * 4.0 * dx
    worst = np.linalg.norm(out, axis=1).max()
    assert worst < 1e-4 * scale
    # the plain Morris form c (v_i - v_j) would not vanish at the rim: the check separates it
    plain = np.zeros_like(v)
    for i in range(len(pts)):
        for k in range(off[i], off[i + 1]):
            plain[i] += 4 * w[k] * (v[i] - v[nb[k]])
    assert np.linalg.norm(plain, axis=1).max() > 1e-2 * scale
    shear = np.stack([rel[:, 1], np.zeros(len(rel))], 1)
    out_s = orc.c3_viscous_of(cfg, pts, shear, False, mode)
    assert float((out_s * shear).sum()) > 0.0


def _brute_system(pts, m, mu, h, dt):
    """FP64 (I + dt L) over all pairs within h, from the reference formulas (pbf.rs:312-407)."""
    n = len(pts)
    k6, ks = 4.0 / (np.pi * h ** 8), 30.0 / (np.pi * h ** 5)
    d = pts[:, None, :] - pts[None, :, :]
    r2 = (d ** 2).sum(-1)
    near = (r2 <= h * h) & ~np.eye(n, dtype=bool)
    w = np.where(near, k6 * np.clip(h * h - r2, 0, None) ** 3, 0.0)
    rho = m * (k6 * h ** 6 + w.sum(1))
    r = np.sqrt(r2)
    q = np.where(near & (r > 1e-9) & (r < h), ks * (h - r) ** 2, 0.0)
    c = np.where(near, m * (2 * mu) / (rho[:, None] * rho[None, :]) * r * q / (r2 + ETA * h * h), 0.0)
    A = np.eye(2 * n)
    for i in range(n):
        for j in range(n):
            if not near[i, j] or r[i, j] <= 1e-9:
                continue
            e = d[i, j] / r[i, j]
            blk = dt * 4 * c[i, j] * np.outer(e, e)
            A[2 * i:2 * i + 2, 2 * i:2 * i + 2] += blk
            A[2 * i:2 * i + 2, 2 * j:2 * j + 2] -= blk
    return A, rho, c, near


@pytest.mark.parametrize("tol,max_cg", [(1e-13, 500), (1e-3, 50)])
def test_cg_solution_against_dense_solve(orc, tol, max_cg):
    cfg = dict(CFG, mu=1.0, dt=1.0 / 1440.0, tol=tol, max_cg=max_cg)
    box = orc.c3_box(cfg)
    h, m = box["h"], box["mass"]
    rng = np.random.default_rng(3)
    g = np.stack(np.meshgrid(np.arange(7), np.arange(6)), -1).reshape(-1, 2) * (0.5 * h)
    pts = CENTER + g + rng.uniform(-0.1, 0.1, g.shape) * h
    b = rng.normal(0, 0.01, pts.shape)
    A, rho, c, near = _brute_system(pts, m, 1.0, h, np.float64(np.float32(cfg["dt"])))
    # the oracle's densities and weights equal the brute-force ones
    dens, off, nb, w = orc.c3_weights_of(cfg, pts)
    assert np.allclose(dens, rho, rtol=1e-12, atol=0)
    for i in range(len(pts)):
        row = nb[off[i]:off[i + 1]]
        assert set(row.tolist()) == set(np.nonzero(near[i])[0].tolist())
        assert np.allclose(w[off[i]:off[i + 1]], c[i, row], rtol=1e-11, atol=0)
    v, st = orc.c3_viscous_of(cfg, pts, b, True, "fp64")
    bf = b.reshape(-1)
    res = np.linalg.norm(bf - A @ v.reshape(-1))
    assert res <= tol * np.linalg.norm(bf) * (1 + 1e-6) + 1e-18
    if tol < 1e-10:
        exact = np.linalg.solve(A, bf)
        assert np.abs(v.reshape(-1) - exact).max() <= 1e-9 * np.abs(exact).max()
        assert st["cg_iterations"] > 3
    assert np.linalg.norm(v) < np.linalg.norm(b)  # the implicit step dissipates
