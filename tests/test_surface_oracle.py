"""Pins of the oracle's extract_surface restatement (surface.rs:39-420): the reference's own tests
(surface.rs:437-560: the splat kernel's support, the interior field against the analytic
phi_full on a uniform lattice, no fluid -> no surface, a pool's single closed contour of about the
pool's area inside the drum) plus closed forms: a lone particle's contour is symmetric about it
(the fallback threshold at half the measured peak), and no contour point lies inside a lifter."""
import math

import numpy as np


def test_splat_kernel_support(orc):
    # surface.rs:437-443
    assert orc.splat_kernel(0.0, 1.0) > 0.0 and orc.splat_kernel(0.5, 1.0) > 0.0
    assert orc.splat_kernel(1.0, 1.0) == 0.0 and orc.splat_kernel(2.0, 1.0) == 0.0
    assert orc.splat_kernel(0.0, 1.0) == 1.0 and orc.splat_kernel(0.75, 1.0) == np.float32(0.25) ** 3


def test_interior_field_matches_analytic_phi_full(orc):
    # surface.rs:445-503: a square lattice at the equilibrium number density 1/dx^2, far from edges
    R, res = 2.0, 40
    dx = np.float32(R / res)
    h = float(np.float32(2.0) * dx)  # seed_lattice's h = 2 dx (pbf.rs:1044-1119)
    mass = float(np.float32(1000.0) * dx * dx)
    k = np.arange(-20, 21, dtype=np.float32)
    gx, gy = np.meshgrid(k * dx, k * dx)
    pts = np.stack([gx.ravel(), gy.ravel()], 1).astype(np.float32)
    _, field = orc.extract_surface(orc.drum_array(R), 0.0, pts, h, 1000.0, mass, with_field=True)
    rs = 1.5 * h
    phi = (1000.0 / mass) * (math.pi * rs * rs / 4.0)
    for dj in (-1, 0, 1):
        for di in (-1, 0, 1):
            assert abs(field[64 + dj, 64 + di] - phi) / phi < 0.1


def test_no_fluid_gives_no_surface(orc):
    assert orc.extract_surface(orc.drum_array(0.5), 0.0, np.zeros((0, 2)), 0.02, 1000.0, 0.1) == []


def _pool(orc, radius=0.5, res=24, fill=0.2, lifters=None):
    params = {"mill": {"diameter_m": 2 * radius}, "media": {"fill_fraction": 0.0},
              "slurry": {"fill_fraction": fill}, "simulation": {"resolution": res}}
    if lifters:
        params["lifters"] = lifters
    o = orc.OracleSim(params, "ref")
    return o, o.state()["fluid_x"].astype(np.float32)


def _area(poly):
    x, y = poly[:, 0].astype(np.float64), poly[:, 1].astype(np.float64)
    return 0.5 * abs(np.dot(x, np.roll(y, -1)) - np.dot(y, np.roll(x, -1)))


def test_a_pool_has_one_closed_contour_of_plausible_area(orc):
    # surface.rs:520-575 (on the seeded pool rather than a settled one)
    R = 0.5
    o, fx = _pool(orc, R)
    assert len(fx) > 100
    polys = orc.extract_surface(orc.drum_array(R), 0.0, fx, o.h, 1000.0, o.fluid_mass)
    assert len(polys) == 1 and len(polys[0]) >= 8
    assert abs(_area(polys[0]) - 0.2 * math.pi * R * R) / (0.2 * math.pi * R * R) < 0.35
    assert (np.linalg.norm(polys[0], axis=1) <= R + 1e-4).all()


def test_a_lone_particle_contour_is_symmetric_about_it(orc):
    R, h = 0.5, 0.02
    cell = 2 * R / 127
    p = np.array([[-R + 64 * cell, -R + 60 * cell]], np.float32)  # on a grid point
    polys = orc.extract_surface(orc.drum_array(R), 0.0, p, h, 1000.0, 1.0)
    assert len(polys) == 1
    c = polys[0].astype(np.float64).mean(0)
    assert np.abs(c - p[0]).max() < 1e-6
    r = np.linalg.norm(polys[0] - p[0], axis=1)
    assert r.max() < 1.5 * h  # inside the splat radius


def test_no_contour_point_inside_a_lifter(orc):
    # surface.rs:718-748 (the projection keeps every point in free space)
    lifters = {"count": 8, "height_m": 0.03, "base_width_m": 0.04, "top_width_m": 0.02, "phase_deg": 0.0}
    R = 0.5
    o, fx = _pool(orc, R, fill=0.3, lifters=lifters)
    drum = orc.drum_array(R, 0.0, lifters)
    polys = orc.extract_surface(drum, 0.3, fx, o.h, 1000.0, o.fluid_mass)
    assert polys
    for poly in polys:
        for q in poly:
            assert orc.sdf_world(drum, q, 0.3)[0] >= -1e-4
