"""GPU parity of surface::extract_surface (surface.rs:233-277) through mg_surface_extract.

Bar: bit-exact against the oracle's restatement (oracle/mg_oracle.cpp surf::extract, pinned by
the reference's surface tests in tests/test_surface_oracle.py): every drum's masked 128 x 128
field equals the oracle's as f32 bits, and the polygons match in number, length and every point's
f32 bits. Drums, extracted together in one batch and then again with new inputs: the seeded pool
of a plain drum, a lifter drum turned by 0.3 rad (mask and wall projection), the reference test's
41 x 41 lattice in a 2 m drum, a drum without fluid, a lone particle (the measured-peak fallback),
scattered blobs and droplets (several contours, saddle cells), and a large seeded pool (many
particle chunks per tile).
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LIFTERS = {"count": 8, "height_m": 0.03, "base_width_m": 0.04, "top_width_m": 0.02, "phase_deg": 10.0}


@pytest.fixture(scope="module")
def sf():
    from milldynamics3_amd import surface
    return surface


def _seeded(orc, radius, res, fill, lifters=None):
    params = {"mill": {"diameter_m": 2 * radius}, "media": {"fill_fraction": 0.0},
              "slurry": {"fill_fraction": fill}, "simulation": {"resolution": res}}
    if lifters:
        params["lifters"] = lifters
    o = orc.OracleSim(params, "ref")
    return {"radius_m": radius, "h": o.h, "rest_density": 1000.0, "particle_mass": o.fluid_mass,
            "fluid_x": o.state()["fluid_x"].astype(np.float32), "lifters": lifters}


def _cases(orc, seed):
    rng = np.random.default_rng(seed)
    cases = []
    d = _seeded(orc, 0.5, 24, 0.2)
    d["fluid_x"] = d["fluid_x"] + rng.normal(0, 1e-3, d["fluid_x"].shape).astype(np.float32)
    cases.append(("pool", d))
    d = _seeded(orc, 0.5, 24, 0.3, LIFTERS)
    d["drum_angle"] = 0.3 + 0.1 * seed
    cases.append(("lifters", d))
    R, dx = 2.0, np.float32(2.0 / 40)
    k = np.arange(-20, 21, dtype=np.float32)
    gx, gy = np.meshgrid(k * dx, k * dx)
    cases.append(("lattice", {"radius_m": R, "h": float(np.float32(2) * dx), "rest_density": 1000.0,
                              "particle_mass": float(np.float32(1000) * dx * dx),
                              "fluid_x": np.stack([gx.ravel(), gy.ravel()], 1).astype(np.float32)}))
    cases.append(("empty", {"radius_m": 0.5, "h": 0.02, "rest_density": 1000.0, "particle_mass": 0.1,
                            "fluid_x": np.zeros((0, 2), np.float32)}))
    cases.append(("lone", {"radius_m": 0.5, "h": 0.02, "rest_density": 1000.0, "particle_mass": 1.0,
                           "fluid_x": rng.uniform(-0.2, 0.2, (1, 2)).astype(np.float32)}))
    blobs = [rng.normal(c, 0.03, (150, 2)) for c in rng.uniform(-0.3, 0.3, (4, 2))]
    drops = rng.uniform(-0.45, 0.45, (60, 2))
    cases.append(("blobs", {"radius_m": 0.5, "h": 0.0125, "rest_density": 1000.0, "particle_mass": 0.15,
                            "fluid_x": np.concatenate(blobs + [drops]).astype(np.float32)}))
    d = _seeded(orc, 1.0, 100, 0.35)
    cases.append(("large pool", d))
    return cases


def _oracle(orc, d):
    drum = orc.drum_array(d["radius_m"], 0.0, d.get("lifters"))
    return orc.extract_surface(drum, d.get("drum_angle", 0.0), d["fluid_x"], d["h"], d["rest_density"],
                               d["particle_mass"], with_field=True)


def _bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("seed", [0])
def test_batch_matches_oracle(sf, orc, seed):
    batch = None
    for rnd in range(2):
        cases = _cases(orc, seed + rnd)
        if batch is None:
            batch = sf.SurfaceBatch(len(cases), sum(len(d["fluid_x"]) for _, d in cases) + 1000)
        print(f"surface extraction {rnd}: " + ", ".join(f"{n} {len(d['fluid_x'])}" for n, d in cases), flush=True)
        got = batch.extract([d for _, d in cases])
        for k, ((name, d), polys) in enumerate(zip(cases, got)):
            rp, rf = _oracle(orc, d)
            gf = batch.field(k)
            bad = np.argwhere(_bits(gf) != _bits(rf))
            assert bad.size == 0, f"{name}: field differs at {bad[:5].tolist()} gpu {gf[tuple(bad[0])]} oracle {rf[tuple(bad[0])]}"
            assert len(polys) == len(rp), f"{name}: {len(polys)} polygons vs {len(rp)}"
            for q, (a, b) in enumerate(zip(polys, rp)):
                assert a.shape == b.shape, f"{name} polygon {q}: {a.shape} vs {b.shape}"
                assert (_bits(a) == _bits(b)).all(), f"{name} polygon {q}: points differ"
            if name == "empty":
                assert polys == [] and not gf.any()
            if name in ("pool", "lattice", "large pool"):
                assert len(polys) == 1
            if name == "blobs":
                assert len(polys) >= 2
    batch.close()


def test_many_drums_walked_on_several_threads(sf, orc):
    # 48 drums: the host chain walk runs on a pool of threads, each over a range of drums
    base = _seeded(orc, 0.5, 24, 0.25, LIFTERS)
    cases = [dict(base, drum_angle=0.05 * k) for k in range(48)]
    batch = sf.SurfaceBatch(len(cases), sum(len(d["fluid_x"]) for d in cases))
    got = batch.extract(cases)
    for k, (d, polys) in enumerate(zip(cases, got)):
        rp, rf = _oracle(orc, d)
        assert (_bits(batch.field(k)) == _bits(rf)).all(), f"drum {k}: field differs"
        assert len(polys) == len(rp), f"drum {k}: {len(polys)} polygons vs {len(rp)}"
        for a, b in zip(polys, rp):
            assert a.shape == b.shape and (_bits(a) == _bits(b)).all(), f"drum {k}: points differ"
    batch.close()


def test_too_many_particles_is_an_argument_error(sf):
    batch = sf.SurfaceBatch(1, 10)
    with pytest.raises(ValueError):
        batch.extract([{"radius_m": 0.5, "h": 0.02, "rest_density": 1000.0, "particle_mass": 0.1,
                        "fluid_x": np.zeros((11, 2), np.float32)}])
    batch.close()
