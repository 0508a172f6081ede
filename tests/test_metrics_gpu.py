"""GPU parity of the on-demand drum metrics (metrics.rs) against the oracle: the mixing index is the
reference's f32 arithmetic, so GPU and oracle agree bit for bit (None where the reference returns
None), for one drum and for many drums in one launch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REALTIME = {"simulation": {"max_balls": 300, "resolution": 25}}
C2 = {"mill": {"diameter_m": 0.1397, "speed_mode": "percent_critical", "speed_value": 75.0},
      "media": {"density_kg_m3": 7800.0}, "simulation": {"resolution": 100, "max_balls": 2000}}


@pytest.fixture(scope="module")
def mx():
    from milldynamics3_amd import metrics
    return metrics


def _cases(orc):
    rng = np.random.default_rng(5)
    out = []
    for params in (REALTIME, C2):
        o = orc.OracleSim(params, "ref")
        st = o.state()
        R = float(params.get("mill", {}).get("diameter_m", 0.063)) / 2
        out.append((st["fluid_x"].astype(np.float32), st["dye"].astype(np.float32), R))
        # the same particles, dye shuffled (a mixed field)
        out.append((st["fluid_x"].astype(np.float32), rng.permutation(st["dye"]).astype(np.float32), R))
    R = 0.0315
    for n in (1, 255, 256, 257, 5000, 300000):
        r = R * np.sqrt(rng.random(n))
        a = rng.random(n) * 2 * np.pi
        pts = np.stack([r * np.cos(a), r * np.sin(a)], 1).astype(np.float32)
        out.append((pts, (pts[:, 0] < 0).astype(np.float32), R))
        out.append((pts, rng.random(n).astype(np.float32), R))
    pts = (rng.random((3000, 2)) * 0.2 - 0.1).astype(np.float32)  # many outside [-R, R]: clamped cells
    pts[::97] = np.nan
    out.append((pts, (rng.random(3000) < 0.3).astype(np.float32), R))
    out.append((pts[:10] * 0 + 0.001, np.arange(10, dtype=np.float32) % 2, R))  # one cell: None
    out.append((pts[:50], np.ones(50, np.float32), R))                          # no variance: None
    out.append((pts[:50], np.zeros(50, np.float32), 0.0))                       # no drum: None
    return out


def test_mixing_index_matches_oracle(mx, orc):
    for k, (pts, dye, R) in enumerate(_cases(orc)):
        print(f"mixing index case {k}: {len(pts)} particles", flush=True)
        got = mx.mixing_index(pts, dye, R)
        ref = orc.mixing_index_of(pts, dye, R)
        assert (got is None) == (ref is None), k
        if ref is not None:
            assert np.float32(got) == np.float32(ref), (k, got, ref)


def test_mixing_index_batch_matches_single_calls(mx, orc):
    cases = _cases(orc)
    got = mx.mixing_index_batch(cases)
    for k, (pts, dye, R) in enumerate(cases):
        ref = orc.mixing_index_of(pts, dye, R)
        assert (got[k] is None) == (ref is None), k
        if ref is not None:
            assert np.float32(got[k]) == np.float32(ref), k
