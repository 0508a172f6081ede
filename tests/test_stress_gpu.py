"""GPU parity of StressAccum (stress.rs:83-254) through mg_stress_update / mg_stress_state /
mg_stress_get_metrics.

Bar: bit-exact against the oracle's restatement (oracle/mg_oracle.cpp Stress, pinned by the
reference's own tests in tests/test_stress_oracle.py): all 49 accumulator values of every drum
after every update, and the StressMetrics fields. Record streams: 37 drums over 40 sub-steps, a
small ball population so pairs and wall contacts recur inside and across debounce windows, pair
order reversed at random, energies spread over all 28 bins and placed next to the bin edges, per-drum
debounce factors (0 = off, the default 0.25, 2.0), repeated clocks. Also: a full debounce table is
reported, a clock going back is rejected.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

WALL = 0xFFFFFFFF


@pytest.fixture(scope="module")
def st():
    from milldynamics3_amd import stress
    return stress


def _same(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return (a.view(np.uint64) == b.view(np.uint64)) | (a == b)


def _stream(rng, n, m3, scale, d_ball):
    """n records: ball-ball among 12 balls (random order of the pair) and ball-wall; speeds chosen
    so e_hat covers the bins, a quarter placed next to a bin edge 10^(-3 + k / 4)."""
    kind = (rng.random(n) < 0.3).astype(np.uint32)
    ball = rng.integers(0, 12, n).astype(np.uint32)
    other = (ball + 1 + rng.integers(0, 11, n)).astype(np.uint32) % 12
    partner = np.where(kind == 1, WALL, other).astype(np.uint32)
    m_red = np.where(kind == 0, 0.5 * m3, m3)
    e_hat = 10.0 ** rng.uniform(-3.5, 4.5, n)
    edge = rng.random(n) < 0.25
    e_hat[edge] = 10.0 ** (-3.0 + rng.integers(1, 28, edge.sum()) / 4.0)
    frac = rng.random(n)  # split the energy between the normal and the slip speed
    v2 = 2.0 * e_hat * scale / m_red
    vn = np.sqrt(v2 * frac).astype(np.float32)
    vs = np.sqrt(v2 * (1.0 - frac)).astype(np.float32)
    toward = np.where(rng.random(n) < 0.5, np.inf, -np.inf).astype(np.float32)  # one ulp either side
    vn = np.where(edge & (rng.random(n) < 0.7), np.nextafter(vn, toward), vn).astype(np.float32)
    return kind, ball, partner, vn, vs


def test_many_drums_match_oracle(st, orc):
    rng = np.random.default_rng(5)
    n_drums, steps = 37, 40
    m3, d_ball = 2.0e-5, 0.002
    scale = m3 * 9.81 * d_ball
    gpu = st.StressAccumBatch(n_drums, table_capacity=256)
    refs = [orc.StressAccum() for _ in range(n_drums)]
    for k in range(n_drums):
        f = (0.0, 0.25, 2.0)[k % 3]
        if k % 3 != 1:
            gpu.set_debounce_factor(k, f)
            refs[k].set_debounce_factor(f)
    alphas = rng.uniform(0.01, 1.0, n_drums)
    sub_dt = 1e-4
    for s in range(steps):
        recs, steps_ = [], []
        now = (s - s % 3) * sub_dt * 7.0  # the clock repeats for three updates at a time
        for k in range(n_drums):
            cols = _stream(rng, int(rng.integers(0, 40)), m3, scale, d_ball)
            recs.append(st.records(*cols))
            ww = float(rng.uniform(0.0, 1e-6))
            steps_.append({"m3": m3, "scale_j": scale, "d": d_ball, "now_s": now, "sub_dt": sub_dt,
                           "alpha": float(alphas[k]), "wall_work_j": ww})
            refs[k].update(*cols, m3, scale, d_ball, now, sub_dt, float(alphas[k]), ww)
        print(f"stress update {s}: {sum(len(r) for r in recs)} records", flush=True)
        gpu.update(recs, steps_)
        for k in range(n_drums):
            g, r = gpu.state(k), refs[k].state()
            bad = np.flatnonzero(~_same(g, r))
            assert bad.size == 0, f"update {s} drum {k}: values {bad.tolist()} gpu {g[bad]} oracle {r[bad]}"
    for k in range(0, n_drums, 5):
        g = gpu.metrics(k, 150 + k, m3, scale, d_ball, 3.5)
        r = refs[k].metrics(150 + k, m3, scale, d_ball, 3.5)
        for key, v in r.items():
            assert _same(g[key], v), f"drum {k}: {key} gpu {g[key]!r} oracle {v!r}"
    gpu.close()


def test_full_table_and_clock_going_back_are_errors(st):
    m3, d_ball = 2.0e-5, 0.002
    scale = m3 * 9.81 * d_ball
    gpu = st.StressAccumBatch(1, table_capacity=64)
    n = 80  # 80 distinct ball-wall keys inside one window: more than the 64 slots
    recs = st.records([1] * n, list(range(n)), [WALL] * n, [0.5] * n, [0.0] * n)
    step = {"m3": m3, "scale_j": scale, "d": d_ball, "now_s": 1.0, "sub_dt": 1e-4, "alpha": 1.0, "wall_work_j": 0.0}
    with pytest.raises(OverflowError):
        gpu.update([recs], [step])
    with pytest.raises(ValueError):
        gpu.update([recs[:1]], [dict(step, now_s=0.5)])
    gpu.close()
