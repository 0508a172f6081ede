"""Pins of the oracle's StressAccum restatement (stress.rs:83-254): the reference's own tests
(stress.rs:290-392) ported onto the standalone accumulator, plus the key rules of the debounce
(a pair is one key whatever its order; a ball's wall contacts are a key of their own)."""
import math

import numpy as np
import pytest

M3, D = 2.0e-5, 0.002
SCALE = M3 * 9.81 * D
WALL = 0xFFFFFFFF


def _rec(n, kind=0, vn=0.5, vs=0.0, ball=0, partner=1):
    return ([kind] * n, [ball] * n, [partner] * n, [vn] * n, [vs] * n)


def test_bins_are_log_spaced(orc):
    # stress.rs:291-299
    assert orc.se_bin(1e-9) == 0 and orc.se_bin(1e9) == 27 and orc.se_bin(1.0) == 12
    assert orc.se_bin(0.0) == 0 and orc.se_bin(-1.0) == 0 and orc.se_bin(float("nan")) == 0
    for k in range(1, 28):
        edge = 10.0 ** (-3.0 + k / 4.0)
        assert orc.se_bin(edge * (1 + 1e-9)) == min(k, 27)
        assert orc.se_bin(edge * (1 - 1e-9)) == k - 1


def test_stress_energy_uses_the_reduced_3d_mass(orc):
    # stress.rs:301-312 through one update (alpha 1, sub_dt 1: the rates are the instants)
    for kind, vn, vs, cls, m_red in ((0, 2.0, 0.0, 0, 0.5 * M3), (1, 1.0, 3.0, 3, M3), (0, 1.0, 3.0, 1, 0.5 * M3),
                                     (1, 2.0, 1.0, 2, M3)):
        acc = orc.StressAccum()
        acc.set_debounce_factor(0.0)
        acc.update(*_rec(1, kind, vn, vs), M3, SCALE, D, 0.0, 1.0, 1.0, 0.0)
        st = acc.state()
        se = 0.5 * m_red * (vn * vn + vs * vs)
        assert st[cls] == 1.0 and st[:4].sum() == 1.0
        assert st[4 + cls] == pytest.approx(se / SCALE, rel=1e-12)
        assert st[43] == pytest.approx(se, rel=1e-12)  # cumulative energy above the lowest threshold


def test_ema_converges_to_the_per_second_rate_and_normalises_per_ball(orc):
    # stress.rs:314-351
    acc = orc.StressAccum()
    acc.set_debounce_factor(0.0)
    sub_dt = 1e-3
    alpha = 1.0 - math.exp(-sub_dt / 1.0)
    for _ in range(20000):
        acc.update(*_rec(2), M3, SCALE, D, 0.0, sub_dt, alpha, 0.0)
    m = acc.metrics(100, M3, SCALE, D, 0.0)
    assert abs(m["impacts_per_ball_per_s_0"] - 20.0) < 0.05
    e_hat = 0.5 * 0.5 * M3 * 0.25 / SCALE
    assert abs(m["mean_se_scaled_0"] - e_hat) / e_hat < 1e-4
    assert m["mean_se_scaled_1"] == 0.0
    n_imp, se_each = 2.0 * 20000.0, 0.5 * 0.5 * M3 * 0.25
    assert abs(m["cumulative_effective_energy_per_ball_j_0"] * 100.0 - n_imp * se_each) / (n_imp * se_each) < 1e-9
    assert abs(acc.state()[8:36].sum() - 2000.0) < 5.0


def test_debounce_counts_one_impact_per_episode(orc):
    # stress.rs:353-374
    window = 0.25 * math.sqrt(D / 9.81)
    acc = orc.StressAccum()
    sub_dt = 1e-4
    t = 0.0
    while t < 2.0 * window - sub_dt:
        acc.update(*_rec(1), M3, SCALE, D, t, sub_dt, 1.0, 0.0)
        t += sub_dt
    assert acc.state()[8:36].sum() * sub_dt <= 1.0 + 1e-9
    se_each = 0.5 * 0.5 * M3 * 0.25
    n_counted = acc.metrics(1, M3, SCALE, D, 0.0)["cumulative_effective_energy_per_ball_j_0"] / se_each
    assert abs(n_counted - 2.0) < 1e-6


def test_specific_power_is_per_ball_slab_over_3d_mass(orc):
    # stress.rs:376-381
    m = orc.StressAccum().metrics(200, M3, M3 * 9.81 * D, D, 4.0)
    assert abs(m["specific_power_w_per_kg"] - 2.0) < 1e-5


def test_debounce_keys(orc):
    se_each = 0.5 * 0.5 * M3 * 0.25
    acc = orc.StressAccum()
    # (5, 2) and (2, 5) are one pair; ball 5 against the wall is another key; ball 2 against the
    # wall a third: three episodes in one sub-step, the repeats debounced
    kind = [0, 0, 1, 1, 1]
    ball = [5, 2, 5, 5, 2]
    partner = [2, 5, WALL, WALL, WALL]
    acc.update(kind, ball, partner, [0.5] * 5, [0.0] * 5, M3, SCALE, D, 1.0, 1e-4, 1.0, 0.0)
    st = acc.state()
    assert st[0] * 1e-4 == pytest.approx(1.0) and st[2] * 1e-4 == pytest.approx(2.0)
    assert st[43] == pytest.approx(se_each + 2 * 0.5 * M3 * 0.25)
    # past the window every key counts again
    acc.update(kind, ball, partner, [0.5] * 5, [0.0] * 5, M3, SCALE, D, 1.0 + 1.01 * 0.25 * math.sqrt(D / 9.81), 1e-4,
               1.0, 0.0)
    assert acc.state()[43] == pytest.approx(2 * (se_each + 2 * 0.5 * M3 * 0.25))
