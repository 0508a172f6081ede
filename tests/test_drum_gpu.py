"""Drum path on the GPU through the C ABI.

The initial state is seeded on the host and uploaded; it must equal the oracle's bit for bit
(dem.rs:355-409, pbf.rs:1044-1119, rng.rs). Substeps are not run on the GPU in this build: an
earlier single-kernel form of the substep engine raised a GPU memory aperture violation on its first
substep and no mistake in the code has been shown to cause it, so the step entry points return
MG_E_DISABLED and the substep parity tests stay skipped until a cause is shown and fixed.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REALTIME = {"simulation": {"max_balls": 300, "resolution": 25}}
STEP_SKIP = ("drum substeps are switched off in this build: an earlier form faulted on the GPU "
             "(memory aperture violation) and its cause has not been shown")


@pytest.fixture(scope="module")
def mg():
    import milldynamics3_amd as m
    return m


def test_initial_state_matches_oracle(mg, orc):
    g = mg.Simulation(REALTIME)
    o = orc.OracleSim(REALTIME, "mixed")
    st = o.state()
    gb, gf = g.balls(), g.fluid()
    assert (g.n_balls, g.n_fluid) == (407, 794)
    assert np.array_equal(gb.x.astype(np.float64), st["ball_x"])
    assert np.array_equal(gf.x.astype(np.float64), st["fluid_x"])
    assert np.array_equal(gb.v.astype(np.float64), st["ball_v"])


def test_step_reports_disabled_engine(mg):
    if mg.drum_engine_enabled():
        pytest.skip("engine enabled in this build")
    g = mg.Simulation(REALTIME)
    with pytest.raises(RuntimeError, match="switched off"):
        g.step_fixed(1)


@pytest.mark.skip(reason=STEP_SKIP)
def test_one_substep_matches_oracle(mg, orc):
    g = mg.Simulation(REALTIME)
    o = orc.OracleSim(REALTIME, "mixed")
    g.step_fixed(1)
    o.step_fixed(1)
    st = o.state()
    R = 0.0315
    vref = abs(o.omega) * R + np.sqrt(9.81 * R)
    assert np.abs(g.balls().x - st["ball_x"]).max() / R < 1e-5
    assert np.abs(g.fluid().x - st["fluid_x"]).max() / R < 1e-5
    assert np.abs(g.fluid().v - st["fluid_v"]).max() / vref < 1e-5
