"""The C-ABI library without a GPU: it loads, exports every function include/millgpu.h declares,
and its host-only entry points behave (no kernel is launched here)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    src = open(os.path.join(ROOT, "include", "millgpu.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"\b(mg_[a-z0-9_]+)\s*\(", src)
    return sorted(set(names))


@pytest.fixture(scope="module")
def L():
    from milldynamics3_amd import sim
    return sim.lib()


def test_library_exports_every_declared_function(L):
    names = _declared_functions()
    assert len(names) > 50
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_drum_engine_switch_is_reported(L):
    from milldynamics3_amd import drum_engine_enabled
    assert drum_engine_enabled() in (True, False)
    assert L.mg_drum_engine_enabled() == (1 if drum_engine_enabled() else 0)


def test_box_geometry_matches_oracle(orc):
    from milldynamics3_amd.box import box_geometry
    for cfg in ({"side": 1024}, {"side": 40, "dx": 2.5e-3}, {"side": 64, "dx": 0.125}):
        assert box_geometry(cfg) == orc.c3_box(cfg)


def test_box_geometry_rejects_bad_config():
    from milldynamics3_amd.box import box_geometry
    with pytest.raises(ValueError):
        box_geometry({"side": 4})
    with pytest.raises(ValueError):
        box_geometry({"dx": -1.0})
    for bad in (-1.0, float("inf"), float("nan")):
        with pytest.raises(ValueError):
            box_geometry({"yield_stress": bad})


def test_params_validation_messages_without_device(L):
    err = C.create_string_buffer(512)
    h = C.c_void_p()
    bad = b'{"mill": {"diameter_m": -1.0}}'
    rc = L.mg_sim_create(bad, C.byref(h), err, 512)
    assert rc == -1
    assert err.value


def test_server_config_is_checked_before_the_device(L):
    from milldynamics3_amd import server
    for bad in ((0, 2048), (4097, 2048), (4, 0), (4, 2049)):
        with pytest.raises(RuntimeError):
            server.Server(max_drums=bad[0], max_balls=bad[1])
