"""Grinding-proxy stress accumulators on the GPU over ``include/millgpu.h`` (crates/mill-core/src/stress.rs).

``StressAccumBatch(n_drums)`` holds one ``StressAccum`` (stress.rs:83-254) per drum in device
memory; ``update(records, steps)`` is ``StressAccum::update`` (stress.rs:110-171) for every drum with
one sub-step's impact records, ``state(drum)`` the 49 accumulator values, ``metrics(...)``
``StressAccum::metrics`` (stress.rs:186-238). There is no CPU path.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import sim as _sim
from .grid import DeviceArray, _lib as _grid_lib

WALL = 0xFFFFFFFF


class _Step(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("m3", "scale_j", "d", "now_s", "sub_dt", "alpha", "wall_work_j")]


class _Metrics(C.Structure):
    _fields_ = [("ball_mass_3d_kg", C.c_float), ("scale_energy_j", C.c_float),
                ("impacts_per_ball_per_s", C.c_float * 4), ("impacts_total_per_ball_per_s", C.c_float),
                ("mean_se_scaled", C.c_float * 4), ("mean_se_scaled_drop", C.c_float),
                ("mean_se_scaled_shear", C.c_float), ("stress_intensity_scaled", C.c_float),
                ("wall_power_per_ball_w", C.c_float), ("specific_power_w_per_kg", C.c_float),
                ("thresholds_scaled", C.c_float * 6), ("effective_power_per_ball_w", C.c_float * 6),
                ("cumulative_wall_energy_per_ball_j", C.c_double),
                ("cumulative_specific_energy_j_per_kg", C.c_double),
                ("cumulative_effective_energy_per_ball_j", C.c_double * 6)]


RECORD_DTYPE = np.dtype([("kind", np.uint32), ("ball", np.uint32), ("partner", np.uint32),
                         ("normal_speed", np.float32), ("slip_speed", np.float32), ("restitution", np.float32)])


def _lib():
    L = _grid_lib()
    if getattr(L, "_mg_stress_ready", False):
        return L
    P = C.c_void_p
    L.mg_stress_create.argtypes = [C.c_uint32, C.c_uint32, C.POINTER(P), C.c_char_p, C.c_size_t]
    L.mg_stress_destroy.argtypes = [P]
    L.mg_stress_set_debounce_factor.argtypes = [P, C.c_uint32, C.c_double]
    L.mg_stress_update.argtypes = [P, P, P, C.POINTER(_Step), C.c_char_p, C.c_size_t]
    L.mg_stress_state.argtypes = [P, C.c_uint32, C.POINTER(C.c_double)]
    L.mg_stress_get_metrics.argtypes = [P, C.c_uint32, C.c_uint32] + [C.c_double] * 4 + [C.POINTER(_Metrics)]
    L._mg_stress_ready = True
    return L


def records(kind, ball, partner, normal_speed, slip_speed, restitution=None):
    """ImpactRecord array (dem.rs:152-169) from columns; partner WALL for a wall impact."""
    n = len(kind)
    r = np.zeros(n, dtype=RECORD_DTYPE)
    r["kind"], r["ball"], r["partner"] = kind, ball, partner
    r["normal_speed"], r["slip_speed"] = normal_speed, slip_speed
    r["restitution"] = 0.0 if restitution is None else restitution
    return r


class StressAccumBatch:
    def __init__(self, n_drums: int, table_capacity: int = 4096):
        self._L = _lib()
        self._h = C.c_void_p()
        err = C.create_string_buffer(256)
        rc = self._L.mg_stress_create(int(n_drums), int(table_capacity), C.byref(self._h), err, 256)
        if rc != 0:
            raise (ValueError if rc == -4 else RuntimeError)(f"mg_stress_create: {err.value.decode()}")
        self.n = int(n_drums)

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self._L.mg_stress_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        self.close()

    def set_debounce_factor(self, drum: int, factor: float):
        _sim._check(self._L.mg_stress_set_debounce_factor(self._h, int(drum), float(factor)),
                    what="mg_stress_set_debounce_factor")

    def update(self, per_drum_records, steps):
        """per_drum_records: one RECORD_DTYPE array per drum; steps: one dict per drum with m3,
        scale_j, d, now_s, sub_dt, alpha, wall_work_j."""
        if len(per_drum_records) != self.n or len(steps) != self.n:
            raise ValueError("one record array and one step per drum")
        recs = np.concatenate([np.asarray(r, dtype=RECORD_DTYPE) for r in per_drum_records])
        off = np.zeros(self.n + 1, dtype=np.uint32)
        off[1:] = np.cumsum([len(r) for r in per_drum_records])
        d_rec = DeviceArray.from_numpy(recs.view(np.uint8)) if len(recs) else DeviceArray(16)
        d_off = DeviceArray.from_numpy(off)
        st = (_Step * self.n)()
        for k, s in enumerate(steps):
            for f, _ in _Step._fields_:
                setattr(st[k], f, float(s[f]))
        err = C.create_string_buffer(256)
        rc = self._L.mg_stress_update(self._h, d_rec.ptr, d_off.ptr, st, err, 256)
        if rc != 0:
            raise {-4: ValueError, -3: OverflowError}.get(rc, RuntimeError)(f"mg_stress_update: {err.value.decode()}")

    def state(self, drum: int) -> np.ndarray:
        o = np.zeros(49)
        _sim._check(self._L.mg_stress_state(self._h, int(drum), o.ctypes.data_as(C.POINTER(C.c_double))),
                    what="mg_stress_state")
        return o

    def metrics(self, drum, n_balls, m3, scale_j, d, power_draw_w) -> dict:
        m = _Metrics()
        _sim._check(self._L.mg_stress_get_metrics(self._h, int(drum), int(n_balls), float(m3), float(scale_j), float(d),
                                                  float(power_draw_w), C.byref(m)), what="mg_stress_get_metrics")
        out = {}
        for name, typ in _Metrics._fields_:
            v = getattr(m, name)
            if hasattr(v, "__len__"):
                for i, x in enumerate(v):
                    out[f"{name}_{i}"] = float(x)
            else:
                out[name] = float(v)
        return out
