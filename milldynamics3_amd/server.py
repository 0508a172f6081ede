"""Served sweeps over ``include/millgpu.h``: many independent drum runs stepped together.

``Server(max_drums, max_balls, hbm_budget_bytes)`` owns the slots; ``submit(params, revolutions)``
queues a run of one drum (Params as for ``Simulation``) for that many drum revolutions from its
seeded initial state; ``poll(ticks)`` runs ticks — admissions in arrival order within the HBM
budget, one substep of every running drum as one batch, per-revolution records, retirement of
finished runs; ``result(id)`` and ``stream(id)`` read a run's outcome and its per-revolution
records. A run's results equal the same drum stepped alone (drums never exchange data). The
reference has no server; its sweep analog runs seeds one after another (batch.rs:40-106). There
is no CPU path.
"""
from __future__ import annotations

import ctypes as C
import json

from .sim import _check, lib, params_json


class _Config(C.Structure):
    _fields_ = [("max_drums", C.c_uint32), ("max_balls", C.c_uint32), ("hbm_budget_bytes", C.c_uint64)]


class _Status(C.Structure):
    _fields_ = [("queued", C.c_uint32), ("running", C.c_uint32), ("done", C.c_uint32),
                ("ticks", C.c_uint64), ("drum_substeps", C.c_uint64), ("used_bytes", C.c_uint64),
                ("now_s", C.c_double)]


class _Result(C.Structure):
    _fields_ = [("state", C.c_int32), ("n_balls", C.c_uint32), ("n_fluid", C.c_uint32),
                ("revolutions_done", C.c_uint32), ("substeps_total", C.c_uint64), ("substeps_done", C.c_uint64),
                ("footprint_bytes", C.c_uint64), ("arrival_s", C.c_double), ("start_s", C.c_double),
                ("done_s", C.c_double), ("power_draw_w", C.c_float), ("torque_nm", C.c_float),
                ("collision_rate_per_s", C.c_float), ("dissipated_power_w", C.c_float)]


STATES = {0: "queued", 1: "running", 2: "done"}


def _lib():
    L = lib()
    if getattr(L, "_mg_server_ready", False):
        return L
    P, ch, sz = C.c_void_p, C.c_char_p, C.c_size_t
    L.mg_server_create.argtypes = [C.POINTER(_Config), C.POINTER(P), ch, sz]
    L.mg_server_destroy.argtypes = [P]
    L.mg_server_now.argtypes = [P]
    L.mg_server_now.restype = C.c_double
    L.mg_server_submit.argtypes = [P, ch, C.c_double, C.c_double, C.POINTER(C.c_uint64), ch, sz]
    L.mg_server_poll.argtypes = [P, C.c_uint32, C.POINTER(_Status), ch, sz]
    L.mg_server_result.argtypes = [P, C.c_uint64, C.POINTER(_Result)]
    L.mg_server_stream_json.argtypes = [P, C.c_uint64, ch, sz]
    L.mg_server_release.argtypes = [P, C.c_uint64]
    L._mg_server_ready = True
    return L


class Server:
    def __init__(self, max_drums: int = 256, max_balls: int = 2048, hbm_budget_bytes: int = 0):
        self._L = _lib()
        cfg = _Config(int(max_drums), int(max_balls), int(hbm_budget_bytes))
        err = C.create_string_buffer(512)
        h = C.c_void_p()
        rc = self._L.mg_server_create(C.byref(cfg), C.byref(h), err, 512)
        if rc == -3:
            raise MemoryError(err.value.decode())
        _check(rc, err, "mg_server_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._L.mg_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def now(self) -> float:
        """The server clock (seconds since creation)."""
        return float(self._L.mg_server_now(self._h))

    def submit(self, params, revolutions: float, arrival_s: float = -1.0) -> int:
        """Queues a run; returns its id. ValueError for invalid Params (the reference's validate
        message), a drum that does not turn or a bad revolution count; MemoryError if the drum
        cannot fit the server (balls or footprint)."""
        err = C.create_string_buffer(1024)
        rid = C.c_uint64()
        rc = self._L.mg_server_submit(self._h, params_json(params).encode(), float(revolutions), float(arrival_s),
                                      C.byref(rid), err, 1024)
        if rc == -3:
            raise MemoryError(err.value.decode())
        _check(rc, err, "mg_server_submit")
        return int(rid.value)

    def poll(self, ticks: int = 1) -> dict:
        st = _Status()
        err = C.create_string_buffer(1024)
        _check(self._L.mg_server_poll(self._h, int(ticks), C.byref(st), err, 1024), err, "mg_server_poll")
        return {name: getattr(st, name) for name, _ in _Status._fields_}

    def result(self, rid: int) -> dict:
        r = _Result()
        _check(self._L.mg_server_result(self._h, int(rid), C.byref(r)), what="mg_server_result")
        out = {name: getattr(r, name) for name, _ in _Result._fields_}
        out["state"] = STATES.get(out["state"], out["state"])
        return out

    def stream(self, rid: int) -> list:
        """The run's per-revolution records (power and torque EMAs, collision rate, dissipated
        power, mixing index or None, the revolution's impact-energy histogram)."""
        n = self._L.mg_server_stream_json(self._h, int(rid), None, 0)
        if n >= 0:
            raise RuntimeError("mg_server_stream_json")
        buf = C.create_string_buffer(-n)
        n = self._L.mg_server_stream_json(self._h, int(rid), buf, -n)
        if n < 0:
            raise RuntimeError("mg_server_stream_json")
        return json.loads(buf.value.decode())

    def release(self, rid: int):
        _check(self._L.mg_server_release(self._h, int(rid)), what="mg_server_release")
