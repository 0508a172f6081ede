"""Ball populations stepped on the GPU over ``include/millgpu.h`` (crates/mill-core/src/dem.rs).

``DemBatch(n_populations, max_balls)`` owns the workspace; ``step(populations)`` is
``DemState::step_recorded`` (dem.rs:530-1066) for every population at once. A population is a dict
with the drum (radius_m, omega, drum_angle, lifters), the shared ball radius / mass / inertia,
the MediaParams the DEM reads, iterations, dt, and the state arrays x (n, 2), v (n, 2), theta (n,),
omega_b (n,) — updated in place — plus optional coupling impulses ext_impulse (n, 2) and
ext_angular (n,). There is no CPU path.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import sim as _sim
from .grid import DeviceArray, _lib as _grid_lib
from .stress import RECORD_DTYPE

MEDIA_KEYS = ("restitution_ball_ball", "restitution_ball_wall", "friction_ball_ball", "friction_ball_wall",
              "friction_ball_ball_static", "friction_ball_wall_static", "friction_velocity_scale_m_s",
              "rolling_friction")


class _Drum(C.Structure):
    _fields_ = ([("radius_m", C.c_float), ("omega", C.c_float), ("drum_angle", C.c_float),
                 ("lifter_count", C.c_uint32), ("lifter_height_m", C.c_float), ("lifter_base_width_m", C.c_float),
                 ("lifter_top_width_m", C.c_float), ("lifter_phase_deg", C.c_float), ("ball_radius", C.c_float),
                 ("ball_mass", C.c_float), ("ball_inertia", C.c_float)] +
                [(k, C.c_float) for k in MEDIA_KEYS] +
                [("iterations", C.c_uint32), ("dt", C.c_float), ("n_balls", C.c_uint32), ("has_external", C.c_uint32),
                 ("energy_f64", C.c_uint32)])


class _Stats(C.Structure):
    _fields_ = [("wall_work_j", C.c_float), ("collision_count", C.c_uint32),
                ("impact_energy_histogram", C.c_uint32 * 12), ("dissipated_energy_j", C.c_float),
                ("max_substep_displacement_over_diameter", C.c_float), ("coupling_work_j", C.c_float),
                ("n_pairs", C.c_uint32), ("n_impacts", C.c_uint32), ("n_swept", C.c_uint32),
                ("projection_restarts", C.c_uint32)]


def _lib():
    L = _grid_lib()
    if getattr(L, "_mg_dem_ready", False):
        return L
    P = C.c_void_p
    L.mg_dem_create.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(P), C.c_char_p, C.c_size_t]
    L.mg_dem_destroy.argtypes = [P]
    L.mg_dem_step.argtypes = [P, C.POINTER(_Drum), P, P, P, P, P, P, C.POINTER(_Stats), C.c_char_p, C.c_size_t]
    L.mg_dem_impacts.argtypes = [P, C.c_uint32, P, C.c_uint32, C.POINTER(C.c_uint32)]
    L._mg_dem_ready = True
    return L


class DemBatch:
    def __init__(self, n_populations: int, max_balls: int, pairs_per_ball: int = 24, max_impacts: int = 4096):
        self._L = _lib()
        self._h = C.c_void_p()
        err = C.create_string_buffer(256)
        rc = self._L.mg_dem_create(int(n_populations), int(max_balls), int(pairs_per_ball), int(max_impacts),
                                   C.byref(self._h), err, 256)
        if rc != 0:
            raise (ValueError if rc == -4 else RuntimeError)(f"mg_dem_create: {err.value.decode()}")
        self.n = int(n_populations)
        self.max_impacts = int(max_impacts)

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self._L.mg_dem_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        self.close()

    def step(self, pops):
        """One step of every population (a list of dicts, see the module doc); the state arrays of
        each dict are replaced by the stepped ones. Returns one stats dict per population, with its
        impact records (RECORD_DTYPE) under "impacts"."""
        if len(pops) != self.n:
            raise ValueError("one dict per population")
        recs = (_Drum * self.n)()
        xs, vs, ths, oms, eis, eas = [], [], [], [], [], []
        any_ext = False
        for k, p in enumerate(pops):
            r = recs[k]
            r.radius_m, r.omega, r.drum_angle = p["radius_m"], p["omega"], p["drum_angle"]
            li = p.get("lifters") or {}
            r.lifter_count = int(li.get("count", 0))
            r.lifter_height_m = float(li.get("height_m", 0.0))
            r.lifter_base_width_m = float(li.get("base_width_m", 0.0))
            r.lifter_top_width_m = float(li.get("top_width_m", 0.0))
            r.lifter_phase_deg = float(li.get("phase_deg", 0.0))
            r.ball_radius, r.ball_mass, r.ball_inertia = p["ball_radius"], p["ball_mass"], p["ball_inertia"]
            for key in MEDIA_KEYS:
                setattr(r, key, float(p[key]))
            r.iterations, r.dt = int(p["iterations"]), float(p["dt"])
            r.energy_f64 = 1 if p.get("energy_f64") else 0
            x = np.asarray(p["x"], np.float32).reshape(-1, 2)
            r.n_balls = len(x)
            xs.append(x)
            vs.append(np.asarray(p["v"], np.float32).reshape(-1, 2))
            ths.append(np.asarray(p["theta"], np.float32).reshape(-1))
            oms.append(np.asarray(p["omega_b"], np.float32).reshape(-1))
            has = p.get("ext_impulse") is not None
            r.has_external = 1 if has else 0
            any_ext |= has
            eis.append(np.asarray(p["ext_impulse"], np.float32).reshape(-1, 2) if has else np.zeros((len(x), 2), np.float32))
            eas.append(np.asarray(p["ext_angular"], np.float32).reshape(-1) if has else np.zeros(len(x), np.float32))
        def cat(parts):
            return np.ascontiguousarray(np.concatenate(parts).astype(np.float32)) if parts else np.zeros(0, np.float32)
        dx, dv, dth, dom = (DeviceArray.from_numpy(cat(a)) for a in (xs, vs, ths, oms))
        dei = DeviceArray.from_numpy(cat(eis)) if any_ext else None
        dea = DeviceArray.from_numpy(cat(eas)) if any_ext else None
        st = (_Stats * self.n)()
        err = C.create_string_buffer(512)
        rc = self._L.mg_dem_step(self._h, recs, dx.ptr, dv.ptr, dth.ptr, dom.ptr, dei.ptr if dei else None,
                                 dea.ptr if dea else None, st, err, 512)
        if rc != 0:
            raise {-4: ValueError, -3: OverflowError}.get(rc, RuntimeError)(f"mg_dem_step: {err.value.decode()}")
        counts = [len(x) for x in xs]
        off = np.concatenate([[0], np.cumsum(counts)]).astype(int)
        X = dx.to_numpy(np.float32, 2 * off[-1]).reshape(-1, 2)
        V = dv.to_numpy(np.float32, 2 * off[-1]).reshape(-1, 2)
        TH = dth.to_numpy(np.float32, off[-1])
        OM = dom.to_numpy(np.float32, off[-1])
        out = []
        for k, p in enumerate(pops):
            a, b = off[k], off[k + 1]
            p["x"], p["v"], p["theta"], p["omega_b"] = X[a:b].copy(), V[a:b].copy(), TH[a:b].copy(), OM[a:b].copy()
            s = st[k]
            d = {name: (list(getattr(s, name)) if name == "impact_energy_histogram" else getattr(s, name))
                 for name, _ in _Stats._fields_}
            n_imp = C.c_uint32()
            buf = np.zeros(max(min(s.n_impacts, self.max_impacts), 1), dtype=RECORD_DTYPE)
            _sim._check(self._L.mg_dem_impacts(self._h, k, buf.ctypes.data_as(C.c_void_p), len(buf), C.byref(n_imp)),
                        what="mg_dem_impacts")
            d["impacts"] = buf[: min(n_imp.value, self.max_impacts)]
            out.append(d)
        return out
