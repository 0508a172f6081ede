"""Slurry steps on the GPU over ``include/millgpu.h`` (crates/mill-core/src/pbf.rs).

``FluidBatch(n_drums)`` owns the workspace; ``step(drums)`` is ``FluidParticles::step_coupled``
(pbf.rs:1174-2365) for every drum at once. A drum is a dict with the drum (radius_m, omega,
drum_angle, lifters), the FluidParticles scalars (h, rest_density, particle_mass), the ball scalars
(ball_radius, ball_mass, ball_inertia, ball_true_radius), the SlurryParams the step reads
(viscosity_pa_s, rheology 0/1, yield_stress_pa, wall_no_slip, ball_no_slip, wettability,
surface_tension_n_m), iterations (pbf_iterations) and dt; the fluid state x (n, 2), v (n, 2) and the
warm start guess (n, 2) with has_guess — updated in place — and the balls ball_x (m, 2), ball_v
(m, 2), ball_omega (m,), which the step reads only. ``step`` returns per drum the coupling impulses
(impulse (m, 2), angular (m,)) and the FluidStepStats. There is no CPU path.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .grid import DeviceArray, _lib as _grid_lib

PARAM_KEYS = ("h", "rest_density", "particle_mass", "ball_radius", "ball_mass", "ball_inertia", "ball_true_radius",
              "viscosity_pa_s")
SLURRY_KEYS = ("yield_stress_pa", "wall_no_slip", "ball_no_slip", "wettability", "surface_tension_n_m")


class _Drum(C.Structure):
    _fields_ = ([("radius_m", C.c_float), ("omega", C.c_float), ("drum_angle", C.c_float),
                 ("lifter_count", C.c_uint32), ("lifter_height_m", C.c_float), ("lifter_base_width_m", C.c_float),
                 ("lifter_top_width_m", C.c_float), ("lifter_phase_deg", C.c_float)] +
                [(k, C.c_float) for k in PARAM_KEYS] + [("rheology", C.c_uint32)] +
                [(k, C.c_float) for k in SLURRY_KEYS] +
                [("iterations", C.c_uint32), ("dt", C.c_float), ("n_fluid", C.c_uint32), ("n_balls", C.c_uint32),
                 ("has_guess", C.c_int32)])


class _Stats(C.Structure):
    _fields_ = [("viscosity_iterations", C.c_uint32), ("density_iterations", C.c_uint32),
                ("mean_shear_rate_per_s", C.c_float), ("wall_work_j", C.c_float),
                ("wall_impulse_x", C.c_double), ("wall_impulse_y", C.c_double), ("wall_angular_impulse", C.c_double),
                ("wall_angle_bins", C.c_float * 32),
                ("ke_delta_noslip_j", C.c_double), ("ke_delta_drag_j", C.c_double),
                ("ke_delta_buoyancy_j", C.c_double), ("ke_delta_viscosity_j", C.c_double),
                ("ke_delta_clamp_j", C.c_double),
                ("coupling_clamp_hits", C.c_uint32), ("n_groups", C.c_uint32), ("n_links", C.c_uint32),
                ("links_in_range", C.c_uint32), ("joint_solve", C.c_int32), ("error_bits", C.c_uint32)]


def _lib():
    L = _grid_lib()
    if getattr(L, "_mg_fluid_ready", False):
        return L
    P = C.c_void_p
    L.mg_fluid_create.argtypes = [C.c_uint32, C.POINTER(P), C.c_char_p, C.c_size_t]
    L.mg_fluid_destroy.argtypes = [P]
    L.mg_fluid_force_iterations.argtypes = [P, C.c_int, C.c_uint32, C.c_uint32]
    L.mg_fluid_step.argtypes = [P, C.POINTER(_Drum)] + [P] * 8 + [C.POINTER(_Stats), C.c_char_p, C.c_size_t]
    L._mg_fluid_ready = True
    return L


def _cat(parts, width):
    if not parts:
        return np.zeros(0, np.float32)
    return np.ascontiguousarray(np.concatenate([np.asarray(p, np.float32).reshape(-1, width) for p in parts])
                                .reshape(-1))


class FluidBatch:
    def __init__(self, n_drums: int):
        self._L = _lib()
        self._h = C.c_void_p()
        err = C.create_string_buffer(256)
        rc = self._L.mg_fluid_create(int(n_drums), C.byref(self._h), err, 256)
        if rc != 0:
            raise (ValueError if rc == -4 else RuntimeError)(f"mg_fluid_create: {err.value.decode()}")
        self.n = int(n_drums)

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self._L.mg_fluid_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        self.close()

    def force_iterations(self, density_passes: int, cg_iterations: int, on: bool = True):
        rc = self._L.mg_fluid_force_iterations(self._h, 1 if on else 0, int(density_passes), int(cg_iterations))
        if rc != 0:
            raise ValueError("mg_fluid_force_iterations: bad pass count")

    def step(self, drums):
        """One slurry step of every drum (a list of dicts, see the module doc); x, v, guess and
        has_guess of each dict are replaced by the stepped ones. Returns per drum a dict with
        "impulse", "angular" and the FluidStepStats fields."""
        if len(drums) != self.n:
            raise ValueError("one dict per drum")
        recs = (_Drum * self.n)()
        xs, vs, gs, bxs, bvs, bos = [], [], [], [], [], []
        for k, p in enumerate(drums):
            r = recs[k]
            r.radius_m, r.omega, r.drum_angle = p["radius_m"], p["omega"], p["drum_angle"]
            li = p.get("lifters") or {}
            r.lifter_count = int(li.get("count", 0))
            r.lifter_height_m = float(li.get("height_m", 0.0))
            r.lifter_base_width_m = float(li.get("base_width_m", 0.0))
            r.lifter_top_width_m = float(li.get("top_width_m", 0.0))
            r.lifter_phase_deg = float(li.get("phase_deg", 0.0))
            for key in PARAM_KEYS + SLURRY_KEYS:
                setattr(r, key, float(p[key]))
            r.rheology = int(p.get("rheology", 0))
            r.iterations, r.dt = int(p["iterations"]), float(p["dt"])
            x = np.asarray(p["x"], np.float32).reshape(-1, 2)
            bx = np.asarray(p["ball_x"], np.float32).reshape(-1, 2)
            r.n_fluid, r.n_balls = len(x), len(bx)
            r.has_guess = 1 if p.get("has_guess") else 0
            xs.append(x)
            vs.append(np.asarray(p["v"], np.float32).reshape(-1, 2))
            g = p.get("guess")
            gs.append(np.asarray(g, np.float32).reshape(-1, 2) if g is not None and len(g) == len(x)
                      else np.zeros((len(x), 2), np.float32))
            bxs.append(bx)
            bvs.append(np.asarray(p["ball_v"], np.float32).reshape(-1, 2))
            bos.append(np.asarray(p["ball_omega"], np.float32).reshape(-1))
        dx, dv, dg = (DeviceArray.from_numpy(_cat(a, 2)) for a in (xs, vs, gs))
        dbx, dbv = DeviceArray.from_numpy(_cat(bxs, 2)), DeviceArray.from_numpy(_cat(bvs, 2))
        dbo = DeviceArray.from_numpy(_cat(bos, 1))
        nb_tot = sum(len(b) for b in bxs)
        dimp = DeviceArray.from_numpy(np.zeros(2 * nb_tot + 2, np.float32))
        dang = DeviceArray.from_numpy(np.zeros(nb_tot + 1, np.float32))
        st = (_Stats * self.n)()
        err = C.create_string_buffer(512)
        rc = self._L.mg_fluid_step(self._h, recs, dx.ptr, dv.ptr, dg.ptr, dbx.ptr, dbv.ptr, dbo.ptr, dimp.ptr,
                                   dang.ptr, st, err, 512)
        if rc != 0:
            raise {-4: ValueError, -3: OverflowError}.get(rc, RuntimeError)(f"mg_fluid_step: {err.value.decode()}")
        nf = [len(x) for x in xs]
        nbs = [len(b) for b in bxs]
        fo = np.concatenate([[0], np.cumsum(nf)]).astype(int)
        bo = np.concatenate([[0], np.cumsum(nbs)]).astype(int)
        X = dx.to_numpy(np.float32, 2 * fo[-1]).reshape(-1, 2)
        V = dv.to_numpy(np.float32, 2 * fo[-1]).reshape(-1, 2)
        G = dg.to_numpy(np.float32, 2 * fo[-1]).reshape(-1, 2)
        IMP = dimp.to_numpy(np.float32, 2 * bo[-1]).reshape(-1, 2)
        ANG = dang.to_numpy(np.float32, bo[-1])
        out = []
        for k, p in enumerate(drums):
            a, b = fo[k], fo[k + 1]
            s = st[k]
            p["x"], p["v"] = X[a:b].copy(), V[a:b].copy()
            if s.joint_solve:
                p["guess"], p["has_guess"] = G[a:b].copy(), True
            d = {name: (np.array(getattr(s, name)[:], np.float32) if name == "wall_angle_bins" else getattr(s, name))
                 for name, _ in _Stats._fields_}
            d["impulse"] = IMP[bo[k]:bo[k + 1]].copy()
            d["angular"] = ANG[bo[k]:bo[k + 1]].copy()
            out.append(d)
        return out
