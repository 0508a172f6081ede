"""Fluid free surfaces extracted on the GPU over ``include/millgpu.h`` (crates/mill-core/src/surface.rs).

``SurfaceBatch(n_drums, max_fluid)`` owns the workspace; ``extract(drums)`` is
``surface::extract_surface`` (surface.rs:233-277) for every drum at once, as
``Simulation::fluid_surface`` (lib.rs:353-360) calls it. A drum is a dict with radius_m, drum_angle,
lifters (LiftersParams keys, optional), the FluidParticles scalars h, rest_density and
particle_mass, and fluid_x (n, 2). Each drum's result is a list of (k, 2) float32 polygons, as the
reference's Vec<Polygon>; ``field(d)`` gives drum d's masked 128 x 128 occupancy field.
There is no CPU path.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .grid import DeviceArray, _lib as _grid_lib

G = 128


class _Drum(C.Structure):
    _fields_ = [("radius_m", C.c_float), ("drum_angle", C.c_float), ("lifter_count", C.c_uint32),
                ("lifter_height_m", C.c_float), ("lifter_base_width_m", C.c_float),
                ("lifter_top_width_m", C.c_float), ("lifter_phase_deg", C.c_float), ("h", C.c_float),
                ("rest_density", C.c_float), ("particle_mass", C.c_float), ("n_fluid", C.c_uint32)]


def _lib():
    L = _grid_lib()
    if getattr(L, "_mg_surface_ready", False):
        return L
    P = C.c_void_p
    L.mg_surface_create.argtypes = [C.c_uint32, C.c_uint32, C.POINTER(P), C.c_char_p, C.c_size_t]
    L.mg_surface_destroy.argtypes = [P]
    L.mg_surface_extract.argtypes = [P, C.POINTER(_Drum), P, C.c_char_p, C.c_size_t]
    L.mg_surface_polygons.argtypes = [P, C.c_uint32, P, C.c_int64]
    L.mg_surface_polygons.restype = C.c_int64
    L.mg_surface_field.argtypes = [P, C.c_uint32, P]
    L._mg_surface_ready = True
    return L


def unflatten(flat):
    """flatten_polygons' layout (surface.rs:420-430) back into a list of (k, 2) arrays."""
    flat = np.asarray(flat, np.float32)
    polys, k = [], 1
    for _ in range(int(flat[0])):
        n = int(flat[k])
        polys.append(flat[k + 1:k + 1 + 2 * n].reshape(n, 2).copy())
        k += 1 + 2 * n
    return polys


class SurfaceBatch:
    def __init__(self, n_drums: int, max_fluid: int):
        self._L = _lib()
        self._h = C.c_void_p()
        err = C.create_string_buffer(256)
        rc = self._L.mg_surface_create(int(n_drums), int(max_fluid), C.byref(self._h), err, 256)
        if rc != 0:
            raise (ValueError if rc == -4 else RuntimeError)(f"mg_surface_create: {err.value.decode()}")
        self.n = int(n_drums)
        self._dx = None

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self._L.mg_surface_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        self.close()

    def upload(self, drums):
        """Packs the drums' fluid positions into device memory; returns the descriptor array."""
        if len(drums) != self.n:
            raise ValueError("one dict per drum")
        recs = (_Drum * self.n)()
        xs = []
        for k, d in enumerate(drums):
            r = recs[k]
            r.radius_m, r.drum_angle = float(d["radius_m"]), float(d.get("drum_angle", 0.0))
            li = d.get("lifters") or {}
            r.lifter_count = int(li.get("count", 0))
            r.lifter_height_m = float(li.get("height_m", 0.0))
            r.lifter_base_width_m = float(li.get("base_width_m", 0.0))
            r.lifter_top_width_m = float(li.get("top_width_m", 0.0))
            r.lifter_phase_deg = float(li.get("phase_deg", 0.0))
            r.h, r.rest_density, r.particle_mass = float(d["h"]), float(d["rest_density"]), float(d["particle_mass"])
            x = np.asarray(d["fluid_x"], np.float32).reshape(-1, 2)
            r.n_fluid = len(x)
            xs.append(x)
        flat = np.ascontiguousarray(np.concatenate(xs)) if xs else np.zeros((0, 2), np.float32)
        if self._dx is not None:
            self._dx.free()
        self._dx = DeviceArray.from_numpy(flat)
        return recs

    def run(self, recs):
        """mg_surface_extract over the uploaded positions (the timed part of the bench)."""
        err = C.create_string_buffer(512)
        rc = self._L.mg_surface_extract(self._h, recs, self._dx.ptr, err, 512)
        if rc != 0:
            raise (ValueError if rc == -4 else RuntimeError)(f"mg_surface_extract: {err.value.decode()}")

    def polygons(self, d: int):
        n = self._L.mg_surface_polygons(self._h, int(d), None, 0)
        if n < 0:
            raise ValueError(f"mg_surface_polygons: error {n}")
        buf = np.zeros(int(n), np.float32)
        self._L.mg_surface_polygons(self._h, int(d), buf.ctypes.data_as(C.c_void_p), int(n))
        return unflatten(buf)

    def field(self, d: int):
        out = np.zeros(G * G, np.float32)
        rc = self._L.mg_surface_field(self._h, int(d), out.ctypes.data_as(C.c_void_p))
        if rc != 0:
            raise RuntimeError(f"mg_surface_field: error {rc}")
        return out.reshape(G, G)

    def extract(self, drums):
        """extract_surface for every drum: a list (per drum) of lists of (k, 2) polygons."""
        self.run(self.upload(drums))
        return [self.polygons(d) for d in range(self.n)]
