"""On-demand drum metrics on the GPU over ``include/millgpu.h`` (crates/mill-core/src/metrics.rs).

``mixing_index(points, dye, radius)`` mirrors ``metrics::mixing_index`` (metrics.rs:506-576) and
returns None where the reference does; ``mixing_index_batch`` evaluates many drums in one launch.
Inputs are copied into device memory of the library's own HIP runtime; there is no CPU path.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import sim as _sim
from .grid import DeviceArray, _lib as _grid_lib

_ready = set()


def _lib():
    L = _grid_lib()
    if id(L) in _ready:
        return L
    P = C.c_void_p
    L.mg_mixing_index.argtypes = [P, P, C.c_uint32, C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_int)]
    L.mg_mixing_index_batch.argtypes = [P, P, P, P, C.c_uint32, P, P, P]
    _ready.add(id(L))
    return L


def mixing_index(points, dye, radius: float):
    """metrics::mixing_index of one drum's fluid: the Lacey index, or None."""
    L = _lib()
    pts = np.ascontiguousarray(np.asarray(points, dtype=np.float32).reshape(-1, 2))
    d = np.ascontiguousarray(np.asarray(dye, dtype=np.float32).reshape(-1))
    if d.size != pts.shape[0]:
        raise ValueError("one dye value per particle")
    dx, dd = DeviceArray.from_numpy(pts), DeviceArray.from_numpy(d)
    out, ok = C.c_float(), C.c_int()
    _sim._check(L.mg_mixing_index(dx.ptr, dd.ptr, d.size, float(radius), C.byref(out), C.byref(ok)),
                what="mg_mixing_index")
    return float(out.value) if ok.value else None


def mixing_index_batch(drums):
    """Mixing indices of several drums in one launch: drums = [(points, dye, radius), ...]."""
    L = _lib()
    pts = [np.asarray(p, dtype=np.float32).reshape(-1, 2) for p, _, _ in drums]
    dye = [np.asarray(d, dtype=np.float32).reshape(-1) for _, d, _ in drums]
    off = np.zeros(len(drums) + 1, dtype=np.uint32)
    off[1:] = np.cumsum([len(d) for d in dye])
    rad = np.array([r for _, _, r in drums], dtype=np.float32)
    allp = np.concatenate(pts) if pts else np.zeros((0, 2), np.float32)
    alld = np.concatenate(dye) if dye else np.zeros(0, np.float32)
    dx, dd, doff, drad = (DeviceArray.from_numpy(a) for a in (allp, alld, off, rad))
    dout, dok = DeviceArray(4 * len(drums)), DeviceArray(4 * len(drums))
    _sim._check(L.mg_mixing_index_batch(dx.ptr, dd.ptr, doff.ptr, drad.ptr, len(drums), dout.ptr, dok.ptr, None),
                what="mg_mixing_index_batch")
    out = dout.to_numpy(np.float32, len(drums))
    ok = dok.to_numpy(np.int32, len(drums))
    return [float(v) if k else None for v, k in zip(out, ok)]


class _MetricsDrum(C.Structure):
    _fields_ = [("radius_m", C.c_float), ("omega", C.c_float), ("drum_angle", C.c_float),
                ("lifter_count", C.c_uint32), ("lifter_height_m", C.c_float), ("lifter_base_width_m", C.c_float),
                ("lifter_top_width_m", C.c_float), ("lifter_phase_deg", C.c_float), ("ball_radius", C.c_float),
                ("ball_mass", C.c_float), ("ball_inertia", C.c_float), ("h", C.c_float),
                ("rest_density", C.c_float), ("n_balls", C.c_uint32), ("n_fluid", C.c_uint32)]


class _DrumMetrics(C.Structure):
    _fields_ = [("toe_angle_rad", C.c_float), ("shoulder_angle_rad", C.c_float), ("has_toe_shoulder", C.c_int32),
                ("centroid_x", C.c_float), ("centroid_y", C.c_float), ("has_centroid", C.c_int32),
                ("pool_angle_min_rad", C.c_float), ("pool_angle_max_rad", C.c_float),
                ("has_pool_extent", C.c_int32), ("free_surface_angle_rad", C.c_float),
                ("free_surface_offset_m", C.c_float), ("has_free_surface", C.c_int32),
                ("pool_depth_m", C.c_float), ("has_pool_depth", C.c_int32), ("mixing_index", C.c_float),
                ("has_mixing_index", C.c_int32), ("total_kinetic_energy_j", C.c_float),
                ("max_ball_overlap_fraction", C.c_float), ("max_ball_wall_overlap_fraction", C.c_float),
                ("max_fluid_density_error_fraction", C.c_float), ("mean_fluid_density_error_fraction", C.c_float),
                ("max_fluid_compression_error_fraction", C.c_float),
                ("mean_fluid_compression_error_fraction", C.c_float), ("has_density_errors", C.c_int32),
                ("drum_omega_rad_s", C.c_float), ("fluid_particle_count", C.c_uint32)]


def _metrics_lib():
    L = _lib()
    if getattr(L, "_mg_drum_metrics_ready", False):
        return L
    P = C.c_void_p
    L.mg_drum_metrics_batch.argtypes = [C.POINTER(_MetricsDrum), C.c_uint32, P, P, P, P, P, P,
                                        C.POINTER(_DrumMetrics), C.c_char_p, C.c_size_t]
    fp = C.POINTER(C.c_float)
    L.mg_atan2f_ref.argtypes = [fp, fp, fp, C.c_uint64]
    L.mg_atan2f_ref.restype = None
    L._mg_drum_metrics_ready = True
    return L


def atan2f_ref(y, x):
    """The library's atan2f (the host libm's algorithm, which the metrics kernels use for wall
    angles) evaluated on the host: float32 arrays in, float32 out."""
    L = _metrics_lib()
    y = np.ascontiguousarray(np.asarray(y, dtype=np.float32).reshape(-1))
    x = np.ascontiguousarray(np.asarray(x, dtype=np.float32).reshape(-1))
    if y.size != x.size:
        raise ValueError("y and x differ in length")
    out = np.zeros_like(y)
    fp = C.POINTER(C.c_float)
    L.mg_atan2f_ref(y.ctypes.data_as(fp), x.ctypes.data_as(fp), out.ctypes.data_as(fp), y.size)
    return out


_DRUM_KEYS = ("radius_m", "omega", "drum_angle", "ball_radius", "ball_mass", "ball_inertia", "h", "rest_density")


def drum_metrics_batch(drums):
    """metrics::compute (metrics.rs:195-254) of several drum states in one call. Each drum is a dict:
    radius_m, omega, drum_angle, lifters ({count, height_m, base_width_m, top_width_m, phase_deg},
    optional), ball_radius, ball_mass, ball_inertia, h, rest_density, ball_x (n, 2), ball_v (n, 2),
    ball_omega (n,), fluid_x (m, 2), dye (m,), density (m,). Returns one dict per drum with the
    Metrics fields; a field the reference reports as None is None."""
    L = _metrics_lib()
    n = len(drums)
    recs = (_MetricsDrum * max(n, 1))()
    bx, bv, bw, fx, fd, fr = [], [], [], [], [], []
    for k, d in enumerate(drums):
        r = recs[k]
        for key in _DRUM_KEYS:
            setattr(r, key, float(d[key]))
        li = d.get("lifters") or {}
        r.lifter_count = int(li.get("count", 0))
        r.lifter_height_m = float(li.get("height_m", 0.0))
        r.lifter_base_width_m = float(li.get("base_width_m", 0.0))
        r.lifter_top_width_m = float(li.get("top_width_m", 0.0))
        r.lifter_phase_deg = float(li.get("phase_deg", 0.0))
        b = np.asarray(d["ball_x"], dtype=np.float32).reshape(-1, 2)
        f = np.asarray(d["fluid_x"], dtype=np.float32).reshape(-1, 2)
        r.n_balls, r.n_fluid = len(b), len(f)
        bx.append(b)
        bv.append(np.asarray(d["ball_v"], dtype=np.float32).reshape(-1, 2))
        bw.append(np.asarray(d["ball_omega"], dtype=np.float32).reshape(-1))
        fx.append(f)
        fd.append(np.asarray(d["dye"], dtype=np.float32).reshape(-1))
        fr.append(np.asarray(d["density"], dtype=np.float32).reshape(-1))
        if len(bv[-1]) != len(b) or len(bw[-1]) != len(b) or len(fd[-1]) != len(f) or len(fr[-1]) != len(f):
            raise ValueError(f"drum {k}: array lengths differ")

    def dev(parts, cols):
        a = np.concatenate(parts) if parts else np.zeros((0, cols) if cols > 1 else 0, np.float32)
        return DeviceArray.from_numpy(a.astype(np.float32))
    arrays = [dev(bx, 2), dev(bv, 2), dev(bw, 1), dev(fx, 2), dev(fd, 1), dev(fr, 1)]
    out = (_DrumMetrics * max(n, 1))()
    err = C.create_string_buffer(512)
    rc = L.mg_drum_metrics_batch(recs, n, *[a.ptr for a in arrays], out, err, 512)
    if rc != 0:
        raise (ValueError if rc == -4 else RuntimeError)(f"mg_drum_metrics_batch: {err.value.decode()}")
    res = []
    pairs = (("has_toe_shoulder", ("toe_angle_rad", "shoulder_angle_rad")),
             ("has_centroid", ("centroid_x", "centroid_y")),
             ("has_pool_extent", ("pool_angle_min_rad", "pool_angle_max_rad")),
             ("has_free_surface", ("free_surface_angle_rad", "free_surface_offset_m")),
             ("has_pool_depth", ("pool_depth_m",)), ("has_mixing_index", ("mixing_index",)),
             ("has_density_errors", ("max_fluid_density_error_fraction", "mean_fluid_density_error_fraction",
                                     "max_fluid_compression_error_fraction", "mean_fluid_compression_error_fraction")))
    for k in range(n):
        o = out[k]
        d = {name: getattr(o, name) for name, _ in _DrumMetrics._fields_ if not name.startswith("has_")}
        for flag, names in pairs:
            if not getattr(o, flag):
                for nm in names:
                    d[nm] = None
        res.append(d)
    return res
