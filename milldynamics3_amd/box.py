"""Periodic-box PBF workload (kernel isolation of the fluid hot path) over ``include/millgpu.h``.

``Box(cfg)`` holds ``side**2`` particles on the GPU. One ``step()`` applies the reference's fluid
arithmetic (crates/mill-core/src/pbf.rs: predict :1225-1235, neighbour lists :262-288, density
passes :1257-1334, velocity :1656-1674, density recompute :2136-2144, implicit viscosity CG
:312-480) in a periodic box with minimum-image pair vectors. The compute path is the HIP library;
there is no CPU path.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import sim as _sim

BOX_DEFAULTS = {"side": 1024, "dx": 1.0e-3, "rest": 1000.0, "mu": 1.0e-3, "dt": 1.0 / 1440.0,
                "passes": 4, "tol": 1.0e-6, "max_cg": 500, "seed": 1, "fp64": 0, "yield_stress": 0.0,
                "adaptive": 0, "gravity": 0.0, "speed_clamp": 0.0, "energy_stats": 0,
                "density_ratio": 0.0}


class BoxConfig(C.Structure):
    _fields_ = [("side", C.c_uint32), ("dx", C.c_float), ("rest", C.c_float), ("mu", C.c_float),
                ("dt", C.c_float), ("passes", C.c_uint32), ("tol", C.c_float), ("max_cg", C.c_uint32),
                ("seed", C.c_uint64), ("fp64", C.c_uint32), ("yield_stress", C.c_float),
                ("adaptive", C.c_uint32), ("gravity", C.c_float), ("speed_clamp", C.c_float),
                ("energy_stats", C.c_uint32), ("density_ratio", C.c_float)]


def box_config(overrides=None) -> dict:
    c = dict(BOX_DEFAULTS)
    c.update(overrides or {})
    return c


def _cfg_struct(cfg) -> BoxConfig:
    c = box_config(cfg)
    return BoxConfig(int(c["side"]), float(c["dx"]), float(c["rest"]), float(c["mu"]), float(c["dt"]),
                     int(c["passes"]), float(c["tol"]), int(c["max_cg"]), int(c["seed"]), int(c["fp64"]),
                     float(c["yield_stress"]), int(c["adaptive"]), float(c["gravity"]), float(c["speed_clamp"]),
                     int(c["energy_stats"]), float(c["density_ratio"]))


_ready = set()


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _lib():
    L = _sim.lib()
    if id(L) in _ready:
        return L
    P = C.c_void_p
    fp = C.POINTER(C.c_float)
    dp = C.POINTER(C.c_double)
    u32p = C.POINTER(C.c_uint32)
    cp = C.POINTER(BoxConfig)
    L.mg_box_geometry.argtypes = [cp, dp]
    L.mg_box_create.argtypes = [cp, C.POINTER(P), C.c_char_p, C.c_size_t]
    L.mg_box_destroy.argtypes = [P]
    L.mg_box_count.argtypes = [P]
    L.mg_box_count.restype = C.c_size_t
    L.mg_box_step.argtypes = [P, C.c_uint32]
    L.mg_box_last_step_ms.argtypes = [P]
    L.mg_box_last_step_ms.restype = C.c_float
    L.mg_box_copy_state.argtypes = [P, fp, fp]
    L.mg_box_set_state.argtypes = [P, fp, fp]
    L.mg_box_density.argtypes = [P, fp]
    L.mg_box_copy_state_f64.argtypes = [P, dp, dp]
    L.mg_box_set_state_f64.argtypes = [P, dp, dp]
    L.mg_box_density_f64.argtypes = [P, dp]
    L.mg_box_stats.argtypes = [P, dp]
    L.mg_box_neighbors.argtypes = [P, u32p, u32p, C.c_int64]
    L.mg_box_neighbors.restype = C.c_int64
    L.mg_box_set_timing.argtypes = [P, C.c_uint32]
    L.mg_box_total_cg_iterations.argtypes = [P]
    L.mg_box_total_cg_iterations.restype = C.c_uint64
    L.mg_box_set_cg_prediction.argtypes = [P, C.c_uint32]
    L.mg_box_kernel_times.argtypes = [P, dp, C.POINTER(C.c_uint64), C.c_int]
    L.mg_box_replay_ms.argtypes = [P, C.c_uint32, C.c_uint32, dp]
    L.mg_box_mean_shear_rate.argtypes = [P, dp]
    L.mg_box_density_passes.argtypes = [P]
    L.mg_box_density_passes.restype = C.c_uint32
    L.mg_box_energy.argtypes = [P, dp]
    _ready.add(id(L))
    return L


def box_geometry(cfg=None) -> dict:
    out = np.zeros(5)
    s = _cfg_struct(cfg)
    _sim._check(_lib().mg_box_geometry(C.byref(s), out.ctypes.data_as(C.POINTER(C.c_double))), what="geometry")
    return {"h": out[0], "L": out[1], "cw": out[2], "mass": out[3], "ncx": int(out[4])}


class Box:
    """``side**2`` PBF particles in a periodic box, resident on the GPU (``fp64=1``: FP64 state;
    its state and density come back as float64, exactly)."""

    def __init__(self, cfg=None):
        self.cfg = box_config(cfg)
        self._L = _lib()
        self._h = C.c_void_p()
        err = C.create_string_buffer(512)
        s = _cfg_struct(self.cfg)
        _sim._check(self._L.mg_box_create(C.byref(s), C.byref(self._h), err, 512), err, "mg_box_create")
        self.n = int(self._L.mg_box_count(self._h))
        self.fp64 = bool(self.cfg["fp64"])

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self._L.mg_box_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        self.close()

    def step(self, n: int = 1):
        _sim._check(self._L.mg_box_step(self._h, n), what="mg_box_step")

    @property
    def total_cg_iterations(self) -> int:
        return int(self._L.mg_box_total_cg_iterations(self._h))

    def set_cg_prediction(self, iterations: int):
        """CG iteration count the next multi-step window assumes (see include/millgpu.h)."""
        _sim._check(self._L.mg_box_set_cg_prediction(self._h, iterations), what="set_cg_prediction")

    @property
    def last_ms(self) -> float:
        return float(self._L.mg_box_last_step_ms(self._h))

    def set_timing(self, class_mask: int):
        """HIP-event timing of the kernel classes in class_mask (see include/millgpu.h)."""
        _sim._check(self._L.mg_box_set_timing(self._h, class_mask), what="set_timing")

    def kernel_times(self):
        """(total ms, launches) per kernel class since the last set_timing."""
        ms = np.zeros(8)
        n = np.zeros(8, dtype=np.uint64)
        _sim._check(self._L.mg_box_kernel_times(self._h, ms.ctypes.data_as(C.POINTER(C.c_double)),
                                                n.ctypes.data_as(C.POINTER(C.c_uint64)), 8), what="kernel_times")
        return ms, n

    def replay_ms(self, kernel_class: int, reps: int) -> float:
        """Average ms of one launch of kernel class 1 (density pass A) or 2 (density pass B): the
        last step's last launch of the class run again `reps` times back to back (state unchanged)."""
        ms = C.c_double()
        _sim._check(self._L.mg_box_replay_ms(self._h, kernel_class, reps, C.byref(ms)), what="replay_ms")
        return ms.value

    def state(self):
        if self.fp64:
            x = np.zeros(2 * self.n)
            v = np.zeros(2 * self.n)
            _sim._check(self._L.mg_box_copy_state_f64(self._h, _dp(x), _dp(v)), what="copy_state_f64")
        else:
            x = np.zeros(2 * self.n, dtype=np.float32)
            v = np.zeros(2 * self.n, dtype=np.float32)
            _sim._check(self._L.mg_box_copy_state(self._h, _sim._fp(x), _sim._fp(v)), what="copy_state")
        return x.reshape(self.n, 2), v.reshape(self.n, 2)

    def set_state(self, x, v):
        dt = np.float64 if self.fp64 else np.float32
        x = np.ascontiguousarray(x, dtype=dt).reshape(-1)
        v = np.ascontiguousarray(v, dtype=dt).reshape(-1)
        if x.size != 2 * self.n or v.size != 2 * self.n:
            raise ValueError("state arrays must hold 2*n values")
        if self.fp64:
            _sim._check(self._L.mg_box_set_state_f64(self._h, _dp(x), _dp(v)), what="set_state_f64")
        else:
            _sim._check(self._L.mg_box_set_state(self._h, _sim._fp(x), _sim._fp(v)), what="set_state")

    def density(self) -> np.ndarray:
        if self.fp64:
            d = np.zeros(self.n)
            _sim._check(self._L.mg_box_density_f64(self._h, _dp(d)), what="density_f64")
        else:
            d = np.zeros(self.n, dtype=np.float32)
            _sim._check(self._L.mg_box_density(self._h, _sim._fp(d)), what="density")
        return d

    def mean_shear_rate(self) -> float:
        """Mean shear rate (1/s) of the last step (see include/millgpu.h), 0 before the first step."""
        out = C.c_double()
        _sim._check(self._L.mg_box_mean_shear_rate(self._h, C.byref(out)), what="mean_shear_rate")
        return out.value

    def density_passes(self) -> int:
        """Density passes the last step ran."""
        return int(self._L.mg_box_density_passes(self._h))

    def energy(self) -> dict:
        """Kinetic energies of the last step (needs energy_stats=1): before the CG, after it, after the clamp."""
        o = np.zeros(3)
        _sim._check(self._L.mg_box_energy(self._h, _dp(o)), what="energy")
        return {"ke_before_cg": o[0], "ke_after_cg": o[1], "ke_after_clamp": o[2]}

    def stats(self) -> dict:
        o = np.zeros(3)
        _sim._check(self._L.mg_box_stats(self._h, o.ctypes.data_as(C.POINTER(C.c_double))), what="stats")
        return {"cg_iterations": int(o[0]), "residual": o[1], "b_norm": o[2]}

    def neighbors(self):
        off = np.zeros(self.n + 1, dtype=np.uint32)
        u32p = C.POINTER(C.c_uint32)
        e = self._L.mg_box_neighbors(self._h, off.ctypes.data_as(u32p), None, 0)
        if e < 0:
            _sim._check(int(e), what="neighbors")
        nb = np.zeros(max(int(e), 1), dtype=np.uint32)
        e2 = self._L.mg_box_neighbors(self._h, off.ctypes.data_as(u32p), nb.ctypes.data_as(u32p), int(e))
        if e2 != e:
            raise RuntimeError("neighbour export changed size")
        return off, nb[: int(e)]
