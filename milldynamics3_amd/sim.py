"""Python host mirror of the reference's ``mill_core::Simulation`` API over the C ABI.

Names, argument meanings and error behaviour follow crates/mill-core/src/lib.rs:82-566 (and the
WASM binding crates/mill-wasm/src/lib.rs:15-204): ``Simulation(params)`` raises ``ValueError``
with the reference's ``Params::validate`` message for invalid parameters, ``step(dt)`` and
``step_fixed()`` advance the coupled slurry-plus-media state, and ``balls()`` / ``fluid()``
return copies of the state arrays.

The compute path is ``libmillgpu.so`` (HIP kernels for gfx950). There is no CPU fallback: if the
library is missing or no GPU is present, construction raises.
"""
from __future__ import annotations

import copy
import ctypes as C
import json
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libmillgpu.so")
_lib = None

# Params::default() (crates/mill-core/src/params.rs:62-73, :200-228, :351-376, :419-429, :486-501)
DEFAULT_PARAMS = {
    "mill": {"diameter_m": 0.063, "speed_mode": "rpm", "speed_value": 100.0,
             "direction": "counter_clockwise"},
    "media": {"ball_diameter_m": 0.002, "fill_fraction": 0.50, "packing_fraction_2d": 0.82,
              "density_kg_m3": 6000.0, "restitution_ball_ball": 0.7, "restitution_ball_wall": 0.5,
              "friction_ball_ball": 0.25, "friction_ball_wall": 0.35,
              "friction_ball_ball_static": 0.25, "friction_ball_wall_static": 0.35,
              "friction_velocity_scale_m_s": 0.02, "rolling_friction": 0.01},
    "slurry": {"enabled": True, "fill_fraction": 0.35, "density_kg_m3": 1000.0,
               "viscosity_pa_s": 0.001, "rheology": "newtonian", "yield_stress_pa": 0.0,
               "wall_no_slip": 1.0, "ball_no_slip": 1.0, "wettability": 0.6,
               "surface_tension_n_m": 0.072, "dye_pattern": "left_right"},
    "lifters": {"count": 0, "height_m": 0.020, "base_width_m": 0.030, "top_width_m": 0.020,
                "phase_deg": 0.0},
    "simulation": {"resolution": 40, "substeps": 8, "pbf_iterations": 3, "dem_iterations": 2,
                   "time_scale": 1.0, "frame_budget_ms": 12.0, "max_balls": 600, "seed": 1,
                   "coarse_graining_mode": "off", "coarse_graining_k": 1.0},
}


def params_json(overrides=None) -> str:
    """Full Params JSON: the defaults with ``overrides`` (a nested dict) merged in."""
    def merge(base, over):
        out = copy.deepcopy(base)
        for k, v in (over or {}).items():
            out[k] = merge(out[k], v) if isinstance(v, dict) and isinstance(out.get(k), dict) else v
        return out
    if isinstance(overrides, str):
        return overrides
    return json.dumps(merge(DEFAULT_PARAMS, overrides or {}))


def lib():
    """Loads libmillgpu.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        raise RuntimeError(f"{_LIB_PATH} is missing: run milldynamics3_amd.build.build() first "
                           "(there is no CPU fallback)")
    L = C.CDLL(_LIB_PATH)
    P, ch, sz = C.c_void_p, C.c_char_p, C.c_size_t
    fp = C.POINTER(C.c_float)
    dp = C.POINTER(C.c_double)
    u32p = C.POINTER(C.c_uint32)
    L.mg_drum_engine_enabled.argtypes = []
    L.mg_drum_engine_enabled.restype = C.c_int
    L.mg_sim_create.argtypes = [ch, C.POINTER(P), ch, sz]
    L.mg_sim_create_with_balls.argtypes = [ch, fp, sz, C.POINTER(P), ch, sz]
    L.mg_sim_destroy.argtypes = [P]
    L.mg_sim_set_params.argtypes = [P, ch, ch, sz]
    L.mg_sim_step.argtypes = [P, C.c_float]
    L.mg_sim_step_fixed.argtypes = [P]
    L.mg_sim_step_fixed_n.argtypes = [P, C.c_uint32]
    L.mg_sim_reset_frame_stats.argtypes = [P]
    L.mg_sim_fixed_sub_dt.argtypes = [P]
    L.mg_sim_fixed_sub_dt.restype = C.c_float
    L.mg_sim_effective_substeps.argtypes = [P]
    L.mg_sim_effective_substeps.restype = C.c_uint32
    for f in ("mg_sim_ball_count", "mg_sim_fluid_count"):
        getattr(L, f).argtypes = [P]
        getattr(L, f).restype = sz
    L.mg_sim_copy_balls.argtypes = [P, fp, fp, fp, fp]
    L.mg_sim_copy_fluid.argtypes = [P, fp, fp, fp]
    L.mg_sim_set_state.argtypes = [P, fp, fp, fp, fp, fp, fp, fp, C.c_int]
    L.mg_sim_copy_guess.argtypes = [P, fp, C.POINTER(C.c_int)]
    L.mg_sim_set_clock.argtypes = [P, C.c_float, C.c_double]
    L.mg_sim_drum_angle.argtypes = [P]
    L.mg_sim_drum_angle.restype = C.c_float
    L.mg_sim_sim_time.argtypes = [P]
    L.mg_sim_sim_time.restype = C.c_double
    for f, rt in (("mg_sim_viscosity_iterations", C.c_uint32), ("mg_sim_mean_shear_rate_per_s", C.c_float),
                  ("mg_sim_power_draw_w", C.c_float), ("mg_sim_torque_nm", C.c_float),
                  ("mg_sim_collision_rate_per_s", C.c_float), ("mg_sim_dissipated_power_w", C.c_float),
                  ("mg_sim_max_substep_displacement_over_diameter", C.c_float),
                  ("mg_sim_coupling_clamp_hits", C.c_uint32), ("mg_sim_error_bits", C.c_uint32)):
        getattr(L, f).argtypes = [P]
        getattr(L, f).restype = rt
    L.mg_sim_stats.argtypes = [P, dp, sz]
    L.mg_sim_stress_state.argtypes = [P, dp]
    L.mg_sim_impacts.argtypes = [P, dp, sz, C.POINTER(sz)]
    L.mg_sim_metrics_json.argtypes = [P, ch, sz]
    L.mg_sim_mixing_index.argtypes = [P, fp, C.POINTER(C.c_int)]
    L.mg_sim_energy_budget.argtypes = [P, dp]
    L.mg_sim_reset_energy_budget.argtypes = [P]
    L.mg_sim_set_stress_debounce_factor.argtypes = [P, C.c_double]
    L.mg_sim_save_state.argtypes = [P, C.c_void_p, sz, C.POINTER(sz)]
    L.mg_sim_load_state.argtypes = [P, C.c_void_p, sz]
    L.mg_sim_force_iterations.argtypes = [P, C.c_uint32, C.c_uint32]
    L.mg_batch_create.argtypes = [C.POINTER(ch), C.c_uint32, C.POINTER(P), ch, sz]
    L.mg_batch_destroy.argtypes = [P]
    L.mg_batch_size.argtypes = [P]
    L.mg_batch_size.restype = C.c_uint32
    L.mg_batch_step.argtypes = [P, C.c_uint32, C.c_int]
    L.mg_batch_drum.argtypes = [P, C.c_uint32]
    L.mg_batch_drum.restype = P
    L.mg_batch_sync.argtypes = [P, ch, sz]
    L.mg_batch_last_step_ms.argtypes = [P]
    L.mg_batch_last_step_ms.restype = C.c_float
    L.mg_batch_set_timing.argtypes = [P, C.c_int]
    L.mg_batch_kernel_times.argtypes = [P, C.c_char_p, C.c_size_t]
    L.mg_version.restype = ch
    _lib = L
    return L


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float)) if a is not None else None


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _check(rc, err=None, what="call"):
    if rc != 0:
        msg = err.value.decode() if err is not None and err.value else f"{what} failed ({rc})"
        if rc == -1:
            raise ValueError(msg)
        if rc == -5:
            raise RuntimeError(f"{what}: the drum substep engine is switched off in this build "
                               "(an earlier form faulted on the GPU; cause not yet shown)")
        raise RuntimeError(msg)


@dataclass
class Balls:
    """dem::Balls (dem.rs:277-309) as array copies."""
    x: np.ndarray
    v: np.ndarray
    theta: np.ndarray
    omega: np.ndarray


@dataclass
class Fluid:
    """pbf::FluidParticles (pbf.rs:1018-1028) as array copies."""
    x: np.ndarray
    v: np.ndarray
    dye: np.ndarray


STAT_KEYS = ["drum_angle", "sim_time", "power_draw_w", "torque_nm", "collision_rate_per_s",
             "dissipated_power_w", "max_substep_displacement_over_diameter", "coupling_clamp_hits",
             "viscosity_iterations", "density_iterations", "mean_shear_rate_per_s",
             "fluid_wall_work_j", "dem_wall_work_j", "dem_dissipated_energy_j", "dem_coupling_work_j",
             "collision_count", "n_groups", "n_links", "links_in_range", "n_ball_pairs",
             "ke_delta_noslip_j", "ke_delta_drag_j", "ke_delta_buoyancy_j", "ke_delta_viscosity_j",
             "ke_delta_clamp_j", "wall_impulse_x", "wall_impulse_y", "wall_angular_impulse",
             "error_bits", "n_impacts"] + [f"impact_bin_{k}" for k in range(12)] + ["projection_restarts"]
ENERGY_KEYS = ["elapsed_s", "shaft_work_j", "dem_wall_work_j", "fluid_wall_work_j",
               "ball_contact_dissipation_j", "fluid_wall_slip_j", "fluid_viscous_j",
               "fluid_clamp_removed_j", "interface_created_j", "delta_mechanical_j", "unattributed_j",
               "fluid_wall_angular_impulse"]


class Simulation:
    """mill_core::Simulation (lib.rs:82-566) backed by the MI355X engine."""

    def __init__(self, params=None, initial_balls=None, _handle=None, _owner=None):
        L = lib()
        self._L = L
        self._owner = _owner
        if _handle is not None:
            self._h = _handle
            self._owned = False
        else:
            err = C.create_string_buffer(512)
            h = C.c_void_p()
            pj = None if params is None else params_json(params).encode()
            if initial_balls is None:
                rc = L.mg_sim_create(pj, C.byref(h), err, 512)
            else:
                ib = np.ascontiguousarray(np.asarray(initial_balls, dtype=np.float32).reshape(-1))
                rc = L.mg_sim_create_with_balls(pj, _fp(ib), ib.size // 2, C.byref(h), err, 512)
            _check(rc, err, "Simulation::new")
            self._h = h
            self._owned = True
        self.n_balls = int(L.mg_sim_ball_count(self._h))
        self.n_fluid = int(L.mg_sim_fluid_count(self._h))

    @classmethod
    def with_initial_balls(cls, params, positions):
        return cls(params, initial_balls=positions)

    def close(self):
        if getattr(self, "_owned", False) and self._h:
            self._L.mg_sim_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- stepping ---------------------------------------------------------------------------
    def set_params(self, params):
        err = C.create_string_buffer(512)
        _check(self._L.mg_sim_set_params(self._h, params_json(params).encode(), err, 512), err, "set_params")

    def step(self, dt: float):
        _check(self._L.mg_sim_step(self._h, dt), what="step")

    def step_fixed(self, n: int = 1):
        _check(self._L.mg_sim_step_fixed_n(self._h, n), what="step_fixed")

    def reset_frame_stats(self):
        self._L.mg_sim_reset_frame_stats(self._h)

    def fixed_sub_dt(self) -> float:
        return float(self._L.mg_sim_fixed_sub_dt(self._h))

    def effective_substeps(self) -> int:
        return int(self._L.mg_sim_effective_substeps(self._h))

    # -- state --------------------------------------------------------------------------------
    def balls(self) -> Balls:
        n = self.n_balls
        x, v = np.zeros(2 * n, np.float32), np.zeros(2 * n, np.float32)
        th, om = np.zeros(n, np.float32), np.zeros(n, np.float32)
        _check(self._L.mg_sim_copy_balls(self._h, _fp(x), _fp(v), _fp(th), _fp(om)), what="copy_balls")
        return Balls(x.reshape(-1, 2), v.reshape(-1, 2), th, om)

    def fluid(self) -> Fluid:
        n = self.n_fluid
        x, v, d = np.zeros(2 * n, np.float32), np.zeros(2 * n, np.float32), np.zeros(n, np.float32)
        _check(self._L.mg_sim_copy_fluid(self._h, _fp(x), _fp(v), _fp(d)), what="copy_fluid")
        return Fluid(x.reshape(-1, 2), v.reshape(-1, 2), d)

    def viscous_guess(self):
        g = np.zeros(2 * self.n_fluid, np.float32)
        hg = C.c_int(0)
        self._L.mg_sim_copy_guess(self._h, _fp(g), C.byref(hg))
        return g.reshape(-1, 2), bool(hg.value)

    def set_state(self, ball_x, ball_v, ball_theta, ball_omega, fluid_x, fluid_v, guess=None,
                  has_guess=False):
        f = lambda a: None if a is None else np.ascontiguousarray(np.asarray(a, np.float32).reshape(-1))
        arrs = [f(ball_x), f(ball_v), f(ball_theta), f(ball_omega), f(fluid_x), f(fluid_v), f(guess)]
        _check(self._L.mg_sim_set_state(self._h, *[_fp(a) for a in arrs], 1 if has_guess else 0),
               what="set_state")

    def set_clock(self, drum_angle: float, sim_time: float):
        self._L.mg_sim_set_clock(self._h, drum_angle, sim_time)

    def save_state(self) -> bytes:
        need = C.c_size_t(0)
        self._L.mg_sim_save_state(self._h, None, 0, C.byref(need))
        buf = C.create_string_buffer(need.value)
        _check(self._L.mg_sim_save_state(self._h, buf, need.value, C.byref(need)), what="save_state")
        return buf.raw

    def load_state(self, blob: bytes):
        _check(self._L.mg_sim_load_state(self._h, blob, len(blob)), what="load_state")

    # -- getters (lib.rs:276-349) ------------------------------------------------------------------
    def drum_angle(self) -> float:
        return float(self._L.mg_sim_drum_angle(self._h))

    def sim_time(self) -> float:
        return float(self._L.mg_sim_sim_time(self._h))

    def viscosity_iterations(self) -> int:
        return int(self._L.mg_sim_viscosity_iterations(self._h))

    def mean_shear_rate_per_s(self) -> float:
        return float(self._L.mg_sim_mean_shear_rate_per_s(self._h))

    def power_draw_w(self) -> float:
        return float(self._L.mg_sim_power_draw_w(self._h))

    def torque_nm(self) -> float:
        return float(self._L.mg_sim_torque_nm(self._h))

    def collision_rate_per_s(self) -> float:
        return float(self._L.mg_sim_collision_rate_per_s(self._h))

    def dissipated_power_w(self) -> float:
        return float(self._L.mg_sim_dissipated_power_w(self._h))

    def max_substep_displacement_over_diameter(self) -> float:
        return float(self._L.mg_sim_max_substep_displacement_over_diameter(self._h))

    def coupling_clamp_hits(self) -> int:
        return int(self._L.mg_sim_coupling_clamp_hits(self._h))

    def error_bits(self) -> int:
        return int(self._L.mg_sim_error_bits(self._h))

    def stats(self) -> dict:
        o = np.zeros(64)
        n = self._L.mg_sim_stats(self._h, _dp(o), 64)
        d = dict(zip(STAT_KEYS, o[:n].tolist()))
        d["impact_energy_histogram"] = [int(d.pop(f"impact_bin_{k}")) for k in range(12)]
        return d

    def stress_state(self) -> np.ndarray:
        o = np.zeros(49)
        self._L.mg_sim_stress_state(self._h, _dp(o))
        return o

    def impacts(self) -> np.ndarray:
        cap = 8 * (self.n_balls + 16)
        o = np.zeros(6 * cap)
        n = C.c_size_t(0)
        self._L.mg_sim_impacts(self._h, _dp(o), cap, C.byref(n))
        return o[: 6 * min(n.value, cap)].reshape(-1, 6)

    def mixing_index(self):
        out, ok = C.c_float(0), C.c_int(0)
        self._L.mg_sim_mixing_index(self._h, C.byref(out), C.byref(ok))
        return float(out.value) if ok.value else None

    def metrics(self) -> dict:
        buf = C.create_string_buffer(16384)
        n = self._L.mg_sim_metrics_json(self._h, buf, 16384)
        if n < 0:
            raise RuntimeError("metrics buffer too small")
        return json.loads(buf.value.decode())

    def energy_budget(self) -> dict:
        o = np.zeros(12)
        _check(self._L.mg_sim_energy_budget(self._h, _dp(o)), what="energy_budget")
        return dict(zip(ENERGY_KEYS, o.tolist()))

    def reset_energy_budget(self):
        self._L.mg_sim_reset_energy_budget(self._h)

    def set_stress_debounce_factor(self, factor: float):
        self._L.mg_sim_set_stress_debounce_factor(self._h, factor)

    # -- parity/debug ---------------------------------------------------------------------------
    def force_iterations(self, density_passes: int = 0, cg_iterations: int = 0):
        self._L.mg_sim_force_iterations(self._h, density_passes, cg_iterations)


class Batch:
    """Many independent drums advanced together: every substep runs the same kernel launches for
    all drums (the fluid operator, the DEM operator, the stress update)."""

    def __init__(self, params_list):
        L = lib()
        self._L = L
        js = [params_json(p).encode() for p in params_list]
        arr = (C.c_char_p * len(js))(*js)
        err = C.create_string_buffer(1024)
        h = C.c_void_p()
        _check(L.mg_batch_create(arr, len(js), C.byref(h), err, 1024), err, "mg_batch_create")
        self._h = h
        self.n = len(js)
        self.drums = [Simulation(_handle=L.mg_batch_drum(h, i), _owner=self) for i in range(self.n)]

    def step(self, substeps: int, use_graph: bool = True):
        _check(self._L.mg_batch_step(self._h, substeps, 1 if use_graph else 0), what="mg_batch_step")

    def sync(self):
        err = C.create_string_buffer(512)
        _check(self._L.mg_batch_sync(self._h, err, 512), err, "mg_batch_sync")

    def last_step_ms(self) -> float:
        return float(self._L.mg_batch_last_step_ms(self._h))

    def set_timing(self, on: bool):
        """Per-kernel HIP-event timing of the following substeps (on clears the totals)."""
        _check(self._L.mg_batch_set_timing(self._h, 1 if on else 0), what="mg_batch_set_timing")

    def kernel_times(self) -> dict:
        """{kernel: (total_ms, launches)} since timing was switched on."""
        buf = C.create_string_buffer(1 << 16)
        n = self._L.mg_batch_kernel_times(self._h, buf, 1 << 16)
        if n < 0:
            raise RuntimeError("kernel time buffer too small")
        return {k: (float(v[0]), int(v[1])) for k, v in json.loads(buf.value.decode()).items()}

    def close(self):
        if getattr(self, "_h", None):
            self.drums = []
            self._L.mg_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def drum_engine_enabled() -> bool:
    """True if this build runs drum substeps on the GPU (``mg_drum_engine_enabled``)."""
    return bool(lib().mg_drum_engine_enabled())
