"""ctypes wrapper around ``oracle/_build/libmgoracle.so`` (built by ``oracle/Makefile``).

Test infrastructure only. ``default_params`` restates the serde defaults of the reference's
``Params`` (crates/mill-core/src/params.rs:62-73, :200-228, :351-376, :419-429, :486-501, plus the
``#[serde(default = ...)]`` fields at :162-175 and :339) so the oracle can be driven from the same
JSON the product's C-ABI accepts.
"""
from __future__ import annotations

import copy
import ctypes as C
import json
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libmgoracle.so")
_lib = None

MODES = {"ref": 0, "mixed": 1, "fp64": 2}


def oracle_lib():
    """Loads (building first if needed) the oracle shared library."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", _HERE])
    lib = C.CDLL(_LIB_PATH)
    P = C.c_void_p
    dp = C.POINTER(C.c_double)
    lib.orc_create.restype = P
    lib.orc_create.argtypes = [dp, C.c_uint64, C.c_int, C.POINTER(C.c_float), C.c_size_t]
    lib.orc_destroy.argtypes = [P]
    for fn in ("orc_sizes", "orc_scalars", "orc_energy_budget", "orc_stress", "orc_fluid_bins"):
        getattr(lib, fn).argtypes = [P, dp]
    lib.orc_step_fixed.argtypes = [P]
    lib.orc_step.argtypes = [P, C.c_double]
    lib.orc_advance.argtypes = [P, C.c_double]
    lib.orc_set_debug.argtypes = [P, C.c_int]
    lib.orc_force_iters.argtypes = [P, C.c_int, C.c_uint32, C.c_uint32]
    lib.orc_get_state.argtypes = [P] + [dp] * 8 + [C.POINTER(C.c_int)]
    lib.orc_set_state.argtypes = [P] + [dp] * 7 + [C.c_int]
    lib.orc_set_scalars.argtypes = [P, C.c_double, C.c_double]
    lib.orc_impacts.restype = C.c_int
    lib.orc_impacts.argtypes = [P, dp, C.c_int]
    lib.orc_mixing_index.restype = C.c_int
    lib.orc_mixing_index.argtypes = [P, dp]
    u32p = C.POINTER(C.c_uint32)
    i64p = C.POINTER(C.c_int64)
    lib.orc_dbg_grid.restype = C.c_int
    lib.orc_dbg_grid.argtypes = [P, i64p, u32p, C.c_int, u32p, C.c_int]
    lib.orc_dbg_nbrs.restype = C.c_int
    lib.orc_dbg_nbrs.argtypes = [P, u32p, C.c_int, u32p, C.c_int, i64p]
    fp = C.POINTER(C.c_float)
    lib.orc_grid_build.restype = C.c_int
    lib.orc_grid_build.argtypes = [fp, C.c_size_t, C.c_float, i64p, u32p, C.c_int, u32p, C.c_int]
    lib.orc_grid_pairs.restype = C.c_int64
    lib.orc_grid_pairs.argtypes = [fp, C.c_size_t, C.c_float, u32p, C.c_int64]
    lib.orc_grid_keys.restype = C.c_int64
    lib.orc_grid_keys.argtypes = [fp, C.c_size_t, C.c_float, C.POINTER(C.c_int32), C.c_int64]
    lib.orc_mixing_index_of.restype = C.c_int
    lib.orc_mixing_index_of.argtypes = [fp, fp, C.c_size_t, C.c_float, fp]
    lib.orc_dem_step.argtypes = [P, C.c_double, C.c_uint32, C.c_double, fp, fp, dp]
    lib.orc_surface.restype = C.c_int64
    lib.orc_surface.argtypes = [dp, C.c_float, fp, C.c_size_t, C.c_float, C.c_float, C.c_float, fp, C.c_int64, fp]
    lib.orc_splat_kernel.restype = C.c_float
    lib.orc_splat_kernel.argtypes = [C.c_float, C.c_float]
    lib.orc_dem_params.argtypes = [P, dp]
    lib.orc_fluid_step.argtypes = [P, C.c_double, C.c_uint32, C.c_double, dp, dp, dp]
    lib.orc_fluid_params.argtypes = [P, dp]
    lib.orc_stress_new.restype = P
    lib.orc_stress_new.argtypes = []
    lib.orc_stress_free.argtypes = [P]
    lib.orc_stress_set_factor.argtypes = [P, C.c_double]
    lib.orc_stress_update.argtypes = [P, C.POINTER(C.c_int32), u32p, u32p, fp, fp, C.c_size_t] + [C.c_double] * 7
    lib.orc_stress_state.argtypes = [P, dp]
    lib.orc_stress_metrics.argtypes = [P, C.c_size_t, C.c_double, C.c_double, C.c_double, C.c_double, dp]
    lib.orc_se_bin.restype = C.c_size_t
    lib.orc_se_bin.argtypes = [C.c_double]
    lib.orc_drum_metrics.restype = C.c_int
    lib.orc_drum_metrics.argtypes = [dp, C.c_float, fp, fp, fp, C.c_size_t, C.c_float, C.c_float, C.c_float,
                                     fp, fp, fp, C.c_size_t, C.c_float, C.c_float, dp]
    lib.orc_grid_near.restype = C.c_int64
    lib.orc_grid_near.argtypes = [fp, C.c_size_t, C.c_float, fp, C.c_size_t, u32p, u32p, C.c_int64]
    lib.orc_neighbor_lists.restype = C.c_int
    lib.orc_neighbor_lists.argtypes = [fp, C.c_size_t, C.c_float, u32p, u32p, C.c_int64, i64p]
    lib.orc_rng.argtypes = [C.c_uint64, C.POINTER(C.c_uint64), fp, C.c_int]
    lib.orc_sdf_world.argtypes = [dp, C.c_double, C.c_double, C.c_double, dp]
    lib.orc_toi_swept.restype = C.c_int
    lib.orc_toi_swept.argtypes = [dp, dp, dp, C.c_double, C.c_double, C.c_double, dp]
    lib.orc_squeeze_coefficient.restype = C.c_double
    lib.orc_squeeze_coefficient.argtypes = [C.c_double] * 4
    lib.orc_resolved_ratio.restype = C.c_double
    lib.orc_resolved_ratio.argtypes = [C.c_double]
    _lib = lib
    return lib


_DEFAULTS = {
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


def default_params() -> dict:
    return copy.deepcopy(_DEFAULTS)


def _merge(base: dict, over: dict) -> dict:
    out = copy.deepcopy(base)
    for k, v in (over or {}).items():
        if isinstance(v, dict) and isinstance(out.get(k), dict):
            out[k] = _merge(out[k], v)
        else:
            out[k] = v
    return out


def flatten_params(params) -> tuple[np.ndarray, int]:
    """Flattens a Params dict (or JSON string) to the oracle's 41-double layout plus the seed."""
    if isinstance(params, str):
        params = json.loads(params)
    p = _merge(_DEFAULTS, params or {})
    m, me, s, li, si = p["mill"], p["media"], p["slurry"], p["lifters"], p["simulation"]
    enum = {
        "rpm": 0, "percent_critical": 1, "counter_clockwise": 0, "clockwise": 1,
        "newtonian": 0, "bingham": 1, "left_right": 0, "top_bottom": 1, "none": 2,
        "auto": 0, "off": 1, "manual": 2,
    }
    flat = [
        m["diameter_m"], enum[m["speed_mode"]], m["speed_value"], enum[m["direction"]],
        me["ball_diameter_m"], me["fill_fraction"], me["packing_fraction_2d"], me["density_kg_m3"],
        me["restitution_ball_ball"], me["restitution_ball_wall"], me["friction_ball_ball"],
        me["friction_ball_wall"], me["friction_ball_ball_static"], me["friction_ball_wall_static"],
        me["friction_velocity_scale_m_s"], me["rolling_friction"],
        1.0 if s["enabled"] else 0.0, s["fill_fraction"], s["density_kg_m3"], s["viscosity_pa_s"],
        enum[s["rheology"]], s["yield_stress_pa"], s["wall_no_slip"], s["ball_no_slip"],
        s["wettability"], s["surface_tension_n_m"], enum[s["dye_pattern"]],
        li["count"], li["height_m"], li["base_width_m"], li["top_width_m"], li["phase_deg"],
        si["resolution"], si["substeps"], si["pbf_iterations"], si["dem_iterations"],
        si["time_scale"], si["frame_budget_ms"], si["max_balls"],
        enum[si["coarse_graining_mode"]], si["coarse_graining_k"],
    ]
    # f32 fields are rounded to f32 exactly as serde parses them into f32.
    arr = np.array(flat, dtype=np.float64)
    return arr, int(si["seed"])


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class OracleSim:
    """The oracle's Simulation (lib.rs:82-566 restated). ``mode``: 'ref' | 'mixed' | 'fp64'."""

    def __init__(self, params=None, mode: str = "mixed", initial_balls=None):
        lib = oracle_lib()
        self._lib = lib
        flat, seed = flatten_params(params)
        self._flat = flat
        init_ptr, n_init = None, 0
        if initial_balls is not None:
            ib = np.ascontiguousarray(np.asarray(initial_balls, dtype=np.float32).reshape(-1))
            self._ib = ib
            init_ptr, n_init = ib.ctypes.data_as(C.POINTER(C.c_float)), ib.size // 2
        self._h = lib.orc_create(_dp(flat), seed, MODES[mode], init_ptr, n_init)
        self.mode = mode
        sz = np.zeros(16)
        lib.orc_sizes(self._h, _dp(sz))
        self.n_balls, self.n_fluid = int(sz[0]), int(sz[1])
        self.effective_substeps, self.effective_resolution = int(sz[2]), int(sz[3])
        self.true_ball_count = int(sz[4])
        self.fixed_sub_dt = float(sz[5])
        self.h, self.fluid_mass = float(sz[6]), float(sz[7])
        self.ball_radius, self.ball_mass, self.ball_inertia = float(sz[8]), float(sz[9]), float(sz[10])
        self.omega = float(sz[11])

    def __del__(self):
        if getattr(self, "_h", None):
            self._lib.orc_destroy(self._h)
            self._h = None

    def step_fixed(self, n: int = 1):
        for _ in range(n):
            self._lib.orc_step_fixed(self._h)

    def step(self, dt: float):
        self._lib.orc_step(self._h, dt)

    def advance(self, sub_dt: float):
        self._lib.orc_advance(self._h, sub_dt)

    def set_debug(self, on=True):
        self._lib.orc_set_debug(self._h, 1 if on else 0)

    def force_iterations(self, density: int, cg: int):
        self._lib.orc_force_iters(self._h, 1, density, cg)

    def release_iterations(self):
        self._lib.orc_force_iters(self._h, 0, 0, 0)

    def state(self) -> dict:
        nb, nf = self.n_balls, self.n_fluid
        bx, bv = np.zeros(2 * nb), np.zeros(2 * nb)
        bth, bom = np.zeros(nb), np.zeros(nb)
        fx, fv, dye, g = np.zeros(2 * nf), np.zeros(2 * nf), np.zeros(nf), np.zeros(2 * nf)
        hg = C.c_int(0)
        self._lib.orc_get_state(self._h, _dp(bx), _dp(bv), _dp(bth), _dp(bom), _dp(fx), _dp(fv),
                                _dp(dye), _dp(g), C.byref(hg))
        return {"ball_x": bx.reshape(-1, 2), "ball_v": bv.reshape(-1, 2), "ball_theta": bth,
                "ball_omega": bom, "fluid_x": fx.reshape(-1, 2), "fluid_v": fv.reshape(-1, 2),
                "dye": dye, "guess": g.reshape(-1, 2), "has_guess": bool(hg.value)}

    def set_state(self, st: dict):
        f = lambda k: np.ascontiguousarray(np.asarray(st[k], dtype=np.float64).reshape(-1))
        arrs = [f("ball_x"), f("ball_v"), f("ball_theta"), f("ball_omega"), f("fluid_x"),
                f("fluid_v"), f("guess")]
        self._keep = arrs
        self._lib.orc_set_state(self._h, *[_dp(a) for a in arrs], 1 if st.get("has_guess") else 0)

    def set_scalars(self, drum_angle: float, sim_time: float):
        self._lib.orc_set_scalars(self._h, drum_angle, sim_time)

    SCALAR_KEYS = ["drum_angle", "sim_time", "power_draw_w", "torque_nm", "collision_rate_per_s",
                   "dissipated_power_w", "max_substep_displacement_over_diameter",
                   "coupling_clamp_hits", "viscosity_iterations", "density_iterations",
                   "mean_shear_rate_per_s", "fluid_wall_work_j", "dem_wall_work_j",
                   "dem_dissipated_energy_j", "dem_coupling_work_j", "collision_count",
                   "n_groups", "n_links", "links_in_range", "n_ball_pairs", "ke_delta_noslip_j",
                   "ke_delta_drag_j", "ke_delta_buoyancy_j", "ke_delta_viscosity_j",
                   "ke_delta_clamp_j", "wall_impulse_x", "wall_impulse_y", "wall_angular_impulse"]

    def scalars(self) -> dict:
        o = np.zeros(32)
        self._lib.orc_scalars(self._h, _dp(o))
        return dict(zip(self.SCALAR_KEYS, o.tolist()))

    ENERGY_KEYS = ["elapsed_s", "shaft_work_j", "dem_wall_work_j", "fluid_wall_work_j",
                   "ball_contact_dissipation_j", "fluid_wall_slip_j", "fluid_viscous_j",
                   "fluid_clamp_removed_j", "interface_created_j", "delta_mechanical_j",
                   "unattributed_j", "fluid_wall_angular_impulse"]

    def energy_budget(self) -> dict:
        o = np.zeros(12)
        self._lib.orc_energy_budget(self._h, _dp(o))
        return dict(zip(self.ENERGY_KEYS, o.tolist()))

    def stress(self) -> np.ndarray:
        o = np.zeros(49)
        self._lib.orc_stress(self._h, _dp(o))
        return o

    DEM_STATS_KEYS = (["wall_work_j", "dissipated_energy_j", "coupling_work_j", "collision_count"] +
                      [f"impact_energy_histogram_{b}" for b in range(12)] +
                      ["max_substep_displacement_over_diameter", "n_pairs"])
    DEM_PARAM_KEYS = ("radius_m", "omega", "lifter_count", "lifter_height_m", "lifter_base_width_m",
                      "lifter_top_width_m", "lifter_phase_deg", "ball_radius", "ball_mass", "ball_inertia",
                      "restitution_ball_ball", "restitution_ball_wall", "friction_ball_ball", "friction_ball_wall",
                      "friction_ball_ball_static", "friction_ball_wall_static", "friction_velocity_scale_m_s",
                      "rolling_friction")

    def dem_step(self, drum_angle, iterations, dt, ext_imp=None, ext_ang=None) -> dict:
        """DemState::step_recorded (dem.rs:530-1066) on this simulation's balls alone (the fluid is
        untouched); returns DemStepStats, the records are in impacts()."""
        st = np.zeros(len(self.DEM_STATS_KEYS))
        ei = ea = None
        if ext_imp is not None:
            self._ext = (np.ascontiguousarray(np.asarray(ext_imp, np.float32).reshape(-1)),
                         np.ascontiguousarray(np.asarray(ext_ang, np.float32).reshape(-1)))
            ei, ea = _fp(self._ext[0]), _fp(self._ext[1])
        self._lib.orc_dem_step(self._h, float(drum_angle), int(iterations), float(dt), ei, ea, _dp(st))
        return dict(zip(self.DEM_STATS_KEYS, st.tolist()))

    FLUID_STATS_KEYS = ["viscosity_iterations", "density_iterations", "mean_shear_rate_per_s", "wall_work_j",
                        "wall_impulse_x", "wall_impulse_y", "wall_angular_impulse", "ke_delta_noslip_j",
                        "ke_delta_drag_j", "ke_delta_buoyancy_j", "ke_delta_viscosity_j", "ke_delta_clamp_j",
                        "coupling_clamp_hits", "n_groups", "n_links", "links_in_range", "has_guess", "_pad"]
    FLUID_PARAM_KEYS = ["h", "rest_density", "particle_mass", "ball_radius", "ball_mass", "ball_inertia",
                        "ball_true_radius", "viscosity_pa_s", "rheology", "yield_stress_pa", "wall_no_slip",
                        "ball_no_slip", "wettability", "surface_tension_n_m", "pbf_iterations"]

    def fluid_step(self, drum_angle, iterations, dt):
        """FluidParticles::step_coupled (pbf.rs:1174-2365) on this simulation's fluid with its balls
        as they are (the balls are not moved); returns (impulse (n, 2), angular (n,), stats dict with
        the 32 wall-angle bins under "wall_angle_bins")."""
        imp, ang = np.zeros(2 * self.n_balls + 2), np.zeros(self.n_balls + 1)
        st = np.zeros(len(self.FLUID_STATS_KEYS) + 32)
        self._lib.orc_fluid_step(self._h, float(drum_angle), int(iterations), float(dt), _dp(imp), _dp(ang), _dp(st))
        d = dict(zip(self.FLUID_STATS_KEYS, st[: len(self.FLUID_STATS_KEYS)].tolist()))
        d["wall_angle_bins"] = st[len(self.FLUID_STATS_KEYS):].copy()
        return imp[: 2 * self.n_balls].reshape(-1, 2), ang[: self.n_balls], d

    def fluid_params(self) -> dict:
        o = np.zeros(len(self.FLUID_PARAM_KEYS))
        self._lib.orc_fluid_params(self._h, _dp(o))
        return dict(zip(self.FLUID_PARAM_KEYS, o.tolist()))

    def dem_params(self) -> dict:
        o = np.zeros(len(self.DEM_PARAM_KEYS))
        self._lib.orc_dem_params(self._h, _dp(o))
        return dict(zip(self.DEM_PARAM_KEYS, o.tolist()))

    def impacts(self) -> np.ndarray:
        cap = 4 * (self.n_balls + 8) + 64
        o = np.zeros(6 * cap)
        n = self._lib.orc_impacts(self._h, _dp(o), cap)
        return o[: 6 * min(n, cap)].reshape(-1, 6)

    def wall_bins(self) -> np.ndarray:
        o = np.zeros(32)
        self._lib.orc_fluid_bins(self._h, _dp(o))
        return o

    def mixing_index(self):
        o = np.zeros(1)
        ok = self._lib.orc_mixing_index(self._h, _dp(o))
        return float(o[0]) if ok else None

    def debug_grid(self):
        meta = np.zeros(8, dtype=np.int64)
        cs = np.zeros(4 * self.n_fluid + 1024, dtype=np.uint32)
        idx = np.zeros(self.n_fluid + 1, dtype=np.uint32)
        u32 = lambda a: a.ctypes.data_as(C.POINTER(C.c_uint32))
        ok = self._lib.orc_dbg_grid(self._h, meta.ctypes.data_as(C.POINTER(C.c_int64)), u32(cs),
                                    cs.size, u32(idx), idx.size)
        assert ok
        return {"dense": bool(meta[0]), "min_cx": int(meta[1]), "min_cy": int(meta[2]),
                "width": int(meta[3]), "height": int(meta[4]),
                "cell_start": cs[: meta[5]].copy(), "indices": idx[: meta[6]].copy()}

    def debug_neighbors(self):
        off = np.zeros(self.n_fluid + 2, dtype=np.uint32)
        nb = np.zeros(64 * self.n_fluid + 64, dtype=np.uint32)
        sizes = np.zeros(2, dtype=np.int64)
        u32 = lambda a: a.ctypes.data_as(C.POINTER(C.c_uint32))
        ok = self._lib.orc_dbg_nbrs(self._h, u32(off), off.size, u32(nb), nb.size,
                                    sizes.ctypes.data_as(C.POINTER(C.c_int64)))
        assert ok
        return off[: sizes[0]].copy(), nb[: sizes[1]].copy()


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _u32(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def grid_build(points, cell: float):
    """UniformGrid::build (grid.rs:101-153) over float points; returns the CSR layout."""
    lib = oracle_lib()
    pts = np.ascontiguousarray(np.asarray(points, dtype=np.float32).reshape(-1))
    n = pts.size // 2
    meta = np.zeros(8, dtype=np.int64)
    cs = np.zeros(2 * n + 1024, dtype=np.uint32)
    idx = np.zeros(n + 1, dtype=np.uint32)
    ok = lib.orc_grid_build(_fp(pts), n, cell, meta.ctypes.data_as(C.POINTER(C.c_int64)), _u32(cs),
                            cs.size, _u32(idx), idx.size)
    assert ok
    return {"dense": bool(meta[0]), "min_cx": int(meta[1]), "min_cy": int(meta[2]),
            "width": int(meta[3]), "height": int(meta[4]),
            "cell_start": cs[: meta[5]].copy(), "indices": idx[: meta[6]].copy()}


def mixing_index_of(points, dye, radius: float):
    """metrics::mixing_index (metrics.rs:506-576) of caller points and dye; None when undefined."""
    lib = oracle_lib()
    pts = np.ascontiguousarray(np.asarray(points, dtype=np.float32).reshape(-1))
    d = np.ascontiguousarray(np.asarray(dye, dtype=np.float32).reshape(-1))
    out = np.zeros(1, dtype=np.float32)
    ok = lib.orc_mixing_index_of(_fp(pts), _fp(d), d.size, float(radius), _fp(out))
    return float(out[0]) if ok else None


def grid_keys(points, cell: float) -> np.ndarray:
    """Sparse-layout cell keys (cx, cy) in cell order (grid.rs:204-226); empty for a Dense grid."""
    lib = oracle_lib()
    pts = np.ascontiguousarray(np.asarray(points, dtype=np.float32).reshape(-1))
    n = pts.size // 2
    out = np.zeros(2 * n + 2, dtype=np.int32)
    k = lib.orc_grid_keys(_fp(pts), n, cell, out.ctypes.data_as(C.POINTER(C.c_int32)), n + 1)
    return out[: 2 * k].reshape(-1, 2)


def grid_near(points, cell: float, queries):
    """for_each_near (grid.rs:371-399) of each query point: (offsets, candidate indices)."""
    lib = oracle_lib()
    pts = np.ascontiguousarray(np.asarray(points, dtype=np.float32).reshape(-1))
    q = np.ascontiguousarray(np.asarray(queries, dtype=np.float32).reshape(-1))
    n, nq = pts.size // 2, q.size // 2
    off = np.zeros(nq + 1, dtype=np.uint32)
    k = lib.orc_grid_near(_fp(pts), n, cell, _fp(q), nq, _u32(off), None, 0)
    out = np.zeros(max(k, 1), dtype=np.uint32)
    lib.orc_grid_near(_fp(pts), n, cell, _fp(q), nq, _u32(off), _u32(out), k)
    return off, out[:k]


def grid_pairs(points, cell: float) -> np.ndarray:
    """Candidate pairs in emission order (grid.rs:291-362)."""
    lib = oracle_lib()
    pts = np.ascontiguousarray(np.asarray(points, dtype=np.float32).reshape(-1))
    n = pts.size // 2
    k = lib.orc_grid_pairs(_fp(pts), n, cell, None, 0)
    out = np.zeros(2 * max(k, 1), dtype=np.uint32)
    lib.orc_grid_pairs(_fp(pts), n, cell, _u32(out), k)
    return out[: 2 * k].reshape(-1, 2)


def neighbor_lists(points, h: float):
    """build_neighbor_lists (pbf.rs:262-288)."""
    lib = oracle_lib()
    pts = np.ascontiguousarray(np.asarray(points, dtype=np.float32).reshape(-1))
    n = pts.size // 2
    off = np.zeros(n + 1, dtype=np.uint32)
    cap = 64 * n + 64
    nb = np.zeros(cap, dtype=np.uint32)
    ne = np.zeros(1, dtype=np.int64)
    ok = lib.orc_neighbor_lists(_fp(pts), n, h, _u32(off), _u32(nb), cap,
                                ne.ctypes.data_as(C.POINTER(C.c_int64)))
    assert ok
    return off, nb[: ne[0]].copy()


def rng_stream(seed: int, n: int):
    lib = oracle_lib()
    u = np.zeros(n, dtype=np.uint64)
    f = np.zeros(n, dtype=np.float32)
    lib.orc_rng(seed, u.ctypes.data_as(C.POINTER(C.c_uint64)), _fp(f), n)
    return u, f


def drum_array(radius, omega=0.0, lifters=None):
    li = lifters or {}
    return np.array([radius, omega, li.get("count", 0), li.get("height_m", 0.02),
                     li.get("base_width_m", 0.03), li.get("top_width_m", 0.02),
                     li.get("phase_deg", 0.0)], dtype=np.float64)


def sdf_world(drum, p, angle):
    lib = oracle_lib()
    out = np.zeros(3)
    lib.orc_sdf_world(_dp(drum), float(p[0]), float(p[1]), float(angle), _dp(out))
    return float(out[0]), out[1:].copy()


# orc_drum_metrics output slots (metrics.rs:78-189 field order, None as a has_* flag)
DRUM_METRICS_KEYS = ("toe_angle_rad", "shoulder_angle_rad", "has_toe_shoulder", "centroid_x", "centroid_y",
                     "has_centroid", "pool_angle_min_rad", "pool_angle_max_rad", "has_pool_extent",
                     "free_surface_angle_rad", "free_surface_offset_m", "has_free_surface", "pool_depth_m",
                     "has_pool_depth", "mixing_index", "has_mixing_index", "total_kinetic_energy_j",
                     "max_ball_overlap_fraction", "max_ball_wall_overlap_fraction",
                     "max_fluid_density_error_fraction", "mean_fluid_density_error_fraction",
                     "max_fluid_compression_error_fraction", "mean_fluid_compression_error_fraction",
                     "has_density_errors", "drum_omega_rad_s", "fluid_particle_count")


def _f32(a, cols=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    return a.reshape(-1, cols) if cols else a.reshape(-1)


def drum_metrics(drum, angle, ball_x, ball_v, ball_omega, ball_radius, ball_mass, ball_inertia,
                 fluid_x, dye, density, h, rest_density):
    """metrics::compute (metrics.rs:195-254) of caller arrays: a dict keyed by DRUM_METRICS_KEYS
    (f32 values), or ValueError where the reference panics (NaN angle in a sorted wall list)."""
    lib = oracle_lib()
    bx, bv, bw = _f32(ball_x, 2), _f32(ball_v, 2), _f32(ball_omega)
    fx, dy, de = _f32(fluid_x, 2), _f32(dye), _f32(density)
    out = np.zeros(len(DRUM_METRICS_KEYS))
    ok = lib.orc_drum_metrics(_dp(np.asarray(drum, np.float64)), float(angle), _fp(bx), _fp(bv), _fp(bw), len(bx),
                              float(ball_radius), float(ball_mass), float(ball_inertia), _fp(fx), _fp(dy), _fp(de),
                              len(fx), float(h), float(rest_density), _dp(out))
    if not ok:
        raise ValueError("non-finite angle in a sorted wall list (the reference panics)")
    return dict(zip(DRUM_METRICS_KEYS, out.tolist()))


STRESS_METRICS_KEYS = (["ball_mass_3d_kg", "scale_energy_j"] + [f"impacts_per_ball_per_s_{c}" for c in range(4)] +
                       ["impacts_total_per_ball_per_s"] + [f"mean_se_scaled_{c}" for c in range(4)] +
                       ["mean_se_scaled_drop", "mean_se_scaled_shear", "stress_intensity_scaled",
                        "wall_power_per_ball_w", "specific_power_w_per_kg"] +
                       [f"thresholds_scaled_{t}" for t in range(6)] + [f"effective_power_per_ball_w_{t}" for t in range(6)] +
                       ["cumulative_wall_energy_per_ball_j", "cumulative_specific_energy_j_per_kg"] +
                       [f"cumulative_effective_energy_per_ball_j_{t}" for t in range(6)])


class StressAccum:
    """stress.rs:83-254 restated: update() with impact records, state() as the 49 accumulator
    doubles [count 4, se 4, hist 28, above 6, cum_wall, cum_above 6], metrics() as a dict."""

    def __init__(self):
        self._lib = oracle_lib()
        self._h = self._lib.orc_stress_new()

    def __del__(self):
        if getattr(self, "_h", None):
            self._lib.orc_stress_free(self._h)
            self._h = None

    def set_debounce_factor(self, factor):
        self._lib.orc_stress_set_factor(self._h, float(factor))

    def update(self, kind, ball, partner, vn, slip, m3, scale_j, d, now, sub_dt, alpha, wall_work):
        k = np.ascontiguousarray(np.asarray(kind, np.int32).reshape(-1))
        b = np.ascontiguousarray(np.asarray(ball, np.uint32).reshape(-1))
        p = np.ascontiguousarray(np.asarray(partner, np.uint32).reshape(-1))
        v = np.ascontiguousarray(np.asarray(vn, np.float32).reshape(-1))
        s = np.ascontiguousarray(np.asarray(slip, np.float32).reshape(-1))
        self._lib.orc_stress_update(self._h, k.ctypes.data_as(C.POINTER(C.c_int32)), _u32(b), _u32(p), _fp(v), _fp(s),
                                    len(k), m3, scale_j, d, now, sub_dt, alpha, wall_work)

    def state(self):
        o = np.zeros(49)
        self._lib.orc_stress_state(self._h, _dp(o))
        return o

    def metrics(self, n_balls, m3, scale_j, d, power):
        o = np.zeros(len(STRESS_METRICS_KEYS))
        self._lib.orc_stress_metrics(self._h, int(n_balls), m3, scale_j, d, power, _dp(o))
        return dict(zip(STRESS_METRICS_KEYS, o.tolist()))


def se_bin(e):
    """stress.rs:58-64 as the oracle evaluates it."""
    return int(oracle_lib().orc_se_bin(float(e)))


def unflatten_polygons(flat):
    """The flatten_polygons layout (surface.rs:420-431) back into a list of (k, 2) arrays."""
    flat = np.asarray(flat, np.float32)
    n, k, out = int(flat[0]), 1, []
    for _ in range(n):
        m = int(flat[k])
        out.append(flat[k + 1: k + 1 + 2 * m].reshape(-1, 2).copy())
        k += 1 + 2 * m
    return out


def extract_surface(drum, angle, fluid_x, h, rest_density, particle_mass, with_field=False):
    """extract_surface (surface.rs:233-277): the contour polygons, and the masked splat field."""
    lib = oracle_lib()
    fx = _f32(fluid_x, 2)
    field = np.zeros(128 * 128, np.float32) if with_field else None
    cap = 1 << 16
    while True:
        out = np.zeros(cap, np.float32)
        need = lib.orc_surface(_dp(np.asarray(drum, np.float64)), float(angle), _fp(fx), len(fx), float(h),
                               float(rest_density), float(particle_mass), _fp(out), cap,
                               _fp(field) if with_field else None)
        if need <= cap:
            break
        cap = int(need)
    polys = unflatten_polygons(out[:need])
    return (polys, field.reshape(128, 128)) if with_field else polys


def splat_kernel(r2, radius):
    return float(oracle_lib().orc_splat_kernel(float(r2), float(radius)))


def toi_swept(drum, x0, x1, a0, a1, radius):
    lib = oracle_lib()
    t = np.zeros(1)
    ok = lib.orc_toi_swept(_dp(drum), _dp(np.asarray(x0, float)), _dp(np.asarray(x1, float)),
                           a0, a1, radius, _dp(t))
    return float(t[0]) if ok else None


def squeeze_coefficient(k1, k2, h, nu):
    return oracle_lib().orc_squeeze_coefficient(k1, k2, h, nu)


def resolved_ratio(hc):
    return oracle_lib().orc_resolved_ratio(hc)


# ---- kernel-isolation workload (periodic box), c3_oracle.cpp --------------------------------
C3_DEFAULTS = {"side": 1024, "dx": 1.0e-3, "rest": 1000.0, "mu": 1.0e-3, "dt": 1.0 / 1440.0,
               "passes": 4, "tol": 1.0e-6, "max_cg": 500, "seed": 1, "yield_stress": 0.0,
               "adaptive": 0, "gravity": 0.0, "speed_clamp": 0.0, "density_ratio": 0.0}
_C3_KEYS = ["side", "dx", "rest", "mu", "dt", "passes", "tol", "max_cg", "seed", "yield_stress", "adaptive",
            "gravity", "speed_clamp", "density_ratio"]
C3_MODES = {"ref": 0, "mixed": 1, "fp64": 2}


def c3_config(overrides=None) -> dict:
    c = dict(C3_DEFAULTS)
    c.update(overrides or {})
    return c


def _c3_arr(cfg) -> np.ndarray:
    c = c3_config(cfg)
    return np.array([float(c[k]) for k in _C3_KEYS], dtype=np.float64)


def _c3_lib():
    L = oracle_lib()
    if getattr(L, "_c3_ready", False):
        return L
    P = C.c_void_p
    dp = C.POINTER(C.c_double)
    fp = C.POINTER(C.c_float)
    u32p = C.POINTER(C.c_uint32)
    L.c3_create.restype = P
    L.c3_create.argtypes = [dp, C.c_int]
    L.c3_destroy.argtypes = [P]
    L.c3_box.argtypes = [dp, dp]
    L.c3_step.argtypes = [P, C.c_int]
    L.c3_count.restype = C.c_size_t
    L.c3_count.argtypes = [P]
    L.c3_state.argtypes = [P, dp, dp]
    L.c3_set_state.argtypes = [P, dp, dp]
    L.c3_stats.argtypes = [P, dp]
    L.c3_step_stats.argtypes = [P, dp]
    L.c3_density.argtypes = [P, dp]
    L.c3_neighbors.restype = C.c_int64
    L.c3_neighbors.argtypes = [P, u32p, u32p, C.c_int64]
    L.c3_neighbors_of.restype = C.c_int64
    L.c3_neighbors_of.argtypes = [dp, fp, C.c_size_t, u32p, u32p, C.c_int64]
    L.c3_seed_state.argtypes = [dp, fp, fp]
    L.c3_mean_shear.restype = C.c_double
    L.c3_mean_shear.argtypes = [P]
    L.c3_shear_rates_of.argtypes = [dp, fp, fp, fp, C.c_size_t, C.c_float, u32p, u32p, fp]
    L.c3_bingham_of.argtypes = [fp, C.c_size_t, C.c_float, C.c_float, fp]
    L.c3_pass_of.argtypes = [dp, C.c_int, dp, C.c_size_t, dp, dp, dp]
    L.c3_visc_of.argtypes = [dp, C.c_int, dp, dp, C.c_size_t, C.c_int, dp, dp]
    L.c3_weights_of.restype = C.c_int64
    L.c3_weights_of.argtypes = [dp, dp, C.c_size_t, dp, u32p, u32p, dp, C.c_int64]
    L._c3_ready = True
    return L


def c3_box(cfg=None) -> dict:
    """Box geometry: kernel radius h, side L, cell width cw, particle mass, cells per side."""
    out = np.zeros(5)
    _c3_lib().c3_box(_dp(_c3_arr(cfg)), _dp(out))
    return {"h": out[0], "L": out[1], "cw": out[2], "mass": out[3], "ncx": int(out[4])}


def c3_seed_state(cfg=None):
    c = c3_config(cfg)
    n = int(c["side"]) ** 2
    x = np.zeros(2 * n, dtype=np.float32)
    v = np.zeros(2 * n, dtype=np.float32)
    _c3_lib().c3_seed_state(_dp(_c3_arr(c)), _fp(x), _fp(v))
    return x.reshape(n, 2), v.reshape(n, 2)


def c3_neighbors_of(cfg, points):
    """Neighbour CSR (offsets, neighbours) of caller points in the workload's periodic box."""
    pts = np.ascontiguousarray(points, dtype=np.float32).reshape(-1)
    n = pts.size // 2
    off = np.zeros(n + 1, dtype=np.uint32)
    cap = 64 * n + 64
    nb = np.zeros(cap, dtype=np.uint32)
    e = _c3_lib().c3_neighbors_of(_dp(_c3_arr(cfg)), _fp(pts), n, _u32(off), _u32(nb), cap)
    assert e <= cap
    return off, nb[:e].copy()


def c3_shear_rates_of(cfg, points, velocities, densities, mass, lists=None):
    """Per-particle shear rates (f32, pbf.rs:821-849) of caller points, velocities and densities
    over `lists` (neighbour CSR offsets, entries) or their own neighbour lists in the workload's
    periodic box."""
    pts = np.ascontiguousarray(points, dtype=np.float32).reshape(-1)
    vel = np.ascontiguousarray(velocities, dtype=np.float32).reshape(-1)
    n = pts.size // 2
    dens = np.ascontiguousarray(densities, dtype=np.float32).reshape(-1)
    assert vel.size == 2 * n and dens.size == n
    out = np.zeros(n, dtype=np.float32)
    off = nb = None
    if lists is not None:
        off = np.ascontiguousarray(lists[0], dtype=np.uint32)
        nb = np.ascontiguousarray(lists[1], dtype=np.uint32)
        assert off.size == n + 1 and nb.size >= off[-1]
    _c3_lib().c3_shear_rates_of(_dp(_c3_arr(cfg)), _fp(pts), _fp(vel), _fp(dens), n, float(mass),
                                None if off is None else _u32(off), None if nb is None else _u32(nb), _fp(out))
    return out


def c3_bingham_of(rates, mu, tau_y):
    """Bingham effective viscosity (f32, pbf.rs:854-868) of each rate."""
    r = np.ascontiguousarray(rates, dtype=np.float32).reshape(-1)
    out = np.zeros(r.size, dtype=np.float32)
    _c3_lib().c3_bingham_of(_fp(r), r.size, float(mu), float(tau_y), _fp(out))
    return out


def _pts64(points):
    return np.ascontiguousarray(points, dtype=np.float64).reshape(-1)


def c3_pass_of(cfg, points, mode: str = "fp64"):
    """One density pass (pbf.rs:1257-1334) over caller points in the workload's box: densities,
    lambdas and the Jacobi displacements (positions are not moved). mode "ref" (f32) or "fp64"."""
    pts = _pts64(points)
    n = pts.size // 2
    dens, lam, dp = np.zeros(n), np.zeros(n), np.zeros(2 * n)
    _c3_lib().c3_pass_of(_dp(_c3_arr(cfg)), 2 if mode == "fp64" else 0, _dp(pts), n, _dp(dens), _dp(lam), _dp(dp))
    return dens, lam, dp.reshape(n, 2)


def c3_viscous_of(cfg, points, velocities, solve: bool, mode: str = "fp64"):
    """Over caller points with cfg's mu everywhere: the central-force viscous operator applied to
    the velocities (solve=False, pbf.rs:359-377) or the velocities after the implicit viscosity CG
    with cfg's dt, tol and max_cg (solve=True, pbf.rs:423-480; also returns iterations, residual
    norm and |b|). Densities are those at the points (pbf.rs:2136-2144)."""
    pts = _pts64(points)
    vel = _pts64(velocities)
    n = pts.size // 2
    out = np.zeros(2 * n)
    st = np.zeros(3)
    _c3_lib().c3_visc_of(_dp(_c3_arr(cfg)), 2 if mode == "fp64" else 0, _dp(pts), _dp(vel), n, 1 if solve else 0,
                         _dp(out), _dp(st))
    if solve:
        return out.reshape(n, 2), {"cg_iterations": int(st[0]), "residual": st[1], "b_norm": st[2]}
    return out.reshape(n, 2)


def c3_weights_of(cfg, points):
    """FP64 densities at caller points and the Morris weight of every edge of their neighbour
    lists (pbf.rs:312-344): (density, offsets, neighbours, weights)."""
    pts = _pts64(points)
    n = pts.size // 2
    dens = np.zeros(n)
    off = np.zeros(n + 1, dtype=np.uint32)
    cap = 64 * n + 64
    nb = np.zeros(cap, dtype=np.uint32)
    w = np.zeros(cap)
    e = _c3_lib().c3_weights_of(_dp(_c3_arr(cfg)), _dp(pts), n, _dp(dens), _u32(off), _u32(nb), _dp(w), cap)
    assert e <= cap
    return dens, off, nb[:e].copy(), w[:e].copy()


class C3Oracle:
    """The periodic-box kernel-isolation workload on the CPU (test infrastructure only)."""

    def __init__(self, cfg=None, mode: str = "mixed"):
        self.cfg = c3_config(cfg)
        self._L = _c3_lib()
        self._h = self._L.c3_create(_dp(_c3_arr(self.cfg)), C3_MODES[mode])
        self.n = int(self._L.c3_count(self._h))

    def __del__(self):
        if getattr(self, "_h", None):
            self._L.c3_destroy(self._h)
            self._h = None

    def step(self, n: int = 1):
        self._L.c3_step(self._h, n)

    def state(self):
        x = np.zeros(2 * self.n)
        v = np.zeros(2 * self.n)
        self._L.c3_state(self._h, _dp(x), _dp(v))
        return x.reshape(self.n, 2), v.reshape(self.n, 2)

    def set_state(self, x, v):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
        self._L.c3_set_state(self._h, _dp(x), _dp(v))

    def mean_shear_rate(self) -> float:
        """Mean shear rate of the last step (pbf.rs:800-818), 0 before the first step."""
        return float(self._L.c3_mean_shear(self._h))

    def stats(self) -> dict:
        o = np.zeros(3)
        self._L.c3_stats(self._h, _dp(o))
        return {"cg_iterations": int(o[0]), "residual": o[1], "b_norm": o[2]}

    def step_stats(self) -> dict:
        """Density passes of the last step and its f64 kinetic energies around the CG and the clamp."""
        o = np.zeros(4)
        self._L.c3_step_stats(self._h, _dp(o))
        return {"density_passes": int(o[0]), "ke_before_cg": o[1], "ke_after_cg": o[2], "ke_after_clamp": o[3]}

    def density(self) -> np.ndarray:
        o = np.zeros(self.n)
        self._L.c3_density(self._h, _dp(o))
        return o

    def neighbors(self):
        off = np.zeros(self.n + 1, dtype=np.uint32)
        cap = 64 * self.n + 64
        nb = np.zeros(cap, dtype=np.uint32)
        e = self._L.c3_neighbors(self._h, _u32(off), _u32(nb), cap)
        assert e <= cap
        return off, nb[:e].copy()
