#!/usr/bin/env python3
"""Benchmark driver: prints one JSON line.

Headline (drum-substeps/s, batched FP32 vs plain): the C4 batched ensemble, 256 drums drawn from a
seeded stream (2 mm balls, N_b U{100..2000} with the drum diameter that seeds them, fluid
resolution U{25..75}, media fill U[0.2, 0.4], 50-90 % of critical speed, friction U[0.2, 0.5]),
every drum advanced one substep per step (the fluid step, the DEM step with the fluid's coupling
impulses, the power/torque/energy accumulators and the stress update), state resident in HBM, f32
state with f64 CG dot products and energy sums. value = drums x substeps / wall time of the timed
substeps; ms_per_step = one substep of all drums. Beside it: the plain single-drum path (the C2
2000-disc drum, the C1 realtime preset), the power draw's error against the FP64 oracle after 480
realtime substeps, the canonical algorithmic bytes of a substep (N_f (286 + 32 n_d + 84 n_cg) +
N_b (48 (8 + n_dem) + 144)) and the kernel table from HIP event pairs over five more substeps.

roofline: the dominant kernel of the batched substep by total time, its algorithmic bytes per launch
(bench.kernel_bytes: the arrays it streams at FP32 widths) over its average launch time from a HIP
event pair per launch on the stream it runs on; `traffic` (HBM bytes per launch) is read from the
committed rocprofv3 summary profiles/drums_c4x256_kernel_trace.md (two --pmc runs, FETCH_SIZE and
WRITE_SIZE, scaled by a plain copy's counters) and names that source; null for another kernel or
batch size.
cpu_baseline: the C++ restatement of the reference (oracle/, f32 reference policy, 1 core) on a
bounded sample of the same ensemble (drums in order, two substeps each, about --cpu-seconds).
p50/p99 latency: the served sweep (--no-served skips it): requests drawn as the ensemble's drums
with Poisson arrivals at 50/s, submitted to mg_server_* as the wall clock reaches their arrival
times; latency = completion - scheduled arrival. Their durations are U[2, 20] revolutions scaled by
--served-scale (default 0.02, so a run of the sweep finishes in seconds; 1.0 is the full spec).

kernel_isolation (skipped with --no-box): the periodic PBF box of side x side particles (default
1024^2), per step the grid build + cell reorder, four density passes, velocity update, density
recompute and the implicit viscosity CG to 1e-6 relative residual, with its own roofline and CPU
baseline, FP64 and Bingham lines, and the standalone operators (uniform grid, mixing index, drum
metrics, stress accumulators, DEM steps of 256 populations, free surfaces), each with the
restatement's rate beside it where one exists.
--part drums|ensemble|grid|mixing|... times only that part (for a profiler run) and prints its JSON;
ensemble is the batched drums' timed substeps alone.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "drum-substeps/s (batched FP32 vs plain) at ≤1% power err; p50/p99 latency"
HBM_PEAK_GBS = 8000.0
# canonical algorithmic bytes per particle (FP32 state) per launch of each timed kernel class:
# grid build + reorder (predict 40, bin 16, scan 2, reorder 80), density pass A (R x, W lambda),
# pass B (R x, lambda; W x), velocity + density recompute (R x, x0; W v, rho), CG set-up 52, the
# 84 bytes of a CG iteration split over its two kernels: direction + operator apply (R x, rho, r,
# p; W p, Ap = 44) and vector update (R v, r, Ap; W v, r = 40), and the first pass A of a step,
# which also builds the pair table (the table is not canonical: 12)
CLASS_NAMES = ["grid+reorder", "density_A", "density_B", "velocity+density", "cg_setup", "cg_dir_apply",
               "cg_update", "pairs+density_A1"]
CLASS_BYTES = [138.0, 12.0, 20.0, 28.0, 52.0, 44.0, 40.0, 12.0]
CLASS_KERNELS = [None, "bx_lambda<float>", "bx_dp<float>", "bx_vel_density<float>", "bx_cg_init<float>",
                 "bx_cg_dir<float>", "bx_cg_upd<float>", "bx_pairs_lambda<float>"]
REPLAY_CLASSES = (1, 2)  # the classes mg_box_replay_ms can relaunch (density passes A and B)
CG_CLASSES = (5, 6)  # the classes whose launches per step exceed their working launches (CG iterations)
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "box_1024_pmc_traffic.md")
VALU_PROFILE = os.path.join(ROOT, "profiles", "box_1024_pmc_valu.md")
C2_PARAMS = {"mill": {"diameter_m": 0.1397, "speed_mode": "percent_critical", "speed_value": 75.0},
             "media": {"density_kg_m3": 7800.0},
             "simulation": {"resolution": 100, "max_balls": 2000}}


def measured_traffic(kernel, side):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary
    (calibrated against a plain copy), or None when no summary covers this workload."""
    if kernel is None or side != 1024 or not os.path.exists(TRAFFIC_PROFILE):
        return None
    for line in open(TRAFFIC_PROFILE):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 5 and cells[0] == kernel:
            try:
                return float(cells[4]) * 1024 * 1024
            except ValueError:
                return None
    return None


DRUM_PROFILE = os.path.join(ROOT, "profiles", "drums_c4x256_kernel_trace.md")


def measured_drum_traffic(kernel):
    """HBM bytes per launch (fetch + write) of a batched-substep kernel from the committed
    rocprofv3 summary of the 256-drum ensemble (calibrated against a plain copy), or None."""
    if kernel is None or not os.path.exists(DRUM_PROFILE):
        return None
    for line in open(DRUM_PROFILE):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 7 and cells[0] == f"`{kernel}`":
            try:
                return (float(cells[5]) + float(cells[6])) * 1024 * 1024
            except ValueError:
                return None
    return None


def measured_valu_busy(kernel, side):
    """VALU busy fraction of `kernel` from the committed rocprofv3 SQ/GRBM summary, or None."""
    if kernel is None or side != 1024 or not os.path.exists(VALU_PROFILE):
        return None
    for line in open(VALU_PROFILE):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 4 and cells[0] == kernel:
            try:
                return round(float(cells[3]) / 100.0, 3)
            except ValueError:
                return None
    return None


def cpu_count_used() -> int:
    return 1


def sync():
    """Both HIP runtimes of the process: torch's and the library's own (the box, grid and metrics
    kernels run on the library's streams)."""
    import torch
    from milldynamics3_amd.grid import synchronize
    synchronize()
    torch.cuda.synchronize()


def cpu_baseline_drum(seconds: float) -> dict:
    from oracle import oracle as orc
    s = orc.OracleSim(C2_PARAMS, "ref")
    s.step_fixed(1)
    n = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        s.step_fixed(1)
        n += 1
    dt = time.perf_counter() - t0
    return {"value": n / dt, "unit": "drum-substeps/s", "cores": cpu_count_used(), "kind": "port",
            "sample": f"{n} substeps of C2 (2000 discs, 12697 fluid, D=139.7 mm) on the C++ restatement "
                      f"of the reference (f32), 1 core, {dt:.1f} s"}


def cpu_baseline_box(cfg: dict, seconds: float) -> dict:
    from oracle import oracle as orc
    s = orc.C3Oracle(cfg, "mixed")
    n = 0
    t0 = time.perf_counter()
    while n == 0 or time.perf_counter() - t0 < seconds:
        s.step()
        n += 1
    dt = time.perf_counter() - t0
    npart = cfg["side"] ** 2
    return {"value": n * npart / dt, "unit": "particle-steps/s", "cores": cpu_count_used(), "kind": "port",
            "sample": f"{n} steps of the {cfg['side']}^2 periodic box on the C++ restatement (f32, f64 CG dots), "
                      f"1 core, {dt:.1f} s"}


def _box_worker(cfg, seconds):
    from oracle import oracle as orc
    s = orc.C3Oracle(cfg, "mixed")
    n = 0
    t0 = time.perf_counter()
    while n == 0 or time.perf_counter() - t0 < seconds:
        s.step()
        n += 1
    return n, time.perf_counter() - t0


def cpu_baseline_box_cores(cfg: dict, seconds: float, cores: int) -> dict:
    """The same restatement on `cores` host cores at once, one independent box per core (the
    path shards that way across cores; the restatement itself is single-threaded)."""
    import concurrent.futures as cf
    import multiprocessing as mp
    with cf.ProcessPoolExecutor(max_workers=cores, mp_context=mp.get_context("spawn")) as ex:
        res = list(ex.map(_box_worker, [cfg] * cores, [seconds] * cores))
    npart = cfg["side"] ** 2
    value = sum(n * npart / dt for n, dt in res)
    return {"value": value, "unit": "particle-steps/s", "cores": cores, "kind": "port",
            "sample": f"{cores} independent {cfg['side']}^2 boxes on the C++ restatement, one per core, "
                      f"{sum(n for n, _ in res)} steps in {max(dt for _, dt in res):.1f} s"}


def bench_fp64(args, cfg) -> dict:
    """The same workload with FP64 state and arithmetic: K steps in one call after the warm-up."""
    import torch
    from milldynamics3_amd.box import Box
    box = Box(dict(cfg, fp64=1))
    n = box.n
    box.step(args.warmup)
    sync()
    cg0 = box.total_cg_iterations
    t0 = time.perf_counter()
    box.step(args.steps)
    sync()
    wall = time.perf_counter() - t0
    avg_cg = (box.total_cg_iterations - cg0) / args.steps
    box.close()
    ms = 1000.0 * wall / args.steps
    # FP64 canonical bytes per particle: 434 + 64 n_d + 168 n_cg (every stored value twice as wide)
    gbs = n * (434.0 + 64.0 * 4 + 168.0 * avg_cg) / (ms * 1e-3) / 1e9
    return {"metric": "particle-steps/s", "value": round(n * args.steps / wall, 1), "ms_per_step": round(ms, 4),
            "hbm_gbs_algorithmic": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
            "cg_iterations_per_step": avg_cg, "dtype": "f64 state and arithmetic",
            "fp32_vs_fp64": precision_vs_fp64(cfg, args.warmup + args.steps)}


def bench_bingham(args, cfg, yield_stress=1.0) -> dict:
    """The same workload as a Bingham fluid (yield stress in Pa): the shear-rate / effective-viscosity
    kernel before the CG set-up; K steps in one call after the warm-up; the last step's mean shear
    rate."""
    import torch
    from milldynamics3_amd.box import Box
    box = Box(dict(cfg, yield_stress=yield_stress))
    n = box.n
    box.step(args.warmup)
    sync()
    cg0 = box.total_cg_iterations
    t0 = time.perf_counter()
    box.step(args.steps)
    sync()
    wall = time.perf_counter() - t0
    out = {"metric": "particle-steps/s", "value": round(n * args.steps / wall, 1),
           "ms_per_step": round(1000.0 * wall / args.steps, 4), "yield_stress_pa": yield_stress,
           "cg_iterations_per_step": (box.total_cg_iterations - cg0) / args.steps,
           "mean_shear_rate_per_s": box.mean_shear_rate(), "dtype": "f32 state, f64 CG reductions"}
    box.close()
    return out


def bench_grid(args, cpu=True) -> dict:
    """The reference-layout uniform grid (grid.rs:101-153) over the box's initial points at cell h:
    the build (one host wait for the layout, then the radix-sorted cell order and cell starts) and
    the neighbour lists (pbf.rs:262-288), each timed over K calls after one warm-up call.
    Algorithmic bytes of a build: the points read twice (8 + 8 B) and the indices written (4 B) per
    point, plus 4 B per cell start. cpu_baseline: the oracle's restatement of the same build
    (1 core) on the same points."""
    import numpy as np
    from milldynamics3_amd.box import Box
    from milldynamics3_amd.grid import UniformGrid
    b = Box({"side": args.side})
    x, _ = b.state()
    b.close()
    h = float(np.float32(2.0e-3))
    g = UniformGrid(len(x))
    g.build(x, h)
    sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        g._check(g._L.mg_grid_build(g._h, g.points.ptr, g.n, h, None), "mg_grid_build")
    sync()
    build_ms = 1000.0 * (time.perf_counter() - t0) / args.steps
    md = g.meta()
    g.neighbor_lists(h)
    t0 = time.perf_counter()
    for _ in range(max(args.steps // 4, 1)):
        off, nb = g.neighbor_lists(h)
    nbr_ms = 1000.0 * (time.perf_counter() - t0) / max(args.steps // 4, 1)
    g.close()
    bytes_build = len(x) * 20.0 + 4.0 * md["n_cell_start"]
    out = {"points": len(x), "cell": h, "dense": md["dense"], "cells": md["n_cell_start"] - 1,
           "build_ms": round(build_ms, 4), "points_per_s": round(len(x) / (build_ms * 1e-3), 1),
           "build_gbs_algorithmic": round(bytes_build / (build_ms * 1e-3) / 1e9, 1),
           "neighbor_lists_ms_incl_copy": round(nbr_ms, 3), "edges": int(len(nb)),
           "timing": "host wall clock around K synchronous builds (a build waits once for its layout)"}
    if cpu:
        from oracle import oracle as orc
        orc.grid_build(x, h)
        reps, t0 = 0, time.perf_counter()
        while reps == 0 or time.perf_counter() - t0 < min(args.cpu_seconds, 3.0):
            orc.grid_build(x, h)
            reps += 1
        cpu_ms = 1000.0 * (time.perf_counter() - t0) / reps
        out["cpu_baseline"] = {"value": round(len(x) / (cpu_ms * 1e-3), 1), "unit": "points/s", "cores": 1,
                               "kind": "port", "build_ms": round(cpu_ms, 3),
                               "sample": f"{reps} builds of the same {len(x)} points on the C++ restatement "
                                         "of UniformGrid::build, 1 core"}
    return out


def bench_mixing(args, drums=1024, per_drum=12697, cpu=True) -> dict:
    """metrics::mixing_index (metrics.rs:506-576) over a batch of drums in one launch: synthetic
    drum states (uniform points in a disc of the C2 drum's radius, dye 0/1 by half-plane with 10 %
    flipped), `per_drum` fluid particles each (C2's fluid count). Algorithmic bytes: position and
    dye read once (12 B per particle)."""
    import numpy as np
    from milldynamics3_amd import metrics
    from milldynamics3_amd.grid import DeviceArray
    rng = np.random.default_rng(5)
    R = np.float32(0.1397 / 2)
    n = drums * per_drum
    r = R * np.sqrt(rng.random(n, dtype=np.float32))
    a = np.float32(2 * np.pi) * rng.random(n, dtype=np.float32)
    pts = np.stack([r * np.cos(a), r * np.sin(a)], 1).astype(np.float32)
    dye = (pts[:, 0] < 0).astype(np.float32)
    flip = rng.random(n) < 0.1
    dye[flip] = 1.0 - dye[flip]
    off = (np.arange(drums + 1, dtype=np.uint64) * per_drum).astype(np.uint32)
    rad = np.full(drums, R, dtype=np.float32)
    L = metrics._lib()
    dx, dd, doff, drad = (DeviceArray.from_numpy(v) for v in (pts, dye, off, rad))
    dout, dok = DeviceArray(4 * drums), DeviceArray(4 * drums)

    def launch():
        rc = L.mg_mixing_index_batch(dx.ptr, dd.ptr, doff.ptr, drad.ptr, drums, dout.ptr, dok.ptr, None)
        if rc != 0:
            raise RuntimeError(f"mg_mixing_index_batch failed ({rc})")
    launch()
    sync()
    reps = max(args.steps, 1)
    t0 = time.perf_counter()
    for _ in range(reps):
        launch()
    sync()
    ms = 1000.0 * (time.perf_counter() - t0) / reps
    out = {"drums": drums, "particles_per_drum": per_drum, "batch_ms": round(ms, 4),
           "drums_per_s": round(drums / (ms * 1e-3), 1),
           "gbs_algorithmic": round(12.0 * n / (ms * 1e-3) / 1e9, 1),
           "timing": f"host wall clock around {reps} back-to-back batch launches"}
    if cpu:
        from oracle import oracle as orc
        k, t0 = 0, time.perf_counter()
        while k == 0 or (time.perf_counter() - t0 < min(args.cpu_seconds, 2.0) and k < drums):
            s0 = k * per_drum
            orc.mixing_index_of(pts[s0:s0 + per_drum], dye[s0:s0 + per_drum], float(R))
            k += 1
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(k / dt, 1), "unit": "drums/s", "cores": 1, "kind": "port",
                               "sample": f"{k} drums of the same batch on the C++ restatement, 1 core"}
    return out


def bench_drum_metrics(args, drums=256, cpu=True) -> dict:
    """metrics::compute (metrics.rs:195-254) for a batch of drum states in one call: `drums` copies
    of the 2000-disc drum (12,697 fluid particles) stepped 12 substeps on the restatement, each
    turned by its own drum angle. Timed: the whole synchronous call (upload of the per-drum
    records, five kernels, read-back); the drum states are resident in device memory."""
    import numpy as np
    from milldynamics3_amd import metrics
    from milldynamics3_amd.grid import DeviceArray
    from oracle import oracle as orc
    o = orc.OracleSim(C2_PARAMS, "ref")
    o.step_fixed(12)
    st = o.state()
    nb, nf = len(st["ball_x"]), len(st["fluid_x"])
    rng = np.random.default_rng(3)
    dens = (1000.0 * (1.0 + 0.03 * rng.standard_normal(nf))).astype(np.float32)
    L = metrics._metrics_lib()
    recs = (metrics._MetricsDrum * drums)()
    for k in range(drums):
        r = recs[k]
        r.radius_m, r.omega, r.drum_angle = 0.1397 / 2, o.omega, float(k) * 0.01
        r.ball_radius, r.ball_mass, r.ball_inertia = o.ball_radius, o.ball_mass, o.ball_inertia
        r.h, r.rest_density, r.n_balls, r.n_fluid = o.h, 1000.0, nb, nf
    tile = lambda a, c: DeviceArray.from_numpy(np.tile(np.asarray(a, np.float32).reshape(-1, c), (drums, 1)))
    arrays = [tile(st["ball_x"], 2), tile(st["ball_v"], 2), tile(st["ball_omega"], 1), tile(st["fluid_x"], 2),
              tile(st["dye"], 1), tile(dens, 1)]
    out = (metrics._DrumMetrics * drums)()
    err = __import__("ctypes").create_string_buffer(256)

    def call():
        rc = L.mg_drum_metrics_batch(recs, drums, *[a.ptr for a in arrays], out, err, 256)
        if rc != 0:
            raise RuntimeError(err.value.decode())
    call()
    reps = max(args.steps // 4, 1)
    t0 = time.perf_counter()
    for _ in range(reps):
        call()
    ms = 1000.0 * (time.perf_counter() - t0) / reps
    res = {"drums": drums, "balls_per_drum": nb, "fluid_per_drum": nf, "batch_ms": round(ms, 3),
           "drums_per_s": round(drums / (ms * 1e-3), 1),
           "timing": f"host wall clock around {reps} synchronous batch calls"}
    if cpu:
        drum = orc.drum_array(0.1397 / 2, o.omega)
        k, t0 = 0, time.perf_counter()
        while k == 0 or (time.perf_counter() - t0 < min(args.cpu_seconds, 2.0) and k < drums):
            orc.drum_metrics(drum, float(k) * 0.01, st["ball_x"], st["ball_v"], st["ball_omega"], o.ball_radius,
                             o.ball_mass, o.ball_inertia, st["fluid_x"], st["dye"], dens, o.h, 1000.0)
            k += 1
        dt = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": round(k / dt, 1), "unit": "drums/s", "cores": 1, "kind": "port",
                               "sample": f"{k} drums of the same batch on the C++ restatement, 1 core"}
    return res


def bench_stress(args, drums=256, records=40) -> dict:
    """StressAccum::update (stress.rs:110-171) of `drums` accumulators per call, `records` impact
    records per drum per sub-step (pairs among 64 balls and wall contacts), K sub-steps. Timed: the
    synchronous update calls with the records already in device memory."""
    import numpy as np
    from milldynamics3_amd import stress
    from milldynamics3_amd.grid import DeviceArray
    import ctypes as C
    rng = np.random.default_rng(4)
    n = drums * records
    kind = (rng.random(n) < 0.3).astype(np.uint32)
    ball = rng.integers(0, 64, n).astype(np.uint32)
    partner = np.where(kind == 1, stress.WALL, (ball + 1 + rng.integers(0, 63, n)) % 64).astype(np.uint32)
    rec = stress.records(kind, ball, partner, rng.uniform(0.05, 2.0, n), rng.uniform(0.0, 1.0, n))
    acc = stress.StressAccumBatch(drums, table_capacity=4096)
    L = stress._lib()
    d_rec = DeviceArray.from_numpy(rec.view(np.uint8))
    d_off = DeviceArray.from_numpy((np.arange(drums + 1) * records).astype(np.uint32))
    steps = (stress._Step * drums)()
    err = C.create_string_buffer(256)
    m3, d = 2.0e-5, 0.002
    reps = max(args.steps, 1)
    t0 = None
    for s in range(reps + 1):
        for k in range(drums):
            st = steps[k]
            st.m3, st.scale_j, st.d, st.now_s = m3, m3 * 9.81 * d, d, s * 1e-4
            st.sub_dt, st.alpha, st.wall_work_j = 1e-4, 0.01, 0.0
        if s == 1:
            t0 = time.perf_counter()
        rc = L.mg_stress_update(acc._h, d_rec.ptr, d_off.ptr, steps, err, 256)
        if rc != 0:
            raise RuntimeError(err.value.decode())
    ms = 1000.0 * (time.perf_counter() - t0) / reps
    acc.close()
    return {"drums": drums, "records_per_drum": records, "update_ms": round(ms, 4),
            "drum_updates_per_s": round(drums / (ms * 1e-3), 1),
            "timing": f"host wall clock around {reps} synchronous updates (host step records included)"}


def bench_dem(args, pops=256, cpu=True) -> dict:
    """DemState::step_recorded (dem.rs:530-1066) for `pops` ball populations of the 2000-disc drum
    (dry: 2000 balls at 75 % critical speed) per call, K steps after one warm-up step. Timed: the
    synchronous mg_dem_step calls with the state resident in device memory (the host evaluates
    the rare swept clamp in between). cpu_baseline: the oracle's DEM step of one population."""
    import ctypes as C
    import numpy as np
    from milldynamics3_amd import dem
    from milldynamics3_amd.grid import DeviceArray
    from oracle import oracle as orc
    params = dict(C2_PARAMS, slurry={"enabled": False})
    o = orc.OracleSim(params, "ref")
    st, pr = o.state(), o.dem_params()
    n = len(st["ball_x"])
    iters = orc._merge(orc._DEFAULTS, params)["simulation"]["dem_iterations"]
    L = dem._lib()
    b = dem.DemBatch(pops, n)
    recs = (dem._Drum * pops)()
    for k in range(pops):
        r = recs[k]
        r.radius_m, r.omega, r.drum_angle = pr["radius_m"], pr["omega"], 0.001 * k
        r.ball_radius, r.ball_mass, r.ball_inertia = pr["ball_radius"], pr["ball_mass"], pr["ball_inertia"]
        for key in dem.MEDIA_KEYS:
            setattr(r, key, pr[key])
        r.iterations, r.dt, r.n_balls, r.has_external = iters, o.fixed_sub_dt, n, 0
    tile = lambda a, c: DeviceArray.from_numpy(np.tile(np.asarray(a, np.float32).reshape(-1, c), (pops, 1)))
    dx, dv, dth, dom = tile(st["ball_x"], 2), tile(st["ball_v"], 2), tile(st["ball_theta"], 1), tile(st["ball_omega"], 1)
    stats = (dem._Stats * pops)()
    err = C.create_string_buffer(512)

    def call():
        rc = L.mg_dem_step(b._h, recs, dx.ptr, dv.ptr, dth.ptr, dom.ptr, None, None, stats, err, 512)
        if rc != 0:
            raise RuntimeError(err.value.decode())
    call()
    reps = max(args.steps // 2, 1)
    t0 = time.perf_counter()
    for _ in range(reps):
        call()
    ms = 1000.0 * (time.perf_counter() - t0) / reps
    b.close()
    res = {"populations": pops, "balls_per_population": n, "iterations": iters, "step_ms": round(ms, 3),
           "population_steps_per_s": round(pops / (ms * 1e-3), 1), "pairs_per_population": int(stats[0].n_pairs),
           "timing": f"host wall clock around {reps} synchronous batch steps"}
    if cpu:
        k, t0 = 0, time.perf_counter()
        while k == 0 or time.perf_counter() - t0 < min(args.cpu_seconds, 2.0):
            o.dem_step(0.0, iters, o.fixed_sub_dt)
            k += 1
        dt = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": round(k / dt, 1), "unit": "population-steps/s", "cores": 1, "kind": "port",
                               "sample": f"{k} DEM steps of one 2000-ball population on the C++ restatement, 1 core"}
    return res


def bench_surface(args, drums=256, cpu=True) -> dict:
    """surface::extract_surface (surface.rs:233-277) for `drums` drum states per call: copies of the
    2000-disc drum (12,697 fluid particles) stepped 12 substeps on the restatement, each turned by
    its own drum angle. Timed: the whole synchronous mg_surface_extract (per-drum records, four
    kernels, the segment read-back and the host chain walk) with the positions resident in device
    memory. cpu_baseline: the restatement's extract_surface on one drum."""
    import numpy as np
    from milldynamics3_amd import surface
    from oracle import oracle as orc
    o = orc.OracleSim(C2_PARAMS, "ref")
    o.step_fixed(12)
    fx = o.state()["fluid_x"].astype(np.float32)
    R = 0.1397 / 2
    batch = surface.SurfaceBatch(drums, drums * len(fx))
    recs = batch.upload([{"radius_m": R, "drum_angle": 0.01 * k, "h": o.h, "rest_density": 1000.0,
                          "particle_mass": o.fluid_mass, "fluid_x": fx} for k in range(drums)])
    batch.run(recs)
    reps = max(args.steps // 4, 1)
    t0 = time.perf_counter()
    for _ in range(reps):
        batch.run(recs)
    ms = 1000.0 * (time.perf_counter() - t0) / reps
    polys = batch.polygons(0)
    batch.close()
    res = {"drums": drums, "fluid_per_drum": len(fx), "polygons_drum0": len(polys),
           "points_drum0": int(sum(len(q) for q in polys)), "batch_ms": round(ms, 3),
           "drums_per_s": round(drums / (ms * 1e-3), 1),
           "timing": f"host wall clock around {reps} synchronous batch extractions"}
    if cpu:
        drum = orc.drum_array(R)
        k, t0 = 0, time.perf_counter()
        while k == 0 or (time.perf_counter() - t0 < min(args.cpu_seconds, 2.0) and k < drums):
            orc.extract_surface(drum, 0.01 * k, fx, o.h, 1000.0, o.fluid_mass)
            k += 1
        dt = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": round(k / dt, 1), "unit": "drums/s", "cores": 1, "kind": "port",
                               "sample": f"{k} drums of the same batch on the C++ restatement, 1 core"}
    return res


def precision_vs_fp64(cfg, steps) -> dict:
    """Error of the FP32-state box's outputs against the FP64-state box after the same `steps` steps
    from the same initial state (the FP64 box is the oracle's f64 policy on the GPU): positions in
    lattice spacings (minimum image), velocities and densities relative, kinetic energy relative."""
    import numpy as np
    from milldynamics3_amd.box import Box, box_geometry
    b32, b64 = Box(cfg), Box(dict(cfg, fp64=1))
    L = box_geometry(cfg)["L"]
    out, done = {}, 0
    for k in sorted({1, steps}):
        b32.step(k - done)
        b64.step(k - done)
        done = k
        x32, v32 = b32.state()
        x64, v64 = b64.state()
        r32, r64 = b32.density().astype(np.float64), b64.density()
        dx = x32.astype(np.float64) - x64
        dx -= L * np.round(dx / L)
        v32 = v32.astype(np.float64)
        ke32, ke64 = float((v32 * v32).sum()), float((v64 * v64).sum())
        out[f"after_{k}_steps"] = {
            "x_max_err_spacings": float(np.abs(dx).max() / (L / cfg["side"])),
            "v_rel_l2": float(np.linalg.norm(v32 - v64) / max(np.linalg.norm(v64), 1e-300)),
            "density_max_rel": float(np.abs(r32 - r64).max() / np.abs(r64).max()),
            "kinetic_energy_rel": abs(ke32 - ke64) / max(ke64, 1e-300)}
    b32.close()
    b64.close()
    return out



# ---- drums: the headline ------------------------------------------------------------------------
C1_PARAMS = {"simulation": {"max_balls": 300, "resolution": 25}}


def c4_params(n=256, seed=2026):
    """The batched ensemble: drum k draws N_b ~ U{100..2000}, resolution ~ U{25..75}, media fill
    J ~ U[0.20, 0.40], speed ~ U[50, 90] % of critical and ball friction ~ U[0.2, 0.5] (ball-ball and
    ball-wall, static = kinetic) from one seeded stream; the drum diameter d sqrt(N_b / (0.82 J)) makes
    the seeded population N_b balls of 2 mm; seed k + 1."""
    import math

    import numpy as np
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        nb = int(rng.integers(100, 2001))
        res = int(rng.integers(25, 76))
        fill = float(rng.uniform(0.20, 0.40))
        nc = float(rng.uniform(50.0, 90.0))
        mu = float(rng.uniform(0.2, 0.5))
        dia = 0.002 * math.sqrt(nb / (0.82 * fill))
        out.append({"mill": {"diameter_m": dia, "speed_mode": "percent_critical", "speed_value": nc},
                    "media": {"fill_fraction": fill, "friction_ball_ball": mu, "friction_ball_wall": mu,
                              "friction_ball_ball_static": mu, "friction_ball_wall_static": mu},
                    "simulation": {"resolution": res, "max_balls": nb, "seed": k + 1}})
    return out


def kernel_bytes(name, n_f, n_b, n_d, n_dem):
    """Algorithmic bytes of one launch of a drum-step kernel over n_f fluid particles and n_b balls
    (FP32 storage widths: Vec2 8 B, scalar and index 4 B; each array read or written once per element;
    neighbour gathers, per-edge tables and merged item lists not counted). Density passes and CG
    kernels follow the canonical per-particle table (pass A 12, pass B 20; CG apply 28, update 56);
    the others count the arrays they stream."""
    table = {
        "fl_predict": 40 * n_f,                     # R x, v; W x, v, x0
        "fl_nbr": 16 * n_f,                         # R x; W count / entries index
        "fl_dens_a": 12 * n_f, "fl_dens_b": 20 * n_f,
        "fl_dens_c": 24 * n_f,                      # R x, dp; W x
        "fl_push": 40 * n_f,                        # R x, v; W x, excess; records
        "fl_cohesion": 16 * n_f, "fl_wall": 16 * n_f + 4 * n_f,
        "fl_vel": 56 * n_f,                         # R x, x0, excess, coh, adh; W v
        "fl_gs": 16 * n_f + 36 * n_b,               # R, W v; per ball R x, v, omega, RW impulse
        "fl_gs_pre": 8 * n_f + 24 * n_b,
        "fl_dens2": 12 * n_f, "fl_mu": 12 * n_f, "fl_weights": 24 * n_f,
        "fl_cg_apply": 28 * n_f, "fl_cg_upd": 56 * n_f, "fl_cg_dir": 24 * n_f, "fl_cg_setup": 52 * n_f,
        "fl_shear": 28 * n_f, "fl_clamp": 16 * n_f, "fl_cg_out": 24 * n_f,
        "dem_project": 48 * n_dem * n_b, "dem_predict": 64 * n_b, "dem_friction": 48 * n_b,
        "dem_restitution": 48 * n_b, "dem_finish": 40 * n_b, "dem_sort": 24 * n_b, "dem_grid": 16 * n_b,
    }
    return table.get(name)


def bench_drums(args, cpu=True, batch_only=False) -> dict:
    """The batched ensemble (headline), the plain single drums, the power-draw error of the FP32
    engine against the FP64 oracle, the dominant kernel's roofline and the restatement's rate.
    batch_only: the ensemble's timed substeps alone (no kernel timers, for a profiler run)."""
    import numpy as np

    from milldynamics3_amd.sim import Batch, Simulation
    plist = c4_params(args.drums)
    b = Batch(plist)
    nf = np.array([d.n_fluid for d in b.drums])
    nb = np.array([d.n_balls for d in b.drums])
    b.step(max(args.warmup, 1))
    b.sync()
    sync()
    t0 = time.perf_counter()
    b.step(args.steps)
    b.sync()
    wall = time.perf_counter() - t0
    rate = len(plist) * args.steps / wall
    if batch_only:
        out = {"drum_substeps_per_s": round(rate, 1), "ms_per_substep": round(1000.0 * wall / args.steps, 3),
               "drums": len(plist), "fluid_particles": int(nf.sum()), "balls": int(nb.sum()), "steps": args.steps}
        b.close()
        return out
    st = [d.stats() for d in b.drums]
    n_d = np.array([s["density_iterations"] for s in st])
    n_cg = np.array([s["viscosity_iterations"] for s in st])
    n_dem = 2
    step_bytes = float(np.sum(nf * (286.0 + 32.0 * n_d + 84.0 * n_cg) + nb * (48.0 * (8 + n_dem) + 144.0)))
    # per-kernel HIP-event timing over K more substeps: the dominant kernel's average launch
    k2 = max(1, min(args.steps, 5))
    b.set_timing(True)
    b.step(k2)
    b.sync()
    kt = b.kernel_times()
    b.set_timing(False)
    dom = max(kt, key=lambda k: kt[k][0])
    avg_ms = kt[dom][0] / max(kt[dom][1], 1)
    bpl = kernel_bytes(dom, int(nf.sum()), int(nb.sum()), float(n_d.mean()), n_dem)
    achieved = None if bpl is None else bpl / (avg_ms * 1e-3) / 1e9
    traffic = measured_drum_traffic(dom)
    roofline = {"bound": "hbm", "kernel": dom, "achieved": None if achieved is None else round(achieved, 2),
                "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": None if achieved is None else round(achieved / HBM_PEAK_GBS, 5),
                "traffic": None if traffic is None or args.drums != 256 else round(traffic),
                "traffic_source": ("profiles/drums_c4x256_kernel_trace.md (rocprofv3 FETCH_SIZE + WRITE_SIZE per "
                                   "dispatch, copy-calibrated; not measured in this run)"),
                "bytes_per_launch": bpl, "avg_launch_us": round(1000.0 * avg_ms, 2), "launches": kt[dom][1],
                "timing": f"HIP event pair per launch on the null stream over {k2} batched substeps",
                "workload": f"C4 batched ensemble, {len(plist)} drums"}
    total_ms = sum(v[0] for v in kt.values())
    kernels = {k: {"total_ms": round(v[0], 3), "launches": v[1], "avg_us": round(1000.0 * v[0] / max(v[1], 1), 2),
                   "share": round(v[0] / total_ms, 4)}
               for k, v in sorted(kt.items(), key=lambda kv: -kv[1][0])[:48]}
    b.close()
    # the plain single-drum path
    plain = {}
    for label, p, k in (("C2", C2_PARAMS, max(4, args.steps // 2)), ("C1", C1_PARAMS, args.steps)):
        s = Simulation(p)
        s.step_fixed(2)
        t0 = time.perf_counter()
        s.step_fixed(k)
        dt = time.perf_counter() - t0
        plain[label] = {"drum_substeps_per_s": round(k / dt, 2), "ms_per_substep": round(1000.0 * dt / k, 3),
                        "balls": s.n_balls, "fluid": s.n_fluid}
        s.close()
    out = {"drum_substeps_per_s": round(rate, 1), "ms_per_substep": round(1000.0 * wall / args.steps, 3),
           "drums": len(plist), "fluid_particles": int(nf.sum()), "balls": int(nb.sum()),
           "steps": args.steps, "canonical_bytes_per_substep": step_bytes,
           "hbm_gbs_canonical": round(step_bytes * args.steps / wall / 1e9, 2),
           "density_passes_mean": float(n_d.mean()), "cg_iterations_mean": float(n_cg.mean()),
           "plain": plain, "batched_vs_plain_c2": round(rate / plain["C2"]["drum_substeps_per_s"], 1),
           "roofline": roofline, "kernels": kernels}
    out["precision"] = power_error(args.precision_substeps)
    if cpu:
        out["cpu_baseline"] = cpu_baseline_c4(plist, args.cpu_seconds)
    return out


def power_error(substeps) -> dict:
    """The FP32 engine (f32 state, f64 CG dot products and energy sums) against the FP64 oracle on
    the realtime preset from the same initial state: power, torque and collision-rate EMAs after
    `substeps` substeps (the reference's 1 s time constant)."""
    from milldynamics3_amd.sim import Simulation
    from oracle import oracle as orc
    g = Simulation(C1_PARAMS)
    o = orc.OracleSim(C1_PARAMS, "fp64")
    g.step_fixed(substeps)
    o.step_fixed(substeps)
    gs, os_ = g.stats(), o.scalars()
    g.close()
    out = {"workload": f"realtime preset (407 balls, 794 fluid), {substeps} substeps",
           "reference": "FP64 oracle (f64 state and sums)"}
    for k in ("power_draw_w", "torque_nm", "dissipated_power_w"):
        ref = os_[k]
        out[k] = {"fp32": gs[k], "fp64": ref, "rel_err": abs(gs[k] - ref) / max(abs(ref), 1e-30)}
    out["power_within_1pct"] = out["power_draw_w"]["rel_err"] <= 0.01
    return out


def c5_requests(n, scale, seed=7):
    """The served sweep's requests: Poisson arrivals at 50/s (exponential inter-arrival times) from
    one seeded stream, each drawing its drum as the batched ensemble does (same ranges) and a
    duration of U[2, 20] revolutions times `scale`; seed k + 1."""
    import math

    import numpy as np
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    for k in range(n):
        t += float(rng.exponential(1.0 / 50.0))
        nb = int(rng.integers(100, 2001))
        res = int(rng.integers(25, 76))
        fill = float(rng.uniform(0.20, 0.40))
        nc = float(rng.uniform(50.0, 90.0))
        mu = float(rng.uniform(0.2, 0.5))
        revs = float(rng.uniform(2.0, 20.0)) * scale
        dia = 0.002 * math.sqrt(nb / (0.82 * fill))
        p = {"mill": {"diameter_m": dia, "speed_mode": "percent_critical", "speed_value": nc},
             "media": {"fill_fraction": fill, "friction_ball_ball": mu, "friction_ball_wall": mu,
                       "friction_ball_ball_static": mu, "friction_ball_wall_static": mu},
             "simulation": {"resolution": res, "max_balls": nb, "seed": k + 1}}
        out.append((t, p, revs))
    return out


def bench_served(args) -> dict:
    """The served sweep (continuous batching through mg_server_*): requests submitted as the wall
    clock reaches their arrival times, the server ticked in between; latency = completion minus
    scheduled arrival. Aggregate rate = drum-substeps over the wall time from the first arrival to
    the last completion."""
    import numpy as np

    from milldynamics3_amd.server import Server
    reqs = c5_requests(args.served_requests, args.served_scale)
    s = Server(max_drums=256, max_balls=2048, hbm_budget_bytes=64 << 30)
    warm = s.submit(reqs[0][1], 0.01)  # the operators' first workspaces, outside the timed run
    while s.poll(8)["running"]:
        pass
    s.release(warm)
    t0 = s.now() + 0.05
    ids, i, st = [], 0, s.poll(0)
    base_steps = st["drum_substeps"]
    peak_running = 0
    while i < len(reqs) or st["running"] or st["queued"]:
        now = s.now()
        while i < len(reqs) and t0 + reqs[i][0] <= now:
            ids.append(s.submit(reqs[i][1], reqs[i][2], t0 + reqs[i][0]))
            i += 1
        st = s.poll(1)
        peak_running = max(peak_running, st["running"])
        if not st["running"] and not st["queued"] and i < len(reqs):
            time.sleep(max(0.0, t0 + reqs[i][0] - s.now()))
    res = [s.result(r) for r in ids]
    end = max(r["done_s"] for r in res)
    lat = np.array([r["done_s"] - r["arrival_s"] for r in res])
    wait = np.array([r["start_s"] - r["arrival_s"] for r in res])
    steps = st["drum_substeps"] - base_steps
    out = {"workload": f"{len(reqs)} requests, Poisson arrivals at 50/s (seed 7), drums drawn as the batched "
                       f"ensemble, durations U[2, 20] revolutions x {args.served_scale} "
                       f"({min(r['substeps_total'] for r in res)}-{max(r['substeps_total'] for r in res)} substeps)",
           "latency_p50_s": round(float(np.percentile(lat, 50)), 4),
           "latency_p99_s": round(float(np.percentile(lat, 99)), 4),
           "queue_wait_p99_s": round(float(np.percentile(wait, 99)), 4),
           "drum_substeps_per_s": round(steps / (end - t0), 1), "drum_substeps": int(steps),
           "ticks": int(st["ticks"]), "peak_running": int(peak_running),
           "hbm_budget_bytes": 64 << 30,
           "revolution_records": int(sum(r["revolutions_done"] for r in res))}
    s.close()
    return out


def cpu_baseline_c4(plist, seconds) -> dict:
    """The C++ restatement (oracle, f32 'ref' policy, 1 core) on a bounded sample of the same
    ensemble: drums in order, two substeps each, until `seconds` have passed."""
    from oracle import oracle as orc
    start = time.perf_counter()
    stepping, done, drums = 0.0, 0, 0
    for p in plist:
        o = orc.OracleSim(p, "ref")
        t1 = time.perf_counter()
        o.step_fixed(2)
        stepping += time.perf_counter() - t1
        done += 2
        drums += 1
        if time.perf_counter() - start > seconds:
            break
    return {"value": round(done / stepping, 2), "unit": "drum-substeps/s", "cores": 1, "kind": "port",
            "sample": f"{done} substeps: the first {drums} drums of the ensemble, 2 substeps each, on the C++ "
                      "restatement (oracle/, f32 reference policy), 1 core; stepping time only (each drum's "
                      "seeding excluded)"}


def bench_box(args) -> dict:
    """The kernel-isolation workload (periodic PBF box): the box kernels' table and roofline, FP64 and
    Bingham lines, and the standalone operators (grid, mixing index, drum metrics, stress, DEM,
    free surface)."""
    import numpy as np

    from milldynamics3_amd.box import Box
    cfg = {"side": args.side, "mu": args.mu}
    box = Box(cfg)
    n = box.n
    for _ in range(args.warmup):
        box.step()
    # 1) short pass with every kernel class timed: per-kernel table and the dominant kernel
    box.set_timing(0xFF)
    probe = max(1, min(args.steps, 5))
    for _ in range(probe):
        box.step()
    ms_all, launches_all = box.kernel_times()
    k = int(np.argmax(ms_all[1:8])) + 1  # grid build is several launches together: not a single kernel
    # 2) the timed region: K steps, no per-kernel events
    box.set_timing(0)
    sync()
    cg0 = box.total_cg_iterations
    t0 = time.perf_counter()
    box.step(args.steps)  # K steps in one call (the library waits for each step's CG to converge)
    sync()
    wall = time.perf_counter() - t0
    n_cg = box.total_cg_iterations - cg0
    # 3) the same K steps again with HIP events around the dominant kernel only (its average launch)
    box.set_timing(1 << k)
    cg1 = box.total_cg_iterations
    for _ in range(args.steps):
        box.step()
    ms_k, launches_k = box.kernel_times()
    working_k = box.total_cg_iterations - cg1
    box.set_timing(0)
    ms_per_step = 1000.0 * wall / args.steps
    n_d = 4
    avg_cg = n_cg / args.steps
    bytes_per_step = n * (226.0 + 32.0 * n_d + 84.0 * avg_cg)
    gbs = bytes_per_step / (ms_per_step * 1e-3) / 1e9
    event_pair_ms = ms_k[k] / max(int(launches_k[k]), 1)
    # the density passes: back-to-back launches of the step's last launch of the class between one
    # event pair (no dispatch gap inside the pair, so it compares with a kernel trace's durations)
    replay_reps = int(launches_k[k]) if k in REPLAY_CLASSES else 0
    avg_launch_ms = box.replay_ms(k, replay_reps) if replay_reps else event_pair_ms
    # A CG kernel's launches include the one per step that finds the CG converged and returns without
    # moving data: its algorithmic bytes per launch are the iteration's bytes times the fraction of
    # launches that iterate, so `achieved` is the class's total algorithmic bytes over its total time.
    work_frac = 1.0
    if k in CG_CLASSES and int(launches_k[k]) > 0:
        work_frac = min(1.0, working_k / int(launches_k[k]))
    bytes_per_launch = CLASS_BYTES[k] * n * work_frac

    def scaled(v, f):  # measured traffic is a working launch's: weighted like the bytes
        return None if v is None else round(v * f, 1)
    achieved = bytes_per_launch / (avg_launch_ms * 1e-3) / 1e9
    kernels = {CLASS_NAMES[i]: {"total_ms": round(float(ms_all[i]), 4), "launches": int(launches_all[i]),
                                "avg_us": round(1000.0 * float(ms_all[i]) / max(int(launches_all[i]), 1), 2)}
               for i in range(8)}
    roofline = {"bound": "hbm", "kernel": CLASS_KERNELS[k], "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": scaled(measured_traffic(CLASS_KERNELS[k], args.side), work_frac),
                "traffic_source": os.path.relpath(TRAFFIC_PROFILE, ROOT) + " (rocprofv3 FETCH_SIZE/WRITE_SIZE "
                                  "per dispatch, copy-calibrated; not measured in this run)",
                "valu_busy": measured_valu_busy(CLASS_KERNELS[k], args.side),
                "valu_source": os.path.relpath(VALU_PROFILE, ROOT) + " (not measured in this run)",
                "workload": f"periodic PBF box {args.side}x{args.side}",
                "bytes_per_launch": round(bytes_per_launch, 1), "avg_launch_us": round(1000.0 * avg_launch_ms, 2),
                "launches": replay_reps or int(launches_k[k]),
                "timing": "back-to-back relaunches, one event pair" if replay_reps else "event pair per launch",
                "event_pair_avg_us": round(1000.0 * event_pair_ms, 2)}
    if k in CG_CLASSES:
        roofline["iterating_launches"] = int(working_k)
    iso = {"metric": "particle-steps/s", "value": round(n * args.steps / wall, 1), "ms_per_step": round(ms_per_step, 4),
           "hbm_gbs_algorithmic": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
           "cg_iterations_per_step": avg_cg, "density_passes": n_d,
           "config": {"workload": f"periodic PBF box {args.side}x{args.side}", "particles": n, "mu": args.mu,
                      "dtype": "f32 state, f64 CG reductions"},
           "kernels_probe": kernels}
    iso["mean_shear_rate_per_s"] = box.mean_shear_rate()
    if not args.no_fp64:
        iso["fp64"] = bench_fp64(args, cfg)
    if not args.no_bingham:
        iso["bingham"] = bench_bingham(args, cfg)
    iso["grid"] = bench_grid(args, cpu=not args.no_cpu)
    iso["mixing_index"] = bench_mixing(args, cpu=not args.no_cpu)
    iso["drum_metrics"] = bench_drum_metrics(args, cpu=not args.no_cpu)
    iso["stress"] = bench_stress(args)
    iso["dem"] = bench_dem(args, cpu=not args.no_cpu)
    iso["surface"] = bench_surface(args, cpu=not args.no_cpu)
    box_cpu, drum_cpu = None, None
    if not args.no_cpu:
        box_cpu = cpu_baseline_box(cfg, args.cpu_seconds)
        drum_cpu = cpu_baseline_drum(args.cpu_seconds)
        cores = min(16, os.cpu_count() or 1)  # a one-GPU box's share of the host
        if cores > 1:
            iso["cpu_baseline_all_cores"] = cpu_baseline_box_cores(cfg, min(args.cpu_seconds, 5.0), cores)
    iso["roofline"] = roofline
    iso["cpu_baseline"] = box_cpu
    iso["cpu_baseline_drum_c2"] = drum_cpu
    box.close()
    return iso


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--side", type=int, default=1024)
    ap.add_argument("--mu", type=float, default=1.0e-3)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64-state box measurement")
    ap.add_argument("--no-bingham", action="store_true", help="skip the Bingham-fluid box measurement")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="also run one plain device copy of this size (calibrates PMC byte counters)")
    ap.add_argument("--drums", type=int, default=256, help="drums of the batched ensemble (headline)")
    ap.add_argument("--precision-substeps", type=int, default=480,
                    help="substeps of the FP32-vs-FP64 power comparison (realtime preset)")
    ap.add_argument("--no-box", action="store_true", help="skip the periodic-box kernel-isolation measurements")
    ap.add_argument("--served-requests", type=int, default=250, help="requests of the served sweep")
    ap.add_argument("--served-scale", type=float, default=0.02,
                    help="the served sweep's durations: U[2, 20] revolutions times this factor")
    ap.add_argument("--no-served", action="store_true", help="skip the served sweep")
    ap.add_argument("--part", choices=("all", "drums", "ensemble", "served", "grid", "mixing", "drum_metrics", "stress",
                                       "dem", "surface"),
                    default="all",
                    help="grid / mixing: time only that operator (for a profiler run) and print its JSON")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(drums and boxes shard as independent units, one process per GPU)", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from milldynamics3_amd import drum_engine_enabled
    from milldynamics3_amd.box import Box

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        return 1
    torch.cuda.set_device(0)

    if args.copy_mib > 0:
        src = torch.ones(args.copy_mib * 262144, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if args.part != "all":
        part = {"grid": lambda: bench_grid(args, cpu=not args.no_cpu),
                "mixing": lambda: bench_mixing(args, cpu=not args.no_cpu),
                "drum_metrics": lambda: bench_drum_metrics(args, cpu=not args.no_cpu),
                "stress": lambda: bench_stress(args),
                "dem": lambda: bench_dem(args, cpu=not args.no_cpu),
                "surface": lambda: bench_surface(args, cpu=not args.no_cpu),
                "drums": lambda: bench_drums(args, cpu=not args.no_cpu),
                "ensemble": lambda: bench_drums(args, cpu=False, batch_only=True),
                "served": lambda: bench_served(args)}[args.part]()
        print(json.dumps({"part": args.part, **part}), flush=True)
        return 0
    drums = bench_drums(args, cpu=not args.no_cpu)
    served = None if args.no_served else bench_served(args)
    iso = None if args.no_box else bench_box(args)
    line = {
        "metric": METRIC,
        "value": drums["drum_substeps_per_s"],
        "unit": "drum-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": drums["ms_per_substep"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"C4 batched ensemble: {drums['drums']} drums ({drums['balls']} balls, "
                               f"{drums['fluid_particles']} fluid particles) seeded from Rng 2026 (N_b U{{100..2000}}, "
                               "resolution U{25..75}, media fill U[0.2, 0.4], 50-90 % critical speed, friction "
                               "U[0.2, 0.5]); one substep of every drum per step (fluid step, DEM step with the "
                               "coupling impulses, accumulators, stress update), state resident in HBM",
                   "precision": "f32 state, f64 CG dot products and energy/power sums",
                   "timed": f"{args.steps} batched substeps after {max(args.warmup, 1)} warm-up substeps; "
                            "ms_per_step is one substep of all drums",
                   "plain_c2_drum_substeps_per_s": drums["plain"]["C2"]["drum_substeps_per_s"],
                   "batched_vs_plain_c2": drums["batched_vs_plain_c2"],
                   "power_rel_err_vs_fp64": drums["precision"]["power_draw_w"]["rel_err"],
                   "p50_p99_latency": None if served is None else [served["latency_p50_s"],
                                                                     served["latency_p99_s"]],
                   "p50_p99_workload": None if served is None else served["workload"]},
        "roofline": drums["roofline"],
        "cpu_baseline": drums.get("cpu_baseline"),
        "drums": drums,
        "served": served,
        "kernel_isolation": iso,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
