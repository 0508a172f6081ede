/* millgpu — MI355X-native per-substep slurry-plus-media engine (C ABI).
 *
 * Drop-in boundary for the reference's mill_core::Simulation step API
 * (crates/mill-core/src/lib.rs:82-566) and its WASM binding
 * (crates/mill-wasm/src/lib.rs:15-204). Params cross as the same serde JSON; arrays cross as
 * caller-owned buffers laid out like the binding's Float32Array copies ([x0, y0, x1, y1, ...]).
 *
 * Conventions: every fallible entry point returns 0 on success and a negative MG_E* code on
 * failure, writing a NUL-terminated message into (err, errlen) when err is non-null. No C++
 * exception crosses this boundary. A handle is driven by one thread at a time; distinct
 * handles are independent. The library owns all device memory. */
#ifndef MILLGPU_H
#define MILLGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MG_OK 0
#define MG_E_PARAMS -1   /* JSON parse or Params::validate failure (params.rs:115-530 messages) */
#define MG_E_DEVICE -2   /* HIP runtime error or no GPU */
#define MG_E_CAPACITY -3 /* a per-drum scratch capacity was exceeded (see mg_sim_error_bits) */
#define MG_E_ARG -4      /* invalid argument (null handle, short buffer, bad index) */
#define MG_E_DISABLED -5 /* reserved: a switched-off engine (mg_drum_engine_enabled() == 0); not returned by this build */

typedef struct mg_sim mg_sim;
typedef struct mg_batch mg_batch;

/* 1: this build runs drum substeps on the GPU (0 would mean the step entry points return
 * MG_E_DISABLED). No reference counterpart: a build-level property of this library. */
int mg_drum_engine_enabled(void);

/* EnergyBudget (lib.rs:61-78), f64 throughout. */
typedef struct mg_energy_budget {
  double elapsed_s, shaft_work_j, dem_wall_work_j, fluid_wall_work_j, ball_contact_dissipation_j,
      fluid_wall_slip_j, fluid_viscous_j, fluid_clamp_removed_j, interface_created_j,
      delta_mechanical_j, unattributed_j, fluid_wall_angular_impulse;
} mg_energy_budget;

/* ---- single drum: Simulation::new / with_initial_balls / set_params (lib.rs:126-274) ---- */
/* Replaces Simulation::new (lib.rs:126) and the wasm constructor (mill-wasm/src/lib.rs:24-34):
 * params_json == NULL means Params::default(). */
int mg_sim_create(const char* params_json, mg_sim** out, char* err, size_t errlen);
/* Replaces Simulation::with_initial_balls (lib.rs:134): xy holds n_balls (x, y) pairs. */
int mg_sim_create_with_balls(const char* params_json, const float* xy, size_t n_balls, mg_sim** out,
                             char* err, size_t errlen);
void mg_sim_destroy(mg_sim* sim);
/* Replaces Simulation::set_params (lib.rs:270): hot-applies without reseeding. */
int mg_sim_set_params(mg_sim* sim, const char* params_json, char* err, size_t errlen);

/* ---- stepping (lib.rs:426-509) ---- */
int mg_sim_step(mg_sim* sim, float dt);             /* Simulation::step (lib.rs:456) */
int mg_sim_step_fixed(mg_sim* sim);                 /* Simulation::step_fixed (lib.rs:438) */
int mg_sim_step_fixed_n(mg_sim* sim, uint32_t n);   /* n x step_fixed, one host synchronisation */
void mg_sim_reset_frame_stats(mg_sim* sim);          /* Simulation::reset_frame_stats (lib.rs:449) */
float mg_sim_fixed_sub_dt(const mg_sim* sim);        /* Simulation::fixed_sub_dt (lib.rs:426) */
uint32_t mg_sim_effective_substeps(const mg_sim* sim); /* Params::effective_substeps (params.rs:705) */

/* ---- state copies (Simulation::balls / fluid, lib.rs:287-292; wasm copies :103-138) ---- */
size_t mg_sim_ball_count(const mg_sim* sim);
size_t mg_sim_fluid_count(const mg_sim* sim);
/* Any output pointer may be NULL. x2n/v2n hold 2*count floats. */
int mg_sim_copy_balls(mg_sim* sim, float* x2n, float* v2n, float* theta, float* omega);
int mg_sim_copy_fluid(mg_sim* sim, float* x2n, float* v2n, float* dye);
/* Overwrites the dynamic state (parity harness / state injection). guess2n may be NULL. */
int mg_sim_set_state(mg_sim* sim, const float* ball_x2n, const float* ball_v2n, const float* theta,
                     const float* omega, const float* fluid_x2n, const float* fluid_v2n,
                     const float* guess2n, int has_guess);
int mg_sim_copy_guess(mg_sim* sim, float* guess2n, int* has_guess);
/* Sets the drum clock (drum_angle, sim_time) used by the next substep. */
int mg_sim_set_clock(mg_sim* sim, float drum_angle, double sim_time);

/* ---- read-only getters (lib.rs:276-349) ---- */
float mg_sim_drum_angle(const mg_sim* sim);
double mg_sim_sim_time(const mg_sim* sim);
uint32_t mg_sim_viscosity_iterations(mg_sim* sim);
float mg_sim_mean_shear_rate_per_s(mg_sim* sim);
float mg_sim_power_draw_w(mg_sim* sim);
float mg_sim_torque_nm(mg_sim* sim);
float mg_sim_collision_rate_per_s(mg_sim* sim);
float mg_sim_dissipated_power_w(mg_sim* sim);
float mg_sim_max_substep_displacement_over_diameter(mg_sim* sim);
uint32_t mg_sim_coupling_clamp_hits(mg_sim* sim);
/* Last-substep solver statistics as doubles (order documented in milldynamics3_amd/sim.py). */
int mg_sim_stats(mg_sim* sim, double* out, size_t cap);
/* StressAccum state: count 4, se 4, hist 28, above 6, cum_wall 1, cum_above 6 (stress.rs:83-99). */
int mg_sim_stress_state(mg_sim* sim, double* out49);
/* Impact records of the last substep: 6 doubles each (kind, ball, partner|-1, vn, slip, e). */
int mg_sim_impacts(mg_sim* sim, double* out, size_t cap_records, size_t* n_records);
/* Simulation::metrics (lib.rs:370-412) as JSON (Metrics keys this engine computes). */
int mg_sim_metrics_json(mg_sim* sim, char* buf, size_t cap);
/* metrics::mixing_index (metrics.rs:506-576); returns 0 and sets *ok = 0 for None. */
int mg_sim_mixing_index(mg_sim* sim, float* out, int* ok);
int mg_sim_energy_budget(mg_sim* sim, mg_energy_budget* out);   /* lib.rs:211-229 */
int mg_sim_reset_energy_budget(mg_sim* sim);                    /* lib.rs:205-208 */
int mg_sim_set_stress_debounce_factor(mg_sim* sim, double factor); /* lib.rs:244 */

/* ---- additions for this engine ---- */
/* Whole-state snapshot (state + hidden CG warm start + EMAs + stress accumulator + energy budget). */
int mg_sim_save_state(mg_sim* sim, void* buf, size_t cap, size_t* needed);
int mg_sim_load_state(mg_sim* sim, const void* buf, size_t len);
/* Forces the density pass count and CG iteration count of following substeps, as the reference's
 * tests do; in a batch the first forced drum's counts apply to every drum. */
int mg_sim_force_iterations(mg_sim* sim, uint32_t density_passes, uint32_t cg_iterations);
uint32_t mg_sim_error_bits(mg_sim* sim);

/* ---- batched drums: every drum of the batch advances in the same launches ----------------------
 * A substep of a batch is one fluid step, one DEM step and one stress update for all drums at once;
 * drum d's state follows drum d-1's in the batch's device arrays. */
int mg_batch_create(const char* const* params_json, uint32_t n_drums, mg_batch** out, char* err, size_t errlen);
void mg_batch_destroy(mg_batch* batch);
uint32_t mg_batch_size(const mg_batch* batch);
/* Advances every drum by `substeps` of its own fixed substep (step_fixed each); waits for the
 * device. use_graph is accepted for compatibility and has no effect in this build. */
int mg_batch_step(mg_batch* batch, uint32_t substeps, int use_graph);
/* Borrowed per-drum view (valid until the batch is destroyed; do not destroy it). */
mg_sim* mg_batch_drum(mg_batch* batch, uint32_t index);
/* Waits for the device; returns the last mg_batch_step's status (MG_E_CAPACITY if a drum exceeded a
 * workspace capacity) and its message. */
int mg_batch_sync(mg_batch* batch, char* err, size_t errlen);
/* Host wall-clock time (ms) of the last mg_batch_step (it waits for the device). */
float mg_batch_last_step_ms(const mg_batch* batch);
/* Per-kernel timing of the batch's following substeps (the fluid and DEM operators' kernels), as
 * mg_fluid_set_timing / mg_fluid_kernel_times. No reference counterpart: measurement. */
int mg_batch_set_timing(mg_batch* batch, int on);
int mg_batch_kernel_times(mg_batch* batch, char* json, size_t cap);

/* ---- served sweeps: continuous batching of independent drum runs --------------------------------
 * A request is one drum (Params JSON, validated as mg_sim_create does) run from its seeded initial
 * state for `revolutions` drum revolutions: ceil(revolutions * 2 pi / |omega| / dt) substeps of its
 * fixed sub-step dt (lib.rs:426-438). Each tick of the server admits queued requests in arrival
 * order while a slot is free and the running drums' HBM footprints stay within the budget (the
 * queue's head waits; no request overtakes it), runs one substep of every running drum as one batch
 * (the batch form's operators), records each drum's per-revolution results (power and torque EMAs,
 * collision rate, dissipated power, mixing index, the revolution's impact-energy histogram) and
 * retires the finished runs. A run's results equal the same drum stepped alone with mg_sim_step_fixed
 * (drums never exchange data). The reference has no server; its sweep analog runs seeds one after
 * another (batch.rs:40-106). Times are seconds on the server's clock (a steady clock from create). */
typedef struct mg_server mg_server;
typedef struct mg_server_config {
  uint32_t max_drums;        /* slots: drums stepped together at most (1..4096) */
  uint32_t max_balls;        /* balls per drum at most (1..2048) */
  uint64_t hbm_budget_bytes; /* the slots' fixed workspaces plus the running drums' footprints */
} mg_server_config;
typedef struct mg_server_status {
  uint32_t queued, running, done;
  uint64_t ticks, drum_substeps, used_bytes;
  double now_s;
} mg_server_status;
typedef struct mg_request_result {
  int32_t state; /* 0 queued, 1 running, 2 done */
  uint32_t n_balls, n_fluid, revolutions_done;
  uint64_t substeps_total, substeps_done, footprint_bytes;
  double arrival_s, start_s, done_s;
  float power_draw_w, torque_nm, collision_rate_per_s, dissipated_power_w; /* at the last substep run */
} mg_request_result;
/* MG_E_CAPACITY if the budget does not hold the slots' fixed workspaces. */
int mg_server_create(const mg_server_config* config, mg_server** out, char* err, size_t errlen);
void mg_server_destroy(mg_server* server);
double mg_server_now(const mg_server* server);
/* Queues a request; arrival_s < 0 means now (a driver replaying an arrival schedule passes the
 * scheduled time, so the wait before the call counts in the latency). MG_E_PARAMS for invalid
 * Params, a drum that does not turn or a non-positive revolution count; MG_E_CAPACITY if the drum
 * has more balls than max_balls or its footprint alone exceeds the budget. */
int mg_server_submit(mg_server* server, const char* params_json, double revolutions, double arrival_s, uint64_t* id,
                     char* err, size_t errlen);
/* Runs `ticks` ticks (fewer if nothing is queued or running); status may be NULL. */
int mg_server_poll(mg_server* server, uint32_t ticks, mg_server_status* status, char* err, size_t errlen);
int mg_server_result(mg_server* server, uint64_t id, mg_request_result* out);
/* The request's per-revolution records as a JSON array; returns the length, or -(bytes needed)
 * when cap is too small. */
int mg_server_stream_json(mg_server* server, uint64_t id, char* buf, size_t cap);
/* Forgets a finished request (its result and stream). */
int mg_server_release(mg_server* server, uint64_t id);

const char* mg_version(void);


/* ---- Periodic-box PBF workload: kernel isolation of the fluid hot path ---------------------
 * No reference counterpart (the reference has no periodic domain). One step applies the
 * reference's fluid arithmetic to a periodic box with minimum-image pair vectors: predict
 * (pbf.rs:1225-1235), grid + neighbour pairs (grid.rs:101-399, pbf.rs:262-288), `passes` density
 * passes (pbf.rs:1257-1334), v = (x - x0)/dt (pbf.rs:1656-1674), density recompute
 * (pbf.rs:2136-2144) and the implicit viscosity CG (pbf.rs:312-480) to tol * |b|. */
typedef struct mg_box mg_box;
typedef struct mg_box_config {
  uint32_t side;   /* side x side particles on a square lattice */
  float dx;        /* lattice spacing (m); kernel radius h = 2 dx; box side L = side * dx */
  float rest;      /* rest density */
  float mu;        /* viscosity (Pa s); 0 skips the CG */
  float dt;        /* step (s) */
  uint32_t passes; /* density passes per step */
  float tol;       /* the CG stops at tol * |b| */
  uint32_t max_cg; /* CG iteration cap */
  uint64_t seed;   /* Rng seed of the initial state (rng.rs) */
  uint32_t fp64;   /* 0: FP32 state with FP64 CG reductions; 1: FP64 state and arithmetic
                      (same f32 initial state and geometry, promoted) */
  float yield_stress; /* Bingham yield stress (Pa, the reference's yield_stress_pa with rheology
                         "bingham"); 0: Newtonian */
  uint32_t adaptive;  /* 0: exactly `passes` density passes; 1: the reference's adaptive count
                         (pbf.rs:1261-1333): at least max(passes, 1), stopping after a pass whose
                         largest constraint is <= 0.02, at most 6 more (passes <= 10) */
  float gravity;      /* added to v.y * dt in the predict (pbf.rs:1225-1235; the reference's
                         GRAVITY is -9.81); 0 in the kernel-isolation workload */
  float speed_clamp;  /* v_max of the fluid speed clamp after the CG (clamp_fluid_speed,
                         pbf.rs:987-996); 0: off */
  uint32_t energy_stats; /* 1: f64 kinetic energies of every step (mg_box_energy) */
  float density_ratio; /* density of the unjittered lattice over rest (the particle mass follows);
                          0: 1.02, the kernel-isolation workload */
} mg_box_config;

/* h, L, cell width, particle mass, cells per side. */
int mg_box_geometry(const mg_box_config* cfg, double* out5);
int mg_box_create(const mg_box_config* cfg, mg_box** out, char* err, size_t errlen);
void mg_box_destroy(mg_box* box);
size_t mg_box_count(const mg_box* box);
/* Runs `steps` steps and waits for them; mg_box_last_step_ms is their wall time on the stream. */
int mg_box_step(mg_box* box, uint32_t steps);
float mg_box_last_step_ms(const mg_box* box);
/* Positions and velocities in the original particle order (2n floats each; either may be null).
 * An FP64 box rounds them to f32 here; the _f64 forms copy either precision exactly. */
int mg_box_copy_state(mg_box* box, float* x2n, float* v2n);
int mg_box_set_state(mg_box* box, const float* x2n, const float* v2n);
int mg_box_copy_state_f64(mg_box* box, double* x2n, double* v2n);
int mg_box_set_state_f64(mg_box* box, const double* x2n, const double* v2n);
/* Density at the final positions of the last step, original order. */
int mg_box_density(mg_box* box, float* out_n);
int mg_box_density_f64(mg_box* box, double* out_n);
/* CG iterations, final residual norm and |b| of the last step. */
int mg_box_stats(mg_box* box, double* out3);
/* Mean shear rate (1/s) of the last step: the mean over particles of sqrt(2 D:D) from the SPH
 * velocity gradient at the step's final positions, velocities and densities over its neighbour
 * lists, as the reference reports each substep (FluidStepStats::mean_shear_rate_per_s,
 * pbf.rs:800-818, :2303-2304); the sum is FP64, rounded to f32 for an FP32 box. 0 before the first
 * step. Computed on request (one kernel), not inside the step. */
int mg_box_mean_shear_rate(mg_box* box, double* out);
/* Density passes the last step ran (the adaptive count when cfg.adaptive). */
uint32_t mg_box_density_passes(mg_box* box);
/* Kinetic energy 0.5 m sum |v|^2 (f64 sum of the per-particle |v|^2, kinetic_energy_f64,
 * pbf.rs:924-926) of the last step before the CG, after it and after the speed clamp; needs
 * cfg.energy_stats = 1 (else MG_E_ARG). */
int mg_box_energy(mg_box* box, double* out3);
/* CG iterations summed over every step this box has run. */
uint64_t mg_box_total_cg_iterations(const mg_box* box);
/* Multi-step calls run windows of up to 16 steps between host waits, each step's CG given the
 * largest iteration count the previous window saw (plus a margin while the counts change); a step
 * that needs more stops the rest of its window and is re-run from its checkpoint (same results),
 * and the next window is half as long. This sets the count the next window assumes and marks the
 * counts as steady (a small value exercises the re-run path). */
int mg_box_set_cg_prediction(mg_box* box, uint32_t iterations);
/* Self-check of the square root and the quotient by r that the box's pair loops use (the
 * correctly rounded sequences without range scaling, valid in the pair loops' operand range)
 * against the IEEE operations, on `samples` seeded random operands covering that range and its
 * edges. mismatches4: {f32 square root, f32 quotient, f64 square root, f64 quotient}; all must
 * be 0. No box needed. */
int mg_box_check_arith(uint64_t samples, uint64_t seed, uint64_t* mismatches4);
/* Neighbour CSR of the last step in traversal order (original ids); returns the edge count,
 * copying the entries only when they fit `cap`. */
int64_t mg_box_neighbors(mg_box* box, uint32_t* off_n1, uint32_t* nb, int64_t cap);
/* Per-kernel timing with HIP events on the box's stream for the classes whose bit is set in
 * class_mask (0 = off; resets the totals). Classes: 0 grid build and reorder (all their launches
 * together; also the pair table when a step has no density pass), 1 density pass A (passes 2..),
 * 2 density pass B, 3 velocity + density recompute, 4 CG set-up, 5 CG direction + operator
 * apply, 6 CG vector update, 7 pair table + first density pass A. */
int mg_box_set_timing(mg_box* box, uint32_t class_mask);
int mg_box_kernel_times(mg_box* box, double* ms, uint64_t* launches, int nclass);
/* Average duration of one launch of kernel class cls (1: density pass A, 2: density pass B; other
 * classes return MG_E_ARG): the last step's last launch of that class is launched again `reps`
 * times back to back on the box's stream between one HIP event pair (the launches rewrite the
 * values they wrote, so the box's state is unchanged). No counterpart in the reference: a
 * measurement hook. */
int mg_box_replay_ms(mg_box* box, uint32_t cls, uint32_t reps, double* ms_per_launch);

/* ---- Uniform grid in the reference's layout (grid.rs) --------------------------------------
 * A grid handle owns the device workspace and outputs of builds over up to max_points points.
 * Points are 2n floats [x0, y0, x1, y1, ...] in device memory; `stream` is a hipStream_t (NULL:
 * the null stream). A build waits once for the layout decision; the queries wait for their
 * totals. Outputs stay valid until the next build on the same handle. */
typedef struct mg_grid mg_grid;
int mg_grid_create(uint32_t max_points, mg_grid** out, char* err, size_t errlen);
void mg_grid_destroy(mg_grid* grid);
/* Replaces UniformGrid::build (grid.rs:101-153): cell keys (floor(x / cell) as i32, :280-285) of
 * the finite points, the Dense layout (x-major k = (cx - min_cx) * height + (cy - min_cy),
 * :232-236) when width * height <= 2 m + 512, else the Sparse layout (:204-226); points of a cell
 * in ascending index. MG_E_ARG for a non-positive or non-finite cell_size (the reference asserts,
 * grid.rs:102-105) or more points than the handle's capacity. */
int mg_grid_build(mg_grid* grid, const float* d_xy, uint32_t n, float cell_size, void* stream);
/* dense, min_cx, min_cy, width, height, cell_start entries, index entries (finite points), n. */
int mg_grid_meta(const mg_grid* grid, int64_t* meta8);
/* Host copies of cell_start, indices and (Sparse) the (cx, cy) keys of the occupied cells; any
 * pointer may be NULL. */
int mg_grid_copy(mg_grid* grid, uint32_t* cell_start, size_t cs_cap, uint32_t* indices, size_t idx_cap,
                 int32_t* keys2, size_t keys_cap);
/* Device views of the same arrays (valid until the next build). */
const uint32_t* mg_grid_cell_start_device(const mg_grid* grid);
const uint32_t* mg_grid_indices_device(const mg_grid* grid);
const int32_t* mg_grid_keys_device(const mg_grid* grid);
/* Replaces UniformGrid::for_each_candidate_pair (grid.rs:291-362) as a list: pairs (i, j), i < j,
 * in the reference's emission order into d_pairs2 (2 x cap u32, device); *n_pairs = all pairs
 * (those at or beyond cap are not written). */
int mg_grid_candidate_pairs(mg_grid* grid, uint32_t* d_pairs2, uint64_t cap, uint64_t* n_pairs, void* stream);
/* Replaces build_neighbor_lists (pbf.rs:262-288) over the points the grid was built from (same
 * d_xy): per point, the partners with |x_i - x_j|^2 <= h^2 in the order the reference appends
 * them. d_offsets: n + 1 u32 (device); d_nbrs: cap u32 (device); *n_edges = all entries. */
int mg_grid_neighbor_lists(mg_grid* grid, const float* d_xy, float h, uint32_t* d_offsets, uint32_t* d_nbrs,
                           uint64_t cap, uint64_t* n_edges, void* stream);
/* Replaces UniformGrid::for_each_near (grid.rs:371-399) for nq query points: CSR of the candidate
 * indices in visiting order. d_offsets: nq + 1 u32 (device). */
int mg_grid_near(mg_grid* grid, const float* d_queries, uint32_t nq, uint32_t* d_offsets, uint32_t* d_out,
                 uint64_t cap, uint64_t* n_out, void* stream);
/* Message of the handle's last failure. */
const char* mg_grid_last_error(const mg_grid* grid);
/* No counterpart in the reference: mg_grid_build for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). It
 * allocates nothing: its scratch is one fixed device array (about 33 MiB on each GPU it runs on,
 * held until the process ends) sized for at most 1,048,576 points (more: MG_E_ARG). The caller
 * passes the points and the outputs in device memory: cell_start (cs_cap entries; a Dense grid
 * needs width * height + 1 <= 2 m + 513, a Sparse one at most m + 1), indices (idx_cap >= m, the
 * finite points) and keys2 (keys_cap int32, 2 per occupied cell; a Sparse grid may need 2 m).
 * Too small an output: MG_E_CAPACITY before any output is written. meta8 (host) receives what
 * mg_grid_meta reports. Every call on one GPU shares that GPU's scratch, so calls that overlap
 * on one GPU race. Waits for `stream` before returning. */
int mg_grid_build_noalloc(const float* d_xy, uint32_t n, float cell_size, int64_t* meta8, uint32_t* d_cell_start,
                          uint64_t cs_cap, uint32_t* d_indices, uint64_t idx_cap, int32_t* d_keys2, uint64_t keys_cap,
                          void* stream, char* err, size_t errlen);

/* ---- on-demand metrics of drum states (metrics.rs) --------------------------------------------
 * Replaces metrics::mixing_index (metrics.rs:506-576): the Lacey index of the dye field of n fluid
 * particles (device xy, 2n floats; device dye, n floats) in a drum of the given radius, bit for bit
 * the reference's f32 sums; *defined = 0 where the reference returns None. */
int mg_mixing_index(const float* d_xy, const float* d_dye, uint32_t n, float radius, float* out, int* defined);
/* The same for n_drums drums in one launch (one workgroup per drum): drum d holds particles
 * [d_offsets[d], d_offsets[d + 1]) of d_xy / d_dye and radius d_radius[d]; results in device
 * d_out[d] / d_ok[d]. */
int mg_mixing_index_batch(const float* d_xy, const float* d_dye, const uint32_t* d_offsets, const float* d_radius,
                          uint32_t n_drums, float* d_out, int32_t* d_ok, void* stream);

/* One drum's scalars for mg_drum_metrics_batch: the Drum (geometry.rs:18-22, LiftersParams), the
 * drum angle, the ball population's shared radius / mass / inertia (dem.rs:277-309), the fluid's
 * smoothing length and rest density (pbf.rs:1018-1028), and the drum's counts. */
typedef struct mg_metrics_drum {
  float radius_m, omega, drum_angle;
  uint32_t lifter_count;
  float lifter_height_m, lifter_base_width_m, lifter_top_width_m, lifter_phase_deg;
  float ball_radius, ball_mass, ball_inertia;
  float h, rest_density;
  uint32_t n_balls, n_fluid;
} mg_metrics_drum;

/* metrics::compute's state-derived fields (metrics.rs:78-189 Metrics, in its field order); each
 * has_* = 0 where the reference returns None. */
typedef struct mg_drum_metrics {
  float toe_angle_rad, shoulder_angle_rad;
  int32_t has_toe_shoulder;
  float centroid_x, centroid_y;
  int32_t has_centroid;
  float pool_angle_min_rad, pool_angle_max_rad;
  int32_t has_pool_extent;
  float free_surface_angle_rad, free_surface_offset_m;
  int32_t has_free_surface;
  float pool_depth_m;
  int32_t has_pool_depth;
  float mixing_index;
  int32_t has_mixing_index;
  float total_kinetic_energy_j, max_ball_overlap_fraction, max_ball_wall_overlap_fraction;
  float max_fluid_density_error_fraction, mean_fluid_density_error_fraction;
  float max_fluid_compression_error_fraction, mean_fluid_compression_error_fraction;
  int32_t has_density_errors;
  float drum_omega_rad_s;
  uint32_t fluid_particle_count;
} mg_drum_metrics;

/* Replaces metrics::compute (metrics.rs:195-254) for n_drums drum states at once: toe/shoulder
 * (:327-357), charge centroid (:362-369), pool angular extent (:377-401), free-surface line
 * (:414-453), pool depth (:463-476), mixing index (:506-576), ball kinetic energy (:581-590), ball
 * and ball-wall overlaps (:599-634), density and compression errors (:649-710). Drum d's balls and
 * fluid follow drum d-1's in the device arrays (d_ball_x / d_ball_v: 2 floats per ball,
 * d_ball_omega: 1; d_fluid_x: 2 per particle, d_fluid_dye / d_fluid_density: 1); drums and out are
 * host arrays of n_drums records. Results equal the reference's f32 arithmetic bit for bit (sums
 * in the reference's order, wall angles by the host libm's atan2f algorithm); the free-surface
 * angle and offset use the host libm's atan2f / sinf / cosf. MG_E_ARG where the reference panics:
 * a non-finite position among at least 6 wall-adjacent balls or 2 wall-adjacent fluid particles
 * (its angle sort unwraps a NaN comparison); err names the drum. */
int mg_drum_metrics_batch(const mg_metrics_drum* drums, uint32_t n_drums, const float* d_ball_x, const float* d_ball_v,
                          const float* d_ball_omega, const float* d_fluid_x, const float* d_fluid_dye,
                          const float* d_fluid_density, mg_drum_metrics* out, char* err, size_t errlen);
/* The library's atan2f (the host libm's algorithm, used on the device for wall angles), on the
 * host over n pairs. No reference counterpart: exported so callers can check it against their
 * libm. */
void mg_atan2f_ref(const float* y, const float* x, float* out, uint64_t n);

/* ---- grinding-proxy stress accumulators (stress.rs) -------------------------------------------
 * ImpactRecord (dem.rs:152-169): kind 0 ball-ball, 1 ball-wall; partner 0xFFFFFFFF
 * (ImpactRecord::WALL) for a wall impact. */
typedef struct mg_impact_record {
  uint32_t kind, ball, partner;
  float normal_speed, slip_speed, restitution;
} mg_impact_record;
/* The arguments of StressAccum::update (stress.rs:110-120) for one drum and one sub-step. */
typedef struct mg_stress_step {
  double m3, scale_j, d, now_s, sub_dt, alpha, wall_work_j;
} mg_stress_step;
/* StressMetrics (stress.rs:243-273). */
typedef struct mg_stress_metrics {
  float ball_mass_3d_kg, scale_energy_j, impacts_per_ball_per_s[4], impacts_total_per_ball_per_s;
  float mean_se_scaled[4], mean_se_scaled_drop, mean_se_scaled_shear, stress_intensity_scaled;
  float wall_power_per_ball_w, specific_power_w_per_kg, thresholds_scaled[6], effective_power_per_ball_w[6];
  double cumulative_wall_energy_per_ball_j, cumulative_specific_energy_j_per_kg;
  double cumulative_effective_energy_per_ball_j[6];
} mg_stress_metrics;
typedef struct mg_stress mg_stress;
/* StressAccum::default (stress.rs:83-99) for n_drums drums, each with a debounce table of at
 * least table_capacity keys (rounded up to a power of two, at most 2^24) in device memory. */
int mg_stress_create(uint32_t n_drums, uint32_t table_capacity, mg_stress** out, char* err, size_t errlen);
void mg_stress_destroy(mg_stress* stress);
/* StressAccum::set_debounce_factor (stress.rs:103-106) of one drum (0 disables the debounce). */
int mg_stress_set_debounce_factor(mg_stress* stress, uint32_t drum, double factor);
/* StressAccum::update (stress.rs:110-171) of every drum with one sub-step's records: drum d's
 * records are d_records[d_offsets[d] .. d_offsets[d + 1]) (device arrays), steps[d] (host) its
 * arguments. Results equal the reference's f64 arithmetic bit for bit (records in order, se_bin by
 * the host libm's log10). Waits for the device. MG_E_ARG if a drum's now_s decreases;
 * MG_E_CAPACITY if a drum's debounce table had no room for a key (its counts may then differ from
 * the reference's from that record on). */
int mg_stress_update(mg_stress* stress, const mg_impact_record* d_records, const uint32_t* d_offsets,
                     const mg_stress_step* steps, char* err, size_t errlen);
/* One drum's 49 accumulator values: count_rate 4, se_hat_rate 4, hist_rate 28, above_rate_j 6,
 * cum_wall_j, cum_above_j 6 (stress.rs:83-99 field order). */
int mg_stress_state(mg_stress* stress, uint32_t drum, double* out49);
/* ---- ball populations: DemState::step_recorded (dem.rs:530-1066) --------------------------------
 * One population's scalars for one step: the Drum (geometry.rs:18-22, LiftersParams), the drum
 * angle, the shared ball radius / mass / inertia (dem.rs:277-309), the MediaParams the DEM reads
 * (params.rs), the solver iterations and the sub-step dt, the ball count, and whether coupling
 * impulses are applied (step_with_external_forces, dem.rs:513-528). */
typedef struct mg_dem_drum {
  float radius_m, omega, drum_angle;
  uint32_t lifter_count;
  float lifter_height_m, lifter_base_width_m, lifter_top_width_m, lifter_phase_deg;
  float ball_radius, ball_mass, ball_inertia;
  float restitution_ball_ball, restitution_ball_wall, friction_ball_ball, friction_ball_wall;
  float friction_ball_ball_static, friction_ball_wall_static, friction_velocity_scale_m_s, rolling_friction;
  uint32_t iterations;
  float dt;
  uint32_t n_balls, has_external;
  /* 0: the energies and the wall work are f32 sums in ball order, as the reference forms them;
   * 1: the same terms summed in f64 (the precision form's f64 energy accumulators). */
  uint32_t energy_f64;
} mg_dem_drum;
/* DemStepStats (dem.rs:175-222) plus the step's candidate-pair and impact-record counts and the
 * number of balls the swept time-of-impact test clamped. */
typedef struct mg_dem_stats {
  float wall_work_j;
  uint32_t collision_count, impact_energy_histogram[12];
  float dissipated_energy_j, max_substep_displacement_over_diameter, coupling_work_j;
  uint32_t n_pairs, n_impacts, n_swept;
  uint32_t projection_restarts; /* 0: the projection walked the near pairs only; 1: their travel
                                   check failed and a wider set passed it; 2: it walked every pair */
} mg_dem_stats;
typedef struct mg_dem mg_dem;
/* Workspace for n_drums ball populations of at most max_balls balls each (at most 2048: a
 * population's ball state is held in LDS during the pair sweeps), pairs_per_ball candidate pairs
 * per ball (0: 24) and max_impacts impact records per population and step. */
int mg_dem_create(uint32_t n_drums, uint32_t max_balls, uint32_t pairs_per_ball, uint32_t max_impacts, mg_dem** out,
                  char* err, size_t errlen);
void mg_dem_destroy(mg_dem* dem);
/* DemState::step_recorded (dem.rs:530-1066) for every population at once: population d's balls
 * follow population d-1's in the device arrays (x, v: 2 floats per ball; theta, omega: 1), updated
 * in place; d_ext_impulse (2 floats per ball) / d_ext_angular are the coupling impulses of the
 * populations with has_external (CouplingImpulses, coupling.rs:36-52), else may be NULL. stats
 * (host, n_drums) receives each population's DemStepStats. Results equal the reference's f32
 * arithmetic bit for bit: the pair sweeps keep its order, the friction blend uses the host libm's
 * expf algorithm, impact bins its log10f, the swept time of impact runs on the host with its sinf /
 * cosf. MG_E_CAPACITY if a population has more candidate pairs than the workspace holds (nothing
 * after the grid ran) or more impact records (the count is exact, the surplus records dropped). */
int mg_dem_step(mg_dem* dem, const mg_dem_drum* drums, float* d_x, float* d_v, float* d_theta, float* d_omega,
                const float* d_ext_impulse, const float* d_ext_angular, mg_dem_stats* stats, char* err,
                size_t errlen);
/* Per-kernel timing of the DEM steps, as mg_fluid_set_timing / mg_fluid_kernel_times. */
int mg_dem_set_timing(mg_dem* dem, int on);
int mg_dem_kernel_times(mg_dem* dem, char* json, size_t cap);
/* The last step's impact records of one population, in the reference's order (ImpactRecord,
 * dem.rs:152-169); *n receives the count. */
int mg_dem_impacts(mg_dem* dem, uint32_t drum, mg_impact_record* out, uint32_t cap, uint32_t* n);

/* ---- slurry: FluidParticles::step_coupled (pbf.rs:1174-2365) --------------------------------------
 * One drum's scalars for one step: the Drum (geometry.rs:18-22, LiftersParams) and its angle, the
 * FluidParticles scalars (h, rest_density, particle_mass; pbf.rs:1018-1028), the ball scalars the
 * coupling reads (dem.rs:277-309: radius, mass, inertia, true radius), the SlurryParams the step reads
 * (params.rs: viscosity, rheology 0 newtonian / 1 bingham, yield stress, wall and ball no-slip,
 * wettability, surface tension), the density iterations (pbf_iterations) and the sub-step dt, the
 * counts, and whether the caller's guess array holds a warm start for this drum (the previous joint
 * viscosity solve's velocity change, FluidParticles::viscosity_guess). */
typedef struct mg_fluid_drum {
  float radius_m, omega, drum_angle;
  uint32_t lifter_count;
  float lifter_height_m, lifter_base_width_m, lifter_top_width_m, lifter_phase_deg;
  float h, rest_density, particle_mass;
  float ball_radius, ball_mass, ball_inertia, ball_true_radius;
  float viscosity_pa_s;
  uint32_t rheology;
  float yield_stress_pa, wall_no_slip, ball_no_slip, wettability, surface_tension_n_m;
  uint32_t iterations;
  float dt;
  uint32_t n_fluid, n_balls;
  int32_t has_guess;
} mg_fluid_drum;
/* FluidStepStats (pbf.rs:1018-1028 and the step's tail) plus the step's structure counts: the wall
 * impulse and angular impulse are f64 sums, the 32 wall-angle bins f32 sums; KE deltas are the f64
 * kinetic-energy changes of the no-slip blend, form drag, buoyancy, viscosity solve and speed clamp.
 * joint_solve = 1 when the viscosity solve included the balls (the guess array then holds the
 * drum's next warm start); error_bits != 0 when a workspace capacity was exceeded. */
typedef struct mg_fluid_stats {
  uint32_t viscosity_iterations, density_iterations;
  float mean_shear_rate_per_s, wall_work_j;
  double wall_impulse_x, wall_impulse_y, wall_angular_impulse;
  float wall_angle_bins[32];
  double ke_delta_noslip_j, ke_delta_drag_j, ke_delta_buoyancy_j, ke_delta_viscosity_j, ke_delta_clamp_j;
  uint32_t coupling_clamp_hits, n_groups, n_links, links_in_range;
  int32_t joint_solve;
  uint32_t error_bits;
} mg_fluid_stats;
typedef struct mg_fluid mg_fluid;
/* Workspace for n_drums drums (at most 65535); its device memory grows with the batch it steps. */
int mg_fluid_create(uint32_t n_drums, mg_fluid** out, char* err, size_t errlen);
void mg_fluid_destroy(mg_fluid* fluid);
/* Per-kernel timing of the following steps (on = 1 starts it and clears the totals): each launch is
 * bracketed by HIP events on its stream. mg_*_kernel_times writes {"kernel": [total_ms, launches],
 * ...} as JSON into json (returns its length, or minus the size needed). No reference counterpart:
 * measurement. */
int mg_fluid_set_timing(mg_fluid* fluid, int on);
int mg_fluid_kernel_times(mg_fluid* fluid, char* json, size_t cap);
/* Fixed density-pass and CG-iteration counts for every drum (on = 1), as the reference's tests
 * force them; on = 0 restores the adaptive passes and the CG tolerance. No reference counterpart. */
int mg_fluid_force_iterations(mg_fluid* fluid, int on, uint32_t density_passes, uint32_t cg_iterations);
/* FluidParticles::step_coupled (pbf.rs:1174-2365) for every drum at once: drum d's fluid follows
 * drum d-1's in d_x / d_v / d_guess (2 floats per particle; x and v updated in place, guess written
 * for drums whose solve was joint), its balls follow drum d-1's in d_ball_x / d_ball_v (2 floats)
 * and d_ball_omega (1), and the coupling impulses it returns (CouplingImpulses, coupling.rs:36-52)
 * are written to d_impulse (2 floats per ball) and d_angular (1). stats (host, n_drums) receives
 * each drum's FluidStepStats. Positions, velocities and impulses up to the viscosity solve equal the
 * reference's f32 arithmetic bit for bit (adhesion, form drag and buoyancy in the reference's ball
 * order; the form-drag and Bingham exponentials by the host libm's expf algorithm, the wall-angle bins
 * by its atan2f); the CG's f64 dot products are summed in another order than the reference's, so the
 * solve agrees to rounding. Waits for the device. MG_E_CAPACITY when a drum exceeded a workspace
 * capacity (stats[d].error_bits says which; that drum's results are not valid). */
int mg_fluid_step(mg_fluid* fluid, const mg_fluid_drum* drums, float* d_x, float* d_v, float* d_guess,
                  const float* d_ball_x, const float* d_ball_v, const float* d_ball_omega, float* d_impulse,
                  float* d_angular, mg_fluid_stats* stats, char* err, size_t errlen);

/* ---- fluid free surface: surface::extract_surface (surface.rs:233-277) ---------------------------
 * One drum's inputs: the Drum (geometry.rs:18-22, LiftersParams), its angle, the FluidParticles
 * scalars the extraction reads (h, rest_density, particle_mass, pbf.rs) and the particle count. */
typedef struct mg_surface_drum {
  float radius_m, drum_angle;
  uint32_t lifter_count;
  float lifter_height_m, lifter_base_width_m, lifter_top_width_m, lifter_phase_deg;
  float h, rest_density, particle_mass;
  uint32_t n_fluid;
} mg_surface_drum;
typedef struct mg_surface mg_surface;
/* Workspace for n_drums drums (at most 65535) holding at most max_fluid fluid particles in total. */
int mg_surface_create(uint32_t n_drums, uint32_t max_fluid, mg_surface** out, char* err, size_t errlen);
void mg_surface_destroy(mg_surface* surface);
/* extract_surface (surface.rs:233-277) for every drum at once; Simulation::fluid_surface
 * (lib.rs:353-360) calls it with the simulation's drum and angle. Drum d's particles follow drum
 * d-1's in d_fluid_x (device, 2 floats each). The device splats the 128 x 128 occupancy field,
 * masks it with the wall SDF, classifies the marching-squares cells and interpolates the crossing
 * points; the host chains the segments, smooths and projects them (the reference's segment order
 * and HashMap walk). The field, segments and points equal the reference's f32 arithmetic bit for
 * bit; rotations use the host libm's sinf / cosf. Waits for the device. */
int mg_surface_extract(mg_surface* surface, const mg_surface_drum* drums, const float* d_fluid_x, char* err,
                       size_t errlen);
/* The last extraction's polygons of one drum flattened as flatten_polygons (surface.rs:420-430)
 * does, [n_polys, len_0, x, y, ..., len_1, ...] (mill-wasm lib.rs:145-152 fluidSurface): copies them
 * to out when they fit in cap floats; returns their length, or a negative MG_E_* code. */
int64_t mg_surface_polygons(mg_surface* surface, uint32_t drum, float* out, int64_t cap);
/* The last extraction's masked occupancy field of one drum (build_field, surface.rs:84-133):
 * 128 x 128 floats, row j = grid y index; zeros for a drum without fluid. */
int mg_surface_field(mg_surface* surface, uint32_t drum, float* out16384);

/* StressAccum::metrics (stress.rs:186-238) of one drum. */
int mg_stress_get_metrics(mg_stress* stress, uint32_t drum, uint32_t n_balls, double m3, double scale_j, double d,
                      double power_draw_w, mg_stress_metrics* out);

/* ---- device memory of this library's HIP runtime ----------------------------------------------
 * No reference counterpart: plumbing for callers without a HIP runtime of their own (bindings,
 * tests). Device pointers of another HIP runtime loaded into the same process (for example one
 * bundled with a framework) are not valid arguments of this library. */
int mg_device_malloc(size_t bytes, void** out);
int mg_device_free(void* ptr);
int mg_memcpy_htod(void* dst, const void* src, size_t bytes);
int mg_memcpy_dtoh(void* dst, const void* src, size_t bytes);
int mg_device_synchronize(void);

#ifdef __cplusplus
}
#endif
#endif /* MILLGPU_H */
