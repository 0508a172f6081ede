// Params (crates/mill-core/src/params.rs:1-745) for the host runtime: serde-JSON parsing with the
// reference's field names, enum spellings and `#[serde(default = ...)]` fields, `validate` with
// the reference's messages, and the derived sizes (effective media, fluid resolution, substeps).
#pragma once
#include <stdint.h>

#include <string>

namespace mg {

struct Params {
  // mill (params.rs:45-71)
  float mill_diameter_m = 0.063f;
  int speed_mode = 0;  // 0 rpm, 1 percent_critical
  float speed_value = 100.0f;
  int direction = 0;  // 0 counter_clockwise, 1 clockwise
  // media (params.rs:150-228)
  float ball_diameter_m = 0.002f, media_fill = 0.50f, packing = 0.82f, media_density = 6000.0f;
  float rest_bb = 0.7f, rest_bw = 0.5f, fric_bb = 0.25f, fric_bw = 0.35f;
  float fric_bb_static = 0.25f, fric_bw_static = 0.35f, fric_vscale = 0.02f, rolling = 0.01f;
  // slurry (params.rs:303-376)
  bool slurry_enabled = true;
  float slurry_fill = 0.35f, slurry_density = 1000.0f, viscosity = 0.001f;
  int rheology = 0;  // 0 newtonian, 1 bingham
  float yield_stress = 0.0f, wall_no_slip = 1.0f, ball_no_slip = 1.0f, wettability = 0.6f;
  float surface_tension = 0.072f;
  int dye_pattern = 0;  // 0 left_right, 1 top_bottom, 2 none
  // lifters (params.rs:409-429)
  uint32_t lifter_count = 0;
  float lifter_height = 0.020f, lifter_base = 0.030f, lifter_top = 0.020f, lifter_phase_deg = 0.0f;
  // simulation (params.rs:455-501)
  uint32_t resolution = 40, substeps = 8, pbf_iterations = 3, dem_iterations = 2;
  float time_scale = 1.0f, frame_budget_ms = 12.0f;
  uint32_t max_balls = 600;
  uint64_t seed = 1;
  int cg_mode = 1;  // 0 auto, 1 off, 2 manual
  float cg_k = 1.0f;

  // Parses serde JSON (all non-defaulted fields required, unknown fields ignored).
  static bool from_json(const char* json, Params& out, std::string& err);
  std::string to_json() const;
  bool validate(std::string& err) const;  // params.rs:115-530

  float critical_speed_rpm() const;
  float rpm() const;
  float omega() const;
  float radius_m() const;
  float true_ball_count() const;
  void effective_media(float& true_d, float& d, float& density, uint32_t& count, float& scale) const;
  uint32_t effective_fluid_resolution() const;
  uint32_t effective_substeps() const;
};

}  // namespace mg
