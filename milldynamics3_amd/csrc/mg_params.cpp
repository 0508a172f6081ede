// Params parsing, validation and derived sizes (crates/mill-core/src/params.rs).
// Compiled with -ffp-contract=off so every f32 expression rounds like the reference's Rust code.
#include "mg_params.h"

#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <vector>

namespace mg {
namespace {

const float PI_F = 3.14159265358979323846f;

// ---- a small JSON reader (RFC 8259 subset sufficient for Params documents) ----
struct JVal {
  enum Kind { Null, Bool, Num, Str, Arr, Obj } kind = Null;
  bool b = false;
  double num = 0.0;
  std::string text;  // number token / string value
  std::vector<JVal> arr;
  std::map<std::string, JVal> obj;
};

struct JParser {
  const char* p;
  std::string err;
  void ws() { while (*p && std::isspace((unsigned char)*p)) p++; }
  bool fail(const char* m) { if (err.empty()) err = m; return false; }
  bool parse_string(std::string& out) {
    if (*p != '"') return fail("expected string");
    p++;
    while (*p && *p != '"') {
      if (*p == '\\') {
        p++;
        char c = *p;
        if (!c) return fail("unterminated string");
        if (c == 'n') out += '\n'; else if (c == 't') out += '\t'; else if (c == 'r') out += '\r';
        else if (c == 'b') out += '\b'; else if (c == 'f') out += '\f';
        else if (c == 'u') {
          unsigned v = 0;
          for (int k = 0; k < 4; k++) {
            p++;
            if (!std::isxdigit((unsigned char)*p)) return fail("bad \\u escape");
            v = v * 16 + (unsigned)(std::isdigit((unsigned char)*p) ? *p - '0' : (std::tolower(*p) - 'a' + 10));
          }
          if (v < 0x80) out += (char)v; else out += '?';
        } else out += c;
        p++;
      } else {
        out += *p++;
      }
    }
    if (*p != '"') return fail("unterminated string");
    p++;
    return true;
  }
  bool parse(JVal& v) {
    ws();
    if (*p == '{') {
      p++;
      v.kind = JVal::Obj;
      ws();
      if (*p == '}') { p++; return true; }
      for (;;) {
        ws();
        std::string key;
        if (!parse_string(key)) return false;
        ws();
        if (*p != ':') return fail("expected ':'");
        p++;
        JVal child;
        if (!parse(child)) return false;
        v.obj[key] = std::move(child);
        ws();
        if (*p == ',') { p++; continue; }
        if (*p == '}') { p++; return true; }
        return fail("expected ',' or '}'");
      }
    }
    if (*p == '[') {
      p++;
      v.kind = JVal::Arr;
      ws();
      if (*p == ']') { p++; return true; }
      for (;;) {
        JVal child;
        if (!parse(child)) return false;
        v.arr.push_back(std::move(child));
        ws();
        if (*p == ',') { p++; continue; }
        if (*p == ']') { p++; return true; }
        return fail("expected ',' or ']'");
      }
    }
    if (*p == '"') { v.kind = JVal::Str; return parse_string(v.text); }
    if (!std::strncmp(p, "true", 4)) { p += 4; v.kind = JVal::Bool; v.b = true; return true; }
    if (!std::strncmp(p, "false", 5)) { p += 5; v.kind = JVal::Bool; v.b = false; return true; }
    if (!std::strncmp(p, "null", 4)) { p += 4; v.kind = JVal::Null; return true; }
    const char* s = p;
    if (*p == '-') p++;
    if (!std::isdigit((unsigned char)*p)) return fail("expected value");
    while (std::isdigit((unsigned char)*p) || *p == '.' || *p == 'e' || *p == 'E' || *p == '+' || *p == '-') p++;
    v.kind = JVal::Num;
    v.text.assign(s, p);
    v.num = std::strtod(v.text.c_str(), nullptr);
    return true;
  }
};

bool is_integer_token(const std::string& t) {
  for (char c : t) if (c == '.' || c == 'e' || c == 'E') return false;
  return true;
}

struct Reader {
  std::string err;
  const JVal* section(const JVal& root, const char* name) {
    auto it = root.obj.find(name);
    if (it == root.obj.end()) { if (err.empty()) err = std::string("missing field `") + name + "`"; return nullptr; }
    if (it->second.kind != JVal::Obj) { if (err.empty()) err = std::string("invalid type for `") + name + "`, expected struct"; return nullptr; }
    return &it->second;
  }
  const JVal* field(const JVal* obj, const char* name, bool required) {
    if (!obj) return nullptr;
    auto it = obj->obj.find(name);
    if (it == obj->obj.end()) {
      if (required && err.empty()) err = std::string("missing field `") + name + "`";
      return nullptr;
    }
    return &it->second;
  }
  void f32(const JVal* obj, const char* name, float& out, bool required = true) {
    const JVal* v = field(obj, name, required);
    if (!v) return;
    if (v->kind != JVal::Num) { if (err.empty()) err = std::string("invalid type for `") + name + "`, expected f32"; return; }
    out = (float)v->num;  // serde_json: decimal -> f64 -> f32
  }
  void u32(const JVal* obj, const char* name, uint32_t& out) {
    const JVal* v = field(obj, name, true);
    if (!v) return;
    if (v->kind != JVal::Num || !is_integer_token(v->text) || v->text[0] == '-' || v->num > 4294967295.0) {
      if (err.empty()) err = std::string("invalid value for `") + name + "`, expected u32";
      return;
    }
    out = (uint32_t)std::strtoull(v->text.c_str(), nullptr, 10);
  }
  void u64(const JVal* obj, const char* name, uint64_t& out) {
    const JVal* v = field(obj, name, true);
    if (!v) return;
    if (v->kind != JVal::Num || !is_integer_token(v->text) || v->text[0] == '-') {
      if (err.empty()) err = std::string("invalid value for `") + name + "`, expected u64";
      return;
    }
    out = std::strtoull(v->text.c_str(), nullptr, 10);
  }
  void boolean(const JVal* obj, const char* name, bool& out) {
    const JVal* v = field(obj, name, true);
    if (!v) return;
    if (v->kind != JVal::Bool) { if (err.empty()) err = std::string("invalid type for `") + name + "`, expected bool"; return; }
    out = v->b;
  }
  void enumeration(const JVal* obj, const char* name, int& out, std::initializer_list<const char*> names) {
    const JVal* v = field(obj, name, true);
    if (!v) return;
    if (v->kind != JVal::Str) { if (err.empty()) err = std::string("invalid type for `") + name + "`, expected string"; return; }
    int k = 0;
    for (const char* n : names) {
      if (v->text == n) { out = k; return; }
      k++;
    }
    if (err.empty()) err = "unknown variant `" + v->text + "`";
  }
};

bool in_closed(float x, float lo, float hi) { return x >= lo && x <= hi; }  // RangeInclusive::contains

std::string fmt(const char* s) { return std::string(s); }

}  // namespace

bool Params::from_json(const char* json, Params& out, std::string& err) {
  Params p;
  JParser jp{json, ""};
  JVal root;
  if (!jp.parse(root)) { err = jp.err; return false; }
  jp.ws();
  if (*jp.p) { err = "trailing characters"; return false; }
  if (root.kind != JVal::Obj) { err = "invalid type: expected struct Params"; return false; }
  Reader r;
  const JVal* m = r.section(root, "mill");
  const JVal* me = r.section(root, "media");
  const JVal* s = r.section(root, "slurry");
  const JVal* l = r.section(root, "lifters");
  const JVal* si = r.section(root, "simulation");
  r.f32(m, "diameter_m", p.mill_diameter_m);
  r.enumeration(m, "speed_mode", p.speed_mode, {"rpm", "percent_critical"});
  r.f32(m, "speed_value", p.speed_value);
  r.enumeration(m, "direction", p.direction, {"counter_clockwise", "clockwise"});
  r.f32(me, "ball_diameter_m", p.ball_diameter_m);
  r.f32(me, "fill_fraction", p.media_fill);
  r.f32(me, "packing_fraction_2d", p.packing);
  r.f32(me, "density_kg_m3", p.media_density);
  r.f32(me, "restitution_ball_ball", p.rest_bb);
  r.f32(me, "restitution_ball_wall", p.rest_bw);
  r.f32(me, "friction_ball_ball", p.fric_bb);
  r.f32(me, "friction_ball_wall", p.fric_bw);
  p.fric_bb_static = 0.25f;  // params.rs:180-197 serde defaults
  p.fric_bw_static = 0.35f;
  p.fric_vscale = 0.02f;
  r.f32(me, "friction_ball_ball_static", p.fric_bb_static, false);
  r.f32(me, "friction_ball_wall_static", p.fric_bw_static, false);
  r.f32(me, "friction_velocity_scale_m_s", p.fric_vscale, false);
  r.f32(me, "rolling_friction", p.rolling);
  r.boolean(s, "enabled", p.slurry_enabled);
  r.f32(s, "fill_fraction", p.slurry_fill);
  r.f32(s, "density_kg_m3", p.slurry_density);
  r.f32(s, "viscosity_pa_s", p.viscosity);
  r.enumeration(s, "rheology", p.rheology, {"newtonian", "bingham"});
  r.f32(s, "yield_stress_pa", p.yield_stress);
  r.f32(s, "wall_no_slip", p.wall_no_slip);
  r.f32(s, "ball_no_slip", p.ball_no_slip);
  r.f32(s, "wettability", p.wettability);
  p.surface_tension = 0.072f;  // params.rs:339-345 serde default
  r.f32(s, "surface_tension_n_m", p.surface_tension, false);
  r.enumeration(s, "dye_pattern", p.dye_pattern, {"left_right", "top_bottom", "none"});
  r.u32(l, "count", p.lifter_count);
  r.f32(l, "height_m", p.lifter_height);
  r.f32(l, "base_width_m", p.lifter_base);
  r.f32(l, "top_width_m", p.lifter_top);
  r.f32(l, "phase_deg", p.lifter_phase_deg);
  r.u32(si, "resolution", p.resolution);
  r.u32(si, "substeps", p.substeps);
  r.u32(si, "pbf_iterations", p.pbf_iterations);
  r.u32(si, "dem_iterations", p.dem_iterations);
  r.f32(si, "time_scale", p.time_scale);
  r.f32(si, "frame_budget_ms", p.frame_budget_ms);
  r.u32(si, "max_balls", p.max_balls);
  r.u64(si, "seed", p.seed);
  r.enumeration(si, "coarse_graining_mode", p.cg_mode, {"auto", "off", "manual"});
  r.f32(si, "coarse_graining_k", p.cg_k);
  if (!r.err.empty()) { err = r.err; return false; }
  out = p;
  return true;
}

std::string Params::to_json() const {
  char b[4096];
  const char* sm[] = {"rpm", "percent_critical"};
  const char* dr[] = {"counter_clockwise", "clockwise"};
  const char* rh[] = {"newtonian", "bingham"};
  const char* dp[] = {"left_right", "top_bottom", "none"};
  const char* cg[] = {"auto", "off", "manual"};
  std::snprintf(b, sizeof b,
                "{\"mill\":{\"diameter_m\":%.9g,\"speed_mode\":\"%s\",\"speed_value\":%.9g,\"direction\":\"%s\"},"
                "\"media\":{\"ball_diameter_m\":%.9g,\"fill_fraction\":%.9g,\"packing_fraction_2d\":%.9g,"
                "\"density_kg_m3\":%.9g,\"restitution_ball_ball\":%.9g,\"restitution_ball_wall\":%.9g,"
                "\"friction_ball_ball\":%.9g,\"friction_ball_wall\":%.9g,\"friction_ball_ball_static\":%.9g,"
                "\"friction_ball_wall_static\":%.9g,\"friction_velocity_scale_m_s\":%.9g,\"rolling_friction\":%.9g},"
                "\"slurry\":{\"enabled\":%s,\"fill_fraction\":%.9g,\"density_kg_m3\":%.9g,\"viscosity_pa_s\":%.9g,"
                "\"rheology\":\"%s\",\"yield_stress_pa\":%.9g,\"wall_no_slip\":%.9g,\"ball_no_slip\":%.9g,"
                "\"wettability\":%.9g,\"surface_tension_n_m\":%.9g,\"dye_pattern\":\"%s\"},"
                "\"lifters\":{\"count\":%u,\"height_m\":%.9g,\"base_width_m\":%.9g,\"top_width_m\":%.9g,\"phase_deg\":%.9g},"
                "\"simulation\":{\"resolution\":%u,\"substeps\":%u,\"pbf_iterations\":%u,\"dem_iterations\":%u,"
                "\"time_scale\":%.9g,\"frame_budget_ms\":%.9g,\"max_balls\":%u,\"seed\":%llu,"
                "\"coarse_graining_mode\":\"%s\",\"coarse_graining_k\":%.9g}}",
                mill_diameter_m, sm[speed_mode], speed_value, dr[direction], ball_diameter_m, media_fill, packing,
                media_density, rest_bb, rest_bw, fric_bb, fric_bw, fric_bb_static, fric_bw_static, fric_vscale, rolling,
                slurry_enabled ? "true" : "false", slurry_fill, slurry_density, viscosity, rh[rheology], yield_stress,
                wall_no_slip, ball_no_slip, wettability, surface_tension, dp[dye_pattern], lifter_count, lifter_height,
                lifter_base, lifter_top, lifter_phase_deg, resolution, substeps, pbf_iterations, dem_iterations,
                time_scale, frame_budget_ms, max_balls, (unsigned long long)seed, cg[cg_mode], cg_k);
  return b;
}

bool Params::validate(std::string& e) const {
  // MillParams::validate (params.rs:115-126)
  if (!(mill_diameter_m > 0.02f && mill_diameter_m < 20.0f)) { e = fmt("mill.diameter_m must be in (0.02, 20.0)"); return false; }
  if (!(std::isfinite(speed_value) && speed_value >= 0.0f)) { e = fmt("mill.speed_value must be finite and >= 0"); return false; }
  // MediaParams::validate (params.rs:231-278)
  if (!(ball_diameter_m > 0.0005f && ball_diameter_m < 1.0f)) { e = fmt("media.ball_diameter_m must be in (0.0005, 1.0)"); return false; }
  if (!in_closed(media_fill, 0.0f, 0.9f)) { e = fmt("media.fill_fraction must be in [0, 0.9]"); return false; }
  if (!in_closed(packing, 0.5f, 0.907f)) { e = fmt("media.packing_fraction_2d must be in [0.5, 0.907]"); return false; }
  if (!(std::isfinite(media_density) && media_density > 0.0f)) { e = fmt("media.density_kg_m3 must be > 0"); return false; }
  if (!in_closed(rest_bb, 0.0f, 1.0f)) { e = fmt("media.restitution_ball_ball must be in [0, 1]"); return false; }
  if (!in_closed(rest_bw, 0.0f, 1.0f)) { e = fmt("media.restitution_ball_wall must be in [0, 1]"); return false; }
  const std::pair<const char*, float> mus[] = {{"friction_ball_ball", fric_bb}, {"friction_ball_wall", fric_bw},
                                               {"friction_ball_ball_static", fric_bb_static},
                                               {"friction_ball_wall_static", fric_bw_static}, {"rolling_friction", rolling}};
  for (auto& m : mus)
    if (!(m.second >= 0.0f && std::isfinite(m.second))) { e = std::string("media.") + m.first + " must be >= 0"; return false; }
  if (fric_bb_static < fric_bb) { e = fmt("media.friction_ball_ball_static must be >= media.friction_ball_ball"); return false; }
  if (fric_bw_static < fric_bw) { e = fmt("media.friction_ball_wall_static must be >= media.friction_ball_wall"); return false; }
  if (!(fric_vscale > 0.0f && std::isfinite(fric_vscale))) { e = fmt("media.friction_velocity_scale_m_s must be > 0"); return false; }
  // SlurryParams::validate (params.rs:379-405)
  if (!in_closed(slurry_fill, 0.0f, 0.9f)) { e = fmt("slurry.fill_fraction must be in [0, 0.9]"); return false; }
  if (!(std::isfinite(slurry_density) && slurry_density > 0.0f)) { e = fmt("slurry.density_kg_m3 must be > 0"); return false; }
  if (!(viscosity >= 0.0f && std::isfinite(viscosity))) { e = fmt("slurry.viscosity_pa_s must be >= 0"); return false; }
  if (!(yield_stress >= 0.0f && std::isfinite(yield_stress))) { e = fmt("slurry.yield_stress_pa must be >= 0"); return false; }
  if (!(surface_tension >= 0.0f && std::isfinite(surface_tension))) { e = fmt("slurry.surface_tension_n_m must be >= 0"); return false; }
  if (!in_closed(wall_no_slip, 0.0f, 1.0f)) { e = fmt("slurry.wall_no_slip must be in [0, 1]"); return false; }
  if (!in_closed(ball_no_slip, 0.0f, 1.0f)) { e = fmt("slurry.ball_no_slip must be in [0, 1]"); return false; }
  if (!in_closed(wettability, 0.0f, 1.0f)) { e = fmt("slurry.wettability must be in [0, 1]"); return false; }
  // LiftersParams::validate (params.rs:432-445)
  if (lifter_count > 0) {
    if (!(lifter_height > 0.0f && std::isfinite(lifter_height))) { e = fmt("lifters.height_m must be > 0"); return false; }
    if (!(lifter_base > 0.0f && std::isfinite(lifter_base))) { e = fmt("lifters.base_width_m must be > 0"); return false; }
    if (!(lifter_top >= 0.0f && std::isfinite(lifter_top))) { e = fmt("lifters.top_width_m must be >= 0"); return false; }
  }
  // SimulationParams::validate (params.rs:504-530)
  if (!(resolution >= 4 && resolution <= 200)) { e = fmt("simulation.resolution must be in [4, 200]"); return false; }
  if (!(substeps >= 1 && substeps <= 64)) { e = fmt("simulation.substeps must be in [1, 64]"); return false; }
  if (!(pbf_iterations >= 1 && pbf_iterations <= 20)) { e = fmt("simulation.pbf_iterations must be in [1, 20]"); return false; }
  if (!(dem_iterations >= 1 && dem_iterations <= 20)) { e = fmt("simulation.dem_iterations must be in [1, 20]"); return false; }
  if (!(time_scale > 0.0f && std::isfinite(time_scale))) { e = fmt("simulation.time_scale must be > 0"); return false; }
  if (!(frame_budget_ms > 0.0f && std::isfinite(frame_budget_ms))) { e = fmt("simulation.frame_budget_ms must be > 0"); return false; }
  if (!(max_balls >= 10 && max_balls <= 50000)) { e = fmt("simulation.max_balls must be in [10, 50000]"); return false; }
  if (!(cg_k >= 1.0f && cg_k <= 1000.0f)) { e = fmt("simulation.coarse_graining_k must be in [1, 1000]"); return false; }
  return true;
}

float Params::critical_speed_rpm() const { return 42.3f / std::sqrt(mill_diameter_m); }
float Params::rpm() const { return speed_mode == 0 ? speed_value : critical_speed_rpm() * speed_value / 100.0f; }
float Params::omega() const {
  float m = rpm() * 2.0f * PI_F / 60.0f;
  return direction == 0 ? m : -m;
}
float Params::radius_m() const { return mill_diameter_m * 0.5f; }
float Params::true_ball_count() const {
  float rt = ball_diameter_m * 0.5f;
  if (rt <= 0.0f) return 0.0f;
  float rr = radius_m();
  float drum_area = PI_F * (rr * rr);
  return (media_fill * packing * drum_area) / (PI_F * rt * rt);
}
static uint32_t as_u32(float x) {
  if (std::isnan(x) || x <= 0.0f) return 0;
  if (x >= 4294967295.0f) return 0xFFFFFFFFu;
  return (uint32_t)x;
}
void Params::effective_media(float& true_d, float& d, float& density, uint32_t& count, float& scale) const {
  true_d = ball_diameter_m;
  density = media_density;
  float n_true = true_ball_count();
  float mb = (float)max_balls;
  float sf = 1.0f;
  if (cg_mode == 2) sf = std::fmax(cg_k, 1.0f);
  else if (cg_mode == 0) sf = (n_true > mb && mb > 0.0f) ? std::sqrt(n_true / mb) : 1.0f;
  if (sf > 1.0f) {
    count = as_u32(std::fmax(std::round(n_true / (sf * sf)), 1.0f));
    d = true_d * sf;
    scale = sf;
  } else {
    count = as_u32(std::fmax(std::round(n_true), 0.0f));
    d = true_d;
    scale = 1.0f;
  }
}
uint32_t Params::effective_fluid_resolution() const {
  float td, d, de, sf;
  uint32_t c;
  effective_media(td, d, de, c, sf);
  float r = radius_m();
  if (!(r > 0.0f && d > 0.0f)) return resolution;
  uint32_t mn = as_u32(std::ceil(r / d));
  uint32_t v = resolution > mn ? resolution : mn;
  return v < 200u ? v : 200u;
}
uint32_t Params::effective_substeps() const {
  float td, d, de, sf;
  uint32_t c;
  effective_media(td, d, de, c, sf);
  if (!(d > 0.0f && mill_diameter_m > 0.0f)) return substeps;
  float df = 1.0f + 2.5f * (media_fill - 0.30f);
  if (df < 1.0f) df = 1.0f;
  if (df > 2.0f) df = 2.0f;
  float v_ref = std::sqrt(2.0f * 9.81f * mill_diameter_m);
  uint32_t mn = as_u32(std::fmax(std::ceil(v_ref * df / (60.0f * 0.5f * d)), 1.0f));
  uint32_t cap = as_u32(std::ceil(16.0f * df));
  uint32_t lim = mn < cap ? mn : cap;
  return substeps > lim ? substeps : lim;
}

}  // namespace mg
