// Host runtime and C ABI (include/millgpu.h) of the drum engine: mill_core::Simulation
// (crates/mill-core/src/lib.rs:82-566) for one drum or a batch of drums.
//
// The initial state is seeded on the host exactly like the reference (dem.rs:355-409,
// pbf.rs:1044-1119, rng.rs:9-47) and uploaded once; a batch keeps every drum's state in device
// arrays (drum d's fluid and balls after drum d-1's). One substep (advance_one_sub_step,
// lib.rs:470-509, and update_grinding_stats, lib.rs:516-565) is, for all drums of the batch at once:
// FluidParticles::step_coupled (the fluid operator, mg_fluid.hip), DemState::step_recorded with the
// fluid's coupling impulses (the DEM operator, mg_dem.hip), the f32 power / torque / dissipation
// EMAs and the f64 energy budget on the host from the two steps' statistics, StressAccum::update
// with the DEM's impact records (mg_stress.hip), and the drum clock. The state never leaves the
// device between substeps; the host reads back only the per-drum statistics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/millgpu.h"
#include "mg_common.h"
#include "mg_internal.h"
#include "mg_params.h"

namespace mg {
namespace {

void set_err(char* err, size_t errlen, const std::string& m) {
  if (err && errlen) {
    size_t n = std::min(errlen - 1, m.size());
    std::memcpy(err, m.data(), n);
    err[n] = 0;
  }
}

int32_t as_i32(float x) {
  if (std::isnan(x)) return 0;
  if (x <= -2147483648.0f) return INT32_MIN;
  if (x >= 2147483647.0f) return INT32_MAX;
  return (int32_t)x;
}
uint64_t as_u64(float x) {
  if (std::isnan(x) || x <= 0.0f) return 0;
  if (x >= 1.8446744073709552e19f) return UINT64_MAX;
  return (uint64_t)x;
}

// Balls::seed_lattice (dem.rs:355-409).
void seed_balls(const Params& P, std::vector<float>& bx, float& r_out, float& mass, float& inertia, float& true_r) {
  float td, d, dens, sf;
  uint32_t count;
  P.effective_media(td, d, dens, count, sf);
  float r = d * 0.5f;
  float rr = d * 0.5f;
  mass = dens * PI_F * rr * rr;
  inertia = 0.5f * mass * r * r;
  r_out = r;
  true_r = td * 0.5f;
  float R = P.radius_m();
  Rng rng(P.seed);
  bx.clear();
  if (r > 0.0f && count > 0) {
    float spacing = 2.0f * r * 1.02f;
    float row_h = spacing * (std::sqrt(3.0f) * 0.5f);
    float fill_r = std::fmax(R - r, 0.0f);
    int32_t n_rows = as_i32(std::ceil(2.0f * fill_r / row_h)) + 1;
    for (int32_t row = -n_rows / 2; row <= n_rows / 2; row++) {
      float y = (float)row * row_h;
      if (std::fabs(y) > fill_r) continue;
      float hw = std::sqrt(std::fmax(fill_r * fill_r - y * y, 0.0f));
      float off = (((row % 2) + 2) % 2 == 0) ? 0.0f : spacing * 0.5f;
      int32_t n_cols = as_i32(std::floor(2.0f * hw / spacing));
      bool done = false;
      for (int32_t col = -n_cols / 2; col <= n_cols / 2; col++) {
        float cx = (float)col * spacing + off;
        if (cx * cx + y * y > fill_r * fill_r) continue;
        float jx = rng.range(-0.05f, 0.05f);
        float jy = rng.range(-0.05f, 0.05f);
        bx.push_back(cx + jx * r);
        bx.push_back(y + jy * r);
        if (bx.size() / 2 >= count) { done = true; break; }
      }
      if (done) break;
    }
  }
}

// FluidParticles::seed_lattice (pbf.rs:1044-1119) and dye_value (pbf.rs:2369-2388).
void seed_fluid(const Params& P, float fill, const std::vector<float>& balls, float ball_r, std::vector<float>& fx,
                std::vector<float>& dye, float& h, float& mass, float& rest) {
  float R = P.radius_m();
  uint32_t res = P.effective_fluid_resolution();
  float dx = R / (float)std::max(res, 1u);
  h = 2.0f * dx;
  rest = P.slurry_density;
  float row_h = dx * (std::sqrt(3.0f) * 0.5f);
  float cell_area = dx * row_h;
  mass = rest * cell_area;
  float target_area = fill * PI_F * R * R;
  uint64_t target = as_u64(std::fmax(std::round(target_area / cell_area), 0.0f));
  fx.clear();
  dye.clear();
  if (target > 0 && dx > 0.0f) {
    int32_t n_rows = as_i32(std::ceil(2.0f * R / row_h)) + 1;
    for (int32_t row = -n_rows / 2; row <= n_rows / 2; row++) {
      float y = (float)row * row_h;
      if (std::fabs(y) > R) continue;
      float hw = std::sqrt(std::fmax(R * R - y * y, 0.0f));
      float off = (((row % 2) + 2) % 2 == 0) ? 0.0f : dx * 0.5f;
      int32_t n_cols = as_i32(std::floor(2.0f * hw / dx));
      bool done = false;
      for (int32_t col = -n_cols / 2; col <= n_cols / 2; col++) {
        float cx = (float)col * dx + off;
        if (cx * cx + y * y > R * R) continue;
        bool ov = false;
        for (size_t b = 0; b < balls.size() / 2; b++) {
          float ex = cx - balls[2 * b], ey = y - balls[2 * b + 1];
          if (std::sqrt(ex * ex + ey * ey) < ball_r) { ov = true; break; }
        }
        if (ov) continue;
        float dv = 0.0f;
        if (P.dye_pattern == 0) dv = cx < 0.0f ? 0.0f : 1.0f;
        else if (P.dye_pattern == 1) dv = y < 0.0f ? 0.0f : 1.0f;
        dye.push_back(dv);
        fx.push_back(cx);
        fx.push_back(y);
        if (fx.size() / 2 >= target) { done = true; break; }
      }
      if (done) break;
    }
  }
}

}  // namespace

// One drum: Params, sizes, clock, accumulators and the last substep's statistics (host side).
struct Drum {
  Params P;
  uint32_t nf = 0, nb = 0, f0 = 0, b0 = 0;
  float h = 0.0f, fmass = 0.0f, rest = 0.0f, br = 0.0f, bmass = 0.0f, binertia = 0.0f, btrue = 0.0f;
  float drum_angle = 0.0f;
  double sim_time = 0.0;
  float power = 0.0f, torque = 0.0f, diss_power = 0.0f, max_disp = 0.0f;
  uint32_t clamp_hits_total = 0;
  // EnergyBudget (lib.rs:61-78): elapsed, shaft, dem wall, fluid wall, contact dissipation, wall
  // slip, viscous, clamp removed, interface, fluid wall angular impulse
  double eb[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  double e0 = 0.0;
  bool has_guess = false;
  bool forced = false;
  uint32_t force_density = 0, force_cg = 0;
  std::vector<float> dye;
  std::vector<float> init_bx, init_bv, init_bom, init_fx, init_fv;  // until uploaded
  mg_fluid_stats fs{};
  mg_dem_stats ds{};
};

struct Batch {
  std::vector<std::unique_ptr<Drum>> drums;
  uint64_t TF = 0, TB = 0;
  char* dev = nullptr;
  float *fx = nullptr, *fv = nullptr, *fg = nullptr, *bx = nullptr, *bv = nullptr, *bth = nullptr, *bom = nullptr;
  float *imp = nullptr, *ang = nullptr;
  uint32_t *rbeg = nullptr, *rend = nullptr;
  uint32_t rstride = 0;
  mg_fluid* fl = nullptr;
  mg_dem* dem = nullptr;
  mg_stress* st = nullptr;
  float last_ms = 0.0f;
  int last_rc = MG_OK;
  std::string msg;
  std::vector<struct ::mg_sim*> views;
  std::vector<mg_fluid_drum> fd;
  std::vector<mg_dem_drum> dd;
  std::vector<mg_fluid_stats> fs;
  std::vector<mg_dem_stats> ds;
  std::vector<mg_stress_step> ss;
  std::vector<uint32_t> hbeg, hend;
  // The operators' handles may hold more drum slots than the batch has drums (a server's batch
  // changes between substeps): slots past the drums step empty drums, and drum d's stress
  // accumulator is slot sslot[d] of the stress handle (its records start at hbeg[sslot[d]]).
  std::vector<uint32_t> sslot;
  mg_fluid_drum fd_empty{};
  mg_dem_drum dd_empty{};
  bool ranges_dirty = false;
  ~Batch();
};

}  // namespace mg

struct mg_sim {
  mg::Batch* batch;
  uint32_t index;
  bool owns_batch;
};
struct mg_batch {
  mg::Batch* b;
};

namespace mg {

Batch::~Batch() {
  if (fl) mg_fluid_destroy(fl);
  if (dem) mg_dem_destroy(dem);
  if (st) mg_stress_destroy(st);
  if (dev) (void)hipFree(dev);
  for (auto* v : views) delete v;
}

namespace {

// Mechanical energy in f64 (dem.rs:265-273 + pbf.rs:1145-1150) from host copies.
double mech_energy(const Drum& D, const std::vector<float>& bx, const std::vector<float>& bv,
                   const std::vector<float>& bom, const std::vector<float>& fx, const std::vector<float>& fv) {
  double e = 0.0;
  for (uint32_t i = 0; i < D.nb; i++) {
    const float vx = bv[2 * i], vy = bv[2 * i + 1];
    e += 0.5 * (double)D.bmass * (double)(vx * vx + vy * vy);
    e += 0.5 * (double)D.binertia * (double)(bom[i] * bom[i]);
    e += (double)D.bmass * (double)9.81f * (double)bx[2 * i + 1];
  }
  double py = 0.0, ke = 0.0;
  for (uint32_t i = 0; i < D.nf; i++) py += (double)fx[2 * i + 1];
  for (uint32_t i = 0; i < D.nf; i++) {
    const float vx = fv[2 * i], vy = fv[2 * i + 1];
    ke += (double)(vx * vx + vy * vy);
  }
  return e + (0.5 * (double)D.fmass * ke + py * (double)D.fmass * (double)9.81f);
}

// Simulation::build (lib.rs:138-207): seeded balls (or the caller's), seeded fluid, sizes.
std::unique_ptr<Drum> make_drum(const Params& P, const float* init, size_t n_init) {
  auto D = std::make_unique<Drum>();
  D->P = P;
  std::vector<float> bx, fx;
  seed_balls(P, bx, D->br, D->bmass, D->binertia, D->btrue);
  if (init) bx.assign(init, init + 2 * n_init);
  if (P.slurry_enabled) seed_fluid(P, P.slurry_fill, bx, D->br, fx, D->dye, D->h, D->fmass, D->rest);
  else seed_fluid(P, 0.0f, {}, 0.0f, fx, D->dye, D->h, D->fmass, D->rest);
  D->nb = (uint32_t)(bx.size() / 2);
  D->nf = (uint32_t)(fx.size() / 2);
  std::vector<float> bv(bx.size(), 0.0f), bom(D->nb, 0.0f), fv(fx.size(), 0.0f);
  if (init) {  // rigid rotation with the drum (lib.rs:134-136)
    const float w = P.omega();
    for (uint32_t i = 0; i < D->nb; i++) {
      bv[2 * i] = w * -bx[2 * i + 1];
      bv[2 * i + 1] = w * bx[2 * i];
      bom[i] = w;
    }
    for (uint32_t i = 0; i < D->nf; i++) {
      fv[2 * i] = w * -fx[2 * i + 1];
      fv[2 * i + 1] = w * fx[2 * i];
    }
  }
  D->e0 = mech_energy(*D, bx, bv, bom, fx, fv);
  D->init_bx = std::move(bx);
  D->init_bv = std::move(bv);
  D->init_bom = std::move(bom);
  D->init_fx = std::move(fx);
  D->init_fv = std::move(fv);
  return D;
}

// Lays the drums out in the batch's device arrays, uploads their initial states and creates the
// three operators' workspaces.
Batch* make_batch(std::vector<std::unique_ptr<Drum>>&& drums, std::string& err) {
  auto B = std::make_unique<Batch>();
  B->drums = std::move(drums);
  const uint32_t nd = (uint32_t)B->drums.size();
  uint32_t max_nb = 1;
  for (auto& D : B->drums) {
    D->f0 = (uint32_t)B->TF;
    D->b0 = (uint32_t)B->TB;
    B->TF += D->nf;
    B->TB += D->nb;
    max_nb = std::max(max_nb, D->nb);
  }
  if (max_nb > 2048) {
    err = "more than 2048 balls in a drum: the DEM step keeps a drum's balls in LDS";
    return nullptr;
  }
  const size_t TF = B->TF, TB = B->TB;
  auto up = [](size_t b) { return ((b + 255) / 256) * 256 + 256; };
  const size_t bytes = 3 * up(TF * 8) + 2 * up(TB * 8) + 2 * up(TB * 4) + up(TB * 8) + up(TB * 4) + 2 * up(nd * 4);
  if (hipMalloc(&B->dev, bytes) != hipSuccess) {
    err = "hipMalloc failed for the batch state";
    return nullptr;
  }
  char* p = B->dev;
  auto take = [&](size_t b) {
    char* q = p;
    p += up(b);
    return q;
  };
  B->fx = (float*)take(TF * 8);
  B->fv = (float*)take(TF * 8);
  B->fg = (float*)take(TF * 8);
  B->bx = (float*)take(TB * 8);
  B->bv = (float*)take(TB * 8);
  B->bth = (float*)take(TB * 4);
  B->bom = (float*)take(TB * 4);
  B->imp = (float*)take(TB * 8);
  B->ang = (float*)take(TB * 4);
  B->rbeg = (uint32_t*)take(nd * 4);
  B->rend = (uint32_t*)take(nd * 4);
  bool ok = hipMemset(B->dev, 0, bytes) == hipSuccess;
  for (auto& D : B->drums) {
    if (!ok) break;
    if (D->nf) {
      ok &= hipMemcpy(B->fx + 2 * (size_t)D->f0, D->init_fx.data(), D->nf * 8, hipMemcpyHostToDevice) == hipSuccess;
      ok &= hipMemcpy(B->fv + 2 * (size_t)D->f0, D->init_fv.data(), D->nf * 8, hipMemcpyHostToDevice) == hipSuccess;
    }
    if (D->nb) {
      ok &= hipMemcpy(B->bx + 2 * (size_t)D->b0, D->init_bx.data(), D->nb * 8, hipMemcpyHostToDevice) == hipSuccess;
      ok &= hipMemcpy(B->bv + 2 * (size_t)D->b0, D->init_bv.data(), D->nb * 8, hipMemcpyHostToDevice) == hipSuccess;
      ok &= hipMemcpy(B->bom + D->b0, D->init_bom.data(), D->nb * 4, hipMemcpyHostToDevice) == hipSuccess;
    }
    std::vector<float>().swap(D->init_bx);
    std::vector<float>().swap(D->init_bv);
    std::vector<float>().swap(D->init_bom);
    std::vector<float>().swap(D->init_fx);
    std::vector<float>().swap(D->init_fv);
  }
  if (!ok) {
    err = "HIP error uploading the initial states";
    return nullptr;
  }
  char e[512] = {0};
  if (mg_fluid_create(nd, &B->fl, e, sizeof e) != MG_OK ||
      mg_dem_create(nd, max_nb, 24, 4 * max_nb + 64, &B->dem, e, sizeof e) != MG_OK ||
      mg_stress_create(nd, 4 * max_nb + 1024, &B->st, e, sizeof e) != MG_OK) {
    err = e;
    return nullptr;
  }
  mg_dem_records_device(B->dem, &B->rstride);
  std::vector<uint32_t> beg(nd);
  for (uint32_t d = 0; d < nd; d++) beg[d] = d * B->rstride;
  if (hipMemcpy(B->rbeg, beg.data(), nd * 4, hipMemcpyHostToDevice) != hipSuccess) {
    err = "HIP copy error";
    return nullptr;
  }
  B->fd.resize(nd);
  B->dd.resize(nd);
  B->fs.resize(nd);
  B->ds.resize(nd);
  B->ss.resize(nd);
  B->hbeg = beg;
  B->hend.resize(nd);
  B->sslot.resize(nd);
  for (uint32_t d = 0; d < nd; d++) B->sslot[d] = d;
  Batch* out = B.release();
  for (uint32_t i = 0; i < nd; i++) out->views.push_back(new ::mg_sim{out, i, false});
  return out;
}

// One substep of every drum (lib.rs:470-565), drum d with sub-step dt[d].
int substep(Batch& B, const std::vector<float>& dt) {
  const uint32_t nd = (uint32_t)B.drums.size();
  char e[512] = {0};
  for (size_t d = nd; d < B.fd.size(); d++) {  // empty slots of a server's operators
    B.fd[d] = B.fd_empty;
    B.dd[d] = B.dd_empty;
  }
  for (size_t q = 0; q < B.ss.size(); q++) {  // stress slots without a drum: no records, no decay
    B.ss[q] = mg_stress_step{};
    B.ss[q].sub_dt = 1.0;
    B.hend[q] = B.hbeg[q];
  }
  const Drum* forced = nullptr;
  for (uint32_t d = 0; d < nd; d++) {
    const Drum& D = *B.drums[d];
    const Params& P = D.P;
    if (D.forced && !forced) forced = &D;
    mg_fluid_drum& f = B.fd[d];
    f.radius_m = P.radius_m();
    f.omega = P.omega();
    f.drum_angle = D.drum_angle;
    f.lifter_count = P.lifter_count;
    f.lifter_height_m = P.lifter_height;
    f.lifter_base_width_m = P.lifter_base;
    f.lifter_top_width_m = P.lifter_top;
    f.lifter_phase_deg = P.lifter_phase_deg;
    f.h = D.h;
    f.rest_density = D.rest;
    f.particle_mass = D.fmass;
    f.ball_radius = D.br;
    f.ball_mass = D.bmass;
    f.ball_inertia = D.binertia;
    f.ball_true_radius = D.btrue;
    f.viscosity_pa_s = P.viscosity;
    f.rheology = (uint32_t)P.rheology;
    f.yield_stress_pa = P.yield_stress;
    f.wall_no_slip = P.wall_no_slip;
    f.ball_no_slip = P.ball_no_slip;
    f.wettability = P.wettability;
    f.surface_tension_n_m = P.surface_tension;
    f.iterations = P.pbf_iterations;
    f.dt = dt[d];
    f.n_fluid = D.nf;
    f.n_balls = D.nb;
    f.has_guess = D.has_guess ? 1 : 0;
    mg_dem_drum& m = B.dd[d];
    m.radius_m = f.radius_m;
    m.omega = f.omega;
    m.drum_angle = D.drum_angle;
    m.lifter_count = P.lifter_count;
    m.lifter_height_m = P.lifter_height;
    m.lifter_base_width_m = P.lifter_base;
    m.lifter_top_width_m = P.lifter_top;
    m.lifter_phase_deg = P.lifter_phase_deg;
    m.ball_radius = D.br;
    m.ball_mass = D.bmass;
    m.ball_inertia = D.binertia;
    m.restitution_ball_ball = P.rest_bb;
    m.restitution_ball_wall = P.rest_bw;
    m.friction_ball_ball = P.fric_bb;
    m.friction_ball_wall = P.fric_bw;
    m.friction_ball_ball_static = P.fric_bb_static;
    m.friction_ball_wall_static = P.fric_bw_static;
    m.friction_velocity_scale_m_s = P.fric_vscale;
    m.rolling_friction = P.rolling;
    m.iterations = P.dem_iterations;
    m.dt = dt[d];
    m.n_balls = D.nb;
    m.has_external = 1;  // coupling::step_recorded always hands the fluid's impulses over (coupling.rs:90-95)
    m.energy_f64 = 1;    // f64 energy and wall-work accumulators (the precision form)
  }
  if (forced) mg_fluid_force_iterations(B.fl, 1, forced->force_density, forced->force_cg);
  else mg_fluid_force_iterations(B.fl, 0, 0, 0);
  int rc = mg_fluid_step(B.fl, B.fd.data(), B.fx, B.fv, B.fg, B.bx, B.bv, B.bom, B.imp, B.ang, B.fs.data(), e, sizeof e);
  if (rc != MG_OK) {
    B.msg = std::string("fluid step: ") + e;
    return rc;
  }
  rc = mg_dem_step(B.dem, B.dd.data(), B.bx, B.bv, B.bth, B.bom, B.imp, B.ang, B.ds.data(), e, sizeof e);
  if (rc != MG_OK) {
    B.msg = std::string("DEM step: ") + e;
    return rc;
  }
  for (uint32_t d = 0; d < nd; d++) {
    Drum& D = *B.drums[d];
    const mg_fluid_stats& fs = B.fs[d];
    const mg_dem_stats& ds = B.ds[d];
    D.fs = fs;
    D.ds = ds;
    if (fs.joint_solve) D.has_guess = true;
    const float sd = dt[d];
    const float omega = D.P.omega();
    D.clamp_hits_total += fs.coupling_clamp_hits;
    D.eb[0] += (double)sd;
    D.eb[2] += (double)ds.wall_work_j;
    D.eb[3] += (double)fs.wall_work_j;
    D.eb[1] += (double)(float)(ds.wall_work_j + fs.wall_work_j);
    D.eb[4] += (double)ds.dissipated_energy_j;
    D.eb[5] += (double)(fs.wall_work_j - (float)fs.ke_delta_noslip_j);
    D.eb[6] -= (double)(float)fs.ke_delta_viscosity_j;
    D.eb[7] -= (double)(float)fs.ke_delta_clamp_j;
    D.eb[8] += (double)(ds.coupling_work_j + (float)fs.ke_delta_drag_j + (float)fs.ke_delta_buoyancy_j);
    D.eb[9] += (double)(float)fs.wall_angular_impulse;
    // update_grinding_stats (lib.rs:516-565)
    const float alpha = 1.0f - std::exp(-sd / 1.0f);
    const float ww = ds.wall_work_j + fs.wall_work_j;
    const float pinst = ww / sd;
    D.power += alpha * (pinst - D.power);
    const float tinst = std::fabs(omega) > 1e-6f ? pinst / omega : 0.0f;
    D.torque += alpha * (tinst - D.torque);
    const double rr = (double)D.br;
    const uint32_t slot = B.sslot[d];
    mg_stress_step& s = B.ss[slot];
    s.m3 = (double)D.bmass * 4.0 / 3.0 * rr;
    s.d = 2.0 * rr;
    s.scale_j = s.m3 * 9.81 * s.d;
    s.now_s = D.sim_time;
    s.sub_dt = (double)sd;
    s.alpha = (double)alpha;
    s.wall_work_j = (double)ww;
    B.hend[slot] = B.hbeg[slot] + std::min(ds.n_impacts, B.rstride);
    const float dpi = std::fmax(ds.dissipated_energy_j, 0.0f) / sd;
    D.diss_power += alpha * (dpi - D.diss_power);
    D.max_disp = std::fmax(D.max_disp, ds.max_substep_displacement_over_diameter);
    const float a = std::fmod(D.drum_angle + omega * sd, TAU_F);  // rem_euclid (lib.rs:507)
    D.drum_angle = a < 0.0f ? a + std::fabs(TAU_F) : a;
    D.sim_time += (double)sd;
  }
  if ((B.ranges_dirty &&
       hipMemcpy(B.rbeg, B.hbeg.data(), B.hbeg.size() * 4, hipMemcpyHostToDevice) != hipSuccess) ||
      hipMemcpy(B.rend, B.hend.data(), B.hend.size() * 4, hipMemcpyHostToDevice) != hipSuccess) {
    B.msg = "HIP copy error";
    return MG_E_DEVICE;
  }
  B.ranges_dirty = false;
  uint32_t stride = 0;
  const mg_impact_record* recs = mg_dem_records_device(B.dem, &stride);
  rc = mg_stress_update_ranges(B.st, recs, B.rbeg, B.rend, B.ss.data(), e, sizeof e);
  if (rc != MG_OK) {
    B.msg = std::string("stress update: ") + e;
    return rc;
  }
  return MG_OK;
}

// k substeps of every drum; dt_each < 0 takes each drum's fixed sub-step (lib.rs:426-438).
int run_substeps(Batch& B, uint32_t k, float dt_each) {
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<float> dt(B.drums.size());
  for (size_t d = 0; d < B.drums.size(); d++)
    dt[d] = dt_each >= 0.0f ? dt_each : 1.0f / (60.0f * (float)std::max(B.drums[d]->P.effective_substeps(), 1u));
  int rc = MG_OK;
  for (uint32_t s = 0; s < k && rc == MG_OK; s++) rc = substep(B, dt);
  B.last_rc = rc;
  B.last_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return rc;
}

// metrics::mixing_index (metrics.rs:506-576): the Lacey index of the dye on a 16 x 16 grid over
// the drum's square, from host copies of the fluid positions (fx: 2 per particle).
void mixing_index_host(const Drum& Dr, const float* fx, float* out, int* okp) {
  const Drum* D = &Dr;  // the drum's sizes, radius and dye
  *okp = 0;
  *out = 0.0f;
  const size_t n = (size_t)D->nf;
  const float r = D->P.radius_m();
  if (n == 0 || r <= 0.0f) return;
  const int G = 16;
  const float cell = 2.0f * r / (float)G;
  std::vector<float> dsum(G * G, 0.0f);
  std::vector<uint32_t> cnt(G * G, 0);
  float tot = 0.0f;
  for (size_t i = 0; i < n; i++) {
    const int gi = std::min(std::max(as_i32(std::floor((fx[2 * i] + r) / cell)), 0), G - 1);
    const int gj = std::min(std::max(as_i32(std::floor((fx[2 * i + 1] + r) / cell)), 0), G - 1);
    const size_t idx = (size_t)gj * G + gi;
    dsum[idx] += D->dye[i];
    cnt[idx]++;
    tot += D->dye[i];
  }
  const float pm = tot / (float)n;
  std::vector<float> means;
  std::vector<uint32_t> cts;
  for (size_t q = 0; q < dsum.size(); q++)
    if (cnt[q]) {
      means.push_back(dsum[q] / (float)cnt[q]);
      cts.push_back(cnt[q]);
    }
  if (means.size() < 2) return;
  uint32_t wsu = 0;
  for (auto c : cts) wsu += c;
  const float ws = (float)wsu;
  float ov = 0.0f;
  for (size_t q = 0; q < means.size(); q++) {
    const float d = means[q] - pm;
    ov += (float)cts[q] * (d * d);
  }
  ov = ov / ws;
  const float s0 = pm * (1.0f - pm);
  float wa = 0.0f;
  for (auto c : cts) wa += (float)c * (float)c;
  wa = wa / ws;
  const float sr = s0 / std::fmax(wa, 1.0f);
  if (s0 < 1e-9f) return;
  const float den = s0 - sr;
  *okp = 1;
  if (std::fabs(den) < 1e-9f) {
    *out = 1.0f;
    return;
  }
  const float mm = (s0 - ov) / den;
  *out = mm < 0.0f ? 0.0f : (mm > 1.0f ? 1.0f : mm);
}

bool download(const Batch& B, const Drum& D, std::vector<float>& bx, std::vector<float>& bv, std::vector<float>& bom,
              std::vector<float>& fx, std::vector<float>& fv) {
  bx.resize(2 * D.nb);
  bv.resize(2 * D.nb);
  bom.resize(D.nb);
  fx.resize(2 * D.nf);
  fv.resize(2 * D.nf);
  bool ok = true;
  if (D.nb) {
    ok &= hipMemcpy(bx.data(), B.bx + 2 * (size_t)D.b0, D.nb * 8, hipMemcpyDeviceToHost) == hipSuccess;
    ok &= hipMemcpy(bv.data(), B.bv + 2 * (size_t)D.b0, D.nb * 8, hipMemcpyDeviceToHost) == hipSuccess;
    ok &= hipMemcpy(bom.data(), B.bom + D.b0, D.nb * 4, hipMemcpyDeviceToHost) == hipSuccess;
  }
  if (D.nf) {
    ok &= hipMemcpy(fx.data(), B.fx + 2 * (size_t)D.f0, D.nf * 8, hipMemcpyDeviceToHost) == hipSuccess;
    ok &= hipMemcpy(fv.data(), B.fv + 2 * (size_t)D.f0, D.nf * 8, hipMemcpyDeviceToHost) == hipSuccess;
  }
  return ok;
}

}  // namespace
}  // namespace mg

using namespace mg;

static Drum* drum_of(const mg_sim* s) { return s ? s->batch->drums[s->index].get() : nullptr; }

extern "C" {

const char* mg_version(void) { return "millgpu 0.2 (gfx950)"; }
int mg_drum_engine_enabled(void) { return 1; }

static int create_impl(const char* json, const float* xy, size_t n, mg_sim** out, char* err, size_t errlen) {
  if (!out) return MG_E_ARG;
  *out = nullptr;
  Params P;
  std::string e;
  if (json && !Params::from_json(json, P, e)) {
    set_err(err, errlen, e);
    return MG_E_PARAMS;
  }
  if (!P.validate(e)) {
    set_err(err, errlen, e);
    return MG_E_PARAMS;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    set_err(err, errlen, "no HIP device");
    return MG_E_DEVICE;
  }
  std::vector<std::unique_ptr<Drum>> v;
  v.push_back(make_drum(P, xy, n));
  Batch* B = make_batch(std::move(v), e);
  if (!B) {
    set_err(err, errlen, e);
    return MG_E_DEVICE;
  }
  *out = new mg_sim{B, 0, true};
  return MG_OK;
}

int mg_sim_create(const char* params_json, mg_sim** out, char* err, size_t errlen) {
  return create_impl(params_json, nullptr, 0, out, err, errlen);
}
int mg_sim_create_with_balls(const char* params_json, const float* xy, size_t n_balls, mg_sim** out, char* err,
                             size_t errlen) {
  if (!xy && n_balls) return MG_E_ARG;
  static const float none[2] = {0.0f, 0.0f};
  return create_impl(params_json, xy ? xy : none, n_balls, out, err, errlen);
}
void mg_sim_destroy(mg_sim* sim) {
  if (!sim) return;
  if (sim->owns_batch) {
    delete sim->batch;
    delete sim;
  }
}
int mg_sim_set_params(mg_sim* sim, const char* params_json, char* err, size_t errlen) {
  Drum* D = drum_of(sim);
  if (!D || !params_json) return MG_E_ARG;
  Params P;
  std::string e;
  if (!Params::from_json(params_json, P, e) || !P.validate(e)) {
    set_err(err, errlen, e);
    return MG_E_PARAMS;
  }
  // Hot-apply (lib.rs:252-274): the populations keep their sizes, radius, masses and h; every
  // other parameter read by the step takes effect from the next substep.
  D->P = P;
  return MG_OK;
}

int mg_sim_step(mg_sim* sim, float dt) {
  Drum* D = drum_of(sim);
  if (!D) return MG_E_ARG;
  if (sim->batch->drums.size() != 1) return MG_E_ARG;  // a batch steps through mg_batch_step
  const uint32_t ns = std::max(D->P.effective_substeps(), 1u);
  mg_sim_reset_frame_stats(sim);
  return run_substeps(*sim->batch, ns, dt / (float)ns);
}
int mg_sim_step_fixed(mg_sim* sim) { return mg_sim_step_fixed_n(sim, 1); }
int mg_sim_step_fixed_n(mg_sim* sim, uint32_t n) {
  if (!drum_of(sim) || sim->batch->drums.size() != 1) return MG_E_ARG;
  return run_substeps(*sim->batch, n, -1.0f);
}
void mg_sim_reset_frame_stats(mg_sim* sim) {  // lib.rs:449-452
  Drum* D = drum_of(sim);
  if (!D) return;
  D->max_disp = 0.0f;
  D->clamp_hits_total = 0;
}
float mg_sim_fixed_sub_dt(const mg_sim* sim) {
  Drum* D = drum_of(sim);
  return D ? 1.0f / (60.0f * (float)std::max(D->P.effective_substeps(), 1u)) : 0.0f;
}
uint32_t mg_sim_effective_substeps(const mg_sim* sim) {
  Drum* D = drum_of(sim);
  return D ? D->P.effective_substeps() : 0u;
}
size_t mg_sim_ball_count(const mg_sim* sim) {
  Drum* D = drum_of(sim);
  return D ? (size_t)D->nb : 0;
}
size_t mg_sim_fluid_count(const mg_sim* sim) {
  Drum* D = drum_of(sim);
  return D ? (size_t)D->nf : 0;
}

int mg_sim_copy_balls(mg_sim* sim, float* x2n, float* v2n, float* theta, float* omega) {
  Drum* D = drum_of(sim);
  if (!D) return MG_E_ARG;
  const Batch& B = *sim->batch;
  bool ok = true;
  if (D->nb) {
    if (x2n) ok &= hipMemcpy(x2n, B.bx + 2 * (size_t)D->b0, D->nb * 8, hipMemcpyDeviceToHost) == hipSuccess;
    if (v2n) ok &= hipMemcpy(v2n, B.bv + 2 * (size_t)D->b0, D->nb * 8, hipMemcpyDeviceToHost) == hipSuccess;
    if (theta) ok &= hipMemcpy(theta, B.bth + D->b0, D->nb * 4, hipMemcpyDeviceToHost) == hipSuccess;
    if (omega) ok &= hipMemcpy(omega, B.bom + D->b0, D->nb * 4, hipMemcpyDeviceToHost) == hipSuccess;
  }
  return ok ? MG_OK : MG_E_DEVICE;
}
int mg_sim_copy_fluid(mg_sim* sim, float* x2n, float* v2n, float* dye) {
  Drum* D = drum_of(sim);
  if (!D) return MG_E_ARG;
  const Batch& B = *sim->batch;
  bool ok = true;
  if (D->nf) {
    if (x2n) ok &= hipMemcpy(x2n, B.fx + 2 * (size_t)D->f0, D->nf * 8, hipMemcpyDeviceToHost) == hipSuccess;
    if (v2n) ok &= hipMemcpy(v2n, B.fv + 2 * (size_t)D->f0, D->nf * 8, hipMemcpyDeviceToHost) == hipSuccess;
    if (dye) std::memcpy(dye, D->dye.data(), D->nf * sizeof(float));
  }
  return ok ? MG_OK : MG_E_DEVICE;
}
int mg_sim_set_state(mg_sim* sim, const float* bx, const float* bv, const float* th, const float* om,
                     const float* fx, const float* fv, const float* guess, int has_guess) {
  Drum* D = drum_of(sim);
  if (!D) return MG_E_ARG;
  const Batch& B = *sim->batch;
  bool ok = true;
  if (D->nb) {
    if (bx) ok &= hipMemcpy(B.bx + 2 * (size_t)D->b0, bx, D->nb * 8, hipMemcpyHostToDevice) == hipSuccess;
    if (bv) ok &= hipMemcpy(B.bv + 2 * (size_t)D->b0, bv, D->nb * 8, hipMemcpyHostToDevice) == hipSuccess;
    if (th) ok &= hipMemcpy(B.bth + D->b0, th, D->nb * 4, hipMemcpyHostToDevice) == hipSuccess;
    if (om) ok &= hipMemcpy(B.bom + D->b0, om, D->nb * 4, hipMemcpyHostToDevice) == hipSuccess;
  }
  if (D->nf) {
    if (fx) ok &= hipMemcpy(B.fx + 2 * (size_t)D->f0, fx, D->nf * 8, hipMemcpyHostToDevice) == hipSuccess;
    if (fv) ok &= hipMemcpy(B.fv + 2 * (size_t)D->f0, fv, D->nf * 8, hipMemcpyHostToDevice) == hipSuccess;
    if (guess) ok &= hipMemcpy(B.fg + 2 * (size_t)D->f0, guess, D->nf * 8, hipMemcpyHostToDevice) == hipSuccess;
  }
  D->has_guess = has_guess != 0;
  return ok ? MG_OK : MG_E_DEVICE;
}
int mg_sim_copy_guess(mg_sim* sim, float* guess2n, int* has_guess) {
  Drum* D = drum_of(sim);
  if (!D) return MG_E_ARG;
  if (has_guess) *has_guess = D->has_guess ? 1 : 0;
  if (guess2n && D->nf &&
      hipMemcpy(guess2n, sim->batch->fg + 2 * (size_t)D->f0, D->nf * 8, hipMemcpyDeviceToHost) != hipSuccess)
    return MG_E_DEVICE;
  return MG_OK;
}
int mg_sim_set_clock(mg_sim* sim, float drum_angle, double sim_time) {
  Drum* D = drum_of(sim);
  if (!D) return MG_E_ARG;
  D->drum_angle = drum_angle;
  D->sim_time = sim_time;
  return MG_OK;
}
float mg_sim_drum_angle(const mg_sim* sim) {
  Drum* D = drum_of(sim);
  return D ? D->drum_angle : 0.0f;
}
double mg_sim_sim_time(const mg_sim* sim) {
  Drum* D = drum_of(sim);
  return D ? D->sim_time : 0.0;
}
uint32_t mg_sim_viscosity_iterations(mg_sim* sim) {
  Drum* D = drum_of(sim);
  return D ? D->fs.viscosity_iterations : 0u;
}
float mg_sim_mean_shear_rate_per_s(mg_sim* sim) {
  Drum* D = drum_of(sim);
  return D ? D->fs.mean_shear_rate_per_s : 0.0f;
}
float mg_sim_power_draw_w(mg_sim* sim) {
  Drum* D = drum_of(sim);
  return D ? D->power : 0.0f;
}
float mg_sim_torque_nm(mg_sim* sim) {
  Drum* D = drum_of(sim);
  return D ? D->torque : 0.0f;
}
// StressAccum::total_rate (stress.rs:176-182) after the last update.
float mg_sim_collision_rate_per_s(mg_sim* sim) {
  Drum* D = drum_of(sim);
  if (!D) return 0.0f;
  double v[49];
  if (mg_stress_state(sim->batch->st, sim->index, v) != MG_OK) return 0.0f;
  return (float)(v[0] + v[1] + v[2] + v[3]);
}
float mg_sim_dissipated_power_w(mg_sim* sim) {
  Drum* D = drum_of(sim);
  return D ? D->diss_power : 0.0f;
}
float mg_sim_max_substep_displacement_over_diameter(mg_sim* sim) {
  Drum* D = drum_of(sim);
  return D ? D->max_disp : 0.0f;
}
uint32_t mg_sim_coupling_clamp_hits(mg_sim* sim) {
  Drum* D = drum_of(sim);
  return D ? D->clamp_hits_total : 0u;
}

int mg_sim_stats(mg_sim* sim, double* o, size_t cap) {
  Drum* D = drum_of(sim);
  if (!D || !o) return MG_E_ARG;
  const mg_fluid_stats& f = D->fs;
  const mg_dem_stats& s = D->ds;
  const double v[] = {(double)D->drum_angle, D->sim_time, D->power, D->torque,
                      (double)mg_sim_collision_rate_per_s(sim), D->diss_power, D->max_disp,
                      (double)D->clamp_hits_total, (double)f.viscosity_iterations, (double)f.density_iterations,
                      f.mean_shear_rate_per_s, f.wall_work_j, s.wall_work_j, s.dissipated_energy_j, s.coupling_work_j,
                      (double)s.collision_count, (double)f.n_groups, (double)f.n_links, (double)f.links_in_range,
                      (double)s.n_pairs, f.ke_delta_noslip_j, f.ke_delta_drag_j, f.ke_delta_buoyancy_j,
                      f.ke_delta_viscosity_j, f.ke_delta_clamp_j, f.wall_impulse_x, f.wall_impulse_y,
                      f.wall_angular_impulse, (double)f.error_bits, (double)s.n_impacts,
                      // DemStepStats::impact_energy_histogram (dem.rs:99-109)
                      (double)s.impact_energy_histogram[0], (double)s.impact_energy_histogram[1],
                      (double)s.impact_energy_histogram[2], (double)s.impact_energy_histogram[3],
                      (double)s.impact_energy_histogram[4], (double)s.impact_energy_histogram[5],
                      (double)s.impact_energy_histogram[6], (double)s.impact_energy_histogram[7],
                      (double)s.impact_energy_histogram[8], (double)s.impact_energy_histogram[9],
                      (double)s.impact_energy_histogram[10], (double)s.impact_energy_histogram[11],
                      (double)s.projection_restarts};
  const size_t n = sizeof(v) / sizeof(v[0]);
  std::memcpy(o, v, std::min(n, cap) * sizeof(double));
  return (int)n;
}
int mg_sim_stress_state(mg_sim* sim, double* o) {
  if (!drum_of(sim) || !o) return MG_E_ARG;
  return mg_stress_state(sim->batch->st, sim->index, o);
}
int mg_sim_impacts(mg_sim* sim, double* o, size_t cap, size_t* n_out) {
  Drum* D = drum_of(sim);
  if (!D) return MG_E_ARG;
  uint32_t n = 0;
  std::vector<mg_impact_record> rec(std::max<size_t>(cap, 1));
  const int rc = mg_dem_impacts(sim->batch->dem, sim->index, rec.data(), (uint32_t)rec.size(), &n);
  if (rc != MG_OK) return rc;
  if (n_out) *n_out = n;
  for (size_t i = 0; i < n && i < cap && o; i++) {
    o[6 * i] = rec[i].kind;
    o[6 * i + 1] = rec[i].ball;
    o[6 * i + 2] = rec[i].partner == 0xFFFFFFFFu ? -1.0 : (double)rec[i].partner;
    o[6 * i + 3] = rec[i].normal_speed;
    o[6 * i + 4] = rec[i].slip_speed;
    o[6 * i + 5] = rec[i].restitution;
  }
  return MG_OK;
}

// metrics::mixing_index (metrics.rs:506-576) over the current fluid positions and dye.
int mg_sim_mixing_index(mg_sim* sim, float* out, int* okp) {
  Drum* D = drum_of(sim);
  if (!D || !out || !okp) return MG_E_ARG;
  *okp = 0;
  *out = 0.0f;
  const size_t n = (size_t)D->nf;
  if (n == 0 || D->P.radius_m() <= 0.0f) return MG_OK;
  std::vector<float> fx(2 * n);
  if (mg_sim_copy_fluid(sim, fx.data(), nullptr, nullptr) != MG_OK) return MG_E_DEVICE;
  mixing_index_host(*D, fx.data(), out, okp);
  return MG_OK;
}

int mg_sim_energy_budget(mg_sim* sim, mg_energy_budget* o) {  // lib.rs:211-229
  Drum* D = drum_of(sim);
  if (!D || !o) return MG_E_ARG;
  std::vector<float> bx, bv, bom, fx, fv;
  if (!download(*sim->batch, *D, bx, bv, bom, fx, fv)) return MG_E_DEVICE;
  o->elapsed_s = D->eb[0];
  o->shaft_work_j = D->eb[1];
  o->dem_wall_work_j = D->eb[2];
  o->fluid_wall_work_j = D->eb[3];
  o->ball_contact_dissipation_j = D->eb[4];
  o->fluid_wall_slip_j = D->eb[5];
  o->fluid_viscous_j = D->eb[6];
  o->fluid_clamp_removed_j = D->eb[7];
  o->interface_created_j = D->eb[8];
  o->fluid_wall_angular_impulse = D->eb[9];
  o->delta_mechanical_j = mech_energy(*D, bx, bv, bom, fx, fv) - D->e0;
  const double expl = o->shaft_work_j - o->ball_contact_dissipation_j - o->fluid_wall_slip_j - o->fluid_viscous_j -
                      o->fluid_clamp_removed_j + o->interface_created_j;
  o->unattributed_j = o->delta_mechanical_j - expl;
  return MG_OK;
}
int mg_sim_reset_energy_budget(mg_sim* sim) {  // lib.rs:205-208
  Drum* D = drum_of(sim);
  if (!D) return MG_E_ARG;
  std::vector<float> bx, bv, bom, fx, fv;
  if (!download(*sim->batch, *D, bx, bv, bom, fx, fv)) return MG_E_DEVICE;
  std::fill(D->eb, D->eb + 10, 0.0);
  D->e0 = mech_energy(*D, bx, bv, bom, fx, fv);
  return MG_OK;
}
int mg_sim_set_stress_debounce_factor(mg_sim* sim, double factor) {
  if (!drum_of(sim)) return MG_E_ARG;
  return mg_stress_set_debounce_factor(sim->batch->st, sim->index, factor);
}

int mg_sim_metrics_json(mg_sim* sim, char* buf, size_t cap) {
  Drum* D = drum_of(sim);
  if (!D || !buf) return MG_E_ARG;
  float mix = 0.0f;
  int ok = 0;
  mg_sim_mixing_index(sim, &mix, &ok);
  double st[49];
  if (mg_stress_state(sim->batch->st, sim->index, st) != MG_OK) return MG_E_DEVICE;
  float td, d, dens, sf;
  uint32_t cnt;
  D->P.effective_media(td, d, dens, cnt, sf);
  const double rr = (double)D->br, m3 = (double)D->bmass * 4.0 / 3.0 * rr, scale = m3 * 9.81 * (2.0 * rr);
  const double n = std::max<uint32_t>(1, D->nb);
  std::string hist, edges;
  for (int k = 0; k < 28; k++) {
    char t[40];
    std::snprintf(t, sizeof t, "%s%.9g", k ? "," : "", (double)(float)st[8 + k]);
    hist += t;
  }
  for (int k = 0; k <= 28; k++) {
    char t[40];
    std::snprintf(t, sizeof t, "%s%.9g", k ? "," : "", (double)(float)(std::pow(10.0, -3.0 + k / 4.0) * scale));
    edges += t;
  }
  const double total = st[0] + st[1] + st[2] + st[3];
  char mixs[32];
  if (ok) std::snprintf(mixs, sizeof mixs, "%.9g", mix);
  else std::snprintf(mixs, sizeof mixs, "null");
  char b[8192];
  const int len = std::snprintf(
      b, sizeof b,
      "{\"power_draw_w\":%.9g,\"torque_nm\":%.9g,\"collision_rate_per_s\":%.9g,\"dissipated_power_w\":%.9g,"
      "\"mixing_index\":%s,\"mean_shear_rate_per_s\":%.9g,\"viscosity_solver_iterations\":%u,"
      "\"coupling_clamp_hits\":%u,\"max_substep_displacement_over_diameter\":%.9g,"
      "\"effective_ball_diameter_m\":%.9g,\"simulated_ball_count\":%u,\"true_ball_count\":%u,"
      "\"coarse_graining_factor\":%.9g,\"impact_energy_histogram\":{\"bin_edges_j\":[%s],\"counts_per_s\":[%s]},"
      "\"stress\":{\"impacts_total_per_ball_per_s\":%.9g,\"ball_mass_3d_kg\":%.9g,\"scale_energy_j\":%.9g}}",
      D->power, D->torque, (double)(float)total, D->diss_power, mixs, D->fs.mean_shear_rate_per_s,
      D->fs.viscosity_iterations, D->clamp_hits_total, D->max_disp, d, cnt,
      (unsigned)std::fmax(std::round(D->P.true_ball_count()), 0.0f), sf, edges.c_str(), hist.c_str(),
      (double)(float)(total / n), (double)(float)m3, (double)(float)scale);
  if (len < 0 || (size_t)len + 1 > cap) return -(len + 1);
  std::memcpy(buf, b, (size_t)len + 1);
  return len;
}

// Snapshot blob: header, the drum's host scalars, state arrays, warm start, stress accumulator.
namespace {
struct SnapHead {
  uint32_t magic, nb, nf, stress_bytes;
  float drum_angle, power, torque, diss_power, max_disp;
  uint32_t clamp_hits_total, has_guess, pad;
  double sim_time, e0, eb[10];
};
}  // namespace
int mg_sim_save_state(mg_sim* sim, void* buf, size_t cap, size_t* needed) {
  Drum* D = drum_of(sim);
  if (!D || !needed) return MG_E_ARG;
  const Batch& B = *sim->batch;
  const size_t sb = mg_stress_blob_bytes(B.st);
  const size_t sz = sizeof(SnapHead) + (size_t)D->nb * 24 + (size_t)D->nf * 24 + sb;
  *needed = sz;
  if (!buf || cap < sz) return MG_E_ARG;
  SnapHead h{};
  h.magic = 0x4D475332u;
  h.nb = D->nb;
  h.nf = D->nf;
  h.stress_bytes = (uint32_t)sb;
  h.drum_angle = D->drum_angle;
  h.power = D->power;
  h.torque = D->torque;
  h.diss_power = D->diss_power;
  h.max_disp = D->max_disp;
  h.clamp_hits_total = D->clamp_hits_total;
  h.has_guess = D->has_guess ? 1u : 0u;
  h.sim_time = D->sim_time;
  h.e0 = D->e0;
  std::memcpy(h.eb, D->eb, sizeof h.eb);
  char* p = static_cast<char*>(buf);
  std::memcpy(p, &h, sizeof h);
  p += sizeof h;
  bool ok = true;
  auto d2h = [&](const void* src, size_t bytes) {
    if (bytes) ok &= hipMemcpy(p, src, bytes, hipMemcpyDeviceToHost) == hipSuccess;
    p += bytes;
  };
  d2h(B.bx + 2 * (size_t)D->b0, D->nb * 8);
  d2h(B.bv + 2 * (size_t)D->b0, D->nb * 8);
  d2h(B.bth + D->b0, D->nb * 4);
  d2h(B.bom + D->b0, D->nb * 4);
  d2h(B.fx + 2 * (size_t)D->f0, D->nf * 8);
  d2h(B.fv + 2 * (size_t)D->f0, D->nf * 8);
  d2h(B.fg + 2 * (size_t)D->f0, D->nf * 8);
  if (ok && mg_stress_save(B.st, sim->index, p) != MG_OK) ok = false;
  return ok ? MG_OK : MG_E_DEVICE;
}
int mg_sim_load_state(mg_sim* sim, const void* buf, size_t len) {
  Drum* D = drum_of(sim);
  if (!D || !buf || len < sizeof(SnapHead)) return MG_E_ARG;
  Batch& B = *sim->batch;
  SnapHead h;
  std::memcpy(&h, buf, sizeof h);
  const size_t sb = mg_stress_blob_bytes(B.st);
  const size_t sz = sizeof(SnapHead) + (size_t)D->nb * 24 + (size_t)D->nf * 24 + sb;
  if (h.magic != 0x4D475332u || h.nb != D->nb || h.nf != D->nf || h.stress_bytes != sb || len < sz) return MG_E_ARG;
  const char* p = static_cast<const char*>(buf) + sizeof h;
  bool ok = true;
  auto h2d = [&](void* dst, size_t bytes) {
    if (bytes) ok &= hipMemcpy(dst, p, bytes, hipMemcpyHostToDevice) == hipSuccess;
    p += bytes;
  };
  h2d(B.bx + 2 * (size_t)D->b0, D->nb * 8);
  h2d(B.bv + 2 * (size_t)D->b0, D->nb * 8);
  h2d(B.bth + D->b0, D->nb * 4);
  h2d(B.bom + D->b0, D->nb * 4);
  h2d(B.fx + 2 * (size_t)D->f0, D->nf * 8);
  h2d(B.fv + 2 * (size_t)D->f0, D->nf * 8);
  h2d(B.fg + 2 * (size_t)D->f0, D->nf * 8);
  if (ok && mg_stress_load(B.st, sim->index, p) != MG_OK) ok = false;
  if (!ok) return MG_E_DEVICE;
  D->drum_angle = h.drum_angle;
  D->power = h.power;
  D->torque = h.torque;
  D->diss_power = h.diss_power;
  D->max_disp = h.max_disp;
  D->clamp_hits_total = h.clamp_hits_total;
  D->has_guess = h.has_guess != 0;
  D->sim_time = h.sim_time;
  D->e0 = h.e0;
  std::memcpy(D->eb, h.eb, sizeof h.eb);
  return MG_OK;
}

int mg_sim_force_iterations(mg_sim* sim, uint32_t density_passes, uint32_t cg_iterations) {
  Drum* D = drum_of(sim);
  if (!D || density_passes > 32) return MG_E_ARG;
  D->forced = true;
  D->force_density = density_passes;
  D->force_cg = cg_iterations;
  return MG_OK;
}
uint32_t mg_sim_error_bits(mg_sim* sim) {
  Drum* D = drum_of(sim);
  return D ? D->fs.error_bits : 0u;
}

int mg_batch_create(const char* const* params_json, uint32_t n, mg_batch** out, char* err, size_t errlen) {
  if (!out || n == 0) return MG_E_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    set_err(err, errlen, "no HIP device");
    return MG_E_DEVICE;
  }
  std::vector<std::unique_ptr<Drum>> drums;
  for (uint32_t i = 0; i < n; i++) {
    Params P;
    std::string e;
    const char* j = params_json ? params_json[i] : nullptr;
    if (j && !Params::from_json(j, P, e)) {
      set_err(err, errlen, "drum " + std::to_string(i) + ": " + e);
      return MG_E_PARAMS;
    }
    if (!P.validate(e)) {
      set_err(err, errlen, "drum " + std::to_string(i) + ": " + e);
      return MG_E_PARAMS;
    }
    drums.push_back(make_drum(P, nullptr, 0));
  }
  std::string e;
  Batch* B = make_batch(std::move(drums), e);
  if (!B) {
    set_err(err, errlen, e);
    return MG_E_DEVICE;
  }
  *out = new mg_batch{B};
  return MG_OK;
}
void mg_batch_destroy(mg_batch* b) {
  if (!b) return;
  delete b->b;
  delete b;
}
uint32_t mg_batch_size(const mg_batch* b) { return b ? (uint32_t)b->b->drums.size() : 0u; }
int mg_batch_step(mg_batch* b, uint32_t substeps, int use_graph) {
  (void)use_graph;
  if (!b) return MG_E_ARG;
  return run_substeps(*b->b, substeps, -1.0f);
}
mg_sim* mg_batch_drum(mg_batch* b, uint32_t i) {
  if (!b || i >= b->b->drums.size()) return nullptr;
  return b->b->views[i];
}
int mg_batch_sync(mg_batch* b, char* err, size_t errlen) {
  if (!b) return MG_E_ARG;
  if (hipDeviceSynchronize() != hipSuccess) {
    set_err(err, errlen, "HIP device error");
    return MG_E_DEVICE;
  }
  if (b->b->last_rc != MG_OK) set_err(err, errlen, b->b->msg);
  return b->b->last_rc;
}
float mg_batch_last_step_ms(const mg_batch* b) { return b ? b->b->last_ms : 0.0f; }
int mg_batch_set_timing(mg_batch* b, int on) {
  if (!b) return MG_E_ARG;
  const int r1 = mg_fluid_set_timing(b->b->fl, on), r2 = mg_dem_set_timing(b->b->dem, on);
  return r1 != MG_OK ? r1 : r2;
}
int mg_batch_kernel_times(mg_batch* b, char* json, size_t cap) {
  if (!b || !json || cap == 0) return MG_E_ARG;
  std::vector<char> f(1 << 16), d(1 << 16);
  if (mg_fluid_kernel_times(b->b->fl, f.data(), f.size()) < 0 || mg_dem_kernel_times(b->b->dem, d.data(), d.size()) < 0)
    return MG_E_ARG;
  std::string fs(f.data()), ds(d.data());
  std::string s = fs.substr(0, fs.size() - 1);
  if (ds.size() > 2) s += (s.size() > 1 ? "," : "") + ds.substr(1);
  else s += "}";
  if (s.size() + 1 > cap) return -(int)(s.size() + 1);
  std::memcpy(json, s.c_str(), s.size() + 1);
  return (int)s.size();
}

}  // extern "C"

// ---- served sweeps (continuous batching) --------------------------------------------------------
namespace mg {
namespace {

struct RevRec {
  uint32_t rev;
  double sim_time;
  float power, torque, coll, diss, mix;
  int mix_ok;
  uint64_t hist[12];
};

struct Request {
  uint64_t id = 0;
  int state = 0;                // 0 queued, 1 running, 2 done
  std::unique_ptr<Drum> owned;  // until admitted (then the batch owns the drum)
  uint64_t total = 0, steps = 0, footprint = 0;
  double arrival = 0.0, start = 0.0, done = 0.0, period = 0.0, next_rev = 0.0;
  float dt = 0.0f;
  uint32_t slot = 0, revs = 0, nb = 0, nf = 0;
  uint64_t hist[12] = {};
  float power = 0.0f, torque = 0.0f, coll = 0.0f, diss = 0.0f;
  std::vector<RevRec> stream;
};

// The device bytes a running drum takes: its state and its share of the fluid operator's
// workspace, which mg_fluid_step lays out per step from the drums' sizes (per particle: the
// neighbour lists at their 48-entry capacity with weights and directions, the push records, the
// gather lists, the CG rows; per ball: the adhesion, boundary-group and form-drag item lists, the
// squeeze links at their capacity, the CG rows; per drum: the cells of its five grids). An upper
// estimate of the layout, not a measurement.
uint64_t drum_footprint(const Drum& D) {
  const double nf = D.nf, nb = D.nb, br = D.br, h = D.h, R = D.P.radius_m();
  const double lr = br > 0.0 ? (2.0 * br + 4.0 * h) / br : 0.0;
  const double lcap = std::ceil(0.95 * lr * lr) + 4.0;
  const double cell = std::max(std::min(h, br > 0.0 ? 2.0 * br : h), 1e-6);
  const double side = std::floor(3.0 * R / cell) + 3.0;
  return (uint64_t)(nf * 1100.0 + nb * (19500.0 + 20.0 * lcap) + 5.0 * 8.0 * side * side + 65536.0);
}

struct PoolView {
  float *fx, *fv, *fg, *bx, *bv, *bth, *bom, *imp, *ang;
};
uint64_t up256(uint64_t b) { return ((b + 255) / 256) * 256; }
uint64_t pool_bytes(uint64_t F, uint64_t Bn) {
  return 3 * up256(F * 8) + 3 * up256(Bn * 8) + 3 * up256(Bn * 4);
}
PoolView pool_view(char* p, uint64_t F, uint64_t Bn) {
  PoolView v;
  auto take = [&](uint64_t b) {
    char* q = p;
    p += up256(b);
    return reinterpret_cast<float*>(q);
  };
  v.fx = take(F * 8);
  v.fv = take(F * 8);
  v.fg = take(F * 8);
  v.bx = take(Bn * 8);
  v.bv = take(Bn * 8);
  v.imp = take(Bn * 8);
  v.bth = take(Bn * 4);
  v.bom = take(Bn * 4);
  v.ang = take(Bn * 4);
  return v;
}

}  // namespace

// A server's batch holds the running drums packed in admission order; two state pools alternate:
// a change of the running set copies the drums that keep running (runs of consecutive drums, one
// device copy per run and array) and uploads the admitted drums' initial states into the other
// pool.
struct Server {
  mg_server_config cfg{};
  std::unique_ptr<Batch> B;
  std::vector<Request*> running;  // in the batch's drum order
  std::deque<Request*> queue;
  std::map<uint64_t, std::unique_ptr<Request>> reqs;
  std::vector<uint32_t> free_slots;
  char* pool[2] = {nullptr, nullptr};
  uint64_t fcap[2] = {0, 0}, bcap[2] = {0, 0};
  int cur = 0;
  uint64_t fixed = 0, used = 0, next_id = 1, ticks = 0, drum_substeps = 0, n_done = 0;
  bool retired = false;
  std::chrono::steady_clock::time_point t0;
  std::string msg;
  double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
  ~Server() {
    B.reset();
    for (char* p : pool)
      if (p) (void)hipFree(p);
  }
};

namespace {

int server_fail(Server& S, const std::string& m, int rc) {
  S.msg = m;
  return rc;
}

// Lays out the drums that keep running followed by the admitted ones in the other pool.
int rebuild(Server& S, std::vector<Request*>& incoming) {
  Batch& B = *S.B;
  std::vector<Request*> keep;
  std::vector<std::unique_ptr<Drum>> drums;
  std::vector<uint32_t> of, ob;  // the kept drums' old offsets
  for (size_t i = 0; i < S.running.size(); i++) {
    if (S.running[i]->state != 1) continue;  // finished: its drum is dropped here
    keep.push_back(S.running[i]);
    of.push_back(B.drums[i]->f0);
    ob.push_back(B.drums[i]->b0);
    drums.push_back(std::move(B.drums[i]));
  }
  uint64_t F = 0, Bn = 0;
  for (auto& d : drums) {
    F += d->nf;
    Bn += d->nb;
  }
  for (auto* r : incoming) {
    F += r->owned->nf;
    Bn += r->owned->nb;
  }
  const int dst = 1 - S.cur;
  if (!S.pool[dst] || F > S.fcap[dst] || Bn > S.bcap[dst]) {
    const uint64_t nf = std::max<uint64_t>({F, 2 * S.fcap[dst], 1u << 16});
    const uint64_t nbc = std::max<uint64_t>({Bn, 2 * S.bcap[dst], 1u << 12});
    if (S.pool[dst]) (void)hipFree(S.pool[dst]);
    S.pool[dst] = nullptr;
    S.fcap[dst] = S.bcap[dst] = 0;
    if (hipMalloc(&S.pool[dst], pool_bytes(nf, nbc)) != hipSuccess)
      return server_fail(S, "hipMalloc failed for the server's state pool", MG_E_DEVICE);
    S.fcap[dst] = nf;
    S.bcap[dst] = nbc;
  }
  const PoolView dv = pool_view(S.pool[dst], S.fcap[dst], S.bcap[dst]);
  bool ok = true;
  uint64_t f0 = 0, b0 = 0;
  if (!drums.empty()) {
    const PoolView sv = pool_view(S.pool[S.cur], S.fcap[S.cur], S.bcap[S.cur]);
    const hipMemcpyKind dd = hipMemcpyDeviceToDevice;
    for (size_t k = 0; k < drums.size();) {
      size_t m = k + 1;  // drums k..m-1 are consecutive in the old layout
      while (m < drums.size() && of[m] == of[m - 1] + drums[m - 1]->nf && ob[m] == ob[m - 1] + drums[m - 1]->nb) m++;
      const uint64_t fn = of[m - 1] + drums[m - 1]->nf - of[k], bn = ob[m - 1] + drums[m - 1]->nb - ob[k];
      if (fn) {
        ok &= hipMemcpyAsync(dv.fx + 2 * f0, sv.fx + 2 * (uint64_t)of[k], fn * 8, dd, nullptr) == hipSuccess;
        ok &= hipMemcpyAsync(dv.fv + 2 * f0, sv.fv + 2 * (uint64_t)of[k], fn * 8, dd, nullptr) == hipSuccess;
        ok &= hipMemcpyAsync(dv.fg + 2 * f0, sv.fg + 2 * (uint64_t)of[k], fn * 8, dd, nullptr) == hipSuccess;
      }
      if (bn) {
        ok &= hipMemcpyAsync(dv.bx + 2 * b0, sv.bx + 2 * (uint64_t)ob[k], bn * 8, dd, nullptr) == hipSuccess;
        ok &= hipMemcpyAsync(dv.bv + 2 * b0, sv.bv + 2 * (uint64_t)ob[k], bn * 8, dd, nullptr) == hipSuccess;
        ok &= hipMemcpyAsync(dv.bth + b0, sv.bth + ob[k], bn * 4, dd, nullptr) == hipSuccess;
        ok &= hipMemcpyAsync(dv.bom + b0, sv.bom + ob[k], bn * 4, dd, nullptr) == hipSuccess;
      }
      for (size_t j = k; j < m; j++) {
        drums[j]->f0 = (uint32_t)(f0 + (of[j] - of[k]));
        drums[j]->b0 = (uint32_t)(b0 + (ob[j] - ob[k]));
      }
      f0 += fn;
      b0 += bn;
      k = m;
    }
  }
  const hipMemcpyKind hd = hipMemcpyHostToDevice;
  for (auto* r : incoming) {
    Drum& D = *r->owned;
    D.f0 = (uint32_t)f0;
    D.b0 = (uint32_t)b0;
    if (D.nf) {
      ok &= hipMemcpy(dv.fx + 2 * f0, D.init_fx.data(), D.nf * 8, hd) == hipSuccess;
      ok &= hipMemcpy(dv.fv + 2 * f0, D.init_fv.data(), D.nf * 8, hd) == hipSuccess;
      ok &= hipMemset(dv.fg + 2 * f0, 0, D.nf * 8) == hipSuccess;
    }
    if (D.nb) {
      ok &= hipMemcpy(dv.bx + 2 * b0, D.init_bx.data(), D.nb * 8, hd) == hipSuccess;
      ok &= hipMemcpy(dv.bv + 2 * b0, D.init_bv.data(), D.nb * 8, hd) == hipSuccess;
      ok &= hipMemcpy(dv.bom + b0, D.init_bom.data(), D.nb * 4, hd) == hipSuccess;
      ok &= hipMemset(dv.bth + b0, 0, D.nb * 4) == hipSuccess;
    }
    std::vector<float>().swap(D.init_bx);
    std::vector<float>().swap(D.init_bv);
    std::vector<float>().swap(D.init_bom);
    std::vector<float>().swap(D.init_fx);
    std::vector<float>().swap(D.init_fv);
    f0 += D.nf;
    b0 += D.nb;
    drums.push_back(std::move(r->owned));
    keep.push_back(r);
  }
  ok &= hipDeviceSynchronize() == hipSuccess;
  if (!ok) return server_fail(S, "HIP error laying out the running drums", MG_E_DEVICE);
  S.cur = dst;
  B.fx = dv.fx;
  B.fv = dv.fv;
  B.fg = dv.fg;
  B.bx = dv.bx;
  B.bv = dv.bv;
  B.bth = dv.bth;
  B.bom = dv.bom;
  B.imp = dv.imp;
  B.ang = dv.ang;
  B.TF = F;
  B.TB = Bn;
  B.drums = std::move(drums);
  S.running = std::move(keep);
  B.sslot.resize(S.running.size());
  for (size_t d = 0; d < S.running.size(); d++) {
    B.sslot[d] = S.running[d]->slot;
    B.hbeg[S.running[d]->slot] = (uint32_t)d * B.rstride;
  }
  B.ranges_dirty = true;
  return MG_OK;
}

// One tick: admissions, one substep of every running drum, per-revolution records, finished runs.
int tick(Server& S) {
  Batch& B = *S.B;
  std::vector<Request*> incoming;
  while (!S.queue.empty() && !S.free_slots.empty()) {
    Request* r = S.queue.front();
    if (S.fixed + S.used + r->footprint > S.cfg.hbm_budget_bytes) break;  // the head waits
    S.queue.pop_front();
    r->slot = S.free_slots.back();
    S.free_slots.pop_back();
    if (mg_stress_reset(B.st, r->slot) != MG_OK) return server_fail(S, "HIP error resetting a stress slot", MG_E_DEVICE);
    r->state = 1;
    r->start = S.now();
    S.used += r->footprint;
    incoming.push_back(r);
  }
  if (!incoming.empty() || S.retired) {
    const int rc = rebuild(S, incoming);
    if (rc != MG_OK) return rc;
    S.retired = false;
  }
  const uint32_t nd = (uint32_t)S.running.size();
  if (nd == 0) return MG_OK;
  std::vector<float> dt(nd);
  for (uint32_t d = 0; d < nd; d++) dt[d] = S.running[d]->dt;
  const int rc = substep(B, dt);
  if (rc != MG_OK) return server_fail(S, B.msg, rc);
  S.ticks++;
  S.drum_substeps += nd;
  const double t = S.now();
  for (uint32_t d = 0; d < nd; d++) {
    Request* r = S.running[d];
    const Drum& D = *B.drums[d];
    r->steps++;
    for (int k = 0; k < 12; k++) r->hist[k] += D.ds.impact_energy_histogram[k];
    r->power = D.power;
    r->torque = D.torque;
    r->diss = D.diss_power;
    const bool fin = r->steps >= r->total, rev = D.sim_time >= r->next_rev;
    if (fin || rev) {  // StressAccum::total_rate of the drum's accumulator (stress.rs:176-182)
      double st[49];
      if (mg_stress_state(B.st, r->slot, st) != MG_OK) return server_fail(S, "HIP copy error", MG_E_DEVICE);
      r->coll = (float)(st[0] + st[1] + st[2] + st[3]);
    }
    if (rev) {
      RevRec q{};
      q.rev = ++r->revs;
      q.sim_time = D.sim_time;
      q.power = r->power;
      q.torque = r->torque;
      q.coll = r->coll;
      q.diss = r->diss;
      std::vector<float> fx(2 * (size_t)D.nf);
      if (D.nf && hipMemcpy(fx.data(), B.fx + 2 * (size_t)D.f0, D.nf * 8, hipMemcpyDeviceToHost) != hipSuccess)
        return server_fail(S, "HIP copy error", MG_E_DEVICE);
      mixing_index_host(D, fx.data(), &q.mix, &q.mix_ok);
      std::memcpy(q.hist, r->hist, sizeof q.hist);
      std::memset(r->hist, 0, sizeof r->hist);
      r->stream.push_back(q);
      r->next_rev += r->period;
    }
    if (fin) {
      r->state = 2;
      r->done = t;
      S.used -= r->footprint;
      S.n_done++;
      S.retired = true;
      if (mg_stress_reset(B.st, r->slot) != MG_OK) return server_fail(S, "HIP error resetting a stress slot", MG_E_DEVICE);
      S.free_slots.push_back(r->slot);
    }
  }
  return MG_OK;
}

}  // namespace
}  // namespace mg

struct mg_server {
  mg::Server s;
};

extern "C" {

int mg_server_create(const mg_server_config* cfg, mg_server** out, char* err, size_t errlen) {
  if (!cfg || !out) return MG_E_ARG;
  *out = nullptr;
  if (cfg->max_drums < 1 || cfg->max_drums > 4096 || cfg->max_balls < 1 || cfg->max_balls > 2048) {
    set_err(err, errlen, "max_drums must be in 1..4096 and max_balls in 1..2048");
    return MG_E_ARG;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    set_err(err, errlen, "no HIP device");
    return MG_E_DEVICE;
  }
  auto h = std::make_unique<mg_server>();
  Server& S = h->s;
  S.cfg = *cfg;
  if (S.cfg.hbm_budget_bytes == 0) S.cfg.hbm_budget_bytes = 16ull << 30;
  const uint32_t cap = cfg->max_drums, mb = cfg->max_balls;
  // the slots' fixed workspaces: the DEM's per-population arrays and the stress accumulators
  uint64_t tbl = 64;
  while (tbl < 4ull * mb + 1024) tbl <<= 1;
  S.fixed = (uint64_t)cap * ((uint64_t)mb * (24 * 32 + 96) + (4ull * mb + 64) * 32 + 49 * 8 + 16 * tbl);
  if (S.fixed > S.cfg.hbm_budget_bytes) {
    set_err(err, errlen, "the budget does not hold the slots' fixed workspaces (" + std::to_string(S.fixed >> 20) +
                             " MiB)");
    return MG_E_CAPACITY;
  }
  S.B = std::make_unique<Batch>();
  Batch& B = *S.B;
  char e[512] = {0};
  if (mg_fluid_create(cap, &B.fl, e, sizeof e) != MG_OK || mg_dem_create(cap, mb, 24, 4 * mb + 64, &B.dem, e, sizeof e) != MG_OK ||
      mg_stress_create(cap, 4 * mb + 1024, &B.st, e, sizeof e) != MG_OK) {
    set_err(err, errlen, e);
    return MG_E_DEVICE;
  }
  const size_t rb = up256((uint64_t)cap * 4);
  if (hipMalloc(&B.dev, 2 * rb) != hipSuccess || hipMemset(B.dev, 0, 2 * rb) != hipSuccess) {
    set_err(err, errlen, "hipMalloc failed for the server");
    return MG_E_DEVICE;
  }
  B.rbeg = reinterpret_cast<uint32_t*>(B.dev);
  B.rend = reinterpret_cast<uint32_t*>(B.dev + rb);
  mg_dem_records_device(B.dem, &B.rstride);
  B.fd.resize(cap);
  B.dd.resize(cap);
  B.fs.resize(cap);
  B.ds.resize(cap);
  B.ss.resize(cap);
  B.hbeg.assign(cap, 0);
  B.hend.assign(cap, 0);
  mg_fluid_drum& f = B.fd_empty;  // a slot without a drum: no particles, no balls
  f.radius_m = 0.1f;
  f.h = 1e-3f;
  f.rest_density = 1000.0f;
  f.particle_mass = 1e-3f;
  f.ball_radius = 1e-3f;
  f.ball_mass = 1e-3f;
  f.ball_inertia = 1e-9f;
  f.ball_true_radius = 1e-3f;
  f.iterations = 3;
  f.dt = 1e-3f;
  mg_dem_drum& m = B.dd_empty;
  m.radius_m = 0.1f;
  m.ball_radius = 1e-3f;
  m.ball_mass = 1e-3f;
  m.ball_inertia = 1e-9f;
  m.iterations = 2;
  m.dt = 1e-3f;
  m.has_external = 1;
  m.energy_f64 = 1;
  for (uint32_t q = cap; q-- > 0;) S.free_slots.push_back(q);
  S.t0 = std::chrono::steady_clock::now();
  *out = h.release();
  return MG_OK;
}

void mg_server_destroy(mg_server* h) { delete h; }

double mg_server_now(const mg_server* h) { return h ? h->s.now() : 0.0; }

int mg_server_submit(mg_server* h, const char* params_json, double revolutions, double arrival_s, uint64_t* id,
                     char* err, size_t errlen) {
  if (!h || !id) return MG_E_ARG;
  Server& S = h->s;
  Params P;
  std::string e;
  if (params_json && !Params::from_json(params_json, P, e)) {
    set_err(err, errlen, e);
    return MG_E_PARAMS;
  }
  if (!P.validate(e)) {
    set_err(err, errlen, e);
    return MG_E_PARAMS;
  }
  const double w = std::fabs((double)P.omega());
  if (!(w > 0.0) || !std::isfinite(w)) {
    set_err(err, errlen, "a drum that does not turn has no revolutions");
    return MG_E_PARAMS;
  }
  if (!(revolutions > 0.0) || !std::isfinite(revolutions)) {
    set_err(err, errlen, "revolutions must be positive and finite");
    return MG_E_PARAMS;
  }
  auto r = std::make_unique<Request>();
  r->owned = make_drum(P, nullptr, 0);
  const Drum& D = *r->owned;
  if (D.nb > S.cfg.max_balls) {
    set_err(err, errlen, std::to_string(D.nb) + " balls: more than the server's max_balls");
    return MG_E_CAPACITY;
  }
  r->footprint = drum_footprint(D);
  if (S.fixed + r->footprint > S.cfg.hbm_budget_bytes) {
    set_err(err, errlen, "the drum's footprint alone exceeds the budget");
    return MG_E_CAPACITY;
  }
  r->dt = 1.0f / (60.0f * (float)std::max(P.effective_substeps(), 1u));
  r->period = 6.283185307179586 / w;
  r->total = std::max<uint64_t>(1, (uint64_t)std::ceil(revolutions * r->period / (double)r->dt));
  r->next_rev = r->period;
  r->nb = D.nb;
  r->nf = D.nf;
  r->id = S.next_id++;
  r->arrival = arrival_s < 0.0 ? S.now() : arrival_s;
  *id = r->id;
  S.queue.push_back(r.get());
  S.reqs[r->id] = std::move(r);
  return MG_OK;
}

int mg_server_poll(mg_server* h, uint32_t ticks, mg_server_status* st, char* err, size_t errlen) {
  if (!h) return MG_E_ARG;
  Server& S = h->s;
  for (uint32_t t = 0; t < ticks; t++) {
    if (S.queue.empty() && S.running.empty()) break;
    const int rc = tick(S);
    if (rc != MG_OK) {
      set_err(err, errlen, S.msg);
      return rc;
    }
  }
  if (st) {
    st->queued = (uint32_t)S.queue.size();
    uint32_t run = 0;
    for (auto* r : S.running) run += r->state == 1 ? 1u : 0u;
    st->running = run;
    st->done = (uint32_t)S.n_done;
    st->ticks = S.ticks;
    st->drum_substeps = S.drum_substeps;
    st->used_bytes = S.fixed + S.used;
    st->now_s = S.now();
  }
  return MG_OK;
}

int mg_server_result(mg_server* h, uint64_t id, mg_request_result* o) {
  if (!h || !o) return MG_E_ARG;
  auto it = h->s.reqs.find(id);
  if (it == h->s.reqs.end()) return MG_E_ARG;
  const Request& r = *it->second;
  std::memset(o, 0, sizeof *o);
  o->state = r.state;
  o->n_balls = r.nb;
  o->n_fluid = r.nf;
  o->revolutions_done = r.revs;
  o->substeps_total = r.total;
  o->substeps_done = r.steps;
  o->footprint_bytes = r.footprint;
  o->arrival_s = r.arrival;
  o->start_s = r.start;
  o->done_s = r.done;
  o->power_draw_w = r.power;
  o->torque_nm = r.torque;
  o->collision_rate_per_s = r.coll;
  o->dissipated_power_w = r.diss;
  return MG_OK;
}

int mg_server_stream_json(mg_server* h, uint64_t id, char* buf, size_t cap) {
  if (!h) return MG_E_ARG;
  auto it = h->s.reqs.find(id);
  if (it == h->s.reqs.end()) return MG_E_ARG;
  std::string s = "[";
  for (const RevRec& q : it->second->stream) {
    char b[1024];
    char mix[32];
    if (q.mix_ok) std::snprintf(mix, sizeof mix, "%.9g", (double)q.mix);
    else std::snprintf(mix, sizeof mix, "null");
    int n = std::snprintf(b, sizeof b,
                          "%s{\"rev\":%u,\"sim_time_s\":%.17g,\"power_draw_w\":%.9g,\"torque_nm\":%.9g,"
                          "\"collision_rate_per_s\":%.9g,\"dissipated_power_w\":%.9g,\"mixing_index\":%s,"
                          "\"impact_energy_histogram\":[",
                          s.size() > 1 ? "," : "", q.rev, q.sim_time, (double)q.power, (double)q.torque,
                          (double)q.coll, (double)q.diss, mix);
    s.append(b, (size_t)n);
    for (int k = 0; k < 12; k++) {
      n = std::snprintf(b, sizeof b, "%s%llu", k ? "," : "", (unsigned long long)q.hist[k]);
      s.append(b, (size_t)n);
    }
    s += "]}";
  }
  s += "]";
  if (!buf || s.size() + 1 > cap) return -(int)(s.size() + 1);
  std::memcpy(buf, s.c_str(), s.size() + 1);
  return (int)s.size();
}

int mg_server_release(mg_server* h, uint64_t id) {
  if (!h) return MG_E_ARG;
  auto it = h->s.reqs.find(id);
  if (it == h->s.reqs.end() || it->second->state != 2) return MG_E_ARG;
  h->s.reqs.erase(it);
  return MG_OK;
}

}  // extern "C"
