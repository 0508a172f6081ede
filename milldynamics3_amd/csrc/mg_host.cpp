// Host runtime and C ABI (include/millgpu.h): drum slots, device arenas, the per-substep
// schedule, batched launches (optionally replayed as a hipGraph) and the Simulation getters.
//
// The initial state is seeded on the host exactly like the reference (dem.rs:355-409,
// pbf.rs:1044-1119, rng.rs:9-47) and uploaded once; afterwards all per-substep work runs on the
// device. The drum clock (angle, time) depends only on Params and the substep sequence, so the
// host advances it in f32 exactly as lib.rs:507-508 does and hands each substep its angle's
// sine/cosine through the schedule.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/millgpu.h"
#include "mg_common.h"
#include "mg_layout.h"
#include "mg_params.h"

extern "C" hipError_t mg_launch_substep(const MgDrum* d_drums, int n_drums, int slot, hipStream_t stream);

namespace mg {
namespace {

// Drum substeps on the GPU. Off: an earlier single-kernel form of this engine raised a GPU memory
// aperture violation on its first substep, and no mistake in the code has been shown to cause it
// (every CPU check of the kernel source is clean). Until one is, the step entry points return
// MG_E_DISABLED instead of launching; creation, state copies and the other host paths still work.
constexpr bool kDrumEngineEnabled = false;


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
double powi3d(double x) { return x * (x * x); }

struct SqueezeFilm {  // lubrication.rs:44-71
  double r, c1, c2, pre;
  SqueezeFilm(double k1, double k2, double nu) {
    r = 1.0 / (k1 + k2);
    double q = powi3d(r) * (powi3d(k1) + powi3d(k2));
    c1 = 0.788 + 0.265 * q;
    c2 = std::fmax(0.006 + 0.057 * (q - 0.25), 0.0);
    pre = 3.0 * std::sqrt(2.0) * M_PI * nu * std::pow(r, 1.5);
  }
};

size_t next_pow2(size_t v) {
  size_t p = 1;
  while (p < v) p <<= 1;
  return p;
}

}  // namespace

// One drum slot: host mirror + device arena.
struct Drum {
  Params P;
  MgDrum rec{};  // constants + device pointers (host copy)
  void* arena = nullptr;
  size_t arena_bytes = 0;
  MgSubstepIn* d_sched = nullptr;
  int sched_cap = 0;
  std::vector<MgSubstepIn> h_sched;
  float drum_angle = 0.0f;
  double sim_time = 0.0;
  double e0 = 0.0;
  double debounce_factor = 0.25;
  uint32_t force_density = 0, force_cg = 0;
  std::vector<float> dye;  // host copy (the dye never changes)
  ~Drum() {
    if (arena) (void)hipFree(arena);
    if (d_sched) (void)hipFree(d_sched);
  }
};

struct Batch {
  std::vector<std::unique_ptr<Drum>> drums;
  MgDrum* d_records = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  float last_ms = 0.0f;
  hipGraphExec_t graph_exec = nullptr;
  hipGraph_t graph = nullptr;
  uint32_t graph_k = 0;
  std::vector<struct ::mg_sim*> views;
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
  if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
  if (graph) (void)hipGraphDestroy(graph);
  if (d_records) (void)hipFree(d_records);
  if (ev0) (void)hipEventDestroy(ev0);
  if (ev1) (void)hipEventDestroy(ev1);
  if (stream) (void)hipStreamDestroy(stream);
  for (auto* v : views) delete v;
}

namespace {

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

// Fills the constant part of a drum record (everything that depends only on Params and counts).
bool fill_constants(Drum& D, std::string& err) {
  const Params& P = D.P;
  MgDrum& m = D.rec;
  float td, d, dens, sf;
  uint32_t cnt;
  P.effective_media(td, d, dens, cnt, sf);
  m.R = P.radius_m();
  m.omega = P.omega();
  m.pbf_iters = P.pbf_iterations;
  m.dem_iters = P.dem_iterations;
  m.surface_tension = P.surface_tension;
  m.wettability = P.wettability;
  m.wall_no_slip = P.wall_no_slip;
  m.ball_no_slip = P.ball_no_slip;
  m.viscosity = P.viscosity;
  m.yield_stress = P.yield_stress;
  m.rheology = P.rheology;
  m.rest_bb = P.rest_bb; m.rest_bw = P.rest_bw;
  m.fric_bb = P.fric_bb; m.fric_bw = P.fric_bw; m.fric_bb_s = P.fric_bb_static; m.fric_bw_s = P.fric_bw_static;
  m.fric_vscale = P.fric_vscale; m.rolling = P.rolling;
  if (P.lifter_count > MG_MAX_LIFTERS) { err = "lifters.count exceeds this engine's limit of 32"; return false; }
  m.lcount = P.lifter_count;
  m.l_height = P.lifter_height;
  m.l_base_half = P.lifter_base * 0.5f;
  m.l_top_half = P.lifter_top * 0.5f;
  float phase = P.lifter_phase_deg * (PI_F / 180.0f);  // f32::to_radians
  for (uint32_t k = 0; k < P.lifter_count; k++) {
    float th = phase + (float)k * TAU_F / (float)P.lifter_count;
    m.lsin[k] = std::sin(th);
    m.lcos[k] = std::cos(th);
  }
  const float h = m.h, br = m.br;
  const float dx = 0.5f * h;
  uint32_t bp = 0;
  if (br > 0.0f && dx > 0.0f) {
    float circ = TAU_F * br;
    float rq = std::round(circ / dx);
    uint32_t c = (std::isnan(rq) || rq <= 0.0f) ? 0u : (rq >= 4294967295.0f ? 0xFFFFFFFFu : (uint32_t)rq);
    bp = std::max(c, 6u);
    if (bp > MG_MAX_BP) { err = "ball circumference needs more than 512 boundary points at this resolution"; return false; }
    for (uint32_t k = 0; k < bp; k++) {
      float a = TAU_F * (float)k / (float)bp;
      m.bp_x[k] = std::cos(a) * br;
      m.bp_y[k] = std::sin(a) * br;
    }
  }
  m.bp_cnt = bp;
  m.k_poly6 = 4.0f / (PI_F * powi8(h));
  m.k_spiky = -30.0f / (PI_F * powi5(h));
  m.k_coh = 35840.0f / (209.0f * PI_F * powi8(h));
  m.h6_64 = powi6(h) / 64.0f;
  m.h_adh = std::fmin(h, 2.0f * br);
  m.k_adh = 35840.0f / (163.0f * PI_F * powi8(m.h_adh));
  m.h_c = br + h;
  m.k_poly6_hc = 4.0f / (PI_F * powi8(m.h_c));
  {
    float hc = m.h_c, h2 = hc * hc, t = h2 - 0.0f;
    float p0 = (0.0f >= hc * hc) ? 0.0f : m.k_poly6_hc * t * t * t;
    m.w_hc0 = std::fmax(p0, 1e-12f);
  }
  m.v_max = 5.0f * (std::fabs(m.omega) * m.R + std::sqrt(4.0f * 9.81f * m.R));
  m.impact_thr = 0.1428f * std::sqrt(9.81f * 2.0f * br);
  double a = (double)br, bD = (double)m.R;
  m.dxd = (double)dx;
  m.reach = (float)(8.0 * m.dxd);
  double mud = (double)std::fmax(P.viscosity, 0.0f);
  if (a > 0.0) {
    SqueezeFilm pf(1.0 / a, 1.0 / a, mud), wf(1.0 / a, -1.0 / bD, mud);
    m.sq_pair[0] = pf.r; m.sq_pair[1] = pf.c1; m.sq_pair[2] = pf.c2; m.sq_pair[3] = pf.pre;
    m.sq_wall[0] = wf.r; m.sq_wall[1] = wf.c1; m.sq_wall[2] = wf.c2; m.sq_wall[3] = wf.pre;
  }
  double rr = (double)br;
  m.m3 = (double)m.bmass * 4.0 / 3.0 * rr;
  m.dd = 2.0 * rr;
  m.scale_j = m.m3 * 9.81 * m.dd;
  m.window = D.debounce_factor * std::sqrt(m.dd / 9.81);
  return true;
}

struct ArenaBuilder {
  size_t off = 0;
  std::vector<std::pair<void**, size_t>> items;
  template <class T> void add(T*& p, size_t count) {
    items.push_back({(void**)&p, off});
    off += ((count * sizeof(T) + 255) / 256) * 256;
    if (count == 0) off += 256;
  }
  void bind(void* base) {
    for (auto& it : items) *it.first = (char*)base + it.second;
  }
};

// Sizes the capacities, allocates the arena and uploads the initial state.
bool alloc_drum(Drum& D, const std::vector<float>& bx, const std::vector<float>& fx, std::string& err) {
  MgDrum& m = D.rec;
  const int nf = (int)(fx.size() / 2), nb = (int)(bx.size() / 2);
  m.nf = nf;
  m.nb = nb;
  // capacities (generous bounds from the geometry; overflows are reported, never silent)
  float dx = 0.5f * m.h;
  double cell_area = dx > 0 ? (double)dx * (double)dx * 0.8660254 : 1.0;
  double rh = (double)m.br + (double)m.h;
  int per_ball_fluid = (int)std::ceil(3.14159265 * rh * rh / cell_area * 2.0) + 64;
  m.item_cap = std::min(std::max(per_ball_fluid, 64), 1 << 16);
  m.e_cap = 64 * nf + 64;
  m.fcell_cap = 2 * nf + 512;
  m.bcell_cap = 2 * nb + 512;
  m.pair_cap = 48 * nb + 1024;
  m.grp_cap = std::max(1, nb) * m.item_cap;
  double lr = m.br > 0 ? (2.0 * m.br + m.reach) / m.br : 4.0;
  int per_ball_links = (int)std::ceil(0.95 * lr * lr) + 4;
  m.link_cap = std::max(1, nb) * per_ball_links;
  m.rec_cap = std::max({8 * nf + 64, std::max(1, nb) * m.item_cap, 2 * m.link_cap, m.pair_cap});
  m.imp_cap = m.pair_cap + nb + 16;
  m.tbl_cap = (int32_t)next_pow2((size_t)(4 * nb + 1024));
  m.fpow2 = (int32_t)next_pow2((size_t)std::max(nf, 1));
  m.bpow2 = (int32_t)next_pow2((size_t)std::max(nb, 1));
  size_t cur_n = (size_t)std::max({m.fcell_cap + 2, m.bcell_cap + 2, nf + 2, nb + 3, m.fpow2 + 1, m.bpow2 + 1});
  size_t len = (size_t)nf + 2 * (size_t)nb;
  ArenaBuilder A;
  A.add(m.fx, nf); A.add(m.fv, nf); A.add(m.fguess, nf); A.add(m.dye, nf);
  A.add(m.x0, nf); A.add(m.dp, nf); A.add(m.excess, nf); A.add(m.coh, nf); A.add(m.adh, nf); A.add(m.lv, nf);
  A.add(m.dens, nf); A.add(m.lam, nf); A.add(m.mu_eff, nf); A.add(m.ffac, nf); A.add(m.touched, nf);
  A.add(m.gA_start, m.fcell_cap + 2); A.add(m.gA_idx, nf); A.add(m.gA_key, nf); A.add(m.gA_meta, 8);
  A.add(m.gB_start, m.fcell_cap + 2); A.add(m.gB_idx, nf); A.add(m.gB_key, nf); A.add(m.gB_meta, 8);
  A.add(m.gA_skey, m.fpow2); A.add(m.gB_skey, m.fpow2); A.add(m.gA_sidx, m.fpow2); A.add(m.gB_sidx, m.fpow2);
  A.add(m.cursor, cur_n);
  A.add(m.noff, nf + 2); A.add(m.nbr, m.e_cap); A.add(m.W, m.e_cap); A.add(m.D, m.e_cap);
  A.add(m.cg_x, len); A.add(m.cg_r, len); A.add(m.cg_z, len); A.add(m.cg_p, len); A.add(m.cg_ap, len);
  A.add(m.cg_inv, len);
  A.add(m.bx, nb); A.add(m.bv, nb); A.add(m.bth, nb); A.add(m.bom, nb);
  A.add(m.bx0, nb); A.add(m.bvpre, nb); A.add(m.imp, nb); A.add(m.bv0, nb);
  A.add(m.bth0, nb); A.add(m.bompre, nb); A.add(m.ang, nb); A.add(m.toi, nb); A.add(m.lam_bw, nb);
  A.add(m.bw, nb); A.add(m.bw0, nb);
  A.add(m.gC_start, m.bcell_cap + 2); A.add(m.gC_idx, nb); A.add(m.gC_key, nb); A.add(m.gC_meta, 8);
  A.add(m.gC_skey, m.bpow2); A.add(m.gC_sidx, m.bpow2);
  A.add(m.pair_i, m.pair_cap); A.add(m.pair_j, m.pair_cap); A.add(m.ct_perm, m.pair_cap); A.add(m.ct_start, nb + 2);
  A.add(m.lam_bb, m.pair_cap);
  A.add(m.g_fluid, m.grp_cap); A.add(m.g_ball, m.grp_cap); A.add(m.g_bstart, nb + 2); A.add(m.gf_start, nf + 2);
  A.add(m.gf_perm, m.grp_cap); A.add(m.g_c, m.grp_cap); A.add(m.g_l2, m.grp_cap); A.add(m.g_lc, m.grp_cap);
  A.add(m.l_ball, m.link_cap); A.add(m.l_other, m.link_cap); A.add(m.l_bstart, nb + 2); A.add(m.li_start, nb + 3);
  A.add(m.li_perm, 2 * (size_t)m.link_cap); A.add(m.l_n, m.link_cap); A.add(m.l_c, m.link_cap);
  size_t items = (size_t)std::max(1, nb) * m.item_cap;
  A.add(m.it_j, items); A.add(m.it_v, items); A.add(m.it_w, items); A.add(m.it_w2, items); A.add(m.it_n, nb + 1);
  A.add(m.rec_key, m.rec_cap); A.add(m.rec_ord, m.rec_cap); A.add(m.rec_perm, m.rec_cap);
  A.add(m.rec_seg, std::max(nf, nb) + 3);
  A.add(m.rec_v0, m.rec_cap); A.add(m.rec_v1, m.rec_cap); A.add(m.rec_v2, m.rec_cap);
  A.add(m.impacts, m.imp_cap);
  A.add(m.tbl_key, 2 * (size_t)m.tbl_cap); A.add(m.tbl_time, 2 * (size_t)m.tbl_cap);
  A.add(m.sc, 1); A.add(m.st, 1);
  D.arena_bytes = A.off;
  if (hipMalloc(&D.arena, D.arena_bytes) != hipSuccess) { err = "hipMalloc failed for the drum arena"; return false; }
  A.bind(D.arena);
  if (hipMemset(D.arena, 0, D.arena_bytes) != hipSuccess) { err = "hipMemset failed"; return false; }
  if (hipMemset(m.tbl_key, 0xFF, 2 * (size_t)m.tbl_cap * sizeof(uint64_t)) != hipSuccess) { err = "hipMemset failed"; return false; }
  if (nf) {
    hipMemcpy(m.fx, fx.data(), nf * sizeof(float2), hipMemcpyHostToDevice);
    hipMemcpy(m.dye, D.dye.data(), nf * sizeof(float), hipMemcpyHostToDevice);
  }
  if (nb) hipMemcpy(m.bx, bx.data(), nb * sizeof(float2), hipMemcpyHostToDevice);
  D.sched_cap = 64;
  if (hipMalloc(&D.d_sched, D.sched_cap * sizeof(MgSubstepIn)) != hipSuccess) { err = "hipMalloc failed"; return false; }
  D.h_sched.resize(D.sched_cap);
  m.sched = D.d_sched;
  return hipGetLastError() == hipSuccess || (err = "HIP error during upload", false);
}

// Host f64 mechanical energy from a state copy (dem.rs:265-273 + pbf.rs:1145-1150).
double mech_energy(const Drum& D, const std::vector<float>& bx, const std::vector<float>& bv,
                   const std::vector<float>& bom, const std::vector<float>& fx, const std::vector<float>& fv) {
  const MgDrum& m = D.rec;
  double e = 0.0;
  for (int i = 0; i < m.nb; i++) {
    float vx = bv[2 * i], vy = bv[2 * i + 1];
    e += 0.5 * (double)m.bmass * (double)(vx * vx + vy * vy);
    e += 0.5 * (double)m.binertia * (double)(bom[i] * bom[i]);
    e += (double)m.bmass * (double)9.81f * (double)bx[2 * i + 1];
  }
  double py = 0.0, ke = 0.0;
  for (int i = 0; i < m.nf; i++) py += (double)fx[2 * i + 1];
  for (int i = 0; i < m.nf; i++) {
    float vx = fv[2 * i], vy = fv[2 * i + 1];
    ke += (double)(vx * vx + vy * vy);
  }
  return e + (0.5 * (double)m.fmass * ke + py * (double)m.fmass * (double)9.81f);
}

bool download(Drum& D, std::vector<float>& bx, std::vector<float>& bv, std::vector<float>& bth,
              std::vector<float>& bom, std::vector<float>& fx, std::vector<float>& fv) {
  const MgDrum& m = D.rec;
  bx.resize(2 * m.nb); bv.resize(2 * m.nb); bth.resize(m.nb); bom.resize(m.nb);
  fx.resize(2 * m.nf); fv.resize(2 * m.nf);
  bool ok = true;
  if (m.nb) {
    ok &= hipMemcpy(bx.data(), m.bx, m.nb * sizeof(float2), hipMemcpyDeviceToHost) == hipSuccess;
    ok &= hipMemcpy(bv.data(), m.bv, m.nb * sizeof(float2), hipMemcpyDeviceToHost) == hipSuccess;
    ok &= hipMemcpy(bth.data(), m.bth, m.nb * sizeof(float), hipMemcpyDeviceToHost) == hipSuccess;
    ok &= hipMemcpy(bom.data(), m.bom, m.nb * sizeof(float), hipMemcpyDeviceToHost) == hipSuccess;
  }
  if (m.nf) {
    ok &= hipMemcpy(fx.data(), m.fx, m.nf * sizeof(float2), hipMemcpyDeviceToHost) == hipSuccess;
    ok &= hipMemcpy(fv.data(), m.fv, m.nf * sizeof(float2), hipMemcpyDeviceToHost) == hipSuccess;
  }
  return ok;
}

// Builds a drum from Params (+ optional initial balls), Simulation::build (lib.rs:138-207).
std::unique_ptr<Drum> make_drum(const Params& P, const float* init, size_t n_init, std::string& err) {
  auto D = std::make_unique<Drum>();
  D->P = P;
  std::vector<float> bx, fx;
  float br, bmass, bin, btrue;
  seed_balls(P, bx, br, bmass, bin, btrue);
  if (init) bx.assign(init, init + 2 * n_init);
  float h, fmass, rest;
  if (P.slurry_enabled) seed_fluid(P, P.slurry_fill, bx, br, fx, D->dye, h, fmass, rest);
  else seed_fluid(P, 0.0f, {}, 0.0f, fx, D->dye, h, fmass, rest);
  MgDrum& m = D->rec;
  m.br = br; m.bmass = bmass; m.binertia = bin; m.btrue = btrue;
  m.h = h; m.fmass = fmass; m.rest = rest;
  if (!fill_constants(*D, err)) return nullptr;
  if (!alloc_drum(*D, bx, fx, err)) return nullptr;
  std::vector<float> bv(bx.size(), 0.0f), bom(bx.size() / 2, 0.0f), fv(fx.size(), 0.0f);
  if (init) {
    float w = P.omega();
    for (size_t i = 0; i < bx.size() / 2; i++) { bv[2 * i] = w * (-bx[2 * i + 1]); bv[2 * i + 1] = w * bx[2 * i]; bom[i] = w; }
    for (size_t i = 0; i < fx.size() / 2; i++) { fv[2 * i] = w * (-fx[2 * i + 1]); fv[2 * i + 1] = w * fx[2 * i]; }
    if (m.nb) {
      hipMemcpy(m.bv, bv.data(), m.nb * sizeof(float2), hipMemcpyHostToDevice);
      hipMemcpy(m.bom, bom.data(), m.nb * sizeof(float), hipMemcpyHostToDevice);
    }
    if (m.nf) hipMemcpy(m.fv, fv.data(), m.nf * sizeof(float2), hipMemcpyHostToDevice);
  }
  D->e0 = mech_energy(*D, bx, bv, bom, fx, fv);
  return D;
}

// Writes the schedule entry for the drum's next substep and advances the host clock
// (lib.rs:507-508: angle = (angle + omega dt).rem_euclid(TAU), time += dt).
MgSubstepIn next_substep(Drum& D, float dt) {
  MgSubstepIn s{};
  float om = D.rec.omega;
  s.dt = dt;
  s.angle = D.drum_angle;
  s.angle_next = D.drum_angle + om * dt;
  s.ca0 = std::cos(-s.angle); s.sa0 = std::sin(-s.angle);
  s.cn = std::cos(-s.angle_next); s.sn = std::sin(-s.angle_next);
  s.cp = std::cos(s.angle_next); s.sp = std::sin(s.angle_next);
  s.alpha = 1.0f - std::exp(-dt / 1.0f);
  s.now = D.sim_time;
  s.force_density = D.force_density;
  s.force_cg = D.force_cg;
  float a = std::fmod(s.angle_next, TAU_F);
  D.drum_angle = a < 0.0f ? a + std::fabs(TAU_F) : a;
  D.sim_time += (double)dt;
  return s;
}

bool upload_records(Batch& B, std::string& err) {
  size_t n = B.drums.size();
  if (!B.d_records) {
    if (hipMalloc(&B.d_records, n * sizeof(MgDrum)) != hipSuccess) { err = "hipMalloc failed"; return false; }
  }
  std::vector<MgDrum> recs(n);
  for (size_t i = 0; i < n; i++) recs[i] = B.drums[i]->rec;
  if (hipMemcpy(B.d_records, recs.data(), n * sizeof(MgDrum), hipMemcpyHostToDevice) != hipSuccess) {
    err = "hipMemcpy failed";
    return false;
  }
  return true;
}

Batch* make_batch(std::vector<std::unique_ptr<Drum>>&& drums, std::string& err) {
  auto* B = new Batch();
  B->drums = std::move(drums);
  if (hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking) != hipSuccess) { err = "hipStreamCreate failed"; delete B; return nullptr; }
  hipEventCreate(&B->ev0);
  hipEventCreate(&B->ev1);
  if (!upload_records(*B, err)) { delete B; return nullptr; }
  for (uint32_t i = 0; i < B->drums.size(); i++) B->views.push_back(new ::mg_sim{B, i, false});
  return B;
}

// Runs `k` substeps for every drum (drum d with dts[d] per substep, or its fixed substep).
int run_substeps(Batch& B, uint32_t k, const float* dt_override, bool use_graph) {
  if (!kDrumEngineEnabled) return MG_E_DISABLED;
  if (k == 0) return MG_OK;
  uint32_t done = 0;
  hipEventRecord(B.ev0, B.stream);
  while (done < k) {
    uint32_t chunk = std::min<uint32_t>(k - done, 64);
    for (auto& Dp : B.drums) {
      Drum& D = *Dp;
      float dt = dt_override ? *dt_override : 1.0f / (60.0f * (float)std::max(D.P.effective_substeps(), 1u));
      for (uint32_t s = 0; s < chunk; s++) D.h_sched[s] = next_substep(D, dt);
      if (hipMemcpyAsync(D.d_sched, D.h_sched.data(), chunk * sizeof(MgSubstepIn), hipMemcpyHostToDevice, B.stream) != hipSuccess)
        return MG_E_DEVICE;
    }
    if (use_graph) {
      if (!B.graph_exec || B.graph_k != chunk) {
        if (B.graph_exec) { hipGraphExecDestroy(B.graph_exec); B.graph_exec = nullptr; }
        if (B.graph) { hipGraphDestroy(B.graph); B.graph = nullptr; }
        hipStream_t cs = nullptr;
        if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) return MG_E_DEVICE;
        if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
          (void)hipStreamDestroy(cs);
          return MG_E_DEVICE;
        }
        bool launched = true;
        for (uint32_t s = 0; s < chunk; s++)
          launched &= mg_launch_substep(B.d_records, (int)B.drums.size(), (int)s, cs) == hipSuccess;
        // End the capture on every path, so the stream never stays in capture mode.
        hipGraph_t g = nullptr;
        const bool ended = hipStreamEndCapture(cs, &g) == hipSuccess;
        (void)hipStreamDestroy(cs);
        if (!ended || !launched) {
          if (g) (void)hipGraphDestroy(g);
          return MG_E_DEVICE;
        }
        B.graph = g;
        if (hipGraphInstantiate(&B.graph_exec, B.graph, nullptr, nullptr, 0) != hipSuccess) return MG_E_DEVICE;
        B.graph_k = chunk;
      }
      if (hipGraphLaunch(B.graph_exec, B.stream) != hipSuccess) return MG_E_DEVICE;
    } else {
      for (uint32_t s = 0; s < chunk; s++)
        if (mg_launch_substep(B.d_records, (int)B.drums.size(), (int)s, B.stream) != hipSuccess) return MG_E_DEVICE;
    }
    done += chunk;
  }
  hipEventRecord(B.ev1, B.stream);
  return MG_OK;
}

int sync_batch(Batch& B, std::string& err) {
  if (hipStreamSynchronize(B.stream) != hipSuccess) { err = "HIP stream error"; return MG_E_DEVICE; }
  hipEventElapsedTime(&B.last_ms, B.ev0, B.ev1);
  for (size_t i = 0; i < B.drums.size(); i++) {
    uint32_t e = 0;
    hipMemcpy(&e, &B.drums[i]->rec.sc->err, sizeof(uint32_t), hipMemcpyDeviceToHost);
    if (e) {
      char b[128];
      std::snprintf(b, sizeof b, "drum %zu: capacity overflow (error bits 0x%x)", i, e);
      err = b;
      return MG_E_CAPACITY;
    }
  }
  return MG_OK;
}

MgScalars read_scalars(Drum& D) {
  MgScalars s{};
  hipMemcpy(&s, D.rec.sc, sizeof(MgScalars), hipMemcpyDeviceToHost);
  return s;
}

}  // namespace
}  // namespace mg

using namespace mg;

static Drum* drum_of(const mg_sim* s) { return s ? s->batch->drums[s->index].get() : nullptr; }
static int sync_one(mg_sim* s) {
  std::string e;
  return sync_batch(*s->batch, e);
}

extern "C" {

const char* mg_version(void) { return "millgpu 0.1 (gfx950)"; }
int mg_drum_engine_enabled(void) { return mg::kDrumEngineEnabled ? 1 : 0; }

static int create_impl(const char* json, const float* xy, size_t n, mg_sim** out, char* err, size_t errlen) {
  if (!out) return MG_E_ARG;
  *out = nullptr;
  Params P;
  std::string e;
  if (json && !Params::from_json(json, P, e)) { set_err(err, errlen, e); return MG_E_PARAMS; }
  if (!P.validate(e)) { set_err(err, errlen, e); return MG_E_PARAMS; }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) { set_err(err, errlen, "no HIP device"); return MG_E_DEVICE; }
  auto D = make_drum(P, xy, n, e);
  if (!D) { set_err(err, errlen, e); return MG_E_DEVICE; }
  std::vector<std::unique_ptr<Drum>> v;
  v.push_back(std::move(D));
  Batch* B = make_batch(std::move(v), e);
  if (!B) { set_err(err, errlen, e); return MG_E_DEVICE; }
  *out = new mg_sim{B, 0, true};
  return MG_OK;
}

int mg_sim_create(const char* params_json, mg_sim** out, char* err, size_t errlen) {
  return create_impl(params_json, nullptr, 0, out, err, errlen);
}
int mg_sim_create_with_balls(const char* params_json, const float* xy, size_t n_balls, mg_sim** out, char* err,
                             size_t errlen) {
  if (!xy && n_balls) return MG_E_ARG;
  return create_impl(params_json, xy ? xy : (const float*)"", n_balls, out, err, errlen);
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
  if (!Params::from_json(params_json, P, e) || !P.validate(e)) { set_err(err, errlen, e); return MG_E_PARAMS; }
  // Hot-apply (lib.rs:252-274): the populations keep their sizes, radius, masses and h; every
  // other parameter read by the step takes effect from the next substep.
  MgDrum keep = D->rec;
  D->P = P;
  if (!fill_constants(*D, e)) { D->rec = keep; set_err(err, errlen, e); return MG_E_PARAMS; }
  if (!upload_records(*sim->batch, e)) { set_err(err, errlen, e); return MG_E_DEVICE; }
  return MG_OK;
}

int mg_sim_step(mg_sim* sim, float dt) {
  Drum* D = drum_of(sim);
  if (!D) return MG_E_ARG;
  uint32_t ns = std::max(D->P.effective_substeps(), 1u);
  float sd = dt / (float)ns;
  mg_sim_reset_frame_stats(sim);
  int rc = run_substeps(*sim->batch, ns, &sd, false);
  if (rc) return rc;
  return sync_one(sim);
}
int mg_sim_step_fixed(mg_sim* sim) { return mg_sim_step_fixed_n(sim, 1); }
int mg_sim_step_fixed_n(mg_sim* sim, uint32_t n) {
  if (!drum_of(sim)) return MG_E_ARG;
  int rc = run_substeps(*sim->batch, n, nullptr, false);
  if (rc) return rc;
  return sync_one(sim);
}
void mg_sim_reset_frame_stats(mg_sim* sim) {
  Drum* D = drum_of(sim);
  if (!D) return;
  float z = 0.0f;
  uint32_t zu = 0;
  hipMemcpy(&D->rec.sc->max_disp, &z, sizeof z, hipMemcpyHostToDevice);
  hipMemcpy(&D->rec.sc->clamp_hits_total, &zu, sizeof zu, hipMemcpyHostToDevice);
}
float mg_sim_fixed_sub_dt(const mg_sim* sim) {
  Drum* D = drum_of(sim);
  return D ? 1.0f / (60.0f * (float)std::max(D->P.effective_substeps(), 1u)) : 0.0f;
}
uint32_t mg_sim_effective_substeps(const mg_sim* sim) {
  Drum* D = drum_of(sim);
  return D ? D->P.effective_substeps() : 0u;
}
size_t mg_sim_ball_count(const mg_sim* sim) { Drum* D = drum_of(sim); return D ? (size_t)D->rec.nb : 0; }
size_t mg_sim_fluid_count(const mg_sim* sim) { Drum* D = drum_of(sim); return D ? (size_t)D->rec.nf : 0; }

int mg_sim_copy_balls(mg_sim* sim, float* x2n, float* v2n, float* theta, float* omega) {
  Drum* D = drum_of(sim);
  if (!D) return MG_E_ARG;
  hipStreamSynchronize(sim->batch->stream);
  const MgDrum& m = D->rec;
  bool ok = true;
  if (m.nb) {
    if (x2n) ok &= hipMemcpy(x2n, m.bx, m.nb * sizeof(float2), hipMemcpyDeviceToHost) == hipSuccess;
    if (v2n) ok &= hipMemcpy(v2n, m.bv, m.nb * sizeof(float2), hipMemcpyDeviceToHost) == hipSuccess;
    if (theta) ok &= hipMemcpy(theta, m.bth, m.nb * sizeof(float), hipMemcpyDeviceToHost) == hipSuccess;
    if (omega) ok &= hipMemcpy(omega, m.bom, m.nb * sizeof(float), hipMemcpyDeviceToHost) == hipSuccess;
  }
  return ok ? MG_OK : MG_E_DEVICE;
}
int mg_sim_copy_fluid(mg_sim* sim, float* x2n, float* v2n, float* dye) {
  Drum* D = drum_of(sim);
  if (!D) return MG_E_ARG;
  hipStreamSynchronize(sim->batch->stream);
  const MgDrum& m = D->rec;
  bool ok = true;
  if (m.nf) {
    if (x2n) ok &= hipMemcpy(x2n, m.fx, m.nf * sizeof(float2), hipMemcpyDeviceToHost) == hipSuccess;
    if (v2n) ok &= hipMemcpy(v2n, m.fv, m.nf * sizeof(float2), hipMemcpyDeviceToHost) == hipSuccess;
    if (dye) std::memcpy(dye, D->dye.data(), m.nf * sizeof(float));
  }
  return ok ? MG_OK : MG_E_DEVICE;
}
int mg_sim_set_state(mg_sim* sim, const float* bx, const float* bv, const float* th, const float* om,
                     const float* fx, const float* fv, const float* guess, int has_guess) {
  Drum* D = drum_of(sim);
  if (!D) return MG_E_ARG;
  hipStreamSynchronize(sim->batch->stream);
  const MgDrum& m = D->rec;
  bool ok = true;
  if (m.nb) {
    if (bx) ok &= hipMemcpy(m.bx, bx, m.nb * sizeof(float2), hipMemcpyHostToDevice) == hipSuccess;
    if (bv) ok &= hipMemcpy(m.bv, bv, m.nb * sizeof(float2), hipMemcpyHostToDevice) == hipSuccess;
    if (th) ok &= hipMemcpy(m.bth, th, m.nb * sizeof(float), hipMemcpyHostToDevice) == hipSuccess;
    if (om) ok &= hipMemcpy(m.bom, om, m.nb * sizeof(float), hipMemcpyHostToDevice) == hipSuccess;
  }
  if (m.nf) {
    if (fx) ok &= hipMemcpy(m.fx, fx, m.nf * sizeof(float2), hipMemcpyHostToDevice) == hipSuccess;
    if (fv) ok &= hipMemcpy(m.fv, fv, m.nf * sizeof(float2), hipMemcpyHostToDevice) == hipSuccess;
    if (guess) ok &= hipMemcpy(m.fguess, guess, m.nf * sizeof(float2), hipMemcpyHostToDevice) == hipSuccess;
  }
  uint32_t hg = has_guess ? 1u : 0u;
  ok &= hipMemcpy(&m.sc->has_guess, &hg, sizeof hg, hipMemcpyHostToDevice) == hipSuccess;
  return ok ? MG_OK : MG_E_DEVICE;
}
int mg_sim_copy_guess(mg_sim* sim, float* guess2n, int* has_guess) {
  Drum* D = drum_of(sim);
  if (!D) return MG_E_ARG;
  hipStreamSynchronize(sim->batch->stream);
  MgScalars s = read_scalars(*D);
  if (has_guess) *has_guess = (int)s.has_guess;
  if (guess2n && D->rec.nf) hipMemcpy(guess2n, D->rec.fguess, D->rec.nf * sizeof(float2), hipMemcpyDeviceToHost);
  return MG_OK;
}
int mg_sim_set_clock(mg_sim* sim, float drum_angle, double sim_time) {
  Drum* D = drum_of(sim);
  if (!D) return MG_E_ARG;
  D->drum_angle = drum_angle;
  D->sim_time = sim_time;
  return MG_OK;
}
float mg_sim_drum_angle(const mg_sim* sim) { Drum* D = drum_of(sim); return D ? D->drum_angle : 0.0f; }
double mg_sim_sim_time(const mg_sim* sim) { Drum* D = drum_of(sim); return D ? D->sim_time : 0.0; }

static MgScalars scal(mg_sim* sim) {
  Drum* D = drum_of(sim);
  hipStreamSynchronize(sim->batch->stream);
  return read_scalars(*D);
}
uint32_t mg_sim_viscosity_iterations(mg_sim* sim) { return drum_of(sim) ? scal(sim).visc_iters : 0u; }
float mg_sim_mean_shear_rate_per_s(mg_sim* sim) { return drum_of(sim) ? scal(sim).mean_shear : 0.0f; }
float mg_sim_power_draw_w(mg_sim* sim) { return drum_of(sim) ? scal(sim).power_draw_w : 0.0f; }
float mg_sim_torque_nm(mg_sim* sim) { return drum_of(sim) ? scal(sim).torque_nm : 0.0f; }
float mg_sim_collision_rate_per_s(mg_sim* sim) { return drum_of(sim) ? scal(sim).collision_rate_per_s : 0.0f; }
float mg_sim_dissipated_power_w(mg_sim* sim) { return drum_of(sim) ? scal(sim).dissipated_power_w : 0.0f; }
float mg_sim_max_substep_displacement_over_diameter(mg_sim* sim) { return drum_of(sim) ? scal(sim).max_disp : 0.0f; }
uint32_t mg_sim_coupling_clamp_hits(mg_sim* sim) { return drum_of(sim) ? scal(sim).clamp_hits_total : 0u; }

int mg_sim_stats(mg_sim* sim, double* o, size_t cap) {
  if (!drum_of(sim) || !o) return MG_E_ARG;
  MgScalars s = scal(sim);
  double v[] = {(double)s.drum_angle, s.sim_time, s.power_draw_w, s.torque_nm, s.collision_rate_per_s,
                s.dissipated_power_w, s.max_disp, (double)s.clamp_hits_total, (double)s.visc_iters,
                (double)s.density_iters, s.mean_shear, s.fluid_wall_work, s.dem_wall_work, s.dem_dissipated,
                s.dem_coupling_work, (double)s.collisions, (double)s.n_groups, (double)s.n_links,
                (double)s.links_in_range, (double)s.n_pairs, s.ke_noslip, s.ke_drag, s.ke_buoy, s.ke_visc,
                s.ke_clamp, s.wall_imp_x, s.wall_imp_y, s.wall_ang, (double)s.err, (double)s.n_impacts};
  size_t n = sizeof(v) / sizeof(v[0]);
  std::memcpy(o, v, std::min(n, cap) * sizeof(double));
  return (int)n;
}
int mg_sim_stress_state(mg_sim* sim, double* o) {
  if (!drum_of(sim) || !o) return MG_E_ARG;
  MgScalars s = scal(sim);
  int k = 0;
  for (int c = 0; c < 4; c++) o[k++] = s.st_count[c];
  for (int c = 0; c < 4; c++) o[k++] = s.st_se[c];
  for (int c = 0; c < 28; c++) o[k++] = s.st_hist[c];
  for (int c = 0; c < 6; c++) o[k++] = s.st_above[c];
  o[k++] = s.st_cum_wall;
  for (int c = 0; c < 6; c++) o[k++] = s.st_cum_above[c];
  return MG_OK;
}
int mg_sim_impacts(mg_sim* sim, double* o, size_t cap, size_t* n_out) {
  Drum* D = drum_of(sim);
  if (!D) return MG_E_ARG;
  MgScalars s = scal(sim);
  std::vector<MgImpact> rec(s.n_impacts);
  if (s.n_impacts) hipMemcpy(rec.data(), D->rec.impacts, s.n_impacts * sizeof(MgImpact), hipMemcpyDeviceToHost);
  if (n_out) *n_out = s.n_impacts;
  for (size_t i = 0; i < rec.size() && i < cap; i++) {
    o[6 * i] = rec[i].kind; o[6 * i + 1] = rec[i].ball;
    o[6 * i + 2] = rec[i].partner == MG_NO_SLOT ? -1.0 : (double)rec[i].partner;
    o[6 * i + 3] = rec[i].vn; o[6 * i + 4] = rec[i].slip; o[6 * i + 5] = rec[i].e;
  }
  return MG_OK;
}

// metrics::mixing_index (metrics.rs:506-576) over the current fluid positions and dye.
int mg_sim_mixing_index(mg_sim* sim, float* out, int* okp) {
  Drum* D = drum_of(sim);
  if (!D || !out || !okp) return MG_E_ARG;
  *okp = 0;
  *out = 0.0f;
  size_t n = (size_t)D->rec.nf;
  float r = D->rec.R;
  if (n == 0 || r <= 0.0f) return MG_OK;
  std::vector<float> fx(2 * n);
  mg_sim_copy_fluid(sim, fx.data(), nullptr, nullptr);
  const int G = 16;
  float cell = 2.0f * r / (float)G;
  std::vector<float> dsum(G * G, 0.0f);
  std::vector<uint32_t> cnt(G * G, 0);
  float tot = 0.0f;
  for (size_t i = 0; i < n; i++) {
    int gi = std::min(std::max(as_i32(std::floor((fx[2 * i] + r) / cell)), 0), G - 1);
    int gj = std::min(std::max(as_i32(std::floor((fx[2 * i + 1] + r) / cell)), 0), G - 1);
    size_t idx = (size_t)gj * G + gi;
    dsum[idx] += D->dye[i];
    cnt[idx]++;
    tot += D->dye[i];
  }
  float pm = tot / (float)n;
  std::vector<float> means;
  std::vector<uint32_t> cts;
  for (size_t q = 0; q < dsum.size(); q++)
    if (cnt[q]) { means.push_back(dsum[q] / (float)cnt[q]); cts.push_back(cnt[q]); }
  if (means.size() < 2) return MG_OK;
  uint32_t wsu = 0;
  for (auto c : cts) wsu += c;
  float ws = (float)wsu;
  float ov = 0.0f;
  for (size_t q = 0; q < means.size(); q++) { float d = means[q] - pm; ov += (float)cts[q] * (d * d); }
  ov = ov / ws;
  float s0 = pm * (1.0f - pm);
  float wa = 0.0f;
  for (auto c : cts) wa += (float)c * (float)c;
  wa = wa / ws;
  float sr = s0 / std::fmax(wa, 1.0f);
  if (s0 < 1e-9f) return MG_OK;
  float den = s0 - sr;
  *okp = 1;
  if (std::fabs(den) < 1e-9f) { *out = 1.0f; return MG_OK; }
  float mm = (s0 - ov) / den;
  *out = mm < 0.0f ? 0.0f : (mm > 1.0f ? 1.0f : mm);
  return MG_OK;
}

int mg_sim_energy_budget(mg_sim* sim, mg_energy_budget* o) {
  Drum* D = drum_of(sim);
  if (!D || !o) return MG_E_ARG;
  MgScalars s = scal(sim);
  std::vector<float> bx, bv, bth, bom, fx, fv;
  download(*D, bx, bv, bth, bom, fx, fv);
  o->elapsed_s = s.eb_elapsed; o->shaft_work_j = s.eb_shaft; o->dem_wall_work_j = s.eb_dem_wall;
  o->fluid_wall_work_j = s.eb_fluid_wall; o->ball_contact_dissipation_j = s.eb_contact_diss;
  o->fluid_wall_slip_j = s.eb_wall_slip; o->fluid_viscous_j = s.eb_viscous; o->fluid_clamp_removed_j = s.eb_clamp_removed;
  o->interface_created_j = s.eb_interface; o->fluid_wall_angular_impulse = s.eb_fluid_wall_ang;
  o->delta_mechanical_j = mech_energy(*D, bx, bv, bom, fx, fv) - D->e0;
  double expl = o->shaft_work_j - o->ball_contact_dissipation_j - o->fluid_wall_slip_j - o->fluid_viscous_j -
                o->fluid_clamp_removed_j + o->interface_created_j;
  o->unattributed_j = o->delta_mechanical_j - expl;
  return MG_OK;
}
int mg_sim_reset_energy_budget(mg_sim* sim) {
  Drum* D = drum_of(sim);
  if (!D) return MG_E_ARG;
  MgScalars s = scal(sim);
  s.eb_elapsed = s.eb_shaft = s.eb_dem_wall = s.eb_fluid_wall = s.eb_contact_diss = s.eb_wall_slip = 0.0;
  s.eb_viscous = s.eb_clamp_removed = s.eb_interface = s.eb_fluid_wall_ang = 0.0;
  hipMemcpy(D->rec.sc, &s, sizeof s, hipMemcpyHostToDevice);
  std::vector<float> bx, bv, bth, bom, fx, fv;
  download(*D, bx, bv, bth, bom, fx, fv);
  D->e0 = mech_energy(*D, bx, bv, bom, fx, fv);
  return MG_OK;
}
int mg_sim_set_stress_debounce_factor(mg_sim* sim, double factor) {
  Drum* D = drum_of(sim);
  if (!D) return MG_E_ARG;
  D->debounce_factor = factor;
  D->rec.window = factor * std::sqrt(D->rec.dd / 9.81);
  std::string e;
  return upload_records(*sim->batch, e) ? MG_OK : MG_E_DEVICE;
}

int mg_sim_metrics_json(mg_sim* sim, char* buf, size_t cap) {
  Drum* D = drum_of(sim);
  if (!D || !buf) return MG_E_ARG;
  MgScalars s = scal(sim);
  float mix = 0.0f;
  int ok = 0;
  mg_sim_mixing_index(sim, &mix, &ok);
  float td, d, dens, sf;
  uint32_t cnt;
  D->P.effective_media(td, d, dens, cnt, sf);
  double n = std::max(1, D->rec.nb);
  std::string hist;
  for (int k = 0; k < 28; k++) {
    char t[40];
    std::snprintf(t, sizeof t, "%s%.9g", k ? "," : "", (double)(float)s.st_hist[k]);
    hist += t;
  }
  std::string edges;
  for (int k = 0; k <= 28; k++) {
    char t[40];
    std::snprintf(t, sizeof t, "%s%.9g", k ? "," : "", (double)(float)(std::pow(10.0, -3.0 + k / 4.0) * D->rec.scale_j));
    edges += t;
  }
  double total = s.st_count[0] + s.st_count[1] + s.st_count[2] + s.st_count[3];
  char mixs[32];
  if (ok) std::snprintf(mixs, sizeof mixs, "%.9g", mix); else std::snprintf(mixs, sizeof mixs, "null");
  char b[8192];
  int len = std::snprintf(
      b, sizeof b,
      "{\"power_draw_w\":%.9g,\"torque_nm\":%.9g,\"collision_rate_per_s\":%.9g,\"dissipated_power_w\":%.9g,"
      "\"mixing_index\":%s,\"mean_shear_rate_per_s\":%.9g,\"viscosity_solver_iterations\":%u,"
      "\"coupling_clamp_hits\":%u,\"max_substep_displacement_over_diameter\":%.9g,"
      "\"effective_ball_diameter_m\":%.9g,\"simulated_ball_count\":%u,\"true_ball_count\":%u,"
      "\"coarse_graining_factor\":%.9g,\"impact_energy_histogram\":{\"bin_edges_j\":[%s],\"counts_per_s\":[%s]},"
      "\"stress\":{\"impacts_total_per_ball_per_s\":%.9g,\"ball_mass_3d_kg\":%.9g,\"scale_energy_j\":%.9g}}",
      s.power_draw_w, s.torque_nm, s.collision_rate_per_s, s.dissipated_power_w, mixs, s.mean_shear, s.visc_iters,
      s.clamp_hits_total, s.max_disp, d, cnt, (unsigned)std::fmax(std::round(D->P.true_ball_count()), 0.0f), sf,
      edges.c_str(), hist.c_str(), (double)(float)(total / n), (double)(float)D->rec.m3, (double)(float)D->rec.scale_j);
  if (len < 0 || (size_t)len + 1 > cap) return -(len + 1);
  std::memcpy(buf, b, (size_t)len + 1);
  return len;
}

// Snapshot blob: header, Params JSON, scalars, state arrays, guess, stress table, clock.
int mg_sim_save_state(mg_sim* sim, void* buf, size_t cap, size_t* needed) {
  Drum* D = drum_of(sim);
  if (!D || !needed) return MG_E_ARG;
  const MgDrum& m = D->rec;
  MgScalars s = scal(sim);
  size_t tb = 2 * (size_t)m.tbl_cap;
  size_t sz = 16 + sizeof(MgScalars) + (size_t)m.nb * 24 + (size_t)m.nf * 24 + tb * 16 + 16 + 8;
  *needed = sz;
  if (!buf || cap < sz) return MG_E_ARG;
  char* p = (char*)buf;
  uint32_t hdr[4] = {0x4D475331u, (uint32_t)m.nb, (uint32_t)m.nf, (uint32_t)m.tbl_cap};
  std::memcpy(p, hdr, 16); p += 16;
  std::memcpy(p, &s, sizeof s); p += sizeof s;
  auto d2h = [&](const void* src, size_t bytes) { if (bytes) hipMemcpy(p, src, bytes, hipMemcpyDeviceToHost); p += bytes; };
  d2h(m.bx, m.nb * 8); d2h(m.bv, m.nb * 8); d2h(m.bth, m.nb * 4); d2h(m.bom, m.nb * 4);
  d2h(m.fx, m.nf * 8); d2h(m.fv, m.nf * 8); d2h(m.fguess, m.nf * 8);
  d2h(m.tbl_key, tb * 8); d2h(m.tbl_time, tb * 8);
  std::memcpy(p, &D->drum_angle, 4); p += 4;
  float pad = 0; std::memcpy(p, &pad, 4); p += 4;
  std::memcpy(p, &D->sim_time, 8); p += 8;
  std::memcpy(p, &D->e0, 8); p += 8;
  return MG_OK;
}
int mg_sim_load_state(mg_sim* sim, const void* buf, size_t len) {
  Drum* D = drum_of(sim);
  if (!D || !buf) return MG_E_ARG;
  const MgDrum& m = D->rec;
  const char* p = (const char*)buf;
  uint32_t hdr[4];
  if (len < 16) return MG_E_ARG;
  std::memcpy(hdr, p, 16); p += 16;
  if (hdr[0] != 0x4D475331u || hdr[1] != (uint32_t)m.nb || hdr[2] != (uint32_t)m.nf || hdr[3] != (uint32_t)m.tbl_cap) return MG_E_ARG;
  size_t tb = 2 * (size_t)m.tbl_cap;
  size_t sz = 16 + sizeof(MgScalars) + (size_t)m.nb * 24 + (size_t)m.nf * 24 + tb * 16 + 16 + 8;
  if (len < sz) return MG_E_ARG;
  hipStreamSynchronize(sim->batch->stream);
  hipMemcpy(m.sc, p, sizeof(MgScalars), hipMemcpyHostToDevice); p += sizeof(MgScalars);
  auto h2d = [&](void* dst, size_t bytes) { if (bytes) hipMemcpy(dst, p, bytes, hipMemcpyHostToDevice); p += bytes; };
  h2d(m.bx, m.nb * 8); h2d(m.bv, m.nb * 8); h2d(m.bth, m.nb * 4); h2d(m.bom, m.nb * 4);
  h2d(m.fx, m.nf * 8); h2d(m.fv, m.nf * 8); h2d(m.fguess, m.nf * 8);
  h2d(m.tbl_key, tb * 8); h2d(m.tbl_time, tb * 8);
  std::memcpy(&D->drum_angle, p, 4); p += 8;
  std::memcpy(&D->sim_time, p, 8); p += 8;
  std::memcpy(&D->e0, p, 8);
  return MG_OK;
}

int mg_sim_debug_grid(mg_sim* sim, int64_t* meta7, uint32_t* cs, size_t cs_cap, uint32_t* idx, size_t idx_cap) {
  Drum* D = drum_of(sim);
  if (!D || !meta7) return MG_E_ARG;
  hipStreamSynchronize(sim->batch->stream);
  int32_t meta[8];
  hipMemcpy(meta, D->rec.gA_meta, sizeof meta, hipMemcpyDeviceToHost);
  int64_t ncs = meta[5] ? (int64_t)meta[5] + 1 : 1;
  meta7[0] = meta[4]; meta7[1] = meta[0]; meta7[2] = meta[1]; meta7[3] = meta[2]; meta7[4] = meta[3];
  meta7[5] = ncs; meta7[6] = D->rec.nf;
  if (cs && (size_t)ncs <= cs_cap) {
    if (meta[5]) hipMemcpy(cs, D->rec.gA_start, ncs * 4, hipMemcpyDeviceToHost); else cs[0] = 0;
  }
  if (idx && (size_t)D->rec.nf <= idx_cap && D->rec.nf) hipMemcpy(idx, D->rec.gA_idx, D->rec.nf * 4, hipMemcpyDeviceToHost);
  return MG_OK;
}
int mg_sim_debug_neighbors(mg_sim* sim, uint32_t* off, size_t off_cap, uint32_t* nb, size_t nb_cap, size_t* n_edges) {
  Drum* D = drum_of(sim);
  if (!D || !n_edges) return MG_E_ARG;
  hipStreamSynchronize(sim->batch->stream);
  size_t n = (size_t)D->rec.nf;
  uint32_t tot = 0;
  if (n) hipMemcpy(&tot, D->rec.noff + n, 4, hipMemcpyDeviceToHost);
  *n_edges = tot;
  if (off && off_cap >= n + 1) hipMemcpy(off, D->rec.noff, (n + 1) * 4, hipMemcpyDeviceToHost);
  if (nb && nb_cap >= tot && tot) hipMemcpy(nb, D->rec.nbr, (size_t)tot * 4, hipMemcpyDeviceToHost);
  return MG_OK;
}
int mg_sim_force_iterations(mg_sim* sim, uint32_t density_passes, uint32_t cg_iterations) {
  Drum* D = drum_of(sim);
  if (!D) return MG_E_ARG;
  D->force_density = density_passes;
  D->force_cg = cg_iterations;
  return MG_OK;
}
uint32_t mg_sim_error_bits(mg_sim* sim) { return drum_of(sim) ? scal(sim).err : 0u; }

int mg_batch_create(const char* const* params_json, uint32_t n, mg_batch** out, char* err, size_t errlen) {
  if (!out || n == 0) return MG_E_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) { set_err(err, errlen, "no HIP device"); return MG_E_DEVICE; }
  std::vector<std::unique_ptr<Drum>> drums;
  for (uint32_t i = 0; i < n; i++) {
    Params P;
    std::string e;
    const char* j = params_json ? params_json[i] : nullptr;
    if (j && !Params::from_json(j, P, e)) { set_err(err, errlen, "drum " + std::to_string(i) + ": " + e); return MG_E_PARAMS; }
    if (!P.validate(e)) { set_err(err, errlen, "drum " + std::to_string(i) + ": " + e); return MG_E_PARAMS; }
    auto D = make_drum(P, nullptr, 0, e);
    if (!D) { set_err(err, errlen, "drum " + std::to_string(i) + ": " + e); return MG_E_DEVICE; }
    drums.push_back(std::move(D));
  }
  std::string e;
  Batch* B = make_batch(std::move(drums), e);
  if (!B) { set_err(err, errlen, e); return MG_E_DEVICE; }
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
  if (!b) return MG_E_ARG;
  return run_substeps(*b->b, substeps, nullptr, use_graph != 0);
}
mg_sim* mg_batch_drum(mg_batch* b, uint32_t i) {
  if (!b || i >= b->b->drums.size()) return nullptr;
  return b->b->views[i];
}
int mg_batch_sync(mg_batch* b, char* err, size_t errlen) {
  if (!b) return MG_E_ARG;
  std::string e;
  int rc = sync_batch(*b->b, e);
  if (rc) set_err(err, errlen, e);
  return rc;
}
float mg_batch_last_step_ms(const mg_batch* b) { return b ? b->b->last_ms : 0.0f; }

}  // extern "C"
