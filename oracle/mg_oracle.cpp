// CPU oracle: a sequential C++ restatement of MillDynamics3's per-substep slurry-plus-media step.
//
// TEST INFRASTRUCTURE ONLY. Nothing in the product links, loads or calls this file; only the
// tests, the smoke check and the CPU-baseline leg of the bench do, and only as the checker (or
// as the timed CPU reference path). It is compiled with -ffp-contract=off and without fast-math so
// that its float arithmetic rounds the way the reference's Rust f32 code does (Rust never
// contracts a*b+c into an FMA).
//
// Three arithmetic policies share this source through templates:
//   Sim<float, float>   "ref"   : restates the reference exactly (f32 state, f32 sums).
//   Sim<float, double>  "mixed" : f32 state; CG dot products / residual norms, wall-work and
//                                 energy sums accumulated in f64 (the precision form the GPU runs).
//   Sim<double, double> "fp64"  : the FP64 accuracy oracle.
//
// Every function cites the reference file:line it restates (paths relative to the reference
// checkout, crates/mill-core/src/...).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <type_traits>
#include <utility>
#include <vector>

namespace orc {

// ------------------------------------------------------------------------------------------
// Scalar helpers restating Rust f32/f64 semantics.
// ------------------------------------------------------------------------------------------
template <class R> static inline R rmax(R a, R b) { return std::fmax(a, b); }  // f32::max (maxNum)
template <class R> static inline R rmin(R a, R b) { return std::fmin(a, b); }
template <class R> static inline R rclamp(R x, R lo, R hi) {  // f32::clamp
  if (x < lo) x = lo;
  if (x > hi) x = hi;
  return x;
}
// `x as i32` (saturating, NaN -> 0).
template <class R> static inline int32_t sat_i32(R x) {
  if (std::isnan(x)) return 0;
  if (x <= (R)-2147483648.0) return INT32_MIN;
  if (x >= (R)2147483647.0) return INT32_MAX;
  return (int32_t)x;
}
// `x as u32` / `as usize` (saturating, NaN -> 0, negative -> 0).
template <class R> static inline uint64_t sat_u64(R x) {
  if (std::isnan(x) || x <= (R)0) return 0;
  if (x >= (R)1.8446744073709552e19) return UINT64_MAX;
  return (uint64_t)x;
}
template <class R> static inline uint32_t sat_u32(R x) {
  if (std::isnan(x) || x <= (R)0) return 0;
  if (x >= (R)4294967295.0) return UINT32_MAX;
  return (uint32_t)x;
}
// f32::powi lowers to binary exponentiation (compiler-rt __powisf2 / LLVM ExpandPowI):
// result = product of the squarings selected by the exponent bits, low bit first.
template <class R> static inline R powi(R a, int b) {
  R res = (R)1;
  bool first = true;
  R cur = a;
  while (b) {
    if (b & 1) {
      res = first ? cur : res * cur;
      first = false;
    }
    b >>= 1;
    if (b) cur = cur * cur;
  }
  return res;
}
template <class R> static inline R rem_euclid(R a, R b) {
  R r = std::fmod(a, b);
  return r < (R)0 ? r + std::fabs(b) : r;
}
template <class R> static inline R signum(R x) {
  if (std::isnan(x)) return x;
  return std::signbit(x) ? (R)-1 : (R)1;
}
template <class R> struct C {
  static constexpr R PI = (R)3.14159265358979323846;
  static constexpr R TAU = (R)6.28318530717958647692;
};
// Every constant written as an f32 literal in the reference is first rounded to f32, then
// promoted (identity for the float policies).
#define F(x) ((R)(float)(x))

// glam::Vec2 restated: component-wise ops in the order glam evaluates them.
template <class R> struct V2 {
  R x, y;
  V2() : x(0), y(0) {}
  V2(R a, R b) : x(a), y(b) {}
  V2 operator+(V2 o) const { return V2(x + o.x, y + o.y); }
  V2 operator-(V2 o) const { return V2(x - o.x, y - o.y); }
  V2 operator-() const { return V2(-x, -y); }
  V2 operator*(R s) const { return V2(x * s, y * s); }
  V2 operator/(R s) const { return V2(x / s, y / s); }
  V2& operator+=(V2 o) { x += o.x; y += o.y; return *this; }
  V2& operator-=(V2 o) { x -= o.x; y -= o.y; return *this; }
  V2& operator*=(R s) { x *= s; y *= s; return *this; }
  R dot(V2 o) const { return x * o.x + y * o.y; }
  R len2() const { return x * x + y * y; }
  R len() const { return std::sqrt(len2()); }
  bool finite() const { return std::isfinite(x) && std::isfinite(y); }
  V2 normalize() const { R rc = (R)1 / len(); return V2(x * rc, y * rc); }
  V2 normalize_or_zero() const {
    R rc = (R)1 / len();
    if (std::isfinite(rc) && rc > (R)0) return V2(x * rc, y * rc);
    return V2();
  }
};
template <class R> static inline V2<R> operator*(R s, V2<R> v) { return V2<R>(s * v.x, s * v.y); }
template <class R> static inline R cross2(V2<R> a, V2<R> b) { return a.x * b.y - a.y * b.x; }
template <class R> static inline V2<R> perp(V2<R> l) { return V2<R>(-l.y, l.x); }
template <class R> static inline V2<R> from_angle(R a) { return V2<R>(std::cos(a), std::sin(a)); }
template <class R> static inline V2<R> rotate(V2<R> v, V2<R> u) {  // geometry.rs:202-204
  return V2<R>(v.x * u.x - v.y * u.y, v.x * u.y + v.y * u.x);
}

// ------------------------------------------------------------------------------------------
// rng.rs:9-47 — xorshift64 (13, 7, 17) then the xorshift64* multiplier.
// ------------------------------------------------------------------------------------------
struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed == 0 ? 0x9E3779B97F4A7C15ull : seed) {}
  uint64_t next_u64() {
    uint64_t x = s;
    x ^= x << 13;
    x ^= x >> 7;
    x ^= x << 17;
    s = x;
    return x * 0x2545F4914F6CDD1Dull;
  }
  float next_f32() { return (float)(next_u64() >> 40) / (float)(1u << 24); }
  float range_f32(float lo, float hi) { return lo + next_f32() * (hi - lo); }
};

// ------------------------------------------------------------------------------------------
// params.rs — the Params fields the step reads, as f32/u32 values (serde parses into f32).
// ------------------------------------------------------------------------------------------
struct Params {
  float mill_diameter_m; int speed_mode; float speed_value; int direction;
  float ball_diameter_m, media_fill, packing, media_density, rest_bb, rest_bw, fric_bb, fric_bw,
      fric_bb_static, fric_bw_static, fric_vscale, rolling;
  int slurry_enabled; float slurry_fill, slurry_density, viscosity; int rheology;
  float yield_stress, wall_no_slip, ball_no_slip, wettability, surface_tension; int dye_pattern;
  uint32_t lifter_count; float lifter_height, lifter_base, lifter_top, lifter_phase_deg;
  uint32_t resolution, substeps, pbf_iterations, dem_iterations; float time_scale, frame_budget_ms;
  uint32_t max_balls; uint64_t seed; int cg_mode; float cg_k;

  // params.rs:51-69
  float critical_speed_rpm() const { return 42.3f / std::sqrt(mill_diameter_m); }
  float rpm() const { return speed_mode == 0 ? speed_value : critical_speed_rpm() * speed_value / 100.0f; }
  float omega_magnitude() const { return rpm() * 2.0f * C<float>::PI / 60.0f; }
  float omega() const { return direction == 0 ? omega_magnitude() : -omega_magnitude(); }
  float radius_m() const { return mill_diameter_m * 0.5f; }
  // params.rs:558-566
  float true_ball_count() const {
    float r_true = ball_diameter_m * 0.5f;
    if (r_true <= 0.0f) return 0.0f;
    float drum_area = C<float>::PI * powi(radius_m(), 2);
    return (media_fill * packing * drum_area) / (C<float>::PI * r_true * r_true);
  }
  // params.rs:584-620
  void effective_media(float& true_d, float& d, float& density, uint32_t& count, float& sf) const {
    true_d = ball_diameter_m;
    density = media_density;
    float n_true = true_ball_count();
    float mb = (float)max_balls;
    float scale = 1.0f;
    if (cg_mode == 2) scale = rmax(cg_k, 1.0f);
    else if (cg_mode == 0) scale = (n_true > mb && mb > 0.0f) ? std::sqrt(n_true / mb) : 1.0f;
    if (scale > 1.0f) {
      count = sat_u32(rmax(std::round(n_true / (scale * scale)), 1.0f));
      d = true_d * scale;
      sf = scale;
    } else {
      count = sat_u32(rmax(std::round(n_true), 0.0f));
      d = true_d;
      sf = 1.0f;
    }
  }
  float eff_diameter() const { float a, d, b, s; uint32_t c; effective_media(a, d, b, c, s); return d; }
  // params.rs:650-658
  uint32_t effective_fluid_resolution() const {
    float r = radius_m(), d = eff_diameter();
    if (!(r > 0.0f && d > 0.0f)) return resolution;
    uint32_t mn = sat_u32(std::ceil(r / d));
    return std::min(std::max(resolution, mn), 200u);
  }
  // params.rs:705-726
  uint32_t effective_substeps() const {
    float d = eff_diameter();
    if (!(d > 0.0f && mill_diameter_m > 0.0f)) return substeps;
    float df = rclamp(1.0f + 2.5f * (media_fill - 0.30f), 1.0f, 2.0f);
    float v_ref = std::sqrt(2.0f * 9.81f * mill_diameter_m);
    uint32_t mn = sat_u32(rmax(std::ceil(v_ref * df / (60.0f * 0.5f * d)), 1.0f));
    uint32_t cap = sat_u32(std::ceil(16.0f * df));
    return std::max(substeps, std::min(mn, cap));
  }
};

// ------------------------------------------------------------------------------------------
// geometry.rs:11-275 — drum SDF with optional trapezoid lifters.
// ------------------------------------------------------------------------------------------
template <class R> struct Drum {
  R radius, omega;
  uint32_t lifter_count;
  R l_height, l_base, l_top, l_phase_deg;

  // geometry.rs:233-262
  static R convex_polygon_sdf(V2<R> p, const V2<R>* v, int n) {
    R min_d2 = std::numeric_limits<R>::max();
    bool inside = true;
    for (int i = 0; i < n; i++) {
      V2<R> a = v[i], b = v[(i + 1) % n];
      V2<R> edge = b - a, to_p = p - a;
      R ls = edge.len2();
      if (ls < F(1e-12)) continue;
      R t = rclamp(to_p.dot(edge) / ls, (R)0, (R)1);
      V2<R> closest = a + edge * t;
      min_d2 = rmin(min_d2, (p - closest).len2());
      R cr = edge.x * to_p.y - edge.y * to_p.x;
      if (cr < (R)0) inside = false;
    }
    R d = std::sqrt(min_d2);
    return inside ? d : -d;
  }
  // geometry.rs:37-72
  R sdf(V2<R> p) const {
    R dist = radius - p.len();
    if (lifter_count == 0) return dist;
    R phase = l_phase_deg * F(0.017453292519943295);  // f32::to_radians
    R max_inside = -std::numeric_limits<R>::infinity();
    for (uint32_t k = 0; k < lifter_count; k++) {
      R th = phase + (R)k * C<R>::TAU / (R)lifter_count;
      R s = std::sin(th), c = std::cos(th);
      R pr = p.x * c + p.y * s;
      R pt = -p.x * s + p.y * c;
      R rt = radius - l_height;
      R bh = l_base * F(0.5), thw = l_top * F(0.5);
      V2<R> verts[4] = {V2<R>(radius, -bh), V2<R>(radius, bh), V2<R>(rt, thw), V2<R>(rt, -thw)};
      max_inside = rmax(max_inside, convex_polygon_sdf(V2<R>(pr, pt), verts, 4));
    }
    return rmin(dist, -max_inside);
  }
  // geometry.rs:87-108
  std::pair<R, V2<R>> sdf_world(V2<R> pw, R angle) const {
    if (!pw.finite()) return {(R)0, V2<R>()};
    V2<R> pl = rotate(pw, from_angle(-angle));
    const R EPS = F(1e-4);
    R d = sdf(pl);
    R dx = sdf(pl + V2<R>(EPS, 0)) - sdf(pl - V2<R>(EPS, 0));
    R dy = sdf(pl + V2<R>(0, EPS)) - sdf(pl - V2<R>(0, EPS));
    V2<R> g = V2<R>(dx, dy) / (F(2.0) * EPS);
    V2<R> nl = g.len2() > F(1e-12) ? g.normalize() : -pl.normalize_or_zero();
    return {d, rotate(nl, from_angle(angle))};
  }
  // geometry.rs:112-114
  V2<R> wall_velocity(V2<R> p) const { return omega * V2<R>(-p.y, p.x); }
  // geometry.rs:146-194
  bool toi_swept(V2<R> x0, V2<R> x1, R a0, R a1, R radius_b, R& out) const {
    V2<R> l0 = rotate(x0, from_angle(-a0));
    if (sdf(l0) - radius_b <= (R)0) return false;
    V2<R> path = x1 - x0;
    R pl = path.len();
    R da = a1 - a0;
    R pos_bound = x0.len() + pl;
    R lsb = pl + std::fabs(da) * pos_bound;
    if (lsb < F(1e-9)) return false;
    R t = 0;
    for (int it = 0; it < 32; it++) {
      R at = a0 + da * t;
      V2<R> xt = x0 + path * t;
      V2<R> loc = rotate(xt, from_angle(-at));
      R d = sdf(loc) - radius_b;
      if (d <= (R)0) { out = t; return true; }
      R st = rmax(d / lsb, F(1e-6));
      t += st;
      if (t >= (R)1) return false;
    }
    out = rmin(t, (R)1);
    return true;
  }
};

// ------------------------------------------------------------------------------------------
// grid.rs:101-399 — counting-sort CSR uniform grid, Dense or Sparse layout.
// ------------------------------------------------------------------------------------------
template <class R> struct Grid {
  R cell = 1;
  bool dense = false;
  int32_t min_cx = 0, min_cy = 0;
  int64_t W = 0, H = 0;
  std::vector<uint32_t> cell_start;
  std::vector<std::pair<int32_t, int32_t>> keys;
  std::vector<uint32_t> indices;

  static std::pair<int32_t, int32_t> key(V2<R> p, R c) {  // grid.rs:280-285
    return {sat_i32(std::floor(p.x / c)), sat_i32(std::floor(p.y / c))};
  }
  void build(const std::vector<V2<R>>& pts, R c) {  // grid.rs:101-153
    cell = c;
    std::vector<std::pair<std::pair<int32_t, int32_t>, uint32_t>> keyed;
    keyed.reserve(pts.size());
    int32_t mnx = INT32_MAX, mny = INT32_MAX, mxx = INT32_MIN, mxy = INT32_MIN;
    for (size_t i = 0; i < pts.size(); i++) {
      if (!pts[i].finite()) continue;
      auto k = key(pts[i], c);
      mnx = std::min(mnx, k.first); mny = std::min(mny, k.second);
      mxx = std::max(mxx, k.first); mxy = std::max(mxy, k.second);
      keyed.push_back({k, (uint32_t)i});
    }
    cell_start.clear(); keys.clear(); indices.clear();
    if (keyed.empty()) { dense = false; cell_start.push_back(0); return; }
    W = (int64_t)mxx - (int64_t)mnx + 1;
    H = (int64_t)mxy - (int64_t)mny + 1;
    unsigned __int128 tc = (unsigned __int128)(uint64_t)W * (uint64_t)H;
    uint64_t total = tc > UINT64_MAX ? UINT64_MAX : (uint64_t)tc;
    uint64_t budget = 2ull * keyed.size() + 512ull;
    if (total <= budget) {  // grid.rs:158-197
      dense = true; min_cx = mnx; min_cy = mny;
      cell_start.assign(total + 1, 0);
      for (auto& kv : keyed) cell_start[dense_index(kv.first) + 1]++;
      for (size_t i = 1; i < cell_start.size(); i++) cell_start[i] += cell_start[i - 1];
      std::vector<uint32_t> cur(cell_start);
      indices.assign(keyed.size(), 0);
      for (auto& kv : keyed) indices[cur[dense_index(kv.first)]++] = kv.second;
    } else {  // grid.rs:204-226
      dense = false;
      std::stable_sort(keyed.begin(), keyed.end(),
                       [](const auto& a, const auto& b) { return a.first < b.first; });
      cell_start.push_back(0);
      size_t i = 0;
      while (i < keyed.size()) {
        auto k = keyed[i].first;
        size_t j = i;
        while (j < keyed.size() && keyed[j].first == k) indices.push_back(keyed[j++].second);
        keys.push_back(k);
        cell_start.push_back((uint32_t)j);
        i = j;
      }
    }
  }
  size_t dense_index(std::pair<int32_t, int32_t> k) const {  // grid.rs:232-236
    return (size_t)(((int64_t)k.first - min_cx) * H + ((int64_t)k.second - min_cy));
  }
  bool range(std::pair<int32_t, int32_t> k, size_t& s, size_t& e) const {  // grid.rs:242-271
    if (dense) {
      int64_t dx = (int64_t)k.first - min_cx, dy = (int64_t)k.second - min_cy;
      if (dx < 0 || dy < 0 || dx >= W || dy >= H) return false;
      size_t kk = (size_t)(dx * H + dy);
      s = cell_start[kk]; e = cell_start[kk + 1];
      return s != e;
    }
    auto it = std::lower_bound(keys.begin(), keys.end(), k);
    if (it == keys.end() || *it != k) return false;
    size_t idx = it - keys.begin();
    s = cell_start[idx]; e = cell_start[idx + 1];
    return true;
  }
  static int32_t sadd(int32_t a, int32_t b) {
    int64_t r = (int64_t)a + b;
    return r > INT32_MAX ? INT32_MAX : (r < INT32_MIN ? INT32_MIN : (int32_t)r);
  }
  template <class Fn> void visit_cell(int32_t cx, int32_t cy, size_t s, size_t e, Fn& f) const {
    static const int32_t OFF[4][2] = {{1, 0}, {1, 1}, {0, 1}, {-1, 1}};
    for (size_t a = s; a < e; a++)
      for (size_t b = a + 1; b < e; b++) {
        uint32_t i = indices[a], j = indices[b];
        f(std::min(i, j), std::max(i, j));
      }
    for (auto& o : OFF) {
      std::pair<int32_t, int32_t> k = {sadd(cx, o[0]), sadd(cy, o[1])};
      if (k.first == cx && k.second == cy) continue;
      size_t os, oe;
      if (!range(k, os, oe)) continue;
      for (size_t a = s; a < e; a++)
        for (size_t b = os; b < oe; b++) {
          uint32_t i = indices[a], j = indices[b];
          f(std::min(i, j), std::max(i, j));
        }
    }
  }
  template <class Fn> void candidate_pairs(Fn f) const {  // grid.rs:291-362
    if (dense) {
      size_t total = (size_t)(W * H);
      for (size_t k = 0; k < total; k++) {
        size_t s = cell_start[k], e = cell_start[k + 1];
        if (s == e) continue;
        int32_t cx = min_cx + (int32_t)((int64_t)k / H), cy = min_cy + (int32_t)((int64_t)k % H);
        visit_cell(cx, cy, s, e, f);
      }
    } else {
      for (size_t i = 0; i < keys.size(); i++) visit_cell(keys[i].first, keys[i].second, cell_start[i], cell_start[i + 1], f);
    }
  }
  template <class Fn> void near(V2<R> p, Fn f) const {  // grid.rs:371-399
    if (!p.finite()) return;
    auto c = key(p, cell);
    std::pair<int32_t, int32_t> ks[9];
    int nk = 0;
    for (int dy = -1; dy <= 1; dy++)
      for (int dx = -1; dx <= 1; dx++) {
        std::pair<int32_t, int32_t> k = {sadd(c.first, dx), sadd(c.second, dy)};
        bool dup = false;
        for (int q = 0; q < nk; q++) dup |= ks[q] == k;
        if (!dup) ks[nk++] = k;
      }
    for (int q = 0; q < nk; q++) {
      size_t s, e;
      if (!range(ks[q], s, e)) continue;
      for (size_t a = s; a < e; a++) f(indices[a]);
    }
  }
};

// pbf.rs:235-288 — mutual CSR neighbour lists, stable by pair-emission order.
struct Nbrs {
  std::vector<uint32_t> off, nb;
};
template <class R> static Nbrs build_neighbor_lists(const std::vector<V2<R>>& x, const Grid<R>& g, R h) {
  size_t n = x.size();
  std::vector<std::pair<uint32_t, uint32_t>> edges;
  g.candidate_pairs([&](uint32_t i, uint32_t j) {
    if ((x[i] - x[j]).len2() <= h * h) { edges.push_back({i, j}); edges.push_back({j, i}); }
  });
  Nbrs L;
  L.off.assign(n + 1, 0);
  for (auto& e : edges) L.off[e.first + 1]++;
  for (size_t k = 1; k < L.off.size(); k++) L.off[k] += L.off[k - 1];
  std::vector<uint32_t> cur(L.off);
  L.nb.assign(edges.size(), 0);
  for (auto& e : edges) L.nb[cur[e.first]++] = e.second;
  return L;
}

// metrics::mixing_index (metrics.rs:506-576): Lacey index of the dye field on a 16 x 16 grid over
// [-radius, radius]^2, count-weighted; f32 sums in particle order as the reference forms them.
template <class R, class D>
static bool mixing_index_of(const V2<R>* x, const D* dye, size_t n, R rad, float& out) {
  if (n == 0 || rad <= (R)0) return false;
  const int G = 16;
  R cell = F(2.0) * rad / (R)G;
  std::vector<R> ds_(G * G, 0);
  std::vector<uint32_t> cnt(G * G, 0);
  R tot = 0;
  for (size_t i = 0; i < n; i++) {
    int gi = std::min(std::max(sat_i32(std::floor((x[i].x + rad) / cell)), 0), G - 1);
    int gj = std::min(std::max(sat_i32(std::floor((x[i].y + rad) / cell)), 0), G - 1);
    size_t idx = (size_t)gj * G + gi;
    ds_[idx] += (R)dye[i];
    cnt[idx]++;
    tot += (R)dye[i];
  }
  R pm = tot / (R)n;
  std::vector<R> means;
  std::vector<uint32_t> counts;
  for (size_t q = 0; q < ds_.size(); q++)
    if (cnt[q] > 0) { means.push_back(ds_[q] / (R)cnt[q]); counts.push_back(cnt[q]); }
  if (means.size() < 2) return false;
  uint32_t wsum_u = 0;
  for (auto c : counts) wsum_u += c;
  R ws = (R)wsum_u;
  R ov = 0;
  for (size_t q = 0; q < means.size(); q++) ov += (R)counts[q] * powi(means[q] - pm, 2);
  ov = ov / ws;
  R s0 = pm * (F(1.0) - pm);
  R wa = 0;
  for (auto c : counts) wa += (R)c * (R)c;
  wa = wa / ws;
  R sr = s0 / rmax(wa, (R)1);
  if (s0 < F(1e-9)) return false;
  R den = s0 - sr;
  if (std::fabs(den) < F(1e-9)) { out = 1.0f; return true; }
  R m = (s0 - ov) / den;
  out = (float)rclamp(m, (R)0, (R)1);
  return true;
}

// ------------------------------------------------------------------------------------------
// surface.rs:39-420 — free-surface extraction (f32): splat field, wall mask, marching squares,
// chains, one Chaikin pass, projection inside the wall. glam's Vec2::lerp is restated as
// a + (b - a) * t (glam 0.29 is not vendored here: the last bit of contour points is unpinned).
// ------------------------------------------------------------------------------------------
namespace surf {
constexpr int G = 128;
inline float splat_kernel(float r2, float radius) {  // surface.rs:39-46
  const float radius2 = radius * radius;
  if (r2 >= radius2) return 0.0f;
  const float t = 1.0f - r2 / radius2;
  return t * t * t;
}
inline size_t grid_index(float world, float R, float cell) {  // surface.rs:140-143
  const float raw = rclamp(std::round((world + R) / cell), 0.0f, (float)(G - 1));
  return std::isnan(raw) ? 0 : (size_t)raw;
}
inline V2<float> grid_point(size_t gi, size_t gj, float R, float cell) {  // surface.rs:145-150
  return V2<float>(-R + (float)gi * cell, -R + (float)gj * cell);
}
inline V2<float> lerp(V2<float> a, V2<float> b, float t) { return a + (b - a) * t; }
// surface.rs:84-133
static std::vector<float> build_field(const std::vector<V2<float>>& x, const Drum<float>& drum, float angle, float rs) {
  const float R = drum.radius;
  std::vector<float> field((size_t)G * G, 0.0f);
  if (x.empty() || R <= 0.0f) return field;
  const float cell = 2.0f * R / ((float)G - 1.0f);
  for (const auto& p : x) {
    const size_t i0 = grid_index(p.x - rs, R, cell), i1 = grid_index(p.x + rs, R, cell);
    const size_t j0 = grid_index(p.y - rs, R, cell), j1 = grid_index(p.y + rs, R, cell);
    for (size_t gj = j0; gj <= j1; gj++)
      for (size_t gi = i0; gi <= i1; gi++) field[gj * G + gi] += splat_kernel((grid_point(gi, gj, R, cell) - p).len2(), rs);
  }
  const V2<float> to_local = from_angle(-angle);
  for (size_t gj = 0; gj < (size_t)G; gj++)
    for (size_t gi = 0; gi < (size_t)G; gi++) {
      float& f = field[gj * G + gi];
      if (f == 0.0f) continue;
      if (drum.sdf(rotate(grid_point(gi, gj, R, cell), to_local)) < 0.0f) f = 0.0f;
    }
  return field;
}
struct Edge {
  int vert, i, j;  // 0: horizontal (i, j)-(i+1, j); 1: vertical (i, j)-(i, j+1)
  bool operator==(const Edge& o) const { return vert == o.vert && i == o.i && j == o.j; }
  bool operator<(const Edge& o) const { return vert != o.vert ? vert < o.vert : (i != o.i ? i < o.i : j < o.j); }
};
inline Edge cell_edge(int ci, int cj, int side) {  // surface.rs:209-217
  switch (side) {
    case 0: return {0, ci, cj};
    case 1: return {1, ci + 1, cj};
    case 2: return {0, ci, cj + 1};
    default: return {1, ci, cj};
  }
}
inline int edge_pairs(int c, int (*out)[2]) {  // surface.rs:194-207
  switch (c) {
    case 1: case 14: out[0][0] = 3; out[0][1] = 0; return 1;
    case 2: case 13: out[0][0] = 0; out[0][1] = 1; return 1;
    case 3: case 12: out[0][0] = 3; out[0][1] = 1; return 1;
    case 4: case 11: out[0][0] = 1; out[0][1] = 2; return 1;
    case 6: case 9: out[0][0] = 0; out[0][1] = 2; return 1;
    case 7: case 8: out[0][0] = 3; out[0][1] = 2; return 1;
    case 5: out[0][0] = 3; out[0][1] = 0; out[1][0] = 1; out[1][1] = 2; return 2;
    case 10: out[0][0] = 0; out[0][1] = 1; out[1][0] = 3; out[1][1] = 2; return 2;
    default: return 0;
  }
}
// surface.rs:279-357
static std::vector<std::vector<V2<float>>> polygons_at(const std::vector<float>& field, float thr, float R, float cell,
                                                      const Drum<float>& drum, float angle) {
  std::vector<std::pair<Edge, Edge>> seg;
  for (int cj = 0; cj < G - 1; cj++)
    for (int ci = 0; ci < G - 1; ci++) {
      auto corner = [&](int di, int dj) { return field[(size_t)(cj + dj) * G + (ci + di)] > thr; };
      const int c = (int)corner(0, 0) | (int)corner(1, 0) << 1 | (int)corner(1, 1) << 2 | (int)corner(0, 1) << 3;
      int pr[2][2];
      const int k = edge_pairs(c, pr);
      for (int q = 0; q < k; q++) seg.push_back({cell_edge(ci, cj, pr[q][0]), cell_edge(ci, cj, pr[q][1])});
    }
  std::vector<std::vector<V2<float>>> polys;
  if (seg.empty()) return polys;
  std::map<Edge, std::vector<size_t>> by_edge;
  for (size_t k = 0; k < seg.size(); k++) {
    by_edge[seg[k].first].push_back(k);
    by_edge[seg[k].second].push_back(k);
  }
  std::vector<bool> visited(seg.size(), false);
  auto extend = [&](std::vector<Edge>& chain, bool fwd) {  // surface.rs:371-401
    for (;;) {
      const Edge tip = fwd ? chain.back() : chain.front();
      auto it = by_edge.find(tip);
      if (it == by_edge.end()) return;
      size_t next = SIZE_MAX;
      for (size_t k : it->second)
        if (!visited[k]) { next = k; break; }
      if (next == SIZE_MAX) return;
      visited[next] = true;
      const Edge other = seg[next].first == tip ? seg[next].second : seg[next].first;
      if (fwd) chain.push_back(other); else chain.insert(chain.begin(), other);
    }
  };
  for (size_t st = 0; st < seg.size(); st++) {
    if (visited[st]) continue;
    visited[st] = true;
    std::vector<Edge> chain = {seg[st].first, seg[st].second};
    extend(chain, true);
    extend(chain, false);
    if (chain.size() > 1 && chain.front() == chain.back()) chain.pop_back();
    std::vector<V2<float>> pts;
    for (const Edge& e : chain) {  // interp_edge (surface.rs:166-192)
      const size_t i0 = e.i, j0 = e.j, i1 = e.vert ? e.i : e.i + 1, j1 = e.vert ? e.j + 1 : e.j;
      const float v0 = field[j0 * G + i0], v1 = field[j1 * G + i1];
      const float t = std::fabs(v1 - v0) > 1e-9f ? rclamp((thr - v0) / (v1 - v0), 0.0f, 1.0f) : 0.5f;
      pts.push_back(lerp(grid_point(i0, j0, R, cell), grid_point(i1, j1, R, cell), t));
    }
    std::vector<V2<float>> sm;  // chaikin_smooth_closed (surface.rs:403-418)
    if (pts.size() < 3) {
      sm = pts;
    } else {
      for (size_t i = 0; i < pts.size(); i++) {
        const V2<float> a = pts[i], b = pts[(i + 1) % pts.size()];
        sm.push_back(lerp(a, b, 0.25f));
        sm.push_back(lerp(a, b, 0.75f));
      }
    }
    for (auto& p : sm) {  // project_inside_wall (surface.rs:359-366)
      const auto dn = drum.sdf_world(p, angle);
      if (dn.first < 0.0f) p = p - dn.second * dn.first;
    }
    polys.push_back(sm);
  }
  return polys;
}
// surface.rs:233-277
static std::vector<std::vector<V2<float>>> extract(const std::vector<V2<float>>& x, const Drum<float>& drum, float angle,
                                                  float h, float rest, float pmass, std::vector<float>* field_out) {
  const float R = drum.radius;
  if (x.empty() || R <= 0.0f) return {};
  const float rs = 1.5f * h;
  const std::vector<float> field = build_field(x, drum, angle, rs);
  if (field_out) *field_out = field;
  const float cell = 2.0f * R / ((float)G - 1.0f);
  const float nd = rest / rmax(pmass, 1e-12f);
  const float phi = nd * (C<float>::PI * rs * rs / 4.0f);
  auto polys = polygons_at(field, 0.5f * phi, R, cell, drum, angle);
  if (!polys.empty()) return polys;
  float peak = 0.0f;
  for (float f : field) peak = rmax(peak, f);
  if (peak <= 0.0f) return {};
  return polygons_at(field, 0.5f * peak, R, cell, drum, angle);
}
}  // namespace surf

// ------------------------------------------------------------------------------------------
// metrics.rs:195-710 — metrics::compute over a drum state (f32 throughout, as the reference).
// Angles are p.y.atan2(p.x) from the host libm, as the reference's f32::atan2.
// ------------------------------------------------------------------------------------------
struct DrumMetrics {
  float toe = 0, shoulder = 0;
  bool has_toe_shoulder = false;
  float cx = 0, cy = 0;
  bool has_centroid = false;
  float pool_lo = 0, pool_hi = 0;
  bool has_pool = false;
  float fs_angle = 0, fs_offset = 0;
  bool has_fs = false;
  float depth = 0;
  bool has_depth = false;
  float mixing = 0;
  bool has_mixing = false;
  float ke = 0, max_overlap = 0, max_wall_overlap = 0;
  float max_derr = 0, mean_derr = 0, max_cerr = 0, mean_cerr = 0;
  bool has_errors = false;
  float omega = 0;
  uint32_t n_fluid = 0;
  bool non_finite_angle = false;  // the reference's sort_by(partial_cmp().unwrap()) panics here
};

// The world angle of p in [0, 2 pi): p.y.atan2(p.x).rem_euclid(TAU) (metrics.rs:341).
static inline float wall_angle(V2<float> p) { return rem_euclid(std::atan2(p.y, p.x), C<float>::TAU); }

// metrics.rs:293-308: bounds of the cluster across the largest angular gap of sorted angles
// (n >= 2): (cluster_start, cluster_end, max_gap).
static void largest_gap_cluster_bounds(const std::vector<float>& a, float& start, float& end, float& gap) {
  const size_t n = a.size();
  float max_gap = a[0] + C<float>::TAU - a[n - 1];
  size_t k = n - 1;
  for (size_t i = 0; i + 1 < n; i++) {
    const float g = a[i + 1] - a[i];
    if (g > max_gap) {
      max_gap = g;
      k = i;
    }
  }
  end = a[k];
  start = a[(k + 1) % n];
  gap = max_gap;
}

// Sorts finite angles ascending; false (the reference would panic) when one is NaN.
static bool sort_angles(std::vector<float>& a) {
  for (float v : a)
    if (std::isnan(v)) return false;
  std::sort(a.begin(), a.end());
  return true;
}

static DrumMetrics drum_metrics(const Drum<float>& drum, float angle, const V2<float>* bx, const V2<float>* bv,
                                const float* bw, size_t nb, float br, float bmass, float binertia, const V2<float>* fx,
                                const float* dye, const float* dens, size_t nf, float h, float rest) {
  DrumMetrics m;
  const float TAU = C<float>::TAU, PI = C<float>::PI;
  m.omega = drum.omega;
  m.n_fluid = (uint32_t)nf;
  // charge_toe_shoulder (metrics.rs:327-357)
  if (nb > 0) {
    const float margin = 2.5f * br;
    std::vector<float> ang;
    for (size_t i = 0; i < nb; i++)
      if (drum.sdf_world(bx[i], angle).first < margin) ang.push_back(wall_angle(bx[i]));
    if (ang.size() >= 6) {
      if (!sort_angles(ang)) {
        m.non_finite_angle = true;
      } else {
        float st, en, gap;
        largest_gap_cluster_bounds(ang, st, en, gap);
        if (gap > PI * 0.3f) {
          m.has_toe_shoulder = true;
          m.toe = drum.omega >= 0.0f ? st : en;
          m.shoulder = drum.omega >= 0.0f ? en : st;
        }
      }
    }
  }
  // charge_centroid (metrics.rs:362-369)
  if (nb > 0) {
    V2<float> sum;
    for (size_t i = 0; i < nb; i++) sum = sum + bx[i];
    m.has_centroid = true;
    m.cx = sum.x / (float)nb;
    m.cy = sum.y / (float)nb;
  }
  // slurry_pool_angular_extent (metrics.rs:377-401)
  if (nf > 0) {
    const float margin = 1.5f * h;
    std::vector<float> ang;
    for (size_t i = 0; i < nf; i++)
      if (drum.sdf_world(fx[i], angle).first < margin) ang.push_back(wall_angle(fx[i]));
    if (ang.size() == 1) {  // returned as is, without the sort
      m.has_pool = true;
      m.pool_lo = m.pool_hi = ang[0];
    } else if (ang.size() > 1) {
      if (!sort_angles(ang)) {
        m.non_finite_angle = true;
      } else {
        float st, en, gap;
        largest_gap_cluster_bounds(ang, st, en, gap);
        m.has_pool = true;
        m.pool_lo = st;
        m.pool_hi = en;
      }
    }
  }
  // fluid_free_surface_line (metrics.rs:414-453)
  if (nf > 0) {
    const float margin = 1.5f * h;
    std::vector<V2<float>> pts;
    for (size_t i = 0; i < nf; i++)
      if (drum.sdf_world(fx[i], angle).first >= margin) pts.push_back(fx[i]);
    if (pts.size() >= 3) {
      const float n = (float)pts.size();
      V2<float> sum;
      for (auto& p : pts) sum = sum + p;
      const V2<float> c(sum.x / n, sum.y / n);
      float cxx = 0, cyy = 0, cxy = 0;
      for (auto& p : pts) {
        const V2<float> d = p - c;
        cxx += d.x * d.x;
        cyy += d.y * d.y;
        cxy += d.x * d.y;
      }
      cxx /= n;
      cyy /= n;
      cxy /= n;
      const float a = 0.5f * std::atan2(2.0f * cxy, cxx - cyy);
      const float sa = std::sin(a), ca = std::cos(a);
      m.has_fs = true;
      m.fs_angle = a;
      m.fs_offset = c.x * -sa + c.y * ca;
    }
  }
  // slurry_pool_depth_at_bottom (metrics.rs:463-476); angular_separation (metrics.rs:282-285)
  if (nf > 0) {
    const float down = 3.0f * (PI * 0.5f);
    const float band = 5.0f * PI / 180.0f;
    bool any = false;
    float mn = 0;
    for (size_t i = 0; i < nf; i++) {
      const float a = wall_angle(fx[i]);
      const float d = rem_euclid(a - down, TAU);
      if (rmin(d, TAU - d) <= band) {
        const float dist = fx[i].len();
        mn = any ? rmin(mn, dist) : dist;
        any = true;
      }
    }
    if (any) {
      m.has_depth = true;
      m.depth = drum.radius - mn;
    }
  }
  // mixing_index (metrics.rs:506-576)
  m.has_mixing = mixing_index_of<float>(fx, dye, nf, drum.radius, m.mixing);
  // total_kinetic_energy_j (metrics.rs:581-590)
  for (size_t i = 0; i < nb; i++) m.ke += 0.5f * bmass * bv[i].len2() + 0.5f * binertia * bw[i] * bw[i];
  // max_ball_overlap_fraction (metrics.rs:599-613), over the grid's candidate pairs
  if (nb >= 2 && br > 0.0f) {
    std::vector<V2<float>> pts(bx, bx + nb);
    Grid<float> g;
    g.build(pts, rmax(2.0f * br * 1.05f, 1e-6f));
    g.candidate_pairs([&](uint32_t i, uint32_t j) {
      const float dist = (bx[i] - bx[j]).len();
      const float ov = rmax(2.0f * br - dist, 0.0f);
      m.max_overlap = rmax(m.max_overlap, ov / br);
    });
  }
  // max_ball_wall_overlap_fraction (metrics.rs:621-634)
  if (nb > 0 && br > 0.0f)
    for (size_t i = 0; i < nb; i++)
      m.max_wall_overlap = rmax(m.max_wall_overlap, rmax(br - drum.sdf_world(bx[i], angle).first, 0.0f) / br);
  // density and compression errors (metrics.rs:649-710)
  if (nf > 0 && rest > 0.0f) {
    m.has_errors = true;
    float se = 0, sc = 0;
    for (size_t i = 0; i < nf; i++) {
      const float e = std::fabs(dens[i] - rest) / rest;
      const float c = rmax(dens[i] / rest - 1.0f, 0.0f);
      m.max_derr = rmax(m.max_derr, e);
      m.max_cerr = rmax(m.max_cerr, c);
      se += e;
      sc += c;
    }
    m.mean_derr = se / (float)nf;
    m.mean_cerr = sc / (float)nf;
  }
  return m;
}

// ------------------------------------------------------------------------------------------
// pbf.rs:119-219 — kernels and boundary sampling.
// ------------------------------------------------------------------------------------------
template <class R> static R poly6(R r2, R h) {
  if (r2 >= h * h || r2 < (R)0) return 0;
  R h2 = h * h, t = h2 - r2;
  return F(4.0) / (C<R>::PI * powi(h, 8)) * t * t * t;
}
template <class R> static V2<R> spiky_grad(V2<R> d, R r, R h) {
  if (r <= F(1e-9) || r >= h) return V2<R>();
  R coeff = F(-30.0) / (C<R>::PI * powi(h, 5)) * (h - r) * (h - r);
  return (d / r) * coeff;
}
template <class R> static R cohesion_kernel(R r, R h) {
  if (r <= (R)0 || r >= h) return 0;
  R coeff = F(35840.0) / (F(209.0) * C<R>::PI * powi(h, 8));
  R term = powi(h - r, 3) * powi(r, 3);
  if (r <= F(0.5) * h) return coeff * (F(2.0) * term - powi(h, 6) / F(64.0));
  return coeff * term;
}
template <class R> static R adhesion_kernel(R r, R h) {
  if (r <= F(0.5) * h || r >= h) return 0;
  R coeff = F(35840.0) / (F(163.0) * C<R>::PI * powi(h, 8));
  return coeff * powi(h - r, 3) * powi(r, 3);
}
template <class R> static uint32_t boundary_points_per_ball(R br, R dx) {
  R circ = C<R>::TAU * br;
  return std::max(sat_u32(std::round(circ / dx)), 6u);
}

// lubrication.rs:44-115 (f64 squeeze film)
struct SqueezeFilm {
  double r, c1, c2, pre;
  SqueezeFilm(double k1, double k2, double nu) {
    r = 1.0 / (k1 + k2);
    double q = powi(r, 3) * (powi(k1, 3) + powi(k2, 3));
    c1 = 0.788 + 0.265 * q;
    c2 = std::fmax(0.006 + 0.057 * (q - 0.25), 0.0);
    pre = 3.0 * std::sqrt(2.0) * M_PI * nu * std::pow(r, 1.5);
  }
  double coefficient(double h) const {
    double eps = h / r;
    return pre / (h * std::sqrt(h)) * (1.0 + c1 * eps + c2 * eps * eps);
  }
};
static double resolved_ratio(double hc) {
  static const double T[6][2] = {{0, 0}, {0.5, 0.06}, {1, 0.25}, {2, 0.9}, {4, 0.99}, {8, 1}};
  if (hc >= 8.0) return 1.0;
  double h = std::fmax(hc, 0.0);
  for (int i = 0; i < 5; i++)
    if (h <= T[i + 1][0]) {
      double t = (h - T[i][0]) / (T[i + 1][0] - T[i][0]);
      return T[i][1] + t * (T[i + 1][1] - T[i][1]);
    }
  return 1.0;
}

// ------------------------------------------------------------------------------------------
// State types (dem.rs:277-309, pbf.rs:1018-1028, coupling.rs:36-52).
// ------------------------------------------------------------------------------------------
template <class R> struct Balls {
  std::vector<V2<R>> x, v;
  std::vector<R> theta, omega;
  R radius = 0, mass = 0, inertia = 0, true_radius = 0;
  size_t n() const { return x.size(); }
};
template <class R> struct Fluid {
  std::vector<V2<R>> x, v, guess;
  std::vector<float> dye;
  R mass = 0, h = 0, rest = 0;
  size_t n() const { return x.size(); }
};
template <class R> struct Coupling {
  std::vector<V2<R>> imp;
  std::vector<R> ang;
  uint32_t clamp_hits = 0;
  V2<R> fmc;
};
struct FluidStats {
  uint32_t visc_iters = 0, density_iters = 0;
  double mean_shear = 0, wall_work = 0;
  double wall_imp[2] = {0, 0}, wall_ang = 0;
  float bins[32] = {0};
  double ke_noslip = 0, ke_drag = 0, ke_buoy = 0, ke_visc = 0, ke_clamp = 0;
  uint32_t clamp_hits = 0, n_groups = 0, n_links = 0, links_in_range = 0;
};
struct DemStats {
  double wall_work = 0, dissipated = 0, coupling_work = 0;
  uint32_t collisions = 0;
  uint32_t hist[12] = {0};
  float max_disp = 0;
  uint32_t n_pairs = 0;
};
struct Impact {
  int kind;  // 0 ball-ball, 1 ball-wall
  uint32_t ball, partner;
  float vn, slip, e;
};

// stress.rs:23-238
struct Stress {
  double count_rate[4] = {0}, se_rate[4] = {0}, hist_rate[28] = {0}, above_rate[6] = {0};
  double cum_wall = 0, cum_above[6] = {0};
  std::map<std::pair<uint32_t, uint32_t>, double> last;
  double factor = 0.25;
  uint64_t updates = 0;
  static constexpr double THR[6] = {0.03, 0.1, 0.3, 1.0, 3.0, 10.0};
  static size_t se_bin(double e) {
    if (std::isnan(e) || e <= 0.0) return 0;
    double t = (std::log10(e) - -3.0) * 4.0;
    return std::min((size_t)sat_u64(std::fmax(t, 0.0)), (size_t)27);
  }
  void update(const std::vector<Impact>& recs, double m3, double scale_j, double d, double now,
              double sub_dt, double alpha, double wall_work) {
    double window = factor * std::sqrt(d / 9.81);
    updates++;
    if (updates % 4096 == 0) {
      for (auto it = last.begin(); it != last.end();) {
        if (!(now - it->second < 4.0 * window)) it = last.erase(it); else ++it;
      }
    }
    double count[4] = {0}, se_hat[4] = {0}, hist[28] = {0}, above[6] = {0};
    for (auto& rc : recs) {
      if (window > 0.0) {
        std::pair<uint32_t, uint32_t> key = {std::min(rc.ball, rc.partner), std::max(rc.ball, rc.partner)};
        if (rc.partner == UINT32_MAX) key = {rc.ball, UINT32_MAX};
        auto it = last.find(key);
        if (it != last.end() && now - it->second < window) continue;
        last[key] = now;
      }
      bool bb = rc.kind == 0;
      double vn = rc.vn, vs = rc.slip;
      int cls = bb ? (vs > vn ? 1 : 0) : (vs > vn ? 3 : 2);
      double m_red = bb ? 0.5 * m3 : m3;
      double se = 0.5 * m_red * (vn * vn + vs * vs);
      double eh = se / scale_j;
      count[cls] += 1.0;
      se_hat[cls] += eh;
      hist[se_bin(eh)] += 1.0;
      for (int t = 0; t < 6; t++) if (eh >= THR[t]) above[t] += se;
    }
    cum_wall += wall_work;
    for (int t = 0; t < 6; t++) cum_above[t] += above[t];
    auto ema = [&](double& s, double inst) { s += alpha * (inst / sub_dt - s); };
    for (int c = 0; c < 4; c++) { ema(count_rate[c], count[c]); ema(se_rate[c], se_hat[c]); }
    for (int b = 0; b < 28; b++) ema(hist_rate[b], hist[b]);
    for (int t = 0; t < 6; t++) ema(above_rate[t], above[t]);
  }
  double total_rate() const { return count_rate[0] + count_rate[1] + count_rate[2] + count_rate[3]; }
  // stress.rs:186-254: StressMetrics as 36 doubles: ball_mass_3d_kg, scale_energy_j,
  // impacts_per_ball_per_s[4], impacts_total_per_ball_per_s, mean_se_scaled[4], mean_se_scaled_drop,
  // mean_se_scaled_shear, stress_intensity_scaled, wall_power_per_ball_w, specific_power_w_per_kg,
  // thresholds_scaled[6], effective_power_per_ball_w[6] (the f32 fields rounded to f32), then the f64
  // cumulative_wall_energy_per_ball_j, cumulative_specific_energy_j_per_kg,
  // cumulative_effective_energy_per_ball_j[6].
  void metrics(size_t n_balls, double m3, double scale_j, double d, double power, double* o) const {
    const double n = (double)std::max<size_t>(n_balls, 1);
    int k = 0;
    o[k++] = (float)m3;
    o[k++] = (float)scale_j;
    for (int c = 0; c < 4; c++) o[k++] = (float)(count_rate[c] / n);
    const double total = total_rate();
    o[k++] = (float)(total / n);
    for (int c = 0; c < 4; c++) o[k++] = count_rate[c] > 0.0 ? (float)(se_rate[c] / count_rate[c]) : 0.0f;
    auto pooled = [&](int a, int b) {
      const double w = count_rate[a] + count_rate[b];
      return w > 0.0 ? (float)((se_rate[a] + se_rate[b]) / w) : 0.0f;
    };
    o[k++] = pooled(0, 2);
    o[k++] = pooled(1, 3);
    o[k++] = total > 0.0 ? (float)((se_rate[0] + se_rate[1] + se_rate[2] + se_rate[3]) / total) : 0.0f;
    const double ppb = power * d / n;
    o[k++] = (float)ppb;
    o[k++] = (float)(ppb / m3);
    for (int t = 0; t < 6; t++) o[k++] = (float)THR[t];
    for (int t = 0; t < 6; t++) o[k++] = (float)(above_rate[t] / n);
    o[k++] = cum_wall * d / n;
    o[k++] = cum_wall * d / n / m3;
    for (int t = 0; t < 6; t++) o[k++] = cum_above[t] / n;
  }
};
constexpr double Stress::THR[6];

// lib.rs:61-78
struct EnergyBudget {
  double elapsed = 0, shaft = 0, dem_wall = 0, fluid_wall = 0, contact_diss = 0, wall_slip = 0,
         viscous = 0, clamp_removed = 0, interface = 0, delta_mech = 0, unattributed = 0, fluid_wall_ang = 0;
};

// ------------------------------------------------------------------------------------------
// The simulation (lib.rs:82-566) with the fluid (pbf.rs) and DEM (dem.rs) steps.
// ------------------------------------------------------------------------------------------
template <class R, class A> struct Sim {
  Params P;
  Fluid<R> fl;
  Balls<R> bl;
  R drum_angle = 0;
  double sim_time = 0;
  R power = 0, torque = 0, coll_rate = 0, diss_power = 0, max_disp = 0;
  uint32_t clamp_hits_total = 0;
  EnergyBudget eb;
  double e0 = 0;
  Stress stress;
  FluidStats fs;
  DemStats ds;
  std::vector<Impact> impacts;
  // Debug captures of the last substep.
  Grid<R> dbg_grid;
  Nbrs dbg_nbrs;
  bool keep_debug = false;
  bool force_iters = false;
  uint32_t forced_density = 0, forced_cg = 0;

  static double ke_f64(const std::vector<V2<R>>& v, R m) {  // pbf.rs:924-926
    double s = 0;
    for (auto& q : v) s += (double)q.len2();
    return 0.5 * (double)m * s;
  }
  static A adot(V2<R> a, V2<R> b) { return (A)a.x * (A)b.x + (A)a.y * (A)b.y; }
  static A dotv(const std::vector<V2<R>>& a, const std::vector<V2<R>>& b) {  // pbf.rs:292-294
    A s = 0;
    for (size_t i = 0; i < a.size(); i++) s += adot(a[i], b[i]);
    return s;
  }

  Drum<R> drum() const {
    Drum<R> d;
    d.radius = (R)P.radius_m(); d.omega = (R)P.omega(); d.lifter_count = P.lifter_count;
    d.l_height = (R)P.lifter_height; d.l_base = (R)P.lifter_base; d.l_top = (R)P.lifter_top;
    d.l_phase_deg = (R)P.lifter_phase_deg;
    return d;
  }

  // dem.rs:355-409
  void seed_balls() {
    float td, d, dens, sf; uint32_t count;
    P.effective_media(td, d, dens, count, sf);
    float r = d * 0.5f;
    float mass = dens * C<float>::PI * r * r;
    float inertia = 0.5f * mass * r * r;
    float R_ = P.radius_m();
    Rng rng(P.seed);
    std::vector<V2<float>> x;
    if (r > 0.0f && count > 0) {
      float spacing = 2.0f * r * 1.02f;
      float row_h = spacing * (std::sqrt(3.0f) * 0.5f);
      float fill_r = rmax(R_ - r, 0.0f);
      int32_t n_rows = sat_i32(std::ceil(2.0f * fill_r / row_h)) + 1;
      bool done = false;
      for (int32_t row = -n_rows / 2; row <= n_rows / 2 && !done; row++) {
        float y = (float)row * row_h;
        if (std::fabs(y) > fill_r) continue;
        float hw = std::sqrt(rmax(fill_r * fill_r - y * y, 0.0f));
        float off = (((row % 2) + 2) % 2 == 0) ? 0.0f : spacing * 0.5f;
        int32_t n_cols = sat_i32(std::floor(2.0f * hw / spacing));
        for (int32_t col = -n_cols / 2; col <= n_cols / 2; col++) {
          float cx = (float)col * spacing + off;
          if (cx * cx + y * y > fill_r * fill_r) continue;
          float jx = rng.range_f32(-0.05f, 0.05f);
          float jy = rng.range_f32(-0.05f, 0.05f);
          V2<float> j = V2<float>(jx, jy) * r;
          x.push_back(V2<float>(cx, y) + j);
          if (x.size() >= count) { done = true; break; }
        }
      }
    }
    if (x.size() > count) x.resize(count);
    size_t n = x.size();
    bl.x.resize(n);
    for (size_t i = 0; i < n; i++) bl.x[i] = V2<R>((R)x[i].x, (R)x[i].y);
    bl.v.assign(n, V2<R>()); bl.theta.assign(n, 0); bl.omega.assign(n, 0);
    bl.radius = (R)r; bl.mass = (R)mass; bl.inertia = (R)inertia; bl.true_radius = (R)(td * 0.5f);
  }
  // pbf.rs:1044-1119 (+ dye_value pbf.rs:2369-2388)
  void seed_fluid(float fill, const std::vector<V2<float>>& balls_f, float ball_r) {
    float R_ = P.radius_m();
    uint32_t res = P.effective_fluid_resolution();
    float dx = R_ / (float)std::max(res, 1u);
    float h = 2.0f * dx;
    float rest = P.slurry_density;
    float row_h = dx * (std::sqrt(3.0f) * 0.5f);
    float cell_area = dx * row_h;
    float mass = rest * cell_area;
    float target_area = fill * C<float>::PI * R_ * R_;
    size_t target = (size_t)sat_u64(rmax(std::round(target_area / cell_area), 0.0f));
    std::vector<V2<float>> x;
    std::vector<float> dye;
    if (target > 0 && dx > 0.0f) {
      int32_t n_rows = sat_i32(std::ceil(2.0f * R_ / row_h)) + 1;
      bool done = false;
      for (int32_t row = -n_rows / 2; row <= n_rows / 2 && !done; row++) {
        float y = (float)row * row_h;
        if (std::fabs(y) > R_) continue;
        float hw = std::sqrt(rmax(R_ * R_ - y * y, 0.0f));
        float off = (((row % 2) + 2) % 2 == 0) ? 0.0f : dx * 0.5f;
        int32_t n_cols = sat_i32(std::floor(2.0f * hw / dx));
        for (int32_t col = -n_cols / 2; col <= n_cols / 2; col++) {
          float cx = (float)col * dx + off;
          if (cx * cx + y * y > R_ * R_) continue;
          V2<float> p(cx, y);
          bool ov = false;
          for (auto& b : balls_f) if ((p - b).len() < ball_r) { ov = true; break; }
          if (ov) continue;
          float dv = 0.0f;
          if (P.dye_pattern == 0) dv = p.x < 0.0f ? 0.0f : 1.0f;
          else if (P.dye_pattern == 1) dv = p.y < 0.0f ? 0.0f : 1.0f;
          dye.push_back(dv);
          x.push_back(p);
          if (x.size() >= target) { done = true; break; }
        }
      }
    }
    size_t n = x.size();
    fl.x.resize(n);
    for (size_t i = 0; i < n; i++) fl.x[i] = V2<R>((R)x[i].x, (R)x[i].y);
    fl.v.assign(n, V2<R>());
    fl.dye = dye;
    fl.guess.clear();
    fl.mass = (R)mass; fl.h = (R)h; fl.rest = (R)rest;
  }
  // lib.rs:138-207
  void init(const Params& p, const float* init_balls, size_t n_init) {
    P = p;
    seed_balls();
    if (init_balls) {
      bl.x.resize(n_init);
      for (size_t i = 0; i < n_init; i++) bl.x[i] = V2<R>((R)init_balls[2 * i], (R)init_balls[2 * i + 1]);
      bl.v.assign(n_init, V2<R>()); bl.theta.assign(n_init, 0); bl.omega.assign(n_init, 0);
    }
    std::vector<V2<float>> bf(bl.n());
    for (size_t i = 0; i < bl.n(); i++) bf[i] = V2<float>((float)bl.x[i].x, (float)bl.x[i].y);
    if (P.slurry_enabled) seed_fluid(P.slurry_fill, bf, (float)bl.radius);
    else seed_fluid(0.0f, {}, 0.0f);
    if (init_balls) {
      R w = (R)P.omega();
      for (size_t i = 0; i < bl.n(); i++) bl.v[i] = w * V2<R>(-bl.x[i].y, bl.x[i].x);
      for (auto& o : bl.omega) o = w;
      for (size_t i = 0; i < fl.n(); i++) fl.v[i] = w * V2<R>(-fl.x[i].y, fl.x[i].x);
    }
    e0 = total_mech();
  }
  double balls_mech_f64() const {  // dem.rs:265-273
    double e = 0;
    for (size_t i = 0; i < bl.n(); i++) {
      e += 0.5 * (double)bl.mass * (double)bl.v[i].len2();
      e += 0.5 * (double)bl.inertia * (double)(bl.omega[i] * bl.omega[i]);
      e += (double)bl.mass * (double)F(9.81) * (double)bl.x[i].y;
    }
    return e;
  }
  double fluid_mech_f64() const {  // pbf.rs:1145-1150
    double py = 0;
    for (auto& p : fl.x) py += (double)p.y;
    double pot = py * (double)fl.mass * (double)F(9.81);
    return ke_f64(fl.v, fl.mass) + pot;
  }
  double total_mech() const { return balls_mech_f64() + fluid_mech_f64(); }
  A balls_mech_j() const {  // dem.rs:253-261 (f32 in the reference)
    A e = 0;
    for (size_t i = 0; i < bl.n(); i++) {
      e += (A)(F(0.5) * bl.mass * bl.v[i].len2());
      e += (A)(F(0.5) * bl.inertia * bl.omega[i] * bl.omega[i]);
      e += (A)(bl.mass * F(9.81) * bl.x[i].y);
    }
    return e;
  }

  // ---------------------------------------------------------------------------------------
  // FluidParticles::step_coupled (pbf.rs:1174-2365)
  // ---------------------------------------------------------------------------------------
  Coupling<R> fluid_step(const Drum<R>& drum, R angle, uint32_t iterations, R dt) {
    const size_t n = fl.n(), nb = bl.n();
    Coupling<R> cp;
    cp.imp.assign(nb, V2<R>());
    cp.ang.assign(nb, 0);
    fs = FluidStats();
    if (n == 0 || dt <= (R)0) return cp;
    auto& x = fl.x;
    auto& v = fl.v;
    // 0. sanitize (pbf.rs:1200-1205)
    for (size_t i = 0; i < n; i++)
      if (!x[i].finite() || !v[i].finite()) { x[i] = V2<R>(); v[i] = V2<R>(); }
    const R h = fl.h, rest = fl.rest, mass = fl.mass;
    const R br = bl.radius;
    // ball grid (pbf.rs:1214-1222)
    Grid<R> ball_grid;
    bool have_bg = nb > 0;
    R contact_radius = br + F(0.05) * h;
    if (have_bg) ball_grid.build(bl.x, rmax(rmax(F(2.0) * br, contact_radius), F(1e-6)));
    // 1. predict (pbf.rs:1225-1235)
    std::vector<V2<R>> x0 = x;
    for (size_t i = 0; i < n; i++) {
      v[i].y += F(-9.81) * dt;
      x[i] += v[i] * dt;
    }
    R angle_next = angle + drum.omega * dt;
    // 2. grid + neighbour lists (pbf.rs:1238-1239)
    Grid<R> grid;
    grid.build(x, h);
    Nbrs nbr = build_neighbor_lists(x, grid, h);
    if (keep_debug) { dbg_grid = grid; dbg_nbrs = nbr; }
    // 3. density solve (pbf.rs:1257-1334)
    std::vector<R> density(n, 0), lambda(n, 0);
    std::vector<V2<R>> dp(n), gcache(nbr.nb.size());
    uint32_t base = std::max(iterations, 1u), it_count = 0;
    for (;;) {
      R max_c = 0;
      for (size_t i = 0; i < n; i++) {
        R rho = mass * poly6((R)0, h);
        V2<R> gs;
        R sgs = 0;
        for (uint32_t k = nbr.off[i]; k < nbr.off[i + 1]; k++) {
          uint32_t j = nbr.nb[k];
          V2<R> d = x[i] - x[j];
          R r2 = d.len2();
          rho += mass * poly6(r2, h);
          R r = d.len();
          V2<R> rg = spiky_grad(d, r, h);
          gcache[k] = rg;
          V2<R> g = rg * (mass / rest);
          gs += g;
          sgs += g.len2();
        }
        density[i] = rho;
        R c = rmax(density[i] / rest - F(1.0), (R)0);
        max_c = rmax(max_c, c);
        sgs += gs.len2();
        lambda[i] = -c / (sgs + F(200.0));
      }
      for (size_t i = 0; i < n; i++) {
        V2<R> s;
        for (uint32_t k = nbr.off[i]; k < nbr.off[i + 1]; k++) {
          uint32_t j = nbr.nb[k];
          s += (lambda[i] + lambda[j]) * gcache[k];
        }
        dp[i] = s * (mass / rest);
      }
      for (size_t i = 0; i < n; i++) x[i] += dp[i];
      it_count++;
      bool conv = max_c <= F(0.02);
      if (force_iters) {
        if (it_count >= forced_density) break;
      } else if ((it_count >= base && conv) || it_count >= base + 6) {
        break;
      }
    }
    fs.density_iters = it_count;
    // 3.5 overlap push (pbf.rs:1361-1444)
    std::vector<V2<R>> excess(n);
    if (have_bg) {
      R red = bl.mass > (R)0 ? (mass * bl.mass) / (mass + bl.mass) : mass;
      std::vector<uint32_t> nearby;
      for (size_t i = 0; i < n; i++) {
        nearby.clear();
        ball_grid.near(x[i], [&](uint32_t b) { nearby.push_back(b); });
        for (uint32_t b : nearby) {
          V2<R> d = x[i] - bl.x[b];
          R dist = d.len();
          if (std::isnan(dist) || dist >= contact_radius) continue;
          V2<R> nh = dist > F(1e-9) ? d / dist : V2<R>(1, 0);
          R pm = rmin(contact_radius - dist, F(0.5) * br);
          x[i] += pm * nh;
          V2<R> lever = nh * br;
          V2<R> vs = bl.v[b] + bl.omega[b] * V2<R>(-lever.y, lever.x);
          R arrest = rmax(-(v[i] - vs).dot(nh), (R)0);
          R full = mass * pm / dt;
          R im = rmin(full, red * arrest);
          excess[i] += ((full - im) / mass) * nh;
          V2<R> react = -im * nh;
          cp.imp[b] += react;
          cp.fmc += im * nh;
          cp.ang[b] += cross2(lever, react);
        }
      }
    }
    // 3.6a cohesion (pbf.rs:1464-1491)
    std::vector<V2<R>> coh(n);
    R coh_acc = (rmax((R)P.surface_tension, (R)0) / F(0.072)) * F(2.0) * F(9.81);
    if (coh_acc > (R)0) {
      for (size_t i = 0; i < n; i++) {
        V2<R> pull;
        for (uint32_t k = nbr.off[i]; k < nbr.off[i + 1]; k++) {
          uint32_t j = nbr.nb[k];
          V2<R> d = x[i] - x[j];
          R r = d.len();
          if (r <= F(1e-9)) continue;
          R w = mass * cohesion_kernel(r, h) / rest;
          pull -= w * (d / r);
        }
        coh[i] = coh_acc * pull * dt;
      }
    }
    // 3.6b adhesion (pbf.rs:1515-1631)
    std::vector<V2<R>> adh(n);
    R wett = rclamp((R)P.wettability, (R)0, (R)1);
    if (wett > (R)0 && nb > 0 && br > (R)0) {
      R dx = F(0.5) * h;
      R ha = rmin(h, F(2.0) * br);
      uint32_t cpb = boundary_points_per_ball(br, dx);
      std::vector<V2<R>> offs(cpb);
      for (uint32_t k = 0; k < cpb; k++) {
        R a = C<R>::TAU * (R)k / (R)cpb;
        offs[k] = V2<R>(std::cos(a), std::sin(a)) * br;
      }
      if (dx > (R)0 && ha > (R)0) {
        R acc = wett * F(2.0) * F(9.81);
        struct Item { uint32_t j; V2<R> dir; R w; };
        std::vector<Item> items;
        std::vector<uint32_t> slot(n, UINT32_MAX), nearby;
        R range2 = ha * ha * (F(1.0) + F(1e-4));
        R inner2 = F(0.25) * ha * ha * (F(1.0) - F(1e-4));
        for (size_t b = 0; b < nb; b++) {
          if (!bl.x[b].finite()) continue;
          items.clear();
          for (uint32_t k = 0; k < cpb; k++) {
            V2<R> bp = bl.x[b] + offs[k];
            nearby.clear();
            grid.near(bp, [&](uint32_t j) { nearby.push_back(j); });
            for (uint32_t j : nearby) {
              V2<R> d = bp - x[j];
              R r2 = d.len2();
              if (r2 >= range2 || r2 <= inner2) continue;
              R r = std::sqrt(r2);
              if (!std::isfinite(r) || r <= F(0.5) * ha || r >= ha) continue;
              R w = mass * adhesion_kernel(r, ha) / rest;
              if (w <= (R)0) continue;
              V2<R> nh = d / r;
              if (slot[j] == UINT32_MAX) {
                slot[j] = (uint32_t)items.size();
                items.push_back({j, w * nh, w});
              } else {
                Item& e = items[slot[j]];
                e.dir += w * nh;
                e.w += w;
              }
            }
          }
          for (auto& it : items) slot[it.j] = UINT32_MAX;
          if (items.empty()) continue;
          R wsum = 0;
          V2<R> pds;
          for (auto& it : items) { pds += it.dir; wsum += it.w; }
          if (wsum <= F(1e-9)) continue;
          R pl = pds.len();
          if (pl <= F(1e-9)) continue;
          V2<R> bdir = -pds / pl;
          R cov = rmin(wsum, (R)1);
          R mc = mass * wsum;
          R att = mc / (bl.mass + mc);
          V2<R> dvb = bdir * (acc * cov * dt * att);
          cp.imp[b] += bl.mass * dvb;
          V2<R> lever = bdir * br;
          cp.ang[b] += bl.mass * cross2(lever, dvb);
          for (auto& it : items) {
            R frac = it.w / wsum;
            V2<R> dvi = -(bl.mass * frac / mass) * dvb;
            adh[it.j] += dvi;
            cp.fmc += mass * dvi;
          }
        }
      }
    }
    // 4. wall projection (pbf.rs:1635-1653)
    std::vector<uint8_t> touched(n, 0);
    double wimp[2] = {0, 0}, wang = 0;
    float bins[32] = {0};
    for (size_t i = 0; i < n; i++) {
      auto dn = drum.sdf_world(x[i], angle_next);
      R d = dn.first;
      if (d < (R)0) {
        x[i] += -d * dn.second;
        touched[i] = 1;
        V2<R> push = mass * (-d * dn.second) / dt;
        wimp[0] += (double)push.x; wimp[1] += (double)push.y;
        wang += (double)cross2(x[i], push);
        R a = std::atan2(x[i].y, x[i].x);
        size_t bin = (size_t)sat_u64(((a + C<R>::PI) / C<R>::TAU) * F(32.0));
        bins[std::min(bin, (size_t)31)] += (float)(mass * -d / dt);
      }
    }
    // 5. velocity reconstruction (pbf.rs:1656-1674)
    for (size_t i = 0; i < n; i++) v[i] = (x[i] - x0[i]) / dt;
    for (size_t i = 0; i < n; i++) v[i] -= excess[i];
    for (size_t i = 0; i < n; i++) v[i] += coh[i];
    for (size_t i = 0; i < n; i++) v[i] += adh[i];
    // 6. no-slip blend (pbf.rs:1677-1700)
    R beta = rclamp((R)P.wall_no_slip, (R)0, (R)1);
    A wall_work = 0;
    double ke_before = ke_f64(v, mass);
    if (beta > (R)0) {
      for (size_t i = 0; i < n; i++) {
        if (!touched[i]) continue;
        V2<R> vw = drum.wall_velocity(x[i]);
        V2<R> vo = v[i];
        v[i] = (F(1.0) - beta) * vo + beta * vw;
        V2<R> dv = v[i] - vo;
        wall_work += (A)(mass * dv.dot(vw));
        wimp[0] += (double)(mass * dv.x); wimp[1] += (double)(mass * dv.y);
        wang += (double)cross2(x[i], mass * dv);
      }
    }
    double ke_mark = ke_f64(v, mass);
    fs.ke_noslip = ke_mark - ke_before;
    // 6.5 form drag (pbf.rs:1777-2013)
    R mu = rmax((R)P.viscosity, (R)0);
    if (nb > 0 && br > (R)0 && bl.true_radius > (R)0 && mu > F(1e-6)) {
      R beta_b = rclamp((R)P.ball_no_slip, (R)0, (R)1);
      R vmax = F(5.0) * (std::fabs(drum.omega) * drum.radius + std::sqrt(F(4.0) * F(9.81) * drum.radius));
      R rot_cap = F(0.2) * vmax;
      R rho_ball = bl.mass / (C<R>::PI * br * br);
      R mass_true = rho_ball * C<R>::PI * bl.true_radius * bl.true_radius;
      R inv_tau_stokes = 0;
      if (beta_b > (R)0) {
        R hc = br + h;
        R w0 = rmax(poly6((R)0, hc), F(1e-12));
        Grid<R> cg;
        cg.build(x, rmax(hc, F(1e-6)));
        R ib = bl.inertia;
        std::vector<uint32_t> nearby;
        struct It { uint32_t j; V2<R> ri; R phi; };
        std::vector<It> items;
        for (size_t b = 0; b < nb; b++) {
          if (!bl.x[b].finite()) continue;
          nearby.clear();
          cg.near(bl.x[b], [&](uint32_t j) { nearby.push_back(j); });
          if (nearby.empty()) continue;
          items.clear();
          R m_ent = 0, i_ent = 0, l_sum = 0;
          V2<R> r_sum, v_sum;
          for (uint32_t j : nearby) {
            V2<R> ri = x[j] - bl.x[b];
            R r2 = ri.len2();
            if (!std::isfinite(r2) || r2 >= hc * hc) continue;
            R phi = poly6(r2, hc) / w0;
            if (phi <= (R)0) continue;
            R w = mass * phi;
            m_ent += w;
            r_sum += w * ri;
            v_sum += w * v[j];
            i_ent += w * r2;
            l_sum += w * cross2(ri, v[j]);
            items.push_back({j, ri, phi});
          }
          if (m_ent <= F(1e-12) || items.empty()) continue;
          V2<R> r_bar = r_sum / m_ent, v_bar = v_sum / m_ent;
          V2<R> v_rel = v_bar - bl.v[b];
          R sp = v_rel.len();
          R inv_tau_form = (sp > F(1e-6) && rest > (R)0)
                               ? F(0.5) * rest * F(1.0) * (F(2.0) * bl.true_radius) * sp / mass_true
                               : (R)0;
          R bt = beta_b * (F(1.0) - std::exp(-dt * (inv_tau_stokes + inv_tau_form)));
          R a_lin = bt * m_ent / (bl.mass + m_ent);
          V2<R> dvb = a_lin * v_rel;
          R dwb = 0, c_rot = 0;
          V2<R> p_bar;
          if (i_ent > F(1e-12)) {
            R om_bar = l_sum / i_ent;
            R a_rot = bt * i_ent / (ib + i_ent);
            R cand = a_rot * (om_bar - bl.omega[b]);
            p_bar = V2<R>(-r_bar.y, r_bar.x);
            R s = 0, s_scale = 0;
            for (auto& it : items) {
              V2<R> pi_ = V2<R>(-it.ri.y, it.ri.x);
              R phi2 = it.phi * it.phi;
              s += mass * phi2 * cross2(it.ri, pi_ - p_bar);
              s_scale += mass * phi2 * it.ri.len2();
            }
            if (std::fabs(s) > F(0.1) * s_scale) {
              R raw = -ib * cand / s;
              R ml = F(2.0) * hc;
              R cap = ml > F(1e-9) ? rot_cap / ml : (R)0;
              c_rot = rclamp(raw, -cap, cap);
              dwb = -c_rot * s / ib;
            }
          }
          for (auto& it : items) {
            V2<R> dvi = (-(bl.mass / m_ent)) * dvb * it.phi;
            if (c_rot != (R)0) {
              V2<R> pi_ = V2<R>(-it.ri.y, it.ri.x);
              dvi += c_rot * it.phi * (pi_ - p_bar);
            }
            v[it.j] += dvi;
            cp.fmc += mass * dvi;
          }
          cp.imp[b] += bl.mass * dvb;
          cp.ang[b] += ib * dwb + bl.mass * cross2(r_bar, dvb);
        }
      }
    }
    double ke_after_drag = ke_f64(v, mass);
    fs.ke_drag = ke_after_drag - ke_mark;
    ke_mark = ke_after_drag;
    // 6.6 buoyancy (pbf.rs:2035-2122) — queries the step-2 grid with current positions.
    if (nb > 0 && br > (R)0) {
      R area = C<R>::PI * br * br;
      std::vector<uint32_t> nearby;
      std::vector<std::pair<uint32_t, R>> ws;
      for (size_t b = 0; b < nb; b++) {
        if (!bl.x[b].finite()) continue;
        nearby.clear();
        grid.near(bl.x[b], [&](uint32_t j) { nearby.push_back(j); });
        if (nearby.empty()) continue;
        R rho_l = 0;
        ws.clear();
        for (uint32_t j : nearby) {
          R r2 = (bl.x[b] - x[j]).len2();
          R w = mass * poly6(r2, h);
          if (w > (R)0) { rho_l += w; ws.push_back({j, w}); }
        }
        if (rho_l <= (R)0) continue;
        R rho_eff = rmin(rho_l, rest);
        R on = 0, od = 0;
        for (auto& q : ws) {
          V2<R> d = x[q.first];
          R d2 = d.len2();
          if (d2 <= F(1e-12)) continue;
          on += q.second * cross2(d, v[q.first]) / d2;
          od += q.second;
        }
        R om_l = od > (R)0 ? rclamp(on / od, -std::fabs(drum.omega), std::fabs(drum.omega)) : (R)0;
        V2<R> g(0, F(-9.81));
        V2<R> cen = (om_l * om_l) * bl.x[b];
        V2<R> ge = g + cen;
        V2<R> imp = ((-rho_eff) * area) * ge * dt;
        cp.imp[b] += imp;
        for (auto& q : ws) {
          R frac = q.second / rho_l;
          V2<R> dv = -(imp * frac) / mass;
          v[q.first] += dv;
          cp.fmc += mass * dv;
        }
      }
    }
    double ke_after_buoy = ke_f64(v, mass);
    fs.ke_buoy = ke_after_buoy - ke_mark;
    ke_mark = ke_after_buoy;
    // 7a density recompute (pbf.rs:2136-2144)
    for (size_t i = 0; i < n; i++) {
      R rho = mass * poly6((R)0, h);
      for (uint32_t k = nbr.off[i]; k < nbr.off[i + 1]; k++) rho += mass * poly6((x[i] - x[nbr.nb[k]]).len2(), h);
      density[i] = rho;
    }
    uint32_t visc_iters = 0;
    if (mu > (R)0) {
      // 7b weights (pbf.rs:2147-2154; shear_rates :821-849; bingham :854-868; morris :312-344)
      std::vector<R> mu_eff(n, mu);
      if (P.rheology == 1 && P.yield_stress > 0.0f) {
        std::vector<R> rates = shear_rates(x, v, nbr, density, mass, h);
        R tau_y = (R)P.yield_stress;
        R g0 = tau_y / (F(100.0) * mu);
        for (size_t i = 0; i < n; i++) {
          R g = std::isfinite(rates[i]) ? rates[i] : (R)0;
          R extra = g < F(1e-4) * g0 ? tau_y / g0 : tau_y * (F(1.0) - std::exp(-g / g0)) / g;
          mu_eff[i] = mu + extra;
        }
      }
      R eta2 = F(0.01) * h * h;
      std::vector<R> W(nbr.nb.size());
      std::vector<V2<R>> D(nbr.nb.size());
      for (size_t i = 0; i < n; i++) {
        R rho_i = rmax(density[i], F(1e-6));
        for (uint32_t k = nbr.off[i]; k < nbr.off[i + 1]; k++) {
          uint32_t j = nbr.nb[k];
          R rho_j = rmax(density[j], F(1e-6));
          V2<R> d = x[i] - x[j];
          R r = d.len();
          V2<R> g = spiky_grad(d, r, h);
          R num = (mass * (mu_eff[i] + mu_eff[j]) / (rho_i * rho_j)) * std::fabs(d.dot(g));
          W[k] = num / (d.len2() + eta2);
          D[k] = r > F(1e-9) ? d / r : V2<R>(1, 0);
        }
      }
      // 7c ball boundary groups (pbf.rs:2159-2232)
      struct Group { uint32_t fluid, ball; R c; V2<R> lever_c; R lever2_c; };
      std::vector<Group> groups;
      if (nb > 0 && br > (R)0) {
        R dx = F(0.5) * h;
        uint32_t cnt = boundary_points_per_ball(br, dx);
        std::vector<V2<R>> offs(cnt);
        for (uint32_t k = 0; k < cnt; k++) {
          R a = C<R>::TAU * (R)k / (R)cnt;
          offs[k] = V2<R>(std::cos(a), std::sin(a)) * br;
        }
        Grid<R> gnow;
        gnow.build(x, h);
        R sn = F(-30.0) / (C<R>::PI * powi(h, 5));
        R range2 = h * h * (F(1.0) + F(1e-4));
        std::vector<R> ff(n);
        for (size_t j = 0; j < n; j++) ff[j] = F(0.5) * (mass * F(2.0) * mu_eff[j] / (rmax(density[j], F(1e-6)) * rest));
        std::vector<uint32_t> slot(n, UINT32_MAX);
        size_t gstart = 0;
        size_t cur_ball = SIZE_MAX;
        if (dx > (R)0) {
          for (size_t k = 0; k < nb * cnt; k++) {
            size_t bi = k / cnt;
            V2<R> bp = bl.x[bi] + offs[k % cnt];
            if (bi != cur_ball) {
              for (size_t q = gstart; q < groups.size(); q++) slot[groups[q].fluid] = UINT32_MAX;
              gstart = groups.size();
              cur_ball = bi;
            }
            if (!bl.x[bi].finite()) continue;
            gnow.near(bp, [&](uint32_t j) {
              V2<R> d = x[j] - bp;
              R r2 = d.len2();
              if (r2 >= range2) return;
              R r = std::sqrt(r2);
              if (!std::isfinite(r) || r >= h || r <= F(1e-9)) return;
              R da = std::fabs(sn * (h - r) * (h - r) * r);
              R c = ff[j] * da / (r2 + eta2);
              if (c > (R)0) {
                V2<R> lever = bp - bl.x[bi];
                V2<R> lc = c * lever;
                R l2 = c * lever.len2();
                uint32_t s = slot[j];
                if (s == UINT32_MAX) {
                  slot[j] = (uint32_t)groups.size();
                  groups.push_back({j, (uint32_t)bi, c, lc, l2});
                } else {
                  groups[s].c += c; groups[s].lever_c += lc; groups[s].lever2_c += l2;
                }
              }
            });
          }
        }
      }
      fs.n_groups = (uint32_t)groups.size();
      // 7d squeeze links (pbf.rs:544-625), active set from velocities after earlier impulses.
      std::vector<V2<R>> bv0(nb);
      for (size_t b = 0; b < nb; b++) bv0[b] = bl.v[b] + cp.imp[b] / bl.mass;
      struct Link { uint32_t ball, other; V2<R> n; R c; };
      std::vector<Link> links;
      uint32_t in_range = 0;
      if (nb > 0 && br > (R)0 && mu > (R)0) {
        double a = (double)br, bD = (double)drum.radius, dxd = (double)(F(0.5) * h), mud = (double)mu;
        R reach = (R)(8.0 * dxd);
        R c_max = F(8.0) * bl.mass / rmax(dt, F(1e-12));
        Grid<R> g2;
        g2.build(bl.x, F(2.0) * br + reach);
        R rng2 = (F(2.0) * br + reach) * (F(2.0) * br + reach) * (F(1.0) + F(1e-4));
        SqueezeFilm pf(1.0 / a, 1.0 / a, mud), wf(1.0 / a, -1.0 / bD, mud);
        for (size_t i = 0; i < nb; i++) {
          V2<R> xi = bl.x[i];
          if (!xi.finite()) continue;
          g2.near(xi, [&](uint32_t j) {
            if ((size_t)j <= i) return;
            V2<R> e = bl.x[j] - xi;
            if (e.len2() >= rng2 || !e.finite()) return;
            R dist = e.len();
            R gap = dist - F(2.0) * br;
            if (gap >= reach || dist < F(1e-9)) return;
            in_range++;
            V2<R> nn = e / dist;
            if (nn.dot(bv0[i] - bv0[j]) <= (R)0) return;
            double wgt = 1.0 - resolved_ratio((double)gap / dxd);
            double hh = std::fmax((double)gap, 0.02 * a);
            double c = wgt * pf.coefficient(hh);
            links.push_back({(uint32_t)i, j, nn, rmin(F(1.0) * (R)(float)c, c_max)});
          });
          R dist = xi.len();
          R gap = drum.radius - br - dist;
          if (gap < reach && dist > F(1e-9)) {
            in_range++;
            V2<R> nn = xi / dist;
            if (nn.dot(bv0[i]) <= (R)0) continue;
            double wgt = 1.0 - resolved_ratio((double)gap / dxd);
            double hh = std::fmax((double)gap, 0.02 * a);
            double c = wgt * wf.coefficient(hh);
            links.push_back({(uint32_t)i, UINT32_MAX, nn, rmin(F(1.0) * (R)(float)c, c_max)});
          }
        }
      }
      fs.n_links = (uint32_t)links.size();
      fs.links_in_range = in_range;
      // 7e CG (pbf.rs:2244-2299)
      auto lap = [&](const V2<R>* vv, V2<R>* out) {  // pbf.rs:359-377
        for (size_t i = 0; i < n; i++) {
          V2<R> s;
          for (uint32_t k = nbr.off[i]; k < nbr.off[i + 1]; k++) {
            uint32_t j = nbr.nb[k];
            V2<R> e = D[k];
            s += (F(4.0) * W[k] * e.dot(vv[i] - vv[j])) * e;
          }
          out[i] = s;
        }
      };
      if (groups.empty() && in_range == 0) {
        // solve_implicit_viscosity (pbf.rs:423-480)
        visc_iters = [&]() -> uint32_t {
          A bn = std::sqrt(dotv(v, v));
          if (bn <= (A)0 || !std::isfinite(bn)) return 0;
          A target = (A)F(1e-3) * bn;
          std::vector<V2<R>> lv(n), ax(n), r(n), p(n), ap(n);
          lap(v.data(), lv.data());
          for (size_t i = 0; i < n; i++) ax[i] = v[i] + dt * lv[i];
          for (size_t i = 0; i < n; i++) r[i] = v[i] - ax[i];
          p = r;
          A rs_old = dotv(r, r);
          if (std::sqrt(rs_old) <= target) return 0;
          uint32_t maxit = force_iters ? forced_cg : 50;
          if (force_iters && maxit == 0) return 0;
          for (uint32_t it = 1; it <= maxit; it++) {
            lap(p.data(), lv.data());
            for (size_t i = 0; i < n; i++) ap[i] = p[i] + dt * lv[i];
            A pap = dotv(p, ap);
            if (pap <= (A)0 || !std::isfinite(pap)) return it;
            R alpha = (R)(rs_old / pap);
            for (size_t i = 0; i < n; i++) { v[i] += alpha * p[i]; r[i] -= alpha * ap[i]; }
            A rs_new = dotv(r, r);
            if (!force_iters && std::sqrt(rs_new) <= target) return it;
            if (force_iters && it == maxit) return it;
            R bta = (R)(rs_new / rs_old);
            for (size_t i = 0; i < n; i++) p[i] = r[i] + bta * p[i];
            rs_old = rs_new;
          }
          return 50;
        }();
      } else {
        std::vector<V2<R>> bv = bv0;
        std::vector<R> bw(nb);
        for (size_t b = 0; b < nb; b++) bw[b] = bl.omega[b] + cp.ang[b] / bl.inertia;
        std::vector<R> bw0 = bw;
        std::vector<V2<R>> guess;
        guess.swap(fl.guess);
        std::vector<V2<R>> v_in = v;
        // solve_viscosity_with_balls (pbf.rs:641-791)
        const size_t len = n + 2 * nb;
        const R mb = bl.mass, Ib = bl.inertia;
        std::vector<V2<R>> X(len), B(len);
        for (size_t i = 0; i < n; i++) X[i] = guess.size() == n ? v[i] + guess[i] : v[i];
        for (size_t b = 0; b < nb; b++) { X[n + b] = bv[b]; X[n + nb + b] = V2<R>(bw[b], 0); }
        for (size_t i = 0; i < n; i++) B[i] = mass * v[i];
        for (size_t b = 0; b < nb; b++) { B[n + b] = mb * bv[b]; B[n + nb + b] = V2<R>(Ib * bw[b], 0); }
        std::vector<R> diag(len, 0);
        for (size_t i = 0; i < n; i++) {
          R sw = 0;
          for (uint32_t k = nbr.off[i]; k < nbr.off[i + 1]; k++) sw += W[k];
          diag[i] = mass * (F(1.0) + dt * (F(0.5) * F(4.0) * sw));
        }
        for (size_t b = 0; b < nb; b++) { diag[n + b] = mb; diag[n + nb + b] = Ib; }
        for (auto& g : groups) {
          R d = dt * mass * g.c;
          diag[g.fluid] += d;
          diag[n + g.ball] += d;
          diag[n + nb + g.ball] += dt * mass * g.lever2_c;
        }
        for (auto& l : links) {
          R d = dt * l.c;
          diag[n + l.ball] += d;
          if (l.other != UINT32_MAX) diag[n + l.other] += d;
        }
        std::vector<R> inv(len);
        for (size_t i = 0; i < len; i++) inv[i] = F(1.0) / rmax(diag[i], F(1e-20));
        std::vector<V2<R>> lv(n);
        auto apply = [&](const std::vector<V2<R>>& xx, std::vector<V2<R>>& out) {
          lap(xx.data(), lv.data());
          for (size_t i = 0; i < n; i++) out[i] = mass * (xx[i] + dt * lv[i]);
          for (size_t b = 0; b < nb; b++) {
            out[n + b] = mb * xx[n + b];
            out[n + nb + b] = V2<R>(Ib * xx[n + nb + b].x, 0);
          }
          for (auto& g : groups) {
            size_t i = g.fluid, bi = g.ball;
            V2<R> vb = xx[n + bi];
            R w = xx[n + nb + bi].x;
            V2<R> rel = xx[i] - vb;
            R k = dt * mass;
            V2<R> d = k * (g.c * rel - w * perp(g.lever_c));
            out[i] += d;
            out[n + bi] -= d;
            out[n + nb + bi].x -= k * (cross2(g.lever_c, rel) - w * g.lever2_c);
          }
          for (auto& l : links) {
            size_t i = l.ball;
            V2<R> rel = xx[n + i];
            if (l.other != UINT32_MAX) rel = rel - xx[n + l.other];
            V2<R> d = (dt * l.c * l.n.dot(rel)) * l.n;
            out[n + i] += d;
            if (l.other != UINT32_MAX) out[n + l.other] -= d;
          }
        };
        auto res_norm = [&](const std::vector<V2<R>>& r) {
          A s = 0;
          for (size_t i = 0; i < len; i++) s += (A)inv[i] * ((A)r[i].x * (A)r[i].x + (A)r[i].y * (A)r[i].y);
          return std::sqrt(s);
        };
        std::vector<V2<R>> ax(len), r(len), z(len);
        apply(X, ax);
        for (size_t i = 0; i < len; i++) r[i] = B[i] - ax[i];
        for (size_t i = 0; i < len; i++) z[i] = r[i] * inv[i];
        A bn = res_norm(B);
        uint32_t iters = 0;
        if (bn > (A)0 && std::isfinite(bn) && res_norm(r) > (A)F(1e-3) * bn) {
          A target = (A)F(1e-3) * bn;
          std::vector<V2<R>> pd = z, ap(len);
          A rz_old = dotv(r, z);
          uint32_t maxit = force_iters ? forced_cg : 200;
          for (uint32_t it = 1; it <= maxit; it++) {
            iters = it;
            apply(pd, ap);
            A pap = dotv(pd, ap);
            if (pap <= (A)0 || !std::isfinite(pap)) break;
            R alpha = (R)(rz_old / pap);
            for (size_t i = 0; i < len; i++) { X[i] += alpha * pd[i]; r[i] -= alpha * ap[i]; }
            if (!force_iters && res_norm(r) <= target) break;
            if (force_iters && it == maxit) break;
            for (size_t i = 0; i < len; i++) z[i] = r[i] * inv[i];
            A rz_new = dotv(r, z);
            R bta = (R)(rz_new / rz_old);
            for (size_t i = 0; i < len; i++) pd[i] = z[i] + bta * pd[i];
            rz_old = rz_new;
          }
        }
        for (size_t i = 0; i < n; i++) v[i] = X[i];
        for (size_t b = 0; b < nb; b++) { bv[b] = X[n + b]; bw[b] = X[n + nb + b].x; }
        visc_iters = iters;
        fl.guess.resize(n);
        for (size_t i = 0; i < n; i++) {
          V2<R> d = v[i] - v_in[i];
          fl.guess[i] = d.finite() ? d : V2<R>();
        }
        for (size_t b = 0; b < nb; b++) {
          V2<R> dv = bv[b] - bv0[b];
          R dw = bw[b] - bw0[b];
          if (dv.finite() && std::isfinite(dw)) {
            cp.imp[b] += mb * dv;
            cp.ang[b] += Ib * dw;
            cp.fmc -= mb * dv;
          }
        }
      }
    }
    fs.visc_iters = visc_iters;
    // mean shear rate (pbf.rs:800-818)
    {
      std::vector<R> rates = shear_rates(x, v, nbr, density, mass, h);
      A s = 0;
      uint32_t c = 0;
      for (R g : rates) if (std::isfinite(g)) { s += (A)g; c++; }
      fs.mean_shear = c == 0 ? 0.0 : (double)(R)(s / (A)c);
    }
    double ke_after_visc = ke_f64(v, mass);
    fs.ke_visc = ke_after_visc - ke_mark;
    // 7.5 speed clamp (pbf.rs:987-996, :2321)
    R vmax = F(5.0) * (std::fabs(drum.omega) * drum.radius + std::sqrt(F(4.0) * F(9.81) * drum.radius));
    for (size_t i = 0; i < n; i++) {
      R sp = v[i].len();
      if (!std::isfinite(sp)) v[i] = V2<R>();
      else if (sp > vmax) v[i] *= vmax / sp;
    }
    fs.ke_clamp = ke_f64(v, mass) - ke_after_visc;
    // 8. impulse clamp (pbf.rs:2330-2345)
    if (nb > 0) {
      R cl = F(60.0) * bl.mass * F(9.81) * dt;
      R acl = cl * br;
      for (size_t b = 0; b < nb; b++) {
        R m = cp.imp[b].len();
        if (m > cl && m > F(1e-9)) { cp.imp[b] *= cl / m; cp.clamp_hits++; }
        cp.ang[b] = rclamp(cp.ang[b], -acl, acl);
      }
    }
    fs.clamp_hits = cp.clamp_hits;
    fs.wall_work = (double)(R)wall_work;
    fs.wall_imp[0] = wimp[0]; fs.wall_imp[1] = wimp[1];
    fs.wall_ang = wang;
    std::memcpy(fs.bins, bins, sizeof(bins));
    return cp;
  }

  static std::vector<R> shear_rates(const std::vector<V2<R>>& x, const std::vector<V2<R>>& v, const Nbrs& nbr,
                                    const std::vector<R>& density, R mass, R h) {  // pbf.rs:821-849
    size_t n = x.size();
    std::vector<R> out(n, 0);
    for (size_t i = 0; i < n; i++) {
      R gxx = 0, gxy = 0, gyx = 0, gyy = 0;
      for (uint32_t k = nbr.off[i]; k < nbr.off[i + 1]; k++) {
        uint32_t j = nbr.nb[k];
        V2<R> d = x[i] - x[j];
        R r = d.len();
        V2<R> g = spiky_grad(d, r, h);
        R coeff = mass / rmax(density[j], F(1e-6));
        V2<R> dv = v[j] - v[i];
        gxx += coeff * dv.x * g.x;
        gxy += coeff * dv.x * g.y;
        gyx += coeff * dv.y * g.x;
        gyy += coeff * dv.y * g.y;
      }
      R dxy = F(0.5) * (gxy + gyx);
      out[i] = std::sqrt(F(2.0) * (gxx * gxx + gyy * gyy + F(2.0) * dxy * dxy));
    }
    return out;
  }

  // ---------------------------------------------------------------------------------------
  // DemState::step_recorded (dem.rs:530-1066)
  // ---------------------------------------------------------------------------------------
  static R angle_diff(R a, R b) {  // dem.rs:1069-1078
    R d = a - b;
    // Bounded like the engine: only a diverged (huge or non-finite) difference needs > 64 turns.
    for (int k = 0; k < 64 && d > C<R>::PI; k++) d -= F(2.0) * C<R>::PI;
    for (int k = 0; k < 64 && d < -C<R>::PI; k++) d += F(2.0) * C<R>::PI;
    return d;
  }
  static R eff_friction(R mk, R ms, R vt, R vs) {  // dem.rs:128-138
    if (ms <= mk) return mk;
    R ratio = vt / vs;
    R w = std::exp(-(ratio * ratio));
    return mk + (ms - mk) * w;
  }
  void dem_step(const Drum<R>& drum, R angle, uint32_t iterations, R dt, const Coupling<R>* ext) {
    ds = DemStats();
    impacts.clear();
    const size_t n = bl.n();
    if (n == 0 || dt <= (R)0) return;
    auto& x = bl.x;
    auto& v = bl.v;
    auto& th = bl.theta;
    auto& om = bl.omega;
    R w = bl.mass > (R)0 ? F(1.0) / bl.mass : (R)0;
    R w_rot = bl.inertia > (R)0 ? F(1.0) / bl.inertia : (R)0;
    R r = bl.radius;
    R thr = F(0.1428) * std::sqrt(F(9.81) * F(2.0) * r);
    for (size_t i = 0; i < n; i++)
      if (!x[i].finite() || !v[i].finite()) { x[i] = V2<R>(); v[i] = V2<R>(); }
    std::vector<V2<R>> x0 = x, v_pre = v;
    std::vector<R> th0 = th, om_pre = om;
    double cw = 0;
    for (size_t i = 0; i < n; i++) {
      v[i].y += F(-9.81) * dt;
      if (ext) {
        double kb = 0.5 * (double)bl.mass * (double)v[i].len2() + 0.5 * (double)bl.inertia * (double)(om[i] * om[i]);
        v[i] += ext->imp[i] * w;
        om[i] += ext->ang[i] * w_rot;
        double ka = 0.5 * (double)bl.mass * (double)v[i].len2() + 0.5 * (double)bl.inertia * (double)(om[i] * om[i]);
        cw += ka - kb;
      }
      x[i] += v[i] * dt;
      th[i] += om[i] * dt;
    }
    A e_after_predict = balls_mech_j();
    R maxd = 0, maxs = 0;
    for (size_t i = 0; i < n; i++) maxd = rmax(maxd, v[i].len() * dt / rmax(F(2.0) * r, F(1e-12)));
    for (size_t i = 0; i < n; i++) maxs = rmax(maxs, v[i].len());
    R margin = (R)std::max(iterations, 1u) * F(0.2) * F(2.0) * r;
    R cell = rmax(rmax(F(2.0) * r * F(1.05), F(2.0) * r + maxs * dt + margin), F(1e-6));
    Grid<R> g;
    g.build(x, cell);
    std::vector<std::pair<uint32_t, uint32_t>> pairs;
    g.candidate_pairs([&](uint32_t i, uint32_t j) { pairs.push_back({i, j}); });
    ds.n_pairs = (uint32_t)pairs.size();
    R an = angle + drum.omega * dt;
    std::vector<R> toi(n, (R)1);
    for (size_t i = 0; i < n; i++) {
      R ce = drum.sdf_world(x[i], an).first - r;
      if (ce >= -F(0.2) * F(2.0) * r) continue;
      R t;
      if (drum.toi_swept(x0[i], x[i], angle, an, r, t)) toi[i] = rmin(toi[i], t);
    }
    for (size_t i = 0; i < n; i++) {
      if (toi[i] < (R)1) {
        R pl = (x[i] - x0[i]).len();
        R skin = F(1e-2) * r;
        R ts = pl > F(1e-9) ? skin / pl : (R)1;
        R tc = rmin(toi[i] + ts, (R)1);
        x[i] = x0[i] + (x[i] - x0[i]) * tc;
      }
    }
    std::vector<R> lam_bb(pairs.size(), 0), lam_bw(n, 0);
    for (uint32_t it = 0; it < std::max(iterations, 1u); it++) {
      for (size_t k = 0; k < pairs.size(); k++) {
        uint32_t i = pairs[k].first, j = pairs[k].second;
        V2<R> d = x[i] - x[j];
        R dist = d.len();
        R c = dist - F(2.0) * r;
        if (c >= (R)0) continue;
        V2<R> nh = dist > F(1e-9) ? d / dist : V2<R>(1, 0);
        R ws = w + w;
        if (ws <= (R)0) continue;
        R ce = rmax(c, -F(0.2) * F(2.0) * r);
        R dl = -ce / ws;
        x[i] += w * dl * nh;
        x[j] -= w * dl * nh;
        lam_bb[k] += dl;
      }
      for (size_t i = 0; i < n; i++) {
        auto dn = drum.sdf_world(x[i], an);
        R c = dn.first - r;
        if (c >= (R)0) continue;
        if (w <= (R)0) continue;
        R ce = rmax(c, -F(0.2) * F(2.0) * r);
        R dl = -ce / w;
        x[i] += w * dl * dn.second;
        lam_bw[i] += dl;
      }
    }
    for (size_t i = 0; i < n; i++) { v[i] = (x[i] - x0[i]) / dt; om[i] = angle_diff(th[i], th0[i]) / dt; }
    // ContactBook::sorted_ball_ball (dem.rs:459-467)
    struct BB { uint32_t i, j; R lam; };
    std::vector<BB> bb(pairs.size());
    for (size_t k = 0; k < pairs.size(); k++) bb[k] = {pairs[k].first, pairs[k].second, lam_bb[k]};
    std::sort(bb.begin(), bb.end(), [](const BB& a, const BB& b) { return a.i != b.i ? a.i < b.i : a.j < b.j; });
    // friction (dem.rs:762-890)
    for (auto& c : bb) {
      if (c.lam <= (R)0) continue;
      uint32_t i = c.i, j = c.j;
      V2<R> d = x[i] - x[j];
      R dist = d.len();
      V2<R> nh = dist > F(1e-9) ? d / dist : V2<R>(1, 0);
      V2<R> th_ = V2<R>(-nh.y, nh.x);
      R vt = (v[i] - v[j]).dot(th_) - r * om[i] - r * om[j];
      R wst = w + w + r * r * w_rot + r * r * w_rot;
      if (wst <= (R)0) continue;
      R raw = -vt * dt / wst;
      R mu = eff_friction((R)P.fric_bb, (R)P.fric_bb_static, vt, (R)P.fric_vscale);
      R mm = mu * c.lam;
      R dlt = rclamp(raw, -mm, mm);
      x[i] += w * dlt * th_;
      x[j] -= w * dlt * th_;
      th[i] -= r * w_rot * dlt;
      th[j] -= r * w_rot * dlt;
      V2<R> dv = w * dlt * th_ / dt;
      R dom = -r * w_rot * dlt / dt;
      v[i] += dv; v[j] -= dv;
      om[i] += dom; om[j] += dom;
    }
    A ww = 0;
    for (size_t i = 0; i < n; i++) {
      R lam = lam_bw[i];
      if (lam <= (R)0) continue;
      V2<R> nh = drum.sdf_world(x[i], an).second;
      V2<R> th_ = V2<R>(-nh.y, nh.x);
      V2<R> vw = drum.wall_velocity(x[i] - r * nh);
      R vt = (v[i] - vw).dot(th_) - r * om[i];
      R wst = w + r * r * w_rot;
      if (wst <= (R)0) continue;
      R raw = -vt * dt / wst;
      R mu = eff_friction((R)P.fric_bw, (R)P.fric_bw_static, vt, (R)P.fric_vscale);
      R mm = mu * lam;
      R dlt = rclamp(raw, -mm, mm);
      x[i] += w * dlt * th_;
      th[i] -= r * w_rot * dlt;
      v[i] += w * dlt * th_ / dt;
      om[i] += -r * w_rot * dlt / dt;
      ww += (A)((dlt / dt) * th_.dot(vw));
      ww += (A)((lam / dt) * nh.dot(vw));
    }
    for (size_t i = 0; i < n; i++) { v[i] = (x[i] - x0[i]) / dt; om[i] = angle_diff(th[i], th0[i]) / dt; }
    // restitution + impacts (dem.rs:907-1019)
    uint32_t coll = 0;
    for (auto& c : bb) {
      if (c.lam <= (R)0) continue;
      uint32_t i = c.i, j = c.j;
      V2<R> d = x[i] - x[j];
      R dist = d.len();
      V2<R> nh = dist > F(1e-9) ? d / dist : V2<R>(1, 0);
      R vnp = (v_pre[i] - v_pre[j]).dot(nh);
      bool fresh = vnp < -F(0.02);
      if (fresh) {
        coll++;
        R ie = F(0.5) * (bl.mass * F(0.5)) * vnp * vnp;
        int bin = impact_bin((float)ie);
        if (bin >= 0) ds.hist[bin]++;
      }
      R e = fresh ? (R)P.rest_bb : (R)0;
      if (-vnp > thr) {
        V2<R> th_ = V2<R>(-nh.y, nh.x);
        R slip = (v_pre[i] - v_pre[j]).dot(th_) - r * om_pre[i] - r * om_pre[j];
        impacts.push_back({0, i, j, (float)-vnp, (float)std::fabs(slip), (float)e});
      }
      R target = rmax(rmax(-e * vnp, (R)0), rmax(vnp, (R)0));
      R vnn = (v[i] - v[j]).dot(nh);
      R dvn = target - vnn;
      R ws = w + w;
      if (ws <= (R)0) continue;
      R im = dvn / ws;
      v[i] += w * im * nh;
      v[j] -= w * im * nh;
    }
    for (size_t i = 0; i < n; i++) {
      R lam = lam_bw[i];
      if (lam <= (R)0) continue;
      V2<R> nh = drum.sdf_world(x[i], an).second;
      V2<R> vw = drum.wall_velocity(x[i] - r * nh);
      R vnp = (v_pre[i] - vw).dot(nh);
      bool fresh = vnp < -F(0.02);
      if (fresh) {
        coll++;
        R ie = F(0.5) * bl.mass * vnp * vnp;
        int bin = impact_bin((float)ie);
        if (bin >= 0) ds.hist[bin]++;
      }
      R e = fresh ? (R)P.rest_bw : (R)0;
      if (-vnp > thr) {
        V2<R> th_ = V2<R>(-nh.y, nh.x);
        R slip = (v_pre[i] - vw).dot(th_) - r * om_pre[i];
        impacts.push_back({1, (uint32_t)i, UINT32_MAX, (float)-vnp, (float)std::fabs(slip), (float)e});
      }
      R target = rmax(rmax(-e * vnp, (R)0), rmax(vnp, (R)0));
      R vnn = (v[i] - vw).dot(nh);
      R dvn = target - vnn;
      if (w <= (R)0) continue;
      v[i] += dvn * nh;
    }
    // rolling resistance (dem.rs:1036-1055)
    if (P.rolling > 0.0f && bl.inertia > (R)0) {
      std::vector<R> tl(n, 0);
      for (auto& c : bb) { tl[c.i] += rmax(c.lam, (R)0); tl[c.j] += rmax(c.lam, (R)0); }
      for (size_t i = 0; i < n; i++) tl[i] += rmax(lam_bw[i], (R)0);
      for (size_t i = 0; i < n; i++) {
        if (tl[i] <= (R)0 || om[i] == (R)0) continue;
        R nf = tl[i] / dt;
        R mdo = (R)P.rolling * nf * r / bl.inertia * dt;
        R dom = rmin(mdo, std::fabs(om[i]));
        om[i] -= signum(om[i]) * dom;
      }
    }
    A e_final = balls_mech_j();
    ds.wall_work = (double)(R)ww;
    ds.collisions = coll;
    ds.dissipated = (double)(R)(e_after_predict + ww - e_final);
    ds.max_disp = (float)maxd;
    ds.coupling_work = (double)(R)cw;
  }
  static int impact_bin(float e) {  // dem.rs:99-109
    if (!std::isfinite(e) || e <= 0.0f) return -1;
    float ec = rclamp(e, 1e-6f, 1.0f);
    float lmin = std::log10(1e-6f), lmax = std::log10(1.0f);
    float t = (std::log10(ec) - lmin) / (lmax - lmin);
    size_t idx = (size_t)sat_u64(t * 12.0f);
    return (int)std::min(idx, (size_t)11);
  }

  // lib.rs:470-565 — one substep plus the accumulators.
  void advance(R sub_dt) {
    R omega = (R)P.omega();
    Drum<R> d = drum();
    Coupling<R> cp = fluid_step(d, drum_angle, P.pbf_iterations, sub_dt);
    dem_step(d, drum_angle, P.dem_iterations, sub_dt, &cp);
    clamp_hits_total += fs.clamp_hits;
    eb.elapsed += (double)sub_dt;
    eb.dem_wall += ds.wall_work;
    eb.fluid_wall += fs.wall_work;
    eb.shaft += (double)(R)((R)ds.wall_work + (R)fs.wall_work);
    eb.contact_diss += ds.dissipated;
    eb.wall_slip += (double)((R)fs.wall_work - (R)fs.ke_noslip);
    eb.viscous -= (double)(R)fs.ke_visc;
    eb.clamp_removed -= (double)(R)fs.ke_clamp;
    eb.interface += (double)((R)ds.coupling_work + (R)fs.ke_drag + (R)fs.ke_buoy);
    eb.fluid_wall_ang += (double)(R)fs.wall_ang;
    // update_grinding_stats (lib.rs:516-565)
    R alpha = F(1.0) - std::exp(-sub_dt / F(1.0));
    R pinst = ((R)ds.wall_work + (R)fs.wall_work) / sub_dt;
    power += alpha * (pinst - power);
    R tinst = std::fabs(omega) > F(1e-6) ? pinst / omega : (R)0;
    torque += alpha * (tinst - torque);
    double rr = (double)bl.radius;
    double m3 = (double)bl.mass * 4.0 / 3.0 * rr;
    double dd = 2.0 * rr;
    stress.update(impacts, m3, m3 * 9.81 * dd, dd, sim_time, (double)sub_dt, (double)alpha,
                  (double)((R)ds.wall_work + (R)fs.wall_work));
    coll_rate = (R)stress.total_rate();
    R dpi = rmax((R)ds.dissipated, (R)0) / sub_dt;
    diss_power += alpha * (dpi - diss_power);
    max_disp = rmax(max_disp, (R)ds.max_disp);
    drum_angle = rem_euclid(drum_angle + omega * sub_dt, C<R>::TAU);
    sim_time += (double)sub_dt;
  }
  R fixed_sub_dt() const { return F(1.0) / (F(60.0) * (R)std::max(P.effective_substeps(), 1u)); }
  void step_fixed() { advance(fixed_sub_dt()); }
  void reset_frame_stats() { max_disp = 0; clamp_hits_total = 0; }
  void step(R dt) {
    uint32_t ns = std::max(P.effective_substeps(), 1u);
    R sd = dt / (R)ns;
    reset_frame_stats();
    for (uint32_t k = 0; k < ns; k++) advance(sd);
  }

  // metrics.rs:506-576
  bool mixing_index(float& out) const {  // metrics.rs:506-576
    return orc::mixing_index_of<R>(fl.x.data(), fl.dye.data(), fl.n(), (R)P.radius_m(), out);
  }
};

}  // namespace orc

// ------------------------------------------------------------------------------------------
// C API for the Python test harness (ctypes). Handles are opaque; arrays are caller-owned.
// Modes: 0 = "ref" (float, float), 1 = "mixed" (float, double), 2 = "fp64" (double, double).
// ------------------------------------------------------------------------------------------
using namespace orc;
struct OrcHandle {
  int mode;
  void* p;
};
template <class S> static S* as(OrcHandle* h) { return (S*)h->p; }
#define DISPATCH(h, ...)                                                       \
  do {                                                                         \
    if ((h)->mode == 0) { auto* s = as<Sim<float, float>>(h); __VA_ARGS__; }   \
    else if ((h)->mode == 1) { auto* s = as<Sim<float, double>>(h); __VA_ARGS__; } \
    else { auto* s = as<Sim<double, double>>(h); __VA_ARGS__; }               \
  } while (0)

static Params params_from_flat(const double* f, uint64_t seed) {
  Params p;
  int k = 0;
  p.mill_diameter_m = (float)f[k++]; p.speed_mode = (int)f[k++]; p.speed_value = (float)f[k++]; p.direction = (int)f[k++];
  p.ball_diameter_m = (float)f[k++]; p.media_fill = (float)f[k++]; p.packing = (float)f[k++]; p.media_density = (float)f[k++];
  p.rest_bb = (float)f[k++]; p.rest_bw = (float)f[k++]; p.fric_bb = (float)f[k++]; p.fric_bw = (float)f[k++];
  p.fric_bb_static = (float)f[k++]; p.fric_bw_static = (float)f[k++]; p.fric_vscale = (float)f[k++]; p.rolling = (float)f[k++];
  p.slurry_enabled = (int)f[k++]; p.slurry_fill = (float)f[k++]; p.slurry_density = (float)f[k++]; p.viscosity = (float)f[k++];
  p.rheology = (int)f[k++]; p.yield_stress = (float)f[k++]; p.wall_no_slip = (float)f[k++]; p.ball_no_slip = (float)f[k++];
  p.wettability = (float)f[k++]; p.surface_tension = (float)f[k++]; p.dye_pattern = (int)f[k++];
  p.lifter_count = (uint32_t)f[k++]; p.lifter_height = (float)f[k++]; p.lifter_base = (float)f[k++]; p.lifter_top = (float)f[k++];
  p.lifter_phase_deg = (float)f[k++];
  p.resolution = (uint32_t)f[k++]; p.substeps = (uint32_t)f[k++]; p.pbf_iterations = (uint32_t)f[k++]; p.dem_iterations = (uint32_t)f[k++];
  p.time_scale = (float)f[k++]; p.frame_budget_ms = (float)f[k++]; p.max_balls = (uint32_t)f[k++];
  p.cg_mode = (int)f[k++]; p.cg_k = (float)f[k++];
  p.seed = seed;
  return p;
}

extern "C" {
int orc_nparams() { return 41; }

void* orc_create(const double* flat, uint64_t seed, int mode, const float* init_balls, size_t n_init) {
  Params p = params_from_flat(flat, seed);
  OrcHandle* h = new OrcHandle{mode, nullptr};
  if (mode == 0) { auto* s = new Sim<float, float>(); s->init(p, init_balls, n_init); h->p = s; }
  else if (mode == 1) { auto* s = new Sim<float, double>(); s->init(p, init_balls, n_init); h->p = s; }
  else { auto* s = new Sim<double, double>(); s->init(p, init_balls, n_init); h->p = s; }
  return h;
}
void orc_destroy(void* hv) {
  OrcHandle* h = (OrcHandle*)hv;
  if (h->mode == 0) delete as<Sim<float, float>>(h);
  else if (h->mode == 1) delete as<Sim<float, double>>(h);
  else delete as<Sim<double, double>>(h);
  delete h;
}
// Derived sizes: [n_balls, n_fluid, effective_substeps, effective_resolution, true_count_rounded]
void orc_sizes(void* hv, double* out) {
  OrcHandle* h = (OrcHandle*)hv;
  DISPATCH(h, {
    out[0] = (double)s->bl.n(); out[1] = (double)s->fl.n();
    out[2] = (double)s->P.effective_substeps(); out[3] = (double)s->P.effective_fluid_resolution();
    out[4] = (double)rmax(std::round(s->P.true_ball_count()), 0.0f);
    out[5] = (double)s->fixed_sub_dt(); out[6] = (double)s->fl.h; out[7] = (double)s->fl.mass;
    out[8] = (double)s->bl.radius; out[9] = (double)s->bl.mass; out[10] = (double)s->bl.inertia;
    out[11] = (double)s->P.omega();
  });
}
void orc_step_fixed(void* hv) { OrcHandle* h = (OrcHandle*)hv; DISPATCH(h, s->step_fixed()); }
void orc_step(void* hv, double dt) { OrcHandle* h = (OrcHandle*)hv; DISPATCH(h, s->step((decltype(s->drum_angle))dt)); }
void orc_advance(void* hv, double sub_dt) { OrcHandle* h = (OrcHandle*)hv; DISPATCH(h, s->advance((decltype(s->drum_angle))sub_dt)); }
void orc_set_debug(void* hv, int keep) { OrcHandle* h = (OrcHandle*)hv; DISPATCH(h, s->keep_debug = keep != 0); }
void orc_force_iters(void* hv, int on, uint32_t nd, uint32_t ncg) {
  OrcHandle* h = (OrcHandle*)hv;
  DISPATCH(h, { s->force_iters = on != 0; s->forced_density = nd; s->forced_cg = ncg; });
}
// State as doubles: balls x(2n) v(2n) theta(n) omega(n); fluid x(2m) v(2m) dye(m) guess(2m, zero if empty)
void orc_get_state(void* hv, double* bx, double* bv, double* bth, double* bom, double* fx, double* fv, double* dye,
                   double* guess, int* has_guess) {
  OrcHandle* h = (OrcHandle*)hv;
  DISPATCH(h, {
    for (size_t i = 0; i < s->bl.n(); i++) {
      bx[2 * i] = s->bl.x[i].x; bx[2 * i + 1] = s->bl.x[i].y;
      bv[2 * i] = s->bl.v[i].x; bv[2 * i + 1] = s->bl.v[i].y;
      bth[i] = s->bl.theta[i]; bom[i] = s->bl.omega[i];
    }
    for (size_t i = 0; i < s->fl.n(); i++) {
      fx[2 * i] = s->fl.x[i].x; fx[2 * i + 1] = s->fl.x[i].y;
      fv[2 * i] = s->fl.v[i].x; fv[2 * i + 1] = s->fl.v[i].y;
      dye[i] = s->fl.dye[i];
      bool hg = s->fl.guess.size() == s->fl.n();
      guess[2 * i] = hg ? (double)s->fl.guess[i].x : 0.0; guess[2 * i + 1] = hg ? (double)s->fl.guess[i].y : 0.0;
    }
    *has_guess = s->fl.guess.size() == s->fl.n() && s->fl.n() > 0;
  });
}
void orc_set_state(void* hv, const double* bx, const double* bv, const double* bth, const double* bom,
                   const double* fx, const double* fv, const double* guess, int has_guess) {
  OrcHandle* h = (OrcHandle*)hv;
  DISPATCH(h, {
    using RR = decltype(s->drum_angle);
    for (size_t i = 0; i < s->bl.n(); i++) {
      s->bl.x[i] = V2<RR>((RR)bx[2 * i], (RR)bx[2 * i + 1]);
      s->bl.v[i] = V2<RR>((RR)bv[2 * i], (RR)bv[2 * i + 1]);
      s->bl.theta[i] = (RR)bth[i]; s->bl.omega[i] = (RR)bom[i];
    }
    for (size_t i = 0; i < s->fl.n(); i++) {
      s->fl.x[i] = V2<RR>((RR)fx[2 * i], (RR)fx[2 * i + 1]);
      s->fl.v[i] = V2<RR>((RR)fv[2 * i], (RR)fv[2 * i + 1]);
    }
    if (has_guess) {
      s->fl.guess.resize(s->fl.n());
      for (size_t i = 0; i < s->fl.n(); i++) s->fl.guess[i] = V2<RR>((RR)guess[2 * i], (RR)guess[2 * i + 1]);
    } else {
      s->fl.guess.clear();
    }
  });
}
void orc_set_scalars(void* hv, double drum_angle, double sim_time) {
  OrcHandle* h = (OrcHandle*)hv;
  DISPATCH(h, { s->drum_angle = (decltype(s->drum_angle))drum_angle; s->sim_time = sim_time; });
}
// Scalars: [drum_angle, sim_time, power, torque, coll_rate, diss_power, max_disp, clamp_hits,
//  visc_iters, density_iters, mean_shear, fluid_wall_work, dem_wall_work, dem_dissipated,
//  dem_coupling_work, collisions, n_groups, n_links, links_in_range, n_pairs, ke_noslip, ke_drag,
//  ke_buoy, ke_visc, ke_clamp, wall_imp_x, wall_imp_y, wall_ang]
void orc_scalars(void* hv, double* o) {
  OrcHandle* h = (OrcHandle*)hv;
  DISPATCH(h, {
    o[0] = s->drum_angle; o[1] = s->sim_time; o[2] = s->power; o[3] = s->torque; o[4] = s->coll_rate;
    o[5] = s->diss_power; o[6] = s->max_disp; o[7] = s->clamp_hits_total; o[8] = s->fs.visc_iters;
    o[9] = s->fs.density_iters; o[10] = s->fs.mean_shear; o[11] = s->fs.wall_work; o[12] = s->ds.wall_work;
    o[13] = s->ds.dissipated; o[14] = s->ds.coupling_work; o[15] = s->ds.collisions; o[16] = s->fs.n_groups;
    o[17] = s->fs.n_links; o[18] = s->fs.links_in_range; o[19] = s->ds.n_pairs; o[20] = s->fs.ke_noslip;
    o[21] = s->fs.ke_drag; o[22] = s->fs.ke_buoy; o[23] = s->fs.ke_visc; o[24] = s->fs.ke_clamp;
    o[25] = s->fs.wall_imp[0]; o[26] = s->fs.wall_imp[1]; o[27] = s->fs.wall_ang;
  });
}
void orc_energy_budget(void* hv, double* o) {
  OrcHandle* h = (OrcHandle*)hv;
  DISPATCH(h, {
    EnergyBudget b = s->eb;
    b.delta_mech = s->total_mech() - s->e0;
    double expl = b.shaft - b.contact_diss - b.wall_slip - b.viscous - b.clamp_removed + b.interface;
    b.unattributed = b.delta_mech - expl;
    double v[12] = {b.elapsed, b.shaft, b.dem_wall, b.fluid_wall, b.contact_diss, b.wall_slip, b.viscous,
                    b.clamp_removed, b.interface, b.delta_mech, b.unattributed, b.fluid_wall_ang};
    std::memcpy(o, v, sizeof(v));
  });
}
// Stress accumulator: [count_rate 4, se_rate 4, hist_rate 28, above_rate 6, cum_wall, cum_above 6]
void orc_stress(void* hv, double* o) {
  OrcHandle* h = (OrcHandle*)hv;
  DISPATCH(h, {
    int k = 0;
    for (int c = 0; c < 4; c++) o[k++] = s->stress.count_rate[c];
    for (int c = 0; c < 4; c++) o[k++] = s->stress.se_rate[c];
    for (int c = 0; c < 28; c++) o[k++] = s->stress.hist_rate[c];
    for (int c = 0; c < 6; c++) o[k++] = s->stress.above_rate[c];
    o[k++] = s->stress.cum_wall;
    for (int c = 0; c < 6; c++) o[k++] = s->stress.cum_above[c];
  });
}
// A standalone StressAccum (stress.rs:83-254) for the accumulator's own checks.
void* orc_stress_new() { return new Stress(); }
void orc_stress_free(void* h) { delete (Stress*)h; }
void orc_stress_set_factor(void* h, double f) { ((Stress*)h)->factor = f; }
// One update with n records: kind (0 ball-ball, 1 ball-wall), ball, partner (UINT32_MAX: wall),
// normal and slip speeds.
void orc_stress_update(void* h, const int32_t* kind, const uint32_t* ball, const uint32_t* partner, const float* vn,
                       const float* slip, size_t n, double m3, double scale_j, double d, double now, double sub_dt,
                       double alpha, double wall_work) {
  std::vector<Impact> r(n);
  for (size_t i = 0; i < n; i++) r[i] = Impact{kind[i], ball[i], partner[i], vn[i], slip[i], 0.0f};
  ((Stress*)h)->update(r, m3, scale_j, d, now, sub_dt, alpha, wall_work);
}
void orc_stress_state(void* h, double* o) {
  const Stress* st = (const Stress*)h;
  int k = 0;
  for (int c = 0; c < 4; c++) o[k++] = st->count_rate[c];
  for (int c = 0; c < 4; c++) o[k++] = st->se_rate[c];
  for (int c = 0; c < 28; c++) o[k++] = st->hist_rate[c];
  for (int c = 0; c < 6; c++) o[k++] = st->above_rate[c];
  o[k++] = st->cum_wall;
  for (int c = 0; c < 6; c++) o[k++] = st->cum_above[c];
}
void orc_stress_metrics(void* h, size_t n_balls, double m3, double scale_j, double d, double power, double* o) {
  ((const Stress*)h)->metrics(n_balls, m3, scale_j, d, power, o);
}
size_t orc_se_bin(double e) { return Stress::se_bin(e); }
int orc_impacts(void* hv, double* o, int cap) {
  OrcHandle* h = (OrcHandle*)hv;
  int n = 0;
  DISPATCH(h, {
    n = (int)s->impacts.size();
    for (int i = 0; i < n && i < cap; i++) {
      auto& r = s->impacts[i];
      o[6 * i] = r.kind; o[6 * i + 1] = r.ball; o[6 * i + 2] = r.partner == UINT32_MAX ? -1.0 : r.partner;
      o[6 * i + 3] = r.vn; o[6 * i + 4] = r.slip; o[6 * i + 5] = r.e;
    }
  });
  return n;
}
// DemState::step_recorded (dem.rs:530-1066) on the handle's balls alone, at drum angle `angle`;
// ext_imp (2 floats per ball) / ext_ang may be null (no coupling impulses). stats[18]: wall work,
// dissipated energy, coupling work, collisions, impact histogram 12, max displacement over
// diameter, candidate pairs. The impacts are read with orc_impacts.
void orc_dem_step(void* hv, double angle, uint32_t iterations, double dt, const float* ext_imp, const float* ext_ang,
                  double* stats) {
  OrcHandle* h = (OrcHandle*)hv;
  DISPATCH(h, {
    using RR = std::remove_reference<decltype(s->drum_angle)>::type;
    const size_t nb = s->bl.n();
    Coupling<RR> cp;
    const Coupling<RR>* ext = nullptr;
    if (ext_imp && ext_ang) {
      cp.imp.resize(nb);
      cp.ang.resize(nb);
      for (size_t i = 0; i < nb; i++) {
        cp.imp[i] = V2<RR>((RR)ext_imp[2 * i], (RR)ext_imp[2 * i + 1]);
        cp.ang[i] = (RR)ext_ang[i];
      }
      ext = &cp;
    }
    s->dem_step(s->drum(), (RR)angle, iterations, (RR)dt, ext);
    int k = 0;
    stats[k++] = s->ds.wall_work; stats[k++] = s->ds.dissipated; stats[k++] = s->ds.coupling_work;
    stats[k++] = s->ds.collisions;
    for (int b = 0; b < 12; b++) stats[k++] = s->ds.hist[b];
    stats[k++] = s->ds.max_disp; stats[k++] = s->ds.n_pairs;
  });
}
// The f32 values the DEM step of this handle uses: drum radius, omega, lifter count, height, base,
// top, phase (deg), ball radius, mass, inertia, restitution bb / bw, friction bb / bw kinetic,
// bb / bw static, friction velocity scale, rolling friction (18 values).
void orc_dem_params(void* hv, double* o) {
  OrcHandle* h = (OrcHandle*)hv;
  DISPATCH(h, {
    const Params& P = s->P;
    double v[18] = {P.radius_m(), P.omega(), (double)P.lifter_count, P.lifter_height, P.lifter_base, P.lifter_top,
                    P.lifter_phase_deg, (double)s->bl.radius, (double)s->bl.mass, (double)s->bl.inertia, P.rest_bb,
                    P.rest_bw, P.fric_bb, P.fric_bw, P.fric_bb_static, P.fric_bw_static, P.fric_vscale, P.rolling};
    std::memcpy(o, v, sizeof v);
  });
}
// FluidParticles::step_coupled (pbf.rs:1174-2365) on the handle's fluid alone, at drum angle
// `angle`, with the handle's balls as they are (unchanged); iterations = pbf_iterations. Writes the
// coupling impulses (imp: 2 per ball, ang: 1 per ball) and the FluidStepStats as 50 doubles:
// visc_iters, density_iters, mean_shear, wall_work, wall_imp x, y, wall_ang, ke_noslip, ke_drag,
// ke_buoy, ke_visc, ke_clamp, clamp_hits, n_groups, n_links, links_in_range, has_guess, 0, bins 32.
void orc_fluid_step(void* hv, double angle, uint32_t iterations, double dt, double* imp, double* ang,
                    double* stats) {
  OrcHandle* h = (OrcHandle*)hv;
  DISPATCH(h, {
    using RR = std::remove_reference<decltype(s->drum_angle)>::type;
    Coupling<RR> cp = s->fluid_step(s->drum(), (RR)angle, iterations, (RR)dt);
    for (size_t i = 0; i < cp.imp.size(); i++) {
      imp[2 * i] = (double)cp.imp[i].x;
      imp[2 * i + 1] = (double)cp.imp[i].y;
      ang[i] = (double)cp.ang[i];
    }
    const FluidStats& f = s->fs;
    double v[18] = {(double)f.visc_iters, (double)f.density_iters, f.mean_shear, f.wall_work, f.wall_imp[0],
                    f.wall_imp[1], f.wall_ang, f.ke_noslip, f.ke_drag, f.ke_buoy, f.ke_visc, f.ke_clamp,
                    (double)f.clamp_hits, (double)f.n_groups, (double)f.n_links, (double)f.links_in_range,
                    s->fl.guess.size() == s->fl.n() && s->fl.n() > 0 ? 1.0 : 0.0, 0.0};
    std::memcpy(stats, v, sizeof v);
    for (int b = 0; b < 32; b++) stats[18 + b] = f.bins[b];
  });
}
// The f32 values the fluid step of this handle reads: h, rest density, particle mass, ball radius,
// mass, inertia, true radius, viscosity, rheology, yield stress, wall no-slip, ball no-slip,
// wettability, surface tension, pbf iterations (15 values).
void orc_fluid_params(void* hv, double* o) {
  OrcHandle* h = (OrcHandle*)hv;
  DISPATCH(h, {
    const Params& P = s->P;
    double v[15] = {(double)s->fl.h, (double)s->fl.rest, (double)s->fl.mass, (double)s->bl.radius,
                    (double)s->bl.mass, (double)s->bl.inertia, (double)s->bl.true_radius, P.viscosity,
                    (double)P.rheology, P.yield_stress, P.wall_no_slip, P.ball_no_slip, P.wettability,
                    P.surface_tension, (double)P.pbf_iterations};
    std::memcpy(o, v, sizeof v);
  });
}
void orc_fluid_bins(void* hv, double* o) {
  OrcHandle* h = (OrcHandle*)hv;
  DISPATCH(h, for (int b = 0; b < 32; b++) o[b] = s->fs.bins[b]);
}
int orc_mixing_index(void* hv, double* out) {
  OrcHandle* h = (OrcHandle*)hv;
  int ok = 0;
  DISPATCH(h, { float m = 0; ok = s->mixing_index(m) ? 1 : 0; *out = m; });
  return ok;
}
// Debug: the step-2 fluid grid / neighbour CSR of the last substep (keep_debug must be set).
int orc_dbg_grid(void* hv, int64_t* meta, uint32_t* cell_start, int cs_cap, uint32_t* indices, int idx_cap) {
  OrcHandle* h = (OrcHandle*)hv;
  int ok = 0;
  DISPATCH(h, {
    auto& g = s->dbg_grid;
    meta[0] = g.dense; meta[1] = g.min_cx; meta[2] = g.min_cy; meta[3] = g.W; meta[4] = g.H;
    meta[5] = (int64_t)g.cell_start.size(); meta[6] = (int64_t)g.indices.size();
    if ((int)g.cell_start.size() <= cs_cap && (int)g.indices.size() <= idx_cap) {
      std::memcpy(cell_start, g.cell_start.data(), g.cell_start.size() * 4);
      std::memcpy(indices, g.indices.data(), g.indices.size() * 4);
      ok = 1;
    }
  });
  return ok;
}
int orc_dbg_nbrs(void* hv, uint32_t* off, int off_cap, uint32_t* nb, int nb_cap, int64_t* sizes) {
  OrcHandle* h = (OrcHandle*)hv;
  int ok = 0;
  DISPATCH(h, {
    auto& L = s->dbg_nbrs;
    sizes[0] = (int64_t)L.off.size(); sizes[1] = (int64_t)L.nb.size();
    if ((int)L.off.size() <= off_cap && (int)L.nb.size() <= nb_cap) {
      std::memcpy(off, L.off.data(), L.off.size() * 4);
      std::memcpy(nb, L.nb.data(), L.nb.size() * 4);
      ok = 1;
    }
  });
  return ok;
}

// Standalone grid builder over caller points (float policy), for grid-level tests.
int orc_grid_build(const float* pts, size_t n, float cell, int64_t* meta, uint32_t* cell_start, int cs_cap,
                   uint32_t* indices, int idx_cap) {
  std::vector<V2<float>> p(n);
  for (size_t i = 0; i < n; i++) p[i] = V2<float>(pts[2 * i], pts[2 * i + 1]);
  Grid<float> g;
  g.build(p, cell);
  meta[0] = g.dense; meta[1] = g.min_cx; meta[2] = g.min_cy; meta[3] = g.W; meta[4] = g.H;
  meta[5] = (int64_t)g.cell_start.size(); meta[6] = (int64_t)g.indices.size();
  if ((int)g.cell_start.size() > cs_cap || (int)g.indices.size() > idx_cap) return 0;
  std::memcpy(cell_start, g.cell_start.data(), g.cell_start.size() * 4);
  std::memcpy(indices, g.indices.data(), g.indices.size() * 4);
  return 1;
}
// Sparse-layout keys (cx, cy) of the occupied cells, in cell order (grid.rs:204-226); returns
// their count (0 for the Dense layout), copying them when they fit.
int64_t orc_grid_keys(const float* pts, size_t n, float cell, int32_t* keys2, int64_t cap) {
  std::vector<V2<float>> p(n);
  for (size_t i = 0; i < n; i++) p[i] = V2<float>(pts[2 * i], pts[2 * i + 1]);
  Grid<float> g;
  g.build(p, cell);
  const int64_t k = (int64_t)g.keys.size();
  if (k <= cap)
    for (int64_t i = 0; i < k; i++) { keys2[2 * i] = g.keys[i].first; keys2[2 * i + 1] = g.keys[i].second; }
  return k;
}
// for_each_near (grid.rs:371-399) of nq query points over the grid of pts: CSR of the candidates
// in visiting order; returns the total, copying the entries when they fit.
int64_t orc_grid_near(const float* pts, size_t n, float cell, const float* q, size_t nq, uint32_t* off, uint32_t* out,
                      int64_t cap) {
  std::vector<V2<float>> p(n);
  for (size_t i = 0; i < n; i++) p[i] = V2<float>(pts[2 * i], pts[2 * i + 1]);
  Grid<float> g;
  g.build(p, cell);
  int64_t k = 0;
  off[0] = 0;
  for (size_t a = 0; a < nq; a++) {
    g.near(V2<float>(q[2 * a], q[2 * a + 1]), [&](uint32_t j) {
      if (k < cap) out[k] = j;
      k++;
    });
    off[a + 1] = (uint32_t)k;
  }
  return k;
}
// metrics::mixing_index over caller points and dye (f32); returns 0 for None.
int orc_mixing_index_of(const float* pts, const float* dye, size_t n, float radius, float* out) {
  std::vector<V2<float>> p(n);
  for (size_t i = 0; i < n; i++) p[i] = V2<float>(pts[2 * i], pts[2 * i + 1]);
  float m = 0.0f;
  const bool ok = mixing_index_of<float>(p.data(), dye, n, radius, m);
  *out = m;
  return ok ? 1 : 0;
}
// metrics::compute (metrics.rs:195-254) of caller arrays. drum7 = radius, omega, lifter count,
// height, base width, top width, phase (deg). out[26]: see oracle.py DRUM_METRICS_KEYS.
// Returns 0 where the reference would panic (a NaN angle in a sorted wall list).
int orc_drum_metrics(const double* drum7, float angle, const float* bx, const float* bv, const float* bw, size_t nb,
                     float br, float bmass, float binertia, const float* fx, const float* dye, const float* dens,
                     size_t nf, float h, float rest, double* out) {
  Drum<float> d;
  d.radius = (float)drum7[0]; d.omega = (float)drum7[1]; d.lifter_count = (uint32_t)drum7[2];
  d.l_height = (float)drum7[3]; d.l_base = (float)drum7[4]; d.l_top = (float)drum7[5]; d.l_phase_deg = (float)drum7[6];
  std::vector<V2<float>> b(nb), v(nb), f(nf);
  for (size_t i = 0; i < nb; i++) { b[i] = V2<float>(bx[2 * i], bx[2 * i + 1]); v[i] = V2<float>(bv[2 * i], bv[2 * i + 1]); }
  for (size_t i = 0; i < nf; i++) f[i] = V2<float>(fx[2 * i], fx[2 * i + 1]);
  const DrumMetrics m = drum_metrics(d, angle, b.data(), v.data(), bw, nb, br, bmass, binertia, f.data(), dye, dens, nf,
                                     h, rest);
  const double o[26] = {m.toe, m.shoulder, (double)m.has_toe_shoulder, m.cx, m.cy, (double)m.has_centroid,
                        m.pool_lo, m.pool_hi, (double)m.has_pool, m.fs_angle, m.fs_offset, (double)m.has_fs,
                        m.depth, (double)m.has_depth, m.mixing, (double)m.has_mixing, m.ke, m.max_overlap,
                        m.max_wall_overlap, m.max_derr, m.mean_derr, m.max_cerr, m.mean_cerr, (double)m.has_errors,
                        m.omega, (double)m.n_fluid};
  std::memcpy(out, o, sizeof o);
  return m.non_finite_angle ? 0 : 1;
}
// extract_surface (surface.rs:233-277) of caller points: the flatten_polygons layout
// (surface.rs:420-431) into out (returns its length, or the length needed when cap is short);
// field (128 * 128 floats) receives the masked splat field when non-null.
int64_t orc_surface(const double* drum7, float angle, const float* fx, size_t n, float h, float rest, float pmass,
                    float* out, int64_t cap, float* field) {
  Drum<float> d;
  d.radius = (float)drum7[0]; d.omega = (float)drum7[1]; d.lifter_count = (uint32_t)drum7[2];
  d.l_height = (float)drum7[3]; d.l_base = (float)drum7[4]; d.l_top = (float)drum7[5]; d.l_phase_deg = (float)drum7[6];
  std::vector<V2<float>> x(n);
  for (size_t i = 0; i < n; i++) x[i] = V2<float>(fx[2 * i], fx[2 * i + 1]);
  std::vector<float> f;
  const auto polys = surf::extract(x, d, angle, h, rest, pmass, field ? &f : nullptr);
  if (field) {
    if (f.empty()) f.assign((size_t)surf::G * surf::G, 0.0f);
    std::memcpy(field, f.data(), f.size() * sizeof(float));
  }
  std::vector<float> flat = {(float)polys.size()};
  for (const auto& p : polys) {
    flat.push_back((float)p.size());
    for (const auto& q : p) { flat.push_back(q.x); flat.push_back(q.y); }
  }
  if ((int64_t)flat.size() <= cap) std::memcpy(out, flat.data(), flat.size() * sizeof(float));
  return (int64_t)flat.size();
}
float orc_splat_kernel(float r2, float radius) { return surf::splat_kernel(r2, radius); }
// Candidate pairs in emission order (grid.rs:291-362).
int64_t orc_grid_pairs(const float* pts, size_t n, float cell, uint32_t* out, int64_t cap) {
  std::vector<V2<float>> p(n);
  for (size_t i = 0; i < n; i++) p[i] = V2<float>(pts[2 * i], pts[2 * i + 1]);
  Grid<float> g;
  g.build(p, cell);
  int64_t k = 0;
  g.candidate_pairs([&](uint32_t i, uint32_t j) {
    if (k < cap) { out[2 * k] = i; out[2 * k + 1] = j; }
    k++;
  });
  return k;
}
// Neighbour CSR over caller points (pbf.rs:262-288).
int orc_neighbor_lists(const float* pts, size_t n, float h, uint32_t* off, uint32_t* nb, int64_t nb_cap, int64_t* n_edges) {
  std::vector<V2<float>> p(n);
  for (size_t i = 0; i < n; i++) p[i] = V2<float>(pts[2 * i], pts[2 * i + 1]);
  Grid<float> g;
  g.build(p, h);
  Nbrs L = build_neighbor_lists(p, g, h);
  *n_edges = (int64_t)L.nb.size();
  std::memcpy(off, L.off.data(), L.off.size() * 4);
  if ((int64_t)L.nb.size() > nb_cap) return 0;
  std::memcpy(nb, L.nb.data(), L.nb.size() * 4);
  return 1;
}
// Rng stream (rng.rs) for known-answer checks.
void orc_rng(uint64_t seed, uint64_t* out_u64, float* out_f32, int n) {
  Rng r(seed), r2(seed);
  for (int i = 0; i < n; i++) { out_u64[i] = r.next_u64(); out_f32[i] = r2.next_f32(); }
}
// Drum SDF probe (geometry.rs) in float: returns d and the world normal.
void orc_sdf_world(const double* drum, double px, double py, double angle, double* out) {
  Drum<float> d;
  d.radius = (float)drum[0]; d.omega = (float)drum[1]; d.lifter_count = (uint32_t)drum[2];
  d.l_height = (float)drum[3]; d.l_base = (float)drum[4]; d.l_top = (float)drum[5]; d.l_phase_deg = (float)drum[6];
  auto r = d.sdf_world(V2<float>((float)px, (float)py), (float)angle);
  out[0] = r.first; out[1] = r.second.x; out[2] = r.second.y;
}
int orc_toi_swept(const double* drum, const double* x0, const double* x1, double a0, double a1, double radius, double* t) {
  Drum<float> d;
  d.radius = (float)drum[0]; d.omega = (float)drum[1]; d.lifter_count = (uint32_t)drum[2];
  d.l_height = (float)drum[3]; d.l_base = (float)drum[4]; d.l_top = (float)drum[5]; d.l_phase_deg = (float)drum[6];
  float tt = 0;
  bool ok = d.toi_swept(V2<float>((float)x0[0], (float)x0[1]), V2<float>((float)x1[0], (float)x1[1]), (float)a0, (float)a1,
                        (float)radius, tt);
  *t = tt;
  return ok ? 1 : 0;
}
double orc_squeeze_coefficient(double k1, double k2, double h, double nu) { return SqueezeFilm(k1, k2, nu).coefficient(h); }
double orc_resolved_ratio(double hc) { return resolved_ratio(hc); }
}

// Kernel-isolation workload (periodic box): see c3_oracle.cpp.
#include "c3_oracle.cpp"
