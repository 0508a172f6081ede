// DemState::step_recorded on the GPU for a batch of ball populations (crates/mill-core/src/dem.rs:530-1066).
//
// The reference advances the balls in this order: sanitize, predict (gravity, optional coupling
// impulses), a uniform grid over the predicted positions and its candidate pairs, a swept
// time-of-impact clamp for balls that end deep in the wall, Gauss-Seidel XPBD projection (pairs in
// the grid's emission order, then walls), velocities from positions, Coulomb friction (pairs
// sorted by (i, j), then walls), velocities again, restitution with impact records (pairs, then
// walls), rolling resistance, and the step's energies. Results depend on the order of the pair
// sweeps, so each population keeps the reference's order: one 256-lane workgroup per population,
// the per-ball stages spread over the lanes and the pair sweeps walked by lane 0 over LDS copies of
// the ball state, with the pair list streamed through LDS in chunks by all lanes. Many populations
// run side by side (one workgroup each): a batch is what fills the GPU.
//
// Kernels per step: dem_predict, dem_grid (grid.rs:101-362 restated in one workgroup: keys,
// bounding box, Dense or Sparse layout, stable counting sort, candidate pairs in emission order),
// dem_ccd_flag (balls whose end position lies deeper than 0.4 r in the wall), dem_project,
// dem_sort (the contact book sorted by (i, j)), dem_friction, dem_restitution, dem_finish (rolling
// resistance and energies). The swept time of impact of the flagged balls (geometry.rs:146-194,
// rare: a ball crossing the wall within one step) is evaluated on the host, because it rotates
// the drum to intermediate angles with the libm's sinf / cosf; the host writes the clamped
// positions back before dem_project.
//
// f32 sums the reference forms in ball order (energies, wall work) are formed in ball order by one
// lane over LDS-staged terms; the friction blend's exp is the host libm's expf algorithm restated
// (mg_expf), and the impact-energy bin (log10-based, dem.rs:99-109) is a count of 11 thresholds the
// host finds once with its own log10f, so device results equal the reference's f32 arithmetic.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/millgpu.h"
#include "mg_drumgeo.h"
#include "mg_expf.h"
#include "mg_internal.h"
#include "mg_ktimer.h"

#define DE_BLOCK 256
#define DE_MAX_BALLS 2048
#define DE_CHUNK 1024  // pairs staged in LDS per chunk of a sequential sweep
#define DE_NO 0xFFFFFFFFu

namespace {

constexpr float DE_PI = 3.14159265358979323846f;
constexpr float DE_G = -9.81f;

// One population's scalars for one step (host-prepared).
struct DeDrum {
  float radius, omega;
  uint32_t nl, lift0;
  float rt, bh, thw;
  float rc0, rs0;    // from_angle(-angle)
  float rc1, rs1;    // from_angle(-angle_next)
  float wc1, ws1;    // from_angle(angle_next)
  float r, mass, inertia, w, w_rot, thr;
  float rest_bb, rest_bw, fric_bb, fric_bw, fric_bb_s, fric_bw_s, fric_vs, rolling;
  uint32_t iters;
  float dt;
  uint32_t b0, n;    // first ball in the caller's arrays, ball count
  uint32_t has_ext;
  uint32_t f64;  // energies and wall work summed in f64 (the precision form) instead of f32
};

// Per population results of a step.
struct DeOut {
  double e_pred, ww, e_final;  // f32 sums (held exactly) or f64 sums, per DeDrum::f64
  float maxd, maxs;
  double cw;
  uint32_t n_pairs, pair_overflow, n_ccd, coll, n_imp;
  uint32_t hist[12];
  uint32_t depth;   // levels of the projection sweep's schedule (DE_LCAP + 1: sequential)
  uint32_t nact;    // active contacts (normal impulse > 0) of the contact book
  uint32_t cdepth;  // levels of the contact sweeps' schedule (DE_LCAP + 1: sequential)
  uint32_t restart; // the projection walked every candidate pair after the near pairs' travel check failed
};

struct DeItem {  // a sorted contact: (i, j) and its accumulated normal impulse
  uint32_t i, j;
  float lam;
};

__device__ __forceinline__ mgg::Geo de_geo(const DeDrum& D, const float2* lift) {
  return mgg::Geo{D.radius, D.omega, D.nl, D.rt, D.bh, D.thw, lift + D.lift0};
}

// angle_diff (dem.rs:1069-1078): a - b wrapped into [-pi, pi] by whole turns (bounded).
__device__ __forceinline__ float de_angle_diff(float a, float b) {
  float d = a - b;
  for (int k = 0; k < 64 && d > DE_PI; k++) d -= 2.0f * DE_PI;
  for (int k = 0; k < 64 && d < -DE_PI; k++) d += 2.0f * DE_PI;
  return d;
}

// impact_energy_bin (dem.rs:99-109) as a count of the host's 11 thresholds; -1 for no bin.
__device__ __forceinline__ int de_ebin(float e, const float* __restrict__ thr) {
  if (!isfinite(e) || e <= 0.0f) return -1;
  int b = 0;
  for (int k = 0; k < 11; k++) b += e >= thr[k] ? 1 : 0;
  return b;
}

// effective_friction (dem.rs:128-138).
__device__ inline float de_friction(float mk, float ms, float vt, float vs) {
  if (ms <= mk) return mk;
  const float ratio = vt / vs;
  return mk + (ms - mk) * mgm::ref_expf(-(ratio * ratio));
}

__device__ __forceinline__ int32_t de_sat_i32(float v) {
  if (isnan(v)) return 0;
  if (v <= -2147483648.0f) return INT32_MIN;
  if (v >= 2147483647.0f) return INT32_MAX;
  return (int32_t)v;
}
__device__ __forceinline__ int32_t de_sadd(int32_t a, int32_t b) {
  const int64_t r = (int64_t)a + b;
  return r > INT32_MAX ? INT32_MAX : (r < INT32_MIN ? INT32_MIN : (int32_t)r);
}
__device__ __forceinline__ float de_wave_max(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ int32_t de_wave_min_i(int32_t v) {
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ int32_t de_wave_max_i(int32_t v) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}

// Exclusive scan of one value per lane over the workgroup; *total gets the sum.
__device__ uint32_t de_block_scan(uint32_t v, uint32_t* s_w, uint32_t* total) {
  const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t inc = v;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t u = __shfl_up(inc, o, 64);
    if (lane >= (uint32_t)o) inc += u;
  }
  if (lane == 63) s_w[w] = inc;
  __syncthreads();
  uint32_t before = 0, tot = 0;
  for (uint32_t q = 0; q < DE_BLOCK / 64; q++) {
    before += q < w ? s_w[q] : 0u;
    tot += s_w[q];
  }
  __syncthreads();
  *total = tot;
  return before + inc - v;
}

// Mechanical energy of one ball's three terms, in the order mechanical_energy_j adds them
// (dem.rs:253-261).
__device__ __forceinline__ void de_mech_terms(const DeDrum& D, float2 x, float2 v, float om, float* t) {
  t[0] = 0.5f * D.mass * (v.x * v.x + v.y * v.y);
  t[1] = 0.5f * D.inertia * om * om;
  t[2] = D.mass * 9.81f * x.y;
}

}  // namespace

// ---- 1. sanitize + predict (dem.rs:555-598) -------------------------------------------------------
extern "C" __global__ void __launch_bounds__(DE_BLOCK)
dem_predict(const DeDrum* __restrict__ drums, float2* __restrict__ x, float2* __restrict__ v, float* __restrict__ th,
            float* __restrict__ om, const float2* __restrict__ ext_imp, const float* __restrict__ ext_ang,
            float2* __restrict__ x0, float2* __restrict__ vpre, float* __restrict__ th0, float* __restrict__ ompre,
            uint32_t stride, DeOut* __restrict__ out) {
  __shared__ __align__(16) float s_e[3 * DE_BLOCK];
  __shared__ double s_cw[DE_BLOCK];
  __shared__ float s_red[DE_BLOCK / 64];
  const DeDrum D = drums[blockIdx.x];
  const uint32_t t = threadIdx.x;
  const size_t sb = (size_t)blockIdx.x * stride;
  float e = 0.0f;   // lane 0: mechanical energy, three terms per ball in ball order
  double ed = 0.0;  // the same in f64 (DeDrum::f64)
  double cw = 0.0;  // lane 0: coupling work in ball order
  float maxd = 0.0f, maxs = 0.0f;
  for (uint32_t base = 0; base < D.n; base += DE_BLOCK) {
    const uint32_t i = base + t;
    if (i < D.n) {
      const uint32_t g = D.b0 + i;
      float2 xi = x[g], vi = v[g];
      float thi = th[g], omi = om[g];
      if (!(isfinite(xi.x) && isfinite(xi.y)) || !(isfinite(vi.x) && isfinite(vi.y))) {
        xi = make_float2(0.0f, 0.0f);
        vi = make_float2(0.0f, 0.0f);
      }
      x0[sb + i] = xi;
      vpre[sb + i] = vi;
      th0[sb + i] = thi;
      ompre[sb + i] = omi;
      vi.y += DE_G * D.dt;
      double c = 0.0;
      if (D.has_ext) {
        const double kb = 0.5 * (double)D.mass * (double)(vi.x * vi.x + vi.y * vi.y) +
                          0.5 * (double)D.inertia * (double)(omi * omi);
        const float2 im = ext_imp[g];
        vi.x += im.x * D.w;
        vi.y += im.y * D.w;
        omi += ext_ang[g] * D.w_rot;
        const double ka = 0.5 * (double)D.mass * (double)(vi.x * vi.x + vi.y * vi.y) +
                          0.5 * (double)D.inertia * (double)(omi * omi);
        c = ka - kb;
      }
      xi.x += vi.x * D.dt;
      xi.y += vi.y * D.dt;
      thi += omi * D.dt;
      x[g] = xi;
      v[g] = vi;
      th[g] = thi;
      om[g] = omi;
      const float sp = sqrtf(vi.x * vi.x + vi.y * vi.y);
      maxd = fmaxf(maxd, sp * D.dt / fmaxf(2.0f * D.r, 1e-12f));
      maxs = fmaxf(maxs, sp);
      de_mech_terms(D, xi, vi, omi, &s_e[3 * t]);
      s_cw[t] = c;
    }
    __syncthreads();
    if (t == 0) {
      const uint32_t m = min(D.n - base, (uint32_t)DE_BLOCK);
      if (D.f64)
        for (uint32_t q = 0; q < 3 * m; q++) ed += (double)s_e[q];
      else
        for (uint32_t q = 0; q < 3 * m; q++) e += s_e[q];
      if (D.has_ext)
        for (uint32_t q = 0; q < m; q++) cw += s_cw[q];
    }
    __syncthreads();
  }
  maxd = de_wave_max(maxd);
  maxs = de_wave_max(maxs);
  if ((t & 63) == 0) s_red[t >> 6] = maxd;
  __syncthreads();
  float md = s_red[0];
  for (int q = 1; q < DE_BLOCK / 64; q++) md = fmaxf(md, s_red[q]);
  __syncthreads();
  if ((t & 63) == 0) s_red[t >> 6] = maxs;
  __syncthreads();
  if (t == 0) {
    float ms = s_red[0];
    for (int q = 1; q < DE_BLOCK / 64; q++) ms = fmaxf(ms, s_red[q]);
    DeOut& o = out[blockIdx.x];
    o.e_pred = D.f64 ? ed : (double)e;
    o.cw = cw;
    o.maxd = md;
    o.maxs = ms;
  }
}

// ---- 2. grid + candidate pairs (grid.rs:101-362 over the predicted positions, dem.rs:600-608) ------
namespace {

struct DeGridLds {
  int2 key[DE_MAX_BALLS];              // cell key per ball (finite balls)
  uint8_t fin[DE_MAX_BALLS];           // 1: finite position
  uint32_t idx[DE_MAX_BALLS];          // ball indices in cell order
  uint32_t start[2 * DE_MAX_BALLS + 514];  // cell starts (Dense: W*H + 1; Sparse: cells + 1)
  int2 ckey[DE_MAX_BALLS];             // Sparse: occupied cells' keys in (cx, cy) order
  uint32_t w[DE_BLOCK / 64];
  int32_t box[4];
  uint32_t m;
};

__device__ __forceinline__ bool de_key_less(int2 a, int2 b) { return a.x < b.x || (a.x == b.x && a.y < b.y); }

// Appends the candidate pair (min, max) with a zero normal impulse; counts pairs beyond the capacity.
__device__ __forceinline__ void de_emit(uint2* P, float* Lm, uint32_t cap, uint32_t& np, uint32_t a, uint32_t b) {
  if (np < cap) {
    P[np] = make_uint2(min(a, b), max(a, b));
    Lm[np] = 0.0f;
  }
  np++;
}

// Range of cell (cx, cy) in idx; false if it is empty or absent (grid.rs:242-271).
__device__ bool de_range(const DeGridLds& L, bool dense, int32_t mnx, int32_t mny, int64_t W, int64_t H, uint32_t nc,
                         int32_t cx, int32_t cy, uint32_t& s, uint32_t& e) {
  if (dense) {
    const int64_t dx = (int64_t)cx - mnx, dy = (int64_t)cy - mny;
    if (dx < 0 || dy < 0 || dx >= W || dy >= H) return false;
    const uint32_t k = (uint32_t)(dx * H + dy);
    s = L.start[k];
    e = L.start[k + 1];
    return s != e;
  }
  uint32_t lo = 0, hi = nc;  // lower bound of (cx, cy) in the sorted keys
  const int2 q = make_int2(cx, cy);
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (de_key_less(L.ckey[mid], q)) lo = mid + 1; else hi = mid;
  }
  if (lo == nc || L.ckey[lo].x != cx || L.ckey[lo].y != cy) return false;
  s = L.start[lo];
  e = L.start[lo + 1];
  return true;
}

}  // namespace

extern "C" __global__ void __launch_bounds__(DE_BLOCK)
dem_grid(const DeDrum* __restrict__ drums, const float2* __restrict__ x, uint2* __restrict__ pairs,
         float* __restrict__ lam, uint32_t pcap, DeOut* __restrict__ out) {
  __shared__ DeGridLds L;
  const DeDrum D = drums[blockIdx.x];
  DeOut& o = out[blockIdx.x];
  const uint32_t t = threadIdx.x, n = D.n;
  const float margin = (float)max(D.iters, 1u) * 0.2f * 2.0f * D.r;
  const float cell = fmaxf(fmaxf(2.0f * D.r * 1.05f, 2.0f * D.r + o.maxs * D.dt + margin), 1e-6f);
  // keys and the bounding box of the finite points
  int32_t mnx = INT32_MAX, mny = INT32_MAX, mxx = INT32_MIN, mxy = INT32_MIN;
  uint32_t cnt = 0;
  for (uint32_t i = t; i < n; i += DE_BLOCK) {
    const float2 p = x[D.b0 + i];
    const bool f = isfinite(p.x) && isfinite(p.y);
    L.fin[i] = f ? 1 : 0;
    if (f) {
      const int2 k = make_int2(de_sat_i32(floorf(p.x / cell)), de_sat_i32(floorf(p.y / cell)));
      L.key[i] = k;
      mnx = min(mnx, k.x); mny = min(mny, k.y); mxx = max(mxx, k.x); mxy = max(mxy, k.y);
      cnt++;
    }
  }
  mnx = de_wave_min_i(mnx); mny = de_wave_min_i(mny); mxx = de_wave_max_i(mxx); mxy = de_wave_max_i(mxy);
  for (int q = 32; q > 0; q >>= 1) cnt += __shfl_xor(cnt, q, 64);
  if (t == 0) { L.box[0] = INT32_MAX; L.box[1] = INT32_MAX; L.box[2] = INT32_MIN; L.box[3] = INT32_MIN; L.m = 0; }
  __syncthreads();
  if ((t & 63) == 0) {
    atomicMin(&L.box[0], mnx); atomicMin(&L.box[1], mny); atomicMax(&L.box[2], mxx); atomicMax(&L.box[3], mxy);
    atomicAdd(&L.m, cnt);
  }
  __syncthreads();
  const uint32_t m = L.m;
  mnx = L.box[0]; mny = L.box[1];
  const int64_t W = m ? (int64_t)L.box[2] - mnx + 1 : 0, H = m ? (int64_t)L.box[3] - mny + 1 : 0;
  const uint64_t budget = 2ull * m + 512ull;  // grid.rs:113-120: Dense iff W * H <= 2 m + 512
  const bool dense = m > 0 && (uint64_t)W <= budget && (uint64_t)H <= budget && (uint64_t)W * (uint64_t)H <= budget;
  const uint64_t total = dense ? (uint64_t)W * (uint64_t)H : 0u;
  uint32_t nc = 0;  // cells of the layout (Dense: W*H; Sparse: occupied)
  if (dense) {
    nc = (uint32_t)total;
    for (uint32_t k = t; k <= nc; k += DE_BLOCK) L.start[k] = 0;
    __syncthreads();
    for (uint32_t i = t; i < n; i += DE_BLOCK)
      if (L.fin[i]) {
        const uint32_t k = (uint32_t)(((int64_t)L.key[i].x - mnx) * H + ((int64_t)L.key[i].y - mny));
        atomicAdd(&L.start[k + 1], 1u);
      }
    __syncthreads();
    uint32_t carry = 0;  // inclusive scan of the counts: start[k] = balls in cells before k
    for (uint32_t b = 1; b <= nc; b += DE_BLOCK) {
      const uint32_t k = b + t;
      const uint32_t c = k <= nc ? L.start[k] : 0u;
      uint32_t tot;
      const uint32_t ex = de_block_scan(c, L.w, &tot);
      if (k <= nc) L.start[k] = carry + ex + c;
      carry += tot;
      __syncthreads();
    }
    // each ball takes a slot of its cell (any order; start[k] advances from the cell's start to its
    // end and is shifted back after), then each cell's balls are put in ascending index order (a
    // cell holds a few balls)
    for (uint32_t i = t; i < n; i += DE_BLOCK)
      if (L.fin[i]) {
        const uint32_t k = (uint32_t)(((int64_t)L.key[i].x - mnx) * H + ((int64_t)L.key[i].y - mny));
        L.idx[atomicAdd(&L.start[k], 1u)] = i;
      }
    __syncthreads();
    for (int32_t b = (int32_t)(nc / DE_BLOCK) * DE_BLOCK; b >= 0; b -= DE_BLOCK) {  // highest block first
      const uint32_t k = (uint32_t)b + t;
      const uint32_t v = (k > 0 && k <= nc) ? L.start[k - 1] : 0u;
      __syncthreads();
      if (k > 0 && k <= nc) L.start[k] = v;
      __syncthreads();
    }
    if (t == 0) L.start[0] = 0;
    __syncthreads();
    for (uint32_t k = t; k < nc; k += DE_BLOCK) {
      const uint32_t s0 = L.start[k], e0 = L.start[k + 1];
      for (uint32_t a = s0 + 1; a < e0; a++) {
        const uint32_t v = L.idx[a];
        uint32_t b = a;
        while (b > s0 && L.idx[b - 1] > v) {
          L.idx[b] = L.idx[b - 1];
          b--;
        }
        L.idx[b] = v;
      }
    }
    __syncthreads();
  } else if (m > 0) {
    if (t == 0) {  // Sparse: finite balls ordered by (key, index) (insertion into runs of equal keys)
      uint32_t c = 0;
      for (uint32_t i = 0; i < n; i++) {
        if (!L.fin[i]) continue;
        const int2 k = L.key[i];
        uint32_t p = c;  // keep idx sorted by key, ascending index inside equal keys
        while (p > 0 && de_key_less(k, L.key[L.idx[p - 1]])) {
          L.idx[p] = L.idx[p - 1];
          p--;
        }
        L.idx[p] = i;
        c++;
      }
      L.start[0] = 0;
      for (uint32_t a = 0; a < c; a++) {
        const int2 k = L.key[L.idx[a]];
        if (a == 0 || k.x != L.ckey[nc - 1].x || k.y != L.ckey[nc - 1].y) {
          if (a > 0) L.start[nc] = a;
          L.ckey[nc++] = k;
        }
      }
      L.start[nc] = c;
      L.w[0] = nc;
    }
    __syncthreads();
    nc = L.w[0];
    __syncthreads();
  }
  // candidate pairs in emission order (grid.rs:291-362): a cell's pairs (inside it, then with the
  // forward cells (1, 0), (1, 1), (0, 1), (-1, 1)) follow the cells before it, so each cell counts
  // its pairs, a scan gives its offset and it writes them in the reference's order
  uint2* P = pairs + (size_t)blockIdx.x * pcap;
  float* Lm = lam + (size_t)blockIdx.x * pcap;
  uint32_t base = 0;
  for (uint32_t b = 0; b < nc && m > 0; b += DE_BLOCK) {  // nc and m are uniform
    const uint32_t k = b + t;
    uint32_t cnt = 0, s0 = 0, e0 = 0, os[4] = {0, 0, 0, 0}, oe[4] = {0, 0, 0, 0};
    if (k < nc) {
      s0 = L.start[k];
      e0 = L.start[k + 1];
      if (s0 != e0) {
        int32_t cx, cy;
        if (dense) {
          cx = (int32_t)((int64_t)mnx + (int64_t)k / H);
          cy = (int32_t)((int64_t)mny + (int64_t)k % H);
        } else {
          cx = L.ckey[k].x;
          cy = L.ckey[k].y;
        }
        const uint32_t c = e0 - s0;
        cnt = c * (c - 1u) / 2u;
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const int32_t ox = q < 2 ? 1 : (q == 2 ? 0 : -1), oy = q == 0 ? 0 : 1;
          const int32_t tx = de_sadd(cx, ox), ty = de_sadd(cy, oy);
          if (tx == cx && ty == cy) continue;
          uint32_t a, e;
          if (!de_range(L, dense, mnx, mny, W, H, nc, tx, ty, a, e)) continue;
          os[q] = a;
          oe[q] = e;
          cnt += c * (e - a);
        }
      }
    }
    uint32_t tot;
    uint32_t np = base + de_block_scan(cnt, L.w, &tot);
    if (cnt) {
      for (uint32_t a = s0; a < e0; a++)
        for (uint32_t b2 = a + 1; b2 < e0; b2++) de_emit(P, Lm, pcap, np, L.idx[a], L.idx[b2]);
#pragma unroll
      for (int q = 0; q < 4; q++)
        for (uint32_t a = s0; a < e0; a++)
          for (uint32_t b2 = os[q]; b2 < oe[q]; b2++) de_emit(P, Lm, pcap, np, L.idx[a], L.idx[b2]);
    }
    base += tot;
    __syncthreads();
  }
  if (t == 0) {
    o.n_pairs = base;
    o.pair_overflow = base > pcap ? 1u : 0u;
  }
}

// ---- 3. balls ending deep in the wall (dem.rs:611-621): listed for the host's swept test ----------
extern "C" __global__ void __launch_bounds__(DE_BLOCK)
dem_ccd_flag(const DeDrum* __restrict__ drums, const float2* __restrict__ lift, const float2* __restrict__ x,
             uint32_t* __restrict__ flagged, uint32_t stride, DeOut* __restrict__ out) {
  __shared__ uint32_t s_n;
  const DeDrum D = drums[blockIdx.x];
  const mgg::Geo g = de_geo(D, lift);
  if (threadIdx.x == 0) s_n = 0;
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < D.n; i += DE_BLOCK) {
    const float2 p = x[D.b0 + i];
    const float c = mgg::sdf_world_d(p.x, p.y, D.rc1, D.rs1, g) - D.r;
    if (c < -0.2f * 2.0f * D.r) flagged[(size_t)blockIdx.x * stride + atomicAdd(&s_n, 1u)] = i;
  }
  __syncthreads();
  if (threadIdx.x == 0) out[blockIdx.x].n_ccd = s_n;
}

// Writes the host's clamped positions: list of (population, ball) and positions.
extern "C" __global__ void dem_ccd_apply(const DeDrum* __restrict__ drums, const uint2* __restrict__ who,
                                        const float2* __restrict__ pos, uint32_t count, float2* __restrict__ x) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= count) return;
  x[drums[who[k].x].b0 + who[k].y] = pos[k];
}

// Level schedule of a sweep over m items (ball pairs i < j) in sweep order, inactive items skipped:
// level(k) = 1 + the largest level of an earlier active item on ball i or on ball j (0 when there is
// none). Items of one level touch distinct balls, and each ball's items have rising levels, so
// running the levels in order (the items of a level side by side) updates every ball exactly as the
// sequential sweep does. Wave 0 assigns the levels, 64 items at a time: lane k starts from the
// balls' last levels (s_last, n entries of LDS) and takes in, in lane order, the levels of the
// earlier lanes of its batch that share a ball. Then all lanes sort the items by level: lev_g[k]
// (global) holds the levels, s_off[0 .. depth] the level starts, and position p of the level order
// gets pk_g[p] = i | j << 16 and k_g[p] = k. Returns the depth, or DE_LCAP + 1 (nothing sorted)
// for a deeper schedule. Every thread of the workgroup calls it.
#define DE_LCAP 4096
template <class ItemFn>
__device__ uint32_t de_levels(ItemFn item, uint32_t m, uint32_t n, int32_t* s_last, uint32_t* s_off, uint32_t* s_w,
                              uint32_t* s_depth, uint32_t* __restrict__ lev_g, uint32_t* __restrict__ pk_g,
                              uint32_t* __restrict__ k_g) {
  const uint32_t t = threadIdx.x, lane = t & 63u;
  for (uint32_t b = t; b < n; b += DE_BLOCK) s_last[b] = -1;
  __syncthreads();
  if (t < 64) {
    int32_t depth = 0;
    uint32_t ni = 0, nj = 0;  // the next chunk's pair, loaded while this chunk runs
    bool nact = lane < m && item(lane, ni, nj);
    for (uint32_t base = 0; base < m; base += 64) {
      const uint32_t k = base + lane;
      uint32_t i = ni, j = nj;
      const bool act = nact;
      nact = k + 64 < m && item(k + 64, ni, nj);
      int32_t l = act ? 1 + max(s_last[i], s_last[j]) : -1;
      uint64_t am = __ballot(act);
      while (am) {
        const int src = __ffsll((unsigned long long)am) - 1;
        am &= am - 1ull;
        const int32_t ls = __builtin_amdgcn_readlane(l, src);
        const uint32_t is = (uint32_t)__builtin_amdgcn_readlane((int)i, src);
        const uint32_t js = (uint32_t)__builtin_amdgcn_readlane((int)j, src);
        if (act && (int)lane > src && (i == is || i == js || j == is || j == js)) l = max(l, ls + 1);
      }
      if (act) {
        atomicMax(&s_last[i], l);
        atomicMax(&s_last[j], l);
      }
      if (k < m) lev_g[k] = act ? (uint32_t)l : DE_NO;
      depth = max(depth, l + 1);
    }
    for (int o = 32; o > 0; o >>= 1) depth = max(depth, __shfl_xor(depth, o, 64));
    if (lane == 0) *s_depth = (uint32_t)depth;
  }
  __syncthreads();
  const uint32_t depth = *s_depth;
  if (depth > DE_LCAP) return DE_LCAP + 1;
  for (uint32_t L = t; L <= depth; L += DE_BLOCK) s_off[L] = 0;
  __syncthreads();
  for (uint32_t k = t; k < m; k += DE_BLOCK) {
    const uint32_t l = lev_g[k];
    if (l != DE_NO) atomicAdd(&s_off[l], 1u);
  }
  __syncthreads();
  uint32_t carry = 0;
  for (uint32_t b = 0; b < depth; b += DE_BLOCK) {  // exclusive scan of the level counts
    const uint32_t L = b + t;
    const uint32_t c = L < depth ? s_off[L] : 0u;
    uint32_t tot;
    const uint32_t ex = de_block_scan(c, s_w, &tot);
    if (L < depth) s_off[L] = carry + ex;
    carry += tot;
    __syncthreads();
  }
  if (t == 0) s_off[depth] = carry;
  __syncthreads();
  for (uint32_t k = t; k < m; k += DE_BLOCK) {  // any order inside a level
    const uint32_t l = lev_g[k];
    if (l == DE_NO) continue;
    uint32_t i = 0, j = 0;
    item(k, i, j);
    const uint32_t p = atomicAdd(&s_off[l], 1u);
    pk_g[p] = i | (j << 16);
    k_g[p] = k;
  }
  __syncthreads();
  // s_off[L] now ends level L: shift back to the starts, the highest block of 256 first (a block
  // reads the entry just below it, which a lower block would overwrite)
  for (int32_t b = (int32_t)(depth / DE_BLOCK) * DE_BLOCK; b >= 0; b -= DE_BLOCK) {
    const uint32_t L = (uint32_t)b + t;
    const uint32_t v = (L > 0 && L <= depth) ? s_off[L - 1] : 0u;
    __syncthreads();
    if (L <= depth && L > 0) s_off[L] = v;
    __syncthreads();
  }
  if (t == 0) s_off[0] = 0;
  __syncthreads();
  return depth;
}

// ---- 4. XPBD projection (dem.rs:710-750) and the velocity rebuild (:752-756) ---------------------
// Pairs in emission order, then walls, per iteration.
// One workgroup per population, the positions in LDS. A candidate pair changes nothing unless its
// balls overlap at its turn, and most candidates (every pair of balls in neighbouring grid cells)
// never do, so the sweep first takes only the pairs closer than 2r + DE_GAP r at the start (found
// by all lanes, each over a contiguous run of the emission order, and compacted in that order): by
// levels of that subset's emission order (de_levels: per level the lanes update distinct pairs side
// by side, with the same f32 operations as the sequential sweep; each pair's accumulated impulse
// kept in level order and written back to emission order at the end), every move's length added to
// its balls' travel. A pair left out is then exact to have skipped if its start distance less both
// balls' travel still exceeds 2r (with margins over the rounding): it could not have overlapped at
// any of its turns. If any left-out pair fails that test, the projection starts again from the
// start positions with the pairs that the first pass's travel (doubled) could have brought within
// the reach, checked the same way, and failing that with every candidate pair. A schedule deeper
// than DE_LCAP levels falls back to lane 0 walking all pairs in order. The gap is kept per drum
// slot between substeps, from the travel seen (any gap gives the same results; it trades near
// pairs against checks and restarts).
#define DE_GAP 0.05f      // the smallest gap (in r) of the near pairs
#define DE_GAP_MAX 0.8f   // the largest
#define DE_PCHUNK 512  // pairs staged in LDS per chunk of the projection
__device__ __forceinline__ void de_project_pair(float2* s_x, float* s_d, uint32_t pk, float& lam, float w, float ws,
                                                float cap_c, float two_r) {
  const uint32_t i = pk & 0xFFFFu, j = pk >> 16;
  const float dx = s_x[i].x - s_x[j].x, dy = s_x[i].y - s_x[j].y;
  const float dist = sqrtf(dx * dx + dy * dy);
  const float c = dist - two_r;
  if (c >= 0.0f) return;
  float nx = 1.0f, ny = 0.0f;
  if (dist > 1e-9f) {
    nx = dx / dist;
    ny = dy / dist;
  }
  if (ws <= 0.0f) return;
  const float ce = fmaxf(c, cap_c);
  const float dl = -ce / ws;
  const float s = w * dl;
  s_x[i].x += s * nx;
  s_x[i].y += s * ny;
  s_x[j].x -= s * nx;
  s_x[j].y -= s * ny;
  lam += dl;
  if (s_d) {
    s_d[i] += fabsf(s);
    s_d[j] += fabsf(s);
  }
}

extern "C" __global__ void __launch_bounds__(DE_BLOCK)
dem_project(const DeDrum* __restrict__ drums, const float2* __restrict__ lift, float2* __restrict__ x,
            float2* __restrict__ v, const float* __restrict__ th, float* __restrict__ om,
            const float2* __restrict__ x0, const float* __restrict__ th0, const uint2* __restrict__ pairs,
            float* __restrict__ lam, float* __restrict__ lam_bw, uint32_t* __restrict__ lev_g,
            uint32_t* __restrict__ pk_g, uint32_t* __restrict__ k_g, float* __restrict__ lam_s,
            uint32_t* __restrict__ cpk_g, uint32_t* __restrict__ ck_g, float* __restrict__ slack_g,
            float* __restrict__ gap_g, uint32_t stride, uint32_t pcap, DeOut* __restrict__ out) {
  __shared__ float2 s_x[DE_MAX_BALLS];
  __shared__ float s_lbw[DE_MAX_BALLS];
  __shared__ int32_t s_last[DE_MAX_BALLS];  // the schedule's scratch, then each ball's travel
  __shared__ float s_d0[DE_MAX_BALLS];      // each ball's travel in the near pass
  __shared__ uint32_t s_off[DE_LCAP + 1];
  __shared__ uint32_t s_pk[DE_PCHUNK];
  __shared__ float s_l[DE_PCHUNK];
  __shared__ uint32_t s_w[DE_BLOCK / 64], s_depth, s_fail, s_maxd;
  const DeDrum D = drums[blockIdx.x];
  const mgg::Geo g = de_geo(D, lift);
  const uint32_t t = threadIdx.x, n = D.n;
  const size_t sb = (size_t)blockIdx.x * stride, pb = (size_t)blockIdx.x * pcap;
  const uint32_t np = min(out[blockIdx.x].n_pairs, pcap);
  const uint2* P = pairs + pb;
  const float2* X = x + D.b0;  // the start positions (written back only at the end)
  const float w = D.w, ws = D.w + D.w, cap_c = -0.2f * 2.0f * D.r, two_r = 2.0f * D.r;
  const float gap = fminf(fmaxf(gap_g[blockIdx.x], DE_GAP), DE_GAP_MAX);
  const float reach = two_r + gap * D.r, near2 = reach * reach, sure = two_r * 1.0001f;
  bool restarted = false;
  float maxd = 0.0f;  // the largest travel of the last checked pass
  float* s_d = reinterpret_cast<float*>(s_last);
  uint32_t* CPK = cpk_g + pb;  // the walked pairs (i | j << 16) in emission order
  uint32_t* CK = ck_g + pb;    // their emission indices
  float* SL = slack_g + pb;    // per emission index: start distance less 2r (with a margin) of a left-out pair
  for (uint32_t i = t; i < n; i += DE_BLOCK) s_x[i] = X[i];
  __syncthreads();
  uint32_t depth = 0, na = 0;
  for (int pass = 0; pass < 3; pass++) {
    // the pairs this pass walks, in emission order: each lane a contiguous run, counted, scanned,
    // written (the test evaluated twice, the same way)
    auto walks = [&](uint32_t k, uint32_t& pk, float& slack) {
      const uint2 p = P[k];
      pk = p.x | (p.y << 16);
      slack = __int_as_float(0x7f800000);
      if (pass == 2) return true;
      const float dx = s_x[p.x].x - s_x[p.y].x, dy = s_x[p.x].y - s_x[p.y].y;
      const float d2 = dx * dx + dy * dy;
      if (pass == 0) {
        if (!(d2 >= near2)) return true;  // near, or a non-finite distance
      } else if (!(sqrtf(d2) - 2.0f * (s_d0[p.x] + s_d0[p.y]) >= reach)) {
        return true;  // the first pass's travel (doubled) could have brought it near
      }
      slack = sqrtf(d2) - sure;
      return false;
    };
    {
      const uint32_t run = (np + DE_BLOCK - 1) / DE_BLOCK, a = min(np, t * run), e = min(np, a + run);
      uint32_t c = 0;
      for (uint32_t k = a; k < e; k++) {
        uint32_t pk;
        float sl;
        c += walks(k, pk, sl) ? 1u : 0u;
      }
      uint32_t tot;
      uint32_t o = de_block_scan(c, s_w, &tot);
      na = tot;
      for (uint32_t k = a; k < e; k++) {
        uint32_t pk;
        float sl;
        const bool wk = walks(k, pk, sl);
        SL[k] = sl;
        if (wk) {
          CPK[o] = pk;
          CK[o] = k;
          o++;
        }
      }
    }
    __syncthreads();
    auto item = [CPK](uint32_t q, uint32_t& i, uint32_t& j) {
      const uint32_t pk = CPK[q];
      i = pk & 0xFFFFu;
      j = pk >> 16;
      return true;
    };
    depth = de_levels(item, na, n, s_last, s_off, s_w, &s_depth, lev_g + pb, pk_g + pb, k_g + pb);
    const bool leveled = depth <= DE_LCAP;
    const bool track = pass < 2;  // a pass that leaves pairs out checks them afterwards
    if (t == 0) out[blockIdx.x].depth = depth;
    for (uint32_t i = t; i < n; i += DE_BLOCK) {
      s_x[i] = X[i];
      s_lbw[i] = 0.0f;
      s_d[i] = 0.0f;
    }
    for (uint32_t p = t; p < na; p += DE_BLOCK) lam_s[pb + p] = 0.0f;
    __syncthreads();
    for (uint32_t it = 0; it < max(D.iters, 1u); it++) {
      uint32_t L = 0;
      for (uint32_t c0 = 0; c0 < na; c0 += DE_PCHUNK) {
        const uint32_t m = min(na - c0, (uint32_t)DE_PCHUNK), c1 = c0 + m;
        for (uint32_t q = t; q < m; q += DE_BLOCK) {
          s_pk[q] = leveled ? pk_g[pb + c0 + q] : CPK[c0 + q];
          s_l[q] = lam_s[pb + c0 + q];
        }
        __syncthreads();
        if (leveled) {
          while (L < depth && s_off[L] < c1) {  // the levels (or their parts) inside this chunk
            const uint32_t a = max(s_off[L], c0), e = min(s_off[L + 1], c1);
            for (uint32_t p = a + t; p < e; p += DE_BLOCK)
              de_project_pair(s_x, track ? s_d : nullptr, s_pk[p - c0], s_l[p - c0], w, ws, cap_c, two_r);
            __syncthreads();
            if (s_off[L + 1] > c1) break;  // the level goes on in the next chunk
            L++;
          }
        } else if (t == 0) {
          for (uint32_t q = 0; q < m; q++)
            de_project_pair(s_x, track ? s_d : nullptr, s_pk[q], s_l[q], w, ws, cap_c, two_r);
        }
        __syncthreads();
        for (uint32_t q = t; q < m; q += DE_BLOCK) lam_s[pb + c0 + q] = s_l[q];
        __syncthreads();
      }
      for (uint32_t i = t; i < n; i += DE_BLOCK) {  // walls: each ball on its own
        float nx, ny;
        const float d = mgg::sdf_world(s_x[i].x, s_x[i].y, D.rc1, D.rs1, D.wc1, D.ws1, g, nx, ny);
        const float c = d - D.r;
        if (c >= 0.0f) continue;
        if (w <= 0.0f) continue;
        const float ce = fmaxf(c, cap_c);
        const float dl = -ce / w;
        const float s = w * dl;
        s_x[i].x += s * nx;
        s_x[i].y += s * ny;
        s_lbw[i] += dl;
        s_d[i] += fabsf(s);
      }
      __syncthreads();
    }
    if (!track) break;  // every candidate pair was walked
    // the left-out pairs: start distance less both balls' travel above 2r, with margins. A left-out
    // pair starts at least the gap less the margin beyond 2r, so when twice the largest travel
    // (with the same margin factor) stays below that, none can fail and the pairs are not read.
    if (t == 0) {
      s_fail = 0;
      s_maxd = 0u;
    }
    __syncthreads();
    {
      float md = 0.0f;
      for (uint32_t i = t; i < n; i += DE_BLOCK) md = fmaxf(md, s_d[i]);
      if (!(md >= 0.0f && md <= 3.0e38f)) md = __int_as_float(0x7f800000);  // a non-finite travel: check
      atomicMax(&s_maxd, __float_as_uint(md));  // non-negative floats order as their bits
    }
    __syncthreads();
    maxd = __uint_as_float(s_maxd);
    if (2.01f * maxd < (gap - 0.001f) * D.r) break;
    uint32_t bad = 0;
    for (uint32_t k0 = 0; k0 < np; k0 += 8 * DE_BLOCK) {  // eight pairs per lane in flight
      float sl[8];
      uint2 pp[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {  // unconditional loads (a clamped index), so none waits for another
        const uint32_t k = min(k0 + (uint32_t)u * DE_BLOCK + t, np - 1);
        sl[u] = SL[k];
        pp[u] = P[k];
      }
#pragma unroll
      for (int u = 0; u < 8; u++)
        if (sl[u] != __int_as_float(0x7f800000) && !(sl[u] - 1.001f * (s_d[pp[u].x] + s_d[pp[u].y]) > 0.0f)) bad = 1;
    }
    if (bad) s_fail = 1;
    __syncthreads();
    if (!s_fail) break;
    // restart from the start positions: with the pairs the travel could have brought near, then
    // with every pair
    if (t == 0) out[blockIdx.x].restart = pass + 1;
    restarted = true;
    if (pass == 0)
      for (uint32_t i = t; i < n; i += DE_BLOCK) s_d0[i] = s_d[i];
    for (uint32_t i = t; i < n; i += DE_BLOCK) s_x[i] = X[i];
    __syncthreads();
  }
  // each walked pair's impulse back to emission order; the others are zero
  for (uint32_t k = t; k < np; k += DE_BLOCK) lam[pb + k] = 0.0f;
  __syncthreads();
  if (depth <= DE_LCAP)
    for (uint32_t p = t; p < na; p += DE_BLOCK) lam[pb + CK[k_g[pb + p]]] = lam_s[pb + p];
  else
    for (uint32_t q = t; q < na; q += DE_BLOCK) lam[pb + CK[q]] = lam_s[pb + q];
  for (uint32_t i = t; i < n; i += DE_BLOCK) {
    const float2 p = s_x[i], p0 = x0[sb + i];
    x[D.b0 + i] = p;
    v[D.b0 + i] = make_float2((p.x - p0.x) / D.dt, (p.y - p0.y) / D.dt);
    om[D.b0 + i] = de_angle_diff(th[D.b0 + i], th0[sb + i]) / D.dt;
    lam_bw[sb + i] = s_lbw[i];
  }
  if (t == 0) {
    // the next substep's gap: three times this one's largest travel (so that its check is
    // usually skipped), shrinking by at most 10% a substep; doubled after a restart
    const float next = restarted ? 2.0f * gap : fmaxf(3.0f * maxd / D.r, 0.9f * gap);
    gap_g[blockIdx.x] = fmaxf(fminf(next, DE_GAP_MAX), DE_GAP);
  }
}

// ---- 5. ContactBook::sorted_ball_ball (dem.rs:459-467): by i (counting sort), then j ----------------
// The sweeps over the sorted book (friction, restitution, the rolling-resistance totals) act only on
// contacts with a positive normal impulse: those are compacted in book order (spk: i | j << 16,
// slam: the impulse) and given a level schedule (de_levels; level order in cpk / clam, level starts
// in coff), which the three sweeps share.
extern "C" __global__ void __launch_bounds__(DE_BLOCK)
dem_sort(const DeDrum* __restrict__ drums, const uint2* __restrict__ pairs, const float* __restrict__ lam,
         DeItem* __restrict__ sorted, uint32_t* __restrict__ spk, float* __restrict__ slam,
         uint32_t* __restrict__ lev_g, uint32_t* __restrict__ cpk, uint32_t* __restrict__ ck,
         float* __restrict__ clam, uint32_t* __restrict__ coff, uint32_t pcap, DeOut* __restrict__ out) {
  __shared__ uint32_t s_c[DE_MAX_BALLS + 1];
  __shared__ int32_t s_last[DE_MAX_BALLS];
  __shared__ uint32_t s_off[DE_LCAP + 1];
  __shared__ uint32_t s_w[DE_BLOCK / 64];
  __shared__ uint32_t s_carry, s_depth;
  const DeDrum D = drums[blockIdx.x];
  const uint32_t t = threadIdx.x, n = D.n;
  const size_t pb = (size_t)blockIdx.x * pcap, ob = (size_t)blockIdx.x * (DE_LCAP + 1);
  const uint32_t np = min(out[blockIdx.x].n_pairs, pcap);
  for (uint32_t i = t; i <= n; i += DE_BLOCK) s_c[i] = 0;
  if (t == 0) s_carry = 0;
  __syncthreads();
  for (uint32_t q = t; q < np; q += DE_BLOCK) atomicAdd(&s_c[pairs[pb + q].x], 1u);
  __syncthreads();
  for (uint32_t b = 0; b < n; b += DE_BLOCK) {  // exclusive scan of the counts in place
    const uint32_t i = b + t;
    const uint32_t c = i < n ? s_c[i] : 0u;
    uint32_t tot;
    const uint32_t ex = de_block_scan(c, s_w, &tot);
    const uint32_t carry = s_carry;
    if (i < n) s_c[i] = carry + ex;
    __syncthreads();
    if (t == 0) s_carry = carry + tot;
    __syncthreads();
  }
  if (t == 0) s_c[n] = s_carry;
  __syncthreads();
  // each pair takes a slot of its bucket (any order; s_c[i] advances from the bucket's start to
  // its end) and each bucket is then sorted by j
  for (uint32_t q = t; q < np; q += DE_BLOCK) {
    const uint2 p = pairs[pb + q];
    const uint32_t slot = atomicAdd(&s_c[p.x], 1u);
    sorted[pb + slot] = DeItem{p.x, p.y, lam[pb + q]};
  }
  __syncthreads();
  // s_c[i] now holds the end of bucket i (= start of bucket i + 1 before the fill): bucket i spans
  // [start, s_c[i]) with start = (i == 0 ? 0 : s_c[i - 1])
  for (uint32_t i = t; i < n; i += DE_BLOCK) {
    const uint32_t s = i == 0 ? 0u : s_c[i - 1], e = s_c[i];
    for (uint32_t a = s + 1; a < e; a++) {  // insertion sort by j (buckets hold a few pairs)
      const DeItem it = sorted[pb + a];
      uint32_t b = a;
      while (b > s && sorted[pb + b - 1].j > it.j) {
        sorted[pb + b] = sorted[pb + b - 1];
        b--;
      }
      sorted[pb + b] = it;
    }
  }
  __syncthreads();
  uint32_t na = 0;  // the active contacts in book order
  for (uint32_t b = 0; b < np; b += DE_BLOCK) {
    const uint32_t k = b + t;
    DeItem it{0u, 0u, 0.0f};
    if (k < np) it = sorted[pb + k];
    const bool act = k < np && it.lam > 0.0f;
    uint32_t tot;
    const uint32_t rank = de_block_scan(act ? 1u : 0u, s_w, &tot);
    if (act) {
      spk[pb + na + rank] = it.i | (it.j << 16);
      slam[pb + na + rank] = it.lam;
    }
    na += tot;
    __syncthreads();
  }
  const uint32_t* SP = spk + pb;
  auto item = [SP](uint32_t k, uint32_t& i, uint32_t& j) {
    const uint32_t v = SP[k];
    i = v & 0xFFFFu;
    j = v >> 16;
    return true;
  };
  const uint32_t depth = de_levels(item, na, n, s_last, s_off, s_w, &s_depth, lev_g + pb, cpk + pb, ck + pb);
  if (depth <= DE_LCAP) {
    for (uint32_t L = t; L <= depth; L += DE_BLOCK) coff[ob + L] = s_off[L];
    for (uint32_t p = t; p < na; p += DE_BLOCK) clam[pb + p] = slam[pb + ck[pb + p]];
  }
  if (t == 0) {
    out[blockIdx.x].nact = na;
    out[blockIdx.x].cdepth = depth;
  }
}

// Runs f(i, j, lam) over the active contacts in book order on the level schedule (all lanes, a
// barrier per level) or, for a schedule deeper than DE_LCAP, on lane 0 in book order. Every thread
// calls it; it ends with a barrier.
template <class F>
__device__ void de_contact_sweep(F f, const DeOut& o, const uint32_t* __restrict__ spk, const float* __restrict__ slam,
                                 const uint32_t* __restrict__ cpk, const float* __restrict__ clam,
                                 const uint32_t* __restrict__ coff) {
  const uint32_t t = threadIdx.x, na = o.nact, depth = o.cdepth;
  if (depth <= DE_LCAP) {
    for (uint32_t L = 0; L < depth; L++) {
      const uint32_t a = coff[L], e = coff[L + 1];
      for (uint32_t p = a + t; p < e; p += DE_BLOCK) {
        const uint32_t v = cpk[p];
        f(v & 0xFFFFu, v >> 16, clam[p]);
      }
      __syncthreads();
    }
  } else {
    if (t == 0)
      for (uint32_t k = 0; k < na; k++) f(spk[k] & 0xFFFFu, spk[k] >> 16, slam[k]);
    __syncthreads();
  }
}

// ---- 6. friction (dem.rs:676-890) and the velocity rebuild (:892-895) ------------------------------
extern "C" __global__ void __launch_bounds__(DE_BLOCK)
dem_friction(const DeDrum* __restrict__ drums, const float2* __restrict__ lift, float2* __restrict__ x,
             float2* __restrict__ v, float* __restrict__ th, float* __restrict__ om, const float2* __restrict__ x0,
             const float* __restrict__ th0, const uint32_t* __restrict__ spk, const float* __restrict__ slam,
             const uint32_t* __restrict__ cpk, const float* __restrict__ clam, const uint32_t* __restrict__ coff,
             const float* __restrict__ lam_bw, uint32_t stride, uint32_t pcap, DeOut* __restrict__ out) {
  __shared__ float2 s_x[DE_MAX_BALLS], s_v[DE_MAX_BALLS];
  __shared__ float s_th[DE_MAX_BALLS], s_om[DE_MAX_BALLS];
  __shared__ float s_ww[2 * DE_BLOCK];
  const DeDrum D = drums[blockIdx.x];
  const mgg::Geo g = de_geo(D, lift);
  const uint32_t t = threadIdx.x, n = D.n;
  const size_t sb = (size_t)blockIdx.x * stride, pb = (size_t)blockIdx.x * pcap, ob = (size_t)blockIdx.x * (DE_LCAP + 1);
  for (uint32_t i = t; i < n; i += DE_BLOCK) {
    s_x[i] = x[D.b0 + i];
    s_v[i] = v[D.b0 + i];
    s_th[i] = th[D.b0 + i];
    s_om[i] = om[D.b0 + i];
  }
  __syncthreads();
  const float r = D.r, w = D.w, wr = D.w_rot, dt = D.dt;
  const float wst_bb = w + w + r * r * wr + r * r * wr;
  auto pair = [&](uint32_t i, uint32_t j, float lamc) {
    const float dx = s_x[i].x - s_x[j].x, dy = s_x[i].y - s_x[j].y;
    const float dist = sqrtf(dx * dx + dy * dy);
    float nx = 1.0f, ny = 0.0f;
    if (dist > 1e-9f) {
      nx = dx / dist;
      ny = dy / dist;
    }
    const float tx = -ny, ty = nx;
    const float vt = ((s_v[i].x - s_v[j].x) * tx + (s_v[i].y - s_v[j].y) * ty) - r * s_om[i] - r * s_om[j];
    if (wst_bb <= 0.0f) return;
    const float raw = -vt * dt / wst_bb;
    const float mm = de_friction(D.fric_bb, D.fric_bb_s, vt, D.fric_vs) * lamc;
    float dlt = raw;
    if (dlt < -mm) dlt = -mm;
    if (dlt > mm) dlt = mm;
    const float s = w * dlt;
    s_x[i].x += s * tx;
    s_x[i].y += s * ty;
    s_x[j].x -= s * tx;
    s_x[j].y -= s * ty;
    const float dth = r * wr * dlt;
    s_th[i] -= dth;
    s_th[j] -= dth;
    const float dvx = s * tx / dt, dvy = s * ty / dt;
    const float dom = -r * wr * dlt / dt;
    s_v[i].x += dvx;
    s_v[i].y += dvy;
    s_v[j].x -= dvx;
    s_v[j].y -= dvy;
    s_om[i] += dom;
    s_om[j] += dom;
  };
  de_contact_sweep(pair, out[blockIdx.x], spk + pb, slam + pb, cpk + pb, clam + pb, coff + ob);
  // walls: each ball on its own; the wall-work terms are added in ball order by lane 0
  float ww = 0.0f;
  double wwd = 0.0;
  const float wst_bw = w + r * r * wr;
  for (uint32_t base = 0; base < n; base += DE_BLOCK) {
    const uint32_t i = base + t;
    float a = 0.0f, b = 0.0f;  // an inactive ball adds +0.0 (the running sum is never -0.0)
    if (i < n) {
      const float lamw = lam_bw[sb + i];
      if (lamw > 0.0f) {
        float nx, ny;
        mgg::sdf_world(s_x[i].x, s_x[i].y, D.rc1, D.rs1, D.wc1, D.ws1, g, nx, ny);
        const float tx = -ny, ty = nx;
        const float cx = s_x[i].x - r * nx, cy = s_x[i].y - r * ny;
        const float vwx = D.omega * -cy, vwy = D.omega * cx;
        const float vt = ((s_v[i].x - vwx) * tx + (s_v[i].y - vwy) * ty) - r * s_om[i];
        if (wst_bw > 0.0f) {
          const float raw = -vt * dt / wst_bw;
          const float mm = de_friction(D.fric_bw, D.fric_bw_s, vt, D.fric_vs) * lamw;
          float dlt = raw;
          if (dlt < -mm) dlt = -mm;
          if (dlt > mm) dlt = mm;
          const float s = w * dlt;
          s_x[i].x += s * tx;
          s_x[i].y += s * ty;
          s_th[i] -= r * wr * dlt;
          s_v[i].x += s * tx / dt;
          s_v[i].y += s * ty / dt;
          s_om[i] += -r * wr * dlt / dt;
          a = (dlt / dt) * (tx * vwx + ty * vwy);
          b = (lamw / dt) * (nx * vwx + ny * vwy);
        }
      }
    }
    s_ww[2 * t] = a;
    s_ww[2 * t + 1] = b;
    __syncthreads();
    if (t == 0) {
      const uint32_t mm = min(n - base, (uint32_t)DE_BLOCK);
      if (D.f64)
        for (uint32_t q = 0; q < 2 * mm; q++) wwd += (double)s_ww[q];
      else
        for (uint32_t q = 0; q < 2 * mm; q++) ww += s_ww[q];
    }
    __syncthreads();
  }
  for (uint32_t i = t; i < n; i += DE_BLOCK) {
    const float2 p = s_x[i], p0 = x0[sb + i];
    x[D.b0 + i] = p;
    v[D.b0 + i] = make_float2((p.x - p0.x) / dt, (p.y - p0.y) / dt);
    th[D.b0 + i] = s_th[i];
    om[D.b0 + i] = de_angle_diff(s_th[i], th0[sb + i]) / dt;
  }
  if (t == 0) out[blockIdx.x].ww = D.f64 ? wwd : (double)ww;
}

// ---- 7. restitution and impact records (dem.rs:897-1034) ------------------------------------------
extern "C" __global__ void __launch_bounds__(DE_BLOCK)
dem_restitution(const DeDrum* __restrict__ drums, const float2* __restrict__ lift, const float2* __restrict__ x,
                float2* __restrict__ v, const float2* __restrict__ vpre, const float* __restrict__ ompre,
                const uint32_t* __restrict__ spk, const float* __restrict__ slam, const uint32_t* __restrict__ cpk,
                const float* __restrict__ clam, const uint32_t* __restrict__ coff, const float* __restrict__ lam_bw,
                const float* __restrict__ ebin, mg_impact_record* __restrict__ rec, uint32_t rcap, uint32_t stride,
                uint32_t pcap, DeOut* __restrict__ out) {
  __shared__ float2 s_x[DE_MAX_BALLS], s_v[DE_MAX_BALLS], s_vp[DE_MAX_BALLS];
  __shared__ float s_op[DE_MAX_BALLS];
  __shared__ uint32_t s_hist[12], s_coll, s_nrec;
  __shared__ uint32_t s_w[DE_BLOCK / 64];
  const DeDrum D = drums[blockIdx.x];
  const mgg::Geo g = de_geo(D, lift);
  const uint32_t t = threadIdx.x, n = D.n;
  const size_t sb = (size_t)blockIdx.x * stride, pb = (size_t)blockIdx.x * pcap, rb = (size_t)blockIdx.x * rcap;
  const size_t ob = (size_t)blockIdx.x * (DE_LCAP + 1);
  const DeOut& o = out[blockIdx.x];
  const uint32_t na = o.nact;
  for (uint32_t i = t; i < n; i += DE_BLOCK) {
    s_x[i] = x[D.b0 + i];
    s_v[i] = v[D.b0 + i];
    s_vp[i] = vpre[sb + i];
    s_op[i] = ompre[sb + i];
  }
  if (t < 12) s_hist[t] = 0;
  if (t == 0) { s_coll = 0; s_nrec = 0; }
  __syncthreads();
  const float r = D.r, w = D.w;
  // the contact's normal, its pre-step normal speed and whether it is a fresh impact: these read
  // positions and pre-step velocities, which the sweep below does not change
  auto normal = [&](uint32_t i, uint32_t j, float& nx, float& ny) {
    const float dx = s_x[i].x - s_x[j].x, dy = s_x[i].y - s_x[j].y;
    const float dist = sqrtf(dx * dx + dy * dy);
    nx = 1.0f;
    ny = 0.0f;
    if (dist > 1e-9f) {
      nx = dx / dist;
      ny = dy / dist;
    }
    return (s_vp[i].x - s_vp[j].x) * nx + (s_vp[i].y - s_vp[j].y) * ny;
  };
  // collisions, the energy histogram and the impact records, in book order (record k goes after
  // the records of the contacts before it)
  for (uint32_t b = 0; b < na; b += DE_BLOCK) {  // na is uniform
    const uint32_t k = b + t;
    uint32_t want = 0;
    mg_impact_record rr{};
    if (k < na) {
      const uint32_t pk = spk[pb + k], i = pk & 0xFFFFu, j = pk >> 16;
      float nx, ny;
      const float vnp = normal(i, j, nx, ny);
      const bool fresh = vnp < -0.02f;
      if (fresh) {
        atomicAdd(&s_coll, 1u);
        const int bin = de_ebin(0.5f * (D.mass * 0.5f) * vnp * vnp, ebin);
        if (bin >= 0) atomicAdd(&s_hist[bin], 1u);
      }
      const float e = fresh ? D.rest_bb : 0.0f;
      if (-vnp > D.thr) {
        const float tx = -ny, ty = nx;
        const float slip = ((s_vp[i].x - s_vp[j].x) * tx + (s_vp[i].y - s_vp[j].y) * ty) - r * s_op[i] - r * s_op[j];
        rr = mg_impact_record{0u, i, j, -vnp, fabsf(slip), e};
        want = 1;
      }
    }
    uint32_t tot;
    const uint32_t rank = de_block_scan(want, s_w, &tot);
    const uint32_t at = s_nrec + rank;
    if (want && at < rcap) rec[rb + at] = rr;
    __syncthreads();
    if (t == 0) s_nrec += tot;
    __syncthreads();
  }
  // the velocity targets, contact after contact on each ball
  auto pair = [&](uint32_t i, uint32_t j, float) {
    float nx, ny;
    const float vnp = normal(i, j, nx, ny);
    const float e = vnp < -0.02f ? D.rest_bb : 0.0f;
    const float target = fmaxf(fmaxf(-e * vnp, 0.0f), fmaxf(vnp, 0.0f));
    const float vnn = (s_v[i].x - s_v[j].x) * nx + (s_v[i].y - s_v[j].y) * ny;
    const float dvn = target - vnn;
    const float ws = w + w;
    if (ws <= 0.0f) return;
    const float im = dvn / ws;
    const float s = w * im;
    s_v[i].x += s * nx;
    s_v[i].y += s * ny;
    s_v[j].x -= s * nx;
    s_v[j].y -= s * ny;
  };
  de_contact_sweep(pair, o, spk + pb, slam + pb, cpk + pb, clam + pb, coff + ob);
  // walls: each ball on its own; records in ball order after the pair records
  for (uint32_t base = 0; base < n; base += DE_BLOCK) {
    const uint32_t i = base + t;
    uint32_t want = 0;
    mg_impact_record rr{};
    if (i < n) {
      const float lamw = lam_bw[sb + i];
      if (lamw > 0.0f) {
        float nx, ny;
        mgg::sdf_world(s_x[i].x, s_x[i].y, D.rc1, D.rs1, D.wc1, D.ws1, g, nx, ny);
        const float cx = s_x[i].x - r * nx, cy = s_x[i].y - r * ny;
        const float vwx = D.omega * -cy, vwy = D.omega * cx;
        const float vnp = (s_vp[i].x - vwx) * nx + (s_vp[i].y - vwy) * ny;
        const bool fresh = vnp < -0.02f;
        if (fresh) {
          atomicAdd(&s_coll, 1u);
          const int b = de_ebin(0.5f * D.mass * vnp * vnp, ebin);
          if (b >= 0) atomicAdd(&s_hist[b], 1u);
        }
        const float e = fresh ? D.rest_bw : 0.0f;
        if (-vnp > D.thr) {
          const float tx = -ny, ty = nx;
          const float slip = ((s_vp[i].x - vwx) * tx + (s_vp[i].y - vwy) * ty) - r * s_op[i];
          rr = mg_impact_record{1u, i, 0xFFFFFFFFu, -vnp, fabsf(slip), e};
          want = 1;
        }
        const float target = fmaxf(fmaxf(-e * vnp, 0.0f), fmaxf(vnp, 0.0f));
        const float vnn = (s_v[i].x - vwx) * nx + (s_v[i].y - vwy) * ny;
        const float dvn = target - vnn;
        if (w > 0.0f) {
          s_v[i].x += dvn * nx;
          s_v[i].y += dvn * ny;
        }
      }
    }
    uint32_t tot;
    const uint32_t rank = de_block_scan(want, s_w, &tot);
    const uint32_t at = s_nrec + rank;
    if (want && at < rcap) rec[rb + at] = rr;
    __syncthreads();
    if (t == 0) s_nrec += tot;
    __syncthreads();
  }
  for (uint32_t i = t; i < n; i += DE_BLOCK) v[D.b0 + i] = s_v[i];
  if (t == 0) {
    DeOut& o = out[blockIdx.x];
    o.coll = s_coll;
    o.n_imp = s_nrec;
    for (int b = 0; b < 12; b++) o.hist[b] = s_hist[b];
  }
}

// ---- 8. rolling resistance (dem.rs:1036-1055) and the final energy (:1057-1065) -------------------
extern "C" __global__ void __launch_bounds__(DE_BLOCK)
dem_finish(const DeDrum* __restrict__ drums, const float2* __restrict__ x, const float2* __restrict__ v,
           float* __restrict__ om, const uint32_t* __restrict__ spk, const float* __restrict__ slam,
           const uint32_t* __restrict__ cpk, const float* __restrict__ clam, const uint32_t* __restrict__ coff,
           const float* __restrict__ lam_bw, uint32_t stride, uint32_t pcap, DeOut* __restrict__ out) {
  __shared__ float s_tl[DE_MAX_BALLS];
  __shared__ __align__(16) float s_e[3 * DE_BLOCK];
  const DeDrum D = drums[blockIdx.x];
  const uint32_t t = threadIdx.x, n = D.n;
  const size_t sb = (size_t)blockIdx.x * stride, pb = (size_t)blockIdx.x * pcap, ob = (size_t)blockIdx.x * (DE_LCAP + 1);
  if (D.rolling > 0.0f && D.inertia > 0.0f) {
    for (uint32_t i = t; i < n; i += DE_BLOCK) s_tl[i] = 0.0f;
    __syncthreads();
    // total normal impulse per ball in contact order (contacts without impulse add max(lam, 0) = +0.0)
    auto pair = [&](uint32_t i, uint32_t j, float l) {
      s_tl[i] += l;
      s_tl[j] += l;
    };
    de_contact_sweep(pair, out[blockIdx.x], spk + pb, slam + pb, cpk + pb, clam + pb, coff + ob);
    for (uint32_t i = t; i < n; i += DE_BLOCK) {
      const float tl = s_tl[i] + fmaxf(lam_bw[sb + i], 0.0f);
      const float o = om[D.b0 + i];
      if (tl <= 0.0f || o == 0.0f) continue;
      const float nf = tl / D.dt;
      const float mdo = D.rolling * nf * D.r / D.inertia * D.dt;
      const float dom = fminf(mdo, fabsf(o));
      const float sg = isnan(o) ? o : (signbit(o) ? -1.0f : 1.0f);
      om[D.b0 + i] = o - sg * dom;
    }
    __syncthreads();
  }
  float e = 0.0f;
  double ed = 0.0;
  for (uint32_t base = 0; base < n; base += DE_BLOCK) {
    const uint32_t i = base + t;
    if (i < n) de_mech_terms(D, x[D.b0 + i], v[D.b0 + i], om[D.b0 + i], &s_e[3 * t]);
    __syncthreads();
    if (t == 0) {
      const uint32_t m = min(n - base, (uint32_t)DE_BLOCK);
      if (D.f64)
        for (uint32_t q = 0; q < 3 * m; q++) ed += (double)s_e[q];
      else
        for (uint32_t q = 0; q < 3 * m; q++) e += s_e[q];
    }
    __syncthreads();
  }
  if (t == 0) out[blockIdx.x].e_final = D.f64 ? ed : (double)e;
}

// ---- host -----------------------------------------------------------------------------------------
struct mg_dem {
  uint32_t nd = 0, max_balls = 0, pcap = 0, rcap = 0;
  char* dev = nullptr;
  DeDrum* drums = nullptr;
  float2* lift = nullptr;
  uint32_t lift_cap = 0;
  DeOut* out = nullptr;
  float2 *x0 = nullptr, *vpre = nullptr;
  float *th0 = nullptr, *ompre = nullptr, *lam_bw = nullptr, *lam = nullptr, *ebin = nullptr, *lam_s = nullptr;
  uint32_t *lev = nullptr, *lpk = nullptr, *lk = nullptr;  // a level schedule (projection, then contacts)
  uint32_t *spk = nullptr, *coff = nullptr;  // active contacts in book order; contact level starts
  float *slam = nullptr, *clam = nullptr;
  float* gap = nullptr;  // the projection's near-pair gap per drum slot (a schedule choice, not state)
  uint32_t* flagged = nullptr;
  uint2* pairs = nullptr;
  DeItem* sorted = nullptr;
  mg_impact_record* rec = nullptr;
  std::vector<DeOut> h_out;
  std::string msg;
  mgk::Timer tm;
};

namespace {

int de_fail(mg_dem* h, char* err, size_t errlen, int rc, const std::string& m) {
  if (h) h->msg = m;
  if (err && errlen) {
    std::strncpy(err, m.c_str(), errlen - 1);
    err[errlen - 1] = 0;
  }
  return rc;
}

// impact_energy_bin (dem.rs:99-109) with the host libm's log10f.
int ebin_host(float e) {
  if (!std::isfinite(e) || e <= 0.0f) return -1;
  const float ec = std::min(std::max(e, 1e-6f), 1.0f);
  const float lmin = std::log10(1e-6f), lmax = std::log10(1.0f);
  const float t = (std::log10(ec) - lmin) / (lmax - lmin);
  const float u = t * 12.0f;
  const uint64_t idx = !(u > 0.0f) ? 0u : (u >= 1.8446744073709552e19f ? UINT64_MAX : (uint64_t)u);
  return (int)std::min<uint64_t>(idx, 11);
}

// thr[k - 1] = the smallest positive finite float whose bin is at least k (k = 1..11).
void ebin_thresholds(float* thr) {
  for (int k = 1; k < 12; k++) {
    uint32_t lo = 1, hi = 0x7f7fffffu;  // bin(lo) < k <= bin(FLT_MAX) = 11
    while (hi - lo > 1) {
      const uint32_t mid = lo + (hi - lo) / 2;
      float f;
      std::memcpy(&f, &mid, 4);
      if (ebin_host(f) >= k) hi = mid; else lo = mid;
    }
    std::memcpy(&thr[k - 1], &hi, 4);
  }
}

// Drum::toi_swept (geometry.rs:146-194) on the host, rotating with the host libm's sinf / cosf.
bool toi_swept_host(const mgg::Geo& g, float x0x, float x0y, float x1x, float x1y, float a0, float a1, float rb,
                    float& out) {
  const float c0 = std::cos(-a0), s0 = std::sin(-a0);
  if (mgg::sdf_local(x0x * c0 - x0y * s0, x0x * s0 + x0y * c0, g) - rb <= 0.0f) return false;
  const float px = x1x - x0x, py = x1y - x0y;
  const float pl = std::sqrt(px * px + py * py);
  const float da = a1 - a0;
  const float pos_bound = std::sqrt(x0x * x0x + x0y * x0y) + pl;
  const float lsb = pl + std::fabs(da) * pos_bound;
  if (lsb < 1e-9f) return false;
  float t = 0.0f;
  for (int it = 0; it < 32; it++) {
    const float at = a0 + da * t;
    const float xtx = x0x + px * t, xty = x0y + py * t;
    const float c = std::cos(-at), s = std::sin(-at);
    const float d = mgg::sdf_local(xtx * c - xty * s, xtx * s + xty * c, g) - rb;
    if (d <= 0.0f) {
      out = t;
      return true;
    }
    t += std::fmax(d / lsb, 1e-6f);
    if (t >= 1.0f) return false;
  }
  out = std::fmin(t, 1.0f);
  return true;
}

}  // namespace

extern "C" {

int mg_dem_create(uint32_t n_drums, uint32_t max_balls, uint32_t pairs_per_ball, uint32_t max_impacts, mg_dem** out,
                  char* err, size_t errlen) {
  if (!out || n_drums == 0) return de_fail(nullptr, err, errlen, MG_E_ARG, "n_drums must be positive");
  if (max_balls == 0 || max_balls > DE_MAX_BALLS)
    return de_fail(nullptr, err, errlen, MG_E_ARG, "max_balls must be in 1..2048 (the ball state of a population "
                                                   "is held in LDS during the pair sweeps)");
  if (pairs_per_ball == 0) pairs_per_ball = 24;
  auto* h = new mg_dem();
  h->nd = n_drums;
  h->max_balls = max_balls;
  h->pcap = std::max<uint32_t>(pairs_per_ball * max_balls, 64);
  h->rcap = std::max<uint32_t>(max_impacts, 1);
  h->lift_cap = 64 * n_drums;
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t nb = (size_t)n_drums * max_balls, np = (size_t)n_drums * h->pcap;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off += up(bytes);
    return o;
  };
  const size_t o_drums = take(n_drums * sizeof(DeDrum)), o_lift = take(h->lift_cap * sizeof(float2));
  const size_t o_out = take(n_drums * sizeof(DeOut)), o_x0 = take(nb * 8), o_vp = take(nb * 8);
  const size_t o_th0 = take(nb * 4), o_op = take(nb * 4), o_lbw = take(nb * 4), o_flag = take(nb * 4);
  const size_t o_ebin = take(16 * 4), o_pairs = take(np * 8), o_lam = take(np * 4), o_sorted = take(np * sizeof(DeItem));
  const size_t o_lev = take(np * 4), o_lpk = take(np * 4), o_lk = take(np * 4), o_lam_s = take(np * 4);
  const size_t o_spk = take(np * 4), o_slam = take(np * 4), o_clam = take(np * 4);
  const size_t o_coff = take((size_t)n_drums * (DE_LCAP + 1) * 4);
  const size_t o_rec = take((size_t)n_drums * h->rcap * sizeof(mg_impact_record));
  const size_t o_gap = take((size_t)n_drums * 4);
  if (hipMalloc(&h->dev, off) != hipSuccess) {
    delete h;
    return de_fail(nullptr, err, errlen, MG_E_DEVICE, "hipMalloc failed for the DEM workspace");
  }
  h->drums = reinterpret_cast<DeDrum*>(h->dev + o_drums);
  h->lift = reinterpret_cast<float2*>(h->dev + o_lift);
  h->out = reinterpret_cast<DeOut*>(h->dev + o_out);
  h->x0 = reinterpret_cast<float2*>(h->dev + o_x0);
  h->vpre = reinterpret_cast<float2*>(h->dev + o_vp);
  h->th0 = reinterpret_cast<float*>(h->dev + o_th0);
  h->ompre = reinterpret_cast<float*>(h->dev + o_op);
  h->lam_bw = reinterpret_cast<float*>(h->dev + o_lbw);
  h->flagged = reinterpret_cast<uint32_t*>(h->dev + o_flag);
  h->ebin = reinterpret_cast<float*>(h->dev + o_ebin);
  h->pairs = reinterpret_cast<uint2*>(h->dev + o_pairs);
  h->lam = reinterpret_cast<float*>(h->dev + o_lam);
  h->lev = reinterpret_cast<uint32_t*>(h->dev + o_lev);
  h->lpk = reinterpret_cast<uint32_t*>(h->dev + o_lpk);
  h->lk = reinterpret_cast<uint32_t*>(h->dev + o_lk);
  h->lam_s = reinterpret_cast<float*>(h->dev + o_lam_s);
  h->spk = reinterpret_cast<uint32_t*>(h->dev + o_spk);
  h->slam = reinterpret_cast<float*>(h->dev + o_slam);
  h->clam = reinterpret_cast<float*>(h->dev + o_clam);
  h->coff = reinterpret_cast<uint32_t*>(h->dev + o_coff);
  h->sorted = reinterpret_cast<DeItem*>(h->dev + o_sorted);
  h->rec = reinterpret_cast<mg_impact_record*>(h->dev + o_rec);
  h->gap = reinterpret_cast<float*>(h->dev + o_gap);
  float thr[11];
  ebin_thresholds(thr);
  const std::vector<float> gap0(n_drums, DE_GAP);
  if (hipMemcpy(h->ebin, thr, sizeof thr, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(h->gap, gap0.data(), gap0.size() * 4, hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(h->dev);
    delete h;
    return de_fail(nullptr, err, errlen, MG_E_DEVICE, "HIP copy error");
  }
  h->h_out.resize(n_drums);
  *out = h;
  return MG_OK;
}

void mg_dem_destroy(mg_dem* h) {
  if (!h) return;
  if (h->dev) (void)hipFree(h->dev);
  delete h;
}

int mg_dem_step(mg_dem* h, const mg_dem_drum* drums, float* d_x, float* d_v, float* d_theta, float* d_omega,
                const float* d_ext_impulse, const float* d_ext_angular, mg_dem_stats* stats, char* err,
                size_t errlen) {
  if (!h || !drums || !stats) return de_fail(h, err, errlen, MG_E_ARG, "null argument");
  const uint32_t nd = h->nd;
  std::vector<DeDrum> hd(nd);
  std::vector<float2> lift;
  uint64_t tb = 0;
  bool any_ext = false;
  for (uint32_t d = 0; d < nd; d++) {
    const mg_dem_drum& m = drums[d];
    if (m.n_balls > h->max_balls)
      return de_fail(h, err, errlen, MG_E_ARG, "drum " + std::to_string(d) + ": more balls than the handle's max_balls");
    DeDrum& D = hd[d];
    D.radius = m.radius_m;
    D.omega = m.omega;
    D.nl = m.lifter_count;
    D.lift0 = (uint32_t)lift.size();
    D.rt = m.radius_m - m.lifter_height_m;
    D.bh = m.lifter_base_width_m * 0.5f;
    D.thw = m.lifter_top_width_m * 0.5f;
    const float phase = m.lifter_phase_deg * (float)0.017453292519943295;  // f32::to_radians
    const float tau = 6.28318530717958647692f;
    for (uint32_t k = 0; k < m.lifter_count; k++) {
      const float a = phase + (float)k * tau / (float)m.lifter_count;
      lift.push_back(make_float2(std::sin(a), std::cos(a)));
    }
    const float an = m.drum_angle + m.omega * m.dt;  // dem.rs:610
    D.rc0 = std::cos(-m.drum_angle);
    D.rs0 = std::sin(-m.drum_angle);
    D.rc1 = std::cos(-an);
    D.rs1 = std::sin(-an);
    D.wc1 = std::cos(an);
    D.ws1 = std::sin(an);
    D.r = m.ball_radius;
    D.mass = m.ball_mass;
    D.inertia = m.ball_inertia;
    D.w = m.ball_mass > 0.0f ? 1.0f / m.ball_mass : 0.0f;        // Balls::inv_mass (dem.rs:335-341)
    D.w_rot = m.ball_inertia > 0.0f ? 1.0f / m.ball_inertia : 0.0f;  // inv_inertia (:343-349)
    D.thr = 0.1428f * std::sqrt(9.81f * 2.0f * m.ball_radius);    // restitution_velocity_threshold (:44-46)
    D.rest_bb = m.restitution_ball_ball;
    D.rest_bw = m.restitution_ball_wall;
    D.fric_bb = m.friction_ball_ball;
    D.fric_bw = m.friction_ball_wall;
    D.fric_bb_s = m.friction_ball_ball_static;
    D.fric_bw_s = m.friction_ball_wall_static;
    D.fric_vs = m.friction_velocity_scale_m_s;
    D.rolling = m.rolling_friction;
    D.iters = m.iterations;
    D.dt = m.dt;
    D.b0 = (uint32_t)tb;
    // an empty population or a non-positive step leaves the state as it is (dem.rs:542-544)
    D.n = (m.n_balls == 0 || !(m.dt > 0.0f)) ? 0u : m.n_balls;
    D.has_ext = m.has_external ? 1u : 0u;
    D.f64 = m.energy_f64 ? 1u : 0u;
    any_ext |= D.has_ext != 0;
    tb += m.n_balls;
  }
  if (tb && (!d_x || !d_v || !d_theta || !d_omega)) return de_fail(h, err, errlen, MG_E_ARG, "null state array");
  if (any_ext && (!d_ext_impulse || !d_ext_angular)) return de_fail(h, err, errlen, MG_E_ARG, "null coupling impulses");
  if (lift.size() > h->lift_cap) return de_fail(h, err, errlen, MG_E_ARG, "more than 64 lifters per drum on average");
  if (lift.empty()) lift.push_back(make_float2(0.0f, 1.0f));
  bool ok = hipMemcpy(h->drums, hd.data(), nd * sizeof(DeDrum), hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(h->lift, lift.data(), lift.size() * sizeof(float2), hipMemcpyHostToDevice) == hipSuccess &&
            hipMemset(h->out, 0, nd * sizeof(DeOut)) == hipSuccess;
  if (!ok) return de_fail(h, err, errlen, MG_E_DEVICE, "HIP copy error");
  auto* x = reinterpret_cast<float2*>(d_x);
  auto* v = reinterpret_cast<float2*>(d_v);
  const uint32_t S = h->max_balls, P = h->pcap;
  const dim3 G(nd), B(DE_BLOCK);
  MG_TIMED_LAUNCH(h->tm, dem_predict, G, B, 0, h->drums, x, v, d_theta, d_omega,
                     reinterpret_cast<const float2*>(d_ext_impulse), d_ext_angular, h->x0, h->vpre, h->th0, h->ompre, S,
                     h->out);
  MG_TIMED_LAUNCH(h->tm, dem_grid, G, B, 0, h->drums, x, h->pairs, h->lam, P, h->out);
  MG_TIMED_LAUNCH(h->tm, dem_ccd_flag, G, B, 0, h->drums, h->lift, x, h->flagged, S, h->out);
  ok = hipGetLastError() == hipSuccess &&
       hipMemcpy(h->h_out.data(), h->out, nd * sizeof(DeOut), hipMemcpyDeviceToHost) == hipSuccess;
  if (!ok) return de_fail(h, err, errlen, MG_E_DEVICE, "HIP error in the DEM predict / grid kernels");
  for (uint32_t d = 0; d < nd; d++)
    if (h->h_out[d].pair_overflow)
      return de_fail(h, err, errlen, MG_E_CAPACITY,
                     "drum " + std::to_string(d) + ": " + std::to_string(h->h_out[d].n_pairs) +
                         " candidate pairs exceed the handle's capacity (create with a larger pairs_per_ball)");
  // swept time of impact of the balls ending deep in the wall, on the host (dem.rs:611-637)
  std::vector<uint2> who;
  std::vector<float2> pos;
  std::vector<uint32_t> swept(nd, 0);
  for (uint32_t d = 0; d < nd; d++) {
    const uint32_t k = h->h_out[d].n_ccd;
    if (!k) continue;
    std::vector<uint32_t> idx(k);
    ok = hipMemcpy(idx.data(), h->flagged + (size_t)d * S, k * 4, hipMemcpyDeviceToHost) == hipSuccess;
    if (!ok) return de_fail(h, err, errlen, MG_E_DEVICE, "HIP copy error");
    std::sort(idx.begin(), idx.end());
    const DeDrum& D = hd[d];
    mgg::Geo g{D.radius, D.omega, D.nl, D.rt, D.bh, D.thw, lift.data() + D.lift0};
    const float an = drums[d].drum_angle + drums[d].omega * drums[d].dt;
    for (uint32_t i : idx) {
      float2 p0, p1;
      ok = hipMemcpy(&p0, h->x0 + (size_t)d * S + i, 8, hipMemcpyDeviceToHost) == hipSuccess &&
           hipMemcpy(&p1, x + D.b0 + i, 8, hipMemcpyDeviceToHost) == hipSuccess;
      if (!ok) return de_fail(h, err, errlen, MG_E_DEVICE, "HIP copy error");
      float tt;
      if (!toi_swept_host(g, p0.x, p0.y, p1.x, p1.y, drums[d].drum_angle, an, D.r, tt)) continue;
      const float toi = std::fmin(1.0f, tt);
      if (!(toi < 1.0f)) continue;
      const float px = p1.x - p0.x, py = p1.y - p0.y;
      const float pl = std::sqrt(px * px + py * py);
      const float skin = 1e-2f * D.r;
      const float ts = pl > 1e-9f ? skin / pl : 1.0f;
      const float tc = std::fmin(toi + ts, 1.0f);
      who.push_back(make_uint2(d, i));
      pos.push_back(make_float2(p0.x + px * tc, p0.y + py * tc));
      swept[d]++;
    }
  }
  if (!who.empty()) {
    void* dw = nullptr;
    ok = hipMalloc(&dw, who.size() * 16) == hipSuccess &&
         hipMemcpy(dw, who.data(), who.size() * 8, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(static_cast<char*>(dw) + who.size() * 8, pos.data(), pos.size() * 8, hipMemcpyHostToDevice) ==
             hipSuccess;
    if (ok)
      hipLaunchKernelGGL(dem_ccd_apply, dim3((uint32_t)((who.size() + 255) / 256)), dim3(256), 0, nullptr, h->drums,
                         static_cast<const uint2*>(dw),
                         reinterpret_cast<const float2*>(static_cast<char*>(dw) + who.size() * 8), (uint32_t)who.size(),
                         x);
    ok = ok && hipDeviceSynchronize() == hipSuccess;
    if (dw) (void)hipFree(dw);
    if (!ok) return de_fail(h, err, errlen, MG_E_DEVICE, "HIP error applying the swept clamp");
  }
  MG_TIMED_LAUNCH(h->tm, dem_project, G, B, 0, h->drums, h->lift, x, v, d_theta, d_omega, h->x0, h->th0, h->pairs,
                     h->lam, h->lam_bw, h->lev, h->lpk, h->lk, h->lam_s, h->spk, reinterpret_cast<uint32_t*>(h->clam),
                     h->slam, h->gap, S, P, h->out);
  MG_TIMED_LAUNCH(h->tm, dem_sort, G, B, 0, h->drums, h->pairs, h->lam, h->sorted, h->spk, h->slam, h->lev, h->lpk,
                     h->lk, h->clam, h->coff, P, h->out);
  MG_TIMED_LAUNCH(h->tm, dem_friction, G, B, 0, h->drums, h->lift, x, v, d_theta, d_omega, h->x0, h->th0, h->spk,
                     h->slam, h->lpk, h->clam, h->coff, h->lam_bw, S, P, h->out);
  MG_TIMED_LAUNCH(h->tm, dem_restitution, G, B, 0, h->drums, h->lift, x, v, h->vpre, h->ompre, h->spk, h->slam,
                     h->lpk, h->clam, h->coff, h->lam_bw, h->ebin, h->rec, h->rcap, S, P, h->out);
  MG_TIMED_LAUNCH(h->tm, dem_finish, G, B, 0, h->drums, x, v, d_omega, h->spk, h->slam, h->lpk, h->clam, h->coff,
                     h->lam_bw, S, P, h->out);
  ok = hipGetLastError() == hipSuccess &&
       hipMemcpy(h->h_out.data(), h->out, nd * sizeof(DeOut), hipMemcpyDeviceToHost) == hipSuccess;
  if (!ok) return de_fail(h, err, errlen, MG_E_DEVICE, "HIP error in the DEM kernels");
  h->tm.collect();
  int rc = MG_OK;
  for (uint32_t d = 0; d < nd; d++) {
    const DeOut& o = h->h_out[d];
    mg_dem_stats& s = stats[d];
    std::memset(&s, 0, sizeof s);
    if (hd[d].n == 0) continue;  // DemStepStats::default()
    s.wall_work_j = (float)o.ww;
    s.collision_count = o.coll;
    for (int b = 0; b < 12; b++) s.impact_energy_histogram[b] = o.hist[b];
    s.dissipated_energy_j = hd[d].f64 ? (float)(o.e_pred + o.ww - o.e_final)
                                      : (float)o.e_pred + (float)o.ww - (float)o.e_final;
    s.max_substep_displacement_over_diameter = o.maxd;
    s.coupling_work_j = (float)o.cw;
    s.n_pairs = o.n_pairs;
    s.n_impacts = o.n_imp;
    s.n_swept = swept[d];
    s.projection_restarts = o.restart;
    if (o.n_imp > h->rcap && rc == MG_OK)
      rc = de_fail(h, err, errlen, MG_E_CAPACITY,
                   "drum " + std::to_string(d) + ": more impact records than the handle holds (the count is exact, "
                   "the records beyond max_impacts are dropped)");
  }
  return rc;
}

int mg_dem_set_timing(mg_dem* h, int on) {
  if (!h) return MG_E_ARG;
  h->tm.set(on != 0);
  return MG_OK;
}

int mg_dem_kernel_times(mg_dem* h, char* json, size_t cap) {
  if (!h || !json || cap == 0) return MG_E_ARG;
  const std::string s = h->tm.json();
  if (s.size() + 1 > cap) return -(int)(s.size() + 1);
  std::memcpy(json, s.c_str(), s.size() + 1);
  return (int)s.size();
}

const mg_impact_record* mg_dem_records_device(const mg_dem* h, uint32_t* stride) {
  if (!h) return nullptr;
  if (stride) *stride = h->rcap;
  return h->rec;
}

int mg_dem_impacts(mg_dem* h, uint32_t drum, mg_impact_record* out, uint32_t cap, uint32_t* n) {
  if (!h || drum >= h->nd || !n) return MG_E_ARG;
  const uint32_t have = std::min(h->h_out[drum].n_imp, h->rcap);
  *n = h->h_out[drum].n_imp;
  if (!out || cap == 0 || have == 0) return MG_OK;
  const uint32_t k = std::min(have, cap);
  return hipMemcpy(out, h->rec + (size_t)drum * h->rcap, k * sizeof(mg_impact_record), hipMemcpyDeviceToHost) ==
                 hipSuccess
             ? MG_OK
             : MG_E_DEVICE;
}

}  // extern "C"
