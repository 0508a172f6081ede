// Fluid free-surface extraction for a batch of drums (crates/mill-core/src/surface.rs:39-420).
//
// The reference splats every fluid particle onto a 128 x 128 grid of points spanning [-R, R]^2
// (kernel (1 - r^2 / rs^2)^3 of radius rs = 1.5 h over the grid-index box round((p -+ rs + R) /
// cell), cell = 2R / 127), zeroes the points outside the wall or inside a lifter, runs marching
// squares at half the analytic bulk value (half the measured peak if that finds no cell), chains
// the segments into polylines through a map from grid edge to segments, smooths each with one
// Chaikin pass and projects points outside the wall back onto it.
//
// Device side, in four launches:
//   sf_boxes   one lane per particle: its grid-index box, four bytes; per 64 particles the mask of
//              the regions their boxes meet.
//   sf_field   one wave per region of 8 x 8 grid points of a drum, one lane per point: the wave
//              takes the particles whose boxes meet the region in index order (batches of 64
//              named by a region mask sf_boxes writes) and each lane adds the kernel values of
//              those whose box holds its point, so each point's sum is formed in particle order as
//              the reference's scatter forms it. Then the wall mask.
//   sf_cells   one workgroup per drum: the 127 x 127 cells in row order, 1024 at a time, each
//              emitting its 0-2 segments (edge ids and interpolated crossing points) at ordered
//              offsets from a workgroup scan, so the segment list is the reference's. No segment
//              at the analytic threshold: the peak by a workgroup max, and the pass again at half
//              of it.
//   sf_pack    the drums' segment lists packed end to end for one copy to the host.
// Host side: the chain walk over the segment list (each grid edge belongs to at most two
// segments, both kept in segment order, which is the reference's HashMap<EdgeId, Vec<usize>>
// order), Chaikin and the wall projection, O(contour length) per drum.
//
// All arithmetic is the reference's f32 in its order (-ffp-contract=off, IEEE division and sqrt);
// the drum-frame rotations take (cos, sin) from the host libm, as the reference's from_angle.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/millgpu.h"
#include "mg_drumgeo.h"

#define SF_G 128
#define SF_R 8  // a region of sf_field: 8 x 8 grid points, one per lane of a wave
#define SF_RX (SF_G / SF_R)
#define SF_REGIONS (SF_RX * SF_RX)
#define SF_MASK_WORDS (SF_REGIONS / 64)
#define SF_NCELL (SF_G - 1)
#define SF_CELLS (SF_NCELL * SF_NCELL)
#define SF_SEG_CAP (2 * SF_CELLS)
#define SF_CELL_BLOCK 1024
#define SF_CELL_WAVES (SF_CELL_BLOCK / 64)
#define SF_EDGE_IDS (2 * SF_G * SF_G)
static_assert(SF_G % SF_R == 0 && SF_REGIONS % 64 == 0, "whole regions");

namespace {

struct SfDrum {
  float radius, cell, rs, thr;  // thr: 0.5 * analytic_phi_full
  float r2, inv_r2;             // rs * rs and its correctly rounded reciprocal
  float rc, rsn, wc, ws;        // from_angle(-drum_angle), from_angle(drum_angle)
  uint32_t nl, lift0;
  float rt, bh, thw;
  uint32_t f0, n;  // first particle and count; n = 0: no extraction (empty fluid or radius <= 0)
  uint32_t b0;     // first 64-particle batch of the drum in the batch region masks
};
// One marching-squares segment between edges ea and eb (edge id = vertical << 14 | j << 7 | i)
// with crossing points a, b (kept apart, as float4 (a, b), for the rare chains that need them) and
// the projected Chaikin points of both directions: q[0], q[1] = project(lerp(a, b, 0.25)),
// project(lerp(a, b, 0.75)); q[2], q[3] the same from b to a.
struct SfSeg {
  uint32_t ea, eb;
  float2 q[4];
};
struct SfOut {
  uint32_t nseg, fallback;
  float thr, peak;
};

// grid_index (surface.rs:140-143): round half away from zero, clamp, NaN -> 0.
__device__ __forceinline__ uint32_t sf_index(float world, float r, float cell) {
  float raw = roundf((world + r) / cell);
  if (raw < 0.0f) raw = 0.0f;
  if (raw > (float)(SF_G - 1)) raw = (float)(SF_G - 1);
  return isnan(raw) ? 0u : (uint32_t)raw;
}
// splat_kernel (surface.rs:39-46) with radius2 = radius * radius and y its correctly rounded
// reciprocal. The quotient r2 / radius2 is formed as q0 = r2 y, q = q0 + (r2 - radius2 q0) y with
// fused operations, which is the correctly rounded quotient when y is (Markstein's theorem: y
// within half an ulp of 1 / radius2, q0 within one ulp of the quotient, the residual exact); the
// residual is exact unless it falls below the normal range, so dividends under 2^-100 take the
// IEEE division. An exhaustive comparison over every dividend in [2^-100, radius2) for a dozen
// divisors, significands of all ones included, found no difference.
__device__ __forceinline__ float sf_quotient(float r2, float radius2, float y) {
  const float q0 = r2 * y;
  return __builtin_fmaf(__builtin_fmaf(-q0, radius2, r2), y, q0);
}
__device__ __forceinline__ mgg::Geo sf_geo(const SfDrum& D, const float2* lift) {
  return mgg::Geo{D.radius, 0.0f, D.nl, D.rt, D.bh, D.thw, lift + D.lift0};
}
// interp_edge (surface.rs:166-192) of edge id e at threshold thr on drum field f.
__device__ __forceinline__ float2 sf_interp(uint32_t e, const float* __restrict__ f, float thr, float r,
                                            float cell) {
  const uint32_t i0 = e & 127u, j0 = (e >> 7) & 127u, vert = e >> 14;
  const uint32_t i1 = vert ? i0 : i0 + 1u, j1 = vert ? j0 + 1u : j0;
  const float v0 = f[j0 * SF_G + i0], v1 = f[j1 * SF_G + i1];
  float t = 0.5f;
  if (fabsf(v1 - v0) > 1e-9f) {
    t = (thr - v0) / (v1 - v0);
    if (t < 0.0f) t = 0.0f;
    if (t > 1.0f) t = 1.0f;
  }
  const float p0x = -r + (float)i0 * cell, p0y = -r + (float)j0 * cell;
  const float p1x = -r + (float)i1 * cell, p1y = -r + (float)j1 * cell;
  return make_float2(p0x + (p1x - p0x) * t, p0y + (p1y - p0y) * t);  // Vec2::lerp
}
// cell_edge_id (surface.rs:209-217) of side 0 bottom, 1 right, 2 top, 3 left.
__device__ __forceinline__ uint32_t sf_edge(uint32_t ci, uint32_t cj, uint32_t side) {
  switch (side) {
    case 0: return (cj << 7) | ci;
    case 1: return (1u << 14) | (cj << 7) | (ci + 1u);
    case 2: return ((cj + 1u) << 7) | ci;
    default: return (1u << 14) | (cj << 7) | ci;
  }
}
// cell_edge_pairs (surface.rs:194-207): up to two (side, side) pairs packed as 4-bit nibbles
// (first pair in the low byte); the count in bits 16-17.
__device__ __forceinline__ uint32_t sf_pairs(uint32_t c) {
  switch (c) {
    case 1: case 14: return (1u << 16) | 0x03u;         // (3, 0)
    case 2: case 13: return (1u << 16) | 0x10u;         // (0, 1)
    case 3: case 12: return (1u << 16) | 0x13u;         // (3, 1)
    case 4: case 11: return (1u << 16) | 0x21u;         // (1, 2)
    case 6: case 9: return (1u << 16) | 0x20u;          // (0, 2)
    case 7: case 8: return (1u << 16) | 0x23u;          // (3, 2)
    case 5: return (2u << 16) | 0x2103u;                // (3, 0), (1, 2)
    case 10: return (2u << 16) | 0x2310u;               // (0, 1), (3, 2)
    default: return 0u;
  }
}

}  // namespace

// Grid-index box of every particle, i0 | i1 << 8 | j0 << 16 | j1 << 24, and for every batch of 64
// particles (one wave) the mask of the 256 regions of 8 x 8 points its boxes meet (bit
// (j / 8) * 16 + i / 8, four words). Grid (ceil(max n / 256), drums).
extern "C" __global__ void __launch_bounds__(256)
sf_boxes(const SfDrum* __restrict__ drums, const float2* __restrict__ x, uint32_t* __restrict__ boxes,
         uint64_t* __restrict__ batch_regions) {
  const SfDrum D = drums[blockIdx.y];
  const uint32_t i = blockIdx.x * 256u + threadIdx.x, lane = threadIdx.x & 63u;
  if ((i & ~63u) >= D.n) return;  // uniform over the wave
  uint64_t m[SF_MASK_WORDS] = {};
  if (i < D.n) {
    const float2 p = x[D.f0 + i];
    const uint32_t i0 = sf_index(p.x - D.rs, D.radius, D.cell), i1 = sf_index(p.x + D.rs, D.radius, D.cell);
    const uint32_t j0 = sf_index(p.y - D.rs, D.radius, D.cell), j1 = sf_index(p.y + D.rs, D.radius, D.cell);
    boxes[D.f0 + i] = i0 | (i1 << 8) | (j0 << 16) | (j1 << 24);
    if (i0 <= i1 && j0 <= j1) {  // an empty box (rs < 0) meets no region
      const uint32_t rx0 = i0 / SF_R, rx1 = i1 / SF_R;
      const uint64_t row = ((1ull << (rx1 - rx0 + 1u)) - 1ull) << rx0;  // 16 regions per row
#pragma unroll
      for (uint32_t w = 0; w < SF_MASK_WORDS; w++) {  // word w holds region rows 4w .. 4w + 3
        for (uint32_t ry = max(j0 / SF_R, 4u * w); ry <= min(j1 / SF_R, 4u * w + 3u); ry++)
          m[w] |= row << ((ry - 4u * w) * SF_RX);
      }
    }
  }
#pragma unroll
  for (uint32_t w = 0; w < SF_MASK_WORDS; w++) {
    uint64_t v = m[w];
    for (uint32_t s = 32; s >= 1; s >>= 1) {
      const uint32_t lo = __shfl_xor((uint32_t)v, (int)s, 64), hi = __shfl_xor((uint32_t)(v >> 32), (int)s, 64);
      v |= ((uint64_t)hi << 32) | lo;
    }
    if (lane == 0) batch_regions[(size_t)(D.b0 + i / 64u) * SF_MASK_WORDS + w] = v;
  }
}

// build_field (surface.rs:84-133) of one region of 8 x 8 grid points, one wave per region and one
// lane per point. The wave takes the particles whose boxes meet the region in index order (the
// batches of 64 whose mask names the region), and every lane adds the kernel value of each one whose
// box holds its point, so each point's sum is formed in particle order as the reference's scatter
// forms it. Then the wall mask. Grid (drums, 256 regions), 64 lanes.
extern "C" __global__ void __launch_bounds__(64)
sf_field(const SfDrum* __restrict__ drums, const float2* __restrict__ lift, const float2* __restrict__ x,
         const uint32_t* __restrict__ boxes, const uint64_t* __restrict__ batch_regions, float* __restrict__ field) {
  const SfDrum D = drums[blockIdx.x];
  const uint32_t lane = threadIdx.x, reg = blockIdx.y, word = reg / 64u, bit = reg % 64u;
  const uint32_t ri0 = (reg % SF_RX) * SF_R, rj0 = (reg / SF_RX) * SF_R;
  const uint32_t gi = ri0 + (lane & 7u), gj = rj0 + (lane >> 3);
  const float gx = -D.radius + (float)gi * D.cell, gy = -D.radius + (float)gj * D.cell;
  float f = 0.0f;
  const uint32_t nb = (D.n + 63u) / 64u;
  for (uint32_t bb0 = 0; bb0 < nb; bb0 += 64) {  // 64 batch masks at a time; nb is uniform
    const bool named = bb0 + lane < nb &&
                       ((batch_regions[(size_t)(D.b0 + bb0 + lane) * SF_MASK_WORDS + word] >> bit) & 1ull);
    uint64_t mb = __ballot(named);
    while (mb) {
      const uint32_t base = (bb0 + (uint32_t)(__ffsll((unsigned long long)mb) - 1)) * 64u;
      mb &= mb - 1ull;
      const uint32_t k = base + lane;
      uint32_t b = 0;
      bool hit = false;
      float2 p = make_float2(0.0f, 0.0f);
      if (k < D.n) {
        b = boxes[D.f0 + k];
        hit = ((b >> 16) & 255u) <= rj0 + SF_R - 1u && (b >> 24) >= rj0 && (b & 255u) <= ri0 + SF_R - 1u &&
              ((b >> 8) & 255u) >= ri0;
        if (hit) p = x[D.f0 + k];
      }
      uint64_t m = __ballot(hit);
      while (m) {
        const int src = __ffsll((unsigned long long)m) - 1;
        m &= m - 1ull;
        const uint32_t bx = (uint32_t)__builtin_amdgcn_readlane((int)b, src);
        const float px = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(p.x), src));
        const float py = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(p.y), src));
        const bool in = gi >= (bx & 255u) && gi <= ((bx >> 8) & 255u) && gj >= ((bx >> 16) & 255u) && gj <= (bx >> 24);
        const float dx = gx - px, dy = gy - py;
        const float r2 = dx * dx + dy * dy;
        float q = sf_quotient(r2, D.r2, D.inv_r2);
        if (r2 < 0x1p-100f && r2 != 0.0f) q = r2 / D.r2;  // below the range of the fused quotient
        const float t = 1.0f - q;
        const float kv = r2 >= D.r2 ? 0.0f : t * t * t;  // splat_kernel (surface.rs:39-46)
        f += in ? kv : 0.0f;
      }
    }
  }
  if (f != 0.0f) {  // the wall / lifter mask (surface.rs:113-130)
    const mgg::Geo g = sf_geo(D, lift);
    if (mgg::sdf_local(gx * D.rc - gy * D.rsn, gx * D.rsn + gy * D.rc, g) < 0.0f) f = 0.0f;
  }
  field[(size_t)blockIdx.x * SF_G * SF_G + gj * SF_G + gi] = f;
}

namespace {

// One marching-squares pass (surface.rs:288-307) at threshold thr; returns the segment count.
// project_inside_wall (surface.rs:359-366)
__device__ __forceinline__ float2 sf_project(float2 p, const SfDrum& D, const mgg::Geo& g) {
  float nx, ny;
  const float d = mgg::sdf_world(p.x, p.y, D.rc, D.rsn, D.wc, D.ws, g, nx, ny);
  return d < 0.0f ? make_float2(p.x - d * nx, p.y - d * ny) : p;
}
__device__ __forceinline__ float2 sf_lerp(float2 a, float2 b, float t) {  // Vec2::lerp
  return make_float2(a.x + (b.x - a.x) * t, a.y + (b.y - a.y) * t);
}

__device__ uint32_t sf_pass(const float* __restrict__ f, float thr, const SfDrum& D, const mgg::Geo& g,
                            SfSeg* __restrict__ seg, float4* __restrict__ ab, uint32_t* s_w) {
  const float r = D.radius, cell = D.cell;
  const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
  uint32_t count = 0;
  for (uint32_t c0 = 0; c0 < SF_CELLS; c0 += SF_CELL_BLOCK) {
    const uint32_t c = c0 + t;
    uint32_t pr = 0, ci = 0, cj = 0;
    if (c < SF_CELLS) {
      cj = c / SF_NCELL;
      ci = c % SF_NCELL;
      const float* row = f + cj * SF_G + ci;
      const uint32_t code = (uint32_t)(row[0] > thr) | (uint32_t)(row[1] > thr) << 1 |
                            (uint32_t)(row[SF_G + 1] > thr) << 2 | (uint32_t)(row[SF_G] > thr) << 3;
      pr = sf_pairs(code);
    }
    const uint32_t ns = pr >> 16;
    const uint64_t m1 = __ballot(ns >= 1u), m2 = __ballot(ns == 2u);
    const uint64_t lt = (1ull << lane) - 1ull;
    const uint32_t pre = __popcll(m1 & lt) + __popcll(m2 & lt);
    if (lane == 0) s_w[w] = (uint32_t)(__popcll(m1) + __popcll(m2));
    __syncthreads();
    uint32_t off = count, tot = 0;
    for (uint32_t q = 0; q < SF_CELL_WAVES; q++) {
      off += q < w ? s_w[q] : 0u;
      tot += s_w[q];
    }
    for (uint32_t s = 0; s < ns; s++) {
      const uint32_t sides = (pr >> (8u * s)) & 255u;
      SfSeg o;
      o.ea = sf_edge(ci, cj, sides & 15u);
      o.eb = sf_edge(ci, cj, sides >> 4);
      const float2 a = sf_interp(o.ea, f, thr, r, cell), b = sf_interp(o.eb, f, thr, r, cell);
      o.q[0] = sf_project(sf_lerp(a, b, 0.25f), D, g);  // chaikin_smooth_closed (surface.rs:403-418)
      o.q[1] = sf_project(sf_lerp(a, b, 0.75f), D, g);
      o.q[2] = sf_project(sf_lerp(b, a, 0.25f), D, g);
      o.q[3] = sf_project(sf_lerp(b, a, 0.75f), D, g);
      seg[off + pre + s] = o;
      ab[off + pre + s] = make_float4(a.x, a.y, b.x, b.y);
    }
    count += tot;
    __syncthreads();
  }
  return count;
}

}  // namespace

// polygons_at_threshold's segment list (surface.rs:279-307) with extract_surface's fallback to the
// measured peak (surface.rs:253-276). Grid (drums), 1024 lanes.
extern "C" __global__ void __launch_bounds__(SF_CELL_BLOCK)
sf_cells(const SfDrum* __restrict__ drums, const float2* __restrict__ lift, const float* __restrict__ field,
         SfSeg* __restrict__ segs, float4* __restrict__ abs, SfOut* __restrict__ out) {
  __shared__ uint32_t s_w[SF_CELL_WAVES];
  __shared__ float s_m[SF_CELL_WAVES];
  const uint32_t d = blockIdx.x, t = threadIdx.x;
  const SfDrum D = drums[d];
  SfOut o{0u, 0u, D.thr, 0.0f};
  if (D.n == 0) {  // uniform over the workgroup
    if (t == 0) out[d] = o;
    return;
  }
  const float* f = field + (size_t)d * SF_G * SF_G;
  SfSeg* seg = segs + (size_t)d * SF_SEG_CAP;
  float4* ab = abs + (size_t)d * SF_SEG_CAP;
  const mgg::Geo g = sf_geo(D, lift);
  o.nseg = sf_pass(f, D.thr, D, g, seg, ab, s_w);
  if (o.nseg == 0) {
    float m = 0.0f;  // fold(0.0, f32::max): NaN never wins
    for (uint32_t k = t; k < SF_G * SF_G; k += SF_CELL_BLOCK) m = fmaxf(m, f[k]);
    for (uint32_t s = 32; s >= 1; s >>= 1) m = fmaxf(m, __shfl_xor(m, (int)s, 64));
    if ((t & 63u) == 0) s_m[t >> 6] = m;
    __syncthreads();
    float peak = 0.0f;
    for (uint32_t q = 0; q < SF_CELL_WAVES; q++) peak = fmaxf(peak, s_m[q]);
    o.peak = peak;
    if (peak > 0.0f) {
      o.fallback = 1u;
      o.thr = 0.5f * peak;
      o.nseg = sf_pass(f, o.thr, D, g, seg, ab, s_w);
    }
  }
  if (t == 0) out[d] = o;
}

// The drums' segment lists end to end: drum d's at the sum of the counts before it. Grid (drums), 256 lanes.
extern "C" __global__ void __launch_bounds__(256)
sf_pack(const SfSeg* __restrict__ segs, const float4* __restrict__ abs, const SfOut* __restrict__ out,
        SfSeg* __restrict__ packed, float4* __restrict__ packed_ab) {
  __shared__ uint32_t s_sum[4];
  const uint32_t d = blockIdx.x, t = threadIdx.x;
  uint32_t part = 0;
  for (uint32_t q = t; q < d; q += 256u) part += out[q].nseg;
  for (uint32_t s = 32; s >= 1; s >>= 1) part += __shfl_xor(part, (int)s, 64);
  if ((t & 63u) == 0) s_sum[t >> 6] = part;
  __syncthreads();
  const uint32_t base = s_sum[0] + s_sum[1] + s_sum[2] + s_sum[3];
  const uint32_t n = out[d].nseg;
  const SfSeg* src = segs + (size_t)d * SF_SEG_CAP;
  const float4* src_ab = abs + (size_t)d * SF_SEG_CAP;
  for (uint32_t k = t; k < n; k += 256u) {
    packed[base + k] = src[k];
    packed_ab[base + k] = src_ab[k];
  }
}

// ---- host -----------------------------------------------------------------------------------------
// One host thread's scratch for the chain walk: the edge map (the first and second segment holding
// each grid edge, -1: none; reset after each drum through the touched list) and the walk's lists.
struct SfWalker {
  std::vector<int32_t> first, second;
  std::vector<uint32_t> touched, fe, fl, be, bl, edges, links;
  std::vector<uint8_t> visited;
};

// Worker threads kept for the handle's life: run(n, job) calls job(0 .. n-1), job(0) on the caller.
class SfPool {
 public:
  explicit SfPool(uint32_t want) {
    for (uint32_t t = 1; t < want; t++) {
      try {
        th_.emplace_back([this, t] { loop(t); });
      } catch (...) {  // fewer threads: the ranges are spread over those there are
        break;
      }
    }
  }
  ~SfPool() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    go_.notify_all();
    for (std::thread& t : th_) t.join();
  }
  uint32_t size() const { return (uint32_t)th_.size() + 1; }
  void run(const std::function<void(uint32_t)>& job) {
    {
      std::lock_guard<std::mutex> lk(mu_);
      job_ = &job;
      pending_ = (uint32_t)th_.size();
      gen_++;
    }
    go_.notify_all();
    job(0);
    std::unique_lock<std::mutex> lk(mu_);
    done_.wait(lk, [this] { return pending_ == 0; });
    job_ = nullptr;
  }

 private:
  void loop(uint32_t t) {
    uint64_t seen = 0;
    for (;;) {
      const std::function<void(uint32_t)>* job;
      {
        std::unique_lock<std::mutex> lk(mu_);
        go_.wait(lk, [&] { return stop_ || gen_ != seen; });
        if (stop_) return;
        seen = gen_;
        job = job_;
      }
      (*job)(t);
      std::lock_guard<std::mutex> lk(mu_);
      if (--pending_ == 0) done_.notify_one();
    }
  }
  std::vector<std::thread> th_;
  std::mutex mu_;
  std::condition_variable go_, done_;
  const std::function<void(uint32_t)>* job_ = nullptr;
  uint32_t pending_ = 0;
  uint64_t gen_ = 0;
  bool stop_ = false;
};

struct mg_surface {
  uint32_t nd = 0, max_fluid = 0, lift_cap = 0, batch_cap = 0;
  size_t hseg_cap = 0, hab_cap = 0;
  char* dev = nullptr;
  SfDrum* drums = nullptr;
  float2* lift = nullptr;
  uint32_t* boxes = nullptr;
  uint64_t* batch_regions = nullptr;
  float* field = nullptr;
  SfSeg *segs = nullptr, *packed = nullptr;
  float4 *abs = nullptr, *packed_ab = nullptr;
  SfOut* out = nullptr;
  SfSeg* h_seg = nullptr;    // pinned host copy of the packed segments
  float4* h_ab = nullptr;    // and of their crossing points, copied when a chain needs them
  std::vector<SfOut> h_out;
  std::vector<std::vector<float>> flat;  // flatten_polygons per drum
  std::vector<SfWalker> walkers;         // one per pool thread
  SfPool* pool = nullptr;
  bool have = false;
};

namespace {

int sf_fail(char* err, size_t errlen, int rc, const std::string& m) {
  if (err && errlen) {
    std::strncpy(err, m.c_str(), errlen - 1);
    err[errlen - 1] = 0;
  }
  return rc;
}

inline float2 h_lerp(float2 a, float2 b, float t) { return make_float2(a.x + (b.x - a.x) * t, a.y + (b.y - a.y) * t); }

float2 h_project(float2 p, const mgg::Geo& g, float rc, float rsn, float wc, float ws) {  // surface.rs:359-366
  float nx, ny;
  const float d = mgg::sdf_world(p.x, p.y, rc, rsn, wc, ws, g, nx, ny);
  return d < 0.0f ? make_float2(p.x - d * nx, p.y - d * ny) : p;
}

// polygons_at_threshold's chaining, smoothing and projection (surface.rs:309-418) of one drum's
// segment list seg[0..m) into out (flatten_polygons layout). A chain's consecutive edges are the
// two edges of the segment the walk crossed, so each smoothed pair of points is that segment's
// q[] in the walk's direction. An open chain's closing pair and a chain of fewer than three points
// need the crossing points themselves (ab); without them (ab null) the walk returns false.
bool sf_polygons(SfWalker& w, const SfSeg* seg, const float4* ab, uint32_t m, const mgg::Geo& g, float rc, float rsn,
                 float wc, float ws, std::vector<float>& out) {
  out.assign(1, 0.0f);
  if (m == 0) return true;
  out.reserve(1 + 6 * (size_t)m);
  for (uint32_t k = 0; k < m; k++) {
    for (uint32_t e : {seg[k].ea, seg[k].eb}) {
      if (w.first[e] < 0) {
        w.first[e] = (int32_t)k;
        w.touched.push_back(e);
      } else {
        w.second[e] = (int32_t)k;
      }
    }
  }
  auto point = [&](uint32_t e) {  // every segment holding edge e carries the same crossing point
    const int32_t k = w.first[e];
    const float4 v = ab[k];
    return seg[k].ea == e ? make_float2(v.x, v.y) : make_float2(v.z, v.w);
  };
  auto next_of = [&](uint32_t tip) -> int32_t {
    const int32_t a = w.first[tip], b = w.second[tip];
    if (a >= 0 && !w.visited[a]) return a;
    if (b >= 0 && !w.visited[b]) return b;
    return -1;
  };
  w.visited.assign(m, 0);
  // edges and links (segment index * 2 + 1 when crossed from eb to ea) of the forward and the
  // backward part of a chain; the backward part nearest first
  bool ok = true;
  uint32_t npoly = 0;
  for (uint32_t st = 0; st < m && ok; st++) {
    if (w.visited[st]) continue;
    w.visited[st] = 1;
    w.fe.assign({seg[st].ea, seg[st].eb});
    w.fl.assign(1, 2u * st);
    w.be.clear();
    w.bl.clear();
    for (;;) {  // extend_chain forward (surface.rs:371-401)
      const uint32_t tip = w.fe.back();
      const int32_t k = next_of(tip);
      if (k < 0) break;
      w.visited[k] = 1;
      const bool from_a = seg[k].ea == tip;
      w.fe.push_back(from_a ? seg[k].eb : seg[k].ea);
      w.fl.push_back(2u * (uint32_t)k + (from_a ? 0u : 1u));
    }
    for (;;) {  // and backward: the new front edge comes before the tip in the chain
      const uint32_t tip = w.be.empty() ? w.fe.front() : w.be.back();
      const int32_t k = next_of(tip);
      if (k < 0) break;
      w.visited[k] = 1;
      const bool from_a = seg[k].ea == tip;  // then the chain runs eb -> ea
      w.be.push_back(from_a ? seg[k].eb : seg[k].ea);
      w.bl.push_back(2u * (uint32_t)k + (from_a ? 1u : 0u));
    }
    w.edges.assign(w.be.rbegin(), w.be.rend());
    w.edges.insert(w.edges.end(), w.fe.begin(), w.fe.end());
    w.links.assign(w.bl.rbegin(), w.bl.rend());
    w.links.insert(w.links.end(), w.fl.begin(), w.fl.end());
    if (w.edges.size() > 1 && w.edges.front() == w.edges.back()) w.edges.pop_back();  // closed: one link per edge
    const size_t n = w.edges.size();
    if ((n < 3 || w.links.size() < n) && !ab) {
      ok = false;
      break;
    }
    if (n < 3) {  // chaikin_smooth_closed leaves fewer than three points as they are
      out.push_back((float)n);
      for (uint32_t e : w.edges) {
        const float2 p = h_project(point(e), g, rc, rsn, wc, ws);
        out.push_back(p.x);
        out.push_back(p.y);
      }
    } else {
      out.push_back((float)(2 * n));
      for (size_t i = 0; i < n; i++) {
        float2 q0, q1;
        if (i < w.links.size()) {
          const SfSeg& s = seg[w.links[i] >> 1];
          const uint32_t o = (w.links[i] & 1u) * 2u;
          q0 = s.q[o];
          q1 = s.q[o + 1];
        } else {  // an open chain's closing pair, last point to first
          const float2 a = point(w.edges[n - 1]), b = point(w.edges[0]);
          q0 = h_project(h_lerp(a, b, 0.25f), g, rc, rsn, wc, ws);
          q1 = h_project(h_lerp(a, b, 0.75f), g, rc, rsn, wc, ws);
        }
        out.push_back(q0.x);
        out.push_back(q0.y);
        out.push_back(q1.x);
        out.push_back(q1.y);
      }
    }
    npoly++;
  }
  out[0] = (float)npoly;
  for (uint32_t e : w.touched) w.first[e] = w.second[e] = -1;
  w.touched.clear();
  return ok;
}

// Grows a pinned host buffer to hold n elements of T.
template <class T>
bool sf_pinned(T*& buf, size_t& cap, size_t n) {
  if (n <= cap) return true;
  if (buf) (void)hipHostFree(buf);
  buf = nullptr;
  cap = 0;
  const size_t c = std::max<size_t>(n + n / 2, 1024);
  if (hipHostMalloc(reinterpret_cast<void**>(&buf), c * sizeof(T), hipHostMallocDefault) != hipSuccess) return false;
  cap = c;
  return true;
}

}  // namespace

extern "C" {

int mg_surface_create(uint32_t n_drums, uint32_t max_fluid, mg_surface** out, char* err, size_t errlen) {
  if (!out || n_drums == 0 || n_drums > 65535)
    return sf_fail(err, errlen, MG_E_ARG, "n_drums must be in 1..65535 (one grid row of workgroups per drum)");
  auto* h = new mg_surface();
  h->nd = n_drums;
  h->max_fluid = std::max<uint32_t>(max_fluid, 1);
  h->lift_cap = 64 * n_drums;
  h->batch_cap = h->max_fluid / 64 + n_drums + 1;  // at least the sum over drums of ceil(n / 64)
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off += up(bytes);
    return o;
  };
  const size_t nd = n_drums, cap = nd * SF_SEG_CAP;
  const size_t o_drums = take(nd * sizeof(SfDrum)), o_lift = take(h->lift_cap * sizeof(float2));
  const size_t o_boxes = take((size_t)h->max_fluid * 4), o_batches = take((size_t)h->batch_cap * 8 * SF_MASK_WORDS);
  const size_t o_field = take(nd * SF_G * SF_G * 4);
  const size_t o_segs = take(cap * sizeof(SfSeg)), o_packed = take(cap * sizeof(SfSeg));
  const size_t o_abs = take(cap * sizeof(float4)), o_packed_ab = take(cap * sizeof(float4));
  const size_t o_out = take(nd * sizeof(SfOut));
  if (hipMalloc(&h->dev, off) != hipSuccess) {
    delete h;
    return sf_fail(err, errlen, MG_E_DEVICE, "hipMalloc failed for the surface workspace");
  }
  h->drums = reinterpret_cast<SfDrum*>(h->dev + o_drums);
  h->lift = reinterpret_cast<float2*>(h->dev + o_lift);
  h->boxes = reinterpret_cast<uint32_t*>(h->dev + o_boxes);
  h->batch_regions = reinterpret_cast<uint64_t*>(h->dev + o_batches);
  h->field = reinterpret_cast<float*>(h->dev + o_field);
  h->segs = reinterpret_cast<SfSeg*>(h->dev + o_segs);
  h->packed = reinterpret_cast<SfSeg*>(h->dev + o_packed);
  h->abs = reinterpret_cast<float4*>(h->dev + o_abs);
  h->packed_ab = reinterpret_cast<float4*>(h->dev + o_packed_ab);
  h->out = reinterpret_cast<SfOut*>(h->dev + o_out);
  h->h_out.resize(n_drums);
  h->flat.assign(n_drums, std::vector<float>(1, 0.0f));
  try {
    const uint32_t hw = std::max(1u, std::thread::hardware_concurrency());
    h->pool = new SfPool(std::min({hw, 16u, (n_drums + 7) / 8}));
    h->walkers.resize(h->pool->size());
    for (SfWalker& w : h->walkers) {
      w.first.assign(SF_EDGE_IDS, -1);
      w.second.assign(SF_EDGE_IDS, -1);
    }
  } catch (...) {
    delete h->pool;
    (void)hipFree(h->dev);
    delete h;
    return sf_fail(err, errlen, MG_E_DEVICE, "host memory or threads for the chain walk");
  }
  *out = h;
  return MG_OK;
}

void mg_surface_destroy(mg_surface* h) {
  if (!h) return;
  delete h->pool;
  if (h->dev) (void)hipFree(h->dev);
  if (h->h_seg) (void)hipHostFree(h->h_seg);
  if (h->h_ab) (void)hipHostFree(h->h_ab);
  delete h;
}

int mg_surface_extract(mg_surface* h, const mg_surface_drum* drums, const float* d_fluid_x, char* err,
                       size_t errlen) {
  if (!h || !drums) return sf_fail(err, errlen, MG_E_ARG, "null argument");
  const uint32_t nd = h->nd;
  std::vector<SfDrum> hd(nd);
  std::vector<float2> lift;
  uint64_t total = 0, batches = 0;
  uint32_t max_n = 0;
  for (uint32_t d = 0; d < nd; d++) {
    const mg_surface_drum& m = drums[d];
    SfDrum& D = hd[d];
    D.radius = m.radius_m;
    D.cell = 2.0f * m.radius_m / ((float)SF_G - 1.0f);
    D.rs = 1.5f * m.h;  // splat_radius_of (surface.rs:48-52)
    D.r2 = D.rs * D.rs;
    D.inv_r2 = 1.0f / D.r2;
    // analytic_phi_full (surface.rs:71-74): rest / max(mass, 1e-12) * (pi rs^2 / 4)
    const float number_density = m.rest_density / std::fmax(m.particle_mass, 1e-12f);
    const float phi = number_density * (3.14159265358979323846f * D.rs * D.rs / 4.0f);
    D.thr = 0.5f * phi;
    D.rc = std::cos(-m.drum_angle);
    D.rsn = std::sin(-m.drum_angle);
    D.wc = std::cos(m.drum_angle);
    D.ws = std::sin(m.drum_angle);
    D.nl = m.lifter_count;
    D.lift0 = (uint32_t)std::min<size_t>(lift.size(), UINT32_MAX);
    D.rt = m.radius_m - m.lifter_height_m;
    D.bh = m.lifter_base_width_m * 0.5f;
    D.thw = m.lifter_top_width_m * 0.5f;
    if (lift.size() + m.lifter_count > h->lift_cap)
      return sf_fail(err, errlen, MG_E_ARG, "more than 64 lifters per drum on average");
    const float phase = m.lifter_phase_deg * (float)0.017453292519943295;  // f32::to_radians
    const float tau = 6.28318530717958647692f;
    for (uint32_t k = 0; k < m.lifter_count; k++) {  // geometry.rs:50-56
      const float a = phase + (float)k * tau / (float)m.lifter_count;
      lift.push_back(make_float2(std::sin(a), std::cos(a)));
    }
    D.f0 = (uint32_t)std::min<uint64_t>(total, UINT32_MAX);
    D.n = (m.n_fluid == 0 || m.radius_m <= 0.0f) ? 0u : m.n_fluid;  // surface.rs:235-237
    D.b0 = (uint32_t)std::min<uint64_t>(batches, UINT32_MAX);
    batches += (D.n + 63u) / 64u;
    total += m.n_fluid;
    max_n = std::max(max_n, D.n);
  }
  if (total > h->max_fluid) return sf_fail(err, errlen, MG_E_ARG, "more fluid particles than the handle's max_fluid");
  if (batches > h->batch_cap) return sf_fail(err, errlen, MG_E_ARG, "batch table overflow");
  if (total && !d_fluid_x) return sf_fail(err, errlen, MG_E_ARG, "null fluid positions");
  if (lift.empty()) lift.push_back(make_float2(0.0f, 1.0f));
  bool ok = hipMemcpy(h->drums, hd.data(), nd * sizeof(SfDrum), hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(h->lift, lift.data(), lift.size() * sizeof(float2), hipMemcpyHostToDevice) == hipSuccess;
  if (!ok) return sf_fail(err, errlen, MG_E_DEVICE, "HIP copy error");
  const auto* x = reinterpret_cast<const float2*>(d_fluid_x);
  if (max_n)
    hipLaunchKernelGGL(sf_boxes, dim3((max_n + 255) / 256, nd), dim3(256), 0, nullptr, h->drums, x, h->boxes,
                       h->batch_regions);
  hipLaunchKernelGGL(sf_field, dim3(nd, SF_REGIONS), dim3(64), 0, nullptr, h->drums, h->lift, x, h->boxes,
                     h->batch_regions, h->field);
  hipLaunchKernelGGL(sf_cells, dim3(nd), dim3(SF_CELL_BLOCK), 0, nullptr, h->drums, h->lift, h->field, h->segs,
                     h->abs, h->out);
  hipLaunchKernelGGL(sf_pack, dim3(nd), dim3(256), 0, nullptr, h->segs, h->abs, h->out, h->packed, h->packed_ab);
  if (hipGetLastError() != hipSuccess) return sf_fail(err, errlen, MG_E_DEVICE, "kernel launch failed");
  if (hipMemcpy(h->h_out.data(), h->out, nd * sizeof(SfOut), hipMemcpyDeviceToHost) != hipSuccess)
    return sf_fail(err, errlen, MG_E_DEVICE, "HIP copy error");
  std::vector<size_t> s0(nd + 1, 0);
  for (uint32_t d = 0; d < nd; d++) s0[d + 1] = s0[d] + h->h_out[d].nseg;
  const size_t nseg = s0[nd];
  if (!sf_pinned(h->h_seg, h->hseg_cap, nseg))
    return sf_fail(err, errlen, MG_E_DEVICE, "hipHostMalloc failed for the segment read-back");
  if (nseg && hipMemcpy(h->h_seg, h->packed, nseg * sizeof(SfSeg), hipMemcpyDeviceToHost) != hipSuccess)
    return sf_fail(err, errlen, MG_E_DEVICE, "HIP copy error");
  // the drums' chain walks are independent: contiguous ranges of drums per pool thread; a drum with
  // an open or short chain is walked again below with the crossing points (the wall mask zeroes
  // every point of the grid's border, which lies outside the drum, so contours close)
  std::vector<uint8_t> again(nd, 0);
  auto walk = [&](uint32_t t, uint32_t nt, const float4* ab) {
    for (uint32_t d = (uint32_t)((uint64_t)nd * t / nt); d < (uint32_t)((uint64_t)nd * (t + 1) / nt); d++) {
      if (ab && !again[d]) continue;
      const SfDrum& D = hd[d];
      const mgg::Geo g{D.radius, 0.0f, D.nl, D.rt, D.bh, D.thw, lift.data() + D.lift0};
      const bool done = sf_polygons(h->walkers[t], h->h_seg + s0[d], ab ? ab + s0[d] : nullptr, h->h_out[d].nseg, g,
                                    D.rc, D.rsn, D.wc, D.ws, h->flat[d]);
      again[d] = !done;
    }
  };
  const uint32_t nt = h->pool->size();
  try {
    const std::function<void(uint32_t)> job = [&](uint32_t t) { walk(t, nt, nullptr); };
    h->pool->run(job);
  } catch (...) {
    return sf_fail(err, errlen, MG_E_DEVICE, "host thread error in the chain walk");
  }
  bool need_ab = false;
  for (uint8_t a : again) need_ab |= a != 0;
  if (need_ab) {
    if (!sf_pinned(h->h_ab, h->hab_cap, nseg))
      return sf_fail(err, errlen, MG_E_DEVICE, "hipHostMalloc failed for the crossing-point read-back");
    if (hipMemcpy(h->h_ab, h->packed_ab, nseg * sizeof(float4), hipMemcpyDeviceToHost) != hipSuccess)
      return sf_fail(err, errlen, MG_E_DEVICE, "HIP copy error");
    walk(0, 1, h->h_ab);
  }
  h->have = true;
  return MG_OK;
}

int64_t mg_surface_polygons(mg_surface* h, uint32_t drum, float* out, int64_t cap) {
  if (!h || drum >= h->nd) return MG_E_ARG;
  const std::vector<float>& f = h->flat[drum];
  if (out && cap >= (int64_t)f.size()) std::memcpy(out, f.data(), f.size() * sizeof(float));
  return (int64_t)f.size();
}

int mg_surface_field(mg_surface* h, uint32_t drum, float* out16384) {
  if (!h || drum >= h->nd || !out16384) return MG_E_ARG;
  if (!h->have) {
    std::memset(out16384, 0, SF_G * SF_G * sizeof(float));
    return MG_OK;
  }
  return hipMemcpy(out16384, h->field + (size_t)drum * SF_G * SF_G, SF_G * SF_G * sizeof(float),
                   hipMemcpyDeviceToHost) == hipSuccess
             ? MG_OK
             : MG_E_DEVICE;
}

}  // extern "C"
