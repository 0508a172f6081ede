// On-demand metrics of drum states on the GPU: metrics::mixing_index and metrics::compute.
//
// metrics::mixing_index (crates/mill-core/src/metrics.rs:506-576): the Lacey index of the dye
// field, particles binned into a 16 x 16 grid over [-R, R]^2 (cell 2R / 16, floor((p + R) / cell)
// clamped into the grid), per-cell dye sums and counts, the particle-count-weighted variance of
// the cell means against the mean dye p, S0^2 = p (1 - p), Sr^2 = S0^2 / max(sum n^2 / sum n, 1),
// M = clamp((S0^2 - S^2) / (S0^2 - Sr^2), 0, 1); undefined for no particles, fewer than two occupied
// cells or S0^2 < 1e-9, and 1 when |S0^2 - Sr^2| < 1e-9.
//
// The reference forms every sum in f32, in particle order (each cell's dye sum, the total dye) and
// in cell order (the variance terms). One workgroup per drum keeps those orders: particles are
// taken in chunks of 256 in index order; inside a chunk the items of one cell are ranked by wave
// ballots (peer masks over the cell index's 8 bits) and placed contiguously in LDS in their index
// order, so the thread owning a cell adds its items in index order; thread 0 adds the chunk's dye
// to the total in index order; the cell-order terms are formed by one thread at the end. Results
// are bit for bit the reference's f32 arithmetic. Drums of a batch are independent workgroups.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/millgpu.h"
#include "mg_atan2f.h"

#define MX_BLOCK 256
#define MX_G 16
#define MX_CELLS (MX_G * MX_G)
#define MX_WAVES (MX_BLOCK / 64)
static_assert(MX_CELLS == MX_BLOCK, "one thread per grid cell");

namespace {

__device__ __forceinline__ int32_t mx_sat_i32(float v) {  // `as i32`: saturating, NaN -> 0
  if (isnan(v)) return 0;
  if (v <= -2147483648.0f) return INT32_MIN;
  if (v >= 2147483647.0f) return INT32_MAX;
  return (int32_t)v;
}
__device__ __forceinline__ int32_t mx_clamp(int32_t v, int32_t lo, int32_t hi) { return v < lo ? lo : (v > hi ? hi : v); }

}  // namespace

// Drum d covers particles [off[d], off[d + 1]) (off null: one drum of n_single particles) with
// drum radius rad[d] (rad_single). out[d] = the index, ok[d] = 0 when it is undefined (None).
extern "C" __global__ void __launch_bounds__(MX_BLOCK)
mx_mixing_index(const float2* __restrict__ xy, const float* __restrict__ dye, const uint32_t* __restrict__ off,
                const float* __restrict__ rad, uint32_t n_single, float rad_single, float* __restrict__ out,
                int32_t* __restrict__ ok) {
  __shared__ float s_sum[MX_CELLS];
  __shared__ uint32_t s_cnt[MX_CELLS];
  __shared__ uint32_t s_wcnt[MX_WAVES][MX_CELLS];
  __shared__ uint32_t s_cstart[MX_CELLS];
  __shared__ uint32_t s_wave[MX_WAVES];
  __shared__ float s_dye[MX_BLOCK];     // the chunk's dye in index order
  __shared__ float s_sorted[MX_BLOCK];  // the chunk's dye grouped by cell, index order inside a cell
  const uint32_t d = blockIdx.x, t = threadIdx.x, lane = t & 63u, w = t >> 6;
  const uint32_t s = off ? off[d] : 0u;
  const uint32_t n = off ? off[d + 1] - s : n_single;
  const float r = off ? rad[d] : rad_single;
  if (n == 0 || !(r > 0.0f)) {  // uniform over the workgroup
    if (t == 0) { out[d] = 0.0f; ok[d] = 0; }
    return;
  }
  const float cell = 2.0f * r / (float)MX_G;
  s_sum[t] = 0.0f;
  s_cnt[t] = 0u;
  float total = 0.0f;  // thread 0's running sum of all dye values in index order
  for (uint32_t base = 0; base < n; base += MX_BLOCK) {
    for (uint32_t a = 0; a < MX_WAVES; a++) s_wcnt[a][t] = 0u;
    const bool live = base + t < n;
    uint32_t k = 0;
    float dv = 0.0f;
    if (live) {
      const float2 p = xy[s + base + t];
      dv = dye[s + base + t];
      const int32_t gi = mx_clamp(mx_sat_i32(floorf((p.x + r) / cell)), 0, MX_G - 1);
      const int32_t gj = mx_clamp(mx_sat_i32(floorf((p.y + r) / cell)), 0, MX_G - 1);
      k = (uint32_t)(gj * MX_G + gi);
    }
    s_dye[t] = dv;
    uint64_t peers = __ballot(live);
    for (int b = 0; b < 8; b++) {
      const uint64_t m = __ballot(live && ((k >> b) & 1u));
      peers &= ((k >> b) & 1u) ? m : ~m;
    }
    const uint32_t rank = (uint32_t)__popcll(peers & ((1ull << lane) - 1ull));
    __syncthreads();  // s_wcnt cleared by every thread
    if (live && rank == 0) s_wcnt[w][k] = (uint32_t)__popcll(peers);
    __syncthreads();
    // chunk count of cell t, then the cells' exclusive prefix over the chunk
    uint32_t c = 0;
    for (uint32_t a = 0; a < MX_WAVES; a++) c += s_wcnt[a][t];
    {
      uint32_t inc = c;
      for (int o = 1; o < 64; o <<= 1) {
        const uint32_t v = __shfl_up(inc, o, 64);
        if (lane >= (uint32_t)o) inc += v;
      }
      if (lane == 63) s_wave[w] = inc;
      __syncthreads();
      uint32_t before = 0;
      for (uint32_t a = 0; a < w; a++) before += s_wave[a];
      s_cstart[t] = before + inc - c;
    }
    __syncthreads();
    if (live) {
      uint32_t pos = s_cstart[k] + rank;
      for (uint32_t a = 0; a < w; a++) pos += s_wcnt[a][k];
      s_sorted[pos] = dv;
    }
    __syncthreads();
    {  // cell t adds its items of this chunk in index order
      float acc = s_sum[t];
      const uint32_t cs = s_cstart[t];
      for (uint32_t q = 0; q < c; q++) acc += s_sorted[cs + q];
      s_sum[t] = acc;
      s_cnt[t] += c;
    }
    if (t == 0) {
      const uint32_t m = n - base < MX_BLOCK ? n - base : MX_BLOCK;
      for (uint32_t q = 0; q < m; q++) total += s_dye[q];
    }
    __syncthreads();
  }
  if (t != 0) return;
  // cell-order terms, as the reference forms them (metrics.rs:532-575)
  const float pm = total / (float)n;
  uint32_t n_cells = 0, wsum_u = 0;
  for (uint32_t q = 0; q < MX_CELLS; q++)
    if (s_cnt[q] > 0) {
      n_cells++;
      wsum_u += s_cnt[q];
    }
  if (n_cells < 2) { out[d] = 0.0f; ok[d] = 0; return; }
  const float ws = (float)wsum_u;
  float ov = 0.0f, wa = 0.0f;
  for (uint32_t q = 0; q < MX_CELLS; q++)
    if (s_cnt[q] > 0) {
      const float mean = s_sum[q] / (float)s_cnt[q];
      const float dm = mean - pm;
      ov += (float)s_cnt[q] * (dm * dm);
    }
  ov = ov / ws;
  const float s0 = pm * (1.0f - pm);
  for (uint32_t q = 0; q < MX_CELLS; q++)
    if (s_cnt[q] > 0) wa += (float)s_cnt[q] * (float)s_cnt[q];
  wa = wa / ws;
  const float sr = s0 / fmaxf(wa, 1.0f);
  if (s0 < 1e-9f) { out[d] = 0.0f; ok[d] = 0; return; }
  const float den = s0 - sr;
  if (fabsf(den) < 1e-9f) { out[d] = 1.0f; ok[d] = 1; return; }
  float m = (s0 - ov) / den;
  m = m < 0.0f ? 0.0f : (m > 1.0f ? 1.0f : m);
  out[d] = m;
  ok[d] = 1;
}

extern "C" {

int mg_mixing_index_batch(const float* d_xy, const float* d_dye, const uint32_t* d_offsets, const float* d_radius,
                          uint32_t n_drums, float* d_out, int32_t* d_ok, void* stream) {
  if (n_drums == 0) return MG_OK;
  if (!d_xy || !d_dye || !d_offsets || !d_radius || !d_out || !d_ok) return MG_E_ARG;
  hipLaunchKernelGGL(mx_mixing_index, dim3(n_drums), dim3(MX_BLOCK), 0, (hipStream_t)stream,
                     reinterpret_cast<const float2*>(d_xy), d_dye, d_offsets, d_radius, 0u, 0.0f, d_out, d_ok);
  return hipGetLastError() == hipSuccess ? MG_OK : MG_E_DEVICE;
}

int mg_mixing_index(const float* d_xy, const float* d_dye, uint32_t n, float radius, float* out, int* defined) {
  if (!out || !defined || (n && (!d_xy || !d_dye))) return MG_E_ARG;
  float* d_res = nullptr;
  if (hipMalloc(&d_res, 2 * sizeof(float)) != hipSuccess) return MG_E_DEVICE;
  hipLaunchKernelGGL(mx_mixing_index, dim3(1), dim3(MX_BLOCK), 0, nullptr, reinterpret_cast<const float2*>(d_xy), d_dye,
                     nullptr, nullptr, n, radius, d_res, reinterpret_cast<int32_t*>(d_res + 1));
  float h[2] = {0.0f, 0.0f};
  const bool good = hipGetLastError() == hipSuccess &&
                    hipMemcpy(h, d_res, sizeof h, hipMemcpyDeviceToHost) == hipSuccess;
  (void)hipFree(d_res);
  if (!good) return MG_E_DEVICE;
  int32_t k;
  static_assert(sizeof(k) == sizeof(float), "flag shares the result buffer");
  std::memcpy(&k, &h[1], sizeof k);
  *out = h[0];
  *defined = k;
  return MG_OK;
}

}  // extern "C"

// ---- metrics::compute (metrics.rs:195-254) ---------------------------------------------------
// Per drum, one 256-lane workgroup per kernel:
//   dm_balls        ball pass: wall SDF, ball-wall overlap (max), the wall-adjacent balls' angles
//                   (toe/shoulder list), and the f32 sums of positions and kinetic energies;
//   dm_ball_overlap all ball pairs closer than 2.12 r (the reference's candidate pairs at cell
//                   2.1 r contain every pair closer than 2 r; others add exactly 0);
//   dm_fluid        fluid pass: wall SDF, the wall-adjacent particles' angles (pool list), the
//                   free-surface points' position sums and then their covariance sums, the pool
//                   depth band's nearest distance, density and compression errors;
//   dm_sort_gap     sorts one angle list (bitonic, by (angle, particle index) so equal angles keep
//                   the reference's stable order) and picks the largest gap by the reference's rule;
//   mx_mixing_index the Lacey index as above.
// Every f32 sum the reference forms in particle order is formed in particle order here: a chunk
// of 256 values is staged in LDS and lane q of wave 0 adds series q of it in index order. The
// host finishes each drum with the reference's scalar formulas (divisions by counts, the
// free-surface angle by atan2f / sinf / cosf of the host libm, the toe/shoulder and None rules).

namespace {

#define DM_BLOCK 256
constexpr float DM_TAU = 6.28318530717958647692f;
constexpr float DM_PI = 3.14159265358979323846f;

struct DmDrum {
  float rc, rs;          // from_angle(-drum_angle): world -> drum frame (geometry.rs:93-94)
  float radius, omega;
  uint32_t nl, lift0;    // lifter count, first entry of the lifter table
  float rt, bh, thw;     // lifter top radius, half base width, half top width (geometry.rs:58-65)
  float br, bmass, binertia, h, rest;
  uint32_t b0, nb, f0, nf;
};

struct DmOut {
  float wall_ov, ke, sbx, sby;           // balls
  float sfx, sfy, cxx, cyy, cxy;         // free-surface sums
  uint32_t fs_n;
  float depth_min;
  uint32_t depth_any;
  float max_e, max_c, sum_e, sum_c;
  float max_overlap;
  uint32_t toe_n, pool_n;
  float toe_start, toe_end, toe_gap;
  float pool_start, pool_end, pool_gap;
  uint32_t nan_toe, nan_pool;
};

struct DmItem {  // one wall-adjacent particle: its angle and index (the sort's tie-break)
  float a;
  uint32_t i;
};

__device__ __forceinline__ bool dm_finite(float2 p) { return isfinite(p.x) && isfinite(p.y); }

// convex_polygon_sdf (geometry.rs:233-262) over the lifter trapezoid (radius, -bh), (radius, bh),
// (rt, thw), (rt, -thw).
__device__ float dm_polygon_sdf(float px, float py, const DmDrum& D) {
  const float vx[4] = {D.radius, D.radius, D.rt, D.rt};
  const float vy[4] = {-D.bh, D.bh, D.thw, -D.thw};
  float min_d2 = 3.40282347e+38f;
  bool inside = true;
#pragma unroll
  for (int e = 0; e < 4; e++) {
    const float ax = vx[e], ay = vy[e], bx = vx[(e + 1) & 3], by = vy[(e + 1) & 3];
    const float ex = bx - ax, ey = by - ay, tx = px - ax, ty = py - ay;
    const float ls = ex * ex + ey * ey;
    if (ls < 1e-12f) continue;
    float t = (tx * ex + ty * ey) / ls;
    if (t < 0.0f) t = 0.0f;
    if (t > 1.0f) t = 1.0f;
    const float qx = px - (ax + ex * t), qy = py - (ay + ey * t);
    min_d2 = fminf(min_d2, qx * qx + qy * qy);
    if (ex * ty - ey * tx < 0.0f) inside = false;
  }
  const float d = sqrtf(min_d2);
  return inside ? d : -d;
}

// Drum::sdf_world's distance (geometry.rs:37-72, :87-108; the normal is not needed here).
__device__ float dm_sdf_world(float2 p, const DmDrum& D, const float2* __restrict__ lift) {
  if (!dm_finite(p)) return 0.0f;
  const float lx = p.x * D.rc - p.y * D.rs, ly = p.x * D.rs + p.y * D.rc;
  const float dist = D.radius - sqrtf(lx * lx + ly * ly);
  if (D.nl == 0) return dist;
  float max_inside = -INFINITY;
  for (uint32_t k = 0; k < D.nl; k++) {
    const float2 sc = lift[D.lift0 + k];  // (sin, cos) of the lifter's angle
    const float pr = lx * sc.y + ly * sc.x;
    const float pt = -lx * sc.x + ly * sc.y;
    max_inside = fmaxf(max_inside, dm_polygon_sdf(pr, pt, D));
  }
  return fminf(dist, -max_inside);
}

// p.y.atan2(p.x).rem_euclid(TAU) (metrics.rs:341): the atan2 lies in [-pi, pi], so the
// remainder by TAU is the value itself and a negative one gains TAU.
__device__ __forceinline__ float dm_wall_angle(float2 p) {
  const float a = mgm::ref_atan2f(p.y, p.x);
  return a < 0.0f ? a + DM_TAU : a;
}

__device__ __forceinline__ float dm_wave_max(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ float dm_wave_min(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ uint32_t dm_wave_or(uint32_t v) {
  for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o, 64);
  return v;
}
// acc + row[0] + row[1] + ... + row[m-1], added one at a time in that order (the LDS row is read
// four values per load so the loads run ahead of the dependent adds).
__device__ __forceinline__ float dm_seq_sum(float acc, const float* row, uint32_t m) {
  uint32_t q = 0;
#pragma unroll 4
  for (; q + 4 <= m; q += 4) {
    const float4 v = *reinterpret_cast<const float4*>(row + q);
    acc += v.x;
    acc += v.y;
    acc += v.z;
    acc += v.w;
  }
  for (; q < m; q++) acc += row[q];
  return acc;
}

// Maximum over the workgroup (every thread must call it); the result is valid in thread 0.
__device__ float dm_block_max(float v, float* s_red) {
  v = dm_wave_max(v);
  if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = v;
  __syncthreads();
  float r = s_red[0];
  for (int w = 1; w < DM_BLOCK / 64; w++) r = fmaxf(r, s_red[w]);
  __syncthreads();
  return r;
}

}  // namespace

extern "C" __global__ void __launch_bounds__(DM_BLOCK)
dm_balls(const DmDrum* __restrict__ drums, const float2* __restrict__ lift, const float2* __restrict__ bx,
         const float2* __restrict__ bv, const float* __restrict__ bw, DmItem* __restrict__ toe, DmOut* __restrict__ out) {
  __shared__ __align__(16) float s_v[3][DM_BLOCK];
  __shared__ float s_red[DM_BLOCK / 64];
  __shared__ uint32_t s_n;
  const DmDrum D = drums[blockIdx.x];
  const uint32_t t = threadIdx.x;
  if (t == 0) s_n = 0;
  __syncthreads();
  const float margin = 2.5f * D.br;
  float wov = 0.0f, acc = 0.0f;  // acc: lane q of wave 0 sums series q (x, y, kinetic energy)
  for (uint32_t base = 0; base < D.nb; base += DM_BLOCK) {
    const uint32_t i = base + t;
    if (i < D.nb) {
      const float2 p = bx[D.b0 + i];
      const float d = dm_sdf_world(p, D, lift);
      if (D.br > 0.0f) wov = fmaxf(wov, fmaxf(D.br - d, 0.0f) / D.br);
      if (d < margin) {
        const uint32_t slot = atomicAdd(&s_n, 1u);
        toe[D.b0 + slot] = DmItem{dm_wall_angle(p), i};
      }
      const float2 v = bv[D.b0 + i];
      const float w = bw[D.b0 + i];
      s_v[0][t] = p.x;
      s_v[1][t] = p.y;
      s_v[2][t] = 0.5f * D.bmass * (v.x * v.x + v.y * v.y) + 0.5f * D.binertia * w * w;
    }
    __syncthreads();
    if (t < 3) acc = dm_seq_sum(acc, s_v[t], min(D.nb - base, (uint32_t)DM_BLOCK));
    __syncthreads();
  }
  wov = dm_block_max(wov, s_red);
  DmOut* o = out + blockIdx.x;
  if (t == 0) { o->wall_ov = wov; o->toe_n = s_n; o->sbx = acc; }
  if (t == 1) o->sby = acc;
  if (t == 2) o->ke = acc;
}

// 1024 lanes per drum (16 waves, so LDS reads of one wave overlap the arithmetic of others).
#define DM_OV_BLOCK 1024
extern "C" __global__ void __launch_bounds__(DM_OV_BLOCK)
dm_ball_overlap(const DmDrum* __restrict__ drums, const float2* __restrict__ bx, DmOut* __restrict__ out) {
  __shared__ __align__(16) float2 s_p[DM_OV_BLOCK];
  __shared__ float s_red[DM_OV_BLOCK / 64];
  const DmDrum D = drums[blockIdx.x];
  const uint32_t t = threadIdx.x;
  float best = 0.0f;
  if (D.nb >= 2 && D.br > 0.0f) {
    const float two_r = 2.0f * D.br, lim = 4.5f * D.br * D.br;
    for (uint32_t ib = 0; ib < D.nb; ib += DM_OV_BLOCK) {
      const uint32_t i = ib + t;
      const float2 pi = i < D.nb ? bx[D.b0 + i] : make_float2(0.0f, 0.0f);
      for (uint32_t jb = ib; jb < D.nb; jb += DM_OV_BLOCK) {  // tiles at or after the i tile
        __syncthreads();
        if (jb + t < D.nb) s_p[t] = bx[D.b0 + jb + t];
        __syncthreads();
        if (i < D.nb) {
          const uint32_t m = min(D.nb - jb, (uint32_t)DM_OV_BLOCK);
          for (uint32_t q = 0; q < m; q += 2) {  // two positions per LDS read
            const float4 two = *reinterpret_cast<const float4*>(&s_p[q]);
            {
              const float dx = pi.x - two.x, dy = pi.y - two.y, d2 = dx * dx + dy * dy;
              if (jb + q > i && d2 < lim) best = fmaxf(best, fmaxf(two_r - sqrtf(d2), 0.0f) / D.br);
            }
            {
              const float dx = pi.x - two.z, dy = pi.y - two.w, d2 = dx * dx + dy * dy;
              if (q + 1 < m && jb + q + 1 > i && d2 < lim) best = fmaxf(best, fmaxf(two_r - sqrtf(d2), 0.0f) / D.br);
            }
          }
        }
      }
    }
  }
  best = dm_wave_max(best);
  if ((t & 63) == 0) s_red[t >> 6] = best;
  __syncthreads();
  if (t == 0) {
    float r = s_red[0];
    for (int w = 1; w < DM_OV_BLOCK / 64; w++) r = fmaxf(r, s_red[w]);
    out[blockIdx.x].max_overlap = r;
  }
}

extern "C" __global__ void __launch_bounds__(DM_BLOCK)
dm_fluid(const DmDrum* __restrict__ drums, const float2* __restrict__ lift, const float2* __restrict__ fx,
         const float* __restrict__ dens, DmItem* __restrict__ pool, DmOut* __restrict__ out) {
  __shared__ __align__(16) float s_v[4][DM_BLOCK];
  __shared__ float s_red[DM_BLOCK / 64];
  __shared__ uint32_t s_n, s_fs;
  __shared__ float s_c[2];
  const DmDrum D = drums[blockIdx.x];
  const uint32_t t = threadIdx.x;
  if (t == 0) { s_n = 0; s_fs = 0; }
  __syncthreads();
  const float margin = 1.5f * D.h;
  const float down = 3.0f * (DM_PI * 0.5f), band = 5.0f * DM_PI / 180.0f;
  const bool errs = D.rest > 0.0f;
  float dmin = INFINITY, max_e = 0.0f, max_c = 0.0f, acc = 0.0f;
  uint32_t any = 0;
  // pass 1: lists, depth band, errors; sums of the free-surface positions and of the errors
  for (uint32_t base = 0; base < D.nf; base += DM_BLOCK) {
    const uint32_t i = base + t;
    if (i < D.nf) {
      const float2 p = fx[D.f0 + i];
      const float d = dm_sdf_world(p, D, lift);
      const float a = dm_wall_angle(p);
      if (d < margin) {
        const uint32_t slot = atomicAdd(&s_n, 1u);
        pool[D.f0 + slot] = DmItem{a, i};
      }
      const float sep = (a - down) < 0.0f ? (a - down) + DM_TAU : (a - down);
      if (fminf(sep, DM_TAU - sep) <= band) {
        dmin = fminf(dmin, sqrtf(p.x * p.x + p.y * p.y));
        any = 1;
      }
      const bool inc = d >= margin;
      float e = 0.0f, c = 0.0f;
      if (errs) {
        const float rho = dens[D.f0 + i];
        e = fabsf(rho - D.rest) / D.rest;
        c = fmaxf(rho / D.rest - 1.0f, 0.0f);
        max_e = fmaxf(max_e, e);
        max_c = fmaxf(max_c, c);
      }
      // an excluded point adds +0.0, which leaves the running sum unchanged (it is never -0.0)
      s_v[0][t] = inc ? p.x : 0.0f;
      s_v[1][t] = inc ? p.y : 0.0f;
      s_v[2][t] = e;
      s_v[3][t] = c;
      const uint64_t b = __ballot(inc);
      if ((t & 63) == 0) atomicAdd(&s_fs, (uint32_t)__popcll(b));
    }
    __syncthreads();
    if (t < 4) acc = dm_seq_sum(acc, s_v[t], min(D.nf - base, (uint32_t)DM_BLOCK));
    __syncthreads();
  }
  DmOut* o = out + blockIdx.x;
  const uint32_t fs_n = s_fs;
  if (t == 0) s_c[0] = fs_n >= 3 ? acc / (float)fs_n : 0.0f;
  if (t == 1) s_c[1] = fs_n >= 3 ? acc / (float)fs_n : 0.0f;
  if (t == 0) o->sfx = acc;
  if (t == 1) o->sfy = acc;
  if (t == 2) o->sum_e = acc;
  if (t == 3) o->sum_c = acc;
  __syncthreads();
  // pass 2: covariance sums of the free-surface points about their centroid
  float cov = 0.0f;
  if (fs_n >= 3) {
    const float cx = s_c[0], cy = s_c[1];
    for (uint32_t base = 0; base < D.nf; base += DM_BLOCK) {
      const uint32_t i = base + t;
      if (i < D.nf) {
        const float2 p = fx[D.f0 + i];
        const bool inc = dm_sdf_world(p, D, lift) >= margin;
        const float dx = p.x - cx, dy = p.y - cy;
        s_v[0][t] = inc ? dx * dx : 0.0f;
        s_v[1][t] = inc ? dy * dy : 0.0f;
        s_v[2][t] = inc ? dx * dy : 0.0f;
      }
      __syncthreads();
      if (t < 3) cov = dm_seq_sum(cov, s_v[t], min(D.nf - base, (uint32_t)DM_BLOCK));
      __syncthreads();
    }
  }
  if (t == 0) o->cxx = cov;
  if (t == 1) o->cyy = cov;
  if (t == 2) o->cxy = cov;
  dmin = dm_wave_min(dmin);
  if ((t & 63) == 0) s_red[t >> 6] = dmin;
  __syncthreads();
  if (t == 0) {
    float r = s_red[0];
    for (int w = 1; w < DM_BLOCK / 64; w++) r = fminf(r, s_red[w]);
    o->depth_min = r;
  }
  __syncthreads();
  any = dm_wave_or(any);
  if ((t & 63) == 0) s_red[t >> 6] = any ? 1.0f : 0.0f;
  __syncthreads();
  if (t == 0) {
    float r = 0.0f;
    for (int w = 0; w < DM_BLOCK / 64; w++) r = fmaxf(r, s_red[w]);
    o->depth_any = r > 0.0f ? 1u : 0u;
  }
  __syncthreads();
  max_e = dm_block_max(max_e, s_red);
  max_c = dm_block_max(max_c, s_red);
  if (t == 0) {
    o->max_e = max_e;
    o->max_c = max_c;
    o->fs_n = fs_n;
    o->pool_n = s_n;
  }
}

namespace {
__device__ __forceinline__ bool dm_less(DmItem a, DmItem b) { return a.a < b.a || (a.a == b.a && a.i < b.i); }
}  // namespace

// blockIdx.y = 0: the toe/shoulder list (sorted when it holds at least 6 angles), 1: the pool list
// (at least 2). A NaN angle in a list that the reference sorts is reported instead (it panics).
extern "C" __global__ void __launch_bounds__(DM_BLOCK)
dm_sort_gap(const DmDrum* __restrict__ drums, DmItem* __restrict__ toe, DmItem* __restrict__ pool,
            DmOut* __restrict__ out) {
  __shared__ float s_g[DM_BLOCK / 64];
  __shared__ uint32_t s_k[DM_BLOCK / 64];
  __shared__ uint32_t s_nan;
  const DmDrum D = drums[blockIdx.x];
  DmOut* o = out + blockIdx.x;
  const bool is_toe = blockIdx.y == 0;
  const uint32_t n = is_toe ? o->toe_n : o->pool_n;
  DmItem* a = is_toe ? toe + D.b0 : pool + D.f0;
  const uint32_t t = threadIdx.x;
  if (n == 1 && !is_toe) {  // a single pool angle is returned as it is
    if (t == 0) { o->pool_start = o->pool_end = a[0].a; o->pool_gap = 0.0f; }
    return;
  }
  if (n < (is_toe ? 6u : 2u)) return;
  if (t == 0) s_nan = 0;
  __syncthreads();
  uint32_t nan = 0;
  for (uint32_t i = t; i < n; i += DM_BLOCK) nan |= isnan(a[i].a) ? 1u : 0u;
  if (nan) atomicOr(&s_nan, 1u);
  __syncthreads();
  if (s_nan) {
    if (t == 0) { if (is_toe) o->nan_toe = 1; else o->nan_pool = 1; }
    return;
  }
  uint32_t P = 1;
  while (P < n) P <<= 1;
  for (uint32_t k = 2; k <= P; k <<= 1) {
    const uint32_t hk = k >> 1;
    for (uint32_t q = t; q < P / 2; q += DM_BLOCK) {  // mirrored compare inside blocks of k
      const uint32_t blk = q / hk, off = q % hk;
      const uint32_t lo = blk * k + off, hi = blk * k + (k - 1 - off);
      if (hi < n) {
        const DmItem x = a[lo], y = a[hi];
        if (dm_less(y, x)) { a[lo] = y; a[hi] = x; }
      }
    }
    __syncthreads();
    for (uint32_t j = k >> 2; j >= 1; j >>= 1) {  // half-cleaners
      for (uint32_t q = t; q < P / 2; q += DM_BLOCK) {
        const uint32_t blk = q / j, off = q % j;
        const uint32_t lo = blk * 2 * j + off, hi = lo + j;
        if (hi < n) {
          const DmItem x = a[lo], y = a[hi];
          if (dm_less(y, x)) { a[lo] = y; a[hi] = x; }
        }
      }
      __syncthreads();
    }
  }
  // largest_gap_cluster_bounds (metrics.rs:293-308): the first largest inner gap if it exceeds the
  // wrap-around gap, else the wrap-around one
  float bg = -INFINITY;
  uint32_t bk = 0xffffffffu;
  for (uint32_t i = t; i + 1 < n; i += DM_BLOCK) {
    const float g = a[i + 1].a - a[i].a;
    if (g > bg) { bg = g; bk = i; }
  }
  for (int o2 = 32; o2 > 0; o2 >>= 1) {
    const float g2 = __shfl_xor(bg, o2, 64);
    const uint32_t k2 = __shfl_xor(bk, o2, 64);
    if (g2 > bg || (g2 == bg && k2 < bk)) { bg = g2; bk = k2; }
  }
  if ((t & 63) == 0) { s_g[t >> 6] = bg; s_k[t >> 6] = bk; }
  __syncthreads();
  if (t == 0) {
    for (int w = 1; w < DM_BLOCK / 64; w++)
      if (s_g[w] > bg || (s_g[w] == bg && s_k[w] < bk)) { bg = s_g[w]; bk = s_k[w]; }
    const float wrap = a[0].a + DM_TAU - a[n - 1].a;
    uint32_t k = n - 1;
    float gap = wrap;
    if (bg > wrap) { k = bk; gap = bg; }
    const float end = a[k].a, start = a[(k + 1) % n].a;
    if (is_toe) { o->toe_start = start; o->toe_end = end; o->toe_gap = gap; }
    else { o->pool_start = start; o->pool_end = end; o->pool_gap = gap; }
  }
}

extern "C" {

void mg_atan2f_ref(const float* y, const float* x, float* out, uint64_t n) {
  for (uint64_t i = 0; i < n; i++) out[i] = mgm::ref_atan2f(y[i], x[i]);
}

int mg_drum_metrics_batch(const mg_metrics_drum* drums, uint32_t n_drums, const float* d_ball_x, const float* d_ball_v,
                          const float* d_ball_omega, const float* d_fluid_x, const float* d_fluid_dye,
                          const float* d_fluid_density, mg_drum_metrics* out, char* err, size_t errlen) {
  auto fail = [&](int rc, const std::string& m) {
    if (err && errlen) {
      std::strncpy(err, m.c_str(), errlen - 1);
      err[errlen - 1] = 0;
    }
    return rc;
  };
  if (n_drums == 0) return MG_OK;
  if (!drums || !out) return fail(MG_E_ARG, "null drums or out");
  // host side: per-drum records, the lifter (sin, cos) table and the offsets, exactly as the
  // reference evaluates them (geometry.rs:50-56, :93-94) with the host libm
  std::vector<DmDrum> hd(n_drums);
  std::vector<float2> lift;
  std::vector<uint32_t> foff(n_drums + 1, 0);
  std::vector<float> rad(n_drums);
  uint64_t tb = 0, tf = 0;
  for (uint32_t d = 0; d < n_drums; d++) {
    const mg_metrics_drum& m = drums[d];
    DmDrum& D = hd[d];
    D.rc = std::cos(-m.drum_angle);
    D.rs = std::sin(-m.drum_angle);
    D.radius = m.radius_m;
    D.omega = m.omega;
    D.nl = m.lifter_count;
    D.lift0 = (uint32_t)lift.size();
    D.rt = m.radius_m - m.lifter_height_m;
    D.bh = m.lifter_base_width_m * 0.5f;
    D.thw = m.lifter_top_width_m * 0.5f;
    const float phase = m.lifter_phase_deg * (float)0.017453292519943295;  // f32::to_radians
    for (uint32_t k = 0; k < m.lifter_count; k++) {
      const float th = phase + (float)k * DM_TAU / (float)m.lifter_count;
      lift.push_back(make_float2(std::sin(th), std::cos(th)));
    }
    D.br = m.ball_radius;
    D.bmass = m.ball_mass;
    D.binertia = m.ball_inertia;
    D.h = m.h;
    D.rest = m.rest_density;
    D.b0 = (uint32_t)tb;
    D.nb = m.n_balls;
    D.f0 = (uint32_t)tf;
    D.nf = m.n_fluid;
    tb += m.n_balls;
    tf += m.n_fluid;
    if (tb > 0xffffffffull || tf > 0xffffffffull) return fail(MG_E_ARG, "more than 2^32 balls or fluid particles");
    foff[d + 1] = (uint32_t)tf;
    rad[d] = m.radius_m;
  }
  if ((tb && (!d_ball_x || !d_ball_v || !d_ball_omega)) || (tf && (!d_fluid_x || !d_fluid_dye || !d_fluid_density)))
    return fail(MG_E_ARG, "null device array");
  if (lift.empty()) lift.push_back(make_float2(0.0f, 1.0f));
  // one device block: drums, lifters, offsets, radii, outputs, mixing results, the two lists
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t o_drums = 0, o_lift = o_drums + up(n_drums * sizeof(DmDrum));
  const size_t o_foff = o_lift + up(lift.size() * sizeof(float2));
  const size_t o_rad = o_foff + up((n_drums + 1) * sizeof(uint32_t));
  const size_t o_out = o_rad + up(n_drums * sizeof(float));
  const size_t o_mix = o_out + up(n_drums * sizeof(DmOut));
  const size_t o_ok = o_mix + up(n_drums * sizeof(float));
  const size_t o_toe = o_ok + up(n_drums * sizeof(int32_t));
  const size_t o_pool = o_toe + up(std::max<uint64_t>(tb, 1) * sizeof(DmItem));
  const size_t total = o_pool + up(std::max<uint64_t>(tf, 1) * sizeof(DmItem));
  char* dev = nullptr;
  if (hipMalloc(&dev, total) != hipSuccess) return fail(MG_E_DEVICE, "hipMalloc failed for the metrics workspace");
  bool ok = hipMemcpy(dev + o_drums, hd.data(), n_drums * sizeof(DmDrum), hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(dev + o_lift, lift.data(), lift.size() * sizeof(float2), hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(dev + o_foff, foff.data(), foff.size() * sizeof(uint32_t), hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(dev + o_rad, rad.data(), rad.size() * sizeof(float), hipMemcpyHostToDevice) == hipSuccess &&
            hipMemset(dev + o_out, 0, n_drums * sizeof(DmOut)) == hipSuccess;
  auto* dd = reinterpret_cast<const DmDrum*>(dev + o_drums);
  auto* dl = reinterpret_cast<const float2*>(dev + o_lift);
  auto* dout = reinterpret_cast<DmOut*>(dev + o_out);
  auto* dtoe = reinterpret_cast<DmItem*>(dev + o_toe);
  auto* dpool = reinterpret_cast<DmItem*>(dev + o_pool);
  if (ok) {
    hipLaunchKernelGGL(dm_balls, dim3(n_drums), dim3(DM_BLOCK), 0, nullptr, dd, dl,
                       reinterpret_cast<const float2*>(d_ball_x), reinterpret_cast<const float2*>(d_ball_v), d_ball_omega,
                       dtoe, dout);
    hipLaunchKernelGGL(dm_ball_overlap, dim3(n_drums), dim3(DM_OV_BLOCK), 0, nullptr, dd,
                       reinterpret_cast<const float2*>(d_ball_x), dout);
    hipLaunchKernelGGL(dm_fluid, dim3(n_drums), dim3(DM_BLOCK), 0, nullptr, dd, dl,
                       reinterpret_cast<const float2*>(d_fluid_x), d_fluid_density, dpool, dout);
    hipLaunchKernelGGL(dm_sort_gap, dim3(n_drums, 2), dim3(DM_BLOCK), 0, nullptr, dd, dtoe, dpool, dout);
    hipLaunchKernelGGL(mx_mixing_index, dim3(n_drums), dim3(MX_BLOCK), 0, nullptr,
                       reinterpret_cast<const float2*>(d_fluid_x), d_fluid_dye,
                       reinterpret_cast<const uint32_t*>(dev + o_foff), reinterpret_cast<const float*>(dev + o_rad), 0u,
                       0.0f, reinterpret_cast<float*>(dev + o_mix), reinterpret_cast<int32_t*>(dev + o_ok));
    ok = hipGetLastError() == hipSuccess;
  }
  std::vector<DmOut> ho(n_drums);
  std::vector<float> hmix(n_drums);
  std::vector<int32_t> hok(n_drums);
  ok = ok && hipMemcpy(ho.data(), dout, n_drums * sizeof(DmOut), hipMemcpyDeviceToHost) == hipSuccess &&
       hipMemcpy(hmix.data(), dev + o_mix, n_drums * sizeof(float), hipMemcpyDeviceToHost) == hipSuccess &&
       hipMemcpy(hok.data(), dev + o_ok, n_drums * sizeof(int32_t), hipMemcpyDeviceToHost) == hipSuccess;
  (void)hipFree(dev);
  if (!ok) return fail(MG_E_DEVICE, "HIP error in the drum metrics kernels");
  // the reference's scalar epilogue per drum (metrics.rs:327-710)
  for (uint32_t d = 0; d < n_drums; d++) {
    const mg_metrics_drum& m = drums[d];
    const DmOut& r = ho[d];
    mg_drum_metrics& M = out[d];
    std::memset(&M, 0, sizeof M);
    if (r.nan_toe || r.nan_pool)
      return fail(MG_E_ARG, "drum " + std::to_string(d) + ": a non-finite " + (r.nan_toe ? "ball" : "fluid") +
                                " position among the wall-adjacent " + (r.nan_toe ? "balls" : "particles") +
                                " (the reference's angle sort panics on it)");
    if (m.n_balls > 0 && r.toe_n >= 6 && r.toe_gap > DM_PI * 0.3f) {
      M.has_toe_shoulder = 1;
      M.toe_angle_rad = m.omega >= 0.0f ? r.toe_start : r.toe_end;
      M.shoulder_angle_rad = m.omega >= 0.0f ? r.toe_end : r.toe_start;
    }
    if (m.n_balls > 0) {
      M.has_centroid = 1;
      M.centroid_x = r.sbx / (float)m.n_balls;
      M.centroid_y = r.sby / (float)m.n_balls;
    }
    if (m.n_fluid > 0 && r.pool_n > 0) {
      M.has_pool_extent = 1;
      M.pool_angle_min_rad = r.pool_start;
      M.pool_angle_max_rad = r.pool_end;
    }
    if (m.n_fluid > 0 && r.fs_n >= 3) {
      const float n = (float)r.fs_n;
      const float cx = r.sfx / n, cy = r.sfy / n;
      const float cxx = r.cxx / n, cyy = r.cyy / n, cxy = r.cxy / n;
      const float a = 0.5f * std::atan2(2.0f * cxy, cxx - cyy);
      const float sa = std::sin(a), ca = std::cos(a);
      M.has_free_surface = 1;
      M.free_surface_angle_rad = a;
      M.free_surface_offset_m = cx * -sa + cy * ca;
    }
    if (m.n_fluid > 0 && r.depth_any) {
      M.has_pool_depth = 1;
      M.pool_depth_m = m.radius_m - r.depth_min;
    }
    M.has_mixing_index = hok[d] ? 1 : 0;
    M.mixing_index = hok[d] ? hmix[d] : 0.0f;
    M.total_kinetic_energy_j = r.ke;
    M.max_ball_overlap_fraction = r.max_overlap;
    M.max_ball_wall_overlap_fraction = m.n_balls > 0 ? r.wall_ov : 0.0f;
    if (m.n_fluid > 0 && m.rest_density > 0.0f) {
      M.has_density_errors = 1;
      M.max_fluid_density_error_fraction = r.max_e;
      M.mean_fluid_density_error_fraction = r.sum_e / (float)m.n_fluid;
      M.max_fluid_compression_error_fraction = r.max_c;
      M.mean_fluid_compression_error_fraction = r.sum_c / (float)m.n_fluid;
    }
    M.drum_omega_rad_s = m.omega;
    M.fluid_particle_count = m.n_fluid;
  }
  return MG_OK;
}

}  // extern "C"
