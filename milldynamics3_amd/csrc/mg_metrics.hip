// On-demand metrics of a drum state on the GPU.
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

#include <cstring>

#include "../../include/millgpu.h"

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
