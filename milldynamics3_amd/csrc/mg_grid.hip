// The reference's uniform grid on the GPU, in the reference's own layout and order.
//
// UniformGrid::build (crates/mill-core/src/grid.rs:101-226):
//   * cell key (floor(x / cell) as i32, floor(y / cell) as i32) of every finite point
//     (grid.rs:280-285; the f32 quotient is the IEEE one, the cast saturates);
//   * the bounding box [min_cx, max_cx] x [min_cy, max_cy] of those keys, width W and height H;
//   * the Dense layout when W * H <= 2 m + 512 (m finite points, grid.rs:44, :48, :138-152):
//     x-major cell index k = (cx - min_cx) * H + (cy - min_cy) (grid.rs:232-236), cell_start with
//     W * H + 1 entries; otherwise the Sparse layout (grid.rs:204-226): the occupied cells only,
//     sorted by (cx, cy), with cell_start parallel to them;
//   * indices: the points of a cell in ascending point index.
// Exports in the reference's emission order: for_each_candidate_pair (grid.rs:291-362),
// for_each_near (grid.rs:371-399) and build_neighbor_lists (crates/mill-core/src/pbf.rs:262-288).
//
// GPU design. The key kernel reads the points once and reduces the bounding box and the finite
// count (wave reductions, one atomic per wave). The layout decision needs that box, so the host
// reads the 48-byte control block back (one wait per build) and sizes the rest of the launches.
// The cell order is a stable LSD radix sort of the sort key (dense: k; sparse: the offsets
// (cx - min_cx, cy - min_cy) packed into one 64-bit word) with the point index as payload, in
// 8-bit digits: a per-workgroup digit histogram, a digit-major exclusive scan, and a scatter that
// ranks equal digits inside a workgroup from wave ballots (peer masks) so equal keys keep their
// index order — the stability the reference's counting sort and stable sort give. The number of
// passes follows from the key's bit width (a 1M-point dense grid: 22 bits, 3 passes). Non-finite
// points are dropped by a stable compaction before the sort. Dense cell starts come from the
// sorted keys (each sorted position fills the starts of the cells between its key and the
// previous one); sparse cells from the key boundaries and an exclusive scan.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "../../include/millgpu.h"

#define GR_BLOCK 256
#define GR_ITEMS 8
#define GR_TILE (GR_BLOCK * GR_ITEMS)  // items per workgroup in the sort and the compaction
#define GR_RADIX_BITS 9  // two passes for up to 2^18 cells (the 1M-point lattice at cell h)
#define GR_BINS (1 << GR_RADIX_BITS)
#define GR_BPT (GR_BINS / GR_BLOCK)  // digits per thread in the histogram and scatter prologues
#define GR_WAVES (GR_BLOCK / 64)
#define GR_SCAN_TILES (GR_BLOCK * 64)  // most tiles one exclusive scan handles (2^25 items)

#define GRK extern "C" __global__ void __launch_bounds__(GR_BLOCK)
static_assert(GR_BPT == 2, "two consecutive digits per thread in the scatter prologue");
#define GDEV __device__ __forceinline__

// Device control block: the bounding box of the finite keys and their count (filled by
// gr_box_reduce), then the layout chosen by the host.
struct GrCtrl {
  int32_t mnx, mny, mxx, mxy;
  uint32_t m;         // finite points
  uint32_t dense;     // 1: Dense layout
  int64_t H;          // bounding-box height (dense index stride)
  uint32_t lo_bits;   // sparse: bits of the (cy - min_cy) part of the packed key
  uint32_t ncells;    // dense: W * H; sparse: occupied cells (written by gr_sparse_cells)
  uint32_t pad[2];
};

namespace {

// `as i32` of an f32 (saturating; the callers pass finite values).
GDEV int32_t sat_i32(float v) {
  if (v <= -2147483648.0f) return INT32_MIN;
  if (v >= 2147483647.0f) return INT32_MAX;
  return (int32_t)v;
}
GDEV bool finite2(float2 p) { return isfinite(p.x) && isfinite(p.y); }
GDEV int2 cell_key(float2 p, float cell) {  // grid.rs:280-285
  return make_int2(sat_i32(floorf(p.x / cell)), sat_i32(floorf(p.y / cell)));
}
GDEV int32_t sadd(int32_t a, int32_t b) {  // i32::saturating_add
  int64_t r = (int64_t)a + b;
  return r > INT32_MAX ? INT32_MAX : (r < INT32_MIN ? INT32_MIN : (int32_t)r);
}
GDEV uint32_t lane_id() { return threadIdx.x & 63u; }
GDEV uint64_t lanes_below() { return (1ull << lane_id()) - 1ull; }

template <class T> GDEV T wave_min(T v) {
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
  return v;
}
template <class T> GDEV T wave_max(T v) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}

// Exclusive scan of one value per thread over the workgroup; returns the workgroup total.
GDEV uint32_t block_scan_excl(uint32_t v, uint32_t& total, uint32_t* s_wave) {
  const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
  uint32_t inc = v;
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t t = __shfl_up(inc, o, 64);
    if (lane >= (uint32_t)o) inc += t;
  }
  if (lane == 63) s_wave[w] = inc;
  __syncthreads();
  uint32_t before = 0, tot = 0;
  for (uint32_t q = 0; q < GR_WAVES; q++) {
    uint32_t s = s_wave[q];
    before += q < w ? s : 0u;
    tot += s;
  }
  __syncthreads();
  total = tot;
  return before + inc - v;
}

}  // namespace

// ---- layout ------------------------------------------------------------------------------------
GRK gr_init(GrCtrl* __restrict__ c) {
  if (threadIdx.x == 0) {
    c->mnx = INT32_MAX; c->mny = INT32_MAX; c->mxx = INT32_MIN; c->mxy = INT32_MIN;
    c->m = 0; c->dense = 0; c->H = 0; c->lo_bits = 0; c->ncells = 0;
  }
}

// Bounding box and count of the finite points' keys (grid.rs:108-129), per tile: tile_box[4 t ..]
// = (min cx, min cy, max cx, max cy) and tile_cnt[t] = finite points (also the compaction's
// counts). gr_box_reduce folds the tiles into the control block; no atomics on one cache line
// (per-wave atomics there serialised: 122 us at 1M points, per-workgroup ones 35 us).
GRK gr_keys(const float2* __restrict__ xy, uint32_t n, float cell, int32_t* __restrict__ tile_box,
            uint32_t* __restrict__ tile_cnt) {
  __shared__ int32_t s_box[GR_WAVES][4];
  __shared__ uint32_t s_cnt[GR_WAVES];
  int32_t mnx = INT32_MAX, mny = INT32_MAX, mxx = INT32_MIN, mxy = INT32_MIN;
  uint32_t cnt = 0;
  const uint32_t base = blockIdx.x * GR_TILE;
  for (uint32_t q = 0; q < GR_ITEMS; q++) {
    uint32_t i = base + q * GR_BLOCK + threadIdx.x;
    if (i >= n) break;
    float2 p = xy[i];
    if (!finite2(p)) continue;
    int2 k = cell_key(p, cell);
    mnx = min(mnx, k.x); mny = min(mny, k.y); mxx = max(mxx, k.x); mxy = max(mxy, k.y);
    cnt++;
  }
  mnx = wave_min(mnx); mny = wave_min(mny); mxx = wave_max(mxx); mxy = wave_max(mxy);
  for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o, 64);
  const uint32_t w = threadIdx.x >> 6;
  if (lane_id() == 0) {
    s_box[w][0] = mnx; s_box[w][1] = mny; s_box[w][2] = mxx; s_box[w][3] = mxy;
    s_cnt[w] = cnt;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    int32_t v = s_box[0][threadIdx.x];
    for (uint32_t q = 1; q < GR_WAVES; q++) v = threadIdx.x < 2 ? min(v, s_box[q][threadIdx.x]) : max(v, s_box[q][threadIdx.x]);
    tile_box[4 * blockIdx.x + threadIdx.x] = v;
  }
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (uint32_t q = 0; q < GR_WAVES; q++) t += s_cnt[q];
    tile_cnt[blockIdx.x] = t;
  }
}

// One workgroup: the bounding box and finite count over nt tiles into the control block.
GRK gr_box_reduce(const int32_t* __restrict__ tile_box, const uint32_t* __restrict__ tile_cnt, uint32_t nt,
                  GrCtrl* __restrict__ c) {
  __shared__ int32_t s_box[GR_WAVES][4];
  __shared__ uint32_t s_cnt[GR_WAVES];
  int32_t mnx = INT32_MAX, mny = INT32_MAX, mxx = INT32_MIN, mxy = INT32_MIN;
  uint32_t cnt = 0;
  for (uint32_t t = threadIdx.x; t < nt; t += GR_BLOCK) {
    const int4 b = reinterpret_cast<const int4*>(tile_box)[t];
    mnx = min(mnx, b.x); mny = min(mny, b.y); mxx = max(mxx, b.z); mxy = max(mxy, b.w);
    cnt += tile_cnt[t];
  }
  mnx = wave_min(mnx); mny = wave_min(mny); mxx = wave_max(mxx); mxy = wave_max(mxy);
  for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o, 64);
  const uint32_t w = threadIdx.x >> 6;
  if (lane_id() == 0) {
    s_box[w][0] = mnx; s_box[w][1] = mny; s_box[w][2] = mxx; s_box[w][3] = mxy;
    s_cnt[w] = cnt;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (uint32_t q = 1; q < GR_WAVES; q++) {
      mnx = min(mnx, s_box[q][0]); mny = min(mny, s_box[q][1]);
      mxx = max(mxx, s_box[q][2]); mxy = max(mxy, s_box[q][3]);
    }
    uint32_t t = 0;
    for (uint32_t q = 0; q < GR_WAVES; q++) t += s_cnt[q];
    c->mnx = mnx; c->mny = mny; c->mxx = mxx; c->mxy = mxy;
    c->m = t;
  }
}

// Exclusive scan of up to GR_BLOCK * 64 per-tile values in one workgroup (tile offsets of the
// compaction, sorted-position offsets); out[nt] receives the total.
GRK gr_scan_small(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, uint32_t nt) {
  __shared__ uint32_t s_wave[GR_WAVES];
  __shared__ uint32_t s_carry;
  if (threadIdx.x == 0) s_carry = 0;
  __syncthreads();
  for (uint32_t b = 0; b < nt; b += GR_BLOCK) {
    uint32_t i = b + threadIdx.x;
    uint32_t v = i < nt ? in[i] : 0u, tot;
    uint32_t ex = block_scan_excl(v, tot, s_wave);
    uint32_t carry = s_carry;
    if (i < nt) out[i] = carry + ex;
    __syncthreads();
    if (threadIdx.x == 0) s_carry = carry + tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) out[nt] = s_carry;
}

// Sort keys of the finite points in index order: dense k = dcx * H + dcy (grid.rs:232-236),
// sparse dcx << lo_bits | dcy (the (cx, cy) order of grid.rs:204-226). `compact`: some point is
// non-finite, so finite points go to tile offset + rank among the tile's finite points.
template <class K>
__device__ void sort_keys_body(const float2* __restrict__ xy, uint32_t n, float cell, const GrCtrl* __restrict__ c,
                               const uint32_t* __restrict__ tile_off, int compact, K* __restrict__ key,
                               uint32_t* __restrict__ idx) {
  __shared__ uint32_t s_wave[GR_WAVES];
  __shared__ uint32_t s_run;
  const int32_t mnx = c->mnx, mny = c->mny;
  const int64_t H = c->H;
  const uint32_t lo = c->lo_bits;
  if (threadIdx.x == 0) s_run = compact ? tile_off[blockIdx.x] : 0u;
  __syncthreads();
  const uint32_t base = blockIdx.x * GR_TILE;
  for (uint32_t q = 0; q < GR_ITEMS; q++) {
    uint32_t i = base + q * GR_BLOCK + threadIdx.x;
    bool ok = false;
    int2 k = make_int2(0, 0);
    if (i < n) {
      float2 p = xy[i];
      ok = finite2(p);
      if (ok) k = cell_key(p, cell);
    }
    uint32_t tot;
    uint32_t pos = block_scan_excl(ok ? 1u : 0u, tot, s_wave);
    uint32_t run = s_run;
    if (ok) {
      uint64_t dcx = (uint64_t)((int64_t)k.x - mnx), dcy = (uint64_t)((int64_t)k.y - mny);
      K sk = sizeof(K) == 4 ? (K)(dcx * (uint64_t)H + dcy) : (K)((dcx << lo) | dcy);
      uint32_t at = compact ? run + pos : i;
      key[at] = sk;
      idx[at] = i;
    }
    __syncthreads();
    if (threadIdx.x == 0) s_run = run + tot;
    __syncthreads();
  }
}
GRK gr_sort_keys32(const float2* __restrict__ xy, uint32_t n, float cell, const GrCtrl* __restrict__ c,
                   const uint32_t* __restrict__ tile_off, int compact, uint32_t* __restrict__ key,
                   uint32_t* __restrict__ idx) {
  sort_keys_body<uint32_t>(xy, n, cell, c, tile_off, compact, key, idx);
}
GRK gr_sort_keys64(const float2* __restrict__ xy, uint32_t n, float cell, const GrCtrl* __restrict__ c,
                   const uint32_t* __restrict__ tile_off, int compact, uint64_t* __restrict__ key,
                   uint32_t* __restrict__ idx) {
  sort_keys_body<uint64_t>(xy, n, cell, c, tile_off, compact, key, idx);
}

// ---- stable LSD radix sort ---------------------------------------------------------------------
// Digit histogram of one tile, stored digit-major: hist[d * ntiles + tile].
template <class K>
__device__ void radix_hist_body(const K* __restrict__ key, uint32_t m, uint32_t shift, uint32_t* __restrict__ hist,
                                uint32_t ntiles) {
  __shared__ uint32_t s_h[GR_BINS];
  for (uint32_t b = threadIdx.x; b < GR_BINS; b += GR_BLOCK) s_h[b] = 0;
  __syncthreads();
  const uint32_t base = blockIdx.x * GR_TILE;
  for (uint32_t q = 0; q < GR_ITEMS; q++) {
    uint32_t i = base + q * GR_BLOCK + threadIdx.x;
    if (i < m) atomicAdd(&s_h[(uint32_t)(key[i] >> shift) & (GR_BINS - 1)], 1u);
  }
  __syncthreads();
  for (uint32_t b = threadIdx.x; b < GR_BINS; b += GR_BLOCK) hist[(size_t)b * ntiles + blockIdx.x] = s_h[b];
}
GRK gr_radix_hist32(const uint32_t* __restrict__ key, uint32_t m, uint32_t shift, uint32_t* __restrict__ hist,
                    uint32_t ntiles) {
  radix_hist_body<uint32_t>(key, m, shift, hist, ntiles);
}
GRK gr_radix_hist64(const uint64_t* __restrict__ key, uint32_t m, uint32_t shift, uint32_t* __restrict__ hist,
                    uint32_t ntiles) {
  radix_hist_body<uint64_t>(key, m, shift, hist, ntiles);
}

// One workgroup per digit: exclusive scan of that digit's tile counts in place; dtot[d] = total.
GRK gr_radix_dscan(uint32_t* __restrict__ hist, uint32_t ntiles, uint32_t* __restrict__ dtot) {
  __shared__ uint32_t s_wave[GR_WAVES];
  __shared__ uint32_t s_carry;
  uint32_t* row = hist + (size_t)blockIdx.x * ntiles;
  if (threadIdx.x == 0) s_carry = 0;
  __syncthreads();
  for (uint32_t b = 0; b < ntiles; b += GR_BLOCK) {
    uint32_t i = b + threadIdx.x;
    uint32_t v = i < ntiles ? row[i] : 0u, tot;
    uint32_t ex = block_scan_excl(v, tot, s_wave);
    uint32_t carry = s_carry;
    if (i < ntiles) row[i] = carry + ex;
    __syncthreads();
    if (threadIdx.x == 0) s_carry = carry + tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) dtot[blockIdx.x] = s_carry;
}

// Stable scatter of one tile: items of a digit keep their order (tile order, then sub-tile,
// then wave, then lane = ascending position).
template <class K>
__device__ void radix_scatter_body(const K* __restrict__ kin, const uint32_t* __restrict__ iin, K* __restrict__ kout,
                                   uint32_t* __restrict__ iout, uint32_t m, uint32_t shift,
                                   const uint32_t* __restrict__ hist, uint32_t ntiles,
                                   const uint32_t* __restrict__ dtot) {
  __shared__ uint32_t s_wave[GR_WAVES];
  __shared__ uint32_t s_base[GR_BINS];
  __shared__ uint32_t s_wcnt[GR_WAVES][GR_BINS];
  const uint32_t t = threadIdx.x, w = t >> 6;
  {  // thread t owns digits 2t and 2t + 1: their start = digits before them + this tile's offset
    const uint32_t d0 = GR_BPT * t;
    const uint32_t a0 = dtot[d0], a1 = dtot[d0 + 1];
    uint32_t tot;
    const uint32_t db = block_scan_excl(a0 + a1, tot, s_wave);
    s_base[d0] = db + hist[(size_t)d0 * ntiles + blockIdx.x];
    s_base[d0 + 1] = db + a0 + hist[(size_t)(d0 + 1) * ntiles + blockIdx.x];
  }
  const uint32_t base = blockIdx.x * GR_TILE;
  for (uint32_t q = 0; q < GR_ITEMS; q++) {
    for (uint32_t a = 0; a < GR_WAVES; a++)
      for (uint32_t b = t; b < GR_BINS; b += GR_BLOCK) s_wcnt[a][b] = 0;
    __syncthreads();
    const uint32_t i = base + q * GR_BLOCK + t;
    const bool ok = i < m;
    K k = 0;
    uint32_t v = 0, d = 0;
    if (ok) {
      k = kin[i];
      v = iin[i];
      d = (uint32_t)(k >> shift) & (GR_BINS - 1);
    }
    uint64_t peers = __ballot(ok);
    for (int b = 0; b < GR_RADIX_BITS; b++) {
      uint64_t s = __ballot(ok && ((d >> b) & 1u));
      peers &= ((d >> b) & 1u) ? s : ~s;
    }
    const uint32_t rank = (uint32_t)__popcll(peers & lanes_below());
    if (ok && rank == 0) s_wcnt[w][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    if (ok) {
      uint32_t pos = s_base[d] + rank;
      for (uint32_t a = 0; a < w; a++) pos += s_wcnt[a][d];
      kout[pos] = k;
      iout[pos] = v;
    }
    __syncthreads();
    for (uint32_t b = t; b < GR_BINS; b += GR_BLOCK) {
      uint32_t add = 0;
      for (uint32_t a = 0; a < GR_WAVES; a++) add += s_wcnt[a][b];
      s_base[b] += add;
    }
    __syncthreads();
  }
}
GRK gr_radix_scatter32(const uint32_t* __restrict__ kin, const uint32_t* __restrict__ iin, uint32_t* __restrict__ kout,
                       uint32_t* __restrict__ iout, uint32_t m, uint32_t shift, const uint32_t* __restrict__ hist,
                       uint32_t ntiles, const uint32_t* __restrict__ dtot) {
  radix_scatter_body<uint32_t>(kin, iin, kout, iout, m, shift, hist, ntiles, dtot);
}
GRK gr_radix_scatter64(const uint64_t* __restrict__ kin, const uint32_t* __restrict__ iin, uint64_t* __restrict__ kout,
                       uint32_t* __restrict__ iout, uint32_t m, uint32_t shift, const uint32_t* __restrict__ hist,
                       uint32_t ntiles, const uint32_t* __restrict__ dtot) {
  radix_scatter_body<uint64_t>(kin, iin, kout, iout, m, shift, hist, ntiles, dtot);
}

// ---- cell starts -------------------------------------------------------------------------------
// Dense: sorted position p (0..m) owns the cells (key[p-1], key[p]] (with key[-1] = -1 and
// key[m] = W*H): their start is p, the number of points with a smaller cell index.
GRK gr_dense_starts(const uint32_t* __restrict__ skey, uint32_t m, uint32_t ncells, uint32_t* __restrict__ cell_start) {
  uint32_t p = blockIdx.x * GR_BLOCK + threadIdx.x;
  if (p > m) return;
  int64_t prev = p > 0 ? (int64_t)skey[p - 1] : -1;
  int64_t cur = p < m ? (int64_t)skey[p] : (int64_t)ncells;
  for (int64_t c = prev + 1; c <= cur; c++) cell_start[c] = p;
}
// Sparse: flags of the first position of each occupied cell.
GRK gr_sparse_flags(const uint64_t* __restrict__ skey, uint32_t m, uint32_t* __restrict__ flag) {
  uint32_t p = blockIdx.x * GR_BLOCK + threadIdx.x;
  if (p >= m) return;
  flag[p] = (p == 0 || skey[p] != skey[p - 1]) ? 1u : 0u;
}
// Per-tile sums of a u32 array (first level of the large exclusive scan).
GRK gr_tile_sums(const uint32_t* __restrict__ v, uint32_t n, uint32_t* __restrict__ sums) {
  __shared__ uint32_t s_wave[GR_WAVES];
  const uint32_t base = blockIdx.x * GR_TILE;
  uint32_t acc = 0;
  for (uint32_t q = 0; q < GR_ITEMS; q++) {
    uint32_t i = base + q * GR_BLOCK + threadIdx.x;
    if (i < n) acc += v[i];
  }
  uint32_t tot;
  block_scan_excl(acc, tot, s_wave);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}
// Exclusive scan of v (tile offsets from gr_scan_small over gr_tile_sums); out[n] = total when
// total_out.
GRK gr_tile_scan(const uint32_t* v, uint32_t n, const uint32_t* __restrict__ tile_off, uint32_t* out) {
  __shared__ uint32_t s_wave[GR_WAVES];
  __shared__ uint32_t s_run;
  if (threadIdx.x == 0) s_run = tile_off[blockIdx.x];
  __syncthreads();
  const uint32_t base = blockIdx.x * GR_TILE;
  for (uint32_t q = 0; q < GR_ITEMS; q++) {
    uint32_t i = base + q * GR_BLOCK + threadIdx.x;
    uint32_t x = i < n ? v[i] : 0u, tot;
    uint32_t ex = block_scan_excl(x, tot, s_wave);
    uint32_t run = s_run;
    if (i < n) out[i] = run + ex;
    __syncthreads();
    if (threadIdx.x == 0) s_run = run + tot;
    __syncthreads();
  }
}
// Sparse cells: key and start of each occupied cell, cell_start[ncells] = m.
GRK gr_sparse_cells(const uint64_t* __restrict__ skey, const uint32_t* __restrict__ flag,
                    const uint32_t* __restrict__ slot, uint32_t m, GrCtrl* __restrict__ c,
                    int2* __restrict__ keys, uint32_t* __restrict__ cell_start) {
  uint32_t p = blockIdx.x * GR_BLOCK + threadIdx.x;
  if (p > m) return;
  if (p == m) {
    uint32_t nc = m ? slot[m - 1] + flag[m - 1] : 0u;
    c->ncells = nc;
    cell_start[nc] = m;
    return;
  }
  if (!flag[p]) return;
  const uint32_t lo = c->lo_bits;
  const uint64_t k = skey[p];
  const uint64_t dcy = lo ? (k & ((1ull << lo) - 1ull)) : 0ull, dcx = k >> lo;
  keys[slot[p]] = make_int2((int32_t)((int64_t)c->mnx + (int64_t)dcx), (int32_t)((int64_t)c->mny + (int64_t)dcy));
  cell_start[slot[p]] = p;
}
GRK gr_set_u32(uint32_t* __restrict__ p, uint32_t v) {
  if (threadIdx.x == 0) *p = v;
}
GRK gr_copy_u32(uint32_t* __restrict__ dst, const uint32_t* __restrict__ src) {
  if (threadIdx.x == 0) *dst = *src;
}
GRK gr_copy_n(uint32_t* __restrict__ dst, const uint32_t* __restrict__ src, uint32_t n) {
  uint32_t i = blockIdx.x * GR_BLOCK + threadIdx.x;
  if (i < n) dst[i] = src[i];
}
// The layout the host chose from the bounding box (kernel arguments, so nothing is read from
// host memory after the launch returns).
GRK gr_set_layout(GrCtrl* __restrict__ c, uint32_t dense, int64_t H, uint32_t lo_bits, uint32_t ncells) {
  if (threadIdx.x == 0) {
    c->dense = dense;
    c->H = H;
    c->lo_bits = lo_bits;
    c->ncells = ncells;
  }
}

// ---- queries in the reference's order ------------------------------------------------------------
struct GridView {
  const uint32_t* start;
  const uint32_t* idx;
  const int2* keys;  // sparse
  int32_t mnx, mny;
  int64_t W, H;
  uint32_t ncells;
  uint32_t dense;
  float cell;
};

namespace {

GDEV bool key_less(int2 a, int2 b) { return a.x < b.x || (a.x == b.x && a.y < b.y); }

// Range [s, e) of cell (cx, cy); false when out of the box or empty (grid.rs:242-271).
GDEV bool cell_range(const GridView& g, int32_t cx, int32_t cy, uint32_t& s, uint32_t& e) {
  if (g.ncells == 0) return false;
  if (g.dense) {
    int64_t dx = (int64_t)cx - g.mnx, dy = (int64_t)cy - g.mny;
    if (dx < 0 || dy < 0 || dx >= g.W || dy >= g.H) return false;
    int64_t k = dx * g.H + dy;
    s = g.start[k];
    e = g.start[k + 1];
    return s != e;
  }
  uint32_t lo = 0, hi = g.ncells;
  const int2 want = make_int2(cx, cy);
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (key_less(g.keys[mid], want)) lo = mid + 1; else hi = mid;
  }
  if (lo >= g.ncells || g.keys[lo].x != cx || g.keys[lo].y != cy) return false;
  s = g.start[lo];
  e = g.start[lo + 1];
  return true;
}

// The forward offsets of for_each_candidate_pair, in the reference's order (grid.rs:300).
__constant__ int kFwdX[4] = {1, 1, 0, -1};
__constant__ int kFwdY[4] = {0, 1, 1, 1};

// Visits the partners of point p (in cell c) in the order its pairs are emitted by
// for_each_candidate_pair (grid.rs:291-362): cells are visited in ascending (cx, cy), so every
// cell q that reaches c through a forward offset lies in the 3x3 block around c and is visited,
// in x-major order, before or after c itself; at c come its own points (ascending index, p
// skipped) and then the points of its four forward cells. f(j) is called once per emitted pair.
// With keys at the i32 saturation boundary a cell can reach c through more than one offset; the
// reference then emits the pair once per offset, and so does this walk.
template <class F>
GDEV void for_each_partner(const GridView& g, uint32_t p, int2 c, F&& f) {
  for (int qx = -1; qx <= 1; qx++) {
    const int64_t ax = (int64_t)c.x + qx;
    if (ax < INT32_MIN || ax > INT32_MAX) continue;
    for (int qy = -1; qy <= 1; qy++) {
      const int64_t ay = (int64_t)c.y + qy;
      if (ay < INT32_MIN || ay > INT32_MAX) continue;
      const int32_t x = (int32_t)ax, y = (int32_t)ay;
      uint32_t s, e;
      if (qx == 0 && qy == 0) {
        if (!cell_range(g, x, y, s, e)) continue;
        for (uint32_t a = s; a < e; a++) {
          uint32_t j = g.idx[a];
          if (j != p) f(j);
        }
        for (int o = 0; o < 4; o++) {
          int32_t tx = sadd(x, kFwdX[o]), ty = sadd(y, kFwdY[o]);
          if (tx == x && ty == y) continue;
          uint32_t ts, te;
          if (!cell_range(g, tx, ty, ts, te)) continue;
          for (uint32_t a = ts; a < te; a++) f(g.idx[a]);
        }
        continue;
      }
      bool have = false;
      for (int o = 0; o < 4; o++) {
        if (sadd(x, kFwdX[o]) != c.x || sadd(y, kFwdY[o]) != c.y) continue;
        if (!have) {
          if (!cell_range(g, x, y, s, e)) break;
          have = true;
        }
        for (uint32_t a = s; a < e; a++) f(g.idx[a]);
      }
    }
  }
}

}  // namespace

// build_neighbor_lists (pbf.rs:262-288): partners with |x_p - x_j|^2 <= h^2 in emission order.
// pass 0 counts into cnt[p]; pass 1 writes at off[p] (entries at or beyond cap are dropped).
GRK gr_neighbors(GridView g, const float2* __restrict__ xy, uint32_t n, float h2, int pass,
                 uint32_t* cnt, const uint32_t* off, uint32_t* __restrict__ nb, uint64_t cap) {
  const uint32_t p = blockIdx.x * GR_BLOCK + threadIdx.x;
  if (p >= n) return;
  const float2 xp = xy[p];
  uint32_t k = 0;
  if (finite2(xp) && g.ncells) {
    const uint64_t base = pass ? off[p] : 0u;
    for_each_partner(g, p, cell_key(xp, g.cell), [&](uint32_t j) {
      float2 xj = xy[j];
      float dx = xp.x - xj.x, dy = xp.y - xj.y;
      if (dx * dx + dy * dy <= h2) {
        if (pass && base + k < cap) nb[base + k] = j;
        k++;
      }
    });
  }
  if (!pass) cnt[p] = k;
}

// for_each_candidate_pair (grid.rs:291-362): pass 0 counts the pairs each cell emits; pass 1
// writes them (i < j) at the cell's offset (pairs at or beyond cap are dropped).
GRK gr_cell_pairs(GridView g, int pass, uint32_t* __restrict__ cnt, const uint32_t* __restrict__ off,
                  uint32_t* __restrict__ out, uint64_t cap) {
  const uint32_t k = blockIdx.x * GR_BLOCK + threadIdx.x;
  if (k >= g.ncells) return;
  const uint32_t s = g.start[k], e = g.start[k + 1];
  if (s == e) {
    if (!pass) cnt[k] = 0;
    return;
  }
  int32_t cx, cy;
  if (g.dense) {
    cx = (int32_t)((int64_t)g.mnx + (int64_t)k / g.H);
    cy = (int32_t)((int64_t)g.mny + (int64_t)k % g.H);
  } else {
    cx = g.keys[k].x;
    cy = g.keys[k].y;
  }
  uint64_t w = pass ? off[k] : 0u;
  uint32_t c = 0;
  auto emit = [&](uint32_t i, uint32_t j) {
    if (pass && w + c < cap) {
      out[2 * (w + c)] = min(i, j);
      out[2 * (w + c) + 1] = max(i, j);
    }
    c++;
  };
  for (uint32_t a = s; a < e; a++)
    for (uint32_t b = a + 1; b < e; b++) emit(g.idx[a], g.idx[b]);
  for (int o = 0; o < 4; o++) {
    int32_t tx = sadd(cx, kFwdX[o]), ty = sadd(cy, kFwdY[o]);
    if (tx == cx && ty == cy) continue;
    uint32_t ts, te;
    if (!cell_range(g, tx, ty, ts, te)) continue;
    for (uint32_t a = s; a < e; a++)
      for (uint32_t b = ts; b < te; b++) emit(g.idx[a], g.idx[b]);
  }
  if (!pass) cnt[k] = c;
}

// for_each_near (grid.rs:371-399): the points of the up-to-9 distinct cells around query q
// (dy-major order, duplicates at the saturation boundary removed); non-finite queries get none.
GRK gr_near(GridView g, const float2* __restrict__ qxy, uint32_t nq, int pass, uint32_t* cnt, const uint32_t* off,
            uint32_t* __restrict__ out, uint64_t cap) {
  const uint32_t q = blockIdx.x * GR_BLOCK + threadIdx.x;
  if (q >= nq) return;
  const float2 p = qxy[q];
  uint32_t c = 0;
  if (finite2(p) && g.ncells) {
    const int2 k = cell_key(p, g.cell);
    const uint64_t base = pass ? off[q] : 0u;
    for (int dy = -1; dy <= 1; dy++)
      for (int dx = -1; dx <= 1; dx++) {
        const int32_t x = sadd(k.x, dx), y = sadd(k.y, dy);
        // earlier (dy', dx') in the same loop order that saturated onto the same cell
        bool dup = false;
        for (int ey = -1; ey <= dy; ey++)
          for (int ex = -1; ex <= 1; ex++) {
            if (ey == dy && ex >= dx) break;
            dup |= sadd(k.x, ex) == x && sadd(k.y, ey) == y;
          }
        if (dup) continue;
        uint32_t s, e;
        if (!cell_range(g, x, y, s, e)) continue;
        for (uint32_t a = s; a < e; a++) {
          if (pass && base + c < cap) out[base + c] = g.idx[a];
          c++;
        }
      }
  }
  if (!pass) cnt[q] = c;
}

// ---- host side -----------------------------------------------------------------------------------
// Device buffers of one build: scratch and outputs. The handle owns all of them.
struct GrBufs {
  GrCtrl* ctrl = nullptr;
  uint32_t* tile_cnt = nullptr;  // per tile (+1 total)
  uint32_t* tile_off = nullptr;
  uint32_t *idxA = nullptr, *idxB = nullptr;
  void *keyA = nullptr, *keyB = nullptr;  // u32 (dense) or u64 (sparse) sort keys
  uint32_t* hist = nullptr;               // GR_BINS x tiles
  uint32_t* dtot = nullptr;               // GR_BINS
  uint32_t* flag = nullptr;               // sparse boundaries / query counts (max(n, cells) + 1)
  uint32_t* slot = nullptr;
  uint32_t* sums = nullptr;               // per-tile sums of the large scans (+1)
  uint32_t* sums_off = nullptr;
  // outputs
  uint32_t* cell_start = nullptr;  // 2 n + 513 entries
  uint32_t* indices = nullptr;     // n entries
  int2* keys = nullptr;            // n entries (sparse)
};

struct mg_grid {
  uint32_t cap = 0;
  GrBufs b;
  void* arena = nullptr;
  GrCtrl* h_ctrl = nullptr;  // pinned mirror
  GridView view{};
  float cell = 0.0f;
  uint32_t n = 0, m = 0, ncells = 0;
  bool dense = false, built = false;
  int64_t W = 0, H = 0;
  int32_t mnx = 0, mny = 0;
  std::string err;
};

namespace {

uint32_t tiles_of(uint64_t n) { return (uint32_t)((n + GR_TILE - 1) / GR_TILE); }
uint32_t blocks_of(uint64_t n) { return (uint32_t)((n + GR_BLOCK - 1) / GR_BLOCK); }
uint32_t bitlen(uint64_t v) { return v ? 64u - (uint32_t)__builtin_clzll(v) : 0u; }

template <class T> size_t carve(size_t off, T*& p, size_t count, char* base) {
  off = (off + 255) & ~(size_t)255;
  p = reinterpret_cast<T*>(base ? base + off : nullptr);
  return off + count * sizeof(T);
}

// Lays out every buffer for up to cap points from `base` (null: only measures).
size_t layout(GrBufs& b, uint32_t cap, char* base) {
  const uint64_t nt = tiles_of(cap) + 1, cells = 2ull * cap + 513;
  size_t off = 0;
  off = carve(off, b.ctrl, 1, base);
  off = carve(off, b.tile_cnt, nt + 1, base);
  off = carve(off, b.tile_off, nt + 1, base);
  off = carve(off, b.idxA, cap + 1, base);
  off = carve(off, b.idxB, cap + 1, base);
  uint64_t* ka; uint64_t* kb;
  off = carve(off, ka, cap + 1, base);
  off = carve(off, kb, cap + 1, base);
  b.keyA = ka; b.keyB = kb;
  off = carve(off, b.hist, (size_t)GR_BINS * nt, base);
  off = carve(off, b.dtot, GR_BINS, base);
  const uint64_t big = std::max<uint64_t>(cells, cap) + 1;
  off = carve(off, b.flag, big, base);
  off = carve(off, b.slot, big, base);
  off = carve(off, b.sums, GR_SCAN_TILES + 2, base);
  off = carve(off, b.sums_off, GR_SCAN_TILES + 2, base);
  off = carve(off, b.cell_start, cells, base);
  off = carve(off, b.indices, cap + 1, base);
  off = carve(off, b.keys, cap + 1, base);
  return off;
}

hipError_t last() { return hipGetLastError(); }

// Exclusive scan of v[0..n) into out[0..n], out[n] = total (tile sums, one-workgroup scan of the
// sums, then the tiles); out may be v (every element is read before its thread writes it).
bool scan_u32(const GrBufs& b, const uint32_t* v, uint32_t n, uint32_t* out, hipStream_t s) {
  const uint32_t nt = std::max(tiles_of(n), 1u);
  if (nt > GR_SCAN_TILES) return false;
  hipLaunchKernelGGL(gr_tile_sums, dim3(nt), dim3(GR_BLOCK), 0, s, v, n, b.sums);
  hipLaunchKernelGGL(gr_scan_small, dim3(1), dim3(GR_BLOCK), 0, s, b.sums, b.sums_off, nt);
  hipLaunchKernelGGL(gr_tile_scan, dim3(nt), dim3(GR_BLOCK), 0, s, v, n, b.sums_off, out);
  // out[n] = total: the one-workgroup scan left it in sums_off[nt]
  hipLaunchKernelGGL(gr_copy_u32, dim3(1), dim3(64), 0, s, out + n, b.sums_off + nt);
  return last() == hipSuccess;
}

// Stable LSD radix sort of (keyA, idxA)[0..m) on the low `bits` bits of the key; the sorted
// indices end in idx_final (the last pass scatters straight into it), the sorted keys in key_out.
template <class K>
bool radix_sort(const GrBufs& b, uint32_t m, uint32_t bits, K*& key_out, uint32_t* idx_final, hipStream_t s) {
  K* kin = static_cast<K*>(b.keyA);
  K* kout = static_cast<K*>(b.keyB);
  uint32_t* iin = b.idxA;
  uint32_t* iout = b.idxB;
  const uint32_t nt = std::max(tiles_of(m), 1u);
  if (bits == 0) {  // one cell: index order is the cell order
    hipLaunchKernelGGL(gr_copy_n, dim3(std::max(blocks_of(m), 1u)), dim3(GR_BLOCK), 0, s, idx_final, iin, m);
    key_out = kin;
    return last() == hipSuccess;
  }
  for (uint32_t shift = 0; shift < bits; shift += GR_RADIX_BITS) {
    if (shift + GR_RADIX_BITS >= bits) iout = idx_final;
    if (sizeof(K) == 4) {
      hipLaunchKernelGGL(gr_radix_hist32, dim3(nt), dim3(GR_BLOCK), 0, s, (const uint32_t*)kin, m, shift, b.hist, nt);
    } else {
      hipLaunchKernelGGL(gr_radix_hist64, dim3(nt), dim3(GR_BLOCK), 0, s, (const uint64_t*)kin, m, shift, b.hist, nt);
    }
    hipLaunchKernelGGL(gr_radix_dscan, dim3(GR_BINS), dim3(GR_BLOCK), 0, s, b.hist, nt, b.dtot);
    if (sizeof(K) == 4) {
      hipLaunchKernelGGL(gr_radix_scatter32, dim3(nt), dim3(GR_BLOCK), 0, s, (const uint32_t*)kin, iin,
                         (uint32_t*)kout, iout, m, shift, b.hist, nt, b.dtot);
    } else {
      hipLaunchKernelGGL(gr_radix_scatter64, dim3(nt), dim3(GR_BLOCK), 0, s, (const uint64_t*)kin, iin,
                         (uint64_t*)kout, iout, m, shift, b.hist, nt, b.dtot);
    }
    std::swap(kin, kout);
    std::swap(iin, iout);
  }
  key_out = kin;
  return last() == hipSuccess;
}

void set_err(char* err, size_t len, const std::string& m) {
  if (err && len) {
    std::strncpy(err, m.c_str(), len - 1);
    err[len - 1] = 0;
  }
}

// Entries the output buffers of a build hold (cell_start, indices, Sparse keys).
struct OutCaps {
  uint64_t cs = UINT64_MAX, idx = UINT64_MAX, keys = UINT64_MAX;
};

// UniformGrid::build over n device points with buffers b (capacity >= n). Fills the host-side
// summary of g. Returns MG_E_CAPACITY, before any output is written, when the layout needs more
// entries than `caps` allows.
int build_into(mg_grid& g, const GrBufs& b, const float2* xy, uint32_t n, float cell, hipStream_t s,
               OutCaps caps = OutCaps()) {
  if (!(cell > 0.0f) || !std::isfinite(cell)) {
    g.err = "cell_size must be positive";
    return MG_E_ARG;
  }
  g.built = false;
  const uint32_t nt = std::max(tiles_of(n), 1u);
  hipLaunchKernelGGL(gr_init, dim3(1), dim3(64), 0, s, b.ctrl);
  if (n) {  // per-tile boxes in the histogram buffer, which the radix sort only uses later
    int32_t* tile_box = reinterpret_cast<int32_t*>(b.hist);
    hipLaunchKernelGGL(gr_keys, dim3(nt), dim3(GR_BLOCK), 0, s, xy, n, cell, tile_box, b.tile_cnt);
    hipLaunchKernelGGL(gr_box_reduce, dim3(1), dim3(GR_BLOCK), 0, s, tile_box, b.tile_cnt, nt, b.ctrl);
  }
  GrCtrl hc;
  if (hipMemcpyAsync(g.h_ctrl ? (void*)g.h_ctrl : (void*)&hc, b.ctrl, sizeof(GrCtrl), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    g.err = "HIP error in the grid key pass";
    return MG_E_DEVICE;
  }
  if (g.h_ctrl) hc = *g.h_ctrl;
  g.n = n;
  g.cell = cell;
  g.m = hc.m;
  const uint32_t m = hc.m;
  if (m == 0) {  // grid.rs:131-139: the empty Sparse layout, cell_start = [0]
    if (caps.cs < 1) { g.err = "cell_start buffer too small (1 entry needed)"; return MG_E_CAPACITY; }
    g.dense = false; g.W = g.H = 0; g.mnx = g.mny = 0; g.ncells = 0;
    hipLaunchKernelGGL(gr_set_u32, dim3(1), dim3(64), 0, s, b.cell_start, 0u);
  } else {
    const int64_t W = (int64_t)hc.mxx - (int64_t)hc.mnx + 1, H = (int64_t)hc.mxy - (int64_t)hc.mny + 1;
    const unsigned __int128 tc = (unsigned __int128)(uint64_t)W * (uint64_t)H;
    const uint64_t total = tc > (unsigned __int128)UINT64_MAX ? UINT64_MAX : (uint64_t)tc;
    const uint64_t budget = 2ull * m + 512ull;
    g.dense = total <= budget;
    const uint64_t need_cs = g.dense ? total + 1 : (uint64_t)m + 1;
    if (caps.cs < need_cs || caps.idx < m || (!g.dense && caps.keys < m)) {
      g.err = "output buffers too small for this grid (cell_start " + std::to_string(need_cs) + ", indices " +
              std::to_string(m) + (g.dense ? std::string() : ", keys " + std::to_string(m)) + " entries needed)";
      return MG_E_CAPACITY;
    }
    g.W = W; g.H = H; g.mnx = hc.mnx; g.mny = hc.mny;
    const uint32_t lo_bits = g.dense ? 0u : bitlen((uint64_t)(H - 1));
    hipLaunchKernelGGL(gr_set_layout, dim3(1), dim3(64), 0, s, b.ctrl, g.dense ? 1u : 0u, H, lo_bits,
                       g.dense ? (uint32_t)total : 0u);
    const int compact = m < n ? 1 : 0;
    if (compact) hipLaunchKernelGGL(gr_scan_small, dim3(1), dim3(GR_BLOCK), 0, s, b.tile_cnt, b.tile_off, nt);
    if (g.dense) {
      g.ncells = (uint32_t)total;
      hipLaunchKernelGGL(gr_sort_keys32, dim3(nt), dim3(GR_BLOCK), 0, s, xy, n, cell, b.ctrl, b.tile_off, compact,
                         (uint32_t*)b.keyA, b.idxA);
      uint32_t* sk;
      if (!radix_sort<uint32_t>(b, m, bitlen(total - 1), sk, b.indices, s)) {
        g.err = "HIP error in the grid sort";
        return MG_E_DEVICE;
      }
      hipLaunchKernelGGL(gr_dense_starts, dim3(blocks_of((uint64_t)m + 1)), dim3(GR_BLOCK), 0, s, sk, m, (uint32_t)total,
                         b.cell_start);
    } else {
      const uint32_t lo = lo_bits, hi = bitlen((uint64_t)(W - 1));
      hipLaunchKernelGGL(gr_sort_keys64, dim3(nt), dim3(GR_BLOCK), 0, s, xy, n, cell, b.ctrl, b.tile_off, compact,
                         (uint64_t*)b.keyA, b.idxA);
      uint64_t* sk;
      if (!radix_sort<uint64_t>(b, m, lo + hi, sk, b.indices, s)) {
        g.err = "HIP error in the grid sort";
        return MG_E_DEVICE;
      }
      hipLaunchKernelGGL(gr_sparse_flags, dim3(blocks_of(m)), dim3(GR_BLOCK), 0, s, sk, m, b.flag);
      if (!scan_u32(b, b.flag, m, b.slot, s)) { g.err = "grid too large for the scan"; return MG_E_ARG; }
      hipLaunchKernelGGL(gr_sparse_cells, dim3(blocks_of((uint64_t)m + 1)), dim3(GR_BLOCK), 0, s, sk, b.flag, b.slot, m,
                         b.ctrl, b.keys, b.cell_start);
      uint32_t nc = 0;
      if (hipMemcpyAsync(&nc, &b.ctrl->ncells, 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
          hipStreamSynchronize(s) != hipSuccess) {
        g.err = "HIP error in the sparse grid";
        return MG_E_DEVICE;
      }
      g.ncells = nc;
    }
  }
  if (last() != hipSuccess) { g.err = "HIP launch error in the grid build"; return MG_E_DEVICE; }
  GridView v{};
  v.start = b.cell_start; v.idx = b.indices; v.keys = b.keys;
  v.mnx = g.mnx; v.mny = g.mny; v.W = g.W; v.H = g.H; v.ncells = g.ncells; v.dense = g.dense ? 1u : 0u; v.cell = cell;
  g.view = v;
  g.built = true;
  return MG_OK;
}

// Count pass into d_off, exclusive scan in place (d_off holds items + 1 entries), fill pass;
// *total = entries (host wait).
template <class Launch>
int two_pass(mg_grid& g, uint32_t items, uint32_t* d_off, uint64_t* total, hipStream_t s, Launch&& launch) {
  const GrBufs& b = g.b;
  launch(0, d_off);
  if (!scan_u32(b, d_off, items, d_off, s)) { g.err = "query too large for the scan"; return MG_E_ARG; }
  launch(1, d_off);
  uint32_t t = 0;
  if (hipMemcpyAsync(&t, d_off + items, 4, hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess ||
      last() != hipSuccess) {
    g.err = "HIP error in a grid query";
    return MG_E_DEVICE;
  }
  *total = t;
  return MG_OK;
}

}  // namespace

extern "C" {

int mg_grid_create(uint32_t max_points, mg_grid** out, char* err, size_t errlen) {
  if (!out) return MG_E_ARG;
  *out = nullptr;
  if (max_points == 0 || max_points > (1u << 30)) {
    set_err(err, errlen, "max_points must be in 1..2^30");
    return MG_E_ARG;
  }
  auto* g = new mg_grid();
  g->cap = max_points;
  const size_t bytes = layout(g->b, max_points, nullptr);
  if (hipMalloc(&g->arena, bytes) != hipSuccess) {
    set_err(err, errlen, "hipMalloc failed for the grid workspace");
    delete g;
    return MG_E_DEVICE;
  }
  layout(g->b, max_points, static_cast<char*>(g->arena));
  if (hipHostMalloc((void**)&g->h_ctrl, sizeof(GrCtrl), hipHostMallocDefault) != hipSuccess) g->h_ctrl = nullptr;
  *out = g;
  return MG_OK;
}

void mg_grid_destroy(mg_grid* g) {
  if (!g) return;
  if (g->arena) (void)hipFree(g->arena);
  if (g->h_ctrl) (void)hipHostFree(g->h_ctrl);
  delete g;
}

int mg_grid_build(mg_grid* g, const float* d_xy, uint32_t n, float cell_size, void* stream) {
  if (!g || (n && !d_xy)) return MG_E_ARG;
  if (n > g->cap) {
    g->err = "more points than the grid's capacity";
    return MG_E_ARG;
  }
  return build_into(*g, g->b, reinterpret_cast<const float2*>(d_xy), n, cell_size, (hipStream_t)stream);
}

int mg_grid_meta(const mg_grid* g, int64_t* meta8) {
  if (!g || !meta8 || !g->built) return MG_E_ARG;
  meta8[0] = g->dense ? 1 : 0;
  meta8[1] = g->mnx;
  meta8[2] = g->mny;
  meta8[3] = g->W;
  meta8[4] = g->H;
  meta8[5] = (int64_t)g->ncells + 1;
  meta8[6] = g->m;
  meta8[7] = g->n;
  return MG_OK;
}

int mg_grid_copy(mg_grid* g, uint32_t* cell_start, size_t cs_cap, uint32_t* indices, size_t idx_cap, int32_t* keys2,
                 size_t keys_cap) {
  if (!g || !g->built) return MG_E_ARG;
  const size_t ncs = (size_t)g->ncells + 1, nk = g->dense ? 0 : g->ncells;
  if ((cell_start && cs_cap < ncs) || (indices && idx_cap < g->m) || (keys2 && keys_cap < nk)) return MG_E_ARG;
  bool ok = true;
  if (cell_start) ok &= hipMemcpy(cell_start, g->b.cell_start, ncs * 4, hipMemcpyDeviceToHost) == hipSuccess;
  if (indices && g->m) ok &= hipMemcpy(indices, g->b.indices, (size_t)g->m * 4, hipMemcpyDeviceToHost) == hipSuccess;
  if (keys2 && nk) ok &= hipMemcpy(keys2, g->b.keys, nk * 8, hipMemcpyDeviceToHost) == hipSuccess;
  return ok ? MG_OK : MG_E_DEVICE;
}

const uint32_t* mg_grid_cell_start_device(const mg_grid* g) { return g ? g->b.cell_start : nullptr; }
const uint32_t* mg_grid_indices_device(const mg_grid* g) { return g ? g->b.indices : nullptr; }
const int32_t* mg_grid_keys_device(const mg_grid* g) { return g ? reinterpret_cast<const int32_t*>(g->b.keys) : nullptr; }

int mg_grid_candidate_pairs(mg_grid* g, uint32_t* d_pairs2, uint64_t cap, uint64_t* n_pairs, void* stream) {
  if (!g || !g->built || !n_pairs || (cap && !d_pairs2)) return MG_E_ARG;
  hipStream_t s = (hipStream_t)stream;
  const uint32_t nc = g->ncells;
  if (nc == 0) {
    *n_pairs = 0;
    return MG_OK;
  }
  uint32_t* off = g->b.slot;
  return two_pass(*g, nc, off, n_pairs, s, [&](int pass, uint32_t* cnt) {
    hipLaunchKernelGGL(gr_cell_pairs, dim3(blocks_of(nc)), dim3(GR_BLOCK), 0, s, g->view, pass, cnt, off, d_pairs2, cap);
  });
}

int mg_grid_neighbor_lists(mg_grid* g, const float* d_xy, float h, uint32_t* d_offsets, uint32_t* d_nbrs, uint64_t cap,
                           uint64_t* n_edges, void* stream) {
  if (!g || !g->built || !d_offsets || !n_edges || (cap && !d_nbrs) || (g->n && !d_xy)) return MG_E_ARG;
  hipStream_t s = (hipStream_t)stream;
  const uint32_t n = g->n;
  const float h2 = h * h;
  if (n == 0) {
    hipLaunchKernelGGL(gr_set_u32, dim3(1), dim3(64), 0, s, d_offsets, 0u);
    *n_edges = 0;
    return hipStreamSynchronize(s) == hipSuccess ? MG_OK : MG_E_DEVICE;
  }
  const float2* xy = reinterpret_cast<const float2*>(d_xy);
  return two_pass(*g, n, d_offsets, n_edges, s, [&](int pass, uint32_t* cnt) {
    hipLaunchKernelGGL(gr_neighbors, dim3(blocks_of(n)), dim3(GR_BLOCK), 0, s, g->view, xy, n, h2, pass, cnt, d_offsets,
                       d_nbrs, cap);
  });
}

int mg_grid_near(mg_grid* g, const float* d_queries, uint32_t nq, uint32_t* d_offsets, uint32_t* d_out, uint64_t cap,
                 uint64_t* n_out, void* stream) {
  if (!g || !g->built || !d_offsets || !n_out || (cap && !d_out) || (nq && !d_queries)) return MG_E_ARG;
  if ((uint64_t)nq >= (uint64_t)GR_SCAN_TILES * GR_TILE) {
    g->err = "too many query points (at most 2^25 - 1 per call)";
    return MG_E_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  if (nq == 0) {
    hipLaunchKernelGGL(gr_set_u32, dim3(1), dim3(64), 0, s, d_offsets, 0u);
    *n_out = 0;
    return hipStreamSynchronize(s) == hipSuccess ? MG_OK : MG_E_DEVICE;
  }
  const float2* q = reinterpret_cast<const float2*>(d_queries);
  return two_pass(*g, nq, d_offsets, n_out, s, [&](int pass, uint32_t* cnt) {
    hipLaunchKernelGGL(gr_near, dim3(blocks_of(nq)), dim3(GR_BLOCK), 0, s, g->view, q, nq, pass, cnt, d_offsets, d_out, cap);
  });
}

const char* mg_grid_last_error(const mg_grid* g) { return g ? g->err.c_str() : "null grid"; }

// Device memory of this library's HIP runtime, for callers that hold no HIP runtime of their own
// (a process may load another HIP runtime, e.g. one bundled with a framework; pointers are not
// shared with it).
int mg_device_malloc(size_t bytes, void** out) {
  if (!out) return MG_E_ARG;
  *out = nullptr;
  if (bytes == 0) return MG_OK;
  return hipMalloc(out, bytes) == hipSuccess ? MG_OK : MG_E_DEVICE;
}
int mg_device_free(void* p) { return (!p || hipFree(p) == hipSuccess) ? MG_OK : MG_E_DEVICE; }
int mg_memcpy_htod(void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return MG_OK;
  if (!dst || !src) return MG_E_ARG;
  return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess ? MG_OK : MG_E_DEVICE;
}
int mg_memcpy_dtoh(void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return MG_OK;
  if (!dst || !src) return MG_E_ARG;
  return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) == hipSuccess ? MG_OK : MG_E_DEVICE;
}
int mg_device_synchronize(void) { return hipDeviceSynchronize() == hipSuccess ? MG_OK : MG_E_DEVICE; }

}  // extern "C"

// ---- the build that allocates nothing ------------------------------------------------------------
// mg_grid_build_noalloc keeps the scratch of a build (control block, tile counts and offsets, the
// radix sort's key/index ping-pong buffers and digit histograms, the sparse boundary flags and
// slots, the scan's tile sums) in one file-scope device array of fixed size, sized for
// GR_NOALLOC_MAX points; the caller passes the points and the output buffers. Every call on a GPU
// uses that GPU's copy of the array: calls that overlap on one GPU race.
#define GR_NOALLOC_MAX (1u << 20)

static __device__ __align__(256) unsigned char scratch[
    // ctrl, tile_cnt, tile_off, idxA, idxB, keyA, keyB, hist, dtot, flag, slot, sums, sums_off
    256 + 2 * (((GR_NOALLOC_MAX / GR_TILE + 3) * 4 + 255) / 256 * 256) +
    2 * (((GR_NOALLOC_MAX + 1) * 4 + 255) / 256 * 256) + 2 * (((GR_NOALLOC_MAX + 1) * 8 + 255) / 256 * 256) +
    (((size_t)GR_BINS * (GR_NOALLOC_MAX / GR_TILE + 2) * 4 + 255) / 256 * 256) + GR_BINS * 4 +
    2 * (((GR_NOALLOC_MAX + 2) * 4 + 255) / 256 * 256) + 2 * (((GR_SCAN_TILES + 2) * 4 + 255) / 256 * 256)];

namespace {

// Carves the scratch buffers of a build over up to cap points from base; returns the bytes used.
size_t noalloc_layout(GrBufs& b, uint32_t cap, char* base) {
  const uint64_t nt = tiles_of(cap) + 1;
  size_t off = 0;
  off = carve(off, b.ctrl, 1, base);
  off = carve(off, b.tile_cnt, nt + 1, base);
  off = carve(off, b.tile_off, nt + 1, base);
  off = carve(off, b.idxA, cap + 1, base);
  off = carve(off, b.idxB, cap + 1, base);
  uint64_t* ka;
  uint64_t* kb;
  off = carve(off, ka, cap + 1, base);
  off = carve(off, kb, cap + 1, base);
  b.keyA = ka;
  b.keyB = kb;
  off = carve(off, b.hist, (size_t)GR_BINS * nt, base);
  off = carve(off, b.dtot, GR_BINS, base);
  off = carve(off, b.flag, (size_t)cap + 2, base);
  off = carve(off, b.slot, (size_t)cap + 2, base);
  off = carve(off, b.sums, GR_SCAN_TILES + 2, base);
  off = carve(off, b.sums_off, GR_SCAN_TILES + 2, base);
  return off;
}

// Device address of `scratch` on each device, looked up once per device with that device current.
std::mutex g_scratch_mu;
void* g_scratch_ptr[64] = {};

int scratch_of_current_device(void** out, std::string& err) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
    err = "no current HIP device";
    return MG_E_DEVICE;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
      err = "hipGetSymbolAddress failed for the grid scratch array";
      return MG_E_DEVICE;
    }
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return MG_OK;
}

}  // namespace

extern "C" {

int mg_grid_build_noalloc(const float* d_xy, uint32_t n, float cell_size, int64_t* meta8, uint32_t* d_cell_start,
                          uint64_t cs_cap, uint32_t* d_indices, uint64_t idx_cap, int32_t* d_keys2, uint64_t keys_cap,
                          void* stream, char* err, size_t errlen) {
  if (!meta8 || (n && !d_xy) || !d_cell_start || (idx_cap && !d_indices) || (keys_cap && !d_keys2)) {
    set_err(err, errlen, "null argument");
    return MG_E_ARG;
  }
  if (n > GR_NOALLOC_MAX) {
    set_err(err, errlen, "mg_grid_build_noalloc takes at most 1048576 points (its scratch array is fixed)");
    return MG_E_ARG;
  }
  std::string e;
  void* base = nullptr;
  int rc = scratch_of_current_device(&base, e);
  if (rc != MG_OK) {
    set_err(err, errlen, e);
    return rc;
  }
  mg_grid g;  // host-side summary only: no device memory of its own
  if (noalloc_layout(g.b, GR_NOALLOC_MAX, static_cast<char*>(base)) > sizeof(scratch)) {
    set_err(err, errlen, "grid scratch layout exceeds the scratch array");
    return MG_E_ARG;
  }
  g.b.cell_start = d_cell_start;
  g.b.indices = d_indices;
  g.b.keys = reinterpret_cast<int2*>(d_keys2);
  OutCaps caps;
  caps.cs = cs_cap;
  caps.idx = idx_cap;
  caps.keys = keys_cap / 2;
  rc = build_into(g, g.b, reinterpret_cast<const float2*>(d_xy), n, cell_size, (hipStream_t)stream, caps);
  if (rc != MG_OK) {
    set_err(err, errlen, g.err);
    return rc;
  }
  if (hipStreamSynchronize((hipStream_t)stream) != hipSuccess) {
    set_err(err, errlen, "HIP error in the grid build");
    return MG_E_DEVICE;
  }
  return mg_grid_meta(&g, meta8);
}

}  // extern "C"
