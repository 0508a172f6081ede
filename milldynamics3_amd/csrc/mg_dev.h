// Device helpers for the per-drum substep kernel: 2D vector arithmetic with the reference's
// evaluation order, workgroup reductions and scans with a fixed (deterministic) combination
// order, and the uniform-grid build/query used by every neighbour stage.
//
// Every helper is always inlined (the stage kernels make no device-function calls and keep no
// per-thread arrays), and the file is compiled with -ffp-contract=off so a*b+c is never fused
// (the reference's Rust f32 code never fuses either); f32 division and sqrt are IEEE-rounded.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mg_layout.h"

#define DEV __device__ __attribute__((always_inline)) inline

// ---------------------------------------------------------------------------------------------
// f32 2D vector operations (glam::Vec2 semantics: component-wise, left to right)
// ---------------------------------------------------------------------------------------------
DEV float2 v2(float x, float y) { return make_float2(x, y); }
DEV float2 vadd(float2 a, float2 b) { return v2(a.x + b.x, a.y + b.y); }
DEV float2 vsub(float2 a, float2 b) { return v2(a.x - b.x, a.y - b.y); }
DEV float2 vneg(float2 a) { return v2(-a.x, -a.y); }
DEV float2 vmul(float2 a, float s) { return v2(a.x * s, a.y * s); }   // Vec2 * f32
DEV float2 smul(float s, float2 a) { return v2(s * a.x, s * a.y); }   // f32 * Vec2
DEV float2 vdiv(float2 a, float s) { return v2(a.x / s, a.y / s); }
DEV float vdot(float2 a, float2 b) { return a.x * b.x + a.y * b.y; }
DEV float vlen2(float2 a) { return a.x * a.x + a.y * a.y; }
DEV float vlen(float2 a) { return sqrtf(vlen2(a)); }
DEV bool vfinite(float2 a) { return isfinite(a.x) && isfinite(a.y); }
DEV float cross2(float2 a, float2 b) { return a.x * b.y - a.y * b.x; }
DEV float2 vperp(float2 l) { return v2(-l.y, l.x); }
DEV float2 vrot(float2 v, float ux, float uy) { return v2(v.x * ux - v.y * uy, v.x * uy + v.y * ux); }
DEV float2 vnormalize(float2 a) {
  float rc = 1.0f / vlen(a);
  return v2(a.x * rc, a.y * rc);
}
DEV float2 vnormalize_or_zero(float2 a) {
  float rc = 1.0f / vlen(a);
  if (isfinite(rc) && rc > 0.0f) return v2(a.x * rc, a.y * rc);
  return v2(0.0f, 0.0f);
}
DEV float fmaxnum(float a, float b) { return fmaxf(a, b); }  // f32::max (NaN-ignoring)
DEV float fminnum(float a, float b) { return fminf(a, b); }
DEV float fclamp(float x, float lo, float hi) {  // f32::clamp
  if (x < lo) x = lo;
  if (x > hi) x = hi;
  return x;
}
// `as i32` saturating cast.
DEV int32_t sat_i32(float x) {
  if (isnan(x)) return 0;
  if (x <= -2147483648.0f) return INT32_MIN;
  if (x >= 2147483647.0f) return INT32_MAX;
  return (int32_t)x;
}
DEV uint32_t sat_u32(float x) {
  if (isnan(x) || x <= 0.0f) return 0u;
  if (x >= 4294967295.0f) return 0xFFFFFFFFu;
  return (uint32_t)x;
}
DEV int32_t sadd_i32(int32_t a, int32_t b) {
  int64_t r = (int64_t)a + b;
  return r > INT32_MAX ? INT32_MAX : (r < INT32_MIN ? INT32_MIN : (int32_t)r);
}
// f32::powi: binary exponentiation, low exponent bit first (what LLVM emits for powi).
DEV float powi3(float x) { return x * (x * x); }
DEV float powi2(float x) { return x * x; }

// Kernels (pbf.rs:119-181) with precomputed leading constants (same f32 value as the reference's
// per-call expression, evaluated once on the host).
DEV float poly6_k(float r2, float h, float k) {
  if (r2 >= h * h || r2 < 0.0f) return 0.0f;
  float h2 = h * h, t = h2 - r2;
  return k * t * t * t;
}
DEV float2 spiky_grad_k(float2 d, float r, float h, float k) {
  if (r <= 1e-9f || r >= h) return v2(0.0f, 0.0f);
  float coeff = k * (h - r) * (h - r);
  return vmul(vdiv(d, r), coeff);
}
DEV float cohesion_kernel_k(float r, float h, float k, float h6_64) {
  if (r <= 0.0f || r >= h) return 0.0f;
  float term = powi3(h - r) * powi3(r);
  if (r <= 0.5f * h) return k * (2.0f * term - h6_64);
  return k * term;
}
DEV float adhesion_kernel_k(float r, float h, float k) {
  if (r <= 0.5f * h || r >= h) return 0.0f;
  return k * powi3(h - r) * powi3(r);
}

// ---------------------------------------------------------------------------------------------
// Workgroup reductions (fixed shuffle order => run-to-run deterministic)
// ---------------------------------------------------------------------------------------------
DEV int tid() { return (int)threadIdx.x; }
DEV int nthr() { return (int)blockDim.x; }
DEV void bsync() { __syncthreads(); }

DEV double block_sum_d(double v) {
  __shared__ double s_part[MG_BLOCK / 64];
  __shared__ double s_out;
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  int w = tid() >> 6, l = tid() & 63;
  if (l == 0) s_part[w] = v;
  bsync();
  if (tid() < 64) {
    double r = tid() < nthr() / 64 ? s_part[tid()] : 0.0;
    for (int o = 32; o > 0; o >>= 1) r += __shfl_down(r, o, 64);
    if (tid() == 0) s_out = r;
  }
  bsync();
  double r = s_out;
  bsync();
  return r;
}
DEV uint32_t block_sum_u(uint32_t v) {
  __shared__ uint32_t s_part[MG_BLOCK / 64];
  __shared__ uint32_t s_out;
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  int w = tid() >> 6, l = tid() & 63;
  if (l == 0) s_part[w] = v;
  bsync();
  if (tid() < 64) {
    uint32_t r = tid() < nthr() / 64 ? s_part[tid()] : 0u;
    for (int o = 32; o > 0; o >>= 1) r += __shfl_down(r, o, 64);
    if (tid() == 0) s_out = r;
  }
  bsync();
  uint32_t r = s_out;
  bsync();
  return r;
}
// f32::max fold from 0.0 (exact regardless of order).
DEV float block_max_f(float v) {
  __shared__ float s_part[MG_BLOCK / 64];
  __shared__ float s_out;
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_down(v, o, 64));
  int w = tid() >> 6, l = tid() & 63;
  if (l == 0) s_part[w] = v;
  bsync();
  if (tid() < 64) {
    float r = tid() < nthr() / 64 ? s_part[tid()] : 0.0f;
    for (int o = 32; o > 0; o >>= 1) r = fmaxf(r, __shfl_down(r, o, 64));
    if (tid() == 0) s_out = r;
  }
  bsync();
  float r = s_out;
  bsync();
  return r;
}

// In-place exclusive prefix sum over a[0..n) (global memory), one contiguous chunk per thread.
DEV void block_scan_excl(uint32_t* a, int n) {
  __shared__ uint32_t s_part[MG_BLOCK];
  int T = nthr();
  int chunk = (n + T - 1) / T;
  int b = tid() * chunk, e = min(n, b + chunk);
  uint32_t sum = 0;
  for (int i = b; i < e; i++) sum += a[i];
  s_part[tid()] = sum;
  bsync();
  for (int off = 1; off < T; off <<= 1) {
    uint32_t v = tid() >= off ? s_part[tid() - off] : 0u;
    bsync();
    s_part[tid()] += v;
    bsync();
  }
  uint32_t base = tid() ? s_part[tid() - 1] : 0u;
  for (int i = b; i < e; i++) {
    uint32_t t = a[i];
    a[i] = base;
    base += t;
  }
  bsync();
}

// Groups record indices 0..nrec-1 by key[r] in 0..K-1 and, within one key, ascending by ord[r].
// Output: seg[0..K] offsets, perm[0..nrec) record indices. cursor: K scratch entries.
DEV void ordered_csr(int nrec, const uint32_t* key, const uint32_t* ord, int K, uint32_t* seg,
                                   uint32_t* perm, uint32_t* cursor) {
  for (int k = tid(); k <= K; k += nthr()) seg[k] = 0;
  bsync();
  for (int r = tid(); r < nrec; r += nthr())
    if (key[r] < (uint32_t)K) atomicAdd(&seg[key[r]], 1u);
  bsync();
  block_scan_excl(seg, K + 1);
  for (int k = tid(); k < K; k += nthr()) cursor[k] = seg[k];
  bsync();
  for (int r = tid(); r < nrec; r += nthr()) {
    if (key[r] >= (uint32_t)K) continue;
    uint32_t p = atomicAdd(&cursor[key[r]], 1u);
    perm[p] = (uint32_t)r;
  }
  bsync();
  for (int k = tid(); k < K; k += nthr()) {
    uint32_t s = seg[k], e = seg[k + 1];
    for (uint32_t a = s + 1; a < e; a++) {
      uint32_t v = perm[a], ov = ord[v];
      uint32_t b = a;
      while (b > s && ord[perm[b - 1]] > ov) {
        perm[b] = perm[b - 1];
        b--;
      }
      perm[b] = v;
    }
  }
  bsync();
}

// ---------------------------------------------------------------------------------------------
// Uniform grid (grid.rs:101-399): counting-sort CSR, Dense layout when the bounding box has at
// most 2n + 512 cells, otherwise the Sparse layout (occupied cells only, lexicographic keys).
// ---------------------------------------------------------------------------------------------


DEV uint64_t pack_key(int32_t cx, int32_t cy) {
  return ((uint64_t)((uint32_t)cx ^ 0x80000000u) << 32) | (uint64_t)((uint32_t)cy ^ 0x80000000u);
}
DEV int32_t key_cx(uint64_t k) { return (int32_t)((uint32_t)(k >> 32) ^ 0x80000000u); }
DEV int32_t key_cy(uint64_t k) { return (int32_t)((uint32_t)k ^ 0x80000000u); }

// Returns the [s, e) range of cell (cx, cy); false if out of range or empty.
DEV bool grid_range(const GridView& g, int32_t cx, int32_t cy, uint32_t& s, uint32_t& e) {
  if (g.ncells == 0) return false;
  if (g.dense) {
    int64_t dx = (int64_t)cx - g.min_cx, dy = (int64_t)cy - g.min_cy;
    if (dx < 0 || dy < 0 || dx >= g.W || dy >= g.H) return false;
    int64_t k = dx * g.H + dy;
    s = g.start[k];
    e = g.start[k + 1];
    return s != e;
  }
  uint64_t key = pack_key(cx, cy);
  int lo = 0, hi = g.ncells;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (g.skeys[mid] < key) lo = mid + 1; else hi = mid;
  }
  if (lo >= g.ncells || g.skeys[lo] != key) return false;
  s = g.start[lo];
  e = g.start[lo + 1];
  return true;
}

DEV void cell_of(float2 p, float c, int32_t& cx, int32_t& cy) {  // grid.rs:280-285
  cx = sat_i32(floorf(p.x / c));
  cy = sat_i32(floorf(p.y / c));
}

// for_each_near (grid.rs:371-399): the up-to-9 distinct cells around a point, visited dy-major
// (dy = -1..1 outer, dx = -1..1 inner). Cell q (0..8) is offset (q % 3 - 1, q / 3 - 1); a cell
// that saturates onto an earlier one at the i32 boundary is skipped, as the reference's dedupe
// does. No per-thread arrays: callers loop q = 0..8 and call near_cell for each.
DEV bool near_dup(int32_t cx, int32_t cy, int q) {
  int dx = q % 3 - 1, dy = q / 3 - 1;
  int32_t ax = sadd_i32(cx, dx), ay = sadd_i32(cy, dy);
  for (int p = 0; p < q; p++) {
    int ex = p % 3 - 1, ey = p / 3 - 1;
    if (sadd_i32(cx, ex) == ax && sadd_i32(cy, ey) == ay) return true;
  }
  return false;
}
DEV bool near_cell(const GridView& g, int32_t cx, int32_t cy, int q, uint32_t& s, uint32_t& e) {
  if (near_dup(cx, cy, q)) return false;
  return grid_range(g, sadd_i32(cx, q % 3 - 1), sadd_i32(cy, q / 3 - 1), s, e);
}

struct GridScratch {
  uint32_t* start;   // >= cap_cells + 1 entries (dense) / n + 1 (sparse)
  uint32_t* idx;     // n entries
  uint32_t* key;     // n entries (linear cell index per point, or UINT32_MAX)
  int32_t* meta;     // 8 ints persisted for later views
  uint32_t* cursor;  // >= max(cap_cells, n) + 1 entries
  uint64_t* skey;    // sparse sort keys (pow2 >= n)
  uint32_t* sidx;    // sparse sort payload (pow2 >= n)
  int32_t cap_cells;
  int32_t pow2;
};

// Block-wide grid build over pts[0..n). Every thread must call it. Returns the view.
DEV GridView grid_build(const float2* pts, int n, float cell, const GridScratch& G, uint32_t* err) {
  __shared__ int32_t s_mnx, s_mny, s_mxx, s_mxy;
  __shared__ uint32_t s_cnt;
  if (tid() == 0) { s_mnx = INT32_MAX; s_mny = INT32_MAX; s_mxx = INT32_MIN; s_mxy = INT32_MIN; s_cnt = 0; }
  bsync();
  {
    int32_t mnx = INT32_MAX, mny = INT32_MAX, mxx = INT32_MIN, mxy = INT32_MIN;
    uint32_t cnt = 0;
    for (int i = tid(); i < n; i += nthr()) {
      float2 p = pts[i];
      if (!vfinite(p)) continue;
      int32_t cx, cy;
      cell_of(p, cell, cx, cy);
      mnx = min(mnx, cx); mny = min(mny, cy); mxx = max(mxx, cx); mxy = max(mxy, cy);
      cnt++;
    }
    if (cnt) {
      atomicMin(&s_mnx, mnx); atomicMin(&s_mny, mny); atomicMax(&s_mxx, mxx); atomicMax(&s_mxy, mxy);
      atomicAdd(&s_cnt, cnt);
    }
  }
  bsync();
  GridView g;
  g.start = G.start; g.idx = G.idx; g.skeys = nullptr; g.cell = cell;
  uint32_t cnt = s_cnt;
  if (cnt == 0) {
    g.min_cx = 0; g.min_cy = 0; g.W = 0; g.H = 0; g.ncells = 0; g.dense = 0;
    bsync();
    return g;
  }
  int64_t W = (int64_t)s_mxx - (int64_t)s_mnx + 1;
  int64_t H = (int64_t)s_mxy - (int64_t)s_mny + 1;
  unsigned __int128 tc = (unsigned __int128)(uint64_t)W * (uint64_t)H;
  uint64_t budget = 2ull * cnt + 512ull;
  bool dense = tc <= (unsigned __int128)budget;
  g.min_cx = s_mnx; g.min_cy = s_mny; g.W = W; g.H = H;
  bsync();
  if (dense && (int64_t)(W * H) <= (int64_t)G.cap_cells) {
    int32_t nc = (int32_t)(W * H);
    g.ncells = nc; g.dense = 1;
    for (int k = tid(); k <= nc; k += nthr()) G.start[k] = 0;
    bsync();
    for (int i = tid(); i < n; i += nthr()) {
      float2 p = pts[i];
      uint32_t k = 0xFFFFFFFFu;
      if (vfinite(p)) {
        int32_t cx, cy;
        cell_of(p, cell, cx, cy);
        k = (uint32_t)(((int64_t)cx - g.min_cx) * H + ((int64_t)cy - g.min_cy));
        atomicAdd(&G.start[k], 1u);
      }
      G.key[i] = k;
    }
    bsync();
    block_scan_excl(G.start, nc + 1);
    for (int k = tid(); k < nc; k += nthr()) G.cursor[k] = G.start[k];
    bsync();
    for (int i = tid(); i < n; i += nthr()) {
      uint32_t k = G.key[i];
      if (k != 0xFFFFFFFFu) G.idx[atomicAdd(&G.cursor[k], 1u)] = (uint32_t)i;
    }
    bsync();
    // Stable order inside each cell: ascending point index (insertion sort, cells hold few points).
    for (int k = tid(); k < nc; k += nthr()) {
      uint32_t s = G.start[k], e = G.start[k + 1];
      for (uint32_t a = s + 1; a < e; a++) {
        uint32_t v = G.idx[a], b = a;
        while (b > s && G.idx[b - 1] > v) { G.idx[b] = G.idx[b - 1]; b--; }
        G.idx[b] = v;
      }
    }
  } else if (dense) {
    // Dense by the reference's rule but larger than this slot's cell capacity.
    if (tid() == 0) atomicOr(err, MG_ERR_SPARSE);
    g.ncells = 0; g.dense = 0;
  } else {
    // Sparse layout (grid.rs:204-226): stable sort (key, index), then group.
    int P = G.pow2;
    if ((int)cnt > P) {
      if (tid() == 0) atomicOr(err, MG_ERR_SPARSE);
      g.ncells = 0; g.dense = 0;
      bsync();
      return g;
    }
    __shared__ uint32_t s_fill;
    if (tid() == 0) s_fill = 0;
    bsync();
    // Compact finite points in index order (ballot prefix per wave keeps the order stable).
    for (int base = 0; base < n; base += nthr()) {
      int i = base + tid();
      bool ok = i < n && vfinite(pts[i]);
      __shared__ uint32_t s_wcnt[MG_BLOCK / 64];
      uint64_t m = __ballot(ok);
      int lane = tid() & 63, w = tid() >> 6;
      uint32_t before = __popcll(m & ((1ull << lane) - 1ull));
      if (lane == 0) s_wcnt[w] = __popcll(m);
      bsync();
      uint32_t woff = 0;
      for (int q = 0; q < w; q++) woff += s_wcnt[q];
      if (ok) {
        int32_t cx, cy;
        cell_of(pts[i], cell, cx, cy);
        uint32_t pos = s_fill + woff + before;
        G.skey[pos] = pack_key(cx, cy);
        G.sidx[pos] = (uint32_t)i;
      }
      bsync();
      if (tid() == 0) { uint32_t t = 0; for (int q = 0; q < nthr() / 64; q++) t += s_wcnt[q]; s_fill += t; }
      bsync();
    }
    for (int i = (int)cnt + tid(); i < P; i += nthr()) { G.skey[i] = ~0ull; G.sidx[i] = 0xFFFFFFFFu; }
    bsync();
    // Bitonic sort by (key, index): equal keys keep ascending index, i.e. a stable sort.
    for (int k = 2; k <= P; k <<= 1)
      for (int j = k >> 1; j > 0; j >>= 1) {
        for (int i = tid(); i < P; i += nthr()) {
          int l = i ^ j;
          if (l > i) {
            uint64_t ka = G.skey[i], kb = G.skey[l];
            uint32_t ia = G.sidx[i], ib = G.sidx[l];
            bool gt = ka > kb || (ka == kb && ia > ib);
            bool up = (i & k) == 0;
            if (gt == up) { G.skey[i] = kb; G.skey[l] = ka; G.sidx[i] = ib; G.sidx[l] = ia; }
          }
        }
        bsync();
      }
    // idx = payload order; unique keys -> start offsets.
    for (int i = tid(); i < (int)cnt; i += nthr()) {
      G.idx[i] = G.sidx[i];
      G.cursor[i] = (i == 0 || G.skey[i] != G.skey[i - 1]) ? 1u : 0u;
    }
    bsync();
    if (tid() == 0) {
      // Sequential compaction of the unique keys (fallback path only).
      uint32_t u = 0;
      for (uint32_t i = 0; i < cnt; i++) {
        if (G.cursor[i]) { G.skey[u] = G.skey[i]; G.start[u] = i; u++; }
      }
      G.start[u] = cnt;
      G.meta[7] = (int32_t)u;
    }
    bsync();
    g.ncells = G.meta[7];
    g.dense = 0;
    g.skeys = G.skey;
  }
  if (tid() == 0) {
    G.meta[0] = g.min_cx; G.meta[1] = g.min_cy; G.meta[2] = (int32_t)g.W; G.meta[3] = (int32_t)g.H;
    G.meta[4] = g.dense; G.meta[5] = g.ncells;
  }
  bsync();
  return g;
}
