// FluidParticles::step_coupled on the GPU for a batch of drums (crates/mill-core/src/pbf.rs:1174-2365).
//
// The reference advances a drum's slurry in this order: sanitize, a grid over the balls, predict
// (gravity), a grid over the predicted positions and the mutual neighbour lists, the density solve
// (fused density/lambda pass, position-correction pass, x += dp; adaptive pass count), the
// ball-overlap push with its momentum-bounded reaction, Akinci cohesion over the neighbour lists,
// the per-ball adhesion pull, the wall projection with its f64 wall impulse and 32 angle bins,
// velocities from positions, the wall no-slip blend, per-ball form drag and buoyancy (each ball
// updating the fluid velocities before the next ball reads them), the density at the final
// positions, Bingham viscosities, Morris weights, ball boundary groups and squeeze links, the
// implicit viscosity CG (fluid only, or jointly with the balls' linear and angular velocities,
// Jacobi-preconditioned), the mean shear rate, the fluid speed clamp and the per-ball impulse clamp.
//
// Mapping. Particle-parallel stages run over tiles of 256 particles of one drum; many drums' tiles
// run side by side in one launch, so a batch fills the GPU and a single drum spreads over as many
// workgroups as it has tiles. Grids are built per drum by one workgroup (keys, bounding box,
// counting sort, each cell's points in ascending index). Stages whose per-ball results depend on
// the order of candidate visits (adhesion items, boundary groups) run one wave per ball, the
// wave's lanes evaluating 64 candidates at a time and the order-sensitive merges walked in
// candidate order from the lanes' values (readlane). The Gauss-Seidel stages (form drag and
// buoyancy read the fluid velocities the previous ball wrote) walk the balls in index order on one
// wave per drum. Sums the reference forms over a foreign index (ball impulses from fluid contacts,
// fluid kicks from several balls) are gathered per target in the reference's order through a
// per-drum counting sort, never with floating-point atomics, so results are deterministic and the
// f32 arithmetic up to the viscosity solve equals the reference's bit for bit. The CG's f64 dot
// products are summed per tile and then over the tiles in a fixed order (the reference sums
// sequentially), so CG results agree to rounding.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/millgpu.h"
#include "mg_atan2f.h"
#include "mg_drumgeo.h"
#include "mg_expf.h"
#include "mg_ktimer.h"

#define FL_T 256       // particles (or CG rows) per tile, lanes per drum kernel
#define FL_PK 8        // ball-push records per particle
#define FL_NO 0xFFFFFFFFu
#define FL_NPASS 32    // density passes a step can run
#define FL_SLOTS 4     // merged items a lane holds: 256 items per ball
#define FL_ITEMS 1024  // form-drag items per ball (LDS)
#define FL_NP 10       // f64 partial sums per fluid tile

namespace {

constexpr float FL_PI = 3.14159265358979323846f;
constexpr float FL_TAU = 6.28318530717958647692f;
constexpr float FL_G = -9.81f;

enum : uint32_t {
  FE_CELLS = 1u,    // a grid's bounding box exceeded the drum's cell capacity
  FE_EDGES = 2u,    // more neighbour-list entries than the drum's capacity
  FE_PUSH = 4u,     // a particle touched more than FL_PK balls
  FE_ADH = 8u,      // a ball's adhesion items exceeded the per-ball capacity
  FE_GROUPS = 16u,  // a ball's boundary groups exceeded the per-ball capacity
  FE_LINKS = 32u,   // a ball's squeeze links exceeded the per-ball capacity
  FE_ITEMS = 64u,   // a ball's form-drag items exceeded FL_ITEMS
};

// f64 partial-sum slots of a fluid tile.
enum { P_KE0 = 0, P_KE1, P_WW, P_WIX, P_WIY, P_WANG, P_KEV, P_KEC, P_SH, P_SHN };

// One drum's scalars for one step (host-prepared, f32 values evaluated as the reference does).
struct FlDrum {
  // drum geometry at angle_next = angle + omega dt (pbf.rs:1236)
  float radius, omega;
  uint32_t nl, lift0;
  float rt, bh, thw;
  float rc1, rs1, wc1, ws1;  // from_angle(-angle_next), from_angle(angle_next)
  float dt;
  uint32_t base, force, forced_d, forced_cg, npass;
  // fluid (pbf.rs:1018-1028)
  uint32_t f0, n;  // first particle in the caller's arrays; count (0: the step does nothing)
  float h, rest, mass, h2, mr;  // h2 = h*h, mr = mass / rest
  float p6c, spc, cohc, h6_64, poly0m;  // kernel constants (pbf.rs:119-160)
  float coh_acc;                        // cohesion acceleration (pbf.rs:1466)
  float ha, adc, ad_range2, ad_inner2, ad_acc;  // adhesion (pbf.rs:1517-1530)
  float beta_w, mu, eta2;
  uint32_t bingham;
  float tau_y, g0;
  // balls (dem.rs:277-309)
  uint32_t b0, nb;
  float br, bmass, binertia, btrue;
  float contact_r, bgc, red;  // overlap push (pbf.rs:1214-1222, :1368)
  uint32_t cpb, off0;         // boundary points per ball and their offset table
  uint32_t adh_on, grp_on, drag_on, buoy_on, links_on;
  // form drag (pbf.rs:1777-1812)
  float beta_b, hc, hc2, w0, p6hc, mass_true, rot_cap, area;
  // squeeze links (pbf.rs:544-625)
  float reach, c_max, rng2, g5c, g_range2;
  double dxd, a2, sqp[4], sqw[4];  // (r, c1, c2, pre) of the pair and wall films
  // clamps
  float vmax, imp_cl, ang_cl;
  // workspace regions of this drum
  uint32_t e0, ecap;            // neighbour entries
  uint32_t c0[5], ccap[5];      // grid cells (0: fluid h predicted, 1: balls push, 2: fluid hc, 3: fluid h, 4: balls links)
  uint32_t l0, lcap;            // squeeze links (per ball lcap slots from l0 + b * lcap)
  uint32_t t0, nt, r0, nr;      // fluid tiles, CG row tiles
  uint32_t has_guess;
};

// A grid's layout (grid.rs:101-197): bounding box of the cell keys and the cell size; W = H = 0
// for an empty grid.
struct FlGrid {
  int32_t mnx, mny;
  uint32_t W, H;
  float cell;
  uint32_t pad[3];
};

// Per drum state and results of a step.
struct FlState {
  uint32_t err;
  uint32_t maxc[FL_NPASS];  // the largest constraint of each density pass (f32 bits)
  uint32_t mode;            // viscosity solve: 0 none, 1 fluid only, 2 joint with the balls
  uint32_t done, it, iters, maxit;
  uint32_t n_groups, n_links, in_range, clamp_hits;
  double rs, target;
  float alpha, beta;
  double ke_drag, ke_buoy;  // sum |v|^2 after form drag / buoyancy
  uint32_t gs_levels;       // levels of the form-drag and buoyancy schedule
};

// ---- small helpers ----------------------------------------------------------------------------
__host__ __device__ inline bool fin2(float2 p) { return __builtin_isfinite(p.x) && __builtin_isfinite(p.y); }

// `x as i32` (saturating, NaN -> 0).
__device__ inline int32_t sat_i32(float x) {
  if (__builtin_isnan(x)) return 0;
  if (x <= -2147483648.0f) return INT32_MIN;
  if (x >= 2147483647.0f) return INT32_MAX;
  return (int32_t)x;
}
__device__ inline int32_t sadd(int32_t a, int32_t b) {
  const int64_t r = (int64_t)a + b;
  return r > INT32_MAX ? INT32_MAX : (r < INT32_MIN ? INT32_MIN : (int32_t)r);
}
__device__ __forceinline__ float rdl(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ uint32_t rdlu(uint32_t v, int l) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, l);
}
__device__ __forceinline__ int first_lane(uint64_t m) { return __builtin_amdgcn_readfirstlane(__builtin_ctzll(m)); }
__device__ inline float fclampf(float x, float lo, float hi) {  // f32::clamp
  if (x < lo) x = lo;
  if (x > hi) x = hi;
  return x;
}

// poly6 (pbf.rs:119-126), spiky gradient (pbf.rs:128-135), cohesion (pbf.rs:137-150) and
// adhesion (pbf.rs:152-160) kernels with the coefficients prepared on the host.
__device__ inline float poly6(float r2, float h2, float c) {
  if (r2 >= h2 || r2 < 0.0f) return 0.0f;
  const float t = h2 - r2;
  return c * t * t * t;
}
__device__ inline float2 spiky(float2 d, float r, float h, float c) {
  if (r <= 1e-9f || r >= h) return make_float2(0.0f, 0.0f);
  const float k = c * (h - r) * (h - r);
  return make_float2(d.x / r * k, d.y / r * k);
}
__device__ inline float pw3(float x) { return x * (x * x); }  // f32::powi(x, 3)
__device__ inline float cohesion_k(float r, float h, float c, float h6_64) {
  if (r <= 0.0f || r >= h) return 0.0f;
  const float term = pw3(h - r) * pw3(r);
  if (r <= 0.5f * h) return c * (2.0f * term - h6_64);
  return c * term;
}
__device__ inline float adhesion_k(float r, float h, float c) {
  if (r <= 0.5f * h || r >= h) return 0.0f;
  return c * pw3(h - r) * pw3(r);
}

// ---- grids (grid.rs:101-399) ------------------------------------------------------------------
struct GView {
  int32_t mnx, mny;
  uint32_t W, H;
  float cell;
  const uint32_t* start;
  const uint32_t* idx;
};
__device__ inline GView gview(const FlGrid* g, const uint32_t* start, const uint32_t* idx) {
  GView v;
  v.mnx = g->mnx;
  v.mny = g->mny;
  v.W = g->W;
  v.H = g->H;
  v.cell = g->cell;
  v.start = start;
  v.idx = idx;
  return v;
}
__device__ inline void gkey(float2 p, float c, int32_t& kx, int32_t& ky) {  // grid.rs:280-285
  kx = sat_i32(floorf(p.x / c));
  ky = sat_i32(floorf(p.y / c));
}
// Points of cell (cx, cy): [s, e) of idx; false for an empty or absent cell (grid.rs:242-271).
__device__ inline bool grange(const GView& g, int32_t cx, int32_t cy, uint32_t& s, uint32_t& e) {
  const int64_t dx = (int64_t)cx - g.mnx, dy = (int64_t)cy - g.mny;
  if (dx < 0 || dy < 0 || dx >= (int64_t)g.W || dy >= (int64_t)g.H) return false;
  const uint32_t k = (uint32_t)(dx * (int64_t)g.H + dy);
  s = g.start[k];
  e = g.start[k + 1];
  return s != e;
}
// grange with its loads issued whatever the cell (an absent cell reads cell 0 and gives the empty
// range), so that the loads of several cells are in flight together.
__device__ inline void grange_any(const GView& g, int32_t cx, int32_t cy, uint32_t& s, uint32_t& e) {
  const int64_t dx = (int64_t)cx - g.mnx, dy = (int64_t)cy - g.mny;
  const bool in = dx >= 0 && dy >= 0 && dx < (int64_t)g.W && dy < (int64_t)g.H;
  const uint32_t k = in ? (uint32_t)(dx * (int64_t)g.H + dy) : 0u;
  const uint32_t a = g.start[k], b = g.start[k + 1];
  s = in ? a : 0u;
  e = in ? b : 0u;
}
// The cells for_each_near visits around p, in its order (rows dy = -1..1, dx = -1..1, repeats
// skipped): up to 9 ranges.
struct Near {
  uint32_t s[9], e[9];
  uint32_t total;
};
__device__ inline void near_cells(const GView& g, float2 p, Near& q) {
  q.total = 0;
#pragma unroll
  for (int r = 0; r < 9; r++) q.s[r] = q.e[r] = 0;
  if (!fin2(p) || g.W == 0) return;
  int32_t kx, ky;
  gkey(p, g.cell, kx, ky);
  int32_t ax[9], ay[9];
#pragma unroll
  for (int r = 0; r < 9; r++) {
    ax[r] = sadd(kx, r % 3 - 1);
    ay[r] = sadd(ky, r / 3 - 1);
    bool dup = false;
#pragma unroll
    for (int u = 0; u < r; u++) dup |= ax[u] == ax[r] && ay[u] == ay[r];
    uint32_t s = 0, e = 0;
    grange_any(g, ax[r], ay[r], s, e);
    q.s[r] = dup ? 0u : s;
    q.e[r] = dup ? 0u : e;
    q.total += q.e[r] - q.s[r];
  }
}
// near_cells for a point that is the same on every lane of the wave (every lane active): lane r < 9
// finds cell r's range, all nine in one round of loads, and every lane reads the nine.
__device__ inline void near_cells_wave(const GView& g, float2 p, Near& q) {
  const uint32_t lane = threadIdx.x & 63u;
  q.total = 0;
#pragma unroll
  for (int r = 0; r < 9; r++) q.s[r] = q.e[r] = 0;
  if (!fin2(p) || g.W == 0) return;
  int32_t kx, ky;
  gkey(p, g.cell, kx, ky);
  uint32_t s = 0, e = 0;
  if (lane < 9) {
    const int32_t r = (int32_t)lane;
    const int32_t ax = sadd(kx, r % 3 - 1), ay = sadd(ky, r / 3 - 1);
    bool dup = false;
#pragma unroll
    for (int32_t u = 0; u < 8; u++) dup |= u < r && sadd(kx, u % 3 - 1) == ax && sadd(ky, u / 3 - 1) == ay;
    uint32_t cs = 0, ce = 0;
    if (!dup && grange(g, ax, ay, cs, ce)) {
      s = cs;
      e = ce;
    }
  }
#pragma unroll
  for (int r = 0; r < 9; r++) {
    q.s[r] = rdlu(s, r);
    q.e[r] = rdlu(e, r);
    q.total += q.e[r] - q.s[r];
  }
}
// Candidate c (0-based, c < q.total) of a near query: its point index.
__device__ inline uint32_t near_at(const GView& g, const Near& q, uint32_t c) {
  uint32_t j = FL_NO;
#pragma unroll
  for (int r = 0; r < 9; r++) {
    const uint32_t len = q.e[r] - q.s[r];
    if (j == FL_NO) {
      if (c < len) j = g.idx[q.s[r] + c];
      else c -= len;
    }
  }
  return j;
}

// Workgroup exclusive scan of one value per lane (FL_T lanes); returns the prefix, total in *tot.
__device__ inline uint32_t wg_exscan(uint32_t v, uint32_t* lds, uint32_t* tot) {
  const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= (uint32_t)o) x += y;
  }
  if (lane == 63) lds[w] = x;
  __syncthreads();
  uint32_t before = 0, all = 0;
  for (uint32_t q = 0; q < FL_T / 64; q++) {
    const uint32_t s = lds[q];
    if (q < w) before += s;
    all += s;
  }
  __syncthreads();
  *tot = all;
  return before + x - v;
}

// Exclusive scan of cnt[0..n) into off[0..n] (off[n] = total) by one workgroup; cnt and off may
// be the same array. Each lane owns a contiguous run of ceil(n / FL_T) entries: the runs' sums
// are scanned once, then each lane writes its run's offsets. Returns the total.
__device__ inline uint32_t wg_scan_array(const uint32_t* cnt, uint32_t* off, uint32_t n, uint32_t* lds) {
  const uint32_t k = (n + FL_T - 1) / FL_T;
  const uint32_t a = min(n, threadIdx.x * k), e = min(n, a + k);
  uint32_t s = 0;
  for (uint32_t i = a; i < e; i++) s += cnt[i];
  uint32_t tot;
  uint32_t run = wg_exscan(s, lds, &tot);
  for (uint32_t i = a; i < e; i++) {
    const uint32_t c = cnt[i];
    off[i] = run;
    run += c;
  }
  if (threadIdx.x == 0) off[n] = tot;
  __syncthreads();
  return tot;
}

// UniformGrid::build (grid.rs:101-197) of points p[0..n) by one workgroup: finite points only,
// cells of the bounding box in x-major order, each cell's points in ascending index. A bounding
// box of more than cap cells leaves the grid empty and sets FE_CELLS.
__device__ void wg_grid_build(const float2* __restrict__ p, uint32_t n, float cell, uint32_t cap,
                              uint32_t* __restrict__ start, uint32_t* __restrict__ cur, uint32_t* __restrict__ idx,
                              FlGrid* __restrict__ g, uint32_t* __restrict__ err, int32_t* lds_i, uint32_t* lds_u) {
  const uint32_t t = threadIdx.x;
  int32_t mnx = INT32_MAX, mny = INT32_MAX, mxx = INT32_MIN, mxy = INT32_MIN;
  uint32_t cnt = 0;
  for (uint32_t i = t; i < n; i += FL_T) {
    const float2 q = p[i];
    if (!fin2(q)) continue;
    int32_t kx, ky;
    gkey(q, cell, kx, ky);
    mnx = min(mnx, kx);
    mny = min(mny, ky);
    mxx = max(mxx, kx);
    mxy = max(mxy, ky);
    cnt++;
  }
  lds_i[t] = mnx;
  lds_i[FL_T + t] = mny;
  lds_i[2 * FL_T + t] = mxx;
  lds_i[3 * FL_T + t] = mxy;
  lds_u[t] = cnt;
  __syncthreads();
  for (uint32_t s = FL_T / 2; s > 0; s >>= 1) {
    if (t < s) {
      lds_i[t] = min(lds_i[t], lds_i[t + s]);
      lds_i[FL_T + t] = min(lds_i[FL_T + t], lds_i[FL_T + t + s]);
      lds_i[2 * FL_T + t] = max(lds_i[2 * FL_T + t], lds_i[2 * FL_T + t + s]);
      lds_i[3 * FL_T + t] = max(lds_i[3 * FL_T + t], lds_i[3 * FL_T + t + s]);
      lds_u[t] += lds_u[t + s];
    }
    __syncthreads();
  }
  mnx = lds_i[0];
  mny = lds_i[FL_T];
  mxx = lds_i[2 * FL_T];
  mxy = lds_i[3 * FL_T];
  cnt = lds_u[0];
  __syncthreads();
  uint32_t W = 0, H = 0;
  if (cnt > 0) {
    const uint64_t w64 = (uint64_t)((int64_t)mxx - (int64_t)mnx + 1), h64 = (uint64_t)((int64_t)mxy - (int64_t)mny + 1);
    if (w64 * h64 <= (uint64_t)cap) {
      W = (uint32_t)w64;
      H = (uint32_t)h64;
    } else if (t == 0) {
      atomicOr(err, FE_CELLS);
    }
  }
  if (t == 0) {
    g->mnx = mnx;
    g->mny = mny;
    g->W = W;
    g->H = H;
    g->cell = cell;
  }
  const uint32_t T = W * H;
  for (uint32_t k = t; k < T; k += FL_T) cur[k] = 0;
  if (T == 0 && t == 0) start[0] = 0;
  __syncthreads();
  if (T == 0) return;
  for (uint32_t i = t; i < n; i += FL_T) {
    const float2 q = p[i];
    if (!fin2(q)) continue;
    int32_t kx, ky;
    gkey(q, cell, kx, ky);
    atomicAdd(&cur[(uint32_t)(kx - mnx) * H + (uint32_t)(ky - mny)], 1u);
  }
  __syncthreads();
  wg_scan_array(cur, start, T, lds_u);
  for (uint32_t k = t; k < T; k += FL_T) cur[k] = start[k];
  __syncthreads();
  for (uint32_t i = t; i < n; i += FL_T) {
    const float2 q = p[i];
    if (!fin2(q)) continue;
    int32_t kx, ky;
    gkey(q, cell, kx, ky);
    const uint32_t pos = atomicAdd(&cur[(uint32_t)(kx - mnx) * H + (uint32_t)(ky - mny)], 1u);
    idx[pos] = i;
  }
  __syncthreads();
  for (uint32_t k = t; k < T; k += FL_T) {  // ascending index inside each cell
    const uint32_t s = start[k], e = start[k + 1];
    for (uint32_t a = s + 1; a < e; a++) {
      const uint32_t v = idx[a];
      uint32_t b = a;
      while (b > s && idx[b - 1] > v) {
        idx[b] = idx[b - 1];
        b--;
      }
      idx[b] = v;
    }
  }
  __syncthreads();
}

// Adaptive density passes (pbf.rs:1326-1333): is pass r (0-based) run? Decided from the largest
// constraints of the passes before it; a forced count runs max(count, 1) passes.
__device__ inline bool pass_active(const FlDrum& D, const FlState& S, uint32_t r) {
  if (D.force) return r < max(D.forced_d, 1u);
  for (uint32_t s = 0; s < r; s++) {
    const uint32_t it = s + 1;
    const bool conv = __uint_as_float(S.maxc[s]) <= 0.02f;
    if ((it >= D.base && conv) || it >= D.base + 6) return false;
  }
  return true;
}

// Device arrays of the workspace (kernel argument by value).
struct FlPtrs {
  const FlDrum* drums;
  FlState* st;
  const float2* lift;   // per lifter (sin, cos)
  const float2* offs;   // boundary-point offsets of each drum
  const uint2* ftiles;  // (drum, first particle) of each fluid tile
  const uint2* rtiles;  // (drum, first row) of each CG row tile
  FlGrid* grids;        // 5 per drum
  uint32_t *gstart[5], *gcur[5], *gidx[5];
  // caller state
  float2 *x, *v, *guess;
  const float2* bx;
  const float2* bv;
  const float* bom;
  float2* imp;
  float* ang;
  // fluid workspace
  float2 *x0, *dp, *excess, *coh, *adh;
  float *dens, *lam, *mu_eff, *ff, *diagf;
  uint32_t *touched, *ncnt, *noff, *nbr;
  float* W;
  float2* Dk;
  uint32_t* pcnt;
  float4* prec;        // push records: react.x, react.y, angular term, ball (bits)
  uint32_t *bcnt, *boff, *bcur, *pids;
  // adhesion items
  uint32_t acap;
  uint32_t *acnt, *aj;
  float2* adv;
  uint32_t *fcnt, *foff, *fcur, *fids;  // per-particle gathers (adhesion items, then groups)
  // boundary groups
  uint32_t gcap;
  uint32_t *gcnt, *gj;
  float4* gc;          // c, lever_c.x, lever_c.y, lever2_c
  uint32_t *fgoff, *fgids;
  // squeeze links
  // form drag and buoyancy items per ball (GS_ITEMS slots each) and their velocity-free sums
  uint32_t *gdj, *gdn, *gbj, *gbn, *glev;
  float4 *gdr, *gds, *gbr;
  float2 *gdb, *gbs;
  uint32_t *lcnt;
  float4* lrec;        // n.x, n.y, c, other (bits)
  uint32_t *locnt, *looff, *locur, *loids;
  // CG rows
  float2 *X, *B, *R, *P, *AP;
  float* inv;
  double* part;        // FL_NP per fluid tile
  float* bins;         // 32 per fluid tile
  double* cpart;       // 4 per row tile
};

// Item merge used by adhesion and the boundary groups: items keyed by a particle index, item k
// held by lane k % 64 in slot k / 64 (NS slots), created in order of first appearance and
// accumulated in candidate order (pbf.rs:1553-1574, :2204-2226). Adhesion merges with one slot
// first and merges a ball again with FL_SLOTS slots when its items outgrow 64.
template <int NS>
struct Items {
  uint32_t j[NS];
  float a[NS], b[NS], c[NS], e[NS];
};
template <int NS>
__device__ inline void items_clear(Items<NS>& it) {
#pragma unroll
  for (int s = 0; s < NS; s++) {
    it.j[s] = FL_NO;
    it.a[s] = it.b[s] = it.c[s] = it.e[s] = 0.0f;
  }
}
// Adds (va, vb, vc, ve) to item jj (uniform values), creating it after the nit existing ones
// (past 64 NS items only nit and `over` change).
template <int NS>
__device__ inline void items_add(Items<NS>& it, uint32_t& nit, bool& over, uint32_t lane, uint32_t jj, float va,
                                 float vb, float vc, float ve) {
  // updates are value selects (no branch per slot), so the slots stay in registers
  bool mine = false;
#pragma unroll
  for (int s = 0; s < NS; s++) {
    const bool m = it.j[s] == jj;
    mine |= m;
    it.a[s] = m ? it.a[s] + va : it.a[s];
    it.b[s] = m ? it.b[s] + vb : it.b[s];
    it.c[s] = m ? it.c[s] + vc : it.c[s];
    it.e[s] = m ? it.e[s] + ve : it.e[s];
  }
  if (__ballot(mine)) return;
  if (nit >= 64u * NS) {
    over = true;
    nit++;
    return;
  }
  const uint32_t ol = nit & 63u, os = nit >> 6;
#pragma unroll
  for (int s = 0; s < NS; s++) {
    const bool w = lane == ol && (uint32_t)s == os;
    it.j[s] = w ? jj : it.j[s];
    it.a[s] = w ? va : it.a[s];
    it.b[s] = w ? vb : it.b[s];
    it.c[s] = w ? vc : it.c[s];
    it.e[s] = w ? ve : it.e[s];
  }
  nit++;
}

// Block-wide f64 sum of one value per lane in a fixed tree order (deterministic).
__device__ inline double wg_sum_d(double v, double* lds) {
  const uint32_t t = threadIdx.x;
  lds[t] = v;
  __syncthreads();
  for (uint32_t s = FL_T / 2; s > 0; s >>= 1) {
    if (t < s) lds[t] += lds[t + s];
    __syncthreads();
  }
  const double r = lds[0];
  __syncthreads();
  return r;
}
__device__ inline double wave_sum_d(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return __shfl(v, 0, 64);
}

// The largest of the wave's 64 values (unsigned), in every lane: prefix maxima inside each row of
// 16 lanes by DPP row shifts (lanes shifted in from outside the row read 0), the rows combined by
// row broadcasts, lane 63 read out.
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, true));  // row_shr:1
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, true));  // row_shr:2
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, true));  // row_shr:4
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, true));  // row_shr:8
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false));  // row_bcast:15
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false));  // row_bcast:31
  return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

// Wave-local memory ordering for LDS written by some lanes and read by others of the same wave.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Ordering of the gathered lists. A counting sort's scatter lists each target's entries in no
// particular order; every list holds distinct ids and is put in ascending id order, which is the
// reference's order.
#define FL_WCAP 256  // entries a wave orders through LDS

// By one lane, insertion in memory (the fallback for long lists).
__device__ inline void sort_insert(uint32_t* v, uint32_t s0, uint32_t s1) {
  for (uint32_t a = s0 + 1; a < s1; a++) {
    const uint32_t x = v[a];
    uint32_t c = a;
    while (c > s0 && v[c - 1] > x) {
      v[c] = v[c - 1];
      c--;
    }
    v[c] = x;
  }
}
// By one lane, m <= N entries: all loaded at once, each entry's rank counted in registers, each
// stored at its rank (no chain of dependent memory accesses).
template <int N>
__device__ inline void sort_ranks(uint32_t* v, uint32_t s0, uint32_t m) {
  uint32_t a[N];
#pragma unroll
  for (int q = 0; q < N; q++) a[q] = (uint32_t)q < m ? v[s0 + q] : FL_NO;
#pragma unroll
  for (int q = 0; q < N; q++) {
    uint32_t r = 0;
#pragma unroll
    for (int u = 0; u < N; u++) r += a[u] < a[q] ? 1u : 0u;
    if ((uint32_t)q < m) v[s0 + r] = a[q];
  }
}
__device__ inline void sort_lane(uint32_t* v, uint32_t s0, uint32_t s1) {
  const uint32_t m = s1 - s0;
  if (m <= 1) return;
  if (m <= 8) sort_ranks<8>(v, s0, m);
  else if (m <= 32) sort_ranks<32>(v, s0, m);
  else sort_insert(v, s0, s1);
}
// Ranks of a wave's list of m <= FL_WCAP distinct ids: lane L holds entries L + 64 q in mine[q]
// (FL_NO past the end); sb is the wave's LDS buffer. Every lane of the wave calls it.
__device__ inline void wave_ranks(const uint32_t* v, uint32_t s0, uint32_t m, uint32_t* sb, uint32_t lane,
                                  uint32_t (&mine)[FL_WCAP / 64], uint32_t (&rank)[FL_WCAP / 64]) {
#pragma unroll
  for (int q = 0; q < FL_WCAP / 64; q++) {
    const uint32_t k = lane + 64u * q;
    mine[q] = k < m ? v[s0 + k] : FL_NO;
    if (k < m) sb[k] = mine[q];
    rank[q] = 0;
  }
  wave_sync();
  for (uint32_t u = 0; u < m; u++) {
    const uint32_t o = sb[u];
#pragma unroll
    for (int q = 0; q < FL_WCAP / 64; q++) rank[q] += o < mine[q] ? 1u : 0u;
  }
}
// By one wave (uniform arguments): up to FL_WCAP entries by their ranks, longer lists by lane 0.
__device__ inline void wave_sort(uint32_t* v, uint32_t s0, uint32_t s1, uint32_t* sb, uint32_t lane) {
  const uint32_t m = s1 - s0;
  if (m <= 1) return;
  if (m > FL_WCAP) {
    if (lane == 0) sort_insert(v, s0, s1);
    return;
  }
  uint32_t mine[FL_WCAP / 64], rank[FL_WCAP / 64];
  wave_ranks(v, s0, m, sb, lane, mine, rank);
#pragma unroll
  for (int q = 0; q < FL_WCAP / 64; q++)
    if (lane + 64u * q < m) v[s0 + rank[q]] = mine[q];
}

}  // namespace

// ================================================================================================
// Kernels. Tile kernels take (drum, first particle) from a tile table; drum kernels run one
// workgroup per drum.
// ================================================================================================

// 0 sanitize, 1 predict (pbf.rs:1200-1235).
extern "C" __global__ void __launch_bounds__(FL_T) fl_predict(FlPtrs P) {
  const uint2 tl = P.ftiles[blockIdx.x];
  const FlDrum& D = P.drums[tl.x];
  const uint32_t i = tl.y + threadIdx.x;
  if (i >= D.n) return;
  const size_t g = (size_t)D.f0 + i;
  float2 p = P.x[g], q = P.v[g];
  if (!fin2(p) || !fin2(q)) {
    p = make_float2(0.0f, 0.0f);
    q = make_float2(0.0f, 0.0f);
  }
  P.x0[g] = p;
  P.fcnt[g] = 0;  // adhesion kicks per particle, counted by fl_adhesion
  q.y += FL_G * D.dt;
  p.x += q.x * D.dt;
  p.y += q.y * D.dt;
  P.x[g] = p;
  P.v[g] = q;
}

// Resets the drum's state; the ball grid (pbf.rs:1214-1222) and the grid over the predicted
// positions (pbf.rs:1238).
extern "C" __global__ void __launch_bounds__(FL_T) fl_grid_a(FlPtrs P) {
  __shared__ int32_t si[4 * FL_T];
  __shared__ uint32_t su[FL_T];
  const uint32_t d = blockIdx.x;
  const FlDrum& D = P.drums[d];
  FlState& S = P.st[d];
  if (threadIdx.x < FL_NPASS) S.maxc[threadIdx.x] = 0;
  if (threadIdx.x == 0) {
    S.err = 0;
    S.mode = 0;
    S.done = 1;
    S.it = S.iters = S.maxit = 0;
    S.n_groups = S.n_links = S.in_range = S.clamp_hits = 0;
    S.rs = S.target = 0.0;
    S.alpha = S.beta = 0.0f;
    S.ke_drag = S.ke_buoy = 0.0;
  }
  __syncthreads();
  if (D.n == 0) return;
  for (uint32_t b = threadIdx.x; b < D.nb; b += FL_T) P.bcnt[D.b0 + b] = 0;  // push records per ball
  FlGrid* G = P.grids + 5 * (size_t)d;
  wg_grid_build(P.x + D.f0, D.n, D.h, D.ccap[0], P.gstart[0] + D.c0[0], P.gcur[0] + D.c0[0], P.gidx[0] + D.f0, G,
                &S.err, si, su);
  if (D.nb > 0)
    wg_grid_build(P.bx + D.b0, D.nb, D.bgc, D.ccap[1], P.gstart[1] + D.c0[1], P.gcur[1] + D.c0[1], P.gidx[1] + D.b0,
                  G + 1, &S.err, si, su);
  else if (threadIdx.x == 0)
    G[1].W = G[1].H = 0;
}

// build_neighbor_lists (pbf.rs:262-288) over the predicted positions: particle i's list is the
// emission order of the candidate pairs that contain it (grid.rs:291-362) — cells (-1,-1), (-1,0),
// (0,-1), then its own cell's other points, then (1,0), (1,1), (0,1), (-1,1), (1,-1) — keeping
// |xi - xj|^2 <= h^2. fill = 0 counts, fill = 1 writes.
extern "C" __global__ void __launch_bounds__(FL_T) fl_nbr(FlPtrs P, int fill) {
  const uint2 tl = P.ftiles[blockIdx.x];
  const FlDrum& D = P.drums[tl.x];
  const uint32_t i = tl.y + threadIdx.x;
  if (i >= D.n) return;
  const float2* x = P.x + D.f0;
  const GView g = gview(P.grids + 5 * (size_t)tl.x, P.gstart[0] + D.c0[0], P.gidx[0] + D.f0);
  const float2 p = x[i];
  uint32_t cnt = 0;
  const uint32_t base = fill ? P.noff[D.f0 + tl.x + i] : 0u;
  uint32_t* out = P.nbr + D.e0;
  if (fin2(p) && g.W > 0) {
    int32_t kx, ky;
    gkey(p, g.cell, kx, ky);
    const int OX[9] = {-1, -1, 0, 0, 1, 1, 0, -1, 1};
    const int OY[9] = {-1, 0, -1, 0, 0, 1, 1, 1, -1};
    uint32_t cs[9], ce[9];  // the nine cells' ranges, their loads in flight together
#pragma unroll
    for (int q = 0; q < 9; q++) grange_any(g, sadd(kx, OX[q]), sadd(ky, OY[q]), cs[q], ce[q]);
#pragma unroll
    for (int q = 0; q < 9; q++) {
      const uint32_t s = cs[q], e = ce[q];
      for (uint32_t a = s; a < e; a++) {
        const uint32_t j = g.idx[a];
        if (j == i) continue;
        const float2 o = x[j];
        const float dx = p.x - o.x, dy = p.y - o.y;
        if (dx * dx + dy * dy <= D.h2) {
          if (fill && base + cnt < D.ecap) out[base + cnt] = j;
          cnt++;
        }
      }
    }
  }
  if (!fill) P.ncnt[D.f0 + i] = cnt;
}

// The drum's list offsets (exclusive scan of the counts); FE_EDGES past the capacity.
extern "C" __global__ void __launch_bounds__(FL_T) fl_nbr_scan(FlPtrs P) {
  __shared__ uint32_t su[FL_T];
  const uint32_t d = blockIdx.x;
  const FlDrum& D = P.drums[d];
  if (D.n == 0) return;
  const uint32_t tot = wg_scan_array(P.ncnt + D.f0, P.noff + D.f0 + d, D.n, su);
  if (threadIdx.x == 0 && tot > D.ecap) atomicOr(&P.st[d].err, FE_EDGES);
}

// Density pass r, part A (pbf.rs:1263-1300): density, constraint and lambda of every particle;
// the pass's largest constraint.
extern "C" __global__ void __launch_bounds__(FL_T) fl_dens_a(FlPtrs P, uint32_t r) {
  const uint2 tl = P.ftiles[blockIdx.x];
  const FlDrum& D = P.drums[tl.x];
  const uint32_t i = tl.y + threadIdx.x;
  if (i >= D.n || !pass_active(D, P.st[tl.x], r)) return;
  const float2* x = P.x + D.f0;
  const uint32_t* nb = P.nbr + D.e0;
  const uint32_t k0 = P.noff[D.f0 + tl.x + i], k1 = min(P.noff[D.f0 + tl.x + i + 1], D.ecap);
  const float2 p = x[i];
  float rho = D.poly0m, gsx = 0.0f, gsy = 0.0f, sgs = 0.0f;
  for (uint32_t k = k0; k < k1; k++) {
    const float2 o = x[nb[k]];
    const float2 dd = make_float2(p.x - o.x, p.y - o.y);
    const float r2 = dd.x * dd.x + dd.y * dd.y;
    rho += D.mass * poly6(r2, D.h2, D.p6c);
    const float rr = sqrtf(r2);
    const float2 rg = spiky(dd, rr, D.h, D.spc);
    const float gx = rg.x * D.mr, gy = rg.y * D.mr;
    gsx += gx;
    gsy += gy;
    sgs += gx * gx + gy * gy;
  }
  P.dens[D.f0 + i] = rho;
  const float c = fmaxf(rho / D.rest - 1.0f, 0.0f);
  sgs += gsx * gsx + gsy * gsy;
  P.lam[D.f0 + i] = -c / (sgs + 200.0f);
  atomicMax(&P.st[tl.x].maxc[r], __float_as_uint(c));
}

// Density pass r, part B (pbf.rs:1302-1318): the position correction from the lambdas.
extern "C" __global__ void __launch_bounds__(FL_T) fl_dens_b(FlPtrs P, uint32_t r) {
  const uint2 tl = P.ftiles[blockIdx.x];
  const FlDrum& D = P.drums[tl.x];
  const uint32_t i = tl.y + threadIdx.x;
  if (i >= D.n || !pass_active(D, P.st[tl.x], r)) return;
  const float2* x = P.x + D.f0;
  const float* lam = P.lam + D.f0;
  const uint32_t* nb = P.nbr + D.e0;
  const uint32_t k0 = P.noff[D.f0 + tl.x + i], k1 = min(P.noff[D.f0 + tl.x + i + 1], D.ecap);
  const float2 p = x[i];
  const float li = lam[i];
  float sx = 0.0f, sy = 0.0f;
  for (uint32_t k = k0; k < k1; k++) {
    const uint32_t j = nb[k];
    const float2 o = x[j];
    const float2 dd = make_float2(p.x - o.x, p.y - o.y);
    const float rr = sqrtf(dd.x * dd.x + dd.y * dd.y);
    const float2 rg = spiky(dd, rr, D.h, D.spc);
    const float l = li + lam[j];
    sx += l * rg.x;
    sy += l * rg.y;
  }
  P.dp[D.f0 + i] = make_float2(sx * D.mr, sy * D.mr);
}

// Density pass r, part C: x += dp (pbf.rs:1320-1324).
extern "C" __global__ void __launch_bounds__(FL_T) fl_dens_c(FlPtrs P, uint32_t r) {
  const uint2 tl = P.ftiles[blockIdx.x];
  const FlDrum& D = P.drums[tl.x];
  const uint32_t i = tl.y + threadIdx.x;
  if (i >= D.n || !pass_active(D, P.st[tl.x], r)) return;
  const size_t g = (size_t)D.f0 + i;
  const float2 p = P.x[g], q = P.dp[g];
  P.x[g] = make_float2(p.x + q.x, p.y + q.y);
}

// 3.5 overlap push (pbf.rs:1361-1444): each particle is pushed out of the balls near it (the
// candidates of its pre-push position, in the ball grid's visiting order), with the momentum-
// bounded reaction recorded per (particle, ball) for the balls' impulses.
extern "C" __global__ void __launch_bounds__(FL_T) fl_push(FlPtrs P) {
  const uint2 tl = P.ftiles[blockIdx.x];
  const FlDrum& D = P.drums[tl.x];
  const uint32_t i = tl.y + threadIdx.x;
  if (i >= D.n) return;
  const size_t g = (size_t)D.f0 + i;
  float2 p = P.x[g];
  const float2 vi = P.v[g];
  float exx = 0.0f, exy = 0.0f;
  uint32_t cnt = 0;
  if (D.nb > 0) {
    const GView gb = gview(P.grids + 5 * (size_t)tl.x + 1, P.gstart[1] + D.c0[1], P.gidx[1] + D.b0);
    Near q;
    near_cells(gb, p, q);
    const float2* bx = P.bx + D.b0;
    const float2* bv = P.bv + D.b0;
    const float* bom = P.bom + D.b0;
    for (uint32_t c = 0; c < q.total; c++) {
      const uint32_t b = near_at(gb, q, c);
      const float2 bp = bx[b];
      const float2 dd = make_float2(p.x - bp.x, p.y - bp.y);
      const float dist = sqrtf(dd.x * dd.x + dd.y * dd.y);
      if (__builtin_isnan(dist) || dist >= D.contact_r) continue;
      const float2 nh = dist > 1e-9f ? make_float2(dd.x / dist, dd.y / dist) : make_float2(1.0f, 0.0f);
      const float pm = fminf(D.contact_r - dist, 0.5f * D.br);
      p.x += pm * nh.x;
      p.y += pm * nh.y;
      const float lx = nh.x * D.br, ly = nh.y * D.br;
      const float om = bom[b];
      const float2 bvb = bv[b];
      const float vsx = bvb.x + om * -ly, vsy = bvb.y + om * lx;
      const float arrest = fmaxf(-((vi.x - vsx) * nh.x + (vi.y - vsy) * nh.y), 0.0f);
      const float full = D.mass * pm / D.dt;
      const float im = fminf(full, D.red * arrest);
      const float s = (full - im) / D.mass;
      exx += s * nh.x;
      exy += s * nh.y;
      const float rx = -im * nh.x, ry = -im * nh.y;
      const float an = lx * ry - ly * rx;
      if (cnt < FL_PK) {
        P.prec[g * FL_PK + cnt] = make_float4(rx, ry, an, __uint_as_float(b));
        atomicAdd(&P.bcnt[D.b0 + b], 1u);
      } else {
        atomicOr(&P.st[tl.x].err, FE_PUSH);
      }
      cnt++;
    }
  }
  P.x[g] = p;
  P.excess[g] = make_float2(exx, exy);
  P.pcnt[g] = min(cnt, (uint32_t)FL_PK);
}

// The balls' push impulses (pbf.rs:1432-1440): each ball's records summed in particle order,
// gathered by a counting sort — fl_push counts the records per ball, fl_push_scan forms each
// ball's range, fl_push_scatter lists the records, fl_push_fin orders and sums them.
extern "C" __global__ void __launch_bounds__(FL_T) fl_push_scan(FlPtrs P) {
  __shared__ uint32_t su[FL_T];
  const uint32_t d = blockIdx.x, t = threadIdx.x;
  const FlDrum& D = P.drums[d];
  const uint32_t nb = D.nb;
  if (nb == 0) return;
  if (D.n == 0) {
    for (uint32_t b = t; b < nb; b += FL_T) {
      P.imp[D.b0 + b] = make_float2(0.0f, 0.0f);
      P.ang[D.b0 + b] = 0.0f;
    }
    return;
  }
  uint32_t* boff = P.boff + D.b0 + d;
  wg_scan_array(P.bcnt + D.b0, boff, nb, su);
  for (uint32_t b = t; b < nb; b += FL_T) P.bcur[D.b0 + b] = boff[b];
}

extern "C" __global__ void __launch_bounds__(FL_T) fl_push_scatter(FlPtrs P) {
  const uint2 tl = P.ftiles[blockIdx.x];
  const FlDrum& D = P.drums[tl.x];
  const uint32_t i = tl.y + threadIdx.x;
  if (i >= D.n) return;
  const size_t g = (size_t)D.f0 + i;
  const uint32_t c = P.pcnt[g];
  uint32_t* ids = P.pids + (size_t)D.f0 * FL_PK;
  for (uint32_t s = 0; s < c; s++) {
    const uint32_t b = __float_as_uint(P.prec[g * FL_PK + s].w);
    ids[atomicAdd(&P.bcur[D.b0 + b], 1u)] = i * FL_PK + s;
  }
}

extern "C" __global__ void __launch_bounds__(FL_T) fl_push_fin(FlPtrs P, const uint2* __restrict__ btiles) {
  __shared__ uint32_t sid[4][FL_WCAP];
  __shared__ float4 srec[4][FL_WCAP];
  const uint2 tl = btiles[blockIdx.x];
  const FlDrum& D = P.drums[tl.x];
  const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6, b = tl.y + w;
  if (b >= D.nb) return;
  uint32_t* ids = P.pids + (size_t)D.f0 * FL_PK;
  const float4* rec = P.prec + (size_t)D.f0 * FL_PK;
  const uint32_t* boff = P.boff + D.b0 + tl.x;
  const uint32_t s0 = boff[b], s1 = boff[b + 1], m = s1 - s0;
  float ix = 0.0f, iy = 0.0f, an = 0.0f;
  if (m <= FL_WCAP) {  // the records placed in LDS at their ranks, then summed in particle order
    uint32_t mine[FL_WCAP / 64], rank[FL_WCAP / 64];
    wave_ranks(ids, s0, m, sid[w], lane, mine, rank);
#pragma unroll
    for (int q = 0; q < FL_WCAP / 64; q++)
      if (lane + 64u * q < m) srec[w][rank[q]] = rec[mine[q]];
    wave_sync();
    if (lane == 0)
      for (uint32_t a = 0; a < m; a++) {
        const float4 r = srec[w][a];
        ix += r.x;
        iy += r.y;
        an += r.z;
      }
  } else if (lane == 0) {
    sort_insert(ids, s0, s1);
    for (uint32_t a = s0; a < s1; a++) {
      const float4 r = rec[ids[a]];
      ix += r.x;
      iy += r.y;
      an += r.z;
    }
  }
  if (lane == 0) {
    P.imp[D.b0 + b] = make_float2(ix, iy);
    P.ang[D.b0 + b] = an;
  }
}

// 3.6a cohesion (pbf.rs:1464-1491) over the neighbour lists at the pushed positions.
extern "C" __global__ void __launch_bounds__(FL_T) fl_cohesion(FlPtrs P) {
  const uint2 tl = P.ftiles[blockIdx.x];
  const FlDrum& D = P.drums[tl.x];
  const uint32_t i = tl.y + threadIdx.x;
  if (i >= D.n) return;
  float px = 0.0f, py = 0.0f;
  if (D.coh_acc > 0.0f) {
    const float2* x = P.x + D.f0;
    const uint32_t* nb = P.nbr + D.e0;
    const uint32_t k0 = P.noff[D.f0 + tl.x + i], k1 = min(P.noff[D.f0 + tl.x + i + 1], D.ecap);
    const float2 p = x[i];
    for (uint32_t k = k0; k < k1; k++) {
      const float2 o = x[nb[k]];
      const float dx = p.x - o.x, dy = p.y - o.y;
      const float r = sqrtf(dx * dx + dy * dy);
      if (r <= 1e-9f) continue;
      const float w = D.mass * cohesion_k(r, D.h, D.cohc, D.h6_64) / D.rest;
      px -= w * (dx / r);
      py -= w * (dy / r);
    }
    px = D.coh_acc * px * D.dt;
    py = D.coh_acc * py * D.dt;
  }
  P.coh[D.f0 + i] = make_float2(px, py);
}

// One ball's adhesion with NS item slots a lane; false (and nothing written) when NS is 1 and
// the ball's items outgrow them.
template <int NS>
__device__ __forceinline__ bool adh_ball(const FlPtrs& P, const FlDrum& D, uint2 tl, uint32_t lane, uint32_t b,
                                         uint32_t* acnt, float2 bxb) {
  const GView g = gview(P.grids + 5 * (size_t)tl.x, P.gstart[0] + D.c0[0], P.gidx[0] + D.f0);
  const float2* x = P.x + D.f0;
  Items<NS> it;
  items_clear<NS>(it);
  uint32_t nit = 0;
  bool over = false;
  for (uint32_t k = 0; k < D.cpb; k++) {
    const float2 o = P.offs[D.off0 + k];
    const float2 bp = make_float2(bxb.x + o.x, bxb.y + o.y);
    Near q;
    near_cells_wave(g, bp, q);
    for (uint32_t c0 = 0; c0 < q.total; c0 += 64) {
      const uint32_t c = c0 + lane;
      bool valid = false;
      uint32_t j = 0;
      float w = 0.0f, nx = 0.0f, ny = 0.0f;
      if (c < q.total) {
        j = near_at(g, q, c);
        const float2 xj = x[j];
        const float dx = bp.x - xj.x, dy = bp.y - xj.y;
        const float r2 = dx * dx + dy * dy;
        if (!(r2 >= D.ad_range2 || r2 <= D.ad_inner2)) {
          const float r = sqrtf(r2);
          if (__builtin_isfinite(r) && !(r <= 0.5f * D.ha) && !(r >= D.ha)) {
            w = D.mass * adhesion_k(r, D.ha, D.adc) / D.rest;
            if (w > 0.0f) {
              valid = true;
              nx = dx / r;
              ny = dy / r;
            }
          }
        }
      }
      uint64_t m = __ballot(valid);
      while (m) {
        const int L = first_lane(m);
        m &= m - 1;
        const float ww = rdl(w, L);
        items_add(it, nit, over, lane, rdlu(j, L), ww * rdl(nx, L), ww * rdl(ny, L), ww, 0.0f);
      }
    }
  }
  if (NS == 1 && nit > 64u) return false;  // merged again with FL_SLOTS slots
  if (over && lane == 0) atomicOr(&P.st[tl.x].err, FE_ADH);
  nit = min(nit, 64u * NS);
  float wsum = 0.0f, pdx = 0.0f, pdy = 0.0f;
#pragma unroll
  for (int s = 0; s < NS; s++) {  // items in order: slot-major, then lane
    const uint32_t lim = nit > 64u * s ? min(64u, nit - 64u * s) : 0u;
    for (uint32_t L = 0; L < lim; L++) {
      pdx += rdl(it.a[s], (int)L);
      pdy += rdl(it.b[s], (int)L);
      wsum += rdl(it.c[s], (int)L);
    }
  }
  const float pl = sqrtf(pdx * pdx + pdy * pdy);
  if (nit == 0 || wsum <= 1e-9f || pl <= 1e-9f) {
    if (lane == 0) acnt[b] = 0;
    return true;
  }
  const float bdx = -pdx / pl, bdy = -pdy / pl;
  const float cov = fminf(wsum, 1.0f);
  const float mc = D.mass * wsum;
  const float att = mc / (D.bmass + mc);
  const float sc = D.ad_acc * cov * D.dt * att;
  const float dvx = bdx * sc, dvy = bdy * sc;
  if (lane == 0) {
    float2* imp = P.imp + D.b0 + b;
    const float2 i0 = *imp;
    *imp = make_float2(i0.x + D.bmass * dvx, i0.y + D.bmass * dvy);
    const float lx = bdx * D.br, ly = bdy * D.br;
    P.ang[D.b0 + b] += D.bmass * (lx * dvy - ly * dvx);
    acnt[b] = nit;
  }
  const size_t base = ((size_t)D.b0 + b) * P.acap;
#pragma unroll
  for (int s = 0; s < NS; s++) {
    const uint32_t k = (uint32_t)s * 64u + lane;
    if (k < nit) {
      const float frac = it.c[s] / wsum;
      const float f = -(D.bmass * frac / D.mass);
      P.aj[base + k] = it.j[s];
      P.adv[base + k] = make_float2(f * dvx, f * dvy);
      atomicAdd(&P.fcnt[D.f0 + it.j[s]], 1u);
    }
  }
  return true;
}

// 3.6b adhesion (pbf.rs:1515-1631), one wave per ball: the boundary points' candidates (the grid
// over the predicted positions, current positions) merged into items in order of first
// appearance, one bounded pull on the ball and its reaction spread over the items.
extern "C" __global__ void __launch_bounds__(FL_T) fl_adhesion(FlPtrs P, const uint2* __restrict__ btiles) {
  const uint2 tl = btiles[blockIdx.x];
  const FlDrum& D = P.drums[tl.x];
  const uint32_t lane = threadIdx.x & 63u, b = tl.y + (threadIdx.x >> 6);
  if (b >= D.nb) return;
  uint32_t* acnt = P.acnt + D.b0;
  const float2 bxb = P.bx[D.b0 + b];
  if (!D.adh_on || D.n == 0 || !fin2(bxb)) {
    if (lane == 0) acnt[b] = 0;
    return;
  }
  if (!adh_ball<1>(P, D, tl, lane, b, acnt, bxb)) adh_ball<FL_SLOTS>(P, D, tl, lane, b, acnt, bxb);
}

// The fluid's adhesion kicks (pbf.rs:1620-1626): each particle's items summed in ball order,
// gathered like the push impulses — fl_adhesion counts the items per particle, fl_adh_scan forms
// the ranges, fl_adh_scatter lists the items, fl_adh_fin orders and sums them.
extern "C" __global__ void __launch_bounds__(FL_T) fl_adh_scan(FlPtrs P) {
  __shared__ uint32_t su[FL_T];
  const uint32_t d = blockIdx.x, t = threadIdx.x;
  const FlDrum& D = P.drums[d];
  if (D.n == 0 || !D.adh_on) return;
  uint32_t* off = P.foff + D.f0 + d;
  wg_scan_array(P.fcnt + D.f0, off, D.n, su);
  for (uint32_t i = t; i < D.n; i += FL_T) P.fcur[D.f0 + i] = off[i];
}

extern "C" __global__ void __launch_bounds__(FL_T) fl_adh_scatter(FlPtrs P, const uint2* __restrict__ btiles) {
  const uint2 tl = btiles[blockIdx.x];
  const FlDrum& D = P.drums[tl.x];
  const uint32_t lane = threadIdx.x & 63u, b = tl.y + (threadIdx.x >> 6);
  if (b >= D.nb || !D.adh_on) return;
  const uint32_t na = P.acnt[D.b0 + b];
  uint32_t* ids = P.fids + (size_t)D.b0 * P.acap;
  const uint32_t* aj = P.aj + (size_t)D.b0 * P.acap;
  for (uint32_t k = lane; k < na; k += 64) {
    const uint32_t id = b * P.acap + k;
    ids[atomicAdd(&P.fcur[D.f0 + aj[id]], 1u)] = id;
  }
}

extern "C" __global__ void __launch_bounds__(FL_T) fl_adh_fin(FlPtrs P) {
  const uint2 tl = P.ftiles[blockIdx.x];
  const FlDrum& D = P.drums[tl.x];
  const uint32_t i = tl.y + threadIdx.x;
  if (i >= D.n) return;
  float sx = 0.0f, sy = 0.0f;
  if (D.adh_on) {
    uint32_t* ids = P.fids + (size_t)D.b0 * P.acap;
    const float2* adv = P.adv + (size_t)D.b0 * P.acap;
    const uint32_t* off = P.foff + D.f0 + tl.x;
    const uint32_t s0 = off[i], s1 = off[i + 1];
    sort_lane(ids, s0, s1);  // ball order
    for (uint32_t a = s0; a < s1; a++) {
      const float2 q = adv[ids[a]];
      sx += q.x;
      sy += q.y;
    }
  }
  P.adh[D.f0 + i] = make_float2(sx, sy);
}

// 4 wall projection (pbf.rs:1635-1653): push inside the wall at angle_next, the f64 wall impulse
// and angular impulse, and the 32 wall-angle bins (tile partial sums).
extern "C" __global__ void __launch_bounds__(FL_T) fl_wall(FlPtrs P) {
  __shared__ double sd[FL_T];
  __shared__ uint32_t sbin[FL_T];
  __shared__ float sval[FL_T];
  const uint2 tl = P.ftiles[blockIdx.x];
  const FlDrum& D = P.drums[tl.x];
  const uint32_t t = threadIdx.x, i = tl.y + t;
  const bool act = i < D.n;
  double wx = 0.0, wy = 0.0, wa = 0.0;
  uint32_t bin = 32;
  float bval = 0.0f;
  if (act) {
    const size_t g = (size_t)D.f0 + i;
    float2 p = P.x[g];
    const mgg::Geo geo{D.radius, D.omega, D.nl, D.rt, D.bh, D.thw, P.lift + D.lift0};
    float nx, ny;
    const float d = mgg::sdf_world(p.x, p.y, D.rc1, D.rs1, D.wc1, D.ws1, geo, nx, ny);
    uint32_t touched = 0;
    if (d < 0.0f) {
      const float mx = -d * nx, my = -d * ny;
      p.x += mx;
      p.y += my;
      touched = 1;
      const float px = D.mass * mx / D.dt, py = D.mass * my / D.dt;
      wx = (double)px;
      wy = (double)py;
      wa = (double)(p.x * py - p.y * px);
      const float a = mgm::ref_atan2f(p.y, p.x);
      const float u = (a + FL_PI) / FL_TAU * 32.0f;
      const uint32_t q = (__builtin_isnan(u) || u <= 0.0f) ? 0u : (u >= 32.0f ? 31u : (uint32_t)u);
      bin = min(q, 31u);
      bval = D.mass * -d / D.dt;
      P.x[g] = p;
    }
    P.touched[g] = touched;
  }
  sbin[t] = bin;
  sval[t] = bval;
  const double sx = wg_sum_d(wx, sd), sy = wg_sum_d(wy, sd), sa = wg_sum_d(wa, sd);
  double* part = P.part + (size_t)blockIdx.x * FL_NP;
  if (t == 0) {
    part[P_WIX] = sx;
    part[P_WIY] = sy;
    part[P_WANG] = sa;
  }
  if (t < 32) {
    float s = 0.0f;
    for (uint32_t k = 0; k < FL_T; k++)
      if (sbin[k] == t) s += sval[k];
    P.bins[(size_t)blockIdx.x * 32 + t] = s;
  }
}

// 5 velocities from positions and the coupling terms (pbf.rs:1656-1674), 6 no-slip blend with the
// wall work (pbf.rs:1677-1700); f64 |v|^2 sums before and after the blend.
extern "C" __global__ void __launch_bounds__(FL_T) fl_vel(FlPtrs P) {
  __shared__ double sd[FL_T];
  const uint2 tl = P.ftiles[blockIdx.x];
  const FlDrum& D = P.drums[tl.x];
  const uint32_t t = threadIdx.x, i = tl.y + t;
  double k0 = 0.0, k1 = 0.0, ww = 0.0, wx = 0.0, wy = 0.0, wa = 0.0;
  if (i < D.n) {
    const size_t g = (size_t)D.f0 + i;
    const float2 p = P.x[g], q0 = P.x0[g], ex = P.excess[g], co = P.coh[g], ad = P.adh[g];
    float vx = (p.x - q0.x) / D.dt, vy = (p.y - q0.y) / D.dt;
    vx -= ex.x;
    vy -= ex.y;
    vx += co.x;
    vy += co.y;
    vx += ad.x;
    vy += ad.y;
    k0 = (double)(vx * vx + vy * vy);
    if (D.beta_w > 0.0f && P.touched[g]) {
      const float wvx = D.omega * -p.y, wvy = D.omega * p.x;
      const float ox = vx, oy = vy;
      vx = (1.0f - D.beta_w) * ox + D.beta_w * wvx;
      vy = (1.0f - D.beta_w) * oy + D.beta_w * wvy;
      const float dvx = vx - ox, dvy = vy - oy;
      ww = (double)(D.mass * (dvx * wvx + dvy * wvy));
      const float mx = D.mass * dvx, my = D.mass * dvy;
      wx = (double)mx;
      wy = (double)my;
      wa = (double)(p.x * my - p.y * mx);
    }
    k1 = (double)(vx * vx + vy * vy);
    P.v[g] = make_float2(vx, vy);
  }
  const double s0 = wg_sum_d(k0, sd), s1 = wg_sum_d(k1, sd), s2 = wg_sum_d(ww, sd);
  const double s3 = wg_sum_d(wx, sd), s4 = wg_sum_d(wy, sd), s5 = wg_sum_d(wa, sd);
  if (t == 0) {
    double* part = P.part + (size_t)blockIdx.x * FL_NP;
    part[P_KE0] = s0;
    part[P_KE1] = s1;
    part[P_WW] = s2;
    part[P_WIX] += s3;
    part[P_WIY] += s4;
    part[P_WANG] += s5;
  }
}

// Grids over the final positions (form drag's at h + r, the boundary groups' at h) and over the
// balls at the squeeze-link reach (pbf.rs:1790, :2176, :557); workgroup (d, k) builds drum d's
// grid 2 + k (the three are independent).
extern "C" __global__ void __launch_bounds__(FL_T) fl_grid_b(FlPtrs P) {
  __shared__ int32_t si[4 * FL_T];
  __shared__ uint32_t su[FL_T];
  const uint32_t d = blockIdx.x, k = blockIdx.y;
  const FlDrum& D = P.drums[d];
  if (D.n == 0) return;
  FlGrid* G = P.grids + 5 * (size_t)d;
  FlState& S = P.st[d];
  if (k == 0 && D.drag_on)
    wg_grid_build(P.x + D.f0, D.n, fmaxf(D.hc, 1e-6f), D.ccap[2], P.gstart[2] + D.c0[2], P.gcur[2] + D.c0[2],
                  P.gidx[2] + D.f0, G + 2, &S.err, si, su);
  if (k == 1 && D.grp_on)
    wg_grid_build(P.x + D.f0, D.n, D.h, D.ccap[3], P.gstart[3] + D.c0[3], P.gcur[3] + D.c0[3], P.gidx[3] + D.f0,
                  G + 3, &S.err, si, su);
  if (k == 2 && (D.links_on || D.drag_on || D.buoy_on))
    wg_grid_build(P.bx + D.b0, D.nb, D.g5c, D.ccap[4], P.gstart[4] + D.c0[4], P.gcur[4] + D.c0[4], P.gidx[4] + D.b0,
                  G + 4, &S.err, si, su);
}

// 6.5 form drag (pbf.rs:1777-2013) then 6.6 buoyancy (pbf.rs:2035-2122): Gauss-Seidel over the
// balls in index order — each ball reads the fluid velocities the balls before it wrote (seeded
// balls are numbered row by row, so neighbouring balls in index order nearly always share fluid
// particles and the chain of dependent balls is as long as the population). Everything a ball
// needs that does not depend on the fluid velocities is computed for all balls at once first
// (fl_gs_pre, one wave per ball): its items in candidate order (drag: particles within h + r, with
// the kernel weight; buoyancy: particles within h, with their weight and position), and the
// order-sensitive f32 sums over them that do not read velocities (entrained mass, first moment,
// inertia and the rotation sums s, s_scale for drag; local density and the weight of the
// off-axis items for buoyancy). The sweep (fl_gs, one wave per drum, the drum's fluid velocities
// held in LDS when they fit) then walks the balls in order: gather the items' velocities, the
// three velocity sums in item order (one per lane), the ball's scalar update, the kicks.
#define GS_ITEMS 224
#define GS_RING 6  // balls' records in flight per wave in a sweep of short lists

// Per-ball items and velocity-independent sums, one wave per ball.
extern "C" __global__ void __launch_bounds__(FL_T) fl_gs_pre(FlPtrs P, const uint2* __restrict__ btiles) {
  __shared__ float s_a[4][GS_ITEMS], s_b[4][GS_ITEMS], s_c[4][GS_ITEMS], s_d[4][GS_ITEMS];
  const uint2 tl = btiles[blockIdx.x];
  const FlDrum& D = P.drums[tl.x];
  const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6, b = tl.y + wv;
  if (b >= D.nb) return;
  const size_t gb = (size_t)D.b0 + b;
  if (lane == 0) P.locnt[gb] = 0;  // squeeze links naming the ball as partner, counted by fl_links
  const float2 c = P.bx[gb];
  const float2* x = P.x + D.f0;
  float *sa = s_a[wv], *sb = s_b[wv], *sc = s_c[wv], *sd = s_d[wv];
  bool over = false;
  uint32_t nit = 0;
  if (D.drag_on && fin2(c)) {
    const GView g = gview(P.grids + 5 * (size_t)tl.x + 2, P.gstart[2] + D.c0[2], P.gidx[2] + D.f0);
    Near q;
    near_cells_wave(g, c, q);
    for (uint32_t c0 = 0; c0 < q.total; c0 += 64) {
      const uint32_t cc = c0 + lane;
      bool valid = false;
      uint32_t j = 0;
      float rix = 0.0f, riy = 0.0f, phi = 0.0f;
      if (cc < q.total) {
        j = near_at(g, q, cc);
        const float2 xj = x[j];
        rix = xj.x - c.x;
        riy = xj.y - c.y;
        const float r2 = rix * rix + riy * riy;
        if (__builtin_isfinite(r2) && !(r2 >= D.hc2)) {
          phi = poly6(r2, D.hc2, D.p6hc) / D.w0;
          valid = phi > 0.0f;
        }
      }
      const uint64_t m = __ballot(valid);
      const uint32_t pos = nit + (uint32_t)__builtin_popcountll(m & ((1ull << lane) - 1ull));
      if (valid && pos < GS_ITEMS) {
        P.gdj[gb * GS_ITEMS + pos] = j;
        P.gdr[gb * GS_ITEMS + pos] = make_float4(rix, riy, phi, D.mass * phi);
        sa[pos] = rix;
        sb[pos] = riy;
        sc[pos] = phi;
      }
      nit += (uint32_t)__builtin_popcountll(m);
    }
    if (nit > GS_ITEMS) {
      over = true;
      nit = GS_ITEMS;
    }
    wave_sync();
    float acc = 0.0f;  // entrained mass, first moment x / y, inertia: one sum per lane, item order
    if (lane < 4) {
      for (uint32_t i = 0; i < nit; i++) {
        const float w = D.mass * sc[i], ax = sa[i], ay = sb[i];
        acc += lane == 0 ? w : lane == 1 ? w * ax : lane == 2 ? w * ay : w * (ax * ax + ay * ay);
      }
    }
    const float m_ent = rdl(acc, 0), rsx = rdl(acc, 1), rsy = rdl(acc, 2), i_ent = rdl(acc, 3);
    float s = 0.0f, ss = 0.0f, rbx = 0.0f, rby = 0.0f;
    if (m_ent > 1e-12f && nit > 0) {
      rbx = rsx / m_ent;
      rby = rsy / m_ent;
      const float pbx = -rby, pby = rbx;
      float s2 = 0.0f;
      if (lane < 2) {
        for (uint32_t i = 0; i < nit; i++) {
          const float ax = sa[i], ay = sb[i], p = sc[i];
          const float phi2 = p * p;
          s2 += lane == 0 ? D.mass * phi2 * (ax * (ax - pby) - ay * (-ay - pbx)) : D.mass * phi2 * (ax * ax + ay * ay);
        }
      }
      s = rdl(s2, 0);
      ss = rdl(s2, 1);
    }
    if (lane == 0) {
      P.gds[gb] = make_float4(m_ent, i_ent, s, ss);
      P.gdb[gb] = make_float2(rbx, rby);
    }
    wave_sync();
  }
  if (lane == 0) {
    P.gdn[gb] = nit;
    if (nit < GS_ITEMS) P.gdj[gb * GS_ITEMS + nit] = FL_NO;  // the list's end for the schedule
  }
  uint32_t nw = 0;
  if (D.buoy_on && fin2(c)) {
    const GView g = gview(P.grids + 5 * (size_t)tl.x, P.gstart[0] + D.c0[0], P.gidx[0] + D.f0);
    Near q;
    near_cells_wave(g, c, q);
    for (uint32_t c0 = 0; c0 < q.total; c0 += 64) {
      const uint32_t cc = c0 + lane;
      uint32_t j = 0;
      float w = 0.0f;
      float2 xj = make_float2(0.0f, 0.0f);
      if (cc < q.total) {
        j = near_at(g, q, cc);
        xj = x[j];
        const float dx = c.x - xj.x, dy = c.y - xj.y;
        w = D.mass * poly6(dx * dx + dy * dy, D.h2, D.p6c);
      }
      const bool valid = w > 0.0f;
      const uint64_t m = __ballot(valid);
      const uint32_t pos = nw + (uint32_t)__builtin_popcountll(m & ((1ull << lane) - 1ull));
      if (valid && pos < GS_ITEMS) {
        const float d2 = xj.x * xj.x + xj.y * xj.y;
        const float use = (d2 <= 1e-12f) ? 0.0f : 1.0f;
        P.gbj[gb * GS_ITEMS + pos] = j;
        P.gbr[gb * GS_ITEMS + pos] = make_float4(w, xj.x, xj.y, use);
        sa[pos] = w;
        sd[pos] = use;
      }
      nw += (uint32_t)__builtin_popcountll(m);
    }
    if (nw > GS_ITEMS) {
      over = true;
      nw = GS_ITEMS;
    }
    wave_sync();
    float acc = 0.0f;  // local density; weight of the items off the axis
    if (lane < 2)
      for (uint32_t i = 0; i < nw; i++)
        if (lane == 0 || sd[i] != 0.0f) acc += sa[i];
    const float od = rdl(acc, 1);
    if (lane == 0) P.gbs[gb] = make_float2(acc, od);
  }
  if (lane == 0) {
    P.gbn[gb] = nw;
    if (nw < GS_ITEMS) P.gbj[gb * GS_ITEMS + nw] = FL_NO;
  }
  if (over && lane == 0) atomicOr(&P.st[tl.x].err, FE_ITEMS);
}

// acc + v[0] + v[1] + ... + v[n-1], added in that order (eight LDS values per two 16-byte reads).
__device__ __forceinline__ float ordered_sum(const float* __restrict__ v, uint32_t n, float acc) {
  const float4* v4 = reinterpret_cast<const float4*>(v);
  uint32_t i = 0;
  for (; i + 8 <= n; i += 8) {
    const float4 a = v4[i >> 2], b = v4[(i >> 2) + 1];
    acc += a.x;
    acc += a.y;
    acc += a.z;
    acc += a.w;
    acc += b.x;
    acc += b.y;
    acc += b.z;
    acc += b.w;
  }
  for (; i < n; i++) acc += v[i];
  return acc;
}

// The sweeps' schedule. Ball b of the reference's in-order walks reads and writes the velocities of
// its items (the particles of its drag and buoyancy lists); two balls sharing no item commute. A
// ball's level is one more than the largest level of the earlier balls it shares an item with
// (the sequential order restricted to dependent balls), so every ball of a level is independent
// of the others and a particle's updates happen in ball order: walking the levels in order, the
// balls of a level side by side, reproduces the in-order walks exactly. One wave per drum finds
// the levels in ball order, keeping the last level of every particle in LDS (16 bits; entry n reads
// as 0 and entry n + 1 takes the writes of idle lanes); a drum with more particles than that table
// holds keeps the plain order (level b + 1). A ball's lists are one item per lane (lists of up to
// 64 items; a longer list takes the wide step, which reads the lists as it goes); the short lists
// and the counts are loaded eight balls ahead, and the levels leave in runs of 64.
struct LvItems {
  uint32_t jd, jb, nd, nw;
};
__device__ __forceinline__ void lv_load(const FlPtrs& P, size_t g0, uint32_t nb, uint32_t b, LvItems& I) {
  const uint32_t lane = threadIdx.x & 63u;
  const size_t g = g0 + (b < nb ? b : nb - 1);
  I.nd = P.gdn[g];
  I.nw = P.gbn[g];
  I.jd = P.gdj[g * GS_ITEMS + lane];
  I.jb = P.gbj[g * GS_ITEMS + lane];
}
// The level of a ball with a list longer than 64 items: the first 64 items of each list from the
// registers, the rest read here, 64 at a time.
__device__ __forceinline__ uint32_t lv_wide(const FlPtrs& P, uint16_t* s_last, const LvItems& I, size_t g, uint32_t nd,
                                            uint32_t nw, uint32_t n) {
  const uint32_t lane = threadIdx.x & 63u;
  uint32_t m = max(s_last[lane < nd ? I.jd : n], s_last[lane < nw ? I.jb : n]);
  for (uint32_t k = lane + 64; k < nd; k += 64) m = max(m, (uint32_t)s_last[P.gdj[g * GS_ITEMS + k]]);
  for (uint32_t k = lane + 64; k < nw; k += 64) m = max(m, (uint32_t)s_last[P.gbj[g * GS_ITEMS + k]]);
  const uint32_t l = wave_max_u32(m) + 1;
  wave_sync();
  s_last[lane < nd ? I.jd : n + 1] = (uint16_t)l;
  s_last[lane < nw ? I.jb : n + 1] = (uint16_t)l;
  for (uint32_t k = lane + 64; k < nd; k += 64) s_last[P.gdj[g * GS_ITEMS + k]] = (uint16_t)l;
  for (uint32_t k = lane + 64; k < nw; k += 64) s_last[P.gbj[g * GS_ITEMS + k]] = (uint16_t)l;
  wave_sync();
  return l;
}
template <bool WIDE>
__device__ __forceinline__ uint32_t lv_step(const FlPtrs& P, uint16_t* s_last, const LvItems& I, size_t g0, uint32_t b,
                                            uint32_t nb, uint32_t n) {
  if (b >= nb) return 0;
  const uint32_t nd = __builtin_amdgcn_readfirstlane(I.nd), nw = __builtin_amdgcn_readfirstlane(I.nw);
  if (WIDE && (nd > 64u || nw > 64u)) return lv_wide(P, s_last, I, g0 + b, nd, nw, n);
  const uint32_t lane = threadIdx.x & 63u;
  const bool vd = lane < nd, vb = lane < nw;
  const uint32_t a = s_last[vd ? I.jd : n], c = s_last[vb ? I.jb : n];
  const uint32_t l = wave_max_u32(max(a, c)) + 1;
  wave_sync();
  s_last[vd ? I.jd : n + 1] = (uint16_t)l;
  s_last[vb ? I.jb : n + 1] = (uint16_t)l;
  wave_sync();
  return l;
}
// The levels of a drum's balls in ball order; WIDE: some list is longer than 64 items (a separate
// loop, so that the short steps' waits for the loads eight balls ahead stay exact).
template <bool WIDE>
__device__ __forceinline__ uint32_t lv_walk(const FlPtrs& P, uint16_t* s_last, uint32_t* lev, size_t g0, uint32_t nb,
                                            uint32_t n) {
  const uint32_t lane = threadIdx.x & 63u;
  LvItems R[8];
#pragma unroll
  for (int r = 0; r < 8; r++) lv_load(P, g0, nb, (uint32_t)r, R[r]);
  uint32_t top = 0, run = 0;  // run: lane q holds the level of ball 64 k + q of the current run
  for (uint32_t b = 0; b < nb; b += 8) {
#pragma unroll
    for (int r = 0; r < 8; r++) {
      const uint32_t l = lv_step<WIDE>(P, s_last, R[r], g0, b + r, nb, n);
      lv_load(P, g0, nb, b + r + 8, R[r]);
      top = max(top, l);
      if (lane == ((b + r) & 63u)) run = l;
    }
    if (((b + 8) & 63u) == 0 || b + 8 >= nb) {
      const uint32_t q = (b & ~63u) + lane;
      if (q < nb) lev[q] = run;
    }
  }
  return top;
}
extern "C" __global__ void __launch_bounds__(64) fl_gs_level(FlPtrs P, uint32_t cap) {
  extern __shared__ uint16_t s_last[];
  const uint32_t d = blockIdx.x, lane = threadIdx.x;
  const FlDrum& D = P.drums[d];
  FlState& S = P.st[d];
  const uint32_t n = D.n, nb = D.nb;
  if (n == 0 || nb == 0) return;
  uint32_t* lev = P.glev + D.b0;
  if (n + 2 > cap) {
    for (uint32_t b = lane; b < nb; b += 64) lev[b] = b + 1;
    if (lane == 0) S.gs_levels = nb;
    return;
  }
  const size_t g0 = (size_t)D.b0;
  uint32_t mx = 0;
  for (uint32_t b = lane; b < nb; b += 64) mx = max(mx, max(P.gdn[g0 + b], P.gbn[g0 + b]));
  for (uint32_t i = lane; i < n + 2; i += 64) s_last[i] = 0;
  const bool wide = wave_max_u32(mx) > 64u;
  wave_sync();
  const uint32_t top = wide ? lv_walk<true>(P, s_last, lev, g0, nb, n) : lv_walk<false>(P, s_last, lev, g0, nb, n);
  if (lane == 0) S.gs_levels = top;
}

// The sweep over the balls in index order (one wave per drum); V is the drum's velocity array
// (LDS when it fits, else global memory). The balls' item counts are copied to LDS before a sweep;
// a ball's items and velocity-free sums do not depend on the balls before it, so they are loaded
// two balls ahead (three rotating buffers) and only the velocity gathers, the three velocity sums,
// the scalar update and the kicks stay on the sweep's critical path. A kick adds to the velocity
// the gather read: no other ball writes the drum's velocities in between, and a ball's items are
// distinct particles.
#define GS_MAXB 2048  // balls per drum (the fluid operator's limit)
#define GS_WAVES 8    // waves of the sweep kernel
#define GS_THREADS (64 * GS_WAVES)
// A ball's record for a sweep; its items NS slots of 64 (one slot when every list of the drum's
// sweep holds at most 64 items, else four).
template <int NS>
struct GsBall {
  float4 st;  // drag: m_ent, i_ent, s, s_scale; buoyancy: rho_l, od, -, -
  float2 rb, bv;
  float bom;
  uint32_t ij[NS];
  float4 ir[NS];
};
template <int NS>
__device__ __forceinline__ void gs_load(const FlPtrs& P, uint32_t b0, uint32_t nb, uint32_t b, bool drag,
                                        GsBall<NS>& B) {
  const uint32_t lane = threadIdx.x & 63u;
  const size_t gs = b < nb ? (size_t)b0 + b : (size_t)b0;  // past the last ball: any valid address
  if (drag) {
    B.st = P.gds[gs];
    B.rb = P.gdb[gs];
    B.bom = P.bom[gs];
  } else {
    const float2 q = P.gbs[gs];
    B.st = make_float4(q.x, q.y, 0.0f, 0.0f);
    B.rb = P.bx[gs];
    B.bom = 0.0f;
  }
  B.bv = P.bv[gs];
  // every slot of the ball's item block is loaded (the block is always allocated), so the loads do
  // not wait for the count; slots past it are never used
#pragma unroll
  for (int s = 0; s < NS; s++) {
    const uint32_t k = (uint32_t)s * 64u + lane;
    B.ij[s] = 0;
    B.ir[s] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (k < GS_ITEMS) {
      B.ij[s] = drag ? P.gdj[gs * GS_ITEMS + k] : P.gbj[gs * GS_ITEMS + k];
      B.ir[s] = drag ? P.gdr[gs * GS_ITEMS + k] : P.gbr[gs * GS_ITEMS + k];
    }
  }
}

// The drum's constants of the sweeps, in registers (the loop's stores would otherwise make the
// compiler reload them, and every reload waits for the prefetched loads). The balls' impulses and
// angular impulses are LDS copies during the sweeps: a global store in the loop would make a later
// load into its data registers wait for every load in flight.
struct GsConst {
  float ib, rest, btrue, mass_true, beta_b, dt, bmass, hc, rot_cap, om_abs, area, mass;
  float2* imp;
  float* ang;
  const uint64_t* exp_tab;  // an LDS copy of the expf table
};

// Form drag of ball b (pbf.rs:1824-2011). The three item sums run over chunks of 64 terms in LDS,
// added in item order by lanes 0, 1, 2.
template <bool LDSV, int NS>
__device__ __forceinline__ void gs_drag_ball(const GsConst& C, float2* V, float* s_t0, float* s_t1, float* s_t2,
                                             uint32_t nit, uint32_t b, const GsBall<NS>& G) {
  const uint32_t lane = threadIdx.x & 63u;
  const float m_ent = G.st.x, i_ent = G.st.y;
  if (nit == 0 || !(m_ent > 1e-12f)) return;
  float2 vj[NS];
#pragma unroll
  for (int s = 0; s < NS; s++) vj[s] = V[(uint32_t)s * 64u + lane < nit ? G.ij[s] : 0u];
  float acc = 0.0f;
  const float* mine = lane == 0 ? s_t0 : lane == 1 ? s_t1 : s_t2;
#pragma unroll
  for (int s = 0; s < NS; s++) {
    if ((uint32_t)s * 64u < nit) {
      const float w = G.ir[s].w;
      s_t0[lane] = w * vj[s].x;
      s_t1[lane] = w * vj[s].y;
      s_t2[lane] = w * (G.ir[s].x * vj[s].y - G.ir[s].y * vj[s].x);
      wave_sync();
      if (lane < 3) acc = ordered_sum(mine, min(64u, nit - (uint32_t)s * 64u), acc);
      wave_sync();
    }
  }
  const float vsx = rdl(acc, 0), vsy = rdl(acc, 1), l_sum = rdl(acc, 2);
  const float rbx = G.rb.x, rby = G.rb.y, vbx = vsx / m_ent, vby = vsy / m_ent;
  const float vrx = vbx - G.bv.x, vry = vby - G.bv.y;
  const float sp = sqrtf(vrx * vrx + vry * vry);
  const float itf = (sp > 1e-6f && C.rest > 0.0f) ? 0.5f * C.rest * 1.0f * (2.0f * C.btrue) * sp / C.mass_true : 0.0f;
  const float bt = C.beta_b * (1.0f - mgm::ref_expf_t(-C.dt * (0.0f + itf), C.exp_tab));
  const float a_lin = bt * m_ent / (C.bmass + m_ent);
  const float dvbx = a_lin * vrx, dvby = a_lin * vry;
  float dwb = 0.0f, c_rot = 0.0f, pbx = 0.0f, pby = 0.0f;
  if (i_ent > 1e-12f) {
    const float om_bar = l_sum / i_ent;
    const float a_rot = bt * i_ent / (C.ib + i_ent);
    const float cand = a_rot * (om_bar - G.bom);
    pbx = -rby;
    pby = rbx;
    const float s = G.st.z, ss = G.st.w;
    if (fabsf(s) > 0.1f * ss) {
      const float raw = -C.ib * cand / s;
      const float ml = 2.0f * C.hc;
      const float cap = ml > 1e-9f ? C.rot_cap / ml : 0.0f;
      c_rot = fclampf(raw, -cap, cap);
      dwb = -c_rot * s / C.ib;
    }
  }
  const float f = -(C.bmass / m_ent);
#pragma unroll
  for (int s = 0; s < NS; s++) {
    const uint32_t k = (uint32_t)s * 64u + lane;
    if (k < nit) {
      const float p = G.ir[s].z;
      float dx = f * dvbx * p, dy = f * dvby * p;
      if (c_rot != 0.0f) {
        const float cp = c_rot * p;
        dx += cp * (-G.ir[s].y - pbx);
        dy += cp * (G.ir[s].x - pby);
      }
      V[G.ij[s]] = make_float2(vj[s].x + dx, vj[s].y + dy);
    }
  }
  if (lane == 0) {
    const float2 ib = C.imp[b];
    C.imp[b] = make_float2(ib.x + C.bmass * dvbx, ib.y + C.bmass * dvby);
    C.ang[b] = C.ang[b] + (C.ib * dwb + C.bmass * (rbx * dvby - rby * dvbx));
  }
  if (!LDSV) __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  wave_sync();
}

// Buoyancy of ball b (pbf.rs:2035-2122); the item sum runs over chunks of 64 terms in LDS, added
// in item order by lane 0.
template <bool LDSV, int NS>
__device__ __forceinline__ void gs_buoy_ball(const GsConst& C, float2* V, float* s_t0, uint32_t nw, uint32_t b,
                                             const GsBall<NS>& G) {
  const uint32_t lane = threadIdx.x & 63u;
  const float rho_l = G.st.x, od = G.st.y;
  if (nw == 0 || !(rho_l > 0.0f)) return;
  float2 vj[NS];
#pragma unroll
  for (int s = 0; s < NS; s++) vj[s] = V[(uint32_t)s * 64u + lane < nw ? G.ij[s] : 0u];
  float acc = 0.0f;
#pragma unroll
  for (int s = 0; s < NS; s++) {
    if ((uint32_t)s * 64u < nw) {
      float term = 0.0f;
      if (G.ir[s].w != 0.0f) {
        const float dx = G.ir[s].y, dy = G.ir[s].z;
        term = G.ir[s].x * (dx * vj[s].y - dy * vj[s].x) / (dx * dx + dy * dy);
      }
      s_t0[lane] = term;
      wave_sync();
      if (lane == 0) acc = ordered_sum(s_t0, min(64u, nw - (uint32_t)s * 64u), acc);
      wave_sync();
    }
  }
  // an item on the axis adds nothing to `on` in the reference (it is skipped); its +0.0 term
  // leaves the running sum unchanged unless the sum is -0.0, which a sum starting at +0.0 never is
  const float on = rdl(acc, 0);
  const float2 c = G.rb;
  const float rho_eff = fminf(rho_l, C.rest);
  const float om_l = od > 0.0f ? fclampf(on / od, -C.om_abs, C.om_abs) : 0.0f;
  const float o2 = om_l * om_l;
  const float gex = 0.0f + o2 * c.x, gey = FL_G + o2 * c.y;
  const float sc = (-rho_eff) * C.area;
  const float ipx = sc * gex * C.dt, ipy = sc * gey * C.dt;
#pragma unroll
  for (int s = 0; s < NS; s++) {
    const uint32_t k = (uint32_t)s * 64u + lane;
    if (k < nw) {
      const float frac = G.ir[s].x / rho_l;
      V[G.ij[s]] = make_float2(vj[s].x + -(ipx * frac) / C.mass, vj[s].y + -(ipy * frac) / C.mass);
    }
  }
  if (lane == 0) {
    const float2 ib = C.imp[b];
    C.imp[b] = make_float2(ib.x + ipx, ib.y + ipy);
  }
  if (!LDSV) __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  wave_sync();
}

// One sweep, level by level: wave w takes entries w, w + GS_WAVES, ... of the level-ordered list
// and meets the other waves at a barrier after every level (L barriers in all, whichever levels
// hold its entries). Its balls' records are loaded K entries ahead (a register ring, rotated by
// the unrolled loop), also across levels: a record holds nothing the sweep writes before the
// ball's own turn.
template <bool LDSV, bool DRAG, int NS, int K>
__device__ inline void gs_sweep(const FlPtrs& P, const FlDrum& D, float2* V, float* s_t0, float* s_t1, float* s_t2,
                                const uint32_t* s_cnt, const uint32_t* s_list, const uint32_t* s_end, uint32_t L,
                                float2* s_imp, float* s_ang, const uint64_t* s_exp) {
  GsConst C;
  C.exp_tab = s_exp;
  C.ib = D.binertia;
  C.rest = D.rest;
  C.btrue = D.btrue;
  C.mass_true = D.mass_true;
  C.beta_b = D.beta_b;
  C.dt = D.dt;
  C.bmass = D.bmass;
  C.hc = D.hc;
  C.rot_cap = D.rot_cap;
  C.om_abs = fabsf(D.omega);
  C.area = D.area;
  C.mass = D.mass;
  C.imp = s_imp;
  C.ang = s_ang;
  const uint32_t nb = D.nb, b0 = D.b0, w = threadIdx.x >> 6;
  GsBall<NS> R[K];
#pragma unroll
  for (int k = 0; k < K; k++) {
    const uint32_t e = w + (uint32_t)k * GS_WAVES;
    gs_load<NS>(P, b0, nb, e < nb ? s_list[e] : nb, DRAG, R[k]);
  }
  uint32_t l = 1;
  for (uint32_t e0 = w; e0 < nb; e0 += K * GS_WAVES) {
#pragma unroll
    for (int k = 0; k < K; k++) {
      const uint32_t e = e0 + (uint32_t)k * GS_WAVES;
      if (e < nb) {
        while (e >= s_end[l]) {
          __syncthreads();  // the level's velocity updates before the next level reads them
          l++;
        }
        const uint32_t b = s_list[e];
        if (DRAG) gs_drag_ball<LDSV, NS>(C, V, s_t0, s_t1, s_t2, s_cnt[b], b, R[k]);
        else gs_buoy_ball<LDSV, NS>(C, V, s_t0, s_cnt[b], b, R[k]);
      }
      // every slot reloads on every pass (past the end: any valid record), so that the number of
      // loads issued after each slot's is the same on every path
      const uint32_t en = e + K * GS_WAVES;
      gs_load<NS>(P, b0, nb, en < nb ? s_list[en] : nb, DRAG, R[k]);
    }
  }
  for (; l <= L; l++) __syncthreads();
}

__device__ inline double gs_ke(const float2* V, uint32_t n, double* s_sum) {
  const uint32_t t = threadIdx.x, lane = t & 63u, wv = t >> 6;
  double s = 0.0;
  for (uint32_t i = t; i < n; i += GS_THREADS) {
    const float2 q = V[i];
    s += (double)(q.x * q.x + q.y * q.y);
  }
  s = wave_sum_d(s);
  if (lane == 0) s_sum[wv] = s;
  __syncthreads();
  double a = 0.0;
  for (uint32_t q = 0; q < GS_WAVES; q++) a += s_sum[q];
  __syncthreads();
  return a;
}

// The sweeps, level by level (GS_WAVES waves). Dynamic LDS: the drum's velocities (lds_cap
// entries, used when the drum's fit) and the balls' impulses and angular impulses (bcap entries).
extern "C" __global__ void __launch_bounds__(GS_THREADS) fl_gs(FlPtrs P, uint32_t lds_cap, uint32_t bcap) {
  extern __shared__ float2 s_v[];
  __shared__ __align__(16) float s_t[GS_WAVES][3][64];
  __shared__ uint32_t s_cnt[GS_MAXB];        // the current sweep's item count of every ball
  __shared__ uint32_t s_list[GS_MAXB];       // the balls in level order
  __shared__ uint32_t s_end[GS_MAXB + 1];    // s_end[l]: one past the last entry of level l
  __shared__ double s_sum[GS_WAVES];
  __shared__ uint32_t s_wide[2];             // the longest drag and buoyancy lists
  __shared__ uint64_t s_exp[32];             // the expf table
  const uint32_t d = blockIdx.x, t = threadIdx.x, w = t >> 6, lane = t & 63u;
  const FlDrum& D = P.drums[d];
  FlState& S = P.st[d];
  if (D.n == 0) return;
  const uint32_t nb = D.nb, L = nb ? min(S.gs_levels, nb) : 0u;
  float2* v = P.v + D.f0;
  const bool lds = D.n <= lds_cap;
  float2* s_imp = s_v + lds_cap;
  float* s_ang = reinterpret_cast<float*>(s_imp + bcap);
  if (lds)
    for (uint32_t i = t; i < D.n; i += GS_THREADS) s_v[i] = v[i];
  for (uint32_t b = t; b < nb; b += GS_THREADS) {
    s_imp[b] = P.imp[D.b0 + b];
    s_ang[b] = P.ang[D.b0 + b];
  }
  // the balls bucketed by level: counts, an exclusive scan (one wave, a run per lane), cursors
  for (uint32_t l = t; l <= L; l += GS_THREADS) s_end[l] = 0;
  if (t < 2) s_wide[t] = 0;
  if (t < 32) s_exp[t] = mgm::kExp2Tab[t];
  __syncthreads();
  {
    uint32_t md = 0, mw = 0;
    for (uint32_t b = t; b < nb; b += GS_THREADS) {
      const uint32_t c = P.gdn[D.b0 + b];
      s_cnt[b] = c;
      md = max(md, c);
      mw = max(mw, P.gbn[D.b0 + b]);
    }
    md = wave_max_u32(md);
    mw = wave_max_u32(mw);
    if (lane == 0) {
      atomicMax(&s_wide[0], md);
      atomicMax(&s_wide[1], mw);
    }
  }
  __syncthreads();
  const uint32_t* lev = P.glev + D.b0;
  for (uint32_t b = t; b < nb; b += GS_THREADS) atomicAdd(&s_end[min(lev[b], L)], 1u);
  __syncthreads();
  if (w == 0) {
    const uint32_t k = (L + 1 + 63) / 64, a = min(L + 1, lane * k), z = min(L + 1, a + k);
    uint32_t sum = 0;
    for (uint32_t i = a; i < z; i++) sum += s_end[i];
    uint32_t x = sum;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(x, o, 64);
      if (lane >= (uint32_t)o) x += y;
    }
    uint32_t run = x - sum;
    for (uint32_t i = a; i < z; i++) {
      const uint32_t c = s_end[i];
      s_end[i] = run;
      run += c;
    }
  }
  __syncthreads();
  for (uint32_t b = t; b < nb; b += GS_THREADS) s_list[atomicAdd(&s_end[min(lev[b], L)], 1u)] = b;
  __syncthreads();
  float* st0 = s_t[w][0];
  float* st1 = s_t[w][1];
  float* st2 = s_t[w][2];
  if (nb > 0) {
    if (s_wide[0] > 64u) {
      if (lds) gs_sweep<true, true, 4, 1>(P, D, s_v, st0, st1, st2, s_cnt, s_list, s_end, L, s_imp, s_ang, s_exp);
      else gs_sweep<false, true, 4, 1>(P, D, v, st0, st1, st2, s_cnt, s_list, s_end, L, s_imp, s_ang, s_exp);
    } else {
      if (lds) gs_sweep<true, true, 1, GS_RING>(P, D, s_v, st0, st1, st2, s_cnt, s_list, s_end, L, s_imp, s_ang, s_exp);
      else gs_sweep<false, true, 1, GS_RING>(P, D, v, st0, st1, st2, s_cnt, s_list, s_end, L, s_imp, s_ang, s_exp);
    }
  }
  __syncthreads();
  const double k1 = gs_ke(lds ? s_v : v, D.n, s_sum);
  for (uint32_t b = t; b < nb; b += GS_THREADS) s_cnt[b] = P.gbn[D.b0 + b];
  __syncthreads();
  if (nb > 0 && D.buoy_on) {
    if (s_wide[1] > 64u) {
      if (lds) gs_sweep<true, false, 4, 1>(P, D, s_v, st0, st1, st2, s_cnt, s_list, s_end, L, s_imp, s_ang, s_exp);
      else gs_sweep<false, false, 4, 1>(P, D, v, st0, st1, st2, s_cnt, s_list, s_end, L, s_imp, s_ang, s_exp);
    } else {
      if (lds) gs_sweep<true, false, 1, GS_RING>(P, D, s_v, st0, st1, st2, s_cnt, s_list, s_end, L, s_imp, s_ang, s_exp);
      else gs_sweep<false, false, 1, GS_RING>(P, D, v, st0, st1, st2, s_cnt, s_list, s_end, L, s_imp, s_ang, s_exp);
    }
  }
  __syncthreads();
  const double k2 = gs_ke(lds ? s_v : v, D.n, s_sum);
  if (lds)
    for (uint32_t i = t; i < D.n; i += GS_THREADS) v[i] = s_v[i];
  for (uint32_t b = t; b < nb; b += GS_THREADS) {
    P.imp[D.b0 + b] = s_imp[b];
    P.ang[D.b0 + b] = s_ang[b];
  }
  if (t == 0) {
    S.ke_drag = k1;
    S.ke_buoy = k2;
  }
}

// 7a density at the final positions (pbf.rs:2136-2144).
extern "C" __global__ void __launch_bounds__(FL_T) fl_dens2(FlPtrs P) {
  const uint2 tl = P.ftiles[blockIdx.x];
  const FlDrum& D = P.drums[tl.x];
  const uint32_t i = tl.y + threadIdx.x;
  if (i >= D.n) return;
  const float2* x = P.x + D.f0;
  const uint32_t* nb = P.nbr + D.e0;
  const uint32_t k0 = P.noff[D.f0 + tl.x + i], k1 = min(P.noff[D.f0 + tl.x + i + 1], D.ecap);
  const float2 p = x[i];
  float rho = D.poly0m;
  for (uint32_t k = k0; k < k1; k++) {
    const float2 o = x[nb[k]];
    const float dx = p.x - o.x, dy = p.y - o.y;
    rho += D.mass * poly6(dx * dx + dy * dy, D.h2, D.p6c);
  }
  P.dens[D.f0 + i] = rho;
}

// shear_rates (pbf.rs:821-849) of particle i over its list.
__device__ inline float shear_rate(const FlDrum& D, const float2* x, const float2* v, const float* dens,
                                   const uint32_t* nb, uint32_t k0, uint32_t k1, uint32_t i) {
  const float2 p = x[i], vi = v[i];
  float gxx = 0.0f, gxy = 0.0f, gyx = 0.0f, gyy = 0.0f;
  for (uint32_t k = k0; k < k1; k++) {
    const uint32_t j = nb[k];
    const float2 o = x[j];
    const float2 dd = make_float2(p.x - o.x, p.y - o.y);
    const float r = sqrtf(dd.x * dd.x + dd.y * dd.y);
    const float2 g = spiky(dd, r, D.h, D.spc);
    const float coeff = D.mass / fmaxf(dens[j], 1e-6f);
    const float2 vj = v[j];
    const float dvx = vj.x - vi.x, dvy = vj.y - vi.y;
    gxx += coeff * dvx * g.x;
    gxy += coeff * dvx * g.y;
    gyx += coeff * dvy * g.x;
    gyy += coeff * dvy * g.y;
  }
  const float dxy = 0.5f * (gxy + gyx);
  return sqrtf(2.0f * (gxx * gxx + gyy * gyy + 2.0f * dxy * dxy));
}

// 7b effective viscosities: Bingham (pbf.rs:854-868) from the shear rates, else mu.
extern "C" __global__ void __launch_bounds__(FL_T) fl_mu(FlPtrs P) {
  const uint2 tl = P.ftiles[blockIdx.x];
  const FlDrum& D = P.drums[tl.x];
  const uint32_t i = tl.y + threadIdx.x;
  if (i >= D.n || !(D.mu > 0.0f)) return;
  float mu = D.mu;
  if (D.bingham) {
    const float rate = shear_rate(D, P.x + D.f0, P.v + D.f0, P.dens + D.f0, P.nbr + D.e0, P.noff[D.f0 + tl.x + i],
                                  min(P.noff[D.f0 + tl.x + i + 1], D.ecap), i);
    const float g = __builtin_isfinite(rate) ? rate : 0.0f;
    const float extra = g < 1e-4f * D.g0 ? D.tau_y / D.g0 : D.tau_y * (1.0f - mgm::ref_expf(-g / D.g0)) / g;
    mu = D.mu + extra;
  }
  P.mu_eff[D.f0 + i] = mu;
}

// 7b Morris weights and unit directions of every list entry (pbf.rs:312-344), the boundary
// groups' per-particle factor (pbf.rs:2180-2183) and the fluid rows' CG diagonal (pbf.rs:681-688).
extern "C" __global__ void __launch_bounds__(FL_T) fl_weights(FlPtrs P) {
  const uint2 tl = P.ftiles[blockIdx.x];
  const FlDrum& D = P.drums[tl.x];
  const uint32_t i = tl.y + threadIdx.x;
  if (i >= D.n || !(D.mu > 0.0f)) return;
  P.fcnt[D.f0 + i] = 0;  // boundary groups per particle, counted by fl_groups
  const float2* x = P.x + D.f0;
  const float* dens = P.dens + D.f0;
  const float* mue = P.mu_eff + D.f0;
  const uint32_t* nb = P.nbr + D.e0;
  const uint32_t k0 = P.noff[D.f0 + tl.x + i], k1 = min(P.noff[D.f0 + tl.x + i + 1], D.ecap);
  const float2 p = x[i];
  const float rho_i = fmaxf(dens[i], 1e-6f), mui = mue[i];
  float sw = 0.0f;
  for (uint32_t k = k0; k < k1; k++) {
    const uint32_t j = nb[k];
    const float rho_j = fmaxf(dens[j], 1e-6f);
    const float2 o = x[j];
    const float2 dd = make_float2(p.x - o.x, p.y - o.y);
    const float r2 = dd.x * dd.x + dd.y * dd.y;
    const float r = sqrtf(r2);
    const float2 g = spiky(dd, r, D.h, D.spc);
    const float num = (D.mass * (mui + mue[j]) / (rho_i * rho_j)) * fabsf(dd.x * g.x + dd.y * g.y);
    const float w = num / (r2 + D.eta2);
    P.W[(size_t)D.e0 + k] = w;
    P.Dk[(size_t)D.e0 + k] = r > 1e-9f ? make_float2(dd.x / r, dd.y / r) : make_float2(1.0f, 0.0f);
    sw += w;
  }
  P.ff[D.f0 + i] = 0.5f * (D.mass * 2.0f * mui / (fmaxf(dens[i], 1e-6f) * D.rest));
  P.diagf[D.f0 + i] = D.mass * (1.0f + D.dt * (2.0f * sw));
}

// 7c ball boundary groups (pbf.rs:2159-2232), one wave per ball: every boundary point's
// candidates (the grid over the final positions) merged per fluid particle in order of first
// appearance.
extern "C" __global__ void __launch_bounds__(FL_T) fl_groups(FlPtrs P, const uint2* __restrict__ btiles) {
  const uint2 tl = btiles[blockIdx.x];
  const FlDrum& D = P.drums[tl.x];
  const uint32_t lane = threadIdx.x & 63u, b = tl.y + (threadIdx.x >> 6);
  if (b >= D.nb) return;
  uint32_t* gcnt = P.gcnt + D.b0;
  const float2 bxb = P.bx[D.b0 + b];
  if (!D.grp_on || D.n == 0 || !fin2(bxb)) {
    if (lane == 0) gcnt[b] = 0;
    return;
  }
  const GView g = gview(P.grids + 5 * (size_t)tl.x + 3, P.gstart[3] + D.c0[3], P.gidx[3] + D.f0);
  const float2* x = P.x + D.f0;
  const float* ff = P.ff + D.f0;
  Items<FL_SLOTS> it;
  items_clear<FL_SLOTS>(it);
  uint32_t nit = 0;
  bool over = false;
  for (uint32_t k = 0; k < D.cpb; k++) {
    const float2 o = P.offs[D.off0 + k];
    const float2 bp = make_float2(bxb.x + o.x, bxb.y + o.y);
    const float lx = bp.x - bxb.x, ly = bp.y - bxb.y;
    const float ll = lx * lx + ly * ly;
    Near q;
    near_cells_wave(g, bp, q);
    for (uint32_t c0 = 0; c0 < q.total; c0 += 64) {
      const uint32_t c = c0 + lane;
      bool valid = false;
      uint32_t j = 0;
      float cf = 0.0f;
      if (c < q.total) {
        j = near_at(g, q, c);
        const float2 xj = x[j];
        const float dx = xj.x - bp.x, dy = xj.y - bp.y;
        const float r2 = dx * dx + dy * dy;
        if (!(r2 >= D.g_range2)) {
          const float r = sqrtf(r2);
          if (__builtin_isfinite(r) && !(r >= D.h) && !(r <= 1e-9f)) {
            const float da = fabsf(D.spc * (D.h - r) * (D.h - r) * r);
            cf = ff[j] * da / (r2 + D.eta2);
            valid = cf > 0.0f;
          }
        }
      }
      uint64_t m = __ballot(valid);
      while (m) {
        const int L = first_lane(m);
        m &= m - 1;
        const float cc = rdl(cf, L);
        items_add(it, nit, over, lane, rdlu(j, L), cc, cc * lx, cc * ly, cc * ll);
      }
    }
  }
  if (over && lane == 0) atomicOr(&P.st[tl.x].err, FE_GROUPS);
  nit = min(nit, min(64u * FL_SLOTS, P.gcap));
  const size_t base = ((size_t)D.b0 + b) * P.gcap;
#pragma unroll
  for (int s = 0; s < FL_SLOTS; s++) {
    const uint32_t k = (uint32_t)s * 64u + lane;
    if (k < nit) {
      P.gj[base + k] = it.j[s];
      P.gc[base + k] = make_float4(it.a[s], it.b[s], it.c[s], it.e[s]);
      atomicAdd(&P.fcnt[D.f0 + it.j[s]], 1u);
    }
  }
  if (lane == 0) gcnt[b] = nit;
}

// Squeeze film resolved fraction and coefficient (lubrication.rs:44-115), f64.
__device__ inline double resolved_ratio(double hc) {
  if (hc >= 8.0) return 1.0;
  const double h = fmax(hc, 0.0);
  const double T0[6] = {0.0, 0.5, 1.0, 2.0, 4.0, 8.0}, T1[6] = {0.0, 0.06, 0.25, 0.9, 0.99, 1.0};
#pragma unroll
  for (int i = 0; i < 5; i++)
    if (h <= T0[i + 1]) return T1[i] + (h - T0[i]) / (T0[i + 1] - T0[i]) * (T1[i + 1] - T1[i]);
  return 1.0;
}
__device__ inline double film_coef(const double* f, double h) {
  const double eps = h / f[0];
  return f[3] / (h * sqrt(h)) * (1.0 + f[1] * eps + f[2] * eps * eps);
}

// 7d squeeze links (pbf.rs:544-625), one wave per ball: approaching ball pairs within the reach
// (the partner after the ball; the lanes take 64 candidates at a time and the links keep the
// candidates' order) and the wall, with the velocities after the earlier impulses.
extern "C" __global__ void __launch_bounds__(FL_T) fl_links(FlPtrs P, const uint2* __restrict__ btiles) {
  const uint2 tl = btiles[blockIdx.x];
  const FlDrum& D = P.drums[tl.x];
  const uint32_t lane = threadIdx.x & 63u, i = tl.y + (threadIdx.x >> 6);
  if (i >= D.nb) return;
  uint32_t* lcnt = P.lcnt + D.b0;
  const float2* bx = P.bx + D.b0;
  const float2 xi = bx[i];
  if (!D.links_on || D.n == 0 || !fin2(xi)) {
    if (lane == 0) lcnt[i] = 0;
    return;
  }
  const float2* bv = P.bv + D.b0;
  const float2* imp = P.imp + D.b0;
  const GView g = gview(P.grids + 5 * (size_t)tl.x + 4, P.gstart[4] + D.c0[4], P.gidx[4] + D.b0);
  const float2 vi0 = make_float2(bv[i].x + imp[i].x / D.bmass, bv[i].y + imp[i].y / D.bmass);
  float4* out = P.lrec + D.l0 + (size_t)i * D.lcap;
  uint32_t cnt = 0, inr = 0;
  Near q;
  near_cells_wave(g, xi, q);
  for (uint32_t c0 = 0; c0 < q.total; c0 += 64) {
    const uint32_t c = c0 + lane;
    bool in = false, emit = false;
    uint32_t j = 0;
    float4 rec = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (c < q.total) {
      j = near_at(g, q, c);
      if (j > i) {
        const float2 xj = bx[j];
        const float ex = xj.x - xi.x, ey = xj.y - xi.y;
        const float e2 = ex * ex + ey * ey;
        if (!(e2 >= D.rng2 || !(__builtin_isfinite(ex) && __builtin_isfinite(ey)))) {
          const float dist = sqrtf(e2);
          const float gap = dist - 2.0f * D.br;
          if (!(gap >= D.reach || dist < 1e-9f)) {
            in = true;
            const float nx = ex / dist, ny = ey / dist;
            const float2 vj0 = make_float2(bv[j].x + imp[j].x / D.bmass, bv[j].y + imp[j].y / D.bmass);
            if (!(nx * (vi0.x - vj0.x) + ny * (vi0.y - vj0.y) <= 0.0f)) {
              emit = true;
              const double wgt = 1.0 - resolved_ratio((double)gap / D.dxd);
              const double hh = fmax((double)gap, D.a2);
              const double cf = wgt * film_coef(D.sqp, hh);
              rec = make_float4(nx, ny, fminf(1.0f * (float)cf, D.c_max), __uint_as_float(j));
            }
          }
        }
      }
    }
    const uint64_t mi = __ballot(in), me = __ballot(emit);
    inr += (uint32_t)__popcll(mi);
    if (emit) {
      const uint32_t pos = cnt + (uint32_t)__popcll(me & ((1ull << lane) - 1ull));
      if (pos < D.lcap) {
        out[pos] = rec;
        atomicAdd(&P.locnt[D.b0 + j], 1u);
      }
    }
    cnt += (uint32_t)__popcll(me);
  }
  if (lane != 0) return;
  const float dist = sqrtf(xi.x * xi.x + xi.y * xi.y);
  const float gap = D.radius - D.br - dist;
  if (gap < D.reach && dist > 1e-9f) {
    inr++;
    const float nx = xi.x / dist, ny = xi.y / dist;
    if (!(nx * vi0.x + ny * vi0.y <= 0.0f)) {
      const double wgt = 1.0 - resolved_ratio((double)gap / D.dxd);
      const double hh = fmax((double)gap, D.a2);
      const double cf = wgt * film_coef(D.sqw, hh);
      if (cnt < D.lcap) out[cnt] = make_float4(nx, ny, fminf(1.0f * (float)cf, D.c_max), __uint_as_float(FL_NO));
      cnt++;
    }
  }
  if (cnt > D.lcap) atomicOr(&P.st[tl.x].err, FE_LINKS);
  lcnt[i] = min(cnt, D.lcap);
  if (inr) atomicAdd(&P.st[tl.x].in_range, inr);
}

// The viscosity solve's form (pbf.rs:2244-2299): none (mu = 0), the fluid alone when no boundary
// group and no ball pair or wall is within the squeeze reach, else jointly with the balls. For the
// joint form each fluid row lists the groups that hold its particle (ball, then slot order) and
// each ball the links that name it as partner (ball, then slot order): fl_groups and fl_links
// count them, fl_cg_prep forms the ranges, fl_cg_scatter lists them and fl_cg_setup orders each
// row's list.
extern "C" __global__ void __launch_bounds__(FL_T) fl_cg_prep(FlPtrs P) {
  __shared__ uint32_t su[FL_T];
  const uint32_t d = blockIdx.x, t = threadIdx.x;
  const FlDrum& D = P.drums[d];
  FlState& S = P.st[d];
  if (D.n == 0 || !(D.mu > 0.0f)) return;
  const uint32_t nb = D.nb;
  uint32_t gs = 0, ls = 0;
  for (uint32_t b = t; b < nb; b += FL_T) {
    gs += P.gcnt[D.b0 + b];
    ls += P.lcnt[D.b0 + b];
  }
  uint32_t gt, lt;
  wg_exscan(gs, su, &gt);
  wg_exscan(ls, su, &lt);
  const uint32_t mode = (gt == 0 && S.in_range == 0) ? 1u : 2u;
  if (t == 0) {
    S.n_groups = gt;
    S.n_links = lt;
    S.mode = mode;
    S.done = 0;
    S.it = 0;
    S.iters = 0;
    S.maxit = D.force ? D.forced_cg : (mode == 1 ? 50u : 200u);
  }
  if (mode == 1) return;
  uint32_t* off = P.fgoff + D.f0 + d;
  wg_scan_array(P.fcnt + D.f0, off, D.n, su);
  for (uint32_t i = t; i < D.n; i += FL_T) P.fcur[D.f0 + i] = off[i];
  uint32_t* lo = P.looff + D.b0 + d;
  wg_scan_array(P.locnt + D.b0, lo, nb, su);
  for (uint32_t b = t; b < nb; b += FL_T) P.locur[D.b0 + b] = lo[b];
}

extern "C" __global__ void __launch_bounds__(FL_T) fl_cg_scatter(FlPtrs P, const uint2* __restrict__ btiles) {
  const uint2 tl = btiles[blockIdx.x];
  const FlDrum& D = P.drums[tl.x];
  const uint32_t lane = threadIdx.x & 63u, b = tl.y + (threadIdx.x >> 6);
  if (b >= D.nb || !(D.mu > 0.0f) || P.st[tl.x].mode != 2) return;
  const uint32_t ng = P.gcnt[D.b0 + b];
  uint32_t* ids = P.fgids + (size_t)D.b0 * P.gcap;
  const uint32_t* gj = P.gj + (size_t)D.b0 * P.gcap;
  for (uint32_t k = lane; k < ng; k += 64) {
    const uint32_t id = b * P.gcap + k;
    ids[atomicAdd(&P.fcur[D.f0 + gj[id]], 1u)] = id;
  }
  const uint32_t nl = P.lcnt[D.b0 + b];
  uint32_t* lids = P.loids + D.l0;
  for (uint32_t k = lane; k < nl; k += 64) {
    const uint32_t o = __float_as_uint(P.lrec[D.l0 + (size_t)b * D.lcap + k].w);
    if (o != FL_NO) lids[atomicAdd(&P.locur[D.b0 + o], 1u)] = b * D.lcap + k;
  }
}

// The joint solve's ball rows: each ball's list of links naming it as partner in link order, one
// wave per ball.
extern "C" __global__ void __launch_bounds__(FL_T) fl_cg_order(FlPtrs P, const uint2* __restrict__ btiles) {
  __shared__ uint32_t sid[4][FL_WCAP];
  const uint2 tl = btiles[blockIdx.x];
  const FlDrum& D = P.drums[tl.x];
  const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6, b = tl.y + w;
  if (b >= D.nb || !(D.mu > 0.0f) || P.st[tl.x].mode != 2) return;
  const uint32_t* lo = P.looff + D.b0 + tl.x;
  wave_sort(P.loids + D.l0, lo[b], lo[b + 1], sid[w], lane);
}

// Rows of a drum's viscosity system: fluid i < n, ball linear n + b, ball angular n + nb + b.
__device__ inline uint32_t cg_len(const FlDrum& D, uint32_t mode) { return mode == 2 ? D.n + 2 * D.nb : D.n; }

// The initial guess, right-hand side and Jacobi diagonal (pbf.rs:651-699).
extern "C" __global__ void __launch_bounds__(FL_T) fl_cg_setup(FlPtrs P) {
  const uint2 tl = P.rtiles[blockIdx.x];
  const FlDrum& D = P.drums[tl.x];
  const FlState& S = P.st[tl.x];
  const uint32_t r = tl.y + threadIdx.x, mode = S.mode;
  if (mode == 0 || S.done || r >= cg_len(D, mode)) return;
  const size_t ro = (size_t)D.f0 + 2 * (size_t)D.b0 + r;
  const float kd = D.dt * D.mass;
  if (r < D.n) {
    const float2 vi = P.v[D.f0 + r];
    if (mode == 1) {
      P.X[ro] = vi;
      P.B[ro] = vi;
      return;
    }
    float dg = P.diagf[D.f0 + r];
    const uint32_t* off = P.fgoff + D.f0 + tl.x;
    uint32_t* ids = P.fgids + (size_t)D.b0 * P.gcap;
    const float4* gc = P.gc + (size_t)D.b0 * P.gcap;
    const uint32_t s0 = off[r], s1 = off[r + 1];
    sort_lane(ids, s0, s1);
    for (uint32_t a = s0; a < s1; a++) dg += kd * gc[ids[a]].x;
    P.inv[ro] = 1.0f / fmaxf(dg, 1e-20f);
    float2 x0 = vi;
    if (D.has_guess) {
      const float2 gu = P.guess[D.f0 + r];
      x0 = make_float2(vi.x + gu.x, vi.y + gu.y);
    }
    P.X[ro] = x0;
    P.B[ro] = make_float2(D.mass * vi.x, D.mass * vi.y);
    return;
  }
  // ball rows: fl_cg_balls
}

// The joint solve's ball rows, one wave per ball (the rows' sums run over the ball's boundary
// group items and squeeze links). The lanes form 64 terms at a time into LDS; lanes 0, 1, 2 then
// add them to the row sums in order (a subtracted term stored negated: a - t is a + (-t)).
// what = 0: initial guess, right-hand side and Jacobi diagonal (pbf.rs:651-699); 1, 2: A times X
// (before the iterations) or P (pbf.rs:700-744) into AP.
extern "C" __global__ void __launch_bounds__(FL_T) fl_cg_balls(FlPtrs P, const uint2* __restrict__ btiles, int what) {
  __shared__ __align__(16) float s_term[FL_T / 64][3][64];
  const uint2 tl = btiles[blockIdx.x];
  const FlDrum& D = P.drums[tl.x];
  const FlState& S = P.st[tl.x];
  const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6, b = tl.y + wv;
  if (b >= D.nb || S.mode != 2 || S.done) return;
  float* t0 = s_term[wv][0];
  float* t1 = s_term[wv][1];
  float* t2 = s_term[wv][2];
  float* mine = s_term[wv][lane < 3 ? lane : 0];
  const uint32_t n = D.n, nb = D.nb;
  const size_t rb = (size_t)D.f0 + 2 * (size_t)D.b0;
  const size_t gb = ((size_t)D.b0 + b) * P.gcap;
  const uint32_t ng = P.gcnt[D.b0 + b];
  const uint32_t* lo = P.looff + D.b0 + tl.x;
  const uint32_t* lids = P.loids + D.l0;
  const uint32_t i0 = lo[b], i1 = lo[b + 1];
  const uint32_t nl = P.lcnt[D.b0 + b];
  const float4* lrec = P.lrec + D.l0;
  const float kd = D.dt * D.mass;
  // lane q < 3 adds its quantity's terms of the chunk (lim of them; quantities >= nq take none)
  auto add_chunk = [&](float& acc, uint32_t lim, uint32_t nq) {
    wave_sync();
    if (lane < nq) acc = ordered_sum(mine, lim, acc);
    wave_sync();
  };
  if (what == 0) {
    float acc = lane == 0 ? D.bmass : D.binertia;  // lane 0: the linear diagonal, lane 1: the angular
    for (uint32_t c0 = 0; c0 < ng; c0 += 64) {
      float tg = 0.0f, ta = 0.0f;
      if (c0 + lane < ng) {
        const float4 q = P.gc[gb + c0 + lane];
        tg = kd * q.x;
        ta = kd * q.w;
      }
      t0[lane] = tg;
      t1[lane] = ta;
      add_chunk(acc, min(64u, ng - c0), 2);
    }
    for (uint32_t c0 = i0; c0 < i1; c0 += 64) {  // links of earlier balls to this one
      t0[lane] = c0 + lane < i1 ? D.dt * lrec[lids[c0 + lane]].z : 0.0f;
      add_chunk(acc, min(64u, i1 - c0), 1);
    }
    for (uint32_t c0 = 0; c0 < nl; c0 += 64) {
      t0[lane] = c0 + lane < nl ? D.dt * lrec[(size_t)b * D.lcap + c0 + lane].z : 0.0f;
      add_chunk(acc, min(64u, nl - c0), 1);
    }
    const float dl = rdl(acc, 0), da = rdl(acc, 1);
    if (lane == 0) {
      P.inv[rb + n + b] = 1.0f / fmaxf(dl, 1e-20f);
      const float2 bvb = P.bv[D.b0 + b], ib = P.imp[D.b0 + b];
      const float2 v0 = make_float2(bvb.x + ib.x / D.bmass, bvb.y + ib.y / D.bmass);
      P.X[rb + n + b] = v0;
      P.B[rb + n + b] = make_float2(D.bmass * v0.x, D.bmass * v0.y);
      P.inv[rb + n + nb + b] = 1.0f / fmaxf(da, 1e-20f);
      const float w0 = P.bom[D.b0 + b] + P.ang[D.b0 + b] / D.binertia;
      P.X[rb + n + nb + b] = make_float2(w0, 0.0f);
      P.B[rb + n + nb + b] = make_float2(D.binertia * w0, 0.0f);
    }
    return;
  }
  const float2* V = (what == 2 ? P.P : P.X) + rb;
  const float2 vb = V[n + b];
  const float w = V[n + nb + b].x;
  // lane 0: the linear row's x, lane 1: its y, lane 2: the angular row
  float acc = lane == 0 ? D.bmass * vb.x : lane == 1 ? D.bmass * vb.y : D.binertia * w;
  for (uint32_t c0 = 0; c0 < ng; c0 += 64) {
    float tx = 0.0f, ty = 0.0f, ta = 0.0f;
    if (c0 + lane < ng) {
      const float4 q = P.gc[gb + c0 + lane];
      const float2 vf = V[P.gj[gb + c0 + lane]];
      const float rx = vf.x - vb.x, ry = vf.y - vb.y;
      tx = kd * (q.x * rx - w * -q.z);
      ty = kd * (q.x * ry - w * q.y);
      ta = kd * ((q.y * ry - q.z * rx) - w * q.w);
    }
    t0[lane] = -tx;
    t1[lane] = -ty;
    t2[lane] = -ta;
    add_chunk(acc, min(64u, ng - c0), 3);
  }
  for (uint32_t c0 = i0; c0 < i1; c0 += 64) {  // links of earlier balls to this one
    float tx = 0.0f, ty = 0.0f;
    if (c0 + lane < i1) {
      const uint32_t id = lids[c0 + lane];
      const float4 l = lrec[id];
      const float2 va = V[n + id / D.lcap];
      const float rx = va.x - vb.x, ry = va.y - vb.y;
      const float f = D.dt * l.z * (l.x * rx + l.y * ry);
      tx = f * l.x;
      ty = f * l.y;
    }
    t0[lane] = -tx;
    t1[lane] = -ty;
    add_chunk(acc, min(64u, i1 - c0), 2);
  }
  for (uint32_t c0 = 0; c0 < nl; c0 += 64) {
    float tx = 0.0f, ty = 0.0f;
    if (c0 + lane < nl) {
      const float4 l = lrec[(size_t)b * D.lcap + c0 + lane];
      const uint32_t o = __float_as_uint(l.w);
      float rx = vb.x, ry = vb.y;
      if (o != FL_NO) {
        const float2 vo = V[n + o];
        rx = rx - vo.x;
        ry = ry - vo.y;
      }
      const float f = D.dt * l.z * (l.x * rx + l.y * ry);
      tx = f * l.x;
      ty = f * l.y;
    }
    t0[lane] = tx;
    t1[lane] = ty;
    add_chunk(acc, min(64u, nl - c0), 2);
  }
  const float lx = rdl(acc, 0), ly = rdl(acc, 1), ax = rdl(acc, 2);
  if (lane == 0) {
    P.AP[rb + n + b] = make_float2(lx, ly);
    P.AP[rb + n + nb + b] = make_float2(ax, 0.0f);
  }
}

// A (times) src for every row (pbf.rs:359-377 Laplacian; fluid only: v + dt L v, pbf.rs:439-441;
// joint: pbf.rs:700-744), into AP; with src = P (iterations) the tile's f64 p.Ap partial.
extern "C" __global__ void __launch_bounds__(FL_T) fl_cg_apply(FlPtrs P, int iter) {
  __shared__ double sd[FL_T];
  const uint2 tl = P.rtiles[blockIdx.x];
  const FlDrum& D = P.drums[tl.x];
  const FlState& S = P.st[tl.x];
  const uint32_t mode = S.mode;
  if (mode == 0 || S.done) return;
  const uint32_t r = tl.y + threadIdx.x;
  const size_t rb = (size_t)D.f0 + 2 * (size_t)D.b0;
  const float2* V = (iter ? P.P : P.X) + rb;
  double pa = 0.0;
  if (r < cg_len(D, mode)) {
    const float kd = D.dt * D.mass;
    const uint32_t n = D.n, nb = D.nb;
    float ox, oy;
    if (r < n) {
      const float2 vi = V[r];
      const uint32_t* nbl = P.nbr + D.e0;
      const uint32_t k0 = P.noff[D.f0 + tl.x + r], k1 = min(P.noff[D.f0 + tl.x + r + 1], D.ecap);
      float sx = 0.0f, sy = 0.0f;
      for (uint32_t k = k0; k < k1; k++) {
        const float2 vj = V[nbl[k]];
        const float2 e = P.Dk[(size_t)D.e0 + k];
        const float f = 4.0f * P.W[(size_t)D.e0 + k] * (e.x * (vi.x - vj.x) + e.y * (vi.y - vj.y));
        sx += f * e.x;
        sy += f * e.y;
      }
      if (mode == 1) {
        ox = vi.x + D.dt * sx;
        oy = vi.y + D.dt * sy;
      } else {
        ox = D.mass * (vi.x + D.dt * sx);
        oy = D.mass * (vi.y + D.dt * sy);
        const uint32_t* off = P.fgoff + D.f0 + tl.x;
        const uint32_t* ids = P.fgids + (size_t)D.b0 * P.gcap;
        const float4* gc = P.gc + (size_t)D.b0 * P.gcap;
        for (uint32_t a = off[r]; a < off[r + 1]; a++) {
          const uint32_t id = ids[a], bi = id / P.gcap;
          const float4 q = gc[id];
          const float2 vb = V[n + bi];
          const float w = V[n + nb + bi].x;
          const float rx = vi.x - vb.x, ry = vi.y - vb.y;
          ox += kd * (q.x * rx - w * -q.z);
          oy += kd * (q.x * ry - w * q.y);
        }
      }
    } else {  // ball rows, formed by fl_cg_balls
      const float2 o = P.AP[rb + r];
      ox = o.x;
      oy = o.y;
    }
    if (r < n) P.AP[rb + r] = make_float2(ox, oy);
    if (iter) {
      const float2 p = V[r];
      pa = (double)p.x * (double)ox + (double)p.y * (double)oy;
    }
  }
  if (iter) {
    const double s = wg_sum_d(pa, sd);
    if (threadIdx.x == 0) P.cpart[(size_t)blockIdx.x * 4] = s;
  }
}

// r = b - A x0, the first direction and the start-up sums (pbf.rs:436-447, :700-716).
extern "C" __global__ void __launch_bounds__(FL_T) fl_cg_init(FlPtrs P) {
  __shared__ double sd[FL_T];
  const uint2 tl = P.rtiles[blockIdx.x];
  const FlDrum& D = P.drums[tl.x];
  const FlState& S = P.st[tl.x];
  const uint32_t mode = S.mode;
  if (mode == 0 || S.done) return;
  const uint32_t r = tl.y + threadIdx.x;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  if (r < cg_len(D, mode)) {
    const size_t ro = (size_t)D.f0 + 2 * (size_t)D.b0 + r;
    const float2 b = P.B[ro], ax = P.AP[ro];
    const float rx = b.x - ax.x, ry = b.y - ax.y;
    P.R[ro] = make_float2(rx, ry);
    if (mode == 1) {
      P.P[ro] = make_float2(rx, ry);
      s0 = (double)b.x * (double)b.x + (double)b.y * (double)b.y;
      s1 = (double)rx * (double)rx + (double)ry * (double)ry;
    } else {
      const float iv = P.inv[ro];
      const float zx = rx * iv, zy = ry * iv;
      P.P[ro] = make_float2(zx, zy);
      s0 = (double)iv * ((double)b.x * (double)b.x + (double)b.y * (double)b.y);
      s1 = (double)iv * ((double)rx * (double)rx + (double)ry * (double)ry);
      s2 = (double)rx * (double)zx + (double)ry * (double)zy;
    }
  }
  const double a = wg_sum_d(s0, sd), c = wg_sum_d(s1, sd), e = wg_sum_d(s2, sd);
  if (threadIdx.x == 0) {
    double* cp = P.cpart + (size_t)blockIdx.x * 4;
    cp[1] = a;
    cp[2] = c;
    cp[3] = e;
  }
}

// Sum of one slot of the drum's row-tile partials, in tile order.
__device__ inline double cg_sum(const FlPtrs& P, const FlDrum& D, int slot) {
  double s = 0.0;
  for (uint32_t k = 0; k < D.nr; k++) s += P.cpart[((size_t)D.r0 + k) * 4 + slot];
  return s;
}

// Start-up decisions (pbf.rs:433-449, :717-724): the tolerance from |b| and whether to iterate.
extern "C" __global__ void fl_cg_start(FlPtrs P) {
  if (threadIdx.x != 0) return;
  const FlDrum& D = P.drums[blockIdx.x];
  FlState& S = P.st[blockIdx.x];
  if (S.mode == 0 || S.done) return;
  const double a = cg_sum(P, D, 1), c = cg_sum(P, D, 2), e = cg_sum(P, D, 3);
  const double tol = (double)1e-3f;
  if (S.mode == 1) {
    const double bn = sqrt(a);
    if (bn <= 0.0 || !isfinite(bn)) {
      S.done = 1;
      return;
    }
    S.target = tol * bn;
    S.rs = c;
    if (sqrt(c) <= S.target || (D.force && S.maxit == 0)) S.done = 1;
  } else {
    const double bn = sqrt(a);
    if (bn > 0.0 && isfinite(bn) && sqrt(c) > tol * bn && S.maxit > 0) {
      S.target = tol * bn;
      S.rs = e;
    } else {
      S.done = 1;
    }
  }
}

// x += alpha p, r -= alpha Ap (pbf.rs:456-467, :759-770); the tile's residual partials.
extern "C" __global__ void __launch_bounds__(FL_T) fl_cg_upd(FlPtrs P) {
  __shared__ double sd[FL_T];
  const uint2 tl = P.rtiles[blockIdx.x];
  const FlDrum& D = P.drums[tl.x];
  const FlState& S = P.st[tl.x];
  const uint32_t mode = S.mode;
  if (mode == 0 || S.done) return;
  const double pap = cg_sum(P, D, 0);
  if (pap <= 0.0 || !isfinite(pap)) return;
  const float alpha = (float)(S.rs / pap);
  const uint32_t r = tl.y + threadIdx.x;
  double s1 = 0.0, s2 = 0.0;
  if (r < cg_len(D, mode)) {
    const size_t ro = (size_t)D.f0 + 2 * (size_t)D.b0 + r;
    const float2 x = P.X[ro], p = P.P[ro], ap = P.AP[ro], rr = P.R[ro];
    P.X[ro] = make_float2(x.x + alpha * p.x, x.y + alpha * p.y);
    const float rx = rr.x - alpha * ap.x, ry = rr.y - alpha * ap.y;
    P.R[ro] = make_float2(rx, ry);
    if (mode == 1) {
      s1 = (double)rx * (double)rx + (double)ry * (double)ry;
    } else {
      const float iv = P.inv[ro];
      s1 = (double)iv * ((double)rx * (double)rx + (double)ry * (double)ry);
      const float zx = rx * iv, zy = ry * iv;
      s2 = (double)rx * (double)zx + (double)ry * (double)zy;
    }
  }
  const double a = wg_sum_d(s1, sd), c = wg_sum_d(s2, sd);
  if (threadIdx.x == 0) {
    double* cp = P.cpart + (size_t)blockIdx.x * 4;
    cp[1] = a;
    cp[2] = c;
  }
}

// Convergence, the iteration cap and beta (pbf.rs:468-478, :771-783), one lane per drum.
extern "C" __global__ void fl_cg_fin(FlPtrs P, uint32_t nd) {
  const uint32_t d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= nd) return;
  const FlDrum& D = P.drums[d];
  FlState& S = P.st[d];
  if (S.mode == 0 || S.done) return;
  const uint32_t it = S.it + 1;
  S.it = it;
  S.iters = it;
  const double pap = cg_sum(P, D, 0);
  if (pap <= 0.0 || !isfinite(pap)) {
    S.done = 1;
    return;
  }
  const double c = cg_sum(P, D, 1);
  if (!D.force && sqrt(c) <= S.target) {
    S.done = 1;
    return;
  }
  if (D.force && it == S.maxit) {
    S.done = 1;
    return;
  }
  const double nw = S.mode == 1 ? c : cg_sum(P, D, 2);
  S.beta = (float)(nw / S.rs);
  S.rs = nw;
  if (it >= S.maxit) S.done = 1;
}

// p = r + beta p (fluid only) or z + beta p with z = r / diag (pbf.rs:476, :779-781).
extern "C" __global__ void __launch_bounds__(FL_T) fl_cg_dir(FlPtrs P) {
  const uint2 tl = P.rtiles[blockIdx.x];
  const FlDrum& D = P.drums[tl.x];
  const FlState& S = P.st[tl.x];
  const uint32_t mode = S.mode;
  if (mode == 0 || S.done) return;
  const uint32_t r = tl.y + threadIdx.x;
  if (r >= cg_len(D, mode)) return;
  const size_t ro = (size_t)D.f0 + 2 * (size_t)D.b0 + r;
  const float2 rr = P.R[ro], p = P.P[ro];
  const float bt = S.beta;
  if (mode == 1) {
    P.P[ro] = make_float2(rr.x + bt * p.x, rr.y + bt * p.y);
  } else {
    const float iv = P.inv[ro];
    P.P[ro] = make_float2(rr.x * iv + bt * p.x, rr.y * iv + bt * p.y);
  }
}

// The solved fluid velocities and, for the joint solve, the next step's warm start (pbf.rs:2285-
// 2299, :786-791).
extern "C" __global__ void __launch_bounds__(FL_T) fl_cg_out(FlPtrs P) {
  const uint2 tl = P.ftiles[blockIdx.x];
  const FlDrum& D = P.drums[tl.x];
  const uint32_t mode = P.st[tl.x].mode;
  const uint32_t i = tl.y + threadIdx.x;
  if (i >= D.n || mode == 0) return;
  const size_t g = (size_t)D.f0 + i;
  const float2 vn = P.X[(size_t)D.f0 + 2 * (size_t)D.b0 + i];
  if (mode == 2) {
    const float2 vi = P.v[g];
    const float dx = vn.x - vi.x, dy = vn.y - vi.y;
    P.guess[g] = (__builtin_isfinite(dx) && __builtin_isfinite(dy)) ? make_float2(dx, dy) : make_float2(0.0f, 0.0f);
  }
  P.v[g] = vn;
}

// The mean shear rate of the solved velocities (pbf.rs:800-818) and |v|^2 before the clamp.
extern "C" __global__ void __launch_bounds__(FL_T) fl_shear(FlPtrs P) {
  __shared__ double sd[FL_T];
  const uint2 tl = P.ftiles[blockIdx.x];
  const FlDrum& D = P.drums[tl.x];
  const uint32_t i = tl.y + threadIdx.x;
  double s = 0.0, c = 0.0, k = 0.0;
  if (i < D.n) {
    const float g = shear_rate(D, P.x + D.f0, P.v + D.f0, P.dens + D.f0, P.nbr + D.e0, P.noff[D.f0 + tl.x + i],
                               min(P.noff[D.f0 + tl.x + i + 1], D.ecap), i);
    if (__builtin_isfinite(g)) {
      s = (double)g;
      c = 1.0;
    }
    const float2 q = P.v[D.f0 + i];
    k = (double)(q.x * q.x + q.y * q.y);
  }
  const double a = wg_sum_d(s, sd), b = wg_sum_d(c, sd), e = wg_sum_d(k, sd);
  if (threadIdx.x == 0) {
    double* part = P.part + (size_t)blockIdx.x * FL_NP;
    part[P_SH] = a;
    part[P_SHN] = b;
    part[P_KEV] = e;
  }
}

// 7.5 fluid speed clamp (pbf.rs:987-996, :2321).
extern "C" __global__ void __launch_bounds__(FL_T) fl_clamp(FlPtrs P) {
  __shared__ double sd[FL_T];
  const uint2 tl = P.ftiles[blockIdx.x];
  const FlDrum& D = P.drums[tl.x];
  const uint32_t i = tl.y + threadIdx.x;
  double k = 0.0;
  if (i < D.n) {
    const size_t g = (size_t)D.f0 + i;
    float2 q = P.v[g];
    const float sp = sqrtf(q.x * q.x + q.y * q.y);
    if (!__builtin_isfinite(sp)) q = make_float2(0.0f, 0.0f);
    else if (sp > D.vmax) q = make_float2(q.x * (D.vmax / sp), q.y * (D.vmax / sp));
    P.v[g] = q;
    k = (double)(q.x * q.x + q.y * q.y);
  }
  const double a = wg_sum_d(k, sd);
  if (threadIdx.x == 0) P.part[(size_t)blockIdx.x * FL_NP + P_KEC] = a;
}

// The balls' share of the joint solve (pbf.rs:2288-2299) and the per-ball impulse clamp
// (pbf.rs:2330-2345), one lane per ball.
extern "C" __global__ void __launch_bounds__(FL_T) fl_ball_out(FlPtrs P, const uint2* __restrict__ ltiles) {
  const uint2 tl = ltiles[blockIdx.x];
  const FlDrum& D = P.drums[tl.x];
  FlState& S = P.st[tl.x];
  const uint32_t b = tl.y + threadIdx.x;
  if (b >= D.nb || D.n == 0) return;
  float2 ib = P.imp[D.b0 + b];
  float an = P.ang[D.b0 + b];
  if (S.mode == 2) {
    const size_t rb = (size_t)D.f0 + 2 * (size_t)D.b0;
    const float2 bvb = P.bv[D.b0 + b];
    const float2 v0 = make_float2(bvb.x + ib.x / D.bmass, bvb.y + ib.y / D.bmass);
    const float w0 = P.bom[D.b0 + b] + an / D.binertia;
    const float2 vn = P.X[rb + D.n + b];
    const float wn = P.X[rb + D.n + D.nb + b].x;
    const float dx = vn.x - v0.x, dy = vn.y - v0.y, dw = wn - w0;
    if (__builtin_isfinite(dx) && __builtin_isfinite(dy) && __builtin_isfinite(dw)) {
      ib = make_float2(ib.x + D.bmass * dx, ib.y + D.bmass * dy);
      an += D.binertia * dw;
    }
  }
  const float m = sqrtf(ib.x * ib.x + ib.y * ib.y);
  if (m > D.imp_cl && m > 1e-9f) {
    ib = make_float2(ib.x * (D.imp_cl / m), ib.y * (D.imp_cl / m));
    atomicAdd(&S.clamp_hits, 1u);
  }
  an = fclampf(an, -D.ang_cl, D.ang_cl);
  P.imp[D.b0 + b] = ib;
  P.ang[D.b0 + b] = an;
}

// Per drum results of the step (FluidStepStats, pbf.rs:1018-1028 and the step's tail).
struct FlOut {
  double ke0, ke1, ww, wix, wiy, wang, kev, kec, sh, shn;
  float bins[32];
};
extern "C" __global__ void __launch_bounds__(64) fl_stats(FlPtrs P, FlOut* __restrict__ out) {
  const uint32_t d = blockIdx.x, t = threadIdx.x;
  const FlDrum& D = P.drums[d];
  FlOut& o = out[d];
  if (t < FL_NP) {
    double s = 0.0;
    for (uint32_t k = 0; k < D.nt; k++) s += P.part[((size_t)D.t0 + k) * FL_NP + t];
    (&o.ke0)[t] = s;
  }
  if (t < 32) {
    float s = 0.0f;
    for (uint32_t k = 0; k < D.nt; k++) s += P.bins[((size_t)D.t0 + k) * 32 + t];
    o.bins[t] = s;
  }
}

// ---- host -----------------------------------------------------------------------------------------
struct mg_fluid {
  uint32_t nd = 0;
  char* dev = nullptr;
  size_t dev_cap = 0;
  FlPtrs P{};
  FlDrum* d_drums = nullptr;
  FlOut* d_out = nullptr;
  float2 *d_lift = nullptr, *d_offs = nullptr;
  uint2 *d_ft = nullptr, *d_rt = nullptr, *d_wt = nullptr, *d_lt = nullptr;
  bool force = false;
  uint32_t forced_d = 0, forced_cg = 0;
  uint32_t cg_chunk = 8;
  // side streams (non-blocking, default priority) for stages that run beside the Gauss-Seidel sweep
  hipStream_t side[2] = {nullptr, nullptr};
  hipEvent_t fork = nullptr, join[2] = {nullptr, nullptr};
  std::vector<FlState> h_st;
  std::vector<FlOut> h_out;
  std::string msg;
  mgk::Timer tm;
};

namespace {

int fl_fail(mg_fluid* h, char* err, size_t errlen, int rc, const std::string& m) {
  if (h) h->msg = m;
  if (err && errlen) {
    std::strncpy(err, m.c_str(), errlen - 1);
    err[errlen - 1] = 0;
  }
  return rc;
}

// f32::powi / f64::powi as LLVM expands them (binary exponentiation, low bit first).
template <class T> T powi_h(T a, int b) {
  T res = 1;
  bool first = true;
  T cur = a;
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
float poly6_h(float r2, float h) {  // pbf.rs:119-126
  if (r2 >= h * h || r2 < 0.0f) return 0.0f;
  const float h2 = h * h, t = h2 - r2;
  return 4.0f / (FL_PI * powi_h(h, 8)) * t * t * t;
}
uint32_t sat_u32_h(float x) {
  if (std::isnan(x) || x <= 0.0f) return 0;
  if (x >= 4294967295.0f) return UINT32_MAX;
  return (uint32_t)x;
}
// SqueezeFilm::new (lubrication.rs:44-79): (r, c1, c2, prefactor).
void squeeze_film(double k1, double k2, double nu, double* out) {
  const double r = 1.0 / (k1 + k2);
  const double q = powi_h(r, 3) * (powi_h(k1, 3) + powi_h(k2, 3));
  out[0] = r;
  out[1] = 0.788 + 0.265 * q;
  out[2] = std::fmax(0.006 + 0.057 * (q - 0.25), 0.0);
  out[3] = 3.0 * std::sqrt(2.0) * M_PI * nu * std::pow(r, 1.5);
}
// Cells of a grid over points within 1.5 drum radii of the centre.
uint32_t cell_cap(float radius, float cell) {
  const double side = std::floor(3.0 * (double)radius / std::max((double)cell, 1e-12)) + 3.0;
  const double c = side * side;
  return c > (double)(1u << 24) ? (1u << 24) : (uint32_t)c;
}

}  // namespace

extern "C" {

int mg_fluid_create(uint32_t n_drums, mg_fluid** out, char* err, size_t errlen) {
  if (!out || n_drums == 0 || n_drums > 65535) return fl_fail(nullptr, err, errlen, MG_E_ARG, "n_drums must be in 1..65535");
  auto* h = new mg_fluid();
  h->nd = n_drums;
  h->h_st.resize(n_drums);
  h->h_out.resize(n_drums);
  *out = h;
  return MG_OK;
}

void mg_fluid_destroy(mg_fluid* h) {
  if (!h) return;
  if (h->dev) (void)hipFree(h->dev);
  for (int k = 0; k < 2; k++) {
    if (h->side[k]) (void)hipStreamDestroy(h->side[k]);
    if (h->join[k]) (void)hipEventDestroy(h->join[k]);
  }
  if (h->fork) (void)hipEventDestroy(h->fork);
  delete h;
}

int mg_fluid_force_iterations(mg_fluid* h, int on, uint32_t density_passes, uint32_t cg_iterations) {
  if (!h) return MG_E_ARG;
  if (on && density_passes > FL_NPASS) return MG_E_ARG;
  h->force = on != 0;
  h->forced_d = density_passes;
  h->forced_cg = cg_iterations;
  return MG_OK;
}

int mg_fluid_step(mg_fluid* h, const mg_fluid_drum* drums, float* d_x, float* d_v, float* d_guess, const float* d_ball_x,
                  const float* d_ball_v, const float* d_ball_omega, float* d_impulse, float* d_angular,
                  mg_fluid_stats* stats, char* err, size_t errlen) {
  if (!h || !drums || !stats) return fl_fail(h, err, errlen, MG_E_ARG, "null argument");
  const uint32_t nd = h->nd;
  std::vector<FlDrum> hd(nd);
  std::vector<float2> lift, offs;
  std::vector<uint2> ft, rt, wt, lt;
  uint64_t TF = 0, TB = 0, TE = 0, TL = 0, TC[5] = {0, 0, 0, 0, 0};
  uint32_t max_rounds = 1;
  for (uint32_t d = 0; d < nd; d++) {
    const mg_fluid_drum& m = drums[d];
    FlDrum& D = hd[d];
    std::memset(&D, 0, sizeof D);
    if (!(m.h > 0.0f) || !(m.rest_density > 0.0f) || !(m.particle_mass > 0.0f) || !(m.radius_m > 0.0f))
      return fl_fail(h, err, errlen, MG_E_ARG, "drum " + std::to_string(d) + ": h, rest density, particle mass and "
                                               "radius must be positive");
    const uint32_t base = std::max(m.iterations, 1u);
    if (!h->force && base + 6 > FL_NPASS)
      return fl_fail(h, err, errlen, MG_E_ARG, "drum " + std::to_string(d) + ": at most 26 density iterations");
    D.radius = m.radius_m;
    D.omega = m.omega;
    D.nl = m.lifter_count;
    D.lift0 = (uint32_t)lift.size();
    D.rt = m.radius_m - m.lifter_height_m;
    D.bh = m.lifter_base_width_m * 0.5f;
    D.thw = m.lifter_top_width_m * 0.5f;
    const float phase = m.lifter_phase_deg * (float)0.017453292519943295;  // f32::to_radians
    for (uint32_t k = 0; k < m.lifter_count; k++) {
      const float a = phase + (float)k * FL_TAU / (float)m.lifter_count;
      lift.push_back(make_float2(std::sin(a), std::cos(a)));
    }
    const float an = m.drum_angle + m.omega * m.dt;  // pbf.rs:1236
    D.rc1 = std::cos(-an);
    D.rs1 = std::sin(-an);
    D.wc1 = std::cos(an);
    D.ws1 = std::sin(an);
    D.dt = m.dt;
    D.base = base;
    D.force = h->force ? 1u : 0u;
    D.forced_d = h->forced_d;
    D.forced_cg = h->forced_cg;
    D.f0 = (uint32_t)TF;
    D.b0 = (uint32_t)TB;
    const bool active = m.n_fluid > 0 && m.dt > 0.0f;  // pbf.rs:1196-1198
    D.n = active ? m.n_fluid : 0u;
    D.nb = m.n_balls;
    if (m.n_balls > 2048)
      return fl_fail(h, err, errlen, MG_E_ARG, "drum " + std::to_string(d) + ": more than 2048 balls");
    const float hh = m.h, rest = m.rest_density, mass = m.particle_mass;
    D.h = hh;
    D.rest = rest;
    D.mass = mass;
    D.h2 = hh * hh;
    D.mr = mass / rest;
    D.p6c = 4.0f / (FL_PI * powi_h(hh, 8));
    D.spc = -30.0f / (FL_PI * powi_h(hh, 5));
    D.cohc = 35840.0f / (209.0f * FL_PI * powi_h(hh, 8));
    D.h6_64 = powi_h(hh, 6) / 64.0f;
    D.poly0m = mass * poly6_h(0.0f, hh);
    D.coh_acc = (std::fmax(m.surface_tension_n_m, 0.0f) / 0.072f) * 2.0f * 9.81f;
    const float br = m.ball_radius;
    const float dxb = 0.5f * hh;
    const float wett = std::min(std::max(m.wettability, 0.0f), 1.0f);
    D.ha = std::fmin(hh, 2.0f * br);
    D.adc = 35840.0f / (163.0f * FL_PI * powi_h(D.ha, 8));
    D.ad_range2 = D.ha * D.ha * (1.0f + 1e-4f);
    D.ad_inner2 = 0.25f * D.ha * D.ha * (1.0f - 1e-4f);
    D.ad_acc = wett * 2.0f * 9.81f;
    D.beta_w = std::min(std::max(m.wall_no_slip, 0.0f), 1.0f);
    const float mu = std::fmax(m.viscosity_pa_s, 0.0f);
    D.mu = mu;
    D.eta2 = 0.01f * hh * hh;
    D.bingham = (m.rheology == 1 && m.yield_stress_pa > 0.0f) ? 1u : 0u;
    D.tau_y = m.yield_stress_pa;
    D.g0 = D.tau_y / (100.0f * mu);
    D.br = br;
    D.bmass = m.ball_mass;
    D.binertia = m.ball_inertia;
    D.btrue = m.ball_true_radius;
    D.contact_r = br + 0.05f * hh;
    D.bgc = std::fmax(std::fmax(2.0f * br, D.contact_r), 1e-6f);
    D.red = m.ball_mass > 0.0f ? (mass * m.ball_mass) / (mass + m.ball_mass) : mass;
    // boundary points (pbf.rs:201-219): offsets br (cos a, sin a) at a = tau k / count
    uint32_t cpb = 0;
    if (m.n_balls > 0 && br > 0.0f && dxb > 0.0f) {
      cpb = std::max(sat_u32_h(std::round(FL_TAU * br / dxb)), 6u);
      if (cpb > 4096) return fl_fail(h, err, errlen, MG_E_ARG, "drum " + std::to_string(d) + ": too many boundary points per ball");
    }
    D.cpb = cpb;
    D.off0 = (uint32_t)offs.size();
    for (uint32_t k = 0; k < cpb; k++) {
      const float a = FL_TAU * (float)k / (float)cpb;
      offs.push_back(make_float2(std::cos(a) * br, std::sin(a) * br));
    }
    const bool balls = m.n_balls > 0 && br > 0.0f;
    D.adh_on = (wett > 0.0f && balls && dxb > 0.0f && D.ha > 0.0f) ? 1u : 0u;
    D.grp_on = (mu > 0.0f && balls && dxb > 0.0f) ? 1u : 0u;
    D.buoy_on = balls ? 1u : 0u;
    D.links_on = (mu > 0.0f && balls) ? 1u : 0u;
    D.beta_b = std::min(std::max(m.ball_no_slip, 0.0f), 1.0f);
    D.drag_on = (balls && m.ball_true_radius > 0.0f && mu > 1e-6f && D.beta_b > 0.0f) ? 1u : 0u;
    D.vmax = 5.0f * (std::fabs(m.omega) * m.radius_m + std::sqrt(4.0f * 9.81f * m.radius_m));
    D.rot_cap = 0.2f * D.vmax;
    const float rho_ball = m.ball_mass / (FL_PI * br * br);
    D.mass_true = rho_ball * FL_PI * m.ball_true_radius * m.ball_true_radius;
    D.hc = br + hh;
    D.hc2 = D.hc * D.hc;
    D.p6hc = 4.0f / (FL_PI * powi_h(D.hc, 8));
    D.w0 = std::fmax(poly6_h(0.0f, D.hc), 1e-12f);
    D.area = FL_PI * br * br;
    const double a = (double)br, bD = (double)m.radius_m, dxd = (double)dxb, mud = (double)mu;
    D.dxd = dxd;
    D.a2 = 0.02 * a;
    D.reach = (float)(8.0 * dxd);
    D.c_max = 8.0f * m.ball_mass / std::fmax(m.dt, 1e-12f);
    D.g5c = 2.0f * br + D.reach;
    D.rng2 = (2.0f * br + D.reach) * (2.0f * br + D.reach) * (1.0f + 1e-4f);
    D.g_range2 = hh * hh * (1.0f + 1e-4f);
    if (a > 0.0) {
      squeeze_film(1.0 / a, 1.0 / a, mud, D.sqp);
      squeeze_film(1.0 / a, -1.0 / bD, mud, D.sqw);
    }
    D.imp_cl = 60.0f * m.ball_mass * 9.81f * m.dt;
    D.ang_cl = D.imp_cl * br;
    D.has_guess = m.has_guess ? 1u : 0u;
    // workspace regions
    D.e0 = (uint32_t)TE;
    D.ecap = D.n ? 48u * D.n + 64u : 0u;
    const float cells[5] = {hh, D.bgc, std::fmax(D.hc, 1e-6f), hh, D.g5c};
    const bool need[5] = {D.n > 0, D.n > 0 && m.n_balls > 0, D.n > 0 && D.drag_on, D.n > 0 && D.grp_on,
                          D.n > 0 && (D.links_on || D.drag_on || D.buoy_on)};
    for (int q = 0; q < 5; q++) {
      D.c0[q] = (uint32_t)TC[q];
      D.ccap[q] = need[q] ? cell_cap(m.radius_m, cells[q]) : 0u;
      TC[q] += D.ccap[q] + 1;
    }
    D.lcap = 4;
    if (D.links_on && br > 0.0f) {
      const double lr = (2.0 * a + (double)D.reach) / a;
      D.lcap = (uint32_t)std::ceil(0.95 * lr * lr) + 4u;
    }
    D.l0 = (uint32_t)TL;
    TL += (uint64_t)D.lcap * m.n_balls;
    D.t0 = (uint32_t)ft.size();
    for (uint32_t s = 0; s < D.n; s += FL_T) ft.push_back(make_uint2(d, s));
    D.nt = (uint32_t)ft.size() - D.t0;
    D.r0 = (uint32_t)rt.size();
    if (D.n && mu > 0.0f)
      for (uint32_t s = 0; s < D.n + 2 * m.n_balls; s += FL_T) rt.push_back(make_uint2(d, s));
    D.nr = (uint32_t)rt.size() - D.r0;
    if (D.n) {
      for (uint32_t s = 0; s < m.n_balls; s += 4) wt.push_back(make_uint2(d, s));
      for (uint32_t s = 0; s < m.n_balls; s += FL_T) lt.push_back(make_uint2(d, s));
    }
    const uint32_t rounds = h->force ? std::max(h->forced_d, 1u) : base + 6;
    if (D.n) max_rounds = std::max(max_rounds, rounds);
    D.npass = rounds;
    TF += m.n_fluid;
    TB += m.n_balls;
    TE += D.ecap;
    if (TF > 0x7fffffffull || TB > 0x7fffffull || TE > 0xffffffffull || TL > 0xffffffffull)
      return fl_fail(h, err, errlen, MG_E_ARG, "the batch is too large for one workspace");
  }
  if (TF && (!d_x || !d_v)) return fl_fail(h, err, errlen, MG_E_ARG, "null fluid state array");
  if (TB && (!d_ball_x || !d_ball_v || !d_ball_omega || !d_impulse || !d_angular))
    return fl_fail(h, err, errlen, MG_E_ARG, "null ball array");
  bool need_guess = false;
  for (uint32_t d = 0; d < nd; d++) need_guess |= hd[d].n > 0 && hd[d].nr > 0 && hd[d].nb > 0;
  if (need_guess && !d_guess) return fl_fail(h, err, errlen, MG_E_ARG, "null guess array");
  if (lift.empty()) lift.push_back(make_float2(0.0f, 1.0f));
  if (offs.empty()) offs.push_back(make_float2(0.0f, 0.0f));
  const uint64_t TR = TF + 2 * TB;
  const uint32_t GC = 64u * FL_SLOTS;
  // workspace layout
  size_t off = 0;
  std::vector<std::pair<void**, size_t>> bind;
  auto take = [&](void** p, size_t bytes) {
    bind.push_back({p, off});
    off += ((bytes + 255) / 256) * 256 + 256;
  };
  FlPtrs& P = h->P;
  P = FlPtrs{};
  P.acap = GC;
  P.gcap = GC;
  take((void**)&h->d_drums, nd * sizeof(FlDrum));
  take((void**)&P.st, nd * sizeof(FlState));
  take((void**)&h->d_out, nd * sizeof(FlOut));
  take((void**)&h->d_lift, lift.size() * sizeof(float2));
  take((void**)&h->d_offs, offs.size() * sizeof(float2));
  take((void**)&h->d_ft, (ft.size() + 1) * sizeof(uint2));
  take((void**)&h->d_rt, (rt.size() + 1) * sizeof(uint2));
  take((void**)&h->d_wt, (wt.size() + 1) * sizeof(uint2));
  take((void**)&h->d_lt, (lt.size() + 1) * sizeof(uint2));
  take((void**)&P.grids, 5 * nd * sizeof(FlGrid));
  for (int q = 0; q < 5; q++) {
    take((void**)&P.gstart[q], TC[q] * 4);
    take((void**)&P.gcur[q], TC[q] * 4);
    take((void**)&P.gidx[q], (q == 1 || q == 4 ? TB : TF) * 4);
  }
  for (float2** p : {&P.x0, &P.dp, &P.excess, &P.coh, &P.adh}) take((void**)p, TF * 8);
  for (float** p : {&P.dens, &P.lam, &P.mu_eff, &P.ff, &P.diagf}) take((void**)p, TF * 4);
  take((void**)&P.touched, TF * 4);
  take((void**)&P.ncnt, TF * 4);
  take((void**)&P.noff, (TF + nd) * 4);
  take((void**)&P.nbr, TE * 4);
  take((void**)&P.W, TE * 4);
  take((void**)&P.Dk, TE * 8);
  take((void**)&P.pcnt, TF * 4);
  take((void**)&P.prec, TF * FL_PK * 16);
  take((void**)&P.bcnt, TB * 4);
  take((void**)&P.boff, (TB + nd) * 4);
  take((void**)&P.bcur, TB * 4);
  take((void**)&P.pids, TF * FL_PK * 4);
  take((void**)&P.acnt, TB * 4);
  take((void**)&P.aj, TB * GC * 4);
  take((void**)&P.adv, TB * GC * 8);
  take((void**)&P.fcnt, TF * 4);
  take((void**)&P.foff, (TF + nd) * 4);
  take((void**)&P.fcur, TF * 4);
  take((void**)&P.fids, TB * GC * 4);
  take((void**)&P.gcnt, TB * 4);
  take((void**)&P.gj, TB * GC * 4);
  take((void**)&P.gc, TB * GC * 16);
  take((void**)&P.fgoff, (TF + nd) * 4);
  take((void**)&P.fgids, TB * GC * 4);
  take((void**)&P.gdj, TB * GS_ITEMS * 4);
  take((void**)&P.gdr, TB * GS_ITEMS * 16);
  take((void**)&P.gbj, TB * GS_ITEMS * 4);
  take((void**)&P.gbr, TB * GS_ITEMS * 16);
  take((void**)&P.gdn, TB * 4);
  take((void**)&P.gbn, TB * 4);
  take((void**)&P.gds, TB * 16);
  take((void**)&P.gdb, TB * 8);
  take((void**)&P.gbs, TB * 8);
  take((void**)&P.glev, TB * 4);
  take((void**)&P.lcnt, TB * 4);
  take((void**)&P.lrec, TL * 16);
  take((void**)&P.locnt, TB * 4);
  take((void**)&P.looff, (TB + nd) * 4);
  take((void**)&P.locur, TB * 4);
  take((void**)&P.loids, TL * 4);
  for (float2** p : {&P.X, &P.B, &P.R, &P.P, &P.AP}) take((void**)p, TR * 8);
  take((void**)&P.inv, TR * 4);
  take((void**)&P.part, (ft.size() + 1) * FL_NP * 8);
  take((void**)&P.bins, (ft.size() + 1) * 32 * 4);
  take((void**)&P.cpart, (rt.size() + 1) * 4 * 8);
  if (off > h->dev_cap) {
    if (h->dev) (void)hipFree(h->dev);
    h->dev = nullptr;
    h->dev_cap = 0;
    const size_t cap = off + off / 4;
    if (hipMalloc(&h->dev, cap) != hipSuccess) {
      h->dev = nullptr;
      return fl_fail(h, err, errlen, MG_E_DEVICE, "hipMalloc failed for the fluid workspace (" +
                                                      std::to_string(cap >> 20) + " MiB)");
    }
    h->dev_cap = cap;
  }
  for (auto& b : bind) *b.first = h->dev + b.second;
  P.drums = h->d_drums;
  P.lift = h->d_lift;
  P.offs = h->d_offs;
  P.ftiles = h->d_ft;
  P.rtiles = h->d_rt;
  P.x = reinterpret_cast<float2*>(d_x);
  P.v = reinterpret_cast<float2*>(d_v);
  P.guess = reinterpret_cast<float2*>(d_guess);
  P.bx = reinterpret_cast<const float2*>(d_ball_x);
  P.bv = reinterpret_cast<const float2*>(d_ball_v);
  P.bom = d_ball_omega;
  P.imp = reinterpret_cast<float2*>(d_impulse);
  P.ang = d_angular;
  bool ok = hipMemcpy(h->d_drums, hd.data(), nd * sizeof(FlDrum), hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(h->d_lift, lift.data(), lift.size() * sizeof(float2), hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(h->d_offs, offs.data(), offs.size() * sizeof(float2), hipMemcpyHostToDevice) == hipSuccess;
  if (ok && !ft.empty()) ok = hipMemcpy(h->d_ft, ft.data(), ft.size() * sizeof(uint2), hipMemcpyHostToDevice) == hipSuccess;
  if (ok && !rt.empty()) ok = hipMemcpy(h->d_rt, rt.data(), rt.size() * sizeof(uint2), hipMemcpyHostToDevice) == hipSuccess;
  if (ok && !wt.empty()) ok = hipMemcpy(h->d_wt, wt.data(), wt.size() * sizeof(uint2), hipMemcpyHostToDevice) == hipSuccess;
  if (ok && !lt.empty()) ok = hipMemcpy(h->d_lt, lt.data(), lt.size() * sizeof(uint2), hipMemcpyHostToDevice) == hipSuccess;
  if (!ok) return fl_fail(h, err, errlen, MG_E_DEVICE, "HIP copy error");
  const dim3 T(FL_T), G(nd);
  const uint32_t nft = (uint32_t)ft.size(), nrt = (uint32_t)rt.size(), nwt = (uint32_t)wt.size(),
                 nlt = (uint32_t)lt.size();
  if (nft) MG_TIMED_LAUNCH(h->tm, fl_predict, dim3(nft), T, 0, P);
  MG_TIMED_LAUNCH(h->tm, fl_grid_a, G, T, 0, P);
  if (nft) {
    MG_TIMED_LAUNCH(h->tm, fl_nbr, dim3(nft), T, 0, P, 0);
    MG_TIMED_LAUNCH(h->tm, fl_nbr_scan, G, T, 0, P);
    MG_TIMED_LAUNCH(h->tm, fl_nbr, dim3(nft), T, 0, P, 1);
    for (uint32_t r = 0; r < max_rounds; r++) {
      MG_TIMED_LAUNCH(h->tm, fl_dens_a, dim3(nft), T, 0, P, r);
      MG_TIMED_LAUNCH(h->tm, fl_dens_b, dim3(nft), T, 0, P, r);
      MG_TIMED_LAUNCH(h->tm, fl_dens_c, dim3(nft), T, 0, P, r);
    }
    MG_TIMED_LAUNCH(h->tm, fl_push, dim3(nft), T, 0, P);
  }
  MG_TIMED_LAUNCH(h->tm, fl_push_scan, G, T, 0, P);
  if (nft) MG_TIMED_LAUNCH(h->tm, fl_push_scatter, dim3(nft), T, 0, P);
  if (nwt) MG_TIMED_LAUNCH(h->tm, fl_push_fin, dim3(nwt), T, 0, P, (const uint2*)h->d_wt);
  if (nft) MG_TIMED_LAUNCH(h->tm, fl_cohesion, dim3(nft), T, 0, P);
  if (nwt) {
    MG_TIMED_LAUNCH(h->tm, fl_adhesion, dim3(nwt), T, 0, P, (const uint2*)h->d_wt);
    MG_TIMED_LAUNCH(h->tm, fl_adh_scan, G, T, 0, P);
    MG_TIMED_LAUNCH(h->tm, fl_adh_scatter, dim3(nwt), T, 0, P, (const uint2*)h->d_wt);
  }
  if (nft) MG_TIMED_LAUNCH(h->tm, fl_adh_fin, dim3(nft), T, 0, P);
  if (nft) {
    MG_TIMED_LAUNCH(h->tm, fl_wall, dim3(nft), T, 0, P);
    MG_TIMED_LAUNCH(h->tm, fl_vel, dim3(nft), T, 0, P);
  }
  MG_TIMED_LAUNCH(h->tm, fl_grid_b, dim3(G.x, 3), T, 0, P);
  if (nwt) MG_TIMED_LAUNCH(h->tm, fl_gs_pre, dim3(nwt), T, 0, P, (const uint2*)h->d_wt);
  {
    // The sweeps (a few waves per drum) leave most of the GPU idle. The density at the final
    // positions, the Morris weights and the boundary groups read neither the velocities nor the
    // impulses the sweeps and the squeeze links write when no drum is Bingham (a Bingham drum's
    // viscosity reads the swept velocities), so they then run beside them on a second stream.
    bool bingham = false;
    for (uint32_t d = 0; d < nd; d++) bingham |= hd[d].bingham != 0 && hd[d].n > 0 && hd[d].mu > 0.0f;
    const bool overlap = !bingham && nft > 0;
    if (overlap && !h->side[0]) {
      bool ok2 = true;
      for (int k = 0; k < 2; k++) {
        ok2 &= hipStreamCreateWithFlags(&h->side[k], hipStreamNonBlocking) == hipSuccess;
        ok2 &= hipEventCreateWithFlags(&h->join[k], hipEventDisableTiming) == hipSuccess;
      }
      ok2 &= hipEventCreateWithFlags(&h->fork, hipEventDisableTiming) == hipSuccess;
      if (!ok2) return fl_fail(h, err, errlen, MG_E_DEVICE, "HIP error creating the side streams");
    }
    uint32_t maxn = 0;
    for (uint32_t d = 0; d < nd; d++) maxn = std::max(maxn, hd[d].n);
    {  // the sweeps' levels (before the fork: one wave per drum, slowed most by the side stream's
       // kernels): a 16-bit last level per particle in LDS (up to 160 KiB)
      const uint32_t cap = (uint32_t)std::min<size_t>((160u * 1024u) / 2, (size_t)maxn + 2);
      const size_t bytes = (size_t)cap * 2;
      if (bytes > 65536)
        (void)hipFuncSetAttribute((const void*)fl_gs_level, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
      MG_TIMED_LAUNCH(h->tm, fl_gs_level, G, dim3(64), bytes, P, cap);
    }
    hipStream_t sg = nullptr, sw = nullptr;
    if (overlap) {
      sg = h->side[0];
      sw = h->side[1];
      (void)hipEventRecord(h->fork, nullptr);
      (void)hipStreamWaitEvent(sg, h->fork, 0);
      (void)hipStreamWaitEvent(sw, h->fork, 0);
    }
    // the drum's fluid velocities stay in LDS during the sweep when they fit (160 KiB per workgroup
    // on gfx950, less the kernel's static LDS)
    // (after the balls' impulse copies, 12 bytes a ball)
    uint32_t bcap = 1;
    for (uint32_t d = 0; d < nd; d++) bcap = std::max(bcap, hd[d].nb);
    const size_t static_lds = 4u * (GS_WAVES * 3u * 64u + 3u * GS_MAXB + 1u) + 8u * GS_WAVES + 8u + 256u;
    const size_t cap_bytes = 160u * 1024u - static_lds - 1024u - 12u * (size_t)bcap;
    const uint32_t lds_cap = (uint32_t)std::min<size_t>(cap_bytes / 8, (size_t)maxn);
    const size_t bytes = (size_t)lds_cap * 8 + 12u * (size_t)bcap;
    if (bytes > 65536) (void)hipFuncSetAttribute((const void*)fl_gs, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    MG_TIMED_LAUNCH_S(h->tm, sg, fl_gs, G, dim3(GS_THREADS), bytes, P, lds_cap, bcap);
    // the squeeze links read the swept impulses and the ball grid only (their partner counts are
    // cleared by fl_gs_pre), so they follow the sweeps on the sweeps' stream
    if (nwt) MG_TIMED_LAUNCH_S(h->tm, sg, fl_links, dim3(nwt), T, 0, P, (const uint2*)h->d_wt);
    if (nft) {
      MG_TIMED_LAUNCH_S(h->tm, sw, fl_dens2, dim3(nft), T, 0, P);
      MG_TIMED_LAUNCH_S(h->tm, sw, fl_mu, dim3(nft), T, 0, P);
      MG_TIMED_LAUNCH_S(h->tm, sw, fl_weights, dim3(nft), T, 0, P);
    }
    if (nwt) MG_TIMED_LAUNCH_S(h->tm, sw, fl_groups, dim3(nwt), T, 0, P, (const uint2*)h->d_wt);
    if (overlap) {
      (void)hipEventRecord(h->join[0], sg);
      (void)hipEventRecord(h->join[1], sw);
      (void)hipStreamWaitEvent(nullptr, h->join[0], 0);
      (void)hipStreamWaitEvent(nullptr, h->join[1], 0);
    }
  }
  MG_TIMED_LAUNCH(h->tm, fl_cg_prep, G, T, 0, P);
  if (nwt && nrt) {
    MG_TIMED_LAUNCH(h->tm, fl_cg_scatter, dim3(nwt), T, 0, P, (const uint2*)h->d_wt);
    MG_TIMED_LAUNCH(h->tm, fl_cg_order, dim3(nwt), T, 0, P, (const uint2*)h->d_wt);
  }
  if (nrt) {
    if (nwt) {
      MG_TIMED_LAUNCH(h->tm, fl_cg_balls, dim3(nwt), T, 0, P, (const uint2*)h->d_wt, 0);
      MG_TIMED_LAUNCH(h->tm, fl_cg_setup, dim3(nrt), T, 0, P);
      MG_TIMED_LAUNCH(h->tm, fl_cg_balls, dim3(nwt), T, 0, P, (const uint2*)h->d_wt, 1);
    } else {
      MG_TIMED_LAUNCH(h->tm, fl_cg_setup, dim3(nrt), T, 0, P);
    }
    MG_TIMED_LAUNCH(h->tm, fl_cg_apply, dim3(nrt), T, 0, P, 0);
    MG_TIMED_LAUNCH(h->tm, fl_cg_init, dim3(nrt), T, 0, P);
    MG_TIMED_LAUNCH(h->tm, fl_cg_start, G, dim3(64), 0, P);
    const dim3 FG((nd + 63) / 64);
    for (uint32_t done_it = 0; done_it < 200;) {
      const uint32_t k = std::min(h->cg_chunk, 200 - done_it);
      for (uint32_t q = 0; q < k; q++) {
        if (nwt) MG_TIMED_LAUNCH(h->tm, fl_cg_balls, dim3(nwt), T, 0, P, (const uint2*)h->d_wt, 2);
        MG_TIMED_LAUNCH(h->tm, fl_cg_apply, dim3(nrt), T, 0, P, 1);
        MG_TIMED_LAUNCH(h->tm, fl_cg_upd, dim3(nrt), T, 0, P);
        MG_TIMED_LAUNCH(h->tm, fl_cg_fin, FG, dim3(64), 0, P, nd);
        MG_TIMED_LAUNCH(h->tm, fl_cg_dir, dim3(nrt), T, 0, P);
      }
      done_it += k;
      if (hipMemcpy(h->h_st.data(), P.st, nd * sizeof(FlState), hipMemcpyDeviceToHost) != hipSuccess)
        return fl_fail(h, err, errlen, MG_E_DEVICE, "HIP error in the viscosity solve");
      bool all = true;
      for (uint32_t d = 0; d < nd; d++) all &= h->h_st[d].mode == 0 || h->h_st[d].done != 0;
      if (all) break;
    }
  }
  if (nft) {
    MG_TIMED_LAUNCH(h->tm, fl_cg_out, dim3(nft), T, 0, P);
    MG_TIMED_LAUNCH(h->tm, fl_shear, dim3(nft), T, 0, P);
    MG_TIMED_LAUNCH(h->tm, fl_clamp, dim3(nft), T, 0, P);
  }
  if (nlt) MG_TIMED_LAUNCH(h->tm, fl_ball_out, dim3(nlt), T, 0, P, (const uint2*)h->d_lt);
  MG_TIMED_LAUNCH(h->tm, fl_stats, G, dim3(64), 0, P, h->d_out);
  ok = hipGetLastError() == hipSuccess &&
       hipMemcpy(h->h_st.data(), P.st, nd * sizeof(FlState), hipMemcpyDeviceToHost) == hipSuccess &&
       hipMemcpy(h->h_out.data(), h->d_out, nd * sizeof(FlOut), hipMemcpyDeviceToHost) == hipSuccess;
  if (!ok) return fl_fail(h, err, errlen, MG_E_DEVICE, "HIP error in the fluid step kernels");
  h->tm.collect();
  int rc = MG_OK;
  for (uint32_t d = 0; d < nd; d++) {
    const FlDrum& D = hd[d];
    const FlState& S = h->h_st[d];
    const FlOut& o = h->h_out[d];
    mg_fluid_stats& s = stats[d];
    std::memset(&s, 0, sizeof s);
    s.error_bits = S.err;
    if (S.err && rc == MG_OK)
      rc = fl_fail(h, err, errlen, MG_E_CAPACITY,
                   "drum " + std::to_string(d) + ": workspace capacity exceeded (error bits 0x" +
                       [&] { char b[16]; std::snprintf(b, sizeof b, "%x", S.err); return std::string(b); }() + ")");
    if (D.n == 0) continue;
    uint32_t passes = 0;
    for (uint32_t r = 0; r < D.npass; r++) {
      bool act = true;
      if (D.force) act = r < std::max(D.forced_d, 1u);
      else
        for (uint32_t q = 0; q < r && act; q++) {
          float c;
          std::memcpy(&c, &S.maxc[q], 4);
          if ((q + 1 >= D.base && c <= 0.02f) || q + 1 >= D.base + 6) act = false;
        }
      if (act) passes = r + 1;
    }
    s.density_iterations = passes;
    s.viscosity_iterations = S.mode ? S.iters : 0u;
    s.joint_solve = S.mode == 2 ? 1 : 0;
    s.mean_shear_rate_per_s = o.shn > 0.0 ? (float)(o.sh / o.shn) : 0.0f;
    s.wall_work_j = (float)o.ww;
    s.wall_impulse_x = o.wix;
    s.wall_impulse_y = o.wiy;
    s.wall_angular_impulse = o.wang;
    for (int b = 0; b < 32; b++) s.wall_angle_bins[b] = o.bins[b];
    const double hm = 0.5 * (double)D.mass;
    const double k0 = hm * o.ke0, k1 = hm * o.ke1, k2 = hm * S.ke_drag, k3 = hm * S.ke_buoy, k4 = hm * o.kev,
                 k5 = hm * o.kec;
    s.ke_delta_noslip_j = k1 - k0;
    s.ke_delta_drag_j = k2 - k1;
    s.ke_delta_buoyancy_j = k3 - k2;
    s.ke_delta_viscosity_j = k4 - k3;
    s.ke_delta_clamp_j = k5 - k4;
    s.coupling_clamp_hits = S.clamp_hits;
    s.n_groups = S.n_groups;
    s.n_links = S.n_links;
    s.links_in_range = S.in_range;
  }
  return rc;
}

int mg_fluid_set_timing(mg_fluid* h, int on) {
  if (!h) return MG_E_ARG;
  h->tm.set(on != 0);
  return MG_OK;
}

int mg_fluid_kernel_times(mg_fluid* h, char* json, size_t cap) {
  if (!h || !json || cap == 0) return MG_E_ARG;
  const std::string s = h->tm.json();
  if (s.size() + 1 > cap) return -(int)(s.size() + 1);
  std::memcpy(json, s.c_str(), s.size() + 1);
  return (int)s.size();
}

}  // extern "C"
