// Periodic-box PBF workload on the GPU (kernel isolation of the fluid hot path).
//
// One step applies the reference's fluid arithmetic to a box of side L with minimum-image pair
// vectors and wrapped positions:
//   bx_predict            x0 = x, x += v dt, cell key, cell histogram      pbf.rs:1225-1235, grid.rs:101-226
//   bx_scan_*             exclusive scan of the histogram -> cell starts   grid.rs:160-200
//   bx_scatter, bx_cellsort, bx_gather
//                         stable counting sort: cell order, ascending particle id inside a
//                         cell, and the particle arrays physically reordered by cell so every
//                         neighbour gather below reads contiguous memory
//   bx_pairs_lambda       the step's pair table: 3x3-cell candidates with |d|^2 <= h^2 at the
//                         predicted positions (the reference's neighbour lists, pbf.rs:262-288),
//                         built inside the first density pass A (bx_pairs when there is none)
//   bx_lambda, bx_dp      density constraint passes (rho, C, lambda; dp, x)  pbf.rs:1257-1334
//   bx_vel_density        v = (x - x0)/dt, density at the final positions    pbf.rs:1656-1674, 2136-2144
//   bx_cg_init            r = b - A v with the Morris weights kept for the step   pbf.rs:312-407
//   bx_cg_dir, bx_cg_upd  one CG iteration (direction + operator; vector update)  pbf.rs:423-480
// Pairs are visited per particle in the reference's order: cells (-1,-1), (-1,0), (0,-1), own,
// (1,0), (1,1), (0,1), (-1,1), (1,-1), ascending id within a cell; every sum over a particle's
// neighbours therefore rounds exactly as the reference's does. CG dot products and norms are FP64
// partial sums per workgroup (fixed tree), summed in block order by every consumer workgroup, so
// the scalars are identical everywhere and results are run-to-run deterministic.
//
// Precision: the state kernels are templates on the real type R. R = float is FP32 state with FP64
// CG reductions (the reference's f32 arithmetic, bit for bit); R = double is FP64 state and
// arithmetic from the same f32 initial state and geometry, promoted (the oracle's f64 policy).
//
// CDNA4 mapping: 256-thread workgroups (4 wave64s; 1024 for the kernels producing FP64 partials),
// one particle per lane; the blockIdx -> tile map gives each of the 8 XCDs one contiguous run of
// cell-sorted particles so the halo rows of neighbouring tiles are read from the same XCD's L2.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <functional>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/millgpu.h"
#include "mg_common.h"

#define BX_BLOCK 256
#define BX_RBLOCK 1024  // workgroup size of the kernels that produce FP64 partial sums
#define BX_SCAN_ITEMS 4   // one 16-byte vector per lane (coalesced)
#define BX_SCAN_CHUNK (BX_BLOCK * BX_SCAN_ITEMS)
#define BX_CG_CHUNK 4
#define BX_NRED 512    // most workgroups of the FP64 reduction kernels (two rounds of one per CU)
#define BX_WINDOW 16  // steps enqueued between host waits in a multi-step call
#define BX_SPARE_IT 1  // CG iterations launched beyond the predicted count in a window
// Loop-shape switch (both forms give identical results; the default is the measured faster one).
#ifndef BX_GATHER
#define BX_GATHER 8    // table entries whose gathers are issued together in the FP32 density passes
#endif
#ifndef BX_BATCH
#define BX_BATCH 1     // issue the gathers of BX_GATHER table entries before using them
#endif
#ifndef BX_KMAX
#define BX_KMAX 32  // pairs kept per particle in the per-step pair table (more: the kernel rescans)
#endif
#ifndef BX_NB_RANGE
#define BX_NB_RANGE 32767  // largest |j - i| a pair-table entry holds (int16; beyond: the kernel rescans)
#endif
// kernel classes of the optional per-kernel timing (mg_box_kernel_times)
#define BX_NCLASS 8
#define BX_MAXPASS 16  // most density passes of one step (adaptive: base + 6)
enum { BXT_GRID = 0, BXT_LAMBDA, BXT_DP, BXT_VEL, BXT_CGINIT, BXT_SPMV, BXT_UPDATE, BXT_PAIRLAM };
// precision-independent kernels keep C names; the state kernels are templates on the real type R
// (float: FP32 state with FP64 reductions; double: FP64 throughout)
#define BXK extern "C" __global__ void __launch_bounds__(BX_BLOCK)
#define BXKR extern "C" __global__ void __launch_bounds__(BX_RBLOCK)
// the 256-lane state kernels fit 8 waves per SIMD in both precisions; the bound keeps them there
#define BXKT template <class R> __global__ void __launch_bounds__(BX_BLOCK, 8)
// Waves per SIMD the 1024-lane kernels are compiled for: two workgroups per CU need 8, i.e. at most
// 64 VGPRs and 80 SGPRs per wave (SGPRs are allocated in blocks of 16, so 81 already allows only
// one workgroup and halves the resident waves). The FP32 kernels fit; the FP64 ones do not (they
// would spill), so they keep one workgroup per CU.
template <class R> struct BxRWaves { static constexpr int value = 1; };
template <> struct BxRWaves<float> { static constexpr int value = 8; };
#define BXKRT template <class R> __global__ void __launch_bounds__(BX_RBLOCK, BxRWaves<R>::value)
#define BDEV __device__ __attribute__((always_inline)) inline
#define BDEVX __device__ __attribute__((always_inline)) inline  // explicit specializations

template <class R> struct VecT;
template <> struct VecT<float> { using type = float2; };
template <> struct VecT<double> { using type = double2; };
template <class R> using Vec = typename VecT<R>::type;
template <class R> struct Vec4T;
template <> struct Vec4T<float> { using type = float4; };
template <> struct Vec4T<double> { using type = double4; };
template <class R> using Vec4 = typename Vec4T<R>::type;
template <class V> BDEV auto v4c(const V& v, int q) { return q == 0 ? v.x : (q == 1 ? v.y : (q == 2 ? v.z : v.w)); }
template <class V, class T> BDEV void set4(V& v, int q, T x) {
  if (q == 0) v.x = x;
  else if (q == 1) v.y = x;
  else if (q == 2) v.z = x;
  else v.w = x;
}

// Geometry and constants of one box, passed by value to every kernel, in the state's precision
// (the f64 constants are the f32 geometry promoted, then the derived ones formed in f64, as the
// oracle's f64 policy forms them). `halt` is the window stop word: once a step of a multi-step
// window has run out of CG iterations before converging, the window's remaining kernels return at
// once (the host then makes that step's start state live again). [lo, hi] = [2h, L - 2h] is the
// interior where pair vectors need no minimum image (see wave_inside).
template <class R> struct BoxGeo {
  uint32_t n, ncx, nblk, xcd_chunk, nred, xcd_chunk_red, ppt;
  R h, h2, L, halfL, cw, mass, rest, dt, k6, ks, w0, m_over_rest, mu2m, eta2, lo, hi;
  R gdt;  // gravity * dt, added to v.y in the predict
  const uint32_t* halt;
  // Adaptive density passes (pbf.rs:1261-1333): pass p (0-based) runs when p < base or the largest
  // constraint of pass p - 1 exceeded 0.02. dens[p] holds that largest constraint's bits (R's bits
  // widened to 64; non-negative values order like their bits), dens[BX_MAXPASS] the passes run.
  uint32_t adaptive, base;
  unsigned long long* dens;
};
// Particles per lane in the reduction kernels. FP64 (one workgroup per CU) caps the workgroups at
// BX_NRED and loops over g.ppt particles per lane; FP32 keeps one particle per lane (a compile-time
// 1: the loop's live values would push its CG kernels over the 64-VGPR budget).
template <class R> BDEV uint32_t red_ppt(const BoxGeo<R>& g) {
  if constexpr (BxRWaves<R>::value == 8) return 1u;
  else return g.ppt;
}
// body(i, k) for the lane's k-th particle i (FP32: the one particle, as a plain branch).
template <class R, class F> BDEV void for_lane_particles(const BoxGeo<R>& g, uint32_t base, F&& body) {
  if constexpr (BxRWaves<R>::value == 8) {
    if (base < g.n) body(base, 0u);
  } else {
    for (uint32_t k = 0; k < g.ppt; k++) {
      const uint32_t i = base + k * BX_RBLOCK;
      if (i >= g.n) break;
      body(i, k);
    }
  }
}

// Device scalars of the CG loop (pbf.rs:423-480).
// rr[u & 1] = r(u).r(u) once a CG kernel has summed it (block 0 stores it for later kernels).
struct BoxScal {
  double bnorm, target, res;
  double rr[2];
  uint32_t it, done, max_cg;
  float tol;
  const void* mu_eff;  // per-particle effective viscosity (Bingham boxes), else null; set at creation
  unsigned long long dens[BX_MAXPASS + 1];  // BoxGeo::dens
};

// ---------------------------------------------------------------------------------------------
// device helpers (no contraction: compiled with -ffp-contract=off; IEEE division and sqrt)
// ---------------------------------------------------------------------------------------------
BDEV float msqrt(float a) { return sqrtf(a); }
BDEV double msqrt(double a) { return sqrt(a); }
BDEV float mfmod(float a, float b) { return fmodf(a, b); }
BDEV double mfmod(double a, double b) { return fmod(a, b); }
BDEV float mfloor(float a) { return floorf(a); }
BDEV float mexp(float a) { return expf(a); }
BDEV double mexp(double a) { return exp(a); }
BDEV double mfloor(double a) { return floor(a); }
BDEV float mfabs(float a) { return fabsf(a); }
BDEV double mfabs(double a) { return fabs(a); }
BDEV float mfmax(float a, float b) { return fmaxf(a, b); }  // f32::max (maxNum)
BDEV double mfmax(double a, double b) { return fmax(a, b); }
// glam Vec2 operations, component-wise in glam's order
#define BX_VEC_OPS(V, S, MK)                                                   \
  BDEV V mk(S x, S y) { return MK(x, y); }                                     \
  BDEV V add2(V a, V b) { return MK(a.x + b.x, a.y + b.y); }                   \
  BDEV V sub2(V a, V b) { return MK(a.x - b.x, a.y - b.y); }                   \
  BDEV V mul2(V a, S s) { return MK(a.x * s, a.y * s); } /* Vec2 * s */        \
  BDEV V smul2(S s, V a) { return MK(s * a.x, s * a.y); } /* s * Vec2 */       \
  BDEV V div2(V a, S s) { return MK(a.x / s, a.y / s); }                       \
  BDEV S dot2(V a, V b) { return a.x * b.x + a.y * b.y; }                      \
  BDEV S len2(V a) { return a.x * a.x + a.y * a.y; }
BX_VEC_OPS(float2, float, make_float2)
BX_VEC_OPS(double2, double, make_double2)
#undef BX_VEC_OPS
// |d| and d / |d| of a pair vector: the compiler's correctly rounded square root and division
// sequences without their range scaling and special-value handling, which are the identity in the
// range the pair loops use them in, so the results are bit for bit those of sqrt and '/', and the
// two divisions by r share one refined reciprocal:
//  - pair_len(r2) equals sqrt(r2) for r2 >= 2^-96 (f32) / 2^-767 (f64). Below that it returns a
//    value < 1e-9, and every caller only compares such an r with 1e-9 (the kernel gradient and the
//    Morris direction are zero / fixed there), so results do not depend on it.
//  - pair_div(d, r) equals (d.x / r, d.y / r) for r in [2^-100, 2^30] (the box side is at most
//    2^23, so every r > 1e-9 is in range) when both |components| are >= 2^-96 (f32) / 2^-900
//    (f64); it takes the IEEE division for smaller ones (in f32 for every component whose square
//    is zero, i.e. below about 2^-75). Callers use the quotient only for r > 1e-9. With
//    guard = false the small-component test is left out: for callers whose nonzero components
//    are known to be >= 2^-96 (f32) / 2^-900 (f64).
// mg_box_check_arith compares them with the IEEE operations on the GPU.
template <class R> BDEV R pair_len(R r2) { return msqrt(r2); }
template <class R> BDEV Vec<R> pair_div(Vec<R> d, R r, bool guard = true) { return div2(d, r); }
BDEV float sqrt_core(float x) {  // v_sqrt (1 ulp) and the two neighbouring candidates, as LLVM lowers sqrtf
  const float s = __builtin_amdgcn_sqrtf(x);
  const float sm = __uint_as_float(__float_as_uint(s) - 1u);
  const float sp = __uint_as_float(__float_as_uint(s) + 1u);
  const float rm = __builtin_fmaf(-sm, s, x);
  const float rp = __builtin_fmaf(-sp, s, x);
  const float t = rm <= 0.0f ? sm : s;
  return rp > 0.0f ? sp : t;
}
template <> BDEVX float pair_len<float>(float r2) { return sqrt_core(r2); }
// f64: v_rsq and two coupled Newton steps on (x·y, y/2), then two residual corrections (LLVM's
// f64 sqrt lowering without its 2^256 scaling below 2^-767)
BDEV double sqrt_core(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y;
  double h = y * 0.5;
  const double r = __builtin_fma(-h, g, 0.5);
  g = __builtin_fma(g, r, g);
  double d = __builtin_fma(-g, g, x);
  h = __builtin_fma(h, r, h);
  g = __builtin_fma(d, h, g);
  d = __builtin_fma(-g, g, x);
  return __builtin_fma(d, h, g);
}
template <> BDEVX double pair_len<double>(double r2) { return sqrt_core(r2); }
template <> BDEVX double2 pair_div<double>(double2 d, double r, bool guard) {
  const bool tiny = (fabs(d.x) < 0x1p-900) | (fabs(d.y) < 0x1p-900);
  if (guard && __builtin_expect(tiny, 0)) return make_double2(d.x / r, d.y / r);
  // reciprocal refined twice, then one residual correction per quotient (the unscaled path of
  // the compiler's f64 division)
  const double y0 = __builtin_amdgcn_rcp(r);
  const double y1 = __builtin_fma(y0, __builtin_fma(-r, y0, 1.0), y0);
  const double y = __builtin_fma(y1, __builtin_fma(-r, y1, 1.0), y1);
  auto q = [&](double a) __attribute__((always_inline)) {
    const double q0 = a * y;
    return __builtin_fma(__builtin_fma(-r, q0, a), y, q0);
  };
  return make_double2(q(d.x), q(d.y));
}
template <> BDEVX float2 pair_div<float>(float2 d, float r, bool guard) {
  // a component below 2^-75 squares to zero (the same products len2 formed): take the IEEE division
  // for those, which covers every component below the core sequence's 2^-96 bound
  const bool tiny = (d.x * d.x == 0.0f) | (d.y * d.y == 0.0f);
  if (guard && __builtin_expect(tiny, 0)) return make_float2(d.x / r, d.y / r);
  // Newton-refined reciprocal, then two residual corrections per quotient (the unscaled path of
  // the compiler's division: rcp, fma, fma | mul, fma, fma, fma, fma)
  const float y0 = __builtin_amdgcn_rcpf(r);
  const float e = __builtin_fmaf(-r, y0, 1.0f);
  const float y = __builtin_fmaf(e, y0, y0);
  auto q = [&](float a) __attribute__((always_inline)) {
    const float q0 = a * y;
    const float r0 = __builtin_fmaf(-r, q0, a);
    const float q1 = __builtin_fmaf(r0, y, q0);
    const float r1 = __builtin_fmaf(-r, q1, a);
    return __builtin_fmaf(r1, y, q1);
  };
  return make_float2(q(d.x), q(d.y));
}
template <class R> BDEV R wrapR(R a, R L) {  // rem_euclid
  R r = mfmod(a, L);
  return r < (R)0 ? r + mfabs(L) : r;
}
// Minimum image: d - L above L/2, d + L below -L/2 (as d - (-L), the same IEEE operation), else d;
// a select, no branch.
BDEV float mcopysign(float a, float b) { return copysignf(a, b); }
BDEV double mcopysign(double a, double b) { return copysign(a, b); }
template <class R> BDEV R mi1(R d, const BoxGeo<R>& g) {
  const R t = d - mcopysign(g.L, d);
  return mfabs(d) > g.halfL ? t : d;
}
// Pair vector a - b with the minimum image.
template <class R> BDEV Vec<R> rel(Vec<R> a, Vec<R> b, const BoxGeo<R>& g) {
  const Vec<R> d = sub2(a, b);
  return mk(mi1(d.x, g), mi1(d.y, g));
}
// Pair loops come in two forms, chosen per wave: Wrap<true> takes the minimum image, Wrap<false>
// the plain difference. The plain form is exact for a particle a in [2h, L - 2h]^2 and any partner
// b in [0, L]^2 (positions are kept wrapped): a component with |a - b| <= L/2 is its own minimum
// image; one with |a - b| > L/2 has a minimum image of length L - |a - b| >= 2h (less one rounding),
// so the pair lies beyond h in both forms and adds exactly nothing in either (poly6, the spiky
// gradient and the Morris weight vanish at r >= h, and a zero weight adds only signed zeros to a
// sum that started at +0). For a pair within h of such a particle both coordinates lie in
// [xi/2, 2 xi] (xi >= 2h), so each component is an exact difference of two floats >= h: zero or
// at least ulp(h) >= h 2^-24, which for h >= 2^-60 (lo > hi below that) is far above the
// quotient's small-component bound, and Wrap<false> leaves that test out. Wrap<false> also means
// that no lane's table row holds an offset taken modulo n (cnt bit 31 clear), so an entry decodes
// as i + offset.
template <bool W> struct Wrap {};
template <bool W> BDEV constexpr bool guarded(Wrap<W>) { return W; }
template <class R> BDEV Vec<R> rel(Vec<R> a, Vec<R> b, const BoxGeo<R>& g, Wrap<true>) { return rel(a, b, g); }
template <class R> BDEV Vec<R> rel(Vec<R> a, Vec<R> b, const BoxGeo<R>& g, Wrap<false>) { return sub2(a, b); }
// True when every active lane's particle lies in [lo, hi]^2 (wave-uniform).
template <class R> BDEV bool wave_inside(Vec<R> x, const BoxGeo<R>& g) {
  const bool in = x.x >= g.lo && x.x <= g.hi && x.y >= g.lo && x.y <= g.hi;
  return __ballot(!in) == 0ull;
}
// Pair-count word: bit 31 marks a table row with an offset taken modulo n.
#define BX_CNT_WRAPPED 0x80000000u
BDEV uint32_t pcount(uint32_t w) { return w & ~BX_CNT_WRAPPED; }
// As wave_inside, and no lane's table row holds a wrapped offset (the Wrap<false> condition).
template <class R> BDEV bool wave_fast(Vec<R> x, uint32_t cntw, const BoxGeo<R>& g) {
  const bool ok = x.x >= g.lo && x.x <= g.hi && x.y >= g.lo && x.y <= g.hi && !(cntw & BX_CNT_WRAPPED);
  return __ballot(!ok) == 0ull;
}
// f(Wrap<false>{}) when the wave is inside, else f(Wrap<true>{}) (a uniform branch).
template <class F> BDEV void by_wrap(bool inside, F&& f) {
  if (inside) f(Wrap<false>{});
  else f(Wrap<true>{});
}
template <class R> BDEV uint32_t cell1(R a, const BoxGeo<R>& g) {
  R q = mfloor(a / g.cw);
  if (!(q >= (R)0)) return 0u;
  if (q >= (R)(g.ncx - 1u)) return g.ncx - 1u;
  return (uint32_t)q;
}
template <class R> BDEV R poly6(R r2, const BoxGeo<R>& g) {  // pbf.rs:119-126
  if (r2 >= g.h2 || r2 < (R)0) return (R)0;
  R t = g.h2 - r2;
  return g.k6 * t * t * t;
}
template <class R> BDEV Vec<R> spiky(Vec<R> d, R r, const BoxGeo<R>& g, bool guard) {  // pbf.rs:131-137
  if (r <= (R)1e-9f || r >= g.h) return mk((R)0, (R)0);
  R coeff = g.ks * (g.h - r) * (g.h - r);
  return mul2(pair_div(d, r, guard), coeff);
}
// XCD-aware tile index: hardware dispatches consecutive workgroups round-robin over the 8 XCDs;
// remap so XCD x processes tiles [x*chunk, (x+1)*chunk).
template <class R> BDEV uint32_t tile_id(const BoxGeo<R>& g) {
  const uint32_t b = blockIdx.x;
  if (g.xcd_chunk == 0u) return b;  // tile count not a multiple of 8: identity map
  return (b & 7u) * g.xcd_chunk + (b >> 3);
}
template <class R> BDEV uint32_t tile_id_red(const BoxGeo<R>& g) {
  const uint32_t b = blockIdx.x;
  if (g.xcd_chunk_red == 0u) return b;
  return (b & 7u) * g.xcd_chunk_red + (b >> 3);
}
// Candidate cell o (0..8) of cell (cx, cy) in the reference's per-particle order, wrapped.
// A cell coordinate c + o (c in [0, ncx), o in {-1, 0, 1}) wrapped into [0, ncx): (c + o) mod ncx
// by two selects (a modulus by the run-time ncx costs a multiply-high sequence per use).
BDEV uint32_t wrap_cell(uint32_t c, int o, uint32_t ncx) {
  if (o < 0) return c == 0u ? ncx - 1u : c - 1u;
  if (o > 0) return c + 1u >= ncx ? 0u : c + 1u;
  return c;
}
BDEV uint32_t nbr_cell(uint32_t cx, uint32_t cy, int o, uint32_t ncx) {
  const int ox = o == 0 ? -1 : (o == 1 ? -1 : (o == 2 ? 0 : (o == 3 ? 0 : (o == 4 ? 1 : (o == 5 ? 1 : (o == 6 ? 0 : (o == 7 ? -1 : 1)))))));
  const int oy = o == 0 ? -1 : (o == 1 ? 0 : (o == 2 ? -1 : (o == 3 ? 0 : (o == 4 ? 0 : (o == 5 ? 1 : (o == 6 ? 1 : (o == 7 ? 1 : -1)))))));
  return wrap_cell(cy, oy, ncx) * ncx + wrap_cell(cx, ox, ncx);
}
// Fixed-order workgroup sum of one double per lane (wave shuffles, then the waves in order);
// every lane receives the total.
template <int NT>
BDEV double block_sum(double v) {
  __shared__ double s_w[NT / 64];
  __shared__ double s_tot;
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) s_w[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = 0.0;
    for (int k = 0; k < NT / 64; k++) r += s_w[k];
    s_tot = r;
  }
  __syncthreads();
  const double r = s_tot;
  __syncthreads();
  return r;
}
// Sum of the per-block partials in a fixed order. Every block that calls it computes the same
// value bit for bit, so the CG scalars can be evaluated redundantly in each consumer block.
template <int NT>
BDEV double sum_parts(const double* __restrict__ part, uint32_t nb) {
  double s = 0.0;
  for (uint32_t k = threadIdx.x; k < nb; k += NT) s += part[k];
  return block_sum<NT>(s);
}
// Two such sums in one workgroup reduction (each component in exactly sum_parts' order).
template <int NT>
BDEV double2 sum_parts2(const double* __restrict__ pa, const double* __restrict__ pb, uint32_t nb) {
  __shared__ double2 s_w2[NT / 64];
  __shared__ double2 s_tot2;
  double2 v = make_double2(0.0, 0.0);
  for (uint32_t k = threadIdx.x; k < nb; k += NT) {
    v.x += pa[k];
    v.y += pb[k];
  }
  for (int o = 32; o > 0; o >>= 1) {
    v.x += __shfl_down(v.x, o, 64);
    v.y += __shfl_down(v.y, o, 64);
  }
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) s_w2[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double2 r = make_double2(0.0, 0.0);
    for (int k = 0; k < NT / 64; k++) {
      r.x += s_w2[k].x;
      r.y += s_w2[k].y;
    }
    s_tot2 = r;
  }
  __syncthreads();
  const double2 r = s_tot2;
  __syncthreads();
  return r;
}

// Per-step pair table: the neighbours of sorted particle i at the predicted positions, in the
// reference's per-particle order, in quads of four entries per particle (see tquad) so
// lane-adjacent particles read adjacent 8-byte words. An entry is the signed 16-bit offset j - i taken modulo n into [-n/2, n/2]
// (after the cell sort a neighbour sits within about two cell rows of i, or as far on the other
// side of the periodic y seam). cnt[i] is the true count; a particle with more than BX_KMAX pairs,
// or with a neighbour whose offset does not fit 16 bits, has cnt[i] > BX_KMAX and is handled by
// rescanning its 3x3 cells (same order, same membership test).
template <class R> struct PairView {
  const Vec<R>* __restrict__ XP;
  const uint32_t* __restrict__ start;
  const int16_t* __restrict__ nb;
  const uint32_t* __restrict__ cnt;
  uint32_t n;
};
// The table is stored in quads: entries 4q .. 4q + 3 of particle i are the 8 bytes of quad (q, i),
// at uint2 index q * n + i, so one coalesced 8-byte load per lane fetches four entries (entry k in
// half-word k & 3, low half first).
// Neighbour gather p[j] with a 32-bit byte offset (n <= 2^26 and elements of at most 16 bytes keep
// it below 2^30), so the load takes the array base in scalar registers and one offset register.
template <class T> BDEV T at(const T* __restrict__ p, uint32_t j) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(p) + j * (uint32_t)sizeof(T));
}
template <class R> BDEV uint2 tquad(const PairView<R>& pv, uint32_t q, uint32_t i) {
  return at(reinterpret_cast<const uint2*>(pv.nb) + (size_t)q * pv.n, i);
}
BDEV int32_t qent(uint2 v, int e) {  // entry e (0..3) of a quad, sign-extended
  const uint32_t w = (e & 2) ? v.y : v.x;
  return (e & 1) ? ((int32_t)w >> 16) : (int32_t)(int16_t)(w & 0xFFFFu);
}
// Table entry -> particle index. i + d leaves [0, n) only across the y seam, and then in one
// direction per particle (below 0 for i < n/2, past n - 1 otherwise, as |d| <= n/2), so one
// per-particle correction (hoisted out of the pair loops) undoes the modulus.
BDEV uint32_t nb_dec(int32_t d, uint32_t i, uint32_t n) {
  const uint32_t t = i + (uint32_t)d;
  const uint32_t fix = i < (n >> 1) ? n : 0u - n;
  return t >= n ? t + fix : t;
}
BDEV uint32_t nb_dec(int32_t d, uint32_t i, uint32_t n, Wrap<true>) { return nb_dec(d, i, n); }
BDEV uint32_t nb_dec(int32_t d, uint32_t i, uint32_t, Wrap<false>) { return i + (uint32_t)d; }
// A copy of a wave-uniform value in a vector register, for the selects of the table build (a select
// reads at most one scalar register, so a scalar operand is otherwise moved to a vector register on
// every use inside the loop).
BDEV int32_t lane_copy(int32_t s) {
  int32_t v;
  asm volatile("v_mov_b32 %0, %1" : "=v"(v) : "s"(s));
  return v;
}
// The table entry of neighbour j of particle i: j - i taken modulo n into [-n/2, n/2]. An entry
// that does not fit 16 bits (|entry| > BX_NB_RANGE) marks its row for the rescan, and an entry
// other than j - i marks the row as wrapped (cnt bit 31).
// n and -n are passed in vector registers (lane_copy).
BDEV int32_t nb_off(int32_t raw, int32_t half, int32_t vn, int32_t vneg_n) {
  return raw + (raw > half ? vneg_n : (raw < -half ? vn : 0));
}
// f(j, d, r2) for every pair, with d = xpi - XP[j] (minimum image) and r2 = |d|^2 as the test used.
template <class R, bool W, class F>
BDEV void scan_pairs_d(const BoxGeo<R>& g, const Vec<R>* __restrict__ XP, const uint32_t* __restrict__ start,
                       uint32_t i, Vec<R> xpi, Wrap<W> w, F&& f) {
  const uint32_t cx = cell1(xpi.x, g), cy = cell1(xpi.y, g);
  // the nine cell ranges are loaded up front (independent loads), then walked in order
  uint32_t cs[9], ce[9];
#pragma unroll
  for (int o = 0; o < 9; o++) {
    const uint32_t k = nbr_cell(cx, cy, o, g.ncx);
    cs[o] = start[k];
    ce[o] = start[k + 1];
  }
#pragma unroll
  for (int o = 0; o < 9; o++) {
    for (uint32_t j = cs[o]; j < ce[o]; j++) {
      // particle i itself is loaded like any candidate and left out by the one membership test
      const Vec<R> d = rel(xpi, at(XP, j), g, w);
      const R r2 = len2(d);
      if (j != i && !(r2 > g.h2)) f(j, d, r2);
    }
  }
}
template <class R, class F>
BDEV void scan_pairs(const BoxGeo<R>& g, const Vec<R>* __restrict__ XP, const uint32_t* __restrict__ start,
                     uint32_t i, Vec<R> xpi, F&& f) {
  scan_pairs_d(g, XP, start, i, xpi, Wrap<true>{}, [&](uint32_t j, Vec<R>, R) __attribute__((always_inline)) { f(j); });
}
template <class R, class F>
BDEV void for_pairs(const BoxGeo<R>& g, const PairView<R>& pv, uint32_t i, F&& f) {
  const uint32_t c = pcount(pv.cnt[i]);
  if (c <= BX_KMAX) {
    for (uint32_t k = 0; k < c; k += 4u) {
      const uint2 v = tquad(pv, k >> 2, i);
#pragma unroll
      for (int e = 0; e < 4; e++)
        if (k + e < c) f(nb_dec(qent(v, e), i, g.n));
    }
  } else {
    scan_pairs(g, pv.XP, pv.start, i, pv.XP[i], f);
  }
}
// As for_pairs, with the per-neighbour gathers split out: load(j) returns what use(j, v) needs,
// and the loads of four table entries are all issued before the first use (memory-level
// parallelism; without it every neighbour costs two dependent round trips).
// cntw = cnt[i] (read by the caller for the wave's Wrap choice). G table entries are gathered
// together: 8 in the FP32 density passes, 4 where more would not fit the 64-VGPR budget of 8
// waves per SIMD (the FP64 kernels and the Morris-weight set-up).
template <class R> struct Gather { static constexpr int value = 4; };
template <> struct Gather<float> { static constexpr int value = BX_GATHER; };
template <int G = 0, class R, bool W, class L, class F>
BDEV void for_pairs_g(const BoxGeo<R>& g, const PairView<R>& pv, uint32_t i, uint32_t cntw, Wrap<W> w, L&& load,
                      F&& use) {
  constexpr int NG = G ? G : Gather<R>::value;
  const uint32_t c = pcount(cntw);
  if (c <= BX_KMAX) {
    static_assert(NG % 4 == 0, "batches are whole quads");
    uint32_t k = 0;
    // whole batches of NG entries, then quads with the entries past the count switched off. With
    // NG = 4 the quad loop takes every row whole: a wave then runs max(ceil(c / 4)) quads, where
    // whole quads first and a partial one after would run max(floor(c / 4)) + 1 (rows of 9 and 12
    // pairs in one wave: 4 instead of 3).
    for (; BX_BATCH && NG > 4 && k + (uint32_t)NG <= c; k += NG) {
      uint2 qq[NG / 4];
#pragma unroll
      for (int t = 0; t < NG / 4; t++) qq[t] = tquad(pv, (k >> 2) + t, i);
      uint32_t jj[NG];
      decltype(load(0u)) aa[NG];
#pragma unroll
      for (int q = 0; q < NG; q++) jj[q] = nb_dec(qent(qq[q >> 2], q & 3), i, g.n, w);
#pragma unroll
      for (int q = 0; q < NG; q++) aa[q] = load(jj[q]);
#pragma unroll
      for (int q = 0; q < NG; q++) use(jj[q], aa[q]);
    }
    for (; k < c; k += 4u) {  // the last whole or partial quads: all four loads before the first use
      const uint2 v = tquad(pv, k >> 2, i);
      uint32_t jj[4];
      decltype(load(0u)) aa[4];
#pragma unroll
      for (int e = 0; e < 4; e++) jj[e] = k + e < c ? nb_dec(qent(v, e), i, g.n, w) : i;  // past the count: i's own
#pragma unroll
      for (int e = 0; e < 4; e++) aa[e] = load(jj[e]);
#pragma unroll
      for (int e = 0; e < 4; e++)
        if (k + e < c) use(jj[e], aa[e]);
    }
  } else {
    scan_pairs(g, pv.XP, pv.start, i, pv.XP[i], [&](uint32_t j) __attribute__((always_inline)) { use(j, load(j)); });
  }
}

// ---------------------------------------------------------------------------------------------
// grid build + reorder
// ---------------------------------------------------------------------------------------------
// Lanes of one wave holding the same key form runs (the particles arrive in the previous step's
// cell order, so most of a wave's keys repeat); one atomic per run instead of one per particle.
// Returns the run's length on its head lane and 0 elsewhere; head = lane of the run's first entry.
BDEV uint32_t wave_run(uint32_t key, uint32_t& head) {
  const int lane = threadIdx.x & 63;
  const uint32_t prev = __shfl_up(key, 1, 64);
  const bool is_head = lane == 0 || prev != key;
  const uint64_t heads = __ballot(is_head);
  const uint64_t upto = heads & (lane == 63 ? ~0ull : ((2ull << lane) - 1ull));
  head = 63u - (uint32_t)__clzll(upto);
  if (!is_head) return 0u;
  const uint64_t above = lane == 63 ? 0ull : (heads & (~0ull << (lane + 1)));
  const uint32_t next = above ? (uint32_t)(__ffsll((unsigned long long)above) - 1) : 64u;
  return next - (uint32_t)lane;
}

BXKT bx_predict(BoxGeo<R> g, Vec<R>* __restrict__ X, const Vec<R>* __restrict__ V, Vec<R>* __restrict__ X0,
                uint32_t* __restrict__ key, uint32_t* __restrict__ count) {
  if (*g.halt) return;
  const uint32_t i = blockIdx.x * BX_BLOCK + threadIdx.x;
  uint32_t k = 0xFFFFFFFFu;
  if (i < g.n) {
    Vec<R> x = X[i];
    Vec<R> v = V[i];
    // step 0: a non-finite particle restarts at rest at the origin (pbf.rs:1200-1205)
    if (!(isfinite(x.x) && isfinite(x.y) && isfinite(v.x) && isfinite(v.y))) {
      x = mk((R)0, (R)0);
      v = mk((R)0, (R)0);
    }
    X0[i] = x;
    const R vy = v.y + g.gdt;  // v.y += g dt; x += v dt (pbf.rs:1225-1235)
    x = mk(wrapR(x.x + v.x * g.dt, g.L), wrapR(x.y + vy * g.dt, g.L));
    X[i] = x;
    k = cell1(x.y, g) * g.ncx + cell1(x.x, g);
    key[i] = k;
  }
  uint32_t head;
  const uint32_t len = wave_run(k, head);
  if (len && k != 0xFFFFFFFFu) atomicAdd(&count[k], len);
}

// Exclusive scan of a[0..m) in two launches: chunk totals, then each chunk's scan offset by the sum
// of the totals before it.
BDEV uint32_t block_excl_scan(uint32_t v, uint32_t& total) {
  __shared__ uint32_t s_w[BX_BLOCK / 64];
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t x = v;
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(x, o, 64);
    if (l >= o) x += y;
  }
  if (l == 63) s_w[w] = x;
  __syncthreads();
  uint32_t before = 0, tot = 0;
  for (int k = 0; k < BX_BLOCK / 64; k++) {
    uint32_t s = s_w[k];
    if (k < w) before += s;
    tot += s;
  }
  __syncthreads();
  total = tot;
  return before + x - v;
}
// Lane t of a chunk holds items [4t, 4t + 4) as one 16-byte vector (the arrays are 256-byte
// aligned and chunks start at multiples of 4 items); a partial vector at the end is read and
// written item by item.
BDEV uint4 scan_load4(const uint32_t* __restrict__ a, uint32_t base, uint32_t m) {
  if (base + 4u <= m) return *reinterpret_cast<const uint4*>(a + base);
  uint4 v = make_uint4(0u, 0u, 0u, 0u);
  if (base < m) v.x = a[base];
  if (base + 1u < m) v.y = a[base + 1u];
  if (base + 2u < m) v.z = a[base + 2u];
  return v;
}
BXK bx_scan_totals(const uint32_t* __restrict__ halt, const uint32_t* __restrict__ a, uint32_t m,
                   uint32_t* __restrict__ bsum) {
  if (*halt) return;
  const uint32_t base = blockIdx.x * BX_SCAN_CHUNK + threadIdx.x * BX_SCAN_ITEMS;
  const uint4 v = scan_load4(a, base, m);
  uint32_t tot;
  block_excl_scan(v.x + v.y + v.z + v.w, tot);
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}
BXK bx_scan_apply(const uint32_t* __restrict__ halt, uint32_t* __restrict__ a, uint32_t m,
                  const uint32_t* __restrict__ bsum) {
  if (*halt) return;
  const uint32_t base = blockIdx.x * BX_SCAN_CHUNK + threadIdx.x * BX_SCAN_ITEMS;
  const uint4 v = scan_load4(a, base, m);  // requested before the totals are summed
  uint32_t before = 0;  // totals of the chunks before this one
  for (uint32_t k = threadIdx.x; k < blockIdx.x; k += BX_BLOCK) before += bsum[k];
  uint32_t chunks_before;
  block_excl_scan(before, chunks_before);
  uint32_t tot;
  const uint32_t off = block_excl_scan(v.x + v.y + v.z + v.w, tot) + chunks_before;
  const uint4 o = make_uint4(off, off + v.x, off + v.x + v.y, off + v.x + v.y + v.z);
  if (base + 4u <= m) {
    *reinterpret_cast<uint4*>(a + base) = o;
  } else {
    if (base < m) a[base] = o.x;
    if (base + 1u < m) a[base + 1u] = o.y;
    if (base + 2u < m) a[base + 2u] = o.z;
  }
}
BXK bx_scatter(const uint32_t* __restrict__ halt, uint32_t n, const uint32_t* __restrict__ key,
               const uint32_t* __restrict__ start, uint32_t* __restrict__ cursor, uint32_t* __restrict__ slot) {
  if (*halt) return;
  const uint32_t i = blockIdx.x * BX_BLOCK + threadIdx.x;
  const uint32_t k = i < n ? key[i] : 0xFFFFFFFFu;
  uint32_t head;
  const uint32_t len = wave_run(k, head);
  uint32_t base = 0;
  if (len && k != 0xFFFFFFFFu) base = start[k] + atomicAdd(&cursor[k], len);
  base = __shfl(base, (int)head, 64);
  if (i < n) slot[base + ((threadIdx.x & 63u) - head)] = i;
}
// One lane per cell: order the cell's slots by particle id (a cell holds a handful of particles).
// Lane 0 of the grid also resets this step's CG scalars (nothing reads them before the CG).
BXK bx_cellsort(const uint32_t* __restrict__ halt, uint32_t ncell, const uint32_t* __restrict__ start,
                uint32_t* __restrict__ slot, const uint32_t* __restrict__ id, BoxScal* __restrict__ sc,
                uint32_t cg_off, uint32_t max_cg, float tol) {
  if (*halt) return;
  const uint32_t c = blockIdx.x * BX_BLOCK + threadIdx.x;
  if (c == 0) {
    for (int p = 0; p <= BX_MAXPASS; p++) sc->dens[p] = 0ull;
    sc->done = cg_off;
    sc->it = 0u;
    sc->max_cg = max_cg;
    sc->tol = tol;
    sc->res = 0.0;
    sc->bnorm = 0.0;
    sc->target = 0.0;
  }
  if (c >= ncell) return;
  const uint32_t s = start[c], e = start[c + 1];
  for (uint32_t a = s + 1; a < e; a++) {
    uint32_t v = slot[a], iv = id[v], b = a;
    while (b > s && id[slot[b - 1]] > iv) {
      slot[b] = slot[b - 1];
      b--;
    }
    slot[b] = v;
  }
}
// Moves every particle to its sorted position (one lane per destination: coalesced writes); x is
// also written to the predicted copy XP. It also clears the next step's cell counts (the other
// cell-start buffer, whose last user was the previous step) and the scatter cursors.
BXKT bx_gather(BoxGeo<R> g, const uint32_t* __restrict__ slot, const Vec<R>* __restrict__ X,
               const Vec<R>* __restrict__ X0, const Vec<R>* __restrict__ V, const uint32_t* __restrict__ id,
               Vec<R>* __restrict__ Xb, Vec<R>* __restrict__ XPb, Vec<R>* __restrict__ X0b, Vec<R>* __restrict__ Vb,
               uint32_t* __restrict__ idb, uint32_t* __restrict__ count_next, uint32_t* __restrict__ cursor,
               uint32_t ncell) {
  if (*g.halt) return;
  const uint32_t q = blockIdx.x * BX_BLOCK + threadIdx.x;
  if (q <= ncell) count_next[q] = 0u;  // ncell + 1 <= n / 4 + 1 <= n
  if (q < ncell) cursor[q] = 0u;
  if (q >= g.n) return;
  const uint32_t i = slot[q];
  const Vec<R> x = X[i];
  Xb[q] = x;
  XPb[q] = x;
  X0b[q] = X0[i];
  Vb[q] = V[i];
  idb[q] = id[i];
}

// Records the pair table of particle i while f(j, d, r2) sees every pair; returns the count word.
// The scan writes each entry to the lane's LDS row (quad q of the row at lrow[q * BX_BLOCK], entry k
// in half-word k & 3) as it is found, then the row's quads go to the table with one 8-byte store
// each. The per-member work runs whenever any lane of the wave has a member (about 50 times per
// wave at 10 members per particle), so it is kept to the encoding, one LDS half-word store and a
// running byte offset; global half-word stores in the scan measured slower. Entries past the count
// in the last quad are left as they are (no reader goes past the count).
template <class R, bool W, class F>
BDEV uint32_t table_row(const BoxGeo<R>& g, const Vec<R>* __restrict__ XP, const uint32_t* __restrict__ start,
                        int16_t* __restrict__ nb, uint32_t* __restrict__ cnt, uint32_t i, Vec<R> xpi, Wrap<W> w,
                        uint2* lrow, F&& f) {
  uint32_t c = 0;
  // extremes of j - i and of the entries over the row: wrapped and fits follow from them after the
  // scan (an entry that does not fit is stored truncated; its row is rescanned, never read)
  int32_t rmin = 0, rmax = 0, emin = 0, emax = 0;
  char* lp = reinterpret_cast<char*>(lrow);  // entry c at byte (c / 4) * 8 * BX_BLOCK + (c % 4) * 2 of the row
  const int32_t n = (int32_t)g.n, vn = lane_copy(n), vneg_n = lane_copy(-n);
  scan_pairs_d(g, XP, start, i, xpi, w, [&](uint32_t j, Vec<R> d, R r2) __attribute__((always_inline)) {
    const int32_t raw = (int32_t)j - (int32_t)i, e = nb_off(raw, n >> 1, vn, vneg_n);
    rmin = raw < rmin ? raw : rmin;
    rmax = raw > rmax ? raw : rmax;
    emin = e < emin ? e : emin;
    emax = e > emax ? e : emax;
    if (c < BX_KMAX) {
      *reinterpret_cast<int16_t*>(lp) = (int16_t)e;
      lp += (c & 3u) == 3u ? BX_BLOCK * 8u - 6u : 2u;
    }
    c++;
    f(j, d, r2);
  });
  const uint32_t nq = ((c < BX_KMAX ? c : BX_KMAX) + 3u) >> 2;
  for (uint32_t q = 0; q < nq; q++)  // quad q of particle i: uint2 index q * n + i (byte offset < 64 n <= 2^32)
    *reinterpret_cast<uint2*>(reinterpret_cast<char*>(nb) + (q * g.n + i) * 8u) = lrow[q * BX_BLOCK];
  const bool wrapped = rmax > (n >> 1) || rmin < -(n >> 1);
  const bool fits = emin >= -BX_NB_RANGE && emax <= BX_NB_RANGE;
  const uint32_t cw = (fits ? c : (c > BX_KMAX ? c : BX_KMAX + 1u)) | (wrapped ? BX_CNT_WRAPPED : 0u);
  cnt[i] = cw;
  return cw;
}
BXKT bx_pairs(BoxGeo<R> g, const Vec<R>* __restrict__ XP, const uint32_t* __restrict__ start,
              int16_t* __restrict__ nb, uint32_t* __restrict__ cnt) {
  __shared__ uint2 s_row[(BX_KMAX / 4) * BX_BLOCK];  // each lane's table row (16 KB per workgroup)
  if (*g.halt) return;
  const uint32_t i = tile_id(g) * BX_BLOCK + threadIdx.x;
  if (i >= g.n) return;
  table_row(g, XP, start, nb, cnt, i, XP[i], Wrap<true>{}, s_row + threadIdx.x,
            [](uint32_t, Vec<R>, R) __attribute__((always_inline)) {});
}

// ---------------------------------------------------------------------------------------------
// density passes (pbf.rs:1257-1334)
// ---------------------------------------------------------------------------------------------
// Density pass A for one particle: rho, C and lambda from the pair vectors d (r2 = |d|^2).
template <class R> struct LamAcc {
  R rho;
  Vec<R> gs;
  R sgs;
};
template <class R, bool W> BDEV void lambda_term(const BoxGeo<R>& g, Wrap<W> w, Vec<R> d, R r2, LamAcc<R>& a) {
  a.rho += g.mass * poly6(r2, g);
  const R r = pair_len(r2);
  const Vec<R> gr = mul2(spiky(d, r, g, guarded(w)), g.m_over_rest);
  a.gs = add2(a.gs, gr);
  a.sgs += len2(gr);
}
template <class R> BDEV R lambda_of(const BoxGeo<R>& g, LamAcc<R> a, R& c) {
  c = mfmax(a.rho / g.rest - (R)1, (R)0);
  a.sgs += len2(a.gs);
  return -c / (a.sgs + (R)200);
}
// Bits of a non-negative real widened to 64 (they order like the values) and back.
BDEV unsigned long long rbits(float v) { return (unsigned long long)__float_as_uint(v); }
BDEV unsigned long long rbits(double v) { return (unsigned long long)__double_as_longlong(v); }
template <class R> BDEV R from_rbits(unsigned long long b);
template <> BDEV float from_rbits<float>(unsigned long long b) { return __uint_as_float((uint32_t)b); }
template <> BDEV double from_rbits<double>(unsigned long long b) { return __longlong_as_double((long long)b); }
// Folds the wave's largest constraint (c >= 0: mfmax against 0 never yields NaN) into the pass's
// slot: every lane of the wave calls it; lane 0 issues one atomic.
template <class R> BDEV void fold_max_c(const BoxGeo<R>& g, uint32_t p, R c) {
  unsigned long long b = rbits(c);
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long t = (unsigned long long)__shfl_xor((long long)b, o, 64);
    b = t > b ? t : b;
  }
  if ((threadIdx.x & 63u) == 0u && b) atomicMax(&g.dens[p], b);
}
// Whether density pass p runs this step (the reference stops after pass p - 1 once it ran at least
// `base` passes and its largest constraint was at most 0.02; a pass that did not run leaves 0).
template <class R> BDEV bool pass_live(const BoxGeo<R>& g, uint32_t p) {
  return !g.adaptive || p < g.base || from_rbits<R>(g.dens[p - 1]) > (R)0.02f;
}
// Pass A's output for pass B. FP32: the particle's position (the one pass A read) and lambda in
// one 16-byte record (x, y, lambda, 0), so pass B gathers a neighbour's position and lambda in one
// load. FP64: lambda alone (a 32-byte record measured slower than the 16 + 8 bytes it replaces).
// Both use the same buffer.
template <class R> struct PackXL { static constexpr bool value = false; };
template <> struct PackXL<float> { static constexpr bool value = true; };
template <class R> BDEV void put_lambda(void* __restrict__ LB, uint32_t i, Vec<R> x, R lam) {
  if constexpr (PackXL<R>::value) {
    Vec4<R> o;
    o.x = x.x; o.y = x.y; o.z = lam; o.w = (R)0;
    static_cast<Vec4<R>*>(LB)[i] = o;
  } else {
    static_cast<R*>(LB)[i] = lam;
  }
}
BXKT bx_lambda(BoxGeo<R> g, PairView<R> pv, const Vec<R>* __restrict__ X, void* __restrict__ LB, uint32_t p) {
  if (*g.halt || !pass_live(g, p)) return;
  const uint32_t i = tile_id(g) * BX_BLOCK + threadIdx.x;
  R c = (R)0;
  if (i < g.n) {
    const Vec<R> xi = X[i];
    const uint32_t cw = pv.cnt[i];
    LamAcc<R> a{g.w0, mk((R)0, (R)0), (R)0};
    by_wrap(wave_fast(xi, cw, g), [&](auto w) __attribute__((always_inline)) {
      for_pairs_g(g, pv, i, cw, w, [&](uint32_t j) __attribute__((always_inline)) { return at(X, j); },
                  [&](uint32_t, Vec<R> xj) __attribute__((always_inline)) {
        const Vec<R> d = rel(xi, xj, g, w);
        lambda_term(g, w, d, len2(d), a);
      });
    });
    put_lambda<R>(LB, i, xi, lambda_of(g, a, c));
  }
  if (g.adaptive) fold_max_c(g, p, c);
}
// First density pass A of a step, fused with the pair-table build: right after the reorder the
// positions X equal the predicted ones XP, so pass A runs on XP (same values, same order as
// bx_pairs followed by bx_lambda). Two phases: the 3x3-cell scan only tests membership and records
// the row (in the table and in LDS), then pass A walks the row from LDS. Summing inside the scan
// would issue the lambda arithmetic for every candidate step in which any lane of the wave has a
// member (about three times the member count).
BXKT bx_pairs_lambda(BoxGeo<R> g, const Vec<R>* __restrict__ XP, const uint32_t* __restrict__ start,
                     int16_t* __restrict__ nb, uint32_t* __restrict__ cnt, void* __restrict__ LB) {
  __shared__ uint2 s_row[(BX_KMAX / 4) * BX_BLOCK];  // each lane's table row (16 KB per workgroup)
  if (*g.halt) return;
  const uint32_t i = tile_id(g) * BX_BLOCK + threadIdx.x;
  R cmax = (R)0;
  if (i < g.n) {  // every lane of the wave reaches the constraint reduction below
  uint2* lrow = s_row + threadIdx.x;
  const Vec<R> xpi = XP[i];
  uint32_t cw = 0;
  by_wrap(wave_inside(xpi, g), [&](auto w) __attribute__((always_inline)) {
    cw = table_row(g, XP, start, nb, cnt, i, xpi, w, lrow, [](uint32_t, Vec<R>, R) __attribute__((always_inline)) {});
  });
  // pass A over the row just recorded, in its order (lanes read only their own LDS column)
  LamAcc<R> a{g.w0, mk((R)0, (R)0), (R)0};
  const uint32_t c = pcount(cw);
  by_wrap(wave_fast(xpi, cw, g), [&](auto w) __attribute__((always_inline)) {
    auto term = [&](Vec<R> xj) __attribute__((always_inline)) {
      const Vec<R> d = rel(xpi, xj, g, w);
      lambda_term(g, w, d, len2(d), a);
    };
    if (c <= BX_KMAX) {
      constexpr int NG = Gather<R>::value;
      uint32_t k = 0;
      for (; k + (uint32_t)NG <= c; k += NG) {
        uint32_t jj[NG];
        Vec<R> xx[NG];
#pragma unroll
        for (int q = 0; q < NG; q++) jj[q] = nb_dec(qent(lrow[((k >> 2) + (q >> 2)) * BX_BLOCK], q & 3), i, g.n, w);
#pragma unroll
        for (int q = 0; q < NG; q++) xx[q] = at(XP, jj[q]);
#pragma unroll
        for (int q = 0; q < NG; q++) term(xx[q]);
      }
      for (; k < c; k += 4u) {  // as in for_pairs_g: the quad's loads first, past the count i's own
        const uint2 v = lrow[(k >> 2) * BX_BLOCK];
        Vec<R> xx[4];
#pragma unroll
        for (int e = 0; e < 4; e++) xx[e] = at(XP, k + e < c ? nb_dec(qent(v, e), i, g.n, w) : i);
#pragma unroll
        for (int e = 0; e < 4; e++)
          if (k + e < c) term(xx[e]);
      }
    } else {
      scan_pairs(g, XP, start, i, xpi, [&](uint32_t j) __attribute__((always_inline)) { term(XP[j]); });
    }
  });
  put_lambda<R>(LB, i, xpi, lambda_of(g, a, cmax));
  }
  if (g.adaptive) fold_max_c(g, 0u, cmax);
}
// Pass B (FP32: positions and lambdas from pass A's records, which hold the positions X).
BXKT bx_dp(BoxGeo<R> g, PairView<R> pv, const Vec<R>* __restrict__ X, const void* __restrict__ LB,
           Vec<R>* __restrict__ Xout, uint32_t p) {
  if (*g.halt) return;
  const uint32_t i = tile_id(g) * BX_BLOCK + threadIdx.x;
  if (!pass_live(g, p)) {  // a pass that does not run leaves the positions as they are
    if (i < g.n) Xout[i] = X[i];
    return;
  }
  if (g.adaptive && blockIdx.x == 0 && threadIdx.x == 0) g.dens[BX_MAXPASS] = p + 1u;
  if (i >= g.n) return;
  Vec<R> xi;
  R li;
  const Vec4<R>* const XL = static_cast<const Vec4<R>*>(LB);
  const R* const LAM = static_cast<const R*>(LB);
  if constexpr (PackXL<R>::value) {
    const Vec4<R> oi = XL[i];
    xi = mk(oi.x, oi.y);
    li = oi.z;
  } else {
    xi = X[i];
    li = LAM[i];
  }
  Vec<R> s = mk((R)0, (R)0);
  const uint32_t cw = pv.cnt[i];
  auto term = [&](auto w, Vec<R> xj, R lj) __attribute__((always_inline)) {
    const Vec<R> d = rel(xi, xj, g, w);
    const R r = pair_len(len2(d));
    s = add2(s, smul2(li + lj, spiky(d, r, g, guarded(w))));
  };
  by_wrap(wave_fast(xi, cw, g), [&](auto w) __attribute__((always_inline)) {
    if constexpr (PackXL<R>::value) {
      for_pairs_g(g, pv, i, cw, w, [&](uint32_t j) __attribute__((always_inline)) { return at(XL, j); },
                  [&](uint32_t, const Vec4<R>& a) __attribute__((always_inline)) { term(w, mk(a.x, a.y), a.z); });
    } else {
      struct Nb { Vec<R> x; R l; };
      for_pairs_g(g, pv, i, cw, w, [&](uint32_t j) __attribute__((always_inline)) { return Nb{at(X, j), at(LAM, j)}; },
                  [&](uint32_t, const Nb& a) __attribute__((always_inline)) { term(w, a.x, a.l); });
    }
  });
  const Vec<R> dp = mul2(s, g.m_over_rest);
  Xout[i] = mk(wrapR(xi.x + dp.x, g.L), wrapR(xi.y + dp.y, g.L));
}

// ---------------------------------------------------------------------------------------------
// velocity, density recompute, implicit viscosity
// ---------------------------------------------------------------------------------------------
// part_ke (optional): per-block sums of the f64-widened |v|^2 (kinetic_energy_f64, pbf.rs:924-926).
BXKRT bx_vel_density(BoxGeo<R> g, PairView<R> pv, const Vec<R>* __restrict__ X, const Vec<R>* __restrict__ X0,
                     Vec<R>* __restrict__ V, R* __restrict__ DENS, double* __restrict__ part_b,
                     double* __restrict__ part_ke) {
  if (*g.halt) return;
  const uint32_t ppt = red_ppt(g);
  const uint32_t base = tile_id_red(g) * BX_RBLOCK * ppt + threadIdx.x;
  double vv = 0.0, ke = 0.0;
  for_lane_particles(g, base, [&](uint32_t i, uint32_t k) __attribute__((always_inline)) {
    const Vec<R> xi = X[i];
    const Vec<R> v = div2(rel(xi, X0[i], g), g.dt);
    V[i] = v;
    vv += (double)v.x * (double)v.x + (double)v.y * (double)v.y;
    ke += (double)len2(v);
    R rho = g.w0;
    const uint32_t cw = pv.cnt[i];
    by_wrap(wave_fast(xi, cw, g), [&](auto w) __attribute__((always_inline)) {
      for_pairs_g(g, pv, i, cw, w, [&](uint32_t j) __attribute__((always_inline)) { return at(X, j); },
                  [&](uint32_t, Vec<R> xj) __attribute__((always_inline)) {
        rho += g.mass * poly6(len2(rel(xi, xj, g, w)), g);
      });
    });
    DENS[i] = rho;
  });
  const double t = block_sum<BX_RBLOCK>(vv);
  if (threadIdx.x == 0) part_b[blockIdx.x] = t;
  if (part_ke) {
    const double e = block_sum<BX_RBLOCK>(ke);
    if (threadIdx.x == 0) part_ke[blockIdx.x] = e;
  }
}

// After the CG: the per-block f64 sums of |v|^2 before (part_a) and after (part_c) the fluid speed
// clamp (clamp_fluid_speed, pbf.rs:987-996: a non-finite speed zeroes v, a speed above v_max scales
// v by v_max / speed), which runs when v_max > 0.
BXKRT bx_clamp_ke(BoxGeo<R> g, Vec<R>* __restrict__ V, R v_max, double* __restrict__ part_a,
                  double* __restrict__ part_c) {
  if (*g.halt) return;
  const uint32_t ppt = red_ppt(g);
  const uint32_t base = tile_id_red(g) * BX_RBLOCK * ppt + threadIdx.x;
  double a = 0.0, c = 0.0;
  for_lane_particles(g, base, [&](uint32_t i, uint32_t) __attribute__((always_inline)) {
    Vec<R> v = V[i];
    a += (double)len2(v);
    if (v_max > (R)0) {
      const R speed = msqrt(len2(v));
      if (!isfinite(speed)) v = mk((R)0, (R)0);
      else if (speed > v_max) v = mul2(v, v_max / speed);
      V[i] = v;
    }
    c += (double)len2(v);
  });
  if (!part_a) return;  // no energy sums asked for (uniform: the whole workgroup returns)
  const double ta = block_sum<BX_RBLOCK>(a);
  const double tc = block_sum<BX_RBLOCK>(c);
  if (threadIdx.x == 0) {
    part_a[blockIdx.x] = ta;
    part_c[blockIdx.x] = tc;
  }
}

// Shear rate sqrt(2 D:D) of particle i from the SPH velocity gradient over the step's pairs, at the
// final positions, the given velocities and the recomputed densities (pbf.rs:821-849), summed in the
// table's (reference) pair order.
template <class R>
BDEV R shear_rate(const BoxGeo<R>& g, const PairView<R>& pv, uint32_t i, const Vec<R>* __restrict__ X,
                  const Vec<R>* __restrict__ V, const R* __restrict__ DENS) {
  const Vec<R> xi = X[i], vi = V[i];
  R gxx = (R)0, gxy = (R)0, gyx = (R)0, gyy = (R)0;
  const uint32_t cw = pv.cnt[i];
  struct Nb { Vec<R> x; Vec<R> v; R d; };
  by_wrap(wave_fast(xi, cw, g), [&](auto w) __attribute__((always_inline)) {
    for_pairs_g<4>(g, pv, i, cw, w,
                   [&](uint32_t j) __attribute__((always_inline)) { return Nb{at(X, j), at(V, j), at(DENS, j)}; },
                   [&](uint32_t, const Nb& a) __attribute__((always_inline)) {
      const Vec<R> d = rel(xi, a.x, g, w);
      const R r = pair_len(len2(d));
      const Vec<R> gr = spiky(d, r, g, guarded(w));
      const R coeff = g.mass / mfmax(a.d, (R)1e-6f);
      const Vec<R> dv = sub2(a.v, vi);
      gxx += coeff * dv.x * gr.x;
      gxy += coeff * dv.x * gr.y;
      gyx += coeff * dv.y * gr.x;
      gyy += coeff * dv.y * gr.y;
    });
  });
  const R dxy = (R)0.5f * (gxy + gyx);
  return msqrt((R)2 * (gxx * gxx + gyy * gyy + (R)2 * dxy * dxy));
}
// Effective viscosity of a Bingham box (pbf.rs:2147-2151): the Papanastasiou-regularised
// mu + tau_y (1 - exp(-g / g0)) / g of each particle's shear rate g, with g0 = tau_y / (100 mu), its
// g -> 0 limit below 1e-4 g0 and a non-finite rate taken as 0 (pbf.rs:854-868, :490).
// (FP64: no 8-waves pin, under which its 40 bytes of gathers in flight per neighbour quad spill)
template <class R> __global__ void __launch_bounds__(BX_BLOCK, BxRWaves<R>::value)
bx_shear_mu(BoxGeo<R> g, PairView<R> pv, const Vec<R>* __restrict__ X, const Vec<R>* __restrict__ V,
                 const R* __restrict__ DENS, R* __restrict__ MU, R mu, R tau_y) {
  if (*g.halt) return;
  const uint32_t i = tile_id(g) * BX_BLOCK + threadIdx.x;
  if (i >= g.n) return;
  R rate = shear_rate(g, pv, i, X, V, DENS);
  const R g0 = tau_y / ((R)100 * mu);
  if (!isfinite(rate)) rate = (R)0;
  const R extra = rate < (R)1e-4f * g0 ? tau_y / g0 : tau_y * ((R)1 - mexp(-rate / g0)) / rate;
  MU[i] = mu + extra;
}
// Mean shear rate of the last step (pbf.rs:800-818, reported at :2303-2304): per reduction block the
// FP64 sum of the finite rates (part_sum) and their count (part_cnt).
BXKRT bx_shear_mean(BoxGeo<R> g, PairView<R> pv, const Vec<R>* __restrict__ X, const Vec<R>* __restrict__ V,
                    const R* __restrict__ DENS, double* __restrict__ part_sum, double* __restrict__ part_cnt) {
  const uint32_t ppt = red_ppt(g);
  const uint32_t base = tile_id_red(g) * BX_RBLOCK * ppt + threadIdx.x;
  double sum = 0.0, cnt = 0.0;
  for_lane_particles(g, base, [&](uint32_t i, uint32_t) __attribute__((always_inline)) {
    const R rate = shear_rate(g, pv, i, X, V, DENS);
    if (isfinite(rate)) {
      sum += (double)rate;
      cnt += 1.0;
    }
  });
  const double ts = block_sum<BX_RBLOCK>(sum);
  const double tc = block_sum<BX_RBLOCK>(cnt);
  if (threadIdx.x == 0) {
    part_sum[blockIdx.x] = ts;
    part_cnt[blockIdx.x] = tc;
  }
}

// Morris weight and unit direction of pair (i, j) (pbf.rs:312-344).
// The quotient d / r is formed once: the kernel gradient (d / r) * coeff and the direction d / r
// use the same rounded value, exactly as two separate divisions would give. pair_unit forms the
// same direction alone (the CG recomputes it from the unchanged positions instead of storing it).
// mu2m = mass * (mu_i + mu_j): g.mu2m for a Newtonian box, else from the effective viscosities.
template <class R, bool W>
BDEV void morris(const BoxGeo<R>& g, Wrap<W> wr, Vec<R> xi, R rho_i, Vec<R> xj, R dens_j, R mu2m, R& w, Vec<R>& e) {
  const R rho_j = mfmax(dens_j, (R)1e-6f);
  const Vec<R> d = rel(xi, xj, g, wr);
  const R r = pair_len(len2(d));
  const Vec<R> dr = pair_div(d, r, guarded(wr));
  const Vec<R> gr = (r <= (R)1e-9f || r >= g.h) ? mk((R)0, (R)0) : mul2(dr, g.ks * (g.h - r) * (g.h - r));
  const R num = (mu2m / (rho_i * rho_j)) * mfabs(dot2(d, gr));
  w = num / (len2(d) + g.eta2);
  e = r > (R)1e-9f ? dr : mk((R)1, (R)0);
}
template <class R> BDEV Vec<R> pair_unit(Vec<R> d, bool guard) {
  const R r = pair_len(len2(d));
  return r > (R)1e-9f ? pair_div(d, r, guard) : mk((R)1, (R)0);
}
// (L v)_i with the weights computed here; the weights of a tabled particle are kept for the
// operator applications of this step's CG, in the table's quad layout: the four weights of one
// quad of table entries leave as one vector store.
// BG: Bingham box, the pair's viscosity sum from the effective viscosities MU (pbf.rs:2147-2154).
template <class R, bool BG>
BDEV Vec<R> lap_row(const BoxGeo<R>& g, uint32_t i, const PairView<R>& pv, const Vec<R>* __restrict__ X,
                    const R* __restrict__ DENS, const Vec<R>* __restrict__ in, R* __restrict__ CW,
                    const R* __restrict__ MU) {
  const Vec<R> xi = X[i], ini = in[i];
  const R rho_i = mfmax(DENS[i], (R)1e-6f);
  R mu_i = (R)0;
  if constexpr (BG) mu_i = MU[i];
  const uint32_t cw = pv.cnt[i], c = pcount(cw);
  Vec<R> s = mk((R)0, (R)0);
  Vec4<R>* const cwq = reinterpret_cast<Vec4<R>*>(CW);
  struct Nb { Vec<R> x; R d; Vec<R> v; R mu; };
  by_wrap(wave_fast(xi, cw, g), [&](auto wr) __attribute__((always_inline)) {
    auto load = [&](uint32_t j) __attribute__((always_inline)) {
      if constexpr (BG) return Nb{at(X, j), at(DENS, j), at(in, j), at(MU, j)};
      else return Nb{at(X, j), at(DENS, j), at(in, j), (R)0};
    };
    auto term = [&](const Nb& a) __attribute__((always_inline)) {
      R w;
      Vec<R> e;
      morris(g, wr, xi, rho_i, a.x, a.d, BG ? g.mass * (mu_i + a.mu) : g.mu2m, w, e);
      s = add2(s, smul2((R)4 * w * dot2(e, sub2(ini, a.v)), e));
      return w;
    };
    if (c <= BX_KMAX) {
      // one quad of table entries per batch, entries past the count switched off (loads of
      // particle i's own data, weight 0), so a wave runs max(ceil(c / 4)) batches; the next batch's
      // quad is requested one batch ahead
      uint2 vnext = tquad(pv, 0u, i);
      for (uint32_t k = 0; k < c; k += 4u) {
        const uint2 v = vnext;
        if (k + 4u < c) vnext = tquad(pv, (k >> 2) + 1u, i);
        Nb aa[4];
#pragma unroll
        for (int q = 0; q < 4; q++) aa[q] = load(k + q < c ? nb_dec(qent(v, q), i, g.n, wr) : i);
        Vec<R>* const wq = reinterpret_cast<Vec<R>*>(cwq + (size_t)(k >> 2) * g.n + i);
        R w0 = (R)0, w1 = (R)0, w2 = (R)0, w3 = (R)0;
        if (k < c) w0 = term(aa[0]);
        if (k + 1u < c) w1 = term(aa[1]);
        wq[0] = mk(w0, w1);
        if (k + 2u < c) w2 = term(aa[2]);
        if (k + 3u < c) w3 = term(aa[3]);
        wq[1] = mk(w2, w3);
      }
    } else {
      scan_pairs(g, pv.XP, pv.start, i, pv.XP[i], [&](uint32_t j) __attribute__((always_inline)) { term(load(j)); });
    }
  });
  return s;
}

template <class R, bool BG>
BDEV void cg_init_body(const BoxGeo<R>& g, const PairView<R>& pv, const Vec<R>* __restrict__ X,
                       const R* __restrict__ DENS, const Vec<R>* __restrict__ V, Vec<R>* __restrict__ Rv,
                       Vec<R>* __restrict__ P, R* __restrict__ CW, double* __restrict__ part_rr0,
                       const BoxScal* __restrict__ sc, const R* __restrict__ MU) {
  if (*g.halt | sc->done) return;  // both words loaded at once
  const uint32_t ppt = red_ppt(g);
  const uint32_t base = tile_id_red(g) * BX_RBLOCK * ppt + threadIdx.x;
  double rr = 0.0;
  for_lane_particles(g, base, [&](uint32_t i, uint32_t k) __attribute__((always_inline)) {
    const Vec<R> v = V[i];
    const Vec<R> lv = lap_row<R, BG>(g, i, pv, X, DENS, V, CW, MU);
    const Vec<R> r = sub2(v, add2(v, mul2(lv, g.dt)));
    Rv[i] = r;
    P[i] = r;
    rr += (double)r.x * (double)r.x + (double)r.y * (double)r.y;
  });
  const double t = block_sum<BX_RBLOCK>(rr);
  if (threadIdx.x == 0) part_rr0[blockIdx.x] = t;
}
BXKRT bx_cg_init(BoxGeo<R> g, PairView<R> pv, const Vec<R>* __restrict__ X, const R* __restrict__ DENS,
                 const Vec<R>* __restrict__ V, Vec<R>* __restrict__ Rv, Vec<R>* __restrict__ P,
                 R* __restrict__ CW, double* __restrict__ part_rr0, const BoxScal* __restrict__ sc) {
  cg_init_body<R, false>(g, pv, X, DENS, V, Rv, P, CW, part_rr0, sc, nullptr);
}
// The Bingham box's set-up: the Morris weights from the effective viscosities MU.
BXKRT bx_cg_init_bingham(BoxGeo<R> g, PairView<R> pv, const Vec<R>* __restrict__ X, const R* __restrict__ DENS,
                         const Vec<R>* __restrict__ V, Vec<R>* __restrict__ Rv, Vec<R>* __restrict__ P,
                         R* __restrict__ CW, double* __restrict__ part_rr0, const BoxScal* __restrict__ sc,
                         const R* __restrict__ MU) {
  cg_init_body<R, true>(g, pv, X, DENS, V, Rv, P, CW, part_rr0, sc, MU);
}

// CG iteration t (pbf.rs:423-480), first half. Every block evaluates the scalars from the partial
// sums of the previous kernels (same fixed order everywhere): for t = 1 the target tol·|b| and the
// initial residual; for t > 1 the convergence test on r(t-1).r(t-1), the iteration cap, and
// beta = rs(t-1)/rs(t-2), with rs(t-2) read from where iteration t-1 stored it. The direction p = r + beta p_old is formed on the fly for the particle
// and its neighbours (no separate direction kernel), then Ap = p + dt L p with the kept weights
// and the unit directions recomputed from the positions (the same values the set-up used).
BXKRT bx_cg_dir(BoxGeo<R> g, PairView<R> pv, const Vec<R>* __restrict__ X, const R* __restrict__ DENS,
                const R* __restrict__ CW, const Vec<R>* __restrict__ Rv,
                const Vec<R>* __restrict__ Pold, Vec<R>* __restrict__ Pnew, Vec<R>* __restrict__ AP,
                const double* __restrict__ part_b, const double* __restrict__ part_rr_prev,
                double* __restrict__ part_pap, BoxScal* __restrict__ sc, uint32_t t) {
  if (*g.halt | sc->done) return;  // both words loaded at once
  const uint32_t ppt = red_ppt(g);
  const uint32_t base = tile_id_red(g) * BX_RBLOCK * ppt + threadIdx.x;
  const bool first = t == 1u;
  // the first particle's own inputs are requested before the scalar reduction (they do not depend
  // on it)
  Vec<R> xi0 = mk((R)0, (R)0), ri0 = xi0, pi0 = xi0;
  uint32_t cw0 = 0;
  if (base < g.n) {
    xi0 = X[base];
    pi0 = Pold[base];
    if (!first) ri0 = Rv[base];
    cw0 = pv.cnt[base];
  }
  double target;
  bool stop = false;
  R beta = (R)0;
  if (first) {
    const double2 s2 = sum_parts2<BX_RBLOCK>(part_b, part_rr_prev, g.nred);
    const double bn = sqrt(s2.x), rs0 = s2.y;
    target = (double)sc->tol * bn;
    if (!(bn > 0.0) || !isfinite(bn) || sqrt(rs0) <= target) stop = true;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      sc->bnorm = bn;
      sc->target = target;
      sc->res = (bn > 0.0 && isfinite(bn)) ? sqrt(rs0) : 0.0;
      sc->rr[0] = rs0;
    }
  } else {
    target = sc->target;
    // r(t-1).r(t-1) from the partials of the last update; r(t-2).r(t-2) as the previous
    // iteration's kernel summed it (the same value, stored in the other slot)
    const double rs2 = sc->rr[t & 1u];
    const double rs1 = sum_parts<BX_RBLOCK>(part_rr_prev, g.nred);
    if (sqrt(rs1) <= target || t - 1u >= sc->max_cg) stop = true;
    beta = (R)(rs1 / rs2);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      sc->res = sqrt(rs1);
      sc->rr[(t - 1u) & 1u] = rs1;
    }
  }
  if (stop) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      sc->it = t - 1u;
      sc->done = 1u;
    }
    return;
  }
  double pap = 0.0;
  for_lane_particles(g, base, [&](uint32_t i, uint32_t kp) __attribute__((always_inline)) {
    const Vec<R> xi = kp ? X[i] : xi0;
    const Vec<R> pi = kp ? Pold[i] : pi0;
    const Vec<R> ri = (kp && !first) ? Rv[i] : ri0;
    const uint32_t cw = kp ? pv.cnt[i] : cw0;
    const Vec<R> p = first ? pi : add2(ri, mul2(pi, beta));
    Vec<R> s = mk((R)0, (R)0);
    const uint32_t c = pcount(cw);
    by_wrap(wave_fast(xi, cw, g), [&](auto wr) __attribute__((always_inline)) {
      if (c <= BX_KMAX) {
        auto pdir = [&](uint32_t j) __attribute__((always_inline)) {
          return first ? at(Pold, j) : add2(at(Rv, j), mul2(at(Pold, j), beta));
        };
        auto term = [&](R wk, Vec<R> pj, Vec<R> xj) __attribute__((always_inline)) {
          const Vec<R> e = pair_unit<R>(rel(xi, xj, g, wr), guarded(wr));
          s = add2(s, smul2((R)4 * wk * dot2(e, sub2(p, pj)), e));
        };
        // the kept weights are in the table's quad layout: one vector load per four pairs
        auto wquad = [&](uint32_t q) __attribute__((always_inline)) {
          return at(reinterpret_cast<const Vec4<R>*>(CW) + (size_t)q * g.n, i);
        };
        // one quad of the table per batch, entries past the count switched off (loads of particle
        // i's own data), so a wave runs max(ceil(c / 4)) batches
        for (uint32_t k = 0; k < c; k += 4u) {
          const uint2 v = tquad(pv, k >> 2, i);
          const Vec4<R> w4 = wquad(k >> 2);
          uint32_t jj[4];
          Vec<R> pp[4], xx[4];
#pragma unroll
          for (int e = 0; e < 4; e++) jj[e] = k + e < c ? nb_dec(qent(v, e), i, g.n, wr) : i;
#pragma unroll
          for (int e = 0; e < 4; e++) pp[e] = pdir(jj[e]);
#pragma unroll
          for (int e = 0; e < 4; e++) xx[e] = at(X, jj[e]);
#pragma unroll
          for (int e = 0; e < 4; e++)
            if (k + e < c) term(v4c(w4, e), pp[e], xx[e]);
        }
      } else {
        const R rho_i = mfmax(DENS[i], (R)1e-6f);
        const R* const MU = static_cast<const R*>(sc->mu_eff);  // a Bingham box's effective viscosities
        scan_pairs(g, pv.XP, pv.start, i, pv.XP[i], [&](uint32_t j) __attribute__((always_inline)) {
          const Vec<R> pj = first ? Pold[j] : add2(Rv[j], mul2(Pold[j], beta));
          R w;
          Vec<R> e;
          morris(g, wr, xi, rho_i, X[j], DENS[j], MU ? g.mass * (MU[i] + MU[j]) : g.mu2m, w, e);
          s = add2(s, smul2((R)4 * w * dot2(e, sub2(p, pj)), e));
        });
      }
    });
    const Vec<R> ap = add2(p, mul2(s, g.dt));
    Pnew[i] = p;
    AP[i] = ap;
    pap += (double)p.x * (double)ap.x + (double)p.y * (double)ap.y;
  });
  const double tt = block_sum<BX_RBLOCK>(pap);
  if (threadIdx.x == 0) part_pap[blockIdx.x] = tt;
  if (blockIdx.x == 0 && threadIdx.x == 0) sc->it = t;
}

// CG iteration t, second half: alpha = rs(t-1)/p.Ap (stop if p.Ap <= 0), v += alpha p,
// r -= alpha Ap, partial sums of r.r for the next iteration's test.
BXKRT bx_cg_upd(BoxGeo<R> g, Vec<R>* __restrict__ V, Vec<R>* __restrict__ Rv, const Vec<R>* __restrict__ P,
                const Vec<R>* __restrict__ AP, const double* __restrict__ part_pap,
                double* __restrict__ part_rr_out, BoxScal* __restrict__ sc, uint32_t t) {
  if (*g.halt | sc->done) return;  // both words loaded at once
  const uint32_t ppt = red_ppt(g);
  const uint32_t base = blockIdx.x * BX_RBLOCK * ppt + threadIdx.x;
  // the first particle's vectors are requested before the scalar reduction (they do not depend on it)
  Vec<R> vi0 = mk((R)0, (R)0), pi0 = vi0, ri0 = vi0, api0 = vi0;
  if (base < g.n) {
    vi0 = V[base];
    pi0 = P[base];
    ri0 = Rv[base];
    api0 = AP[base];
  }
  const double rs_old = sc->rr[(t - 1u) & 1u];  // r(t-1).r(t-1), stored by this iteration's first half
  const double pap = sum_parts<BX_RBLOCK>(part_pap, g.nred);
  if (!(pap > 0.0) || !isfinite(pap)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      sc->it = t;
      sc->done = 1u;
    }
    return;
  }
  const R a = (R)(rs_old / pap);
  double rr = 0.0;
  for_lane_particles(g, base, [&](uint32_t i, uint32_t k) __attribute__((always_inline)) {
    const Vec<R> vi = k ? V[i] : vi0, pi = k ? P[i] : pi0, ri = k ? Rv[i] : ri0, api = k ? AP[i] : api0;
    V[i] = add2(vi, mul2(pi, a));
    const Vec<R> r = sub2(ri, mul2(api, a));
    Rv[i] = r;
    rr += (double)r.x * (double)r.x + (double)r.y * (double)r.y;
  });
  const double tt = block_sum<BX_RBLOCK>(rr);
  if (threadIdx.x == 0) part_rr_out[blockIdx.x] = tt;
}

// End of a chunk of iterations ending with iteration t: the convergence test and the cap on
// r(t).r(t), so the host sees `done` without another iteration being launched.
// In a multi-step window it also records the step's verdict: verdict[1 + s] = iterations, and
// verdict[0] = the first step s whose CG was still running when its launched iterations ran out;
// that step also sets the stop word, so the rest of the window does no work.
BXKR bx_cg_check(uint32_t nred, const double* __restrict__ part_rr_last, BoxScal* __restrict__ sc, uint32_t t,
                 uint32_t* __restrict__ verdict, uint32_t s, uint32_t* __restrict__ halt) {
  if (*halt) return;
  if (!sc->done) {
    const double rs = sum_parts<BX_RBLOCK>(part_rr_last, nred);
    if (threadIdx.x == 0) {
      sc->res = sqrt(rs);
      if (sqrt(rs) <= sc->target || t >= sc->max_cg) {
        sc->it = t;
        sc->done = 1u;
      }
    }
  }
  if (threadIdx.x == 0 && verdict) {
    verdict[1 + s] = sc->it;
    if (!sc->done) {
      verdict[0] = s;
      *halt = 1u;
    }
  }
}

// ---------------------------------------------------------------------------------------------
// multi-step windows: the per-step CG verdict and the stop word
// ---------------------------------------------------------------------------------------------
BXK bx_verdict_reset(uint32_t* __restrict__ verdict, uint32_t* __restrict__ halt) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    verdict[0] = 0xFFFFFFFFu;
    *halt = 0u;
  }
}

// ---------------------------------------------------------------------------------------------
// debug export: neighbour lists in the traversal order, rows and entries as original ids
// ---------------------------------------------------------------------------------------------
BXKT bx_nbr_count(BoxGeo<R> g, PairView<R> pv, const uint32_t* __restrict__ id, uint32_t* __restrict__ cnt) {
  const uint32_t i = blockIdx.x * BX_BLOCK + threadIdx.x;
  if (i >= g.n) return;
  uint32_t c = 0;
  for_pairs(g, pv, i, [&](uint32_t) __attribute__((always_inline)) { c++; });
  cnt[id[i]] = c;
}
BXKT bx_nbr_write(BoxGeo<R> g, PairView<R> pv, const uint32_t* __restrict__ id, const uint32_t* __restrict__ off,
                  uint32_t* __restrict__ nb, uint32_t cap) {
  const uint32_t i = blockIdx.x * BX_BLOCK + threadIdx.x;
  if (i >= g.n) return;
  uint32_t w = off[id[i]];
  for_pairs(g, pv, i, [&](uint32_t j) __attribute__((always_inline)) {
    if (w < cap) nb[w] = id[j];
    w++;
  });
}

// ---------------------------------------------------------------------------------------------
// self-check: the pair loops' f32 square root and quotient against the IEEE operations
// ---------------------------------------------------------------------------------------------
BDEV uint64_t mix64(uint64_t& s) {  // splitmix64
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
// A float with the given exponent range [e0, e1] and a random significand (every 8th draw all
// ones or all zeros, the significands where a reciprocal's rounding is hardest).
BDEV float draw_float(uint64_t& s, int e0, int e1) {
  const uint64_t z = mix64(s);
  const int e = e0 + (int)((z >> 32) % (uint64_t)(e1 - e0 + 1));
  uint32_t m = (uint32_t)z & 0x7FFFFFu;
  const uint32_t sel = (uint32_t)(z >> 23) & 7u;
  if (sel == 0u) m = 0x7FFFFFu;
  if (sel == 1u) m = 0u;
  return __uint_as_float((uint32_t)(e + 127) << 23 | m);
}
BDEV double draw_double(uint64_t& s, int e0, int e1) {
  const uint64_t z = mix64(s), w = mix64(s);
  const int e = e0 + (int)((z >> 32) % (uint64_t)(e1 - e0 + 1));
  uint64_t m = w & 0xFFFFFFFFFFFFFull;
  const uint32_t sel = (uint32_t)z & 7u;
  if (sel == 0u) m = 0xFFFFFFFFFFFFFull;
  if (sel == 1u) m = 0ull;
  return __longlong_as_double((long long)((uint64_t)(e + 1023) << 52 | m));
}
BXK bx_check_arith(uint64_t seed, uint32_t per_lane, unsigned long long* __restrict__ bad) {
  uint64_t s = seed ^ ((uint64_t)(blockIdx.x * BX_BLOCK + threadIdx.x) * 0xD1B54A32D192ED03ull);
  unsigned long long bad_sqrt = 0, bad_div = 0, bad_sqrt64 = 0, bad_div64 = 0;
  for (uint32_t k = 0; k < per_lane; k++) {
    // square root over [2^-96, 2^60], and below 2^-96 the promise that the value stays < 1e-9
    const float x = draw_float(s, -96, 60);
    if (__float_as_uint(pair_len<float>(x)) != __float_as_uint(sqrtf(x))) bad_sqrt++;
    const float xt = draw_float(s, -126, -97);
    if (!(pair_len<float>(xt) < 1e-9f)) bad_sqrt++;
    // quotients d / r with r in [2^-30, 2^29] and |d.x|, |d.y| <= r down to 2^-100 r (and zero)
    const float r = draw_float(s, -30, 29);
    const uint64_t z = mix64(s);
    float ux = draw_float(s, -100, -1), uy = draw_float(s, -100, -1);
    if ((z & 15u) == 0u) ux = 1.0f;
    if ((z & 15u) == 1u) uy = 0.0f;
    const float2 d = make_float2((z & 16u) ? -r * ux : r * ux, (z & 32u) ? -r * uy : r * uy);
    const float2 q = pair_div<float>(d, r);
    if (__float_as_uint(q.x) != __float_as_uint(d.x / r) || __float_as_uint(q.y) != __float_as_uint(d.y / r))
      bad_div++;
    // the same in f64: square root over [2^-767, 2^60], quotients with |d| down to 2^-940 r
    const double x64 = draw_double(s, -767, 60);
    if (__double_as_longlong(pair_len<double>(x64)) != __double_as_longlong(sqrt(x64))) bad_sqrt64++;
    const double xt64 = draw_double(s, -1022, -768);
    if (!(pair_len<double>(xt64) < 1e-9)) bad_sqrt64++;
    const double r64 = draw_double(s, -30, 29);
    const uint64_t z64 = mix64(s);
    double vx = draw_double(s, -940, -1), vy = draw_double(s, -60, -1);
    if ((z64 & 15u) == 0u) vx = 1.0;
    if ((z64 & 15u) == 1u) vy = 0.0;
    const double2 d64 = make_double2((z64 & 16u) ? -r64 * vx : r64 * vx, (z64 & 32u) ? -r64 * vy : r64 * vy);
    const double2 q64 = pair_div<double>(d64, r64);
    if (__double_as_longlong(q64.x) != __double_as_longlong(d64.x / r64) ||
        __double_as_longlong(q64.y) != __double_as_longlong(d64.y / r64))
      bad_div64++;
  }
  if (bad_sqrt) atomicAdd(&bad[0], bad_sqrt);
  if (bad_div) atomicAdd(&bad[1], bad_div);
  if (bad_sqrt64) atomicAdd(&bad[2], bad_sqrt64);
  if (bad_div64) atomicAdd(&bad[3], bad_div64);
}

// ---------------------------------------------------------------------------------------------
// host runtime
// ---------------------------------------------------------------------------------------------
struct mg_box {
  mg_box_config cfg;
  bool f64 = false;       // state precision (cfg.fp64)
  BoxGeo<float> gf;       // geometry in f32 (the workload's definition)
  BoxGeo<double> gd;      // and promoted to f64 for an FP64 box
  uint32_t n = 0, ncell = 0, nscan = 0, nred = 0;
  // state arrays in the box's precision: positions / velocities are Vec<R>, scalars R
  void *X = nullptr, *XP = nullptr, *X0 = nullptr, *V = nullptr, *X2 = nullptr;
  void *Xb = nullptr, *X0b = nullptr, *Vb = nullptr;
  void* VN = nullptr;  // the step's new velocity (velocity update + CG); swapped with V at step end
  uint32_t *id = nullptr, *idb = nullptr, *key = nullptr, *slot = nullptr, *cursor = nullptr, *bsum = nullptr;
  // cell starts of the current step and the buffer the current step clears for the next one
  uint32_t *start = nullptr, *start_next = nullptr;
  void *LAM = nullptr, *DENS = nullptr;  // LAM: pass A's output for pass B (see put_lambda)
  void* MU = nullptr;  // effective viscosity per particle (Bingham boxes: cfg.yield_stress > 0)
  void *R = nullptr, *P = nullptr, *AP = nullptr;
  int16_t* pnb = nullptr;    // per-step pair table (16-bit relative offsets)
  uint32_t* pcnt = nullptr;  // and the per-particle pair counts
  void* cw = nullptr;        // per-step Morris weights of tabled pairs
  double* part_b = nullptr;    // |b|^2 partials (one per reduction block)
  double* part_rr = nullptr;   // r.r partials, two generations: [0, nred) and [nred, 2 nred)
  double* part_pap = nullptr;  // p.Ap partials
  double* part_ke = nullptr;   // energy_stats: |v|^2 partials before the CG, after it, after the clamp
  void* P2 = nullptr;          // second direction buffer (the CG alternates P and P2)
  uint32_t cg_t = 0;           // next CG iteration number of the current step
  BoxScal* sc = nullptr;
  BoxScal* h_sc = nullptr;  // pinned host mirror
  void* arena = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool have_neighbors = false;
  uint32_t last_it = 0;  // CG iterations of the previous step (sizes the first chunk)
  uint64_t total_it = 0;  // CG iterations summed over every step run
  // multi-step windows: the device verdict [first overrun step, iterations of step 0..W-1] and its
  // pinned host copy; where each step's start state lives (see run_window)
  bool cg_stable = false;  // the last window's steps all took the same CG iteration count
  struct Start { void* x; void* v; uint32_t* id; uint32_t* cells; };
  Start starts[BX_WINDOW];
  uint32_t* verdict = nullptr;
  uint32_t* h_verdict = nullptr;
  uint32_t* halt = nullptr;  // window stop word (BoxGeo::halt)
  uint32_t win = BX_WINDOW;  // length of the next window: halved after an overrun, doubled after a clean one
  float last_ms = 0.0f;
  std::string err;
  // optional per-kernel timing: event pairs around launches, accumulated after each sync
  uint32_t timing = 0;  // bit c set: time launches of kernel class c
  struct Pend { int cls; hipEvent_t a, b; };
  std::vector<Pend> pend;
  std::vector<hipEvent_t> pool;
  double k_ms[BX_NCLASS] = {};
  uint64_t k_n[BX_NCLASS] = {};
  // the last launch of the density passes A (after the first) and B with its arguments, for
  // mg_box_replay_ms: each writes only outputs that no later kernel of the step overwrites, from
  // inputs it does not write, so launching it again rewrites the same values
  std::function<void()> replay[BX_NCLASS];
};

namespace {

template <class R> const BoxGeo<R>& geo(const mg_box* B);
template <> const BoxGeo<float>& geo<float>(const mg_box* B) { return B->gf; }
template <> const BoxGeo<double>& geo<double>(const mg_box* B) { return B->gd; }
template <class R> Vec<R>* vp(void* p) { return static_cast<Vec<R>*>(p); }
template <class R> R* sp(void* p) { return static_cast<R*>(p); }

void box_set_err(char* err, size_t errlen, const std::string& m) {
  if (err && errlen) {
    size_t n = std::min(errlen - 1, m.size());
    std::memcpy(err, m.data(), n);
    err[n] = 0;
  }
}

// Geometry of the workload (same f32 expressions as the seeding description in the header).
bool box_geometry(const mg_box_config& c, BoxGeo<float>& g, std::string& err) {
  if (c.side < 8 || c.side > 8192) { err = "side must be in [8, 8192]"; return false; }
  if (!(c.dx > 0.0f) || !(c.rest > 0.0f) || !(c.dt > 0.0f) || !(c.mu >= 0.0f) || !(c.tol > 0.0f)) {
    err = "dx, rest, dt and tol must be positive and mu non-negative";
    return false;
  }
  if (!(c.dx <= 1024.0f)) { err = "dx must be at most 1024 (box side at most 2^23)"; return false; }
  if (c.fp64 > 1u) { err = "fp64 must be 0 or 1"; return false; }
  if (!(c.yield_stress >= 0.0f) || !std::isfinite(c.yield_stress)) {
    err = "yield_stress must be finite and non-negative";
    return false;
  }
  if (c.adaptive > 1u || c.energy_stats > 1u) { err = "adaptive and energy_stats must be 0 or 1"; return false; }
  if (c.adaptive && std::max(c.passes, 1u) + 6u > BX_MAXPASS) {
    err = "adaptive density passes: passes must be at most 10";
    return false;
  }
  if (!c.adaptive && c.passes > 64u) { err = "passes must be at most 64"; return false; }
  if (!std::isfinite(c.gravity) || !(c.speed_clamp >= 0.0f) || !std::isfinite(c.speed_clamp)) {
    err = "gravity must be finite and speed_clamp finite and non-negative";
    return false;
  }
  g.n = c.side * c.side;
  g.h = 2.0f * c.dx;
  g.h2 = g.h * g.h;
  g.L = (float)c.side * c.dx;
  g.halfL = 0.5f * g.L;
  uint32_t n = (uint32_t)std::floor((double)g.L / (double)g.h);
  while (n > 1 && g.L / (float)n < g.h) n--;
  g.ncx = std::max(n, 1u);
  if (g.ncx < 3) { err = "box too small: fewer than 3 cells per side"; return false; }
  g.cw = g.L / (float)g.ncx;
  g.k6 = 4.0f / (mg::PI_F * mg::powi8(g.h));
  g.ks = -30.0f / (mg::PI_F * mg::powi5(g.h));
  float lw = 0.0f;
  for (int q = -2; q <= 2; q++)
    for (int p = -2; p <= 2; p++) {
      float rx = (float)p * c.dx, ry = (float)q * c.dx;
      float r2 = rx * rx + ry * ry;
      if (r2 >= g.h * g.h || r2 < 0.0f) continue;
      float t = g.h * g.h - r2;
      lw += g.k6 * t * t * t;
    }
  if (!(c.density_ratio >= 0.0f) || !std::isfinite(c.density_ratio)) {
    err = "density_ratio must be finite and non-negative";
    return false;
  }
  g.mass = (c.density_ratio > 0.0f ? c.density_ratio : 1.02f) * c.rest / lw;
  g.rest = c.rest;
  g.dt = c.dt;
  {
    float t = g.h * g.h - 0.0f;
    g.w0 = g.mass * (g.k6 * t * t * t);
  }
  g.m_over_rest = g.mass / c.rest;
  g.mu2m = g.mass * (c.mu + c.mu);
  g.eta2 = 0.01f * g.h * g.h;
  g.lo = 2.0f * g.h;
  g.hi = g.L - g.lo;
  if (!(g.h >= 0x1p-60f)) g.hi = -1.0f;  // no interior: every wave takes the minimum image
  g.halt = nullptr;
  g.gdt = c.gravity * c.dt;
  g.adaptive = c.adaptive;
  g.base = std::max(c.passes, 1u);
  g.dens = nullptr;
  g.nblk = (g.n + BX_BLOCK - 1) / BX_BLOCK;
  g.xcd_chunk = (g.nblk % 8u == 0u) ? g.nblk / 8u : 0u;
  g.nred = (g.n + BX_RBLOCK - 1) / BX_RBLOCK;  // FP32 reduction kernels: one particle per lane
  g.ppt = 1;
  g.xcd_chunk_red = (g.nred % 8u == 0u) ? g.nred / 8u : 0u;
  return true;
}

// The FP64 constants: the f32 geometry promoted, the derived ones formed in f64 (pbf.rs:119-137
// and :312-344 evaluated in f64 over the promoted inputs).
BoxGeo<double> promote(const BoxGeo<float>& f, const mg_box_config& c) {
  BoxGeo<double> g;
  g.n = f.n; g.ncx = f.ncx; g.nblk = f.nblk; g.xcd_chunk = f.xcd_chunk; g.nred = f.nred;
  // FP64 reduction kernels: at most BX_NRED workgroups, each lane taking ppt particles, so the
  // partials every CG consumer block sums stay at most BX_NRED and the workgroups fit one round
  g.nred = std::min<uint32_t>(f.nred, BX_NRED);
  g.ppt = (g.n + g.nred * BX_RBLOCK - 1) / (g.nred * BX_RBLOCK);
  g.xcd_chunk_red = (g.nred % 8u == 0u) ? g.nred / 8u : 0u;
  const double pi = 3.14159265358979323846;
  g.h = (double)f.h;
  g.h2 = g.h * g.h;
  g.L = (double)f.L;
  g.halfL = (double)f.halfL;
  g.cw = (double)f.cw;
  double a = g.h * g.h;  // powi(h, 8) and powi(h, 5) by binary exponentiation, low bit first
  a = a * a;
  g.k6 = 4.0 / (pi * (a * a));
  g.ks = -30.0 / (pi * (g.h * a));
  g.mass = (double)f.mass;
  g.rest = (double)c.rest;
  g.dt = (double)c.dt;
  {
    double t = g.h * g.h - 0.0;
    g.w0 = g.mass * (g.k6 * t * t * t);
  }
  g.m_over_rest = g.mass / g.rest;
  const double mu = (double)c.mu;
  g.mu2m = g.mass * (mu + mu);
  g.eta2 = (double)0.01f * g.h * g.h;
  g.lo = 2.0 * g.h;
  g.hi = g.L - g.lo;
  if (!(g.h >= 0x1p-60)) g.hi = -1.0;
  g.halt = nullptr;
  g.gdt = (double)c.gravity * (double)c.dt;
  g.adaptive = f.adaptive;
  g.base = f.base;
  g.dens = nullptr;
  return g;
}

float wrap_host(float a, float L) {
  float r = std::fmod(a, L);
  return r < 0.0f ? r + std::fabs(L) : r;
}

// Lattice + jitter + Box-Muller velocities from one Rng(seed) stream in particle order.
void box_seed(const mg_box_config& c, const BoxGeo<float>& g, std::vector<float>& x, std::vector<float>& v) {
  const size_t n = (size_t)c.side * c.side;
  x.assign(2 * n, 0.0f);
  v.assign(2 * n, 0.0f);
  mg::Rng rng(c.seed);
  const float jit = 0.1f * c.dx;
  const double two_pi = 6.283185307179586476925;
  for (uint32_t iy = 0; iy < c.side; iy++)
    for (uint32_t ix = 0; ix < c.side; ix++) {
      size_t i = (size_t)iy * c.side + ix;
      float jx = (rng.next_f32() * 2.0f - 1.0f) * jit;
      float jy = (rng.next_f32() * 2.0f - 1.0f) * jit;
      double u1 = rng.next_f32(), u2 = rng.next_f32(), u3 = rng.next_f32(), u4 = rng.next_f32();
      double g1 = std::sqrt(-2.0 * std::log(std::max(u1, 1e-12))) * std::cos(two_pi * u2);
      double g2 = std::sqrt(-2.0 * std::log(std::max(u3, 1e-12))) * std::cos(two_pi * u4);
      x[2 * i] = wrap_host(((float)ix + 0.5f) * c.dx + jx, g.L);
      x[2 * i + 1] = wrap_host(((float)iy + 0.5f) * c.dx + jy, g.L);
      v[2 * i] = (float)(0.01 * g1);
      v[2 * i + 1] = (float)(0.01 * g2);
    }
}

template <class T> size_t carve(size_t off, T*& p, size_t count) {
  p = reinterpret_cast<T*>(off);
  return off + ((count * sizeof(T) + 255) / 256) * 256 + 256;
}
size_t carve_bytes(size_t off, void*& p, size_t bytes) {
  p = reinterpret_cast<void*>(off);
  return off + ((bytes + 255) / 256) * 256 + 256;
}

inline dim3 blocks_for(uint32_t n) { return dim3((n + BX_BLOCK - 1) / BX_BLOCK); }

hipEvent_t take_event(mg_box* B) {
  if (!B->pool.empty()) {
    hipEvent_t e = B->pool.back();
    B->pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}
// Brackets the launches between tic and toc with an event pair of class cls (when timing is on).
struct Tic {
  mg_box* B;
  int cls;
  hipEvent_t a = nullptr;
  Tic(mg_box* b, int c) : B(b), cls(c) {
    if (((B->timing >> cls) & 1u) && (a = take_event(B))) (void)hipEventRecord(a, B->stream);
  }
  ~Tic() {
    if (!a) return;
    hipEvent_t b = take_event(B);
    if (!b) { B->pool.push_back(a); return; }
    (void)hipEventRecord(b, B->stream);
    B->pend.push_back({cls, a, b});
  }
};
// After a stream sync: fold the completed event pairs into the per-class totals.
void collect_times(mg_box* B) {
  for (auto& p : B->pend) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      B->k_ms[p.cls] += ms;
      B->k_n[p.cls] += 1;
    }
    B->pool.push_back(p.a);
    B->pool.push_back(p.b);
  }
  B->pend.clear();
}

template <class R> PairView<R> pair_view(const mg_box* B) {
  return PairView<R>{vp<R>(B->XP), B->start, B->pnb, B->pcnt, B->n};
}

// Enqueues one step up to and including the CG set-up; returns false on a launch error. When it
// returns, X0, V and id hold the step's start state (positions, velocities, ids) in the step's cell
// order: nothing later in the step writes them (the new velocity goes to VN).
template <class R> bool enqueue_front(mg_box* B) {
  const BoxGeo<R>& g = geo<R>(B);
  const uint32_t cg_off = (B->cfg.mu > 0.0f && B->cfg.max_cg > 0) ? 0u : 1u;
  hipStream_t st = B->stream;
  const dim3 blk(BX_BLOCK), gp = blocks_for(g.n);
  std::swap(B->start, B->start_next);  // counts start from the buffer the previous step cleared
  {
    Tic t(B, BXT_GRID);
    hipLaunchKernelGGL(bx_predict<R>, gp, blk, 0, st, g, vp<R>(B->X), vp<R>(B->V), vp<R>(B->X0), B->key, B->start);
    hipLaunchKernelGGL(bx_scan_totals, dim3(B->nscan), blk, 0, st, B->halt, B->start, B->ncell + 1, B->bsum);
    hipLaunchKernelGGL(bx_scan_apply, dim3(B->nscan), blk, 0, st, B->halt, B->start, B->ncell + 1, B->bsum);
    hipLaunchKernelGGL(bx_scatter, gp, blk, 0, st, B->halt, g.n, B->key, B->start, B->cursor, B->slot);
    hipLaunchKernelGGL(bx_cellsort, blocks_for(B->ncell), blk, 0, st, B->halt, B->ncell, B->start, B->slot, B->id,
                       B->sc, cg_off, B->cfg.max_cg, B->cfg.tol);
    hipLaunchKernelGGL(bx_gather<R>, gp, blk, 0, st, g, B->slot, vp<R>(B->X), vp<R>(B->X0), vp<R>(B->V), B->id,
                       vp<R>(B->Xb), vp<R>(B->XP), vp<R>(B->X0b), vp<R>(B->Vb), B->idb, B->start_next, B->cursor,
                       B->ncell);
  }
  std::swap(B->X, B->Xb);
  std::swap(B->X0, B->X0b);
  std::swap(B->V, B->Vb);
  std::swap(B->id, B->idb);
  const PairView<R> pv = pair_view<R>(B);
  if (B->cfg.passes == 0) {
    Tic t(B, BXT_GRID);
    hipLaunchKernelGGL(bx_pairs<R>, gp, blk, 0, st, g, vp<R>(B->XP), B->start, B->pnb, B->pcnt);
  }
  const uint32_t npass = B->cfg.adaptive ? std::max(B->cfg.passes, 1u) + 6u : B->cfg.passes;
  for (uint32_t p = 0; p < npass; p++) {
    if (p == 0) {  // the first pass A also builds the pair table
      Tic t(B, BXT_PAIRLAM);
      hipLaunchKernelGGL(bx_pairs_lambda<R>, gp, blk, 0, st, g, vp<R>(B->XP), B->start, B->pnb, B->pcnt,
                         B->LAM);
    } else {
      Tic t(B, BXT_LAMBDA);
      const BoxGeo<R> gc = g;
      Vec<R>* const x = vp<R>(B->X);
      void* const lam = B->LAM;
      B->replay[BXT_LAMBDA] = [=]() { hipLaunchKernelGGL(bx_lambda<R>, gp, blk, 0, st, gc, pv, x, lam, p); };
      B->replay[BXT_LAMBDA]();
    }
    {
      Tic t(B, BXT_DP);
      const BoxGeo<R> gc = g;
      const Vec<R>* const x = vp<R>(B->X);
      const void* const lam = B->LAM;
      Vec<R>* const xo = vp<R>(B->X2);
      B->replay[BXT_DP] = [=]() { hipLaunchKernelGGL(bx_dp<R>, gp, blk, 0, st, gc, pv, x, lam, xo, p); };
      B->replay[BXT_DP]();
    }
    std::swap(B->X, B->X2);
  }
  const dim3 rblk(BX_RBLOCK), gr(g.nred);
  {
    Tic t(B, BXT_VEL);
    hipLaunchKernelGGL(bx_vel_density<R>, gr, rblk, 0, st, g, pv, vp<R>(B->X), vp<R>(B->X0), vp<R>(B->VN),
                       sp<R>(B->DENS), B->part_b, B->part_ke);
  }
  const bool bingham = B->cfg.yield_stress > 0.0f;
  if (!cg_off && bingham) {
    Tic t(B, BXT_VEL);
    hipLaunchKernelGGL(bx_shear_mu<R>, gp, blk, 0, st, g, pv, vp<R>(B->X), vp<R>(B->VN), sp<R>(B->DENS),
                       sp<R>(B->MU), (R)B->cfg.mu, (R)B->cfg.yield_stress);
  }
  if (!cg_off) {
    Tic t(B, BXT_CGINIT);
    if (bingham)
      hipLaunchKernelGGL(bx_cg_init_bingham<R>, gr, rblk, 0, st, g, pv, vp<R>(B->X), sp<R>(B->DENS), vp<R>(B->VN),
                         vp<R>(B->R), vp<R>(B->P), sp<R>(B->cw), B->part_rr, B->sc, sp<R>(B->MU));
    else
      hipLaunchKernelGGL(bx_cg_init<R>, gr, rblk, 0, st, g, pv, vp<R>(B->X), sp<R>(B->DENS), vp<R>(B->VN),
                         vp<R>(B->R), vp<R>(B->P), sp<R>(B->cw), B->part_rr, B->sc);
  }
  B->cg_t = 1;
  B->have_neighbors = true;
  return hipGetLastError() == hipSuccess;
}

// Enqueues CG iterations cg_t .. cg_t + iters - 1 and the end-of-chunk convergence check (which
// also records the window verdict of step `s` when `verdict` is given).
template <class R> void enqueue_cg_chunk(mg_box* B, uint32_t iters, uint32_t* verdict = nullptr, uint32_t s = 0) {
  const BoxGeo<R>& g = geo<R>(B);
  hipStream_t st = B->stream;
  const PairView<R> pv = pair_view<R>(B);
  const dim3 rblk(BX_RBLOCK), gr(g.nred);
  const size_t nr = g.nred;
  for (uint32_t k = 0; k < iters; k++) {
    const uint32_t t = B->cg_t++;
    // iteration t reads the direction of t-1 and writes the other buffer; r.r of iteration u lives
    // in generation u & 1 (the initial residual is generation 0)
    Vec<R>* pold = vp<R>((t & 1u) ? B->P : B->P2);
    Vec<R>* pnew = vp<R>((t & 1u) ? B->P2 : B->P);
    const double* rr_prev = B->part_rr + ((t - 1u) & 1u) * nr;
    double* rr_out = B->part_rr + (t & 1u) * nr;
    {
      Tic tt(B, BXT_SPMV);
      hipLaunchKernelGGL(bx_cg_dir<R>, gr, rblk, 0, st, g, pv, vp<R>(B->X), sp<R>(B->DENS), sp<R>(B->cw),
                         vp<R>(B->R), pold, pnew, vp<R>(B->AP), B->part_b, rr_prev, B->part_pap, B->sc, t);
    }
    {
      Tic tt(B, BXT_UPDATE);
      hipLaunchKernelGGL(bx_cg_upd<R>, gr, rblk, 0, st, g, vp<R>(B->VN), vp<R>(B->R), pnew, vp<R>(B->AP),
                         B->part_pap, rr_out, B->sc, t);
    }
  }
  const uint32_t last = B->cg_t - 1u;
  hipLaunchKernelGGL(bx_cg_check, dim3(1), rblk, 0, st, g.nred, B->part_rr + (last & 1u) * nr, B->sc, last, verdict,
                     s, B->halt);
}

// After a step's CG: the fluid speed clamp and the kinetic-energy sums, when the box has them.
template <class R> void enqueue_epilogue(mg_box* B) {
  if (!(B->cfg.speed_clamp > 0.0f) && !B->cfg.energy_stats) return;
  const BoxGeo<R>& g = geo<R>(B);
  const size_t nr = g.nred;
  double* pa = B->part_ke ? B->part_ke + nr : nullptr;
  double* pc = B->part_ke ? B->part_ke + 2 * nr : nullptr;
  hipLaunchKernelGGL(bx_clamp_ke<R>, dim3(g.nred), dim3(BX_RBLOCK), 0, B->stream, g, vp<R>(B->VN),
                     (R)B->cfg.speed_clamp, pa, pc);
}

// One full step: the front half and a first CG chunk of `first` iterations are enqueued without a
// host wait; further chunks follow only if the device has not converged.
template <class R> int run_step(mg_box* B, uint32_t first) {
  if (!enqueue_front<R>(B)) { B->err = "kernel launch failed"; return MG_E_DEVICE; }
  uint32_t launched = 0;
  if (B->cfg.mu > 0.0f && B->cfg.max_cg > 0) {
    launched = std::min<uint32_t>(std::max<uint32_t>(first, 1u), B->cfg.max_cg);
    enqueue_cg_chunk<R>(B, launched);
  } else {
    launched = B->cfg.max_cg;  // no CG this step
  }
  for (;;) {
    (void)hipMemcpyAsync(B->h_sc, B->sc, sizeof(BoxScal), hipMemcpyDeviceToHost, B->stream);
    if (hipStreamSynchronize(B->stream) != hipSuccess) { B->err = "HIP stream error"; return MG_E_DEVICE; }
    collect_times(B);
    if (B->h_sc->done || launched >= B->cfg.max_cg) break;
    uint32_t k = std::min<uint32_t>(BX_CG_CHUNK, B->cfg.max_cg - launched);
    enqueue_cg_chunk<R>(B, k);
    launched += k;
    if (hipGetLastError() != hipSuccess) { B->err = "kernel launch failed"; return MG_E_DEVICE; }
  }
  enqueue_epilogue<R>(B);
  std::swap(B->V, B->VN);
  B->last_it = B->h_sc->it;
  B->total_it += B->h_sc->it;
  return MG_OK;
}

// Makes *names[0] point at `keep`, which one of the other names holds; they swap.
template <class T, size_t N>
void take_buffer(T* keep, T** const (&names)[N]) {
  for (size_t k = 1; k < N; k++)
    if (*names[k] == keep) {
      std::swap(*names[0], *names[k]);
      return;
    }
}

// Runs `steps` steps with one host wait per window of up to BX_WINDOW steps. Each step's CG is
// launched with the largest iteration count of the previous window plus a margin while the counts
// are changing (launches after convergence return at once). If step b of a window had not converged
// when its iterations ran out, it sets the stop word: the rest of the window does no work. Step b's
// start state is still in the buffers that held it during step b (its X0, V and id, in its cell
// order: a step's result does not depend on the order of its input, which the cell sort removes),
// so those buffers become the live state again and step b is re-run with host-driven chunks
// (run_step); the next window is then half as long. The results are those of running every step
// with run_step: iterations past convergence change nothing.
template <class R> int run_window(mg_box* B, uint32_t steps) {
  const bool cg_on = B->cfg.mu > 0.0f && B->cfg.max_cg > 0;
  uint32_t done = 0;
  while (done < steps) {
    const uint32_t W = std::min<uint32_t>(steps - done, cg_on ? B->win : BX_WINDOW);
    const uint32_t m = std::max<uint32_t>(B->last_it, 1u);
    const uint32_t spare = B->cg_stable ? 0u : BX_SPARE_IT + m / 8u;
    const uint32_t pred = std::min<uint32_t>(m + spare, B->cfg.max_cg);
    if (cg_on && W == 1) {  // a single step: chunks driven from the host
      const int rc = run_step<R>(B, pred);
      if (rc) return rc;
      B->win = B->h_sc->it <= pred ? 2u : 1u;
      B->cg_stable = false;
      done += 1;
      continue;
    }
    if (cg_on) hipLaunchKernelGGL(bx_verdict_reset, dim3(1), dim3(64), 0, B->stream, B->verdict, B->halt);
    for (uint32_t s = 0; s < W; s++) {
      if (!enqueue_front<R>(B)) { B->err = "kernel launch failed"; return MG_E_DEVICE; }
      B->starts[s] = {B->X0, B->V, B->id, B->start};
      if (cg_on) enqueue_cg_chunk<R>(B, pred, B->verdict, s);
      enqueue_epilogue<R>(B);
      std::swap(B->V, B->VN);
    }
    if (cg_on) (void)hipMemcpyAsync(B->h_verdict, B->verdict, (W + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost, B->stream);
    if (hipStreamSynchronize(B->stream) != hipSuccess) { B->err = "HIP stream error"; return MG_E_DEVICE; }
    if (hipGetLastError() != hipSuccess) { B->err = "kernel launch failed"; return MG_E_DEVICE; }
    collect_times(B);
    if (!cg_on) {
      done += W;
      continue;
    }
    const uint32_t bad = B->h_verdict[0];
    const uint32_t good = bad < W ? bad : W;
    bool same = true;
    uint32_t most = 0;
    for (uint32_t s = 0; s < good; s++) {
      B->total_it += B->h_verdict[1 + s];
      same = same && B->h_verdict[1 + s] == B->h_verdict[1];
      most = std::max<uint32_t>(most, B->h_verdict[1 + s]);
    }
    if (good > 0) B->last_it = most;
    B->cg_stable = bad >= W && same;
    done += good;
    if (bad >= W) {
      B->win = std::min<uint32_t>(2u * B->win, BX_WINDOW);
      continue;
    }
    B->win = std::max<uint32_t>(B->win / 2u, 1u);
    const mg_box::Start& st = B->starts[bad];
    void** const pos[] = {&B->X, &B->Xb, &B->X0, &B->X0b, &B->X2};
    void** const vel[] = {&B->V, &B->Vb, &B->VN};
    uint32_t** const ids[] = {&B->id, &B->idb};
    take_buffer(st.x, pos);
    take_buffer(st.v, vel);
    take_buffer(st.id, ids);
    // step b's counts were accumulated into st.cells and its gather cleared the other buffer: the
    // re-run (which swaps first) counts into the cleared one and clears st.cells for its successor
    uint32_t** const cells[] = {&B->start, &B->start_next};
    take_buffer(st.cells, cells);
    hipLaunchKernelGGL(bx_verdict_reset, dim3(1), dim3(64), 0, B->stream, B->verdict, B->halt);
    const int rc = run_step<R>(B, pred + BX_CG_CHUNK);
    if (rc) return rc;
    done += 1;
  }
  return MG_OK;
}

// Host copies of the live state in the original particle order, converted to T.
template <class R, class T> int copy_state_as(mg_box* B, T* x2n, T* v2n) {
  const uint32_t n = B->n;
  std::vector<Vec<R>> x(n), v(n);
  std::vector<uint32_t> id(n);
  bool ok = hipStreamSynchronize(B->stream) == hipSuccess;
  ok = ok && hipMemcpy(x.data(), B->X, (size_t)n * sizeof(Vec<R>), hipMemcpyDeviceToHost) == hipSuccess;
  ok = ok && hipMemcpy(v.data(), B->V, (size_t)n * sizeof(Vec<R>), hipMemcpyDeviceToHost) == hipSuccess;
  ok = ok && hipMemcpy(id.data(), B->id, (size_t)n * 4, hipMemcpyDeviceToHost) == hipSuccess;
  if (!ok) return MG_E_DEVICE;
  for (uint32_t q = 0; q < n; q++) {
    uint32_t i = id[q];
    if (i >= n) return MG_E_DEVICE;
    if (x2n) { x2n[2 * i] = (T)x[q].x; x2n[2 * i + 1] = (T)x[q].y; }
    if (v2n) { v2n[2 * i] = (T)v[q].x; v2n[2 * i + 1] = (T)v[q].y; }
  }
  return MG_OK;
}
template <class R, class T> int set_state_from(mg_box* B, const T* x2n, const T* v2n) {
  const uint32_t n = B->n;
  for (auto& f : B->replay) f = nullptr;  // their inputs no longer match the state
  std::vector<Vec<R>> x(n), v(n);
  std::vector<uint32_t> ids(n);
  for (uint32_t i = 0; i < n; i++) {
    ids[i] = i;
    x[i].x = (R)x2n[2 * i]; x[i].y = (R)x2n[2 * i + 1];
    v[i].x = (R)v2n[2 * i]; v[i].y = (R)v2n[2 * i + 1];
  }
  bool ok = hipStreamSynchronize(B->stream) == hipSuccess;
  ok = ok && hipMemcpy(B->X, x.data(), (size_t)n * sizeof(Vec<R>), hipMemcpyHostToDevice) == hipSuccess;
  ok = ok && hipMemcpy(B->V, v.data(), (size_t)n * sizeof(Vec<R>), hipMemcpyHostToDevice) == hipSuccess;
  ok = ok && hipMemcpy(B->id, ids.data(), (size_t)n * 4, hipMemcpyHostToDevice) == hipSuccess;
  B->have_neighbors = false;
  return ok ? MG_OK : MG_E_DEVICE;
}
template <class R, class T> int density_as(mg_box* B, T* out) {
  const uint32_t n = B->n;
  std::vector<R> d(n);
  std::vector<uint32_t> id(n);
  bool ok = hipStreamSynchronize(B->stream) == hipSuccess;
  ok = ok && hipMemcpy(d.data(), B->DENS, (size_t)n * sizeof(R), hipMemcpyDeviceToHost) == hipSuccess;
  ok = ok && hipMemcpy(id.data(), B->id, (size_t)n * 4, hipMemcpyDeviceToHost) == hipSuccess;
  if (!ok) return MG_E_DEVICE;
  for (uint32_t q = 0; q < n; q++) {
    if (id[q] >= n) return MG_E_DEVICE;
    out[id[q]] = (T)d[q];
  }
  return MG_OK;
}
template <class R> void launch_nbr_count(mg_box* B, uint32_t* d_cnt) {
  hipLaunchKernelGGL(bx_nbr_count<R>, blocks_for(B->n), dim3(BX_BLOCK), 0, B->stream, geo<R>(B), pair_view<R>(B),
                     B->id, d_cnt);
}
template <class R> void launch_nbr_write(mg_box* B, const uint32_t* d_off, uint32_t* d_nb, uint32_t cap) {
  hipLaunchKernelGGL(bx_nbr_write<R>, blocks_for(B->n), dim3(BX_BLOCK), 0, B->stream, geo<R>(B), pair_view<R>(B),
                     B->id, d_off, d_nb, cap);
}

}  // namespace

extern "C" {

int mg_box_geometry(const mg_box_config* cfg, double* out5) {
  if (!cfg || !out5) return MG_E_ARG;
  BoxGeo<float> g;
  std::string e;
  if (!box_geometry(*cfg, g, e)) return MG_E_PARAMS;
  out5[0] = g.h; out5[1] = g.L; out5[2] = g.cw; out5[3] = g.mass; out5[4] = g.ncx;
  return MG_OK;
}

int mg_box_create(const mg_box_config* cfg, mg_box** out, char* err, size_t errlen) {
  if (!cfg || !out) return MG_E_ARG;
  *out = nullptr;
  auto* B = new mg_box();
  B->cfg = *cfg;
  std::string e;
  if (!box_geometry(*cfg, B->gf, e)) { box_set_err(err, errlen, e); delete B; return MG_E_PARAMS; }
  B->f64 = cfg->fp64 != 0u;
  B->gd = promote(B->gf, *cfg);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) { box_set_err(err, errlen, "no HIP device"); delete B; return MG_E_DEVICE; }
  const uint32_t n = B->gf.n;
  B->n = n;
  B->nred = B->gf.nred;
  B->ncell = B->gf.ncx * B->gf.ncx;
  B->nscan = (B->ncell + 1 + BX_SCAN_CHUNK - 1) / BX_SCAN_CHUNK;
  if (B->nscan > (1u << 16)) { box_set_err(err, errlen, "too many cells"); delete B; return MG_E_PARAMS; }
  const size_t rs = B->f64 ? 8 : 4;   // bytes of one real
  const size_t vb = 2 * rs * n;       // bytes of one Vec<R> array
  size_t off = 0;
  off = carve_bytes(off, B->X, vb); off = carve_bytes(off, B->XP, vb); off = carve_bytes(off, B->X0, vb);
  off = carve_bytes(off, B->V, vb); off = carve_bytes(off, B->X2, vb); off = carve_bytes(off, B->Xb, vb);
  off = carve_bytes(off, B->X0b, vb); off = carve_bytes(off, B->Vb, vb); off = carve_bytes(off, B->VN, vb);
  off = carve(off, B->id, n); off = carve(off, B->idb, n); off = carve(off, B->key, n); off = carve(off, B->slot, n);
  off = carve(off, B->start, (size_t)B->ncell + 1); off = carve(off, B->start_next, (size_t)B->ncell + 1);
  off = carve(off, B->cursor, (size_t)B->ncell);
  off = carve(off, B->bsum, (size_t)B->nscan);
  off = carve_bytes(off, B->LAM, 4 * rs * n); off = carve_bytes(off, B->DENS, rs * n);
  off = carve_bytes(off, B->MU, cfg->yield_stress > 0.0f ? rs * n : 0);
  off = carve_bytes(off, B->R, vb); off = carve_bytes(off, B->P, vb); off = carve_bytes(off, B->AP, vb);
  off = carve(off, B->part_b, (size_t)B->nred); off = carve(off, B->part_rr, 2 * (size_t)B->nred);
  off = carve(off, B->part_pap, (size_t)B->nred); off = carve_bytes(off, B->P2, vb);
  off = carve(off, B->part_ke, cfg->energy_stats ? 3 * (size_t)B->nred : 0);
  off = carve(off, B->pnb, (size_t)BX_KMAX * n); off = carve(off, B->pcnt, n);
  off = carve_bytes(off, B->cw, (size_t)BX_KMAX * rs * n);
  off = carve(off, B->verdict, BX_WINDOW + 1);
  off = carve(off, B->halt, 1); off = carve(off, B->sc, 1);
  if (hipMalloc(&B->arena, off) != hipSuccess) { box_set_err(err, errlen, "hipMalloc failed"); delete B; return MG_E_DEVICE; }
  const uintptr_t base = (uintptr_t)B->arena;
  auto fix = [&](auto*& p) { p = reinterpret_cast<std::remove_reference_t<decltype(p)>>(base + (uintptr_t)p); };
  fix(B->X); fix(B->XP); fix(B->X0); fix(B->V); fix(B->X2); fix(B->Xb); fix(B->X0b); fix(B->Vb); fix(B->VN);
  fix(B->id); fix(B->idb); fix(B->key); fix(B->slot); fix(B->start); fix(B->start_next); fix(B->cursor); fix(B->bsum);
  fix(B->LAM); fix(B->DENS); fix(B->MU); fix(B->R); fix(B->P); fix(B->AP); fix(B->part_b); fix(B->part_rr); fix(B->part_pap);
  fix(B->P2); fix(B->sc); fix(B->pnb); fix(B->pcnt); fix(B->cw); fix(B->verdict); fix(B->halt);
  if (cfg->energy_stats) fix(B->part_ke); else B->part_ke = nullptr;
  B->gf.halt = B->halt;
  B->gd.halt = B->halt;
  B->gf.dens = B->gd.dens = B->sc->dens;
  bool ok = hipMemset(B->arena, 0, off) == hipSuccess;
  if (cfg->yield_stress > 0.0f) {  // the CG's rescan path finds the effective viscosities here
    const void* mu_eff = B->MU;
    ok = ok && hipMemcpy(reinterpret_cast<char*>(B->sc) + offsetof(BoxScal, mu_eff), &mu_eff, sizeof mu_eff,
                         hipMemcpyHostToDevice) == hipSuccess;
  }
  ok = ok && hipHostMalloc((void**)&B->h_sc, sizeof(BoxScal), 0) == hipSuccess;
  ok = ok && hipHostMalloc((void**)&B->h_verdict, (BX_WINDOW + 1) * sizeof(uint32_t), 0) == hipSuccess;
  ok = ok && hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking) == hipSuccess;
  ok = ok && hipEventCreate(&B->ev0) == hipSuccess && hipEventCreate(&B->ev1) == hipSuccess;
  if (ok) {
    std::vector<float> x, v;
    box_seed(*cfg, B->gf, x, v);
    ok = (B->f64 ? set_state_from<double, float>(B, x.data(), v.data())
                 : set_state_from<float, float>(B, x.data(), v.data())) == MG_OK;
  }
  if (!ok) { box_set_err(err, errlen, "HIP error while creating the box"); mg_box_destroy(B); return MG_E_DEVICE; }
  *out = B;
  return MG_OK;
}

void mg_box_destroy(mg_box* B) {
  if (!B) return;
  if (B->stream) (void)hipStreamSynchronize(B->stream);
  if (B->arena) (void)hipFree(B->arena);
  if (B->h_sc) (void)hipHostFree(B->h_sc);
  if (B->h_verdict) (void)hipHostFree(B->h_verdict);
  if (B->ev0) (void)hipEventDestroy(B->ev0);
  if (B->ev1) (void)hipEventDestroy(B->ev1);
  collect_times(B);
  for (hipEvent_t e : B->pool) (void)hipEventDestroy(e);
  if (B->stream) (void)hipStreamDestroy(B->stream);
  delete B;
}

size_t mg_box_count(const mg_box* B) { return B ? (size_t)B->n : 0; }

int mg_box_step(mg_box* B, uint32_t steps) {
  if (!B) return MG_E_ARG;
  (void)hipEventRecord(B->ev0, B->stream);
  const int rc = B->f64 ? run_window<double>(B, steps) : run_window<float>(B, steps);
  if (rc) return rc;
  (void)hipEventRecord(B->ev1, B->stream);
  if (hipEventSynchronize(B->ev1) != hipSuccess) return MG_E_DEVICE;
  (void)hipEventElapsedTime(&B->last_ms, B->ev0, B->ev1);
  return MG_OK;
}

float mg_box_last_step_ms(const mg_box* B) { return B ? B->last_ms : 0.0f; }

int mg_box_set_timing(mg_box* B, uint32_t class_mask) {
  if (!B) return MG_E_ARG;
  B->timing = class_mask;
  for (int k = 0; k < BX_NCLASS; k++) { B->k_ms[k] = 0.0; B->k_n[k] = 0; }
  return MG_OK;
}

int mg_box_replay_ms(mg_box* B, uint32_t cls, uint32_t reps, double* ms_per_launch) {
  if (!B || !ms_per_launch || reps == 0 || cls >= BX_NCLASS) return MG_E_ARG;
  if ((cls != BXT_LAMBDA && cls != BXT_DP) || !B->replay[cls]) return MG_E_ARG;
  hipEvent_t a = take_event(B), b = take_event(B);
  if (!a || !b) return MG_E_DEVICE;
  bool ok = hipEventRecord(a, B->stream) == hipSuccess;
  for (uint32_t r = 0; ok && r < reps; r++) B->replay[cls]();
  ok = ok && hipGetLastError() == hipSuccess && hipEventRecord(b, B->stream) == hipSuccess &&
       hipEventSynchronize(b) == hipSuccess;
  float ms = 0.0f;
  ok = ok && hipEventElapsedTime(&ms, a, b) == hipSuccess;
  B->pool.push_back(a);
  B->pool.push_back(b);
  if (!ok) return MG_E_DEVICE;
  *ms_per_launch = (double)ms / reps;
  return MG_OK;
}

int mg_box_kernel_times(mg_box* B, double* ms, uint64_t* launches, int nclass) {
  if (!B || !ms || !launches) return MG_E_ARG;
  for (int k = 0; k < nclass && k < BX_NCLASS; k++) { ms[k] = B->k_ms[k]; launches[k] = B->k_n[k]; }
  return MG_OK;
}

int mg_box_copy_state(mg_box* B, float* x2n, float* v2n) {
  if (!B) return MG_E_ARG;
  return B->f64 ? copy_state_as<double, float>(B, x2n, v2n) : copy_state_as<float, float>(B, x2n, v2n);
}
int mg_box_copy_state_f64(mg_box* B, double* x2n, double* v2n) {
  if (!B) return MG_E_ARG;
  return B->f64 ? copy_state_as<double, double>(B, x2n, v2n) : copy_state_as<float, double>(B, x2n, v2n);
}

int mg_box_set_state(mg_box* B, const float* x2n, const float* v2n) {
  if (!B || !x2n || !v2n) return MG_E_ARG;
  return B->f64 ? set_state_from<double, float>(B, x2n, v2n) : set_state_from<float, float>(B, x2n, v2n);
}
int mg_box_set_state_f64(mg_box* B, const double* x2n, const double* v2n) {
  if (!B || !x2n || !v2n) return MG_E_ARG;
  return B->f64 ? set_state_from<double, double>(B, x2n, v2n) : set_state_from<float, double>(B, x2n, v2n);
}

int mg_box_density(mg_box* B, float* out) {
  if (!B || !out) return MG_E_ARG;
  return B->f64 ? density_as<double, float>(B, out) : density_as<float, float>(B, out);
}
int mg_box_density_f64(mg_box* B, double* out) {
  if (!B || !out) return MG_E_ARG;
  return B->f64 ? density_as<double, double>(B, out) : density_as<float, double>(B, out);
}

extern "C++" template <class R> int mean_shear_as(mg_box* B, double* out) {
  const BoxGeo<R>& g = geo<R>(B);
  const PairView<R> pv = pair_view<R>(B);
  // part_b and part_pap are free between steps (each step rewrites them before reading)
  hipLaunchKernelGGL(bx_shear_mean<R>, dim3(g.nred), dim3(BX_RBLOCK), 0, B->stream, g, pv, vp<R>(B->X), vp<R>(B->V),
                     sp<R>(B->DENS), B->part_b, B->part_pap);
  std::vector<double> ps(g.nred), pc(g.nred);
  bool ok = hipGetLastError() == hipSuccess && hipStreamSynchronize(B->stream) == hipSuccess;
  ok = ok && hipMemcpy(ps.data(), B->part_b, g.nred * sizeof(double), hipMemcpyDeviceToHost) == hipSuccess;
  ok = ok && hipMemcpy(pc.data(), B->part_pap, g.nred * sizeof(double), hipMemcpyDeviceToHost) == hipSuccess;
  if (!ok) return MG_E_DEVICE;
  double sum = 0.0, cnt = 0.0;
  for (uint32_t k = 0; k < g.nred; k++) { sum += ps[k]; cnt += pc[k]; }
  const double mean = cnt > 0.0 ? sum / cnt : 0.0;
  *out = B->f64 ? mean : (double)(float)mean;  // FP32 boxes report the f32 value (pbf.rs:880)
  return MG_OK;
}

int mg_box_mean_shear_rate(mg_box* B, double* out) {
  if (!B || !out) return MG_E_ARG;
  if (!B->have_neighbors) { *out = 0.0; return MG_OK; }
  if (hipStreamSynchronize(B->stream) != hipSuccess) return MG_E_DEVICE;
  return B->f64 ? mean_shear_as<double>(B, out) : mean_shear_as<float>(B, out);
}

int mg_box_stats(mg_box* B, double* out3) {
  if (!B || !out3) return MG_E_ARG;
  if (hipStreamSynchronize(B->stream) != hipSuccess) return MG_E_DEVICE;
  BoxScal s{};
  if (hipMemcpy(&s, B->sc, sizeof s, hipMemcpyDeviceToHost) != hipSuccess) return MG_E_DEVICE;
  out3[0] = s.it; out3[1] = s.res; out3[2] = s.bnorm;
  return MG_OK;
}

uint32_t mg_box_density_passes(mg_box* B) {
  if (!B || !B->have_neighbors) return 0;
  if (!B->cfg.adaptive) return B->cfg.passes;
  unsigned long long p = 0;
  if (hipStreamSynchronize(B->stream) != hipSuccess ||
      hipMemcpy(&p, &B->sc->dens[BX_MAXPASS], sizeof p, hipMemcpyDeviceToHost) != hipSuccess)
    return 0;
  return (uint32_t)p;
}

int mg_box_energy(mg_box* B, double* out3) {
  if (!B || !out3) return MG_E_ARG;
  if (!B->part_ke || !B->have_neighbors) return MG_E_ARG;
  const uint32_t nr = B->f64 ? B->gd.nred : B->gf.nred;
  std::vector<double> part(3 * (size_t)nr);
  if (hipStreamSynchronize(B->stream) != hipSuccess ||
      hipMemcpy(part.data(), B->part_ke, part.size() * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
    return MG_E_DEVICE;
  const double half_m = 0.5 * (B->f64 ? B->gd.mass : (double)B->gf.mass);
  for (int k = 0; k < 3; k++) {
    double s = 0.0;
    for (uint32_t b = 0; b < nr; b++) s += part[(size_t)k * nr + b];
    out3[k] = half_m * s;
  }
  return MG_OK;
}

uint64_t mg_box_total_cg_iterations(const mg_box* B) { return B ? B->total_it : 0; }

int mg_box_set_cg_prediction(mg_box* B, uint32_t iterations) {
  if (!B) return MG_E_ARG;
  B->last_it = iterations;
  B->cg_stable = true;
  return MG_OK;
}

int mg_box_check_arith(uint64_t samples, uint64_t seed, uint64_t* mismatches4) {
  if (!mismatches4 || samples == 0) return MG_E_ARG;
  const uint32_t per_lane = 64;
  const uint64_t lanes = (samples + per_lane - 1) / per_lane;
  if (lanes > (1ull << 31)) return MG_E_ARG;
  unsigned long long* d_bad = nullptr;
  if (hipMalloc(&d_bad, 4 * sizeof(unsigned long long)) != hipSuccess) return MG_E_DEVICE;
  bool ok = hipMemset(d_bad, 0, 4 * sizeof(unsigned long long)) == hipSuccess;
  if (ok) {
    hipLaunchKernelGGL(bx_check_arith, dim3((uint32_t)((lanes + BX_BLOCK - 1) / BX_BLOCK)), dim3(BX_BLOCK), 0, 0,
                       seed, per_lane, d_bad);
    unsigned long long h[4] = {0, 0, 0, 0};
    ok = hipGetLastError() == hipSuccess && hipDeviceSynchronize() == hipSuccess &&
         hipMemcpy(h, d_bad, sizeof h, hipMemcpyDeviceToHost) == hipSuccess;
    for (int k = 0; k < 4; k++) mismatches4[k] = h[k];
  }
  (void)hipFree(d_bad);
  return ok ? MG_OK : MG_E_DEVICE;
}

int64_t mg_box_neighbors(mg_box* B, uint32_t* off, uint32_t* nb, int64_t cap) {
  if (!B || !off) return MG_E_ARG;
  if (!B->have_neighbors) return MG_E_ARG;
  const uint32_t n = B->n;
  uint32_t* d_cnt = nullptr;
  uint32_t* d_nb = nullptr;
  const uint32_t dcap = (uint32_t)std::min<int64_t>(std::max<int64_t>(cap, 1), (int64_t)64 * n + 64);
  if (hipMalloc(&d_cnt, ((size_t)n + 1) * 4) != hipSuccess) return MG_E_DEVICE;
  if (hipMalloc(&d_nb, (size_t)dcap * 4) != hipSuccess) { (void)hipFree(d_cnt); return MG_E_DEVICE; }
  if (B->f64) launch_nbr_count<double>(B, d_cnt);
  else launch_nbr_count<float>(B, d_cnt);
  std::vector<uint32_t> cnt(n);
  bool ok = hipMemcpyAsync(cnt.data(), d_cnt, (size_t)n * 4, hipMemcpyDeviceToHost, B->stream) == hipSuccess &&
            hipStreamSynchronize(B->stream) == hipSuccess;
  int64_t total = 0;
  if (ok) {
    off[0] = 0;
    for (uint32_t i = 0; i < n; i++) { total += cnt[i]; off[i + 1] = (uint32_t)total; }
    ok = hipMemcpy(d_cnt, off, (size_t)n * 4, hipMemcpyHostToDevice) == hipSuccess;
  }
  if (ok) {
    if (B->f64) launch_nbr_write<double>(B, d_cnt, d_nb, dcap);
    else launch_nbr_write<float>(B, d_cnt, d_nb, dcap);
    ok = hipStreamSynchronize(B->stream) == hipSuccess;
  }
  if (ok && nb && total <= cap) ok = hipMemcpy(nb, d_nb, (size_t)total * 4, hipMemcpyDeviceToHost) == hipSuccess;
  (void)hipFree(d_cnt);
  (void)hipFree(d_nb);
  return ok ? total : (int64_t)MG_E_DEVICE;
}

}  // extern "C"
