// expf with the host libm's rounding, for device code.
//
// The reference evaluates f32::exp (the friction blend, dem.rs:128-138; the form-drag relaxation,
// pbf.rs:1905-1907; the Bingham regularisation, pbf.rs:854-868), which on Linux is the C library's
// expf: glibc's exp2-table method (32 table entries of 2^(k/32) with the exponent folded into the
// bits, a degree-3 polynomial in the reduced argument, evaluated in double and rounded once).
// This is that published algorithm restated; every step is an IEEE double operation, so device
// results equal the host libm's bit for bit (the DEM parity tests exercise it through the friction
// blend; the build disables FMA contraction).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mgm {

static __device__ __constant__ uint64_t kExp2Tab[32] = {
    0x3ff0000000000000ull, 0x3fefd9b0d3158574ull, 0x3fefb5586cf9890full, 0x3fef9301d0125b51ull,
    0x3fef72b83c7d517bull, 0x3fef54873168b9aaull, 0x3fef387a6e756238ull, 0x3fef1e9df51fdee1ull,
    0x3fef06fe0a31b715ull, 0x3feef1a7373aa9cbull, 0x3feedea64c123422ull, 0x3feece086061892dull,
    0x3feebfdad5362a27ull, 0x3feeb42b569d4f82ull, 0x3feeab07dd485429ull, 0x3feea47eb03a5585ull,
    0x3feea09e667f3bcdull, 0x3fee9f75e8ec5f74ull, 0x3feea11473eb0187ull, 0x3feea589994cce13ull,
    0x3feeace5422aa0dbull, 0x3feeb737b0cdc5e5ull, 0x3feec49182a3f090ull, 0x3feed503b23e255dull,
    0x3feee89f995ad3adull, 0x3feeff76f2fb5e47ull, 0x3fef199bdd85529cull, 0x3fef3720dcef9069ull,
    0x3fef5818dcfba487ull, 0x3fef7c97337b9b5full, 0x3fefa4afa2a490daull, 0x3fefd0765b6e4540ull};
// tab: the 32 entries of kExp2Tab, wherever the caller keeps them (a kernel whose loop keeps vector
// loads in flight passes an LDS copy: waiting for a table load from global memory waits for every
// load issued before it).
__device__ inline float ref_expf_t(float x, const uint64_t* tab) {
  const uint32_t ux = __float_as_uint(x);
  const uint32_t abstop = (ux >> 20) & 0x7ffu;
  if (abstop >= 0x42bu) {  // |x| >= 88, infinities, NaN
    if (ux == 0xff800000u) return 0.0f;
    if (abstop >= 0x7f8u) return x + x;
    if (x > 0x1.62e42ep6f) return INFINITY;
    if (x < -0x1.9fe368p6f) return 0.0f;
  }
  const double inv_ln2_n = 0x1.71547652b82fep+0 * 32, shift = 0x1.8p+52;
  const double c0 = 0x1.c6af84b912394p-5 / 32 / 32 / 32, c1 = 0x1.ebfce50fac4f3p-3 / 32 / 32,
               c2 = 0x1.62e42ff0c52d6p-1 / 32;
  double z = inv_ln2_n * (double)x;
  double kd = z + shift;
  const uint64_t ki = (uint64_t)__double_as_longlong(kd);
  kd -= shift;
  const double r = z - kd;
  const uint64_t t = tab[ki % 32] + (ki << 47);
  const double s = __longlong_as_double((long long)t);
  z = c0 * r + c1;
  const double r2 = r * r;
  double y = c2 * r + 1.0;
  y = z * r2 + y;
  y = y * s;
  return (float)y;
}
__device__ inline float ref_expf(float x) { return ref_expf_t(x, kExp2Tab); }

}  // namespace mgm
