// atan2f with the host libm's rounding, for host and device code.
//
// The reference computes wall angles as f32::atan2 (metrics.rs:336-337, :388-389, :465), which on
// Linux is the C library's atan2f. glibc's atan2f / atanf (sysdeps/ieee754/flt-32, glibc 2.35) are
// the fdlibm float algorithms (Sun Microsystems, 1993): special cases by IEEE bit patterns, the
// quotient |y/x| reduced into four intervals around 0.5, 1, 1.5 and infinity with split hi/lo
// values of atan there, and an odd degree-23 polynomial in the reduced argument evaluated as two
// interleaved Horner chains. This is that published algorithm restated; its results equal the
// host libm's atan2f bit for bit (checked on 10^8 inputs, a quarter of them random bit patterns
// including NaN, infinities and subnormals; the library exports mg_atan2f_ref for that check).
// Device code gets the same results because every operation is a correctly rounded IEEE f32
// operation and the build disables contraction into FMAs.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define MG_HD __host__ __device__
#else
#define MG_HD
#endif

namespace mgm {

MG_HD inline int32_t f2i(float f) { return __builtin_bit_cast(int32_t, f); }
MG_HD inline float i2f(int32_t i) { return __builtin_bit_cast(float, i); }
MG_HD inline float fabs_bits(float f) { return i2f(f2i(f) & 0x7fffffff); }

// atan of the reduced argument's four anchor points (0.5, 1, 1.5, infinity), split into a float
// head and a small tail, and the 11 odd-power coefficients of the approximating polynomial.
MG_HD inline float atan_hi(int k) {
  return k == 0 ? 4.6364760399e-01f : k == 1 ? 7.8539812565e-01f : k == 2 ? 9.8279368877e-01f : 1.5707962513e+00f;
}
MG_HD inline float atan_lo(int k) {
  return k == 0 ? 5.0121582440e-09f : k == 1 ? 3.7748947079e-08f : k == 2 ? 3.4473217170e-08f : 7.5497894159e-08f;
}

MG_HD inline float ref_atanf(float x) {
  const float c0 = 3.3333334327e-01f, c1 = -2.0000000298e-01f, c2 = 1.4285714924e-01f, c3 = -1.1111110449e-01f,
              c4 = 9.0908870101e-02f, c5 = -7.6918758452e-02f, c6 = 6.6610731184e-02f, c7 = -5.8335702866e-02f,
              c8 = 4.9768779427e-02f, c9 = -3.6531571299e-02f, c10 = 1.6285819933e-02f;
  const int32_t hx = f2i(x), ix = hx & 0x7fffffff;
  int id;
  if (ix >= 0x4c000000) {  // |x| >= 2^25
    if (ix > 0x7f800000) return x + x;  // NaN
    return hx > 0 ? atan_hi(3) + atan_lo(3) : -atan_hi(3) - atan_lo(3);
  }
  if (ix < 0x3ee00000) {  // |x| < 0.4375
    if (ix < 0x31000000) return x;  // |x| < 2^-29
    id = -1;
  } else {
    x = fabs_bits(x);
    if (ix < 0x3f980000) {    // |x| < 1.1875
      if (ix < 0x3f300000) {  // 7/16 <= |x| < 11/16
        id = 0;
        x = (2.0f * x - 1.0f) / (2.0f + x);
      } else {  // 11/16 <= |x| < 19/16
        id = 1;
        x = (x - 1.0f) / (x + 1.0f);
      }
    } else if (ix < 0x401c0000) {  // |x| < 2.4375
      id = 2;
      x = (x - 1.5f) / (1.0f + 1.5f * x);
    } else {  // 2.4375 <= |x| < 2^25
      id = 3;
      x = -1.0f / x;
    }
  }
  const float z = x * x, w = z * z;
  const float s1 = z * (c0 + w * (c2 + w * (c4 + w * (c6 + w * (c8 + w * c10)))));
  const float s2 = w * (c1 + w * (c3 + w * (c5 + w * (c7 + w * c9))));
  if (id < 0) return x - x * (s1 + s2);
  const float r = atan_hi(id) - ((x * (s1 + s2) - atan_lo(id)) - x);
  return hx < 0 ? -r : r;
}

MG_HD inline float ref_atan2f(float y, float x) {
  const float pi_o_4 = 7.8539818525e-01f, pi_o_2 = 1.5707963705e+00f, pi = 3.1415927410e+00f,
              pi_lo = -8.7422776573e-08f, tiny = 1.0e-30f;
  const int32_t hx = f2i(x), ix = hx & 0x7fffffff, hy = f2i(y), iy = hy & 0x7fffffff;
  if (ix > 0x7f800000 || iy > 0x7f800000) return x + y;  // NaN
  if (hx == 0x3f800000) return ref_atanf(y);             // x = 1
  const int m = ((hy >> 31) & 1) | ((hx >> 30) & 2);    // 2 sign(x) + sign(y)
  if (iy == 0) {
    if (m < 2) return y;                 // atan(+-0, +anything) = +-0
    return m == 2 ? pi + tiny : -pi - tiny;  // atan(+-0, -anything) = +-pi
  }
  if (ix == 0) return hy < 0 ? -pi_o_2 - tiny : pi_o_2 + tiny;
  if (ix == 0x7f800000) {
    if (iy == 0x7f800000) {
      switch (m) {
        case 0: return pi_o_4 + tiny;
        case 1: return -pi_o_4 - tiny;
        case 2: return 3.0f * pi_o_4 + tiny;
        default: return -3.0f * pi_o_4 - tiny;
      }
    }
    switch (m) {
      case 0: return 0.0f;
      case 1: return -0.0f;
      case 2: return pi + tiny;
      default: return -pi - tiny;
    }
  }
  if (iy == 0x7f800000) return hy < 0 ? -pi_o_2 - tiny : pi_o_2 + tiny;
  const int32_t k = (iy - ix) >> 23;
  float z;
  if (k > 60) {
    z = pi_o_2 + 0.5f * pi_lo;  // |y/x| > 2^60
  } else if (hx < 0 && k < -60) {
    z = 0.0f;  // |y|/x < -2^60
  } else {
    z = ref_atanf(fabs_bits(y / x));
  }
  switch (m) {
    case 0: return z;
    case 1: return i2f(f2i(z) ^ (int32_t)0x80000000);
    case 2: return pi - (z - pi_lo);
    default: return (z - pi_lo) - pi;
  }
}

}  // namespace mgm
