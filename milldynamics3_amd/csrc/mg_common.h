// Host-side helpers shared by the drum runtime and the periodic-box runtime: the reference's
// random stream and the f32 integer powers its kernel constants use.
#pragma once
#include <stdint.h>

namespace mg {

const float PI_F = 3.14159265358979323846f;
const float TAU_F = 6.28318530717958647692f;

// rng.rs:9-47: xorshift64 (13, 7, 17), then the xorshift64* multiplier; f32 from the top 24 bits.
struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed == 0 ? 0x9E3779B97F4A7C15ull : seed) {}
  uint64_t next_u64() {
    uint64_t x = s;
    x ^= x << 13;
    x ^= x >> 7;
    x ^= x << 17;
    s = x;
    return x * 0x2545F4914F6CDD1Dull;
  }
  float next_f32() { return (float)(next_u64() >> 40) / (float)(1u << 24); }
  float range(float lo, float hi) { return lo + next_f32() * (hi - lo); }
};

// f32::powi as LLVM expands it (binary exponentiation, low bit first).
inline float powi8(float x) { float a = x * x; a = a * a; return a * a; }
inline float powi5(float x) { float a = x * x; return x * (a * a); }
inline float powi6(float x) { float a = x * x; return a * (a * a); }

}  // namespace mg
