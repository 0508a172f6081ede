// Drum wall geometry for host and device code: the signed distance field with trapezoid lifters,
// its numeric-gradient normal and the wall velocity (crates/mill-core/src/geometry.rs:37-114,
// :233-262), in the reference's f32 evaluation order. Rotations between the world and drum frames
// take (cos, sin) pairs computed once per drum on the host with the host libm (the reference's
// Vec2::from_angle), so device and host evaluate the same f32 operations.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mgg {

struct Geo {
  float radius, omega;
  uint32_t nl;         // lifter count
  float rt, bh, thw;   // lifter top radius (radius - height), half base width, half top width
  const float2* lift;  // per lifter (sin, cos) of its angle (geometry.rs:50-56)
};

// convex_polygon_sdf (geometry.rs:233-262) over the trapezoid (radius, -bh), (radius, bh),
// (rt, thw), (rt, -thw) in the lifter's (radial, tangential) frame.
__host__ __device__ inline float lifter_sdf(float px, float py, const Geo& g) {
  const float vx[4] = {g.radius, g.radius, g.rt, g.rt};
  const float vy[4] = {-g.bh, g.bh, g.thw, -g.thw};
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

// Drum::sdf (geometry.rs:37-72) at a drum-frame point.
__host__ __device__ inline float sdf_local(float lx, float ly, const Geo& g) {
  const float dist = g.radius - sqrtf(lx * lx + ly * ly);
  if (g.nl == 0) return dist;
  float max_inside = -INFINITY;
  for (uint32_t k = 0; k < g.nl; k++) {
    const float2 sc = g.lift[k];
    const float pr = lx * sc.y + ly * sc.x;
    const float pt = -lx * sc.x + ly * sc.y;
    max_inside = fmaxf(max_inside, lifter_sdf(pr, pt, g));
  }
  return fminf(dist, -max_inside);
}

// Drum::sdf_world (geometry.rs:87-108): distance and world-frame normal at world point (px, py);
// (rc, rs) = from_angle(-angle), (wc, ws) = from_angle(angle). A non-finite point reads (0, 0).
__host__ __device__ inline float sdf_world(float px, float py, float rc, float rs, float wc, float ws, const Geo& g,
                                           float& nx, float& ny) {
  if (!(__builtin_isfinite(px) && __builtin_isfinite(py))) {
    nx = 0.0f;
    ny = 0.0f;
    return 0.0f;
  }
  const float lx = px * rc - py * rs, ly = px * rs + py * rc;
  const float eps = 1e-4f;
  const float d = sdf_local(lx, ly, g);
  const float dx = sdf_local(lx + eps, ly + 0.0f, g) - sdf_local(lx - eps, ly - 0.0f, g);
  const float dy = sdf_local(lx + 0.0f, ly + eps, g) - sdf_local(lx - 0.0f, ly - eps, g);
  const float gx = dx / (2.0f * eps), gy = dy / (2.0f * eps);
  float nlx, nly;
  if (gx * gx + gy * gy > 1e-12f) {
    const float rcp = 1.0f / sqrtf(gx * gx + gy * gy);
    nlx = gx * rcp;
    nly = gy * rcp;
  } else {  // -p_local.normalize_or_zero()
    const float rcp = 1.0f / sqrtf(lx * lx + ly * ly);
    if (__builtin_isfinite(rcp) && rcp > 0.0f) {
      nlx = -(lx * rcp);
      nly = -(ly * rcp);
    } else {
      nlx = -0.0f;
      nly = -0.0f;
    }
  }
  nx = nlx * wc - nly * ws;
  ny = nlx * ws + nly * wc;
  return d;
}

// The distance alone.
__host__ __device__ inline float sdf_world_d(float px, float py, float rc, float rs, const Geo& g) {
  if (!(__builtin_isfinite(px) && __builtin_isfinite(py))) return 0.0f;
  return sdf_local(px * rc - py * rs, px * rs + py * rc, g);
}

}  // namespace mgg
