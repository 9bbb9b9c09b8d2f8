// Device helpers for the single-pass variants: the log-depth grey output of
// GPU_renderer/kernels/raytracer.cl:63-68 and the stackless restart traversal
// of raytracer_volatile.cl:25-112 over the old link layout
// (ptr = link >> 9, fault = bit 8, no far or duplicate bits).
#pragma once

#include "svo_device.h"

namespace rayt {
rtdev::Lights MakeLights();
}

namespace rtdev {

// Natural logarithm of a float, evaluated in double with a fixed operation
// order and rounded to float, so that every IEEE-754 implementation (the
// oracle on the host, this kernel on gfx950, both without contraction) gives
// the same bits; a libm `log` differs between the two by an ulp now and then.
// Method: the argument reduction t = m * 2^e with m in [sqrt(1/2), sqrt(2))
// and the series ln m = 2 atanh(s) = 2 (s + s^3/3 + s^5/5 + ...),
// s = (m - 1) / (m + 1) (Abramowitz & Stegun 4.1.27), whose terms up to s^23
// bring the truncation below 2^-55 relative for |s| <= 0.1716.
__device__ __forceinline__ float LnFixed(float t) {
  if (!(t > 0.0f)) return t == 0.0f ? -__builtin_inff() : __builtin_nanf("");
  if (t == __builtin_inff()) return t;
  const uint64_t bits = uint64_t(__double_as_longlong(double(t)));
  int e = int((bits >> 52) & 0x7ff) - 1023;  // t is a float, so its double is normal
  double m = __longlong_as_double(int64_t((bits & 0x000fffffffffffffull) | 0x3ff0000000000000ull));
  if (m > 1.4142135623730951) {
    m = m * 0.5;
    e = e + 1;
  }
  const double s = (m - 1.0) / (m + 1.0);
  const double z = s * s;
  double p = 1.0 / 23.0;
  p = p * z + 1.0 / 21.0;
  p = p * z + 1.0 / 19.0;
  p = p * z + 1.0 / 17.0;
  p = p * z + 1.0 / 15.0;
  p = p * z + 1.0 / 13.0;
  p = p * z + 1.0 / 11.0;
  p = p * z + 1.0 / 9.0;
  p = p * z + 1.0 / 7.0;
  p = p * z + 1.0 / 5.0;
  p = p * z + 1.0 / 3.0;
  p = p * z + 1.0;
  const double ln2 = 0.6931471805599453094;
  return float(double(e) * ln2 + 2.0 * (s * p));
}

// Grey = (1 - (ln t - ln 0.01) / (ln 5 - ln 0.01)) * 255 clamped to [0,255]
// and truncated (raytracer.cl:63-68), ln from LnFixed; the two constant
// logarithms are fixed fp32 values.
__device__ __forceinline__ uint32_t DepthGrey(float t) {
  const float ln_near = -4.60517025f;  // (float)log(0.01)
  const float ln_far = 1.60943794f;    // (float)log(5.0)
  const float lt = LnFixed(t);
  float d = (lt - ln_near) / (ln_far - ln_near);
  d = (1.0f - d) * 255.0f;
  const float g = cl_max(0.0f, cl_min(255.0f, d));
  const uint32_t gi = uint32_t(g);
  return gi | (gi << 8) | (gi << 16) | (255u << 24);
}

struct RestartResult {
  float hit_t;
  int hit_node;
  int fault_block;
  int error_code;
};

__device__ __forceinline__ float MinComponent(float x, float y, float z) { return cl_fmin(cl_fmin(x, y), z); }
__device__ __forceinline__ float MaxComponent(float x, float y, float z) { return cl_fmax(cl_fmax(x, y), z); }

// Restart ("kd-restart") traversal in the cube [0,1]^3. A node index at or
// beyond `node_count` (a malformed link buffer) ends the ray with error 2
// instead of reading outside the buffer.
__device__ __forceinline__ RestartResult CastRayRestart(const uint32_t* __restrict__ links, uint32_t node_count,
                                                        uint32_t root, float ox, float oy, float oz, float dx,
                                                        float dy, float dz, float ray_size_coef,
                                                        float ray_size_bias) {
  RestartResult r;
  r.fault_block = -1;
  r.error_code = 0;
  r.hit_t = 0.0f;
  r.hit_node = -1;
  const float epsilon = 1e-6f;
  if (fabsf(dx) < epsilon) dx = copysignf(epsilon, dx);
  if (fabsf(dy) < epsilon) dy = copysignf(epsilon, dy);
  if (fabsf(dz) < epsilon) dz = copysignf(epsilon, dz);
  const float cx = 1.0f / fabsf(dx), cy = 1.0f / fabsf(dy), cz = 1.0f / fabsf(dz);
  float bx = cx * ox, by = cy * oy, bz = cz * oz;
  uint32_t octant_mask = 0;
  if (dx < 0.0f) octant_mask ^= 1u, bx = cx - bx;
  if (dy < 0.0f) octant_mask ^= 2u, by = cy - by;
  if (dz < 0.0f) octant_mask ^= 4u, bz = cz - bz;
  const float t_max = MinComponent(cx - bx, cy - by, cz - bz);
  float t_min = cl_fmax(0.0f, MaxComponent(-bx, -by, -bz));
  int steps = 0;
  while (t_max > t_min) {
    uint32_t node = root;
    float px = 0.0f, py = 0.0f, pz = 0.0f;
    float size = 1.0f;
    const float ray_size = ray_size_bias + ray_size_coef * t_min;
    while (true) {
      if (++steps > 5000) {
        r.error_code = 1;
        return r;
      }
      if (size < ray_size) {
        r.hit_t = t_min;
        r.hit_node = int(node);
        return r;
      }
      if (node >= node_count) {
        r.error_code = 2;
        return r;
      }
      const uint32_t link = links[node];
      const uint32_t children_mask = link & 255u;
      if (!children_mask) {
        r.hit_t = t_min;
        r.hit_node = int(node);
        return r;
      }
      size = size * 0.5f;
      const float nx = px + size, ny = py + size, nz = pz + size;
      const float tcx = nx * cx - bx, tcy = ny * cy - by, tcz = nz * cz - bz;
      uint32_t index = octant_mask;
      if (tcx < t_min) index ^= 1u, px = nx;
      if (tcy < t_min) index ^= 2u, py = ny;
      if (tcz < t_min) index ^= 4u, pz = nz;
      const uint32_t index_mask = (1u << index) & 255u;
      if (!(children_mask & index_mask)) {
        t_min = MinComponent((px + size) * cx - bx, (py + size) * cy - by, (pz + size) * cz - bz) + epsilon;
        break;
      }
      if (link & (1u << 8)) {
        r.hit_t = t_min;
        r.hit_node = int(node);
        r.fault_block = int(link >> 9);
        return r;
      }
      node = (link >> 9) + __popc(children_mask & (index_mask - 1u) & 255u);
    }
  }
  r.hit_t = 1e30f;
  r.hit_node = -1;
  return r;
}

}  // namespace rtdev
