// Per-pixel shading helpers shared by the plain and batched renderers:
// ProcessHit colour lookup (GPU_renderer/kernels/shading.cl:4-33), the
// shadow-ray extension and the feedback code of a hit
// (raytracing_pass.cl:51, feedback_compression.cl:11).
#pragma once

#include "svo_device.h"

namespace rtdev {

__device__ __forceinline__ uint32_t LoadU32(const uint32_t* p, int64_t i, uint32_t fallback) {
  return p ? p[i] : fallback;
}

// Colour of ProcessHit for `node`, optionally with light 1 occluded.
__device__ __forceinline__ uint32_t ShadeNode(const uint32_t* normals, const uint32_t* colors, int node,
                                              const Lights& lt, bool occluded) {
  const uint32_t n = LoadU32(normals, node, 0u);
  const uint32_t c = LoadU32(colors, node, 0xffffffffu);
  return ShadeColor(n, c, lt, occluded);
}

// Shadow-ray test for a primary hit (extension; no reference counterpart).
// Cast from just before the hit point (a quarter of the hit voxel back along
// the primary ray, which is empty space) towards light 1, without LOD.
template <class Stack, bool COUNT, bool FETCHES = COUNT>
__device__ __forceinline__ bool ShadowOccluded(const PoolView& pool, uint32_t root,
                                               float ox, float oy, float oz, float dx, float dy, float dz,
                                               const CastResult& hit, const Lights& lt, Stack& stack,
                                               uint32_t& fetches, uint32_t& shadow_rays) {
  const int nd = hit.node_depth;
  const float eps = __uint_as_float(uint32_t(127 - nd - 2) << 23);
  const float tb = cl_max(0.0f, hit.hit_t - eps);
  const float sx = ox + dx * tb, sy = oy + dy * tb, sz = oz + dz * tb;
  CastResult s;
  CastRay<Stack, FETCHES, false>(pool, root, sx, sy, sz, lt.shadow_dir[0], lt.shadow_dir[1], lt.shadow_dir[2],
                                 0.0f, 0.0f, stack, s, fetches);
  if (COUNT) ++shadow_rays;
  return s.error_code == 0 && s.fault_block == -1 && s.hit_node != -1;
}

__device__ __forceinline__ int HitCode(int hit_node, int dup, int nodes_in_block) {
  const int unit = nodes_in_block > 0 ? hit_node / nodes_in_block : hit_node;
  return ((unit + 1) << 1) | dup;
}

struct PixelCounters {
  uint32_t primaries = 0, shadow_rays = 0, fetches = 0, hits = 0;
};

// One pixel of a single-pass frame: InitTracingFrame, one RaytracingPass
// (with the optional shadow ray) and FinishTracingFrame. Writes the RGBA8
// pixel and the feedback code (block-level when nodes_in_block > 0).
// LOD = false: lod == 0 (the primary ray has no LOD stop).
// COUNT counts the rays and hits; FETCHES (which needs COUNT) also the node
// fetches, which costs instructions in every traversal step.
template <class Stack, bool COUNT, bool LOD = true, bool FETCHES = COUNT>
__device__ __forceinline__ void RenderPixel(const PoolView& pool, uint32_t root, const uint32_t* normals,
                                            const uint32_t* colors, int nodes_in_block, const float* m, int x, int y,
                                            int W, int H, float lod, const float* bg, int flags, const Lights& lt,
                                            Stack& stack, uint32_t& out_color, int& out_code, PixelCounters& cnt) {
  const PrimaryRay r = InitRay(m, x, y, W, H, lod);
  CastResult res;
  uint32_t fetches = 0, shadow_rays = 0;
  static_assert(COUNT || !FETCHES, "fetches are counted with the rays");
  CastRay<Stack, FETCHES, LOD>(pool, root, r.origin.x, r.origin.y, r.origin.z, r.direction.x, r.direction.y,
                               r.direction.z, r.direction.w, r.origin.w, stack, res, fetches);
  uint32_t color;
  int code;
  if (res.error_code) {
    color = PackRGBA(255, 0, 0, 255);
    code = kNoHit;
  } else if (res.fault_block != -1) {
    // Unfinished after the single pass: FinishTracingFrame shades the fault parent.
    color = ShadeNode(normals, colors, res.hit_node, lt, false);
    code = -res.fault_block - 1;
  } else if (res.hit_node == -1) {
    color = 0u;
    code = kNoHit;
  } else {
    bool occluded = false;
    if ((flags & 1) && normals) {
      const uint32_t n = normals[res.hit_node];
      const float nx = float(int8_t(n & 255u)) / 127.f, ny = float(int8_t((n >> 8) & 255u)) / 127.f,
                  nz = float(int8_t((n >> 16) & 255u)) / 127.f;
      if (cl_max(0.f, LightTerm(nx, ny, nz, lt.l1)) > 0.0f)
        occluded = ShadowOccluded<Stack, COUNT, FETCHES>(pool, root, r.origin.x, r.origin.y, r.origin.z,
                                                         r.direction.x, r.direction.y, r.direction.z, res, lt, stack,
                                                         fetches, shadow_rays);
    }
    color = ShadeNode(normals, colors, res.hit_node, lt, occluded);
    code = HitCode(res.hit_node, res.hit_duplicate, nodes_in_block);
    if (COUNT) ++cnt.hits;
  }
  out_color = Blend(color, bg);
  out_code = code;
  if (COUNT) {
    ++cnt.primaries;
    cnt.shadow_rays += shadow_rays;
    cnt.fetches += fetches;
  }
}

// Counters are kept in kStatSlots copies of {primary rays, shadow rays, node
// fetches, hits}: each wave adds its sums into the copy of its wave index
// (one atomic per counter per wave, spread over the copies, so a launch of
// many one-wave workgroups does not serialise on four words), and the host
// adds the copies.
constexpr int kStatSlots = 64;

__device__ __forceinline__ uint32_t WaveSum(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ void FlushCounters(const PixelCounters& c, unsigned long long* stats) {
  const uint32_t p = WaveSum(c.primaries), sh = WaveSum(c.shadow_rays), f = WaveSum(c.fetches),
                 h = WaveSum(c.hits);
  if ((threadIdx.x & 63) == 0) {
    const uint32_t wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    unsigned long long* s = stats + 4 * (wave % kStatSlots);
    if (p) atomicAdd(&s[0], (unsigned long long)p);
    if (sh) atomicAdd(&s[1], (unsigned long long)sh);
    if (f) atomicAdd(&s[2], (unsigned long long)f);
    if (h) atomicAdd(&s[3], (unsigned long long)h);
  }
}


}  // namespace rtdev
