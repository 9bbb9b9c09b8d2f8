// Renderer kernels for gfx950 and their C-ABI launchers.
//
// Workgroups are 256 lanes covering a 16x16 pixel tile made of four wave64
// 8x8 sub-tiles, so the 64 rays of a wave are spatially coherent and walk the
// same octree nodes. Tiles go out in raster order, so the round-robin
// dispatch over the 8 XCDs gives every XCD every 8th tile: the XCDs then
// share each band of the image, and its cost. Giving each XCD a contiguous
// eighth of the tiles instead (for L2 locality) left the XCDs with the sky
// idle while those with the terrain worked: 0.68 against 0.56-0.57 ms for the
// 1080p frame; groups of 2 or 4 tiles per XCD measured the same as raster.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "block_scan.h"
#include "mark_codes.h"
#include "rt_internal.h"
#include "svo_device.h"
#include "trace_common.h"
#include "render_pixel.h"

using namespace rtdev;

namespace rayt {

Lights MakeLights() {
  // shading.cl:8-10 with normalize(v) = v / sqrt(dot(v,v)) in fp32.
  auto norm = [](float x, float y, float z, float k, float* out) {
    const float s = sqrtf(((x * x + y * y) + z * z) + 0.0f * 0.0f);
    out[0] = (x / s) * k;
    out[1] = (y / s) * k;
    out[2] = (z / s) * k;
  };
  Lights lt;
  norm(2.f, -2.f, 3.f, .6f, lt.l1);
  norm(-3.f, -3.f, -2.f, .5f, lt.l2);
  norm(1.f, 2.5f, -1.f, .3f, lt.l3);
  float u[3];
  norm(2.f, -2.f, 3.f, 1.0f, u);
  lt.shadow_dir[0] = -u[0];
  lt.shadow_dir[1] = -u[1];
  lt.shadow_dir[2] = -u[2];
  return lt;
}

}  // namespace rayt

namespace {

struct Mat16 {
  float m[16];
};
struct Vec4 {
  float v[4];
};

__device__ __forceinline__ void TilePixel(uint32_t tile, uint32_t tiles_x, int& x, int& y) {
  const uint32_t t = threadIdx.x;
  const uint32_t wave = t >> 6, lane = t & 63u;
  x = int((tile % tiles_x) * 16u + (wave & 1u) * 8u + (lane & 7u));
  y = int((tile / tiles_x) * 16u + (wave >> 1) * 8u + (lane >> 3));
}

// ------------------------------------------------------------ InitTracingFrame
__global__ __launch_bounds__(256) void InitTracingFrameKernel(float4* __restrict__ org, float4* __restrict__ dir,
                                                              int4* __restrict__ misc, int W, int H, Mat16 m,
                                                              float lod) {
  const uint32_t tiles_x = (W + 15) / 16;
  int x, y;
  TilePixel(blockIdx.x, tiles_x, x, y);
  if (x >= W || y >= H) return;
  const int64_t i = int64_t(y) * W + x;
  const PrimaryRay r = InitRay(m.m, x, y, W, H, lod);
  org[i] = r.origin;
  dir[i] = r.direction;
  misc[i] = make_int4(-1, 0, int(0xffffffffu), 0);
}

// -------------------------------------------------------------- RaytracingPass
// The stack and the workgroup follow the context's launch shape
// (rt_plain_config): a window of LDS_LEVELS levels from `base`, spilling to a
// private array (SPILL) or alone; WAVES = 4: 16x16-pixel workgroups of four
// waves, WAVES = 1: one wave per 8x8 tile.
template <int WAVES, int LDS_LEVELS, bool SPILL>
__global__ __launch_bounds__(64 * WAVES) void RaytracingPassKernel(
    int* __restrict__ codes, float4* __restrict__ org, const float4* __restrict__ dir, int4* __restrict__ misc, int W,
    int H, PoolView pool, uint32_t root, const uint32_t* __restrict__ normals, const uint32_t* __restrict__ colors,
    int nodes_in_block, int flags, Lights lt, int base) {
  constexpr int BLOCK = 64 * WAVES;
  __shared__ uint2 lds[LDS_LEVELS * BLOCK];
  int x, y;
  if constexpr (WAVES == 4) {
    TilePixel(blockIdx.x, (W + 15) / 16, x, y);
  } else {
    const uint32_t tiles_x = (W + 7) / 8;
    x = int((blockIdx.x % tiles_x) * 8u + (threadIdx.x & 7u));
    y = int((blockIdx.x / tiles_x) * 8u + (threadIdx.x >> 3));
  }
  if (x >= W || y >= H) return;
  const int64_t i = int64_t(y) * W + x;
  int4 s = misc[i];
  if (uint32_t(s.z) == 0u) {
    codes[i] = kNoHit;
    return;
  }
  using Stack = SplitStack<LDS_LEVELS, BLOCK, SPILL>;
  StackSpill<SPILL ? LDS_LEVELS : kStackLevels> spill;
  Stack stack;
  stack.spill = spill.v;
  stack.lds = lds + threadIdx.x;
  stack.set_base(base);
  float4 o = org[i];
  const float4 d = dir[i];
  CastResult res;
  uint32_t fetches = 0, shadow_rays = 0;
  CastRay<Stack, false>(pool, root, o.x, o.y, o.z, d.x, d.y, d.z, d.w, o.w, stack, res, fetches);
  int code;
  if (res.error_code) {
    s.y = int(PackRGBA(255, 0, 0, 255));
    s.z = 0;
    code = kNoHit;
  } else if (res.fault_block != -1) {
    o = make_float4(o.x + d.x * res.hit_t, o.y + d.y * res.hit_t, o.z + d.z * res.hit_t, o.w + d.w * res.hit_t);
    org[i] = o;
    s.x = res.hit_node;
    code = -res.fault_block - 1;
  } else if (res.hit_node == -1) {
    s.z = 0;
    code = kNoHit;
  } else {
    bool occluded = false;
    if ((flags & RT_FLAG_SHADOWS) && normals) {
      const uint32_t n = normals[res.hit_node];
      const float nx = float(int8_t(n & 255u)) / 127.f, ny = float(int8_t((n >> 8) & 255u)) / 127.f,
                  nz = float(int8_t((n >> 16) & 255u)) / 127.f;
      if (cl_max(0.f, LightTerm(nx, ny, nz, lt.l1)) > 0.0f)
        occluded = ShadowOccluded<Stack, false>(pool, root, o.x, o.y, o.z, d.x, d.y, d.z, res, lt, stack, fetches,
                                                shadow_rays);
    }
    s.y = int(ShadeNode(normals, colors, res.hit_node, lt, occluded));
    s.z = 0;
    code = HitCode(res.hit_node, res.hit_duplicate, nodes_in_block);
  }
  misc[i] = s;
  codes[i] = code;
}

// ---------------------------------------------------------- FinishTracingFrame
__global__ __launch_bounds__(256) void FinishTracingFrameKernel(uint32_t* __restrict__ out,
                                                                const int4* __restrict__ misc, int W, int H, Vec4 bg,
                                                                const uint32_t* __restrict__ normals,
                                                                const uint32_t* __restrict__ colors, Lights lt,
                                                                uint32_t channel_nodes) {
  const uint32_t tiles_x = (W + 15) / 16;
  int x, y;
  TilePixel(blockIdx.x, tiles_x, x, y);
  if (x >= W || y >= H) return;
  const int64_t i = int64_t(y) * W + x;
  const int4 s = misc[i];
  uint32_t color = uint32_t(s.y);
  // a fault parent outside the channels (only in states no pass wrote) shades as no node
  if (uint32_t(s.z) != 0u)
    color = s.x >= 0 && uint32_t(s.x) < channel_nodes ? ShadeNode(normals, colors, s.x, lt, false) : 0u;
  out[i] = Blend(color, bg.v);
}

// ---------------------------------------------------------------- plain fused
// WAVES = 4: workgroups of 256 lanes over 16x16 pixels (four 8x8 wave tiles);
// WAVES = 1: one wave per workgroup over an 8x8 tile, so a finished wave frees
// its LDS stack at once and more waves fit a CU.
// mark.fault_bits != nullptr: each wave also marks its (block-level) codes
// into the feedback extractor's bitmaps.
// order / cost (TileSchedule): the tile of workgroup i is order[i] (raster
// order without one), and each wave stores its duration in cycles at
// cost[tile * WAVES + wave].
template <int WAVES, int LDS_LEVELS, bool COUNT, bool SPILL, bool LOD>
__global__ __launch_bounds__(64 * WAVES) void RenderPlainKernel(
    uint32_t* __restrict__ out, int* __restrict__ codes, int W, int H, PoolView pool, uint32_t root,
    const uint32_t* __restrict__ normals, const uint32_t* __restrict__ colors, int nodes_in_block, Mat16 m,
    float lod, Vec4 bg, int flags, Lights lt, int lds_base, unsigned long long* __restrict__ stats,
    rayt::MarkTargets mark, const uint32_t* __restrict__ order, uint32_t* __restrict__ cost) {
  constexpr int BLOCK = 64 * WAVES;
  constexpr int TS = WAVES == 4 ? 16 : 8;
  __shared__ uint2 lds[LDS_LEVELS * BLOCK];
  const uint64_t t_start = cost ? clock64() : 0;
  const uint32_t tiles_x = (W + TS - 1) / TS;
  const uint32_t tile = order ? order[blockIdx.x] : blockIdx.x;
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  const int x = int((tile % tiles_x) * TS + (wave & 1u) * 8u + (lane & 7u));
  const int y = int((tile / tiles_x) * TS + (wave >> 1) * 8u + (lane >> 3));
  PixelCounters cnt;
  int code = 0;
  const bool inside = x < W && y < H;
  if (inside) {
    const int64_t i = int64_t(y) * W + x;
    StackSpill<SPILL ? LDS_LEVELS : kStackLevels> spill;
    SplitStack<LDS_LEVELS, BLOCK, SPILL> stack;
    stack.spill = spill.v;
    stack.lds = lds + threadIdx.x;
    stack.set_base(lds_base);
    uint32_t color;
    RenderPixel<SplitStack<LDS_LEVELS, BLOCK, SPILL>, COUNT, LOD>(pool, root, normals, colors, nodes_in_block, m.m,
                                                                  x, y, W, H, lod, bg.v, flags, lt, stack, color, code,
                                                                  cnt);
    out[i] = color;
    if (codes) codes[i] = code;
  }
  if (mark.fault_bits) rayt::MarkWaveCodes(code, inside, mark);
  if (COUNT) FlushCounters(cnt, stats);
  if (cost && lane == 0) cost[tile * WAVES + wave] = uint32_t(min(clock64() - t_start, uint64_t(0xffffffffu)));
}

// The next launch's tile order from this launch's wave costs: the tiles in
// eight classes by their longest wave (half an octave each below the longest
// tile's, the last class open-ended; costs on a log scale of 1/8 octave),
// longest class first, raster order within a class. One workgroup of 16 waves; wave
// w owns a contiguous run of tiles, reads their costs once (sixteen loads in
// flight per lane, unconditional: an index past the run reads its last tile)
// into one log-cost byte per tile in LDS, and places the tiles by ballots.
constexpr int kOrderWaves = 16;
constexpr int kOrderMaxTiles = rayt::kMaxScheduledTiles;  // LDS holds one byte per tile
constexpr int kOrderClasses = 8;  // half-octave cost classes, the last one open-ended

__device__ __forceinline__ uint32_t LogCost(uint32_t c) {  // 8 x log2(c), from the exponent and 3 mantissa bits
  if (c < 16u) return c >> 1;
  const uint32_t e = 31u - __clz(c);
  return e * 8u + ((c >> (e - 3u)) & 7u);
}

template <int WAVES>
__global__ __launch_bounds__(64 * kOrderWaves) void TileOrderKernel(const uint32_t* __restrict__ cost, int tiles,
                                                                    uint32_t* __restrict__ order) {
  constexpr int kLoads = 16;
  __shared__ uint8_t lc[kOrderMaxTiles];
  __shared__ uint32_t counts[kOrderWaves][kOrderClasses];
  __shared__ uint32_t top;
  const int w = int(threadIdx.x >> 6), lane = int(threadIdx.x & 63);
  if (threadIdx.x == 0) top = 0;
  __syncthreads();
  const int chunk = (((tiles + kOrderWaves - 1) / kOrderWaves) + 63) & ~63;
  const int b = min(tiles, w * chunk), e = min(tiles, b + chunk);
  uint32_t mx = 0;
  for (int t0 = b; t0 < e; t0 += 64 * kLoads) {
    uint32_t c[kLoads];
#pragma unroll
    for (int k = 0; k < kLoads; ++k) {
      const int t = min(t0 + 64 * k + lane, e - 1);
      c[k] = cost[t * WAVES];
#pragma unroll
      for (int v = 1; v < WAVES; ++v) c[k] = max(c[k], cost[t * WAVES + v]);
    }
#pragma unroll
    for (int k = 0; k < kLoads; ++k) {
      const int t = t0 + 64 * k + lane;
      if (t < e) {
        const uint32_t l = LogCost(c[k]);
        lc[t] = uint8_t(l);
        mx = max(mx, l);
      }
    }
  }
  for (int o = 32; o >= 1; o >>= 1) mx = max(mx, uint32_t(__shfl_xor(int(mx), o, 64)));
  if (lane == 0) atomicMax(&top, mx);
  __syncthreads();
  const uint32_t T = top;
  auto cls = [&](int t) {  // half an octave per class below the longest tile
    const uint32_t d = (T - uint32_t(lc[t])) >> 2;
    return int(min(d, uint32_t(kOrderClasses - 1)));
  };
  uint32_t n[kOrderClasses] = {};
  for (int t0 = b; t0 < e; t0 += 64) {
    const int t = t0 + lane;
    const int c = t < e ? cls(t) : kOrderClasses;
#pragma unroll
    for (int k = 0; k < kOrderClasses; ++k) n[k] += __popcll(__ballot(c == k));
  }
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < kOrderClasses; ++k) counts[w][k] = n[k];
  __syncthreads();
  if (threadIdx.x == 0) {  // class-major, then wave (= tile) order: starting positions
    uint32_t run = 0;
    for (int k = 0; k < kOrderClasses; ++k)
      for (int v = 0; v < kOrderWaves; ++v) {
        const uint32_t m = counts[v][k];
        counts[v][k] = run;
        run += m;
      }
  }
  __syncthreads();
  uint32_t base[kOrderClasses];
#pragma unroll
  for (int k = 0; k < kOrderClasses; ++k) base[k] = counts[w][k];
  const uint64_t below = (uint64_t(1) << lane) - 1;
  for (int t0 = b; t0 < e; t0 += 64) {
    const int t = t0 + lane;
    const int c = t < e ? cls(t) : kOrderClasses;
    uint32_t pos = 0;
#pragma unroll
    for (int k = 0; k < kOrderClasses; ++k) {
      const uint64_t m = __ballot(c == k);
      if (c == k) pos = base[k] + uint32_t(__popcll(m & below));
      base[k] += __popcll(m);
    }
    if (t < e) order[pos] = uint32_t(t);
  }
}

// ----------------------------------------------------- RaytraceKernel (depth)
__global__ __launch_bounds__(256) void RaytraceDepthKernel(uint32_t* __restrict__ out, int W, int H,
                                                           PoolView pool, uint32_t root,
                                                           Mat16 m, float lod) {
  __shared__ uint2 lds[kStackLevels * 256];
  const uint32_t tiles_x = (W + 15) / 16;
  int x, y;
  TilePixel(blockIdx.x, tiles_x, x, y);
  if (x >= W || y >= H) return;
  const float fw = float(W), fh = float(H);
  const float u = (float(x) + 0.5f) / fw * 2.0f - 1.0f, v = (float(y) + 0.5f) / fh * 2.0f - 1.0f;
  const float4 origin = MulPoint(m.m, u, v, -1.0f);
  const float4 pf = MulPoint(m.m, u, v, 0.0f);
  float4 d = make_float4(pf.x - origin.x, pf.y - origin.y, pf.z - origin.z, pf.w - origin.w);
  const float len = sqrtf(Dot4(d, d));
  d = make_float4(d.x / len, d.y / len, d.z / len, d.w / len);
  const float x0 = float(x) / fw * 2.0f - 1.0f, y0 = float(y) / fh * 2.0f - 1.0f;
  const float x1 = float(x + 1) / fw * 2.0f - 1.0f, y1 = float(y + 1) / fh * 2.0f - 1.0f;
  const float4 p1n = MulPoint(m.m, x0, y0, -1.0f), p1f = MulPoint(m.m, x0, y0, 0.0f);
  const float4 p2n = MulPoint(m.m, x1, y1, -1.0f), p2f = MulPoint(m.m, x1, y1, 0.0f);
  auto dist = [](float4 a, float4 b) {
    float4 t = make_float4(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w);
    return sqrtf(Dot4(t, t));
  };
  const float bias = lod * dist(p1n, p2n);
  const float coef = (bias < 1e-10f) ? 0.0f : (lod * (dist(p1f, p2f) - dist(p1n, p2n)) / dist(p1n, p1f));
  StackSpill<kStackLevels> spill;
  SplitStack<kStackLevels, 256> stack;
  stack.spill = spill.v;
  stack.lds = lds + threadIdx.x;
  CastResult res;
  uint32_t fetches = 0;
  CastRay<SplitStack<kStackLevels, 256>, false>(pool, root, origin.x, origin.y, origin.z, d.x, d.y, d.z,
                                                coef, bias, stack, res, fetches);
  const int64_t i = int64_t(y) * W + x;
  if (res.error_code || res.fault_block != -1) {
    out[i] = PackRGBA(255, 0, 0, 255);
  } else if (res.hit_node == -1) {
    out[i] = PackRGBA(0, 0, 0, 255);
  } else {
    out[i] = DepthGrey(res.hit_t);
  }
}

// ------------------------------------------- legacy restart traversal (volatile)
__global__ __launch_bounds__(256) void RaytraceRestartKernel(uint32_t* __restrict__ out, int W, int H,
                                                             const uint32_t* __restrict__ links,
                                                             uint32_t node_count, uint32_t root, Mat16 m,
                                                             float lod) {
  const uint32_t tiles_x = (W + 15) / 16;
  int x, y;
  TilePixel(blockIdx.x, tiles_x, x, y);
  if (x >= W || y >= H) return;
  const float fw = float(W), fh = float(H);
  const float u = (float(x) + 0.5f) / fw * 2.0f - 1.0f, v = (float(y) + 0.5f) / fh * 2.0f - 1.0f;
  const float4 origin = MulPoint(m.m, u, v, -1.0f);
  const float4 pf = MulPoint(m.m, u, v, 0.0f);
  float4 d = make_float4(pf.x - origin.x, pf.y - origin.y, pf.z - origin.z, pf.w - origin.w);
  const float len = sqrtf(Dot4(d, d));
  d = make_float4(d.x / len, d.y / len, d.z / len, d.w / len);
  const float x0 = float(x) / fw * 2.0f - 1.0f, y0 = float(y) / fh * 2.0f - 1.0f;
  const float x1 = float(x + 1) / fw * 2.0f - 1.0f, y1 = float(y + 1) / fh * 2.0f - 1.0f;
  const float4 p1n = MulPoint(m.m, x0, y0, -1.0f), p1f = MulPoint(m.m, x0, y0, 0.0f);
  const float4 p2n = MulPoint(m.m, x1, y1, -1.0f), p2f = MulPoint(m.m, x1, y1, 0.0f);
  auto dist = [](float4 a, float4 b) {
    float4 t = make_float4(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w);
    return sqrtf(Dot4(t, t));
  };
  const float bias = lod * dist(p1n, p2n);
  const float coef = (bias < 1e-10f) ? 0.0f : (lod * (dist(p1f, p2f) - dist(p1n, p2n)) / dist(p1n, p1f));
  RestartResult res = CastRayRestart(links, node_count, root, origin.x, origin.y, origin.z, d.x, d.y, d.z, coef,
                                     bias);
  const int64_t i = int64_t(y) * W + x;
  if (res.error_code) {
    out[i] = PackRGBA(255, 0, 0, 255);
  } else {
    const float ht = cl_min(1.0f, res.hit_t / 5.0f);
    const float g = 255.0f - ht * 255.0f;
    const uint32_t gi = uint32_t(g);
    out[i] = PackRGBA(gi, gi, gi, 255);
  }
}

// A far link whose index lies outside the far-pointer table raises the
// overflow flag (the caller reports a range error) instead of reading past it.
__global__ void LegacyLinksKernel(uint32_t* __restrict__ out, const uint32_t* __restrict__ links,
                                  const uint32_t* __restrict__ far_ptrs, uint64_t far_count, uint64_t n,
                                  int* __restrict__ overflow) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t l = links[i];
  const uint32_t mask = l & 255u;
  uint32_t o = 0;
  if (mask) {
    uint64_t target;
    if (l & (1u << 8)) {
      target = l >> 11;  // fault: child block index
      o = (1u << 8);
    } else if (l & (1u << 10)) {
      target = (l >> 11) < far_count ? far_ptrs[l >> 11] : (1ull << 23);
    } else {
      target = uint64_t(int64_t(i) + int64_t(l >> 11) - (1 << 20));
    }
    if (target >= (1ull << 23)) {
      atomicOr(overflow, 1);
      target = 0;
    }
    o |= uint32_t(target << 9) | mask;
  }
  out[i] = o;
}

int CheckTraceArgs(rt_ctx* ctx, int W, int H) {
  RT_REQUIRE(ctx, RT_ERR_INVALID, "null context");
  RT_REQUIRE(W > 0 && H > 0 && int64_t(W) * H <= (int64_t(1) << 30), RT_ERR_RANGE, "bad frame size");
  return RT_OK;
}

uint32_t TileGrid(int W, int H) { return uint32_t((W + 15) / 16) * uint32_t((H + 15) / 16); }

// A normals or colours channel (one 4-byte word per node) must cover every
// node of the link buffer: the shading reads them at hit node indices.
int CheckChannels(const rt_buf* links, const rt_buf* normals, const rt_buf* colors) {
  const uint64_t nodes = links->size / 4;
  RT_REQUIRE(!normals || normals->size / 4 >= nodes, RT_ERR_RANGE, "normals channel smaller than the node links");
  RT_REQUIRE(!colors || colors->size / 4 >= nodes, RT_ERR_RANGE, "colours channel smaller than the node links");
  return RT_OK;
}

// Nodes the finishing shade may read: the smaller of the channels given (no
// bound where neither is: the shade then reads nothing).
uint32_t ChannelNodes(const rt_buf* normals, const rt_buf* colors) {
  uint64_t n = ~uint64_t(0);
  if (normals) n = std::min<uint64_t>(n, normals->size / 4);
  if (colors) n = std::min<uint64_t>(n, colors->size / 4);
  return uint32_t(std::min<uint64_t>(n, 0xffffffffu));
}

const uint32_t* U32(const rt_buf* b) { return b ? static_cast<const uint32_t*>(b->ptr) : nullptr; }

PoolView View(const rt_buf* links, const rt_buf* far_ptrs) {
  PoolView v;
  v.links = U32(links);
  v.far_ptrs = U32(far_ptrs);
  v.nodes = uint32_t(std::min<uint64_t>(links->size / 4, 0xffffffffu));
  v.far_count = uint32_t(std::min<uint64_t>(far_ptrs->size / 4, 0xffffffffu));
  return v;
}

template <int L, bool SP>
void LaunchPassL(int waves, hipStream_t s, int* codes, float4* org, float4* dir, int4* misc, int W, int H,
                 const PoolView& pool, uint32_t root, const uint32_t* nrm, const uint32_t* col,
                 int nib, int flags, const Lights& lt, int base) {
  if (waves == 4)
    RaytracingPassKernel<4, L, SP><<<dim3(TileGrid(W, H)), dim3(256), 0, s>>>(codes, org, dir, misc, W, H, pool,
                                                                                root, nrm, col, nib, flags, lt, base);
  else
    RaytracingPassKernel<1, L, SP><<<dim3(uint32_t((W + 7) / 8) * uint32_t((H + 7) / 8)), dim3(64), 0, s>>>(
        codes, org, dir, misc, W, H, pool, root, nrm, col, nib, flags, lt, base);
}

// The pass kernel for the context's launch shape (waves per workgroup) and stack window (levels, spill, base).
bool DispatchPass(int levels, bool spill, int waves, hipStream_t s, int* codes, float4* org, float4* dir, int4* misc,
                  int W, int H, const PoolView& pool, uint32_t root, const uint32_t* nrm, const uint32_t* col,
                  int nib, int flags, const Lights& lt, int base) {
#define RT_PASS(L)                                                                                           \
  if (levels == L) {                                                                                         \
    if (spill || L >= kStackLevels)                                                                          \
      LaunchPassL<L, true>(waves, s, codes, org, dir, misc, W, H, pool, root, nrm, col, nib, flags, lt, base); \
    else                                                                                                     \
      LaunchPassL<L, L >= kStackLevels>(waves, s, codes, org, dir, misc, W, H, pool, root, nrm, col, nib, flags, \
                                        lt, base);                                                           \
    return true;                                                                                             \
  }
  RT_PASS(8) RT_PASS(10) RT_PASS(12) RT_PASS(14) RT_PASS(16) RT_PASS(22)
#undef RT_PASS
  return false;
}

// The timing variant comes with and without the LOD stop (lod == 0 drops it); the counting variant keeps it.
template <int WV, int L, bool C, bool SP>
void LaunchPlainL(dim3 g, hipStream_t s, uint32_t* out, int* codes, int W, int H, const PoolView& pool,
                  uint32_t root, const uint32_t* nrm, const uint32_t* col, int nib, const Mat16& m, float lod,
                  const Vec4& bg, int flags, const Lights& lt, int base, unsigned long long* st,
                  const rayt::MarkTargets& mk, const uint32_t* order, uint32_t* cost) {
  if constexpr (!C) {
    if (lod == 0.0f) {
      RenderPlainKernel<WV, L, C, SP, false><<<g, dim3(64 * WV), 0, s>>>(out, codes, W, H, pool, root, nrm, col, nib,
                                                                         m, lod, bg, flags, lt, base, st, mk, order,
                                                                         cost);
      return;
    }
  }
  RenderPlainKernel<WV, L, C, SP, true><<<g, dim3(64 * WV), 0, s>>>(out, codes, W, H, pool, root, nrm, col, nib, m,
                                                                    lod, bg, flags, lt, base, st, mk, order, cost);
}

// spill = false selects the window-only stack (the tree fits the LDS window).
template <bool C>
bool DispatchPlain(int waves, int levels, bool spill, hipStream_t s, uint32_t* out, int* codes, int W, int H,
                   const PoolView& pool, uint32_t root, const uint32_t* nrm, const uint32_t* col, int nib,
                   const Mat16& m, float lod, const Vec4& bg, int flags, const Lights& lt, int base,
                   unsigned long long* st, const rayt::MarkTargets& mk, const uint32_t* order = nullptr,
                   uint32_t* cost = nullptr) {
  const int ts = waves == 4 ? 16 : 8;
  const dim3 g(uint32_t((W + ts - 1) / ts) * uint32_t((H + ts - 1) / ts));
#define RT_CASE(WV, L)                                                                                      \
  if (waves == WV && levels == L) {                                                                         \
    if (spill || L >= kStackLevels)                                                                         \
      LaunchPlainL<WV, L, C, true>(g, s, out, codes, W, H, pool, root, nrm, col, nib, m, lod, bg, flags, lt, \
                                   base, st, mk, order, cost);                                              \
    else                                                                                                    \
      LaunchPlainL<WV, L, C, L >= kStackLevels>(g, s, out, codes, W, H, pool, root, nrm, col, nib, m, lod,  \
                                                bg, flags, lt, base, st, mk, order, cost);                  \
    return true;                                                                                            \
  }
  RT_CASE(1, 8) RT_CASE(1, 10) RT_CASE(1, 12) RT_CASE(1, 14) RT_CASE(1, 16) RT_CASE(1, 22)
  RT_CASE(4, 8) RT_CASE(4, 10) RT_CASE(4, 12) RT_CASE(4, 14) RT_CASE(4, 16) RT_CASE(4, 22)
#undef RT_CASE
  return false;
}

}  // namespace

extern "C" {

int rt_init_tracing_frame(rt_ctx* ctx, rt_buf* states, int width, int height, const float vpi[16],
                          float lod_voxel_size) {
  int rc = CheckTraceArgs(ctx, width, height);
  if (rc) return rc;
  RT_REQUIRE(states && vpi, RT_ERR_INVALID, "null argument");
  const uint64_t P = uint64_t(width) * height;
  RT_REQUIRE(states->size >= P * RT_STATE_BYTES_PER_PIXEL, RT_ERR_RANGE, "state buffer too small");
  Mat16 m;
  memcpy(m.m, vpi, sizeof(m.m));
  char* base = static_cast<char*>(states->ptr);
  InitTracingFrameKernel<<<dim3(TileGrid(width, height)), dim3(256), 0, ctx->stream>>>(
      reinterpret_cast<float4*>(base), reinterpret_cast<float4*>(base + 16 * P), reinterpret_cast<int4*>(base + 32 * P),
      width, height, m, lod_voxel_size);
  RT_HIP_CHECK(hipGetLastError());
  return RT_OK;
}

int rt_raytracing_pass(rt_ctx* ctx, rt_buf* faults_and_hits, rt_buf* states, int width, int height,
                       const rt_buf* node_links, const rt_buf* far_pointers, int root_node_index,
                       const rt_buf* normals, const rt_buf* colors, int nodes_in_block, int flags) {
  int rc = CheckTraceArgs(ctx, width, height);
  if (rc) return rc;
  RT_REQUIRE(faults_and_hits && states && node_links && far_pointers, RT_ERR_INVALID, "null argument");
  const uint64_t P = uint64_t(width) * height;
  RT_REQUIRE(states->size >= P * RT_STATE_BYTES_PER_PIXEL && faults_and_hits->size >= P * 4, RT_ERR_RANGE,
             "buffer too small");
  RT_REQUIRE(root_node_index >= 0 && uint64_t(root_node_index) * 4 < node_links->size, RT_ERR_RANGE,
             "root node outside pool");
  RT_REQUIRE(nodes_in_block >= 0, RT_ERR_INVALID, "negative nodes_in_block");
  rc = CheckChannels(node_links, normals, colors);
  if (rc) return rc;
  char* base = static_cast<char*>(states->ptr);
  const Lights lt = rayt::MakeLights();
  const bool ok = DispatchPass(ctx->plain_lds_levels, ctx->plain_spill, ctx->plain_waves == 1 ? 1 : 4, ctx->stream,
                               static_cast<int*>(faults_and_hits->ptr), reinterpret_cast<float4*>(base),
                               reinterpret_cast<float4*>(base + 16 * P), reinterpret_cast<int4*>(base + 32 * P),
                               width, height, View(node_links, far_pointers), uint32_t(root_node_index),
                               U32(normals), U32(colors), nodes_in_block, flags, lt, ctx->plain_lds_base);
  RT_REQUIRE(ok, RT_ERR_STATE, "no pass kernel for this stack window");
  RT_HIP_CHECK(hipGetLastError());
  return RT_OK;
}

int rt_finish_tracing_frame(rt_ctx* ctx, rt_buf* out_image, const rt_buf* states, int width, int height,
                            const float background[4], const rt_buf* normals, const rt_buf* colors) {
  int rc = CheckTraceArgs(ctx, width, height);
  if (rc) return rc;
  RT_REQUIRE(out_image && states && background, RT_ERR_INVALID, "null argument");
  const uint64_t P = uint64_t(width) * height;
  RT_REQUIRE(states->size >= P * RT_STATE_BYTES_PER_PIXEL && out_image->size >= P * 4, RT_ERR_RANGE,
             "buffer too small");
  Vec4 bg;
  memcpy(bg.v, background, sizeof(bg.v));
  const char* base = static_cast<const char*>(states->ptr);
  FinishTracingFrameKernel<<<dim3(TileGrid(width, height)), dim3(256), 0, ctx->stream>>>(
      static_cast<uint32_t*>(out_image->ptr), reinterpret_cast<const int4*>(base + 32 * P), width, height, bg,
      U32(normals), U32(colors), rayt::MakeLights(), ChannelNodes(normals, colors));
  RT_HIP_CHECK(hipGetLastError());
  return RT_OK;
}

int rt_raytrace_depth(rt_ctx* ctx, rt_buf* out_image, int width, int height, const rt_buf* node_links,
                      const rt_buf* far_pointers, int root_node_index, const float vpi[16], float lod_voxel_size) {
  int rc = CheckTraceArgs(ctx, width, height);
  if (rc) return rc;
  RT_REQUIRE(out_image && node_links && far_pointers && vpi, RT_ERR_INVALID, "null argument");
  RT_REQUIRE(out_image->size >= uint64_t(width) * height * 4, RT_ERR_RANGE, "image buffer too small");
  RT_REQUIRE(root_node_index >= 0 && uint64_t(root_node_index) * 4 < node_links->size, RT_ERR_RANGE,
             "root node outside pool");
  Mat16 m;
  memcpy(m.m, vpi, sizeof(m.m));
  RaytraceDepthKernel<<<dim3(TileGrid(width, height)), dim3(256), 0, ctx->stream>>>(
      static_cast<uint32_t*>(out_image->ptr), width, height, View(node_links, far_pointers),
      uint32_t(root_node_index), m, lod_voxel_size);
  RT_HIP_CHECK(hipGetLastError());
  return RT_OK;
}

int rt_raytrace_restart(rt_ctx* ctx, rt_buf* out_image, int width, int height, const rt_buf* legacy_links,
                        int root_node_index, const float vpi[16], float lod_voxel_size) {
  int rc = CheckTraceArgs(ctx, width, height);
  if (rc) return rc;
  RT_REQUIRE(out_image && legacy_links && vpi, RT_ERR_INVALID, "null argument");
  RT_REQUIRE(out_image->size >= uint64_t(width) * height * 4, RT_ERR_RANGE, "image buffer too small");
  RT_REQUIRE(root_node_index >= 0 && uint64_t(root_node_index) * 4 < legacy_links->size, RT_ERR_RANGE,
             "root node outside pool");
  Mat16 m;
  memcpy(m.m, vpi, sizeof(m.m));
  RaytraceRestartKernel<<<dim3(TileGrid(width, height)), dim3(256), 0, ctx->stream>>>(
      static_cast<uint32_t*>(out_image->ptr), width, height, U32(legacy_links),
      uint32_t(std::min<uint64_t>(legacy_links->size / 4, 0xffffffffu)), uint32_t(root_node_index), m,
      lod_voxel_size);
  RT_HIP_CHECK(hipGetLastError());
  return RT_OK;
}

int rt_legacy_links_from_cache(rt_ctx* ctx, rt_buf* legacy_links, const rt_buf* node_links,
                               const rt_buf* far_pointers, uint64_t node_count) {
  RT_REQUIRE(ctx && legacy_links && node_links && far_pointers, RT_ERR_INVALID, "null argument");
  RT_REQUIRE(node_count > 0 && node_count <= legacy_links->size / 4 && node_count <= node_links->size / 4,
             RT_ERR_RANGE, "node count exceeds buffers");
  int* d_over = nullptr;
  RT_HIP_CHECK(hipMalloc(&d_over, sizeof(int)));
  if (const hipError_t e0 = hipMemsetAsync(d_over, 0, sizeof(int), ctx->stream); e0 != hipSuccess) {
    (void)hipFree(d_over);
    return rayt::Fail(RT_ERR_HIP, hipGetErrorString(e0));
  }
  LegacyLinksKernel<<<dim3(unsigned((node_count + 255) / 256)), dim3(256), 0, ctx->stream>>>(
      static_cast<uint32_t*>(legacy_links->ptr), U32(node_links), U32(far_pointers), far_pointers->size / 4,
      node_count, d_over);
  int over = 0;
  hipError_t e = hipMemcpyAsync(&over, d_over, sizeof(int), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  (void)hipFree(d_over);
  if (e != hipSuccess) return rayt::Fail(RT_ERR_HIP, hipGetErrorString(e));
  RT_REQUIRE(!over, RT_ERR_RANGE,
             "pool needs more than 23-bit node indices for the legacy layout, or a far link lies outside the "
             "far-pointer table");
  return RT_OK;
}

int rt_plain_config(rt_ctx* ctx, int lds_stack_levels, int wg_waves, int tree_depth) {
  RT_REQUIRE(ctx, RT_ERR_INVALID, "null context");
  RT_REQUIRE(lds_stack_levels == 8 || lds_stack_levels == 10 || lds_stack_levels == 12 || lds_stack_levels == 14 ||
                 lds_stack_levels == 16 || lds_stack_levels == 22,
             RT_ERR_RANGE, "lds_stack_levels must be one of 8,10,12,14,16,22");
  RT_REQUIRE(wg_waves == 1 || wg_waves == 4, RT_ERR_RANGE, "wg_waves must be 1 or 4");
  RT_REQUIRE(tree_depth >= 0 && tree_depth <= 22, RT_ERR_RANGE, "tree_depth must be in [0, 22]");
  ctx->plain_lds_levels = lds_stack_levels;
  ctx->plain_waves = wg_waves;
  // A depth-D tree pushes stack levels 0 .. D-2; keep the deepest ones in LDS.
  ctx->plain_lds_base = tree_depth > 0 ? std::max(0, (tree_depth - 1) - lds_stack_levels) : 0;
  // A declared depth that fits the window selects the window-only stack.
  ctx->plain_spill = !(tree_depth > 0 && tree_depth - 1 <= lds_stack_levels);
  return RT_OK;
}

int rt_render_plain(rt_ctx* ctx, rt_buf* out_image, rt_buf* codes, int width, int height, const rt_buf* node_links,
                    const rt_buf* far_pointers, int root_node_index, const rt_buf* normals, const rt_buf* colors,
                    int nodes_in_block, const float vpi[16], float lod_voxel_size, const float background[4],
                    int flags, uint64_t* stats) {
  return rayt::RenderPlain(ctx, out_image, codes, width, height, node_links, far_pointers, root_node_index, normals,
                           colors, nodes_in_block, vpi, lod_voxel_size, background, flags, stats, nullptr);
}

}  // extern "C"

int rayt::RenderPlain(rt_ctx* ctx, rt_buf* out_image, rt_buf* codes, int width, int height,
                      const rt_buf* node_links, const rt_buf* far_pointers, int root_node_index,
                      const rt_buf* normals, const rt_buf* colors, int nodes_in_block, const float vpi[16],
                      float lod_voxel_size, const float background[4], int flags, uint64_t* stats,
                      const MarkTargets* mark, const TileSchedule* sched) {
  int rc = CheckTraceArgs(ctx, width, height);
  if (rc) return rc;
  RT_REQUIRE(out_image && node_links && far_pointers && vpi && background, RT_ERR_INVALID, "null argument");
  const uint64_t P = uint64_t(width) * height;
  RT_REQUIRE(out_image->size >= P * 4 && (!codes || codes->size >= P * 4), RT_ERR_RANGE, "buffer too small");
  RT_REQUIRE(root_node_index >= 0 && uint64_t(root_node_index) * 4 < node_links->size, RT_ERR_RANGE,
             "root node outside pool");
  RT_REQUIRE(nodes_in_block >= 0, RT_ERR_INVALID, "negative nodes_in_block");
  rc = CheckChannels(node_links, normals, colors);
  if (rc) return rc;
  RT_REQUIRE(!mark || (codes && nodes_in_block > 0 && !(flags & RT_FLAG_COUNT)), RT_ERR_INVALID,
             "marking needs block-level codes and the timing variant");
  RT_REQUIRE(!sched || (sched->order && sched->cost && !(flags & RT_FLAG_COUNT)), RT_ERR_INVALID,
             "a tile schedule needs its buffers and the timing variant");
  const MarkTargets no_marks;
  const MarkTargets& mk = mark ? *mark : no_marks;
  Mat16 m;
  memcpy(m.m, vpi, sizeof(m.m));
  Vec4 bg;
  memcpy(bg.v, background, sizeof(bg.v));
  const Lights lt = rayt::MakeLights();
  if (flags & RT_FLAG_COUNT) {
    RT_REQUIRE(stats, RT_ERR_INVALID, "counting render needs a stats array");
    unsigned long long* d_stats = nullptr;
    constexpr size_t kStatBytes = 4 * kStatSlots * sizeof(unsigned long long);
    RT_HIP_CHECK(hipMalloc(&d_stats, kStatBytes));
    if (const hipError_t e0 = hipMemsetAsync(d_stats, 0, kStatBytes, ctx->stream); e0 != hipSuccess) {
      (void)hipFree(d_stats);
      return rayt::Fail(RT_ERR_HIP, hipGetErrorString(e0));
    }
    DispatchPlain<true>(ctx->plain_waves, ctx->plain_lds_levels, ctx->plain_spill, ctx->stream,
                        static_cast<uint32_t*>(out_image->ptr),
                        codes ? static_cast<int*>(codes->ptr) : nullptr, width, height, View(node_links, far_pointers),
                        uint32_t(root_node_index), U32(normals), U32(colors), nodes_in_block, m,
                        lod_voxel_size, bg, flags, lt, ctx->plain_lds_base, d_stats, mk);
    hipError_t e = hipGetLastError();
    unsigned long long h[4 * kStatSlots] = {};
    if (e == hipSuccess) e = hipMemcpyAsync(h, d_stats, kStatBytes, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(d_stats);
    if (e != hipSuccess) return rayt::Fail(RT_ERR_HIP, hipGetErrorString(e));
    for (int k = 0; k < 4; ++k) {
      stats[k] = 0;
      for (int j = 0; j < kStatSlots; ++j) stats[k] += h[4 * j + k];
    }
    return RT_OK;
  }
  DispatchPlain<false>(ctx->plain_waves, ctx->plain_lds_levels, ctx->plain_spill, ctx->stream,
                       static_cast<uint32_t*>(out_image->ptr),
                       codes ? static_cast<int*>(codes->ptr) : nullptr, width, height, View(node_links, far_pointers),
                       uint32_t(root_node_index), U32(normals), U32(colors), nodes_in_block, m,
                       lod_voxel_size, bg, flags, lt, ctx->plain_lds_base, nullptr, mk,
                       sched && sched->order_valid ? sched->order : nullptr, sched ? sched->cost : nullptr);
  RT_HIP_CHECK(hipGetLastError());
  return RT_OK;
}

int rayt::OrderTiles(rt_ctx* ctx, const TileSchedule& sched, int width, int height) {
  uint32_t tiles;
  int waves;
  PlainTiles(ctx, width, height, &tiles, &waves);
  RT_REQUIRE(tiles <= uint32_t(kOrderMaxTiles), RT_ERR_RANGE, "too many tiles to schedule");
  if (waves == 4)
    TileOrderKernel<4><<<dim3(1), dim3(64 * kOrderWaves), 0, ctx->stream>>>(sched.cost, int(tiles), sched.order);
  else
    TileOrderKernel<1><<<dim3(1), dim3(64 * kOrderWaves), 0, ctx->stream>>>(sched.cost, int(tiles), sched.order);
  RT_HIP_CHECK(hipGetLastError());
  return RT_OK;
}

void rayt::PlainTiles(const rt_ctx* ctx, int width, int height, uint32_t* tiles, int* waves_per_tile) {
  const int waves = ctx->plain_waves == 4 ? 4 : 1;
  const int ts = waves == 4 ? 16 : 8;
  *tiles = uint32_t((width + ts - 1) / ts) * uint32_t((height + ts - 1) / ts);
  *waves_per_tile = waves;
}
