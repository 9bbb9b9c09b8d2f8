// Batched rendering: many views of any resolution, over one or several
// resident scenes, traced by one launch (no reference counterpart; per view
// the result is the plain single-pass frame of trace.hip, bit for bit).
//
// Two kernels over the same tiles. Every view is cut into wave tiles of 64
// pixels (8x8, 16x4 or 32x2) and the tiles of all views form one sequence.
// BatchTileKernel (wgs_per_cu == 0) runs one wave per tile and leaves the
// balancing to the hardware dispatcher. BatchKernel (wgs_per_cu > 0) is the
// persistent form: a grid of one-wave workgroups, at most wgs_per_cu per
// compute unit, takes the tiles from a shared queue. The queue is eight
// queues, one per XCD (workgroup b runs on XCD b % 8): queue q holds tiles
// q, q + 8, q + 16, ..., the tiles the hardware's round-robin dispatch gives
// that XCD, so neighbouring tiles are traced at the same time on all XCDs
// (contiguous eighths of the sequence ran 7 % slower). A wave claims a group
// of its XCD's tiles with one atomic, steals from the next XCDs' queues when
// its own is exhausted, and exits once all eight are, so the grid drains. Both kernels render a tile the same way (RenderTile): the
// tile's view and scene are wave-uniform (scalar loads) and each lane renders
// its pixel as the plain kernel does (RenderPixel: the primary ray, then the
// shadow ray). The traversal stack keeps the deepest `lds_levels` levels in
// LDS.
#include <algorithm>
#include <cstring>
#include <vector>

#include "batch.h"
#include "cache.h"
#include "render_pixel.h"
#include "trace_common.h"

using namespace rtdev;

namespace {

// Tiles claimed by one atomic: up to this many, fewer when a launch has
// too few tiles per wave for claims of this size to balance the load.
constexpr uint32_t kMaxTilesPerGrab = 2;
// The eight queue counters sit 256 bytes apart, each in a memory line of
// its own, so that the atomics on different queues do not queue up on one
// line.
constexpr uint32_t kCounterStride = 64;

struct DevScene {
  const uint32_t* links;
  const uint32_t* far_ptrs;
  const uint32_t* normals;
  const uint32_t* colors;
  uint32_t nodes, far_count, root;
  int nodes_in_block;
  int lds_base;  // first traversal-stack level kept in LDS for this scene's depth
  int pad;
};

struct DevView {
  float m[16];
  int width, height;
  float lod;
  int flags;
  int tiles_x;
  int scene;
  uint64_t pixel_offset;
};

struct Vec4B {
  float v[4];
};

// Layout of a batch's descriptor block (see rt_batch).
constexpr uint64_t kStatsOffset = sizeof(DevScene) * rayt::kMaxBatchScenes;
constexpr uint64_t kTilePrefixOffset = kStatsOffset + 8ull * 4 * kStatSlots;
static_assert(kStatsOffset % 16 == 0 && kTilePrefixOffset % 16 == 0, "descriptor block alignment");
inline uint64_t ViewsOffset(int nviews) { return (kTilePrefixOffset + 4ull * (nviews + 1) + 15) & ~uint64_t(15); }
inline uint64_t DescBytes(int nviews) { return ViewsOffset(nviews) + sizeof(DevView) * uint64_t(nviews); }

// What a launch counts: nothing (the timing variant), rays, hits and node
// fetches, or rays and hits only (no per-step work: the server's rounds).
constexpr int kNoCount = 0, kCountAll = 1, kCountRays = 2;

// One wave tile of the sequence: the view holding it (a binary search over
// the tile prefix, on wave-uniform values, so scalar loads), then one pixel per
// lane through RenderPixel.
template <int TW, int LDS_LEVELS, int COUNT, bool SPILL, bool LOD>
__device__ __forceinline__ void RenderTile(uint32_t tile, uint32_t lane, const DevView* __restrict__ views,
                                           int nviews, const DevScene* __restrict__ scenes,
                                           const uint32_t* __restrict__ tile_begin, uint32_t* __restrict__ out,
                                           int* __restrict__ codes, const Vec4B& bg, const Lights& lt, uint2* lds,
                                           PixelCounters& cnt) {
  constexpr int TH = 64 / TW;
  int lo = 0, hi = nviews - 1;  // view holding the tile: last v with tile_begin[v] <= tile
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tile_begin[mid] <= tile) lo = mid; else hi = mid - 1;
  }
  const DevView& v = views[lo];
  const DevScene& sc = scenes[v.scene];
  const uint32_t local = tile - tile_begin[lo];
  const int x = int(local % uint32_t(v.tiles_x)) * TW + int(lane % TW);
  const int y = int(local / uint32_t(v.tiles_x)) * TH + int(lane / TW);
  if (x < v.width && y < v.height) {
    PoolView pool;
    pool.links = sc.links;
    pool.far_ptrs = sc.far_ptrs;
    pool.nodes = sc.nodes;
    pool.far_count = sc.far_count;
    StackSpill<SPILL ? LDS_LEVELS : kStackLevels> spill;
    SplitStack<LDS_LEVELS, 64, SPILL> stack;
    stack.spill = spill.v;
    stack.lds = lds + lane;
    stack.set_base(sc.lds_base);
    uint32_t color;
    int code;
    RenderPixel<SplitStack<LDS_LEVELS, 64, SPILL>, COUNT != kNoCount, LOD, COUNT == kCountAll>(
        pool, sc.root, sc.normals, sc.colors, sc.nodes_in_block, v.m, x, y, v.width, v.height, v.lod, bg.v, v.flags,
        lt, stack, color, code, cnt);
    const uint64_t pix = v.pixel_offset + uint64_t(y) * uint32_t(v.width) + uint32_t(x);
    out[pix] = color;
    if (codes) codes[pix] = code;
  }
}

// Persistent waves over the shared tile queue (wgs_per_cu > 0): `grab`
// tiles per claim, from this XCD's queue first, then from the next XCDs'
// queues. Built for eight waves per SIMD: the scalar registers of the tile
// loop would otherwise hold the kernel to seven (3.12 against 3.32 ms on the
// 64-view mix).
// LOD = false: every view of the launch has lod == 0 (no LOD stop test, which
// shadow rays never need either).
template <int TW, int LDS_LEVELS, int COUNT, bool SPILL, bool LOD>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(8))) void BatchKernel(
    const DevView* __restrict__ views, int nviews, const DevScene* __restrict__ scenes,
    const uint32_t* __restrict__ tile_begin, uint32_t total_tiles, uint32_t grab, uint32_t* __restrict__ counters,
    uint32_t* __restrict__ out, int* __restrict__ codes, Vec4B bg, Lights lt, unsigned long long* __restrict__ stats) {
  __shared__ uint2 lds[LDS_LEVELS * 64];
  const uint32_t lane = threadIdx.x;
  const uint32_t xcd = blockIdx.x & 7u;
  PixelCounters cnt;
  uint32_t q_rel = 0;
  while (q_rel < 8u) {
    const uint32_t q = (xcd + q_rel) & 7u;
    const uint32_t nq = total_tiles > q ? (total_tiles - q + 7u) / 8u : 0u;  // tiles q, q + 8, q + 16, ...
    uint32_t t = 0;
    if (lane == 0) t = atomicAdd(&counters[q * kCounterStride], grab);
    t = __builtin_amdgcn_readfirstlane(__shfl(t, 0, 64));
    if (t >= nq) {  // this queue is exhausted: steal from the next one
      ++q_rel;
      continue;
    }
    const uint32_t last = min(nq, t + grab);
    for (uint32_t k = t; k < last; ++k)
      RenderTile<TW, LDS_LEVELS, COUNT, SPILL, LOD>(q + 8u * k, lane, views, nviews, scenes, tile_begin, out, codes,
                                                    bg, lt, lds, cnt);
  }
  if (COUNT != kNoCount) FlushCounters(cnt, stats);
}

// One wave per tile, dispatched by the hardware (wgs_per_cu == 0): workgroup
// g renders tile g and exits; there is no queue and no tail of persistent
// waves.
template <int TW, int LDS_LEVELS, int COUNT, bool SPILL, bool LOD>
__global__ __launch_bounds__(64) void BatchTileKernel(const DevView* __restrict__ views, int nviews,
                                                      const DevScene* __restrict__ scenes,
                                                      const uint32_t* __restrict__ tile_begin,
                                                      uint32_t* __restrict__ out, int* __restrict__ codes, Vec4B bg,
                                                      Lights lt, unsigned long long* __restrict__ stats) {
  __shared__ uint2 lds[LDS_LEVELS * 64];
  PixelCounters cnt;
  RenderTile<TW, LDS_LEVELS, COUNT, SPILL, LOD>(blockIdx.x, threadIdx.x, views, nviews, scenes, tile_begin, out,
                                                codes, bg, lt, lds, cnt);
  if (COUNT != kNoCount) FlushCounters(cnt, stats);
}

__global__ void ZeroCountersKernel(uint32_t* c, int n) {
  for (int i = int(threadIdx.x); i < n; i += int(blockDim.x)) c[i] = 0u;
}

struct LaunchArgs {
  dim3 grid;
  hipStream_t stream;
  const DevView* views;
  int nviews;
  const DevScene* scenes;
  const uint32_t* tile_begin;
  uint32_t total_tiles;
  uint32_t grab;
  uint32_t* counters;
  uint32_t* out;
  int* codes;
  Vec4B bg;
  Lights lt;
  unsigned long long* stats;
  bool lod;        // some view has a nonzero LOD voxel size
  bool per_tile;   // one wave per tile (BatchTileKernel), grid = total_tiles
};

// The timing and ray-counting variants come with and without the LOD stop; the variant that counts fetches keeps it.
template <int TW, int L, int C, bool SP, bool LD>
void LaunchTiles(const LaunchArgs& a) {
  BatchTileKernel<TW, L, C, SP, LD><<<dim3(a.total_tiles), dim3(64), 0, a.stream>>>(
      a.views, a.nviews, a.scenes, a.tile_begin, a.out, a.codes, a.bg, a.lt, a.stats);
}

template <int TW, int L, int C, bool SP>
void Launch(const LaunchArgs& a) {
  if (a.per_tile) {
    if constexpr (C != kCountAll) {
      if (!a.lod) return LaunchTiles<TW, L, C, SP, false>(a);
    }
    return LaunchTiles<TW, L, C, SP, true>(a);
  }
  if constexpr (C != kCountAll) {
    if (!a.lod) {
      BatchKernel<TW, L, C, SP, false><<<a.grid, dim3(64), 0, a.stream>>>(
          a.views, a.nviews, a.scenes, a.tile_begin, a.total_tiles, a.grab, a.counters, a.out, a.codes, a.bg, a.lt,
          a.stats);
      return;
    }
  }
  BatchKernel<TW, L, C, SP, true><<<a.grid, dim3(64), 0, a.stream>>>(a.views, a.nviews, a.scenes, a.tile_begin,
                                                                     a.total_tiles, a.grab, a.counters, a.out,
                                                                     a.codes, a.bg, a.lt, a.stats);
}

// spill = false selects the window-only stack (every scene fits the LDS window).
template <int C>
bool Dispatch(int tw, int levels, bool spill, const LaunchArgs& a) {
#define RT_B(TW, L)                   \
  if (tw == TW && levels == L) {      \
    if (spill)                        \
      Launch<TW, L, C, true>(a);      \
    else                              \
      Launch<TW, L, C, false>(a);     \
    return true;                      \
  }
  RT_B(8, 8) RT_B(8, 10) RT_B(8, 12) RT_B(8, 16)
  RT_B(16, 8) RT_B(16, 10) RT_B(16, 12) RT_B(16, 16)
  RT_B(32, 8) RT_B(32, 10) RT_B(32, 12) RT_B(32, 16)
#undef RT_B
  return false;
}

int TileWidth(int shape) { return shape == 1 ? 16 : shape == 2 ? 32 : 8; }

DevScene SceneOf(rt_cache* c, int tree_depth, int lds_levels) {
  rayt::BlockCacheManager& m = *c->mgr;
  rt_buf* links = rt_cache_channel(c, 0);
  rt_buf* far_ptrs = rt_cache_far_pointers(c);
  rt_buf* normals = rt_cache_channel_by_name(c, "normals");
  rt_buf* colors = rt_cache_channel_by_name(c, "colors");
  DevScene s;
  s.links = static_cast<const uint32_t*>(links->ptr);
  s.far_ptrs = static_cast<const uint32_t*>(far_ptrs->ptr);
  s.normals = normals ? static_cast<const uint32_t*>(normals->ptr) : nullptr;
  s.colors = colors ? static_cast<const uint32_t*>(colors->ptr) : nullptr;
  s.nodes = uint32_t(std::min<uint64_t>(links->size / 4, 0xffffffffu));
  s.far_count = uint32_t(std::min<uint64_t>(far_ptrs->size / 4, 0xffffffffu));
  s.root = uint32_t(m.root_node_index());
  s.nodes_in_block = m.nodes_in_block();
  // Deepest stack levels in LDS: a depth-D tree pushes levels 0 .. D-2.
  s.lds_base = std::max(0, (tree_depth - 1) - lds_levels);
  s.pad = 0;
  return s;
}

}  // namespace

struct rt_batch {
  rt_ctx* ctx = nullptr;
  hipStream_t stream = nullptr;  // the context stream unless the server assigns one
  int max_views = 0;
  int tile_shape = 0;   // 0: 8x8, 1: 16x4, 2: 32x2
  int lds_levels = 12;  // 8, 10, 12 or 16
  int wgs_per_cu = 0;   // 0: one wave per tile; else persistent one-wave workgroups per compute unit
  int tree_depth = 0;   // depth assumed for scenes registered without one (0: unknown)
  std::vector<rt_cache*> scenes;
  std::vector<int> depths;  // per scene; 0 = tree_depth
  // The launch's descriptors in one block, copied to the device in one transfer:
  // [scenes (kMaxBatchScenes) | zeroed statistics (kStatSlots x 4) | tile prefix (nviews + 1) | views (nviews)];
  // the pinned host copy is rewritten only after the previous launch has completed.
  uint8_t* d_block = nullptr;
  uint8_t* h_block = nullptr;
  uint32_t* d_counters = nullptr;
  unsigned long long* h_stats = nullptr;
  bool timing = true;  // record the start event rt_batch_last_ms reads
  bool rays_only = false;  // counting launches leave the node fetches out
  bool last_timed = false;
  bool launched = false;
  bool counted = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

namespace rayt {

int BatchCreate(rt_ctx* ctx, int max_views, rt_batch** out) {
  RT_REQUIRE(ctx && out, RT_ERR_INVALID, "null argument");
  RT_REQUIRE(max_views > 0 && max_views <= (1 << 16), RT_ERR_RANGE, "max_views must be in [1, 65536]");
  rt_batch* b = new rt_batch();
  b->ctx = CtxRetain(ctx);
  b->stream = ctx->stream;
  b->max_views = max_views;
  const uint64_t block = DescBytes(max_views);
  hipError_t e = hipMalloc(reinterpret_cast<void**>(&b->d_block), block);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&b->d_counters), 4 * 8 * kCounterStride);
  if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&b->h_block), block, 0);
  if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&b->h_stats), 8 * 4 * kStatSlots, 0);
  if (e == hipSuccess) memset(b->h_block, 0, block);  // the counters' area stays zero
  if (e == hipSuccess) e = hipEventCreate(&b->ev0);
  if (e == hipSuccess) e = hipEventCreate(&b->ev1);
  if (e != hipSuccess) {
    rt_batch_destroy(b);
    return Fail(RT_ERR_HIP, std::string("batch allocation: ") + hipGetErrorString(e));
  }
  *out = b;
  return RT_OK;
}

void BatchSetStream(rt_batch* b, hipStream_t s) { b->stream = s ? s : b->ctx->stream; }

void BatchSetTiming(rt_batch* b, bool timing) { b->timing = timing; }

void BatchSetRayCounts(rt_batch* b, bool rays_only) { b->rays_only = rays_only; }

hipEvent_t BatchDoneEvent(rt_batch* b) { return b->ev1; }

int BatchSetScenes(rt_batch* b, const std::vector<rt_cache*>& caches, const std::vector<int>& depths) {
  RT_REQUIRE(caches.size() == depths.size() && caches.size() <= size_t(kMaxBatchScenes), RT_ERR_RANGE,
             "bad scene list");
  b->scenes = caches;
  b->depths = depths;
  return RT_OK;
}

int BatchLaunch(rt_batch* b, const rt_view* views, int nviews, rt_buf* out_image, rt_buf* codes, bool count) {
  RT_REQUIRE(b && views && out_image, RT_ERR_INVALID, "null argument");
  RT_REQUIRE(nviews > 0 && nviews <= b->max_views, RT_ERR_RANGE, "nviews out of range");
  RT_REQUIRE(!b->scenes.empty(), RT_ERR_STATE, "batch has no scene");
  const int tw = TileWidth(b->tile_shape), th = 64 / tw;
  // The staging below is read by the previous launch's copies: wait for it.
  if (b->launched) RT_HIP_CHECK(hipEventSynchronize(b->ev1));
  uint64_t tiles = 0;
  bool any_lod = false;
  for (int i = 0; i < nviews; ++i) {
    const rt_view& v = views[i];
    any_lod = any_lod || v.lod_voxel_size != 0.0f;
    RT_REQUIRE(v.width > 0 && v.height > 0 && int64_t(v.width) * v.height <= (int64_t(1) << 28), RT_ERR_RANGE,
               "bad view size");
    RT_REQUIRE(v.scene >= 0 && v.scene < int(b->scenes.size()), RT_ERR_RANGE, "view scene out of range");
    const uint64_t px = uint64_t(v.width) * v.height;
    // in pixels, so that no sum or product can wrap
    const uint64_t out_px = out_image->size / 4, codes_px = codes ? codes->size / 4 : out_px;
    RT_REQUIRE(v.pixel_offset <= out_px && px <= out_px - v.pixel_offset && v.pixel_offset <= codes_px &&
                   px <= codes_px - v.pixel_offset,
               RT_ERR_RANGE, "view outside the output buffers");
    DevView& d = reinterpret_cast<DevView*>(b->h_block + ViewsOffset(nviews))[i];
    memcpy(d.m, v.view_proj_inv, sizeof(d.m));
    d.width = v.width;
    d.height = v.height;
    d.lod = v.lod_voxel_size;
    d.flags = v.flags & RT_FLAG_SHADOWS;
    d.tiles_x = (v.width + tw - 1) / tw;
    d.scene = v.scene;
    d.pixel_offset = v.pixel_offset;
    reinterpret_cast<uint32_t*>(b->h_block + kTilePrefixOffset)[i] = uint32_t(tiles);
    tiles += uint64_t(d.tiles_x) * ((v.height + th - 1) / th);
    RT_REQUIRE(tiles < (1ull << 31), RT_ERR_RANGE, "too many tiles in one batch");
  }
  reinterpret_cast<uint32_t*>(b->h_block + kTilePrefixOffset)[nviews] = uint32_t(tiles);
  // The window-only stack needs every scene's declared depth to fit the LDS window.
  bool spill = false;
  for (size_t i = 0; i < b->scenes.size(); ++i) {
    rt_cache* c = b->scenes[i];
    RT_REQUIRE(c->ctx == b->ctx, RT_ERR_INVALID, "scene belongs to another context");
    RT_REQUIRE(c->mgr->root_node_index() >= 0, RT_ERR_STATE, "scene has no root block in its pool");
    const int depth = b->depths[i] > 0 ? b->depths[i] : b->tree_depth;
    spill = spill || depth <= 0 || depth - 1 > b->lds_levels;
    reinterpret_cast<DevScene*>(b->h_block)[i] = SceneOf(c, depth, b->lds_levels);
  }
  hipStream_t s = b->stream;
  // one transfer: the scenes, the zeroed statistics, the tile prefix and the views
  RT_HIP_CHECK(hipMemcpyAsync(b->d_block, b->h_block, DescBytes(nviews), hipMemcpyHostToDevice, s));
  if (b->wgs_per_cu != 0) ZeroCountersKernel<<<dim3(1), dim3(64), 0, s>>>(b->d_counters, 8 * kCounterStride);
  LaunchArgs a;
  a.stream = s;
  a.views = reinterpret_cast<const DevView*>(b->d_block + ViewsOffset(nviews));
  a.nviews = nviews;
  a.scenes = reinterpret_cast<const DevScene*>(b->d_block);
  a.tile_begin = reinterpret_cast<const uint32_t*>(b->d_block + kTilePrefixOffset);
  a.total_tiles = uint32_t(tiles);
  a.counters = b->d_counters;
  a.out = static_cast<uint32_t*>(out_image->ptr);
  a.codes = codes ? static_cast<int*>(codes->ptr) : nullptr;
  const float kBg[4] = {0.0f, float(216 / 255.), 1.0f, 1.0f};
  memcpy(a.bg.v, kBg, sizeof(a.bg.v));
  a.lt = MakeLights();
  a.per_tile = b->wgs_per_cu == 0;
  // Persistent: enough single-wave workgroups for the work, at most wgs_per_cu per CU, a multiple of 8 (one range
  // per XCD); claims of several tiles only when every wave gets at least 16 of them.
  const uint64_t want = std::min<uint64_t>(uint64_t(std::max(1, b->ctx->compute_units)) * std::max(1, b->wgs_per_cu),
                                           tiles);
  const uint64_t waves = (std::max<uint64_t>(want, 1) + 7) & ~uint64_t(7);
  a.grid = dim3(uint32_t(waves));
  a.grab = uint32_t(std::clamp<uint64_t>(tiles / (waves * 16), 1, kMaxTilesPerGrab));
  a.stats = count ? reinterpret_cast<unsigned long long*>(b->d_block + kStatsOffset) : nullptr;
  a.lod = any_lod;
  b->last_timed = b->timing;
  if (b->timing) RT_HIP_CHECK(hipEventRecord(b->ev0, s));
  const bool ok = !count        ? Dispatch<kNoCount>(tw, b->lds_levels, spill, a)
                  : b->rays_only ? Dispatch<kCountRays>(tw, b->lds_levels, spill, a)
                                 : Dispatch<kCountAll>(tw, b->lds_levels, spill, a);
  RT_REQUIRE(ok, RT_ERR_STATE, "no kernel for this batch configuration");
  RT_HIP_CHECK(hipGetLastError());
  if (count)
    RT_HIP_CHECK(hipMemcpyAsync(b->h_stats, b->d_block + kStatsOffset, 8 * 4 * kStatSlots, hipMemcpyDeviceToHost, s));
  RT_HIP_CHECK(hipEventRecord(b->ev1, s));
  b->launched = true;
  b->counted = count;
  return RT_OK;
}

int BatchStats(rt_batch* b, uint64_t out[4]) {
  RT_REQUIRE(b->launched && b->counted, RT_ERR_STATE, "last launch did not count");
  RT_HIP_CHECK(hipEventSynchronize(b->ev1));
  for (int k = 0; k < 4; ++k) {
    out[k] = 0;
    for (int j = 0; j < kStatSlots; ++j) out[k] += b->h_stats[4 * j + k];
  }
  return RT_OK;
}

}  // namespace rayt

extern "C" {

int rt_batch_create(rt_cache* cache, int max_views, rt_batch** out) {
  RT_REQUIRE(cache && out, RT_ERR_INVALID, "null argument");
  rt_batch* b = nullptr;
  const int rc = rayt::BatchCreate(cache->ctx, max_views, &b);
  if (rc) return rc;
  b->scenes.push_back(cache);
  b->depths.push_back(0);
  *out = b;
  return RT_OK;
}

int rt_batch_destroy(rt_batch* b) {
  if (!b) return RT_OK;
  if (b->launched) (void)hipEventSynchronize(b->ev1);
  (void)hipFree(b->d_block);
  (void)hipFree(b->d_counters);
  if (b->h_block) (void)hipHostFree(b->h_block);
  if (b->h_stats) (void)hipHostFree(b->h_stats);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  rt_ctx* ctx = b->ctx;
  delete b;
  rayt::CtxRelease(ctx);
  return RT_OK;
}

int rt_batch_add_scene(rt_batch* b, rt_cache* cache) {
  RT_REQUIRE(b && cache, RT_ERR_INVALID, "null argument");
  RT_REQUIRE(cache->ctx == b->ctx, RT_ERR_INVALID, "scene belongs to another context");
  RT_REQUIRE(int(b->scenes.size()) < rayt::kMaxBatchScenes, RT_ERR_RANGE, "too many scenes");
  b->scenes.push_back(cache);
  b->depths.push_back(0);
  return int(b->scenes.size()) - 1;
}

int rt_batch_config(rt_batch* b, int tile_shape, int lds_levels, int wgs_per_cu, int tree_depth) {
  RT_REQUIRE(b, RT_ERR_INVALID, "null batch");
  RT_REQUIRE(tile_shape >= 0 && tile_shape <= 2, RT_ERR_RANGE, "tile_shape must be 0 (8x8), 1 (16x4) or 2 (32x2)");
  RT_REQUIRE(lds_levels == 8 || lds_levels == 10 || lds_levels == 12 || lds_levels == 16, RT_ERR_RANGE,
             "lds_levels must be 8, 10, 12 or 16");
  RT_REQUIRE(wgs_per_cu >= 0 && wgs_per_cu <= 32, RT_ERR_RANGE,
             "wgs_per_cu must be in [0, 32] (0: one wave per tile, not persistent)");
  RT_REQUIRE(tree_depth >= 0 && tree_depth <= 22, RT_ERR_RANGE, "tree_depth must be in [0, 22] (0: unknown)");
  b->tile_shape = tile_shape;
  b->lds_levels = lds_levels;
  b->wgs_per_cu = wgs_per_cu;
  b->tree_depth = tree_depth;
  return RT_OK;
}

int rt_batch_render(rt_batch* b, const rt_view* views, int nviews, rt_buf* out_image, rt_buf* codes,
                    uint64_t* stats) {
  const int rc = rayt::BatchLaunch(b, views, nviews, out_image, codes, stats != nullptr);
  if (rc || !stats) return rc;
  return rayt::BatchStats(b, stats);
}

int rt_batch_last_ms(rt_batch* b, double* ms) {
  RT_REQUIRE(b && ms, RT_ERR_INVALID, "null argument");
  RT_REQUIRE(b->launched, RT_ERR_STATE, "nothing launched yet");
  RT_REQUIRE(b->last_timed, RT_ERR_STATE, "the last launch was not timed");
  RT_HIP_CHECK(hipEventSynchronize(b->ev1));
  float f = 0;
  RT_HIP_CHECK(hipEventElapsedTime(&f, b->ev0, b->ev1));
  *ms = f;
  return RT_OK;
}

int rt_batch_autotune(rt_batch* b, const rt_view* views, int nviews, rt_buf* out_image, rt_buf* codes, int repeats,
                      int best[3], double* best_ms) {
  RT_REQUIRE(b && views && out_image && best && best_ms, RT_ERR_INVALID, "null argument");
  RT_REQUIRE(repeats >= 1 && repeats <= 16, RT_ERR_RANGE, "repeats must be in [1, 16]");
  static const int kShapes[3] = {0, 1, 2};
  static const int kLevels[4] = {8, 10, 12, 16};
  // 0: one wave per tile; else persistent one-wave workgroups per CU (the LDS window caps how many fit).
  static const int kWgs[5] = {0, 8, 16, 24, 32};
  double best_t = 1e300;
  int bc[3] = {b->tile_shape, b->lds_levels, b->wgs_per_cu};
  for (int shape : kShapes)
    for (int levels : kLevels)
      for (int wgs : kWgs) {
        // LDS per one-wave workgroup is the stack window; skip shapes that cannot be resident.
        const int lds_bytes = levels * 64 * 8;
        if (wgs * lds_bytes > 160 * 1024) continue;
        b->tile_shape = shape;
        b->lds_levels = levels;
        b->wgs_per_cu = wgs;
        std::vector<double> t;
        for (int r = 0; r < repeats + 1; ++r) {  // the first launch warms the caches and is not counted
          int rc = rt_batch_render(b, views, nviews, out_image, codes, nullptr);
          if (rc) return rc;
          double ms;
          rc = rt_batch_last_ms(b, &ms);
          if (rc) return rc;
          if (r) t.push_back(ms);
        }
        std::sort(t.begin(), t.end());
        const double med = t[t.size() / 2];
        if (med < best_t) {
          best_t = med;
          bc[0] = shape;
          bc[1] = levels;
          bc[2] = wgs;
        }
      }
  b->tile_shape = bc[0];
  b->lds_levels = bc[1];
  b->wgs_per_cu = bc[2];
  for (int k = 0; k < 3; ++k) best[k] = bc[k];
  *best_ms = best_t;
  return RT_OK;
}

}  // extern "C"
