// Batched rendering: many views of any resolution traced by one persistent
// kernel (no reference counterpart; per view the result is the plain
// single-pass frame of trace.hip, bit for bit).
//
// Work layout. Every view is cut into wave tiles of 64 pixels (8x8, 16x4 or
// 32x2); the tiles of all views form one global sequence, split into eight
// contiguous ranges, one per XCD (workgroup b runs on XCD b % 8). A wave
// claims a group of tiles from its XCD's range with one atomic, renders them,
// and when its range is exhausted steals from the next XCD's range; after all
// eight ranges are exhausted every wave exits, so the grid always drains.
// The traversal stack keeps the deepest `lds_levels` levels in LDS.
#include <algorithm>
#include <cstring>
#include <vector>

#include "cache.h"
#include "render_pixel.h"
#include "rt_internal.h"
#include "trace_common.h"

using namespace rtdev;

namespace {

constexpr int kMaxBatchViews = 1024;
constexpr uint32_t kTilesPerGrab = 4;

struct DevView {
  float m[16];
  int width, height;
  float lod;
  int flags;
  int tiles_x;
  int pad;
  uint64_t pixel_offset;
};

struct Vec4B {
  float v[4];
};

// Lanes that need a ray pull the next pixel of the wave's chunk; a chunk is
// kTilesPerGrab wave tiles claimed from the XCD queues. Refills happen when
// at least kRefillMin lanes are idle, so the cost of the bookkeeping is paid
// once per several traversal steps.
constexpr int kRefillMin = 16;

template <int TW, int LDS_LEVELS, bool COUNT>
__global__ __launch_bounds__(256) void BatchKernel(const DevView* __restrict__ views, int nviews,
                                                   const uint32_t* __restrict__ tile_begin, uint32_t total_tiles,
                                                   uint32_t* __restrict__ counters, uint32_t* __restrict__ out,
                                                   int* __restrict__ codes, PoolView pool, uint32_t root,
                                                   const uint32_t* __restrict__ normals,
                                                   const uint32_t* __restrict__ colors, int nodes_in_block, Vec4B bg,
                                                   Lights lt, int lds_base, unsigned long long* __restrict__ stats) {
  constexpr int TH = 64 / TW;
  __shared__ uint2 lds[LDS_LEVELS * 256];
  __shared__ uint32_t begin_s[kMaxBatchViews + 1];
  for (int i = threadIdx.x; i <= nviews; i += blockDim.x) begin_s[i] = tile_begin[i];
  __syncthreads();
  SplitStack<LDS_LEVELS, 256> stack;
  stack.lds = lds + threadIdx.x;
  stack.base = lds_base;
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t lanes_below = (lane == 0) ? 0ull : (~0ull >> (64u - lane));
  const uint32_t xcd = blockIdx.x & 7u;
  const uint32_t per = (total_tiles + 7u) / 8u;
  PixelCounters cnt;
  uint32_t fetches = 0;

  // wave-uniform chunk state
  uint32_t q_rel = 0, chunk_first = 0, chunk_pix = 0, chunk_npix = 0;
  bool queues_done = false;

  // lane state
  bool active = false;
  int phase = 0;  // 0: primary ray, 1: shadow ray
  Trav tr;
  uint64_t pix = 0;
  int flags = 0, code = 0, hit_node = 0;
  float ox = 0, oy = 0, oz = 0, dx = 0, dy = 0, dz = 0;

  while (true) {
    uint64_t need = __ballot(!active);
    if (!queues_done && __popcll(need) >= kRefillMin) {
      while (need && !queues_done) {
        if (chunk_pix >= chunk_npix) {
          while (true) {  // claim the next chunk from this XCD's queue, then steal from the others
            if (q_rel >= 8) {
              queues_done = true;
              break;
            }
            const uint32_t q = (xcd + q_rel) & 7u;
            const uint32_t qbegin = min(total_tiles, q * per), qend = min(total_tiles, qbegin + per);
            uint32_t t = 0;
            if (lane == 0) t = atomicAdd(&counters[q], kTilesPerGrab);
            t = __shfl(t, 0, 64);
            if (t < per && qbegin + t < qend) {
              chunk_first = qbegin + t;
              chunk_npix = (min(qend, chunk_first + kTilesPerGrab) - chunk_first) * 64u;
              chunk_pix = 0;
              break;
            }
            ++q_rel;
          }
          if (queues_done) break;
        }
        const uint32_t avail = chunk_npix - chunk_pix;
        const uint32_t rank = uint32_t(__popcll(need & lanes_below));
        const uint32_t take = min(uint32_t(__popcll(need)), avail);
        if (!active && rank < take) {
          const uint32_t p = chunk_pix + rank;
          const uint32_t tile = chunk_first + p / 64u, local_px = p % 64u;
          int lo = 0, hi = nviews - 1;
          while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (begin_s[mid] <= tile) lo = mid; else hi = mid - 1;
          }
          const DevView& v = views[lo];
          const uint32_t local = tile - begin_s[lo];
          const int x = int(local % uint32_t(v.tiles_x)) * TW + int(local_px % TW);
          const int y = int(local / uint32_t(v.tiles_x)) * TH + int(local_px / TW);
          if (x < v.width && y < v.height) {
            const PrimaryRay r = InitRay(v.m, x, y, v.width, v.height, v.lod);
            ox = r.origin.x, oy = r.origin.y, oz = r.origin.z;
            dx = r.direction.x, dy = r.direction.y, dz = r.direction.z;
            TravInit(tr, root, ox, oy, oz, dx, dy, dz, r.direction.w, r.origin.w);
            pix = v.pixel_offset + uint64_t(y) * v.width + x;
            flags = v.flags;
            phase = 0;
            active = true;
            if (COUNT) ++cnt.primaries;
          }
        }
        chunk_pix += take;
        need = __ballot(!active);
      }
    }
    if (!__any(active)) {
      if (queues_done) break;
      continue;  // every claimed pixel so far was outside its view; claim more
    }
    if (!active) continue;
    CastResult res;
    if (!TravStep<SplitStack<LDS_LEVELS, 256>, COUNT>(tr, pool, stack, res, fetches)) continue;
    uint32_t color;
    if (phase == 0) {
      if (res.error_code) {
        color = PackRGBA(255, 0, 0, 255);
        code = kNoHit;
      } else if (res.fault_block != -1) {
        color = ShadeNode(normals, colors, res.hit_node, lt, false);
        code = -res.fault_block - 1;
      } else if (res.hit_node == -1) {
        color = 0u;
        code = kNoHit;
      } else {
        code = HitCode(res.hit_node, res.hit_duplicate, nodes_in_block);
        hit_node = res.hit_node;
        if (COUNT) ++cnt.hits;
        bool start_shadow = false;
        if ((flags & 1) && normals) {
          const uint32_t n = normals[res.hit_node];
          const float nx = float(int8_t(n & 255u)) / 127.f, ny = float(int8_t((n >> 8) & 255u)) / 127.f,
                      nz = float(int8_t((n >> 16) & 255u)) / 127.f;
          start_shadow = cl_max(0.f, LightTerm(nx, ny, nz, lt.l1)) > 0.0f;
        }
        if (start_shadow) {
          // same origin rule as ShadowOccluded
          const float eps = __uint_as_float(uint32_t(127 - res.node_depth - 2) << 23);
          const float tb = cl_max(0.0f, res.hit_t - eps);
          TravInit(tr, root, ox + dx * tb, oy + dy * tb, oz + dz * tb, lt.shadow_dir[0], lt.shadow_dir[1],
                   lt.shadow_dir[2], 0.0f, 0.0f);
          phase = 1;
          if (COUNT) ++cnt.shadow_rays;
          continue;
        }
        color = ShadeNode(normals, colors, hit_node, lt, false);
      }
    } else {
      const bool occluded = res.error_code == 0 && res.fault_block == -1 && res.hit_node != -1;
      color = ShadeNode(normals, colors, hit_node, lt, occluded);
    }
    out[pix] = Blend(color, bg.v);
    if (codes) codes[pix] = code;
    active = false;
  }
  if (COUNT) {
    cnt.fetches = fetches;
    FlushCounters(cnt, stats);
  }
}

__global__ void ZeroCountersKernel(uint32_t* c, int n) {
  if (int(threadIdx.x) < n) c[threadIdx.x] = 0u;
}

template <int TW, int L, bool C>
void Launch(dim3 grid, hipStream_t s, const DevView* views, int nviews, const uint32_t* tb, uint32_t total,
            uint32_t* counters, uint32_t* out, int* codes, const PoolView& pool, uint32_t root, const uint32_t* nrm,
            const uint32_t* col, int nib, const Vec4B& bg, const Lights& lt, int base, unsigned long long* st) {
  BatchKernel<TW, L, C><<<grid, dim3(256), 0, s>>>(views, nviews, tb, total, counters, out, codes, pool, root, nrm,
                                                   col, nib, bg, lt, base, st);
}

template <bool C>
bool Dispatch(int tw, int levels, dim3 grid, hipStream_t s, const DevView* views, int nviews, const uint32_t* tb,
              uint32_t total, uint32_t* counters, uint32_t* out, int* codes, const PoolView& pool, uint32_t root,
              const uint32_t* nrm, const uint32_t* col, int nib, const Vec4B& bg, const Lights& lt, int base,
              unsigned long long* st) {
#define RT_B(TW, L)                                                                                            \
  if (tw == TW && levels == L) {                                                                               \
    Launch<TW, L, C>(grid, s, views, nviews, tb, total, counters, out, codes, pool, root, nrm, col, nib, bg, lt, \
                     base, st);                                                                                \
    return true;                                                                                               \
  }
  RT_B(8, 8) RT_B(8, 10) RT_B(8, 12) RT_B(8, 16)
  RT_B(16, 8) RT_B(16, 10) RT_B(16, 12) RT_B(16, 16)
  RT_B(32, 8) RT_B(32, 10) RT_B(32, 12) RT_B(32, 16)
#undef RT_B
  return false;
}

int TileWidth(int shape) { return shape == 1 ? 16 : shape == 2 ? 32 : 8; }

}  // namespace

struct rt_batch {
  rt_cache* cache = nullptr;
  rt_ctx* ctx = nullptr;
  int max_views = 0;
  int tile_shape = 0;   // 0: 8x8, 1: 16x4, 2: 32x2
  int lds_levels = 12;  // 8, 10, 12 or 16
  int wgs_per_cu = 4;
  int lds_base = -1;    // -1: derive from tree_depth
  int tree_depth = 11;
  DevView* d_views = nullptr;
  uint32_t* d_tile_begin = nullptr;
  uint32_t* d_counters = nullptr;
  unsigned long long* d_stats = nullptr;
  std::vector<DevView> h_views;
  std::vector<uint32_t> h_tile_begin;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0;
};

extern "C" {

int rt_batch_create(rt_cache* cache, int max_views, rt_batch** out) {
  RT_REQUIRE(cache && out, RT_ERR_INVALID, "null argument");
  RT_REQUIRE(max_views > 0 && max_views <= kMaxBatchViews, RT_ERR_RANGE, "max_views must be in [1, 1024]");
  rt_batch* b = new rt_batch();
  b->cache = cache;
  b->ctx = cache->ctx;
  b->max_views = max_views;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(&b->d_views), sizeof(DevView) * max_views);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&b->d_tile_begin), 4 * (max_views + 1));
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&b->d_counters), 4 * 8);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&b->d_stats), 8 * 4);
  if (e == hipSuccess) e = hipEventCreate(&b->ev0);
  if (e == hipSuccess) e = hipEventCreate(&b->ev1);
  if (e != hipSuccess) {
    rt_batch_destroy(b);
    return rayt::Fail(RT_ERR_HIP, std::string("batch allocation: ") + hipGetErrorString(e));
  }
  *out = b;
  return RT_OK;
}

int rt_batch_destroy(rt_batch* b) {
  if (!b) return RT_OK;
  if (b->ctx) (void)hipStreamSynchronize(b->ctx->stream);
  (void)hipFree(b->d_views);
  (void)hipFree(b->d_tile_begin);
  (void)hipFree(b->d_counters);
  (void)hipFree(b->d_stats);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  delete b;
  return RT_OK;
}

int rt_batch_config(rt_batch* b, int tile_shape, int lds_levels, int wgs_per_cu, int tree_depth) {
  RT_REQUIRE(b, RT_ERR_INVALID, "null batch");
  RT_REQUIRE(tile_shape >= 0 && tile_shape <= 2, RT_ERR_RANGE, "tile_shape must be 0 (8x8), 1 (16x4) or 2 (32x2)");
  RT_REQUIRE(lds_levels == 8 || lds_levels == 10 || lds_levels == 12 || lds_levels == 16, RT_ERR_RANGE,
             "lds_levels must be 8, 10, 12 or 16");
  RT_REQUIRE(wgs_per_cu >= 1 && wgs_per_cu <= 8, RT_ERR_RANGE, "wgs_per_cu must be in [1, 8]");
  RT_REQUIRE(tree_depth >= 1 && tree_depth <= 22, RT_ERR_RANGE, "tree_depth must be in [1, 22]");
  b->tile_shape = tile_shape;
  b->lds_levels = lds_levels;
  b->wgs_per_cu = wgs_per_cu;
  b->tree_depth = tree_depth;
  return RT_OK;
}

int rt_batch_render(rt_batch* b, const rt_view* views, int nviews, rt_buf* out_image, rt_buf* codes,
                    uint64_t* stats) {
  RT_REQUIRE(b && views && out_image, RT_ERR_INVALID, "null argument");
  RT_REQUIRE(nviews > 0 && nviews <= b->max_views, RT_ERR_RANGE, "nviews out of range");
  rayt::BlockCacheManager& m = *b->cache->mgr;
  const int tw = TileWidth(b->tile_shape), th = 64 / tw;
  b->h_views.resize(nviews);
  b->h_tile_begin.resize(nviews + 1);
  uint64_t tiles = 0;
  for (int i = 0; i < nviews; ++i) {
    const rt_view& v = views[i];
    RT_REQUIRE(v.width > 0 && v.height > 0 && int64_t(v.width) * v.height <= (int64_t(1) << 28), RT_ERR_RANGE,
               "bad view size");
    const uint64_t px = uint64_t(v.width) * v.height;
    RT_REQUIRE((v.pixel_offset + px) * 4 <= out_image->size && (!codes || (v.pixel_offset + px) * 4 <= codes->size),
               RT_ERR_RANGE, "view outside the output buffers");
    DevView& d = b->h_views[i];
    memcpy(d.m, v.view_proj_inv, sizeof(d.m));
    d.width = v.width;
    d.height = v.height;
    d.lod = v.lod_voxel_size;
    d.flags = v.flags & RT_FLAG_SHADOWS;
    d.tiles_x = (v.width + tw - 1) / tw;
    d.pad = 0;
    d.pixel_offset = v.pixel_offset;
    b->h_tile_begin[i] = uint32_t(tiles);
    tiles += uint64_t(d.tiles_x) * ((v.height + th - 1) / th);
    RT_REQUIRE(tiles < (1ull << 31), RT_ERR_RANGE, "too many tiles in one batch");
  }
  b->h_tile_begin[nviews] = uint32_t(tiles);
  hipStream_t s = b->ctx->stream;
  RT_HIP_CHECK(hipMemcpyAsync(b->d_views, b->h_views.data(), sizeof(DevView) * nviews, hipMemcpyHostToDevice, s));
  RT_HIP_CHECK(hipMemcpyAsync(b->d_tile_begin, b->h_tile_begin.data(), 4 * (nviews + 1), hipMemcpyHostToDevice, s));
  ZeroCountersKernel<<<dim3(1), dim3(64), 0, s>>>(b->d_counters, 8);
  if (stats) ZeroCountersKernel<<<dim3(1), dim3(64), 0, s>>>(reinterpret_cast<uint32_t*>(b->d_stats), 8);
  rt_buf* links = rt_cache_channel(b->cache, 0);
  rt_buf* far_ptrs = rt_cache_far_pointers(b->cache);
  rt_buf* normals = rt_cache_channel_by_name(b->cache, "normals");
  rt_buf* colors = rt_cache_channel_by_name(b->cache, "colors");
  PoolView pool;
  pool.links = static_cast<const uint32_t*>(links->ptr);
  pool.far_ptrs = static_cast<const uint32_t*>(far_ptrs->ptr);
  pool.nodes = uint32_t(std::min<uint64_t>(links->size / 4, 0xffffffffu));
  pool.far_count = uint32_t(std::min<uint64_t>(far_ptrs->size / 4, 0xffffffffu));
  Vec4B bg;
  const float kBg[4] = {0.0f, float(216 / 255.), 1.0f, 1.0f};
  memcpy(bg.v, kBg, sizeof(bg.v));
  const Lights lt = rayt::MakeLights();
  // Deepest stack levels in LDS: a depth-D tree pushes levels 0 .. D-2.
  const int base = std::max(0, (b->tree_depth - 1) - b->lds_levels);
  const uint32_t wgs = uint32_t(std::max(1, b->ctx->compute_units) * b->wgs_per_cu + 7) & ~7u;
  const dim3 grid(wgs);
  RT_HIP_CHECK(hipEventRecord(b->ev0, s));
  const bool ok =
      stats ? Dispatch<true>(TileWidth(b->tile_shape), b->lds_levels, grid, s, b->d_views, nviews, b->d_tile_begin,
                             uint32_t(tiles), b->d_counters, static_cast<uint32_t*>(out_image->ptr),
                             codes ? static_cast<int*>(codes->ptr) : nullptr, pool, uint32_t(m.root_node_index()),
                             normals ? static_cast<const uint32_t*>(normals->ptr) : nullptr,
                             colors ? static_cast<const uint32_t*>(colors->ptr) : nullptr, m.nodes_in_block(), bg, lt,
                             base, b->d_stats)
            : Dispatch<false>(TileWidth(b->tile_shape), b->lds_levels, grid, s, b->d_views, nviews, b->d_tile_begin,
                              uint32_t(tiles), b->d_counters, static_cast<uint32_t*>(out_image->ptr),
                              codes ? static_cast<int*>(codes->ptr) : nullptr, pool, uint32_t(m.root_node_index()),
                              normals ? static_cast<const uint32_t*>(normals->ptr) : nullptr,
                              colors ? static_cast<const uint32_t*>(colors->ptr) : nullptr, m.nodes_in_block(), bg,
                              lt, base, nullptr);
  RT_REQUIRE(ok, RT_ERR_STATE, "no kernel for this batch configuration");
  RT_HIP_CHECK(hipGetLastError());
  RT_HIP_CHECK(hipEventRecord(b->ev1, s));
  if (stats) {
    unsigned long long h[4];
    RT_HIP_CHECK(hipMemcpyAsync(h, b->d_stats, sizeof(h), hipMemcpyDeviceToHost, s));
    RT_HIP_CHECK(hipStreamSynchronize(s));
    for (int k = 0; k < 4; ++k) stats[k] = h[k];
  }
  return RT_OK;
}

int rt_batch_last_ms(rt_batch* b, double* ms) {
  RT_REQUIRE(b && ms, RT_ERR_INVALID, "null argument");
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
  static const int kWgs[4] = {2, 4, 6, 8};
  const int saved[3] = {b->tile_shape, b->lds_levels, b->wgs_per_cu};
  double best_t = 1e300;
  int bc[3] = {saved[0], saved[1], saved[2]};
  for (int shape : kShapes)
    for (int levels : kLevels)
      for (int wgs : kWgs) {
        // LDS per workgroup: the stack window plus the view table; skip shapes that cannot be resident.
        const int lds_bytes = levels * 256 * 8 + (kMaxBatchViews + 1) * 4;
        if (wgs * lds_bytes > 160 * 1024) continue;
        b->tile_shape = shape;
        b->lds_levels = levels;
        b->wgs_per_cu = wgs;
        std::vector<double> t;
        for (int r = 0; r < repeats + 1; ++r) {  // first launch warms caches and is not counted
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
