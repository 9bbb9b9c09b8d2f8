// Internal definitions shared by the C-ABI implementation files.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <string>

#include "../../include/rayt_hip.h"

// Reference counted: rt_ctx_destroy drops the caller's reference, and every
// object that keeps the context (buffers, caches, tracers, batches, servers)
// holds one more, so the stream outlives all of them whatever the order of
// destruction.
struct rt_ctx {
  std::atomic<int> refs{1};
  int device = 0;
  int compute_units = 0;
  hipStream_t stream = nullptr;
  int plain_lds_levels = 16;  // LDS stack levels for the plain kernel
  int plain_waves = 4;        // waves per workgroup of the plain kernel (1: 8x8 pixels, 4: 16x16)
  int plain_lds_base = 0;     // first stack level held in LDS
  bool plain_spill = true;    // false: the declared tree depth fits the LDS window (window-only stack)
};

struct rt_buf {
  rt_ctx* ctx = nullptr;
  void* ptr = nullptr;
  uint64_t size = 0;
  bool owned = true;
};

struct rt_evt {
  hipEvent_t ev = nullptr;
  std::atomic<int> refs{1};
};

struct rt_tracer;

namespace rayt {

void SetError(const std::string& msg);
int Fail(int code, const std::string& msg);
const char* LastError();

// Context references (see rt_ctx). Release destroys the stream and the
// context with the last reference.
rt_ctx* CtxRetain(rt_ctx* ctx);
void CtxRelease(rt_ctx* ctx);

struct MarkTargets;

// Tile order of a plain launch, for a caller that renders the same view size
// again and again (the tracer). Each wave stores its duration in cycles in
// `cost` (waves of tile t at t * waves_per_tile); after the launch a
// one-workgroup kernel writes into `order` the next launch's tiles, the
// longest (from this launch's costs) first (OrderTiles), so that the frame
// does not end on a long tile started last. order_valid = false: this launch
// takes the tiles in raster order. The output does not depend on the order.
constexpr int kMaxScheduledTiles = 64000;  // larger views keep the raster order
struct TileSchedule {
  uint32_t* order = nullptr;  // one entry per tile
  uint32_t* cost = nullptr;   // one entry per wave
  bool order_valid = false;
};

// rt_render_plain, and with `mark` also the marking of its block-level codes
// into a feedback extractor's bitmaps (FeedbackExtractor::Targets), so that
// the extractor's Process can skip its own pass over the codes; with `sched`
// the tiles are scheduled as TileSchedule says.
int RenderPlain(rt_ctx* ctx, rt_buf* out_image, rt_buf* codes, int width, int height, const rt_buf* node_links,
                const rt_buf* far_pointers, int root_node_index, const rt_buf* normals, const rt_buf* colors,
                int nodes_in_block, const float vpi[16], float lod_voxel_size, const float background[4], int flags,
                uint64_t* stats, const MarkTargets* mark, const TileSchedule* sched = nullptr);

// Writes sched.order for the next launch from the costs a RenderPlain with
// `sched` just stored (queued on the context stream after it).
int OrderTiles(rt_ctx* ctx, const TileSchedule& sched, int width, int height);

// Tiles of one plain launch at the context's launch shape, and the workgroup size in waves.
void PlainTiles(const rt_ctx* ctx, int width, int height, uint32_t* tiles, int* waves_per_tile);

// Looks up, once per device, the device address of the static scratch array
// of the sort that allocates nothing (rt_sort_i32_noalloc). Called by
// rt_ctx_create with `device` current.
int NoAllocScratchInit(int device);

// A tracer whose last frame was pipelined (rt_tracer_render_frame on a fully
// resident pool) still owes that frame's host half: applying its feedback
// lists to the cache's LRU and filling its statistics. The cache and the
// tracer's feedback extractor point at such a tracer, and every entry point
// that reads or changes their state finishes the frame first.
int FlushPending(rt_tracer* tracer);

}  // namespace rayt

#define RT_HIP_CHECK(call)                                                                    \
  do {                                                                                        \
    hipError_t e_ = (call);                                                                   \
    if (e_ != hipSuccess) return rayt::Fail(RT_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define RT_REQUIRE(cond, code, msg) \
  do {                              \
    if (!(cond)) return rayt::Fail((code), (msg)); \
  } while (0)

#define RT_TRY_BEGIN try {
#define RT_TRY_END                                              \
  }                                                             \
  catch (const std::bad_alloc&) {                               \
    return rayt::Fail(RT_ERR_NOMEM, "host allocation failed");  \
  }                                                             \
  catch (const std::exception& e) {                             \
    return rayt::Fail(RT_ERR_STATE, e.what());                  \
  }
