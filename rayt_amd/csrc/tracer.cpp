// Frame renderer: GPURayTracer (GPU_renderer/gpu-ray-tracer.cpp:20-227).
//
// RenderFrame runs InitTracingFrame, then tracing passes each followed by
// feedback extraction and one cache transaction (mark hit blocks, request
// faulted blocks, upload), until a pass faults nothing, the session has
// cycled the whole cache, or the pass budget is spent; FinishTracingFrame
// then composes the image. When the cache holds every block of the tree no
// ray can fault, and the single pass plus finish run as one fused kernel.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <limits>

#include "cache.h"
#include "feedback.h"
#include "tree_handle.h"

struct rt_tracer {
  rt_cache* cache = nullptr;
  rt_ctx* ctx = nullptr;
  int width = 0, height = 0;
  double frame_time_limit = 1.0;  // gpu-ray-tracer.cpp:30
  float lod_voxel_size = 1.5f;    // gpu-ray-tracer.cpp:31
  int max_passes = 0;
  int flags = 0;
  rt_buf* out_image = nullptr;
  rt_buf* codes = nullptr;          // the last frame's codes: one of codes_buf
  rt_buf* codes_buf[2] = {nullptr, nullptr};
  // Resident frames schedule their tiles longest first from the previous
  // frame's wave costs (TileSchedule); sched_waves is the launch shape the
  // order was made for.
  rt_buf* tile_cost = nullptr;
  rt_buf* tile_order = nullptr;
  int sched_waves = 0;
  bool order_ready = false;   // tile_order holds an order for sched_waves (written by a queued rebuild)
  uint64_t sched_frames = 0;  // resident frames since the shape's first
  float order_view[16] = {};  // view of the frame whose costs made the order
  float last_view[16] = {};   // view of the previous resident frame
  rt_buf* states = nullptr;
  rt_feedback* feedback = nullptr;
  uint64_t stats[5] = {0, 0, 0, 0, 0};
  std::vector<int> pass_counts, pass_faults;
  // HIP events around every tracing-pass launch of the last frame, and the
  // summed device time between them.
  std::vector<hipEvent_t> ev;
  double trace_ms = 0.0, frame_ms = 0.0;
  double trace_ms_total = 0.0;  // over every finished frame
  uint64_t frames_total = 0;
  // Pipelined frames (fully resident pools): the frame's trace and feedback
  // walk are queued and the call returns; the host half (the LRU marks of its
  // hit lists, its statistics) runs when the next frame's trace is queued
  // behind it, or when anything reads the state (FlushPending). Two code
  // buffers alternate, so the sort path of a pending frame's feedback still
  // finds its codes while the next frame writes the other buffer.
  bool pipelining = true;
  bool pending = false;
  int pending_buf = 0;
  std::chrono::steady_clock::time_point pending_t0;
  hipEvent_t pev[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};  // trace start / end per code buffer
  // The feedback walk of a pipelined frame runs on walk_stream, reading the
  // bitmap set of the frame's code buffer, while the next frame's trace marks
  // the other set; walk_done[buf] follows the walk of set buf.
  hipStream_t walk_stream = nullptr;
  hipEvent_t walk_done[2] = {nullptr, nullptr};
  bool walk_queued[2] = {false, false};
  hipEvent_t Event(size_t i) {
    while (ev.size() <= i) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) throw std::runtime_error("hipEventCreate failed");
      ev.push_back(e);
    }
    return ev[i];
  }
};

namespace {

const float kBackground[4] = {0.0f, float(216 / 255.), 1.0f, 1.0f};  // gpu-ray-tracer.cpp:190

// The transaction of one pass over the extractor's lists (gpu-ray-tracer.cpp:137-186).
int Transaction(rt_tracer* t, rayt::BlockCacheManager& m, bool& done) {
  rayt::FeedbackExtractor& fx = *t->feedback->fx;
  int rc = m.StartRequestTransaction();
  if (rc) return rc;
  for (int h : fx.hit_blocks())
    if ((rc = m.MarkBlockAsUsed(h))) return rc;
  for (int h : fx.duplicate_hit_blocks())
    if ((rc = m.MarkParentAsUsed(h))) return rc;
  const std::vector<int>& faults = fx.fault_blocks();
  for (size_t i = 0; i < faults.size(); ++i) {
    if (m.TransactionFilledCache()) break;
    if ((rc = m.RequestBlock(faults[i]))) return rc;
  }
  t->pass_counts.push_back(int(faults.size()));
  t->pass_faults.insert(t->pass_faults.end(), faults.begin(), faults.end());
  t->stats[1] += faults.size();
  const uint64_t before = m.cache().uploaded;
  // The uploads stay queued: the next pass follows them on the same stream.
  rc = m.EndRequestTransaction(false);
  if (rc) return rc;
  t->stats[2] += m.cache().uploaded - before;
  done = faults.empty() || m.SessionFilledCache();
  return RT_OK;
}

int FeedbackAndTransaction(rt_tracer* t, rayt::BlockCacheManager& m, bool premarked, bool& done) {
  rayt::FeedbackExtractor& fx = *t->feedback->fx;
  const int rc = fx.ProcessN(static_cast<int*>(t->codes->ptr), fx.count(), m.nodes_in_block(), true, premarked);
  return rc ? rc : Transaction(t, m, done);
}

// Resident frames between rebuilds of the tile order: the rebuild (one
// workgroup, about 30 us on the context stream) follows the first frame of a
// launch shape, the first frame at a view the camera comes to rest on, and
// then every kReorderFrames-th frame; the frames between keep the order. A
// camera at rest gives the same costs frame after frame, so the periodic
// rebuild only refreshes the order.
constexpr uint64_t kReorderFrames = 64;

// The tile schedule of the next resident frame, of view `vpi`. Its order is
// valid when it was made for the same launch shape from a frame of the same
// view: the tiles' costs move with the camera, and an order made for another
// view starts long tiles late (a moving camera renders in raster order).
rayt::TileSchedule Schedule(rt_tracer* t, const float vpi[16]) {
  rayt::TileSchedule s;
  uint32_t tiles;
  int waves;
  rayt::PlainTiles(t->ctx, t->width, t->height, &tiles, &waves);
  s.order = static_cast<uint32_t*>(t->tile_order->ptr);
  s.cost = static_cast<uint32_t*>(t->tile_cost->ptr);
  if (t->sched_waves != waves) {  // a new launch shape starts again from raster order
    t->order_ready = false;
    t->sched_frames = 0;
  }
  s.order_valid = t->order_ready && memcmp(t->order_view, vpi, sizeof(t->order_view)) == 0;
  t->sched_waves = waves;
  if (tiles > uint32_t(rayt::kMaxScheduledTiles)) {  // raster order, no costs: nothing to order
    s.order_valid = false;
    s.order = nullptr;
    s.cost = nullptr;
  }
  return s;
}

// Queues the rebuild of the tile order after a resident frame of view `vpi`
// when it is due, or at once when the camera has come to rest on a view the
// order was not made for.
int MaybeReorder(rt_tracer* t, const rayt::TileSchedule& sched, const float vpi[16]) {
  if (!sched.cost) return RT_OK;
  const bool rested = memcmp(vpi, t->last_view, sizeof(t->last_view)) == 0 &&
                      memcmp(vpi, t->order_view, sizeof(t->order_view)) != 0;
  memcpy(t->last_view, vpi, sizeof(t->last_view));
  const bool due = t->sched_frames % kReorderFrames == 0 || rested;
  ++t->sched_frames;
  if (!due) return RT_OK;
  const int rc = rayt::OrderTiles(t->ctx, sched, t->width, t->height);
  if (rc == RT_OK) {
    t->order_ready = true;
    memcpy(t->order_view, vpi, sizeof(t->order_view));
  }
  return rc;
}

// The trace's start and end events time the trace, and the end event orders
// the feedback walk (another stream of the same device) after it. Neither
// makes the trace's writes visible to the host, so they are recorded with a
// device-scope release instead of the default system-scope fence, whose
// cache writeback put about 11 us between the traces of consecutive frames.
hipEvent_t PipelineEvent(rt_tracer* t, int buf, int which) {
  hipEvent_t& e = t->pev[buf][which];
  if (!e && hipEventCreateWithFlags(&e, hipEventReleaseToDevice) != hipSuccess)
    throw std::runtime_error("hipEventCreate failed");
  return e;
}

// The host half of a pipelined frame: its feedback lists, one transaction, the
// end of its session, its statistics and timing.
int CompletePipelined(rt_tracer* t) {
  t->pending = false;
  if (t->cache->pending == t) t->cache->pending = nullptr;
  t->feedback->pending = nullptr;
  rayt::BlockCacheManager& m = *t->cache->mgr;
  int rc = t->feedback->fx->CompleteN();
  if (rc) return rc;
  bool done = false;
  rc = Transaction(t, m, done);
  if (rc) return rc;
  t->stats[0] = 1;
  t->stats[3] = m.SessionFilledCache() ? 1 : 0;
  t->stats[4] = uint64_t(t->width) * t->height;
  rc = m.EndRequestSession();
  if (rc) return rc;
  float ms = 0.f;
  RT_HIP_CHECK(hipEventSynchronize(t->pev[t->pending_buf][1]));
  RT_HIP_CHECK(hipEventElapsedTime(&ms, t->pev[t->pending_buf][0], t->pev[t->pending_buf][1]));
  t->trace_ms = ms;
  t->trace_ms_total += ms;
  t->frames_total += 1;
  t->frame_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t->pending_t0).count();
  return RT_OK;
}

// A frame over a fully resident pool: the single pass (the fused renderer,
// marking its codes for the feedback) and the feedback walk are queued; the
// previous pipelined frame is finished on the host while this trace runs.
int RenderPipelined(rt_tracer* t, const float vpi[16]) {
  const auto t0 = std::chrono::steady_clock::now();
  rayt::BlockCacheManager& m = *t->cache->mgr;
  rt_cache* c = t->cache;
  const int buf = t->pending ? 1 - t->pending_buf : t->pending_buf;
  rt_buf* codes = t->codes_buf[buf];
  rayt::FeedbackExtractor& fx = *t->feedback->fx;
  const rayt::MarkTargets targets =
      fx.Targets(m.source()->header().blocks_count, 2u * uint32_t(m.cache().max_blocks_count()) + 2u, buf);
  const rayt::TileSchedule sched = Schedule(t, vpi);
  if (!t->walk_stream) {
    RT_HIP_CHECK(hipStreamCreateWithFlags(&t->walk_stream, hipStreamNonBlocking));
    for (hipEvent_t& e : t->walk_done) RT_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  // the walk that last read (and cleared) this bitmap set has finished before the trace marks it (a walk
  // already done needs no wait on the stream, which would put a barrier between the traces)
  if (t->walk_queued[buf]) {
    const hipError_t q = hipEventQuery(t->walk_done[buf]);
    if (q == hipErrorNotReady)
      RT_HIP_CHECK(hipStreamWaitEvent(t->ctx->stream, t->walk_done[buf], 0));
    else
      RT_HIP_CHECK(q);
    t->walk_queued[buf] = false;
  }
  RT_HIP_CHECK(hipEventRecord(PipelineEvent(t, buf, 0), t->ctx->stream));
  int rc = rayt::RenderPlain(t->ctx, t->out_image, codes, t->width, t->height, rt_cache_channel(c, 0),
                             rt_cache_far_pointers(c), m.root_node_index(), rt_cache_channel_by_name(c, "normals"),
                             rt_cache_channel_by_name(c, "colors"), m.nodes_in_block(), vpi, t->lod_voxel_size,
                             kBackground, t->flags, nullptr, &targets, sched.cost ? &sched : nullptr);
  if (rc) return rc;
  RT_HIP_CHECK(hipEventRecord(PipelineEvent(t, buf, 1), t->ctx->stream));
  if ((rc = MaybeReorder(t, sched, vpi))) return rc;
  if (t->pending && (rc = CompletePipelined(t))) return rc;
  for (uint64_t& x : t->stats) x = 0;
  t->pass_counts.clear();
  t->pass_faults.clear();
  t->codes = codes;
  rc = m.StartRequestSession();
  if (rc) return rc;
  rc = fx.EnqueueN(static_cast<int*>(codes->ptr), fx.count(), m.nodes_in_block(), true, true, buf, t->walk_stream,
                   PipelineEvent(t, buf, 1));
  if (rc) return rc;
  RT_HIP_CHECK(hipEventRecord(t->walk_done[buf], t->walk_stream));
  t->walk_queued[buf] = true;
  t->pending = true;
  t->pending_buf = buf;
  t->pending_t0 = t0;
  c->pending = t;
  t->feedback->pending = t;
  return RT_OK;
}

}  // namespace

namespace rayt {

int FlushPending(rt_tracer* t) {
  if (!t || !t->pending) return RT_OK;
  RT_TRY_BEGIN
  return CompletePipelined(t);
  RT_TRY_END
}

}  // namespace rayt

extern "C" {

int rt_tracer_create(rt_cache* cache, int width, int height, rt_tracer** out) {
  RT_REQUIRE(cache && out, RT_ERR_INVALID, "null argument");
  RT_REQUIRE(width > 0 && height > 0 && int64_t(width) * height <= (int64_t(1) << 30), RT_ERR_RANGE,
             "bad frame size");
  rt_tracer* t = new rt_tracer();
  t->cache = cache;
  t->ctx = rayt::CtxRetain(cache->ctx);
  t->width = width;
  t->height = height;
  const uint64_t P = uint64_t(width) * height;
  int rc = rt_buf_create(t->ctx, P * 4, &t->out_image);
  if (!rc) rc = rt_buf_create(t->ctx, P * 4, &t->codes_buf[0]);
  if (!rc) rc = rt_buf_create(t->ctx, P * 4, &t->codes_buf[1]);
  {
    // entries for either launch shape: 8x8 tiles of one wave, 16x16 tiles of four
    const uint64_t t8 = uint64_t((width + 7) / 8) * uint64_t((height + 7) / 8);
    const uint64_t t16 = uint64_t((width + 15) / 16) * uint64_t((height + 15) / 16);
    if (!rc) rc = rt_buf_create(t->ctx, std::max(t8, 4 * t16) * 4, &t->tile_cost);
    if (!rc) rc = rt_buf_create(t->ctx, std::max(t8, t16) * 4, &t->tile_order);
  }
  t->codes = t->codes_buf[0];
  if (!rc) rc = rt_buf_create(t->ctx, P * RT_STATE_BYTES_PER_PIXEL, &t->states);
  if (!rc) rc = rt_feedback_create(t->ctx, int(P), &t->feedback);
  if (rc) {
    rt_tracer_destroy(t);
    return rc;
  }
  *out = t;
  return RT_OK;
}

int rt_tracer_destroy(rt_tracer* t) {
  if (!t) return RT_OK;
  (void)rayt::FlushPending(t);
  if (t->walk_stream) (void)hipStreamSynchronize(t->walk_stream);
  if (t->cache && t->cache->pending == t) t->cache->pending = nullptr;
  rt_buf_destroy(t->out_image);
  rt_buf_destroy(t->codes_buf[0]);
  rt_buf_destroy(t->codes_buf[1]);
  rt_buf_destroy(t->tile_cost);
  rt_buf_destroy(t->tile_order);
  rt_buf_destroy(t->states);
  rt_feedback_destroy(t->feedback);
  for (hipEvent_t e : t->ev) (void)hipEventDestroy(e);
  for (auto& pair : t->pev)
    for (hipEvent_t e : pair)
      if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : t->walk_done)
    if (e) (void)hipEventDestroy(e);
  if (t->walk_stream) (void)hipStreamDestroy(t->walk_stream);
  rt_ctx* ctx = t->ctx;
  delete t;
  rayt::CtxRelease(ctx);
  return RT_OK;
}

int rt_tracer_set_frame_time_limit(rt_tracer* t, double s) {
  RT_REQUIRE(t, RT_ERR_INVALID, "null tracer");
  RT_REQUIRE(s >= 0, RT_ERR_RANGE, "negative time limit");
  t->frame_time_limit = s;
  return RT_OK;
}
int rt_tracer_set_lod_voxel_size(rt_tracer* t, float size) {
  RT_REQUIRE(t, RT_ERR_INVALID, "null tracer");
  RT_REQUIRE(size >= 0, RT_ERR_RANGE, "negative lod voxel size");
  t->lod_voxel_size = size;
  return RT_OK;
}
int rt_tracer_set_max_passes(rt_tracer* t, int passes) {
  RT_REQUIRE(t, RT_ERR_INVALID, "null tracer");
  RT_REQUIRE(passes >= 0, RT_ERR_RANGE, "negative pass budget");
  t->max_passes = passes;
  return RT_OK;
}
int rt_tracer_set_flags(rt_tracer* t, int flags) {
  RT_REQUIRE(t, RT_ERR_INVALID, "null tracer");
  t->flags = flags & RT_FLAG_SHADOWS;
  return RT_OK;
}
int rt_tracer_set_seed(rt_tracer* t, uint64_t seed) {
  RT_REQUIRE(t, RT_ERR_INVALID, "null tracer");
  if (const int rc = rayt::FlushPending(t)) return rc;  // the pending frame draws with the old seed
  t->feedback->fx->set_seed(seed);
  return RT_OK;
}

int rt_tracer_render_frame(rt_tracer* t, const float vpi[16]) {
  RT_REQUIRE(t && vpi, RT_ERR_INVALID, "null argument");
  RT_TRY_BEGIN
  if (t->cache->pending && t->cache->pending != t) {  // another tracer's frame on this cache comes first
    const int rc = rayt::FlushPending(t->cache->pending);
    if (rc) return rc;
  }
  {
    rayt::BlockCacheManager& m = *t->cache->mgr;
    const bool resident = uint32_t(m.cache().loaded_blocks_count()) == m.source()->header().blocks_count;
    if (resident && t->pipelining) return RenderPipelined(t, vpi);
    const int rc = rayt::FlushPending(t);
    if (rc) return rc;
    t->codes = t->codes_buf[t->pending_buf];
  }
  const auto t0 = std::chrono::steady_clock::now();
  rayt::BlockCacheManager& m = *t->cache->mgr;
  rt_cache* c = t->cache;
  for (uint64_t& s : t->stats) s = 0;
  t->pass_counts.clear();
  t->pass_faults.clear();
  const int nib = m.nodes_in_block();
  rt_buf* links = rt_cache_channel(c, 0);
  rt_buf* far_ptrs = rt_cache_far_pointers(c);
  rt_buf* normals = rt_cache_channel_by_name(c, "normals");
  rt_buf* colors = rt_cache_channel_by_name(c, "colors");
  const bool resident = uint32_t(m.cache().loaded_blocks_count()) == m.source()->header().blocks_count;
  int rc;
  if (!resident) {
    rc = rt_init_tracing_frame(t->ctx, t->states, t->width, t->height, vpi, t->lod_voxel_size);
    if (rc) return rc;
  }
  rc = m.StartRequestSession();
  if (rc) return rc;
  int iterations;
  for (iterations = 0;; ++iterations) {
    RT_HIP_CHECK(hipEventRecord(t->Event(2 * size_t(iterations)), t->ctx->stream));
    // The first pass of a resident frame marks its codes for the feedback as it writes them (the
    // feedback always follows the first pass).
    const bool premarked = resident && iterations == 0;
    const rayt::MarkTargets targets =
        t->feedback->fx->Targets(m.source()->header().blocks_count, 2u * uint32_t(m.cache().max_blocks_count()) + 2u);
    rayt::TileSchedule sched;
    if (resident) {
      sched = Schedule(t, vpi);
      rc = rayt::RenderPlain(t->ctx, t->out_image, t->codes, t->width, t->height, links, far_ptrs,
                             m.root_node_index(), normals, colors, nib, vpi, t->lod_voxel_size, kBackground,
                             t->flags, nullptr, premarked ? &targets : nullptr, sched.cost ? &sched : nullptr);
    }
    else
      rc = rt_raytracing_pass(t->ctx, t->codes, t->states, t->width, t->height, links, far_ptrs, m.root_node_index(),
                              normals, colors, nib, t->flags);
    if (rc) return rc;
    RT_HIP_CHECK(hipEventRecord(t->Event(2 * size_t(iterations) + 1), t->ctx->stream));
    if (resident && (rc = MaybeReorder(t, sched, vpi))) return rc;
    if (iterations > 0) {
      const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      if ((t->max_passes > 0 && iterations + 1 >= t->max_passes) || el > t->frame_time_limit) break;
    }
    bool done = false;
    rc = FeedbackAndTransaction(t, m, premarked, done);
    if (rc) return rc;
    if (done) break;
  }
  t->stats[0] = uint64_t(iterations + 1);
  t->stats[3] = m.SessionFilledCache() ? 1 : 0;
  t->stats[4] = uint64_t(t->width) * t->height;
  rc = m.EndRequestSession();
  if (rc) return rc;
  if (!resident) {
    rc = rt_finish_tracing_frame(t->ctx, t->out_image, t->states, t->width, t->height, kBackground, normals, colors);
    if (rc) return rc;
  }
  RT_HIP_CHECK(hipStreamSynchronize(t->ctx->stream));
  t->trace_ms = 0.0;
  for (int i = 0; i <= iterations; ++i) {
    float ms = 0.f;
    RT_HIP_CHECK(hipEventElapsedTime(&ms, t->ev[2 * i], t->ev[2 * i + 1]));
    t->trace_ms += ms;
  }
  t->trace_ms_total += t->trace_ms;
  t->frames_total += 1;
  t->frame_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return RT_OK;
  RT_TRY_END
}

int rt_tracer_set_pipelining(rt_tracer* t, int enabled) {
  RT_REQUIRE(t, RT_ERR_INVALID, "null tracer");
  if (!enabled)
    if (const int rc = rayt::FlushPending(t)) return rc;
  t->pipelining = enabled != 0;
  return RT_OK;
}

int rt_tracer_trace_totals(rt_tracer* t, double* trace_ms, uint64_t* frames) {
  RT_REQUIRE(t, RT_ERR_INVALID, "null tracer");
  if (const int rc = rayt::FlushPending(t)) return rc;
  if (trace_ms) *trace_ms = t->trace_ms_total;
  if (frames) *frames = t->frames_total;
  return RT_OK;
}

int rt_tracer_timings(const rt_tracer* t, double out_ms[2]) {
  RT_REQUIRE(t && out_ms, RT_ERR_INVALID, "null argument");
  if (const int rc = rayt::FlushPending(const_cast<rt_tracer*>(t))) return rc;
  out_ms[0] = t->trace_ms;
  out_ms[1] = t->frame_ms;
  return RT_OK;
}

rt_buf* rt_tracer_out_image(rt_tracer* t) { return t ? t->out_image : nullptr; }
rt_buf* rt_tracer_codes(rt_tracer* t) { return t ? t->codes : nullptr; }
rt_feedback* rt_tracer_feedback(rt_tracer* t) { return t ? t->feedback : nullptr; }

int rt_tracer_frame_stats(const rt_tracer* t, uint64_t out[5]) {
  RT_REQUIRE(t && out, RT_ERR_INVALID, "null argument");
  if (const int rc = rayt::FlushPending(const_cast<rt_tracer*>(t))) return rc;
  memcpy(out, t->stats, sizeof(t->stats));
  return RT_OK;
}

int rt_tracer_pass_faults(const rt_tracer* t, int* counts, int counts_cap, int* faults, int faults_cap) {
  RT_REQUIRE(t, RT_ERR_INVALID, "null tracer");
  if (const int rc = rayt::FlushPending(const_cast<rt_tracer*>(t))) return rc;
  for (int i = 0; i < counts_cap && i < int(t->pass_counts.size()); ++i) counts[i] = t->pass_counts[i];
  for (int i = 0; i < faults_cap && i < int(t->pass_faults.size()); ++i) faults[i] = t->pass_faults[i];
  return int(t->pass_counts.size());
}

}  // extern "C"
