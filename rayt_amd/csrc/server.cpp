// Continuous-batching render server (no reference counterpart; each request's
// result is the plain single-view frame, rendered by the batched kernel).
//
// One server thread owns all device work. Its loop
//   1. retires rounds whose completion event has fired: copies each result
//      from the pinned mirror to the caller's memory, frees its arena range,
//      records its latency;
//   2. when a round slot is free, admits pending requests in FIFO order while
//      the round has room, the request's output fits in the framebuffer
//      arena and its scene is (or can be made) resident, and launches them
//      as one batched launch on the slot's stream, followed by the copies of
//      the results into the pinned mirror and the slot's completion event.
//      A round of at most kDirectRoundPixels pixels has the kernel write its
//      images straight into the pinned mirror (host memory the GPU writes
//      over PCIe), so an image crosses PCIe while the round renders instead
//      of after it (the latency of single requests); a larger round renders
//      into device memory and its images are copied after it, which lets the
//      copy engine overlap the next rounds (the throughput of bulk rounds).
//      The codes stay in device memory, where the feedback extractor reads
//      them, and the requested ones are copied.
// The first pending request that cannot be admitted blocks those behind it
// (no overtaking), so admission is FIFO.
//
// A scene whose block pool is smaller than its tree is rendered out of core.
// When a round retires, the feedback codes of each of its requests on such a
// scene go through the feedback extractor, and the hit and fault lists are
// merged into the scene's pending loads. Once no round of the scene is in
// flight, one request session with one transaction marks the hit blocks used
// and loads the faulted blocks (LRU eviction inside the scene's pool), so the
// pool contents never change under a round that reads them. A request whose
// scene has faults pending waits (with the requests behind it) until the
// scene's rounds in flight have retired and the loads are done.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "batch.h"
#include "cache.h"
#include "feedback.h"
#include "rt_internal.h"
#include "tree_handle.h"

namespace {

constexpr int kMaxSlots = 4;
constexpr uint64_t kBytesPerPixel = 8;  // RGBA8 image word + int32 feedback code
// Rounds of at most this many pixels (two 3840x2160 views) write their images
// into the pinned mirror from the kernel: at 200 req/s over depths 12-14 the
// p99 fell from 2.3 to 1.8 ms that way, while bursts of 64-view rounds ran
// about 10 % faster with the images copied after each round.
constexpr uint64_t kDirectRoundPixels = uint64_t(2) * 3840 * 2160;

using Clock = std::chrono::steady_clock;

// First-fit allocator over [0, capacity) pixel units with coalescing frees.
class ArenaAllocator {
 public:
  explicit ArenaAllocator(uint64_t capacity) : capacity_(capacity) {
    if (capacity) free_[0] = capacity;
  }
  bool Allocate(uint64_t n, uint64_t* offset) {
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < n) continue;
      *offset = it->first;
      const uint64_t rest = it->second - n, at = it->first + n;
      free_.erase(it);
      if (rest) free_[at] = rest;
      return true;
    }
    return false;
  }
  void Release(uint64_t offset, uint64_t n) {
    auto it = free_.emplace(offset, n).first;
    auto next = std::next(it);
    if (next != free_.end() && it->first + it->second == next->first) {
      it->second += next->second;
      free_.erase(next);
    }
    if (it != free_.begin()) {
      auto prev = std::prev(it);
      if (prev->first + prev->second == it->first) {
        prev->second += it->second;
        free_.erase(it);
      }
    }
  }
  uint64_t capacity() const { return capacity_; }

 private:
  uint64_t capacity_;
  std::map<uint64_t, uint64_t> free_;
};

struct Scene {
  rt_tree* tree = nullptr;
  int depth = 11;
  uint64_t pool_bytes = 0;
  int max_blocks = 0;
  bool partial = false;       // the pool holds fewer blocks than the tree: out of core
  rt_cache* cache = nullptr;  // resident when non-null
  int in_flight = 0;          // admitted requests not yet retired
  uint64_t last_use = 0;
  // Feedback of retired rounds not yet applied (slots and tree blocks as of
  // the pool state those rounds saw: no load happens while a round is in flight).
  std::vector<int> pend_hits, pend_dups, pend_faults;
  // The uploads of the last load transaction stay queued on the context
  // stream; `loaded` is recorded behind them and every round of the scene
  // waits for it on its own stream.
  hipEvent_t loaded = nullptr;
  bool load_queued = false;
  void ClearPending() {
    pend_hits.clear();
    pend_dups.clear();
    pend_faults.clear();
  }
};

struct Pending {
  uint64_t id;
  rt_request req;
  double submit_ms;
};

struct Admitted {
  Pending p;
  uint64_t px_offset, px_count;
  double admit_ms;
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  rt_batch* batch = nullptr;
  bool busy = false;
  std::vector<Admitted> reqs;
  std::vector<int> scenes;  // server scene ids of this round, in round scene index order
};

double PercentileNearestRank(std::vector<double> v, double q) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  size_t rank = size_t(q * double(v.size()) + 0.999999999);  // ceil(q n), at least 1
  rank = std::max<size_t>(1, std::min(rank, v.size()));
  return v[rank - 1];
}

}  // namespace

struct rt_server {
  rt_ctx* ctx = nullptr;
  rt_server_config cfg{};
  Clock::time_point t0;
  // device and pinned memory
  rt_buf* arena_image = nullptr;   // device images of large rounds
  rt_buf* pinned_image = nullptr;  // the image half of the pinned mirror (a view of `pinned`, not owned)
  rt_buf* arena_codes = nullptr;
  uint8_t* pinned = nullptr;  // mirror: image words then codes, same pixel offsets
  ArenaAllocator* arena = nullptr;
  Slot slots[kMaxSlots];
  std::vector<Scene> scenes;
  uint64_t pool_used = 0;
  uint64_t use_clock = 0;
  std::unique_ptr<rayt::FeedbackExtractor> fx;  // sized for the largest request of a partial scene so far
  hipStream_t fx_stream = nullptr;  // its stream: not the context's, where block uploads may be queued
  // shared with submitters / pollers
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Pending> pending;
  std::deque<rt_completion> completions;
  uint64_t next_id = 1;
  uint64_t submitted = 0, retired = 0;
  bool stop = false;
  // pinned results held for their callers (id -> arena range), and ranges released by them
  std::map<uint64_t, std::pair<uint64_t, uint64_t>> held;
  std::vector<std::pair<uint64_t, uint64_t>> released;
  // statistics (guarded by mu)
  std::vector<double> latencies;
  rt_server_stats st{};
  double first_admit_ms = -1, last_done_ms = 0;
  uint64_t views_in_rounds = 0;
  std::thread thread;

  double Now() const { return std::chrono::duration<double, std::milli>(Clock::now() - t0).count(); }
  void Loop();
  int MakeResident(int scene_id);
  bool FormRound(Slot& s);
  void Retire(Slot& s, int status);
  int Launch(Slot& s);
  int Feedback(Scene& sc, const Admitted& a, int* fault_blocks);
  int ApplyPending(Scene& sc);
};

int rt_server::MakeResident(int id) {
  Scene& sc = scenes[id];
  if (sc.cache) return RT_OK;
  // Evict idle resident scenes, least recently used first, until the pool fits.
  while (pool_used + sc.pool_bytes > cfg.pool_bytes) {
    int victim = -1;
    for (int i = 0; i < int(scenes.size()); ++i)
      if (scenes[i].cache && scenes[i].in_flight == 0 && (victim < 0 || scenes[i].last_use < scenes[victim].last_use))
        victim = i;
    if (victim < 0) return RT_ERR_NOMEM;  // every resident scene is in use: wait for rounds to retire
    rt_cache_destroy(scenes[victim].cache);  // waits for the uploads queued on the context stream
    scenes[victim].cache = nullptr;
    scenes[victim].load_queued = false;
    scenes[victim].ClearPending();
    pool_used -= scenes[victim].pool_bytes;
    std::lock_guard<std::mutex> g(mu);
    ++st.scene_evictions;
  }
  rt_cache* c = nullptr;
  int rc = rt_cache_create(ctx, sc.tree, sc.max_blocks, &c);
  if (rc == RT_OK) rc = rt_cache_initial_fill(c);
  if (rc == RT_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = RT_ERR_HIP;
  if (rc != RT_OK) {
    if (c) rt_cache_destroy(c);
    return rc;
  }
  sc.cache = c;
  pool_used += sc.pool_bytes;
  std::lock_guard<std::mutex> g(mu);
  ++st.scene_loads;
  return RT_OK;
}

bool rt_server::FormRound(Slot& s) {
  s.reqs.clear();
  s.scenes.clear();
  while (int(s.reqs.size()) < cfg.max_views_per_round) {
    Pending p;
    {
      std::lock_guard<std::mutex> g(mu);
      if (pending.empty()) break;
      p = pending.front();
    }
    const uint64_t npx = uint64_t(p.req.width) * p.req.height;
    Scene& sc = scenes[p.req.scene];
    if (!sc.pend_faults.empty()) {  // loads due: they wait for the scene's rounds in flight
      if (sc.in_flight > 0) break;
      if (ApplyPending(sc) != RT_OK) fprintf(stderr, "rt_server: block loads failed: %s\n", rayt::LastError());
    }
    uint64_t off = 0;
    if (!arena->Allocate(npx, &off)) break;  // framebuffer memory: wait for a round to retire
    const int rc = MakeResident(p.req.scene);
    if (rc == RT_ERR_NOMEM) {  // block-pool memory: wait for a round to retire
      arena->Release(off, npx);
      break;
    }
    {
      std::lock_guard<std::mutex> g(mu);
      pending.pop_front();
    }
    if (rc != RT_OK) {  // the scene cannot be loaded at all: fail this request
      arena->Release(off, npx);
      rt_completion c{p.id, rc, 0, p.submit_ms, Now(), Now()};
      std::lock_guard<std::mutex> g(mu);
      completions.push_back(c);
      ++retired;
      ++st.failed;
      cv_done.notify_all();
      continue;
    }
    ++sc.in_flight;
    sc.last_use = ++use_clock;
    s.reqs.push_back(Admitted{p, off, npx, Now()});
  }
  return !s.reqs.empty();
}

int rt_server::Launch(Slot& s) {
  std::vector<rt_cache*> caches;
  std::vector<int> depths;
  std::vector<rt_view> views(s.reqs.size());
  for (size_t i = 0; i < s.reqs.size(); ++i) {
    const rt_request& r = s.reqs[i].p.req;
    int local = -1;
    for (size_t k = 0; k < s.scenes.size(); ++k)
      if (s.scenes[k] == r.scene) local = int(k);
    if (local < 0) {
      local = int(s.scenes.size());
      s.scenes.push_back(r.scene);
      caches.push_back(scenes[r.scene].cache);
      depths.push_back(scenes[r.scene].depth);
    }
    rt_view& v = views[i];
    memcpy(v.view_proj_inv, r.view_proj_inv, sizeof(v.view_proj_inv));
    v.width = r.width;
    v.height = r.height;
    v.lod_voxel_size = r.lod_voxel_size;
    v.flags = r.flags;
    v.scene = local;
    v.pixel_offset = s.reqs[i].px_offset;
  }
  int rc = rayt::BatchSetScenes(s.batch, caches, depths);
  if (rc) return rc;
  rayt::BatchSetStream(s.batch, s.stream);
  rayt::BatchSetTiming(s.batch, false);
  rayt::BatchSetRayCounts(s.batch, true);  // the server reports rays, not node fetches
  for (int id : s.scenes)  // the scene's last block loads, queued on the context stream
    if (scenes[id].load_queued) RT_HIP_CHECK(hipStreamWaitEvent(s.stream, scenes[id].loaded, 0));
  uint64_t round_px = 0;
  for (const Admitted& a : s.reqs) round_px += a.px_count;
  const bool direct = round_px <= kDirectRoundPixels;
  rc = rayt::BatchLaunch(s.batch, views.data(), int(views.size()), direct ? pinned_image : arena_image, arena_codes,
                         true);
  if (rc) return rc;
  // The results into the pinned mirror (the images of a direct round are there already), requests adjacent in
  // the arena in one copy each.
  const uint64_t cap = arena->capacity();
  for (int which = direct ? 1 : 0; which < 2; ++which) {
    const int flag = which == 0 ? RT_OUT_IMAGE : RT_OUT_CODES;
    std::vector<std::pair<uint64_t, uint64_t>> ranges;
    for (const Admitted& a : s.reqs)
      if (a.p.req.outputs & flag) ranges.push_back({a.px_offset, a.px_count});
    std::sort(ranges.begin(), ranges.end());
    const uint8_t* src = static_cast<const uint8_t*>(which == 0 ? arena_image->ptr : arena_codes->ptr);
    uint8_t* dst = pinned + (which == 0 ? 0 : cap * 4);
    for (size_t i = 0; i < ranges.size();) {
      const uint64_t begin = ranges[i].first;
      uint64_t end = begin + ranges[i].second;
      for (++i; i < ranges.size() && ranges[i].first == end; ++i) end += ranges[i].second;
      RT_HIP_CHECK(hipMemcpyAsync(dst + begin * 4, src + begin * 4, (end - begin) * 4, hipMemcpyDeviceToHost, s.stream));
    }
  }
  RT_HIP_CHECK(hipEventRecord(s.done, s.stream));
  return RT_OK;
}

void rt_server::Retire(Slot& s, int status) {
  uint64_t counts[4] = {0, 0, 0, 0};
  if (status == RT_OK) status = rayt::BatchStats(s.batch, counts);
  const uint64_t cap = arena->capacity();
  std::vector<rt_completion> done;
  std::vector<std::pair<uint64_t, std::pair<uint64_t, uint64_t>>> hold;
  for (const Admitted& a : s.reqs) {
    const rt_request& r = a.p.req;
    int fault_blocks = 0;
    if (status == RT_OK && scenes[r.scene].partial) {
      const int rc = Feedback(scenes[r.scene], a, &fault_blocks);
      if (rc != RT_OK) fprintf(stderr, "rt_server: feedback failed: %s\n", rayt::LastError());
    }
    const uint32_t* img = reinterpret_cast<const uint32_t*>(pinned + a.px_offset * 4);
    const int32_t* cod = reinterpret_cast<const int32_t*>(pinned + (cap + a.px_offset) * 4);
    const bool keep = status == RT_OK && r.delivery == RT_DELIVER_PINNED;
    if (status == RT_OK && r.delivery == RT_DELIVER_COPY) {
      if (r.outputs & RT_OUT_IMAGE) memcpy(r.out_image, img, a.px_count * 4);
      if (r.outputs & RT_OUT_CODES) memcpy(r.out_codes, cod, a.px_count * 4);
    }
    if (keep)
      hold.push_back({a.p.id, {a.px_offset, a.px_count}});
    else
      arena->Release(a.px_offset, a.px_count);
    --scenes[r.scene].in_flight;
    done.push_back(rt_completion{a.p.id, status, int(s.reqs.size()), a.p.submit_ms, a.admit_ms, Now(),
                                 keep && (r.outputs & RT_OUT_IMAGE) ? img : nullptr,
                                 keep && (r.outputs & RT_OUT_CODES) ? cod : nullptr, fault_blocks, 0});
  }
  for (Scene& sc : scenes)
    if (sc.partial && sc.in_flight == 0 && (!sc.pend_faults.empty() || !sc.pend_hits.empty()) &&
        ApplyPending(sc) != RT_OK)
      fprintf(stderr, "rt_server: block loads failed: %s\n", rayt::LastError());
  std::lock_guard<std::mutex> g(mu);
  for (const auto& h : hold) held[h.first] = h.second;
  for (const rt_completion& c : done) {
    completions.push_back(c);
    ++retired;
    if (c.status == RT_OK) {
      ++st.completed;
      latencies.push_back(c.done_ms - c.submit_ms);
      last_done_ms = std::max(last_done_ms, c.done_ms);
    } else {
      ++st.failed;
    }
  }
  for (const Admitted& a : s.reqs) st.pixels += status == RT_OK ? a.px_count : 0;
  st.primary_rays += counts[0];
  st.shadow_rays += counts[1];
  s.reqs.clear();
  s.busy = false;
  cv_done.notify_all();
}

// Feedback of one retired request on a partial scene (its codes are still in
// the framebuffer arena): the extractor's lists join the scene's pending loads.
int rt_server::Feedback(Scene& sc, const Admitted& a, int* fault_blocks) {
  *fault_blocks = 0;
  if (!sc.cache) return RT_OK;
  const int n = int(a.px_count);
  if (!fx || fx->count() < n) {
    fx.reset();
    fx.reset(new rayt::FeedbackExtractor(ctx, n));
    fx->set_stream(fx_stream);
  }
  int* codes = static_cast<int*>(arena_codes->ptr) + a.px_offset;
  const int rc = fx->ProcessN(codes, n, sc.cache->mgr->nodes_in_block(), true);
  if (rc) return rc;
  *fault_blocks = int(fx->fault_blocks().size());
  sc.pend_faults.insert(sc.pend_faults.end(), fx->fault_blocks().begin(), fx->fault_blocks().end());
  sc.pend_hits.insert(sc.pend_hits.end(), fx->hit_blocks().begin(), fx->hit_blocks().end());
  sc.pend_dups.insert(sc.pend_dups.end(), fx->duplicate_hit_blocks().begin(), fx->duplicate_hit_blocks().end());
  return RT_OK;
}

// One request session with one transaction over the merged feedback of the
// scene's retired rounds, as a RenderFrame pass does with one pass's lists
// (gpu-ray-tracer.cpp:141-166): hit blocks marked used, then the faulted
// blocks requested in descending block order until the transaction fills the
// pool or max_block_loads are requested (the others are reported again by
// later requests). Runs only while no round of the scene is in flight. The
// uploads are left queued on the context stream (the scene's next rounds wait
// for them on the device), so the server goes on retiring and launching.
int rt_server::ApplyPending(Scene& sc) {
  auto unique = [](std::vector<int>& v, bool descending) {
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    if (descending) std::reverse(v.begin(), v.end());
  };
  if (!sc.cache) {
    sc.ClearPending();
    return RT_OK;
  }
  unique(sc.pend_hits, false);
  unique(sc.pend_dups, false);
  unique(sc.pend_faults, true);
  rayt::BlockCacheManager& m = *sc.cache->mgr;
  const uint64_t before = m.cache().uploaded;
  int rc = m.StartRequestSession();
  if (rc == RT_OK) rc = m.StartRequestTransaction();
  for (size_t i = 0; rc == RT_OK && i < sc.pend_hits.size(); ++i) rc = m.MarkBlockAsUsed(sc.pend_hits[i]);
  for (size_t i = 0; rc == RT_OK && i < sc.pend_dups.size(); ++i) rc = m.MarkParentAsUsed(sc.pend_dups[i]);
  size_t nfaults = sc.pend_faults.size();
  if (cfg.max_block_loads > 0) nfaults = std::min(nfaults, size_t(cfg.max_block_loads));
  for (size_t i = 0; rc == RT_OK && i < nfaults && !m.TransactionFilledCache(); ++i)
    rc = m.RequestBlock(sc.pend_faults[i]);
  if (rc == RT_OK) rc = m.EndRequestTransaction(false);
  if (rc == RT_OK) rc = m.EndRequestSession();
  if (rc == RT_OK && !sc.loaded && hipEventCreateWithFlags(&sc.loaded, hipEventDisableTiming) != hipSuccess)
    rc = rayt::Fail(RT_ERR_HIP, "load event");
  if (rc == RT_OK) {
    if (hipEventRecord(sc.loaded, ctx->stream) == hipSuccess)
      sc.load_queued = true;
    else
      rc = rayt::Fail(RT_ERR_HIP, "load event record");
  }
  if (rc != RT_OK) (void)hipStreamSynchronize(ctx->stream);  // nothing of a failed transaction is left queued
  sc.ClearPending();
  std::lock_guard<std::mutex> g(mu);
  st.block_uploads += m.cache().uploaded - before;
  ++st.load_transactions;
  return rc;
}

void rt_server::Loop() {
  (void)hipSetDevice(ctx->device);
  while (true) {
    bool progressed = false;
    int busy = 0;
    {
      std::vector<std::pair<uint64_t, uint64_t>> ranges;
      {
        std::lock_guard<std::mutex> g(mu);
        ranges.swap(released);
      }
      for (const auto& r : ranges) arena->Release(r.first, r.second);
      progressed = !ranges.empty();
    }
    for (Slot& s : slots) {
      if (!s.busy) continue;
      const hipError_t q = hipEventQuery(s.done);
      if (q == hipSuccess) {
        Retire(s, RT_OK);
        progressed = true;
      } else if (q != hipErrorNotReady) {
        Retire(s, RT_ERR_HIP);
        progressed = true;
      } else {
        ++busy;
      }
    }
    {
      std::lock_guard<std::mutex> g(mu);
      if (stop && busy == 0) break;
    }
    bool blocked = false;  // the queue head waits for memory that only a release or a retire frees
    for (int k = 0; k < cfg.rounds_in_flight; ++k) {
      Slot& s = slots[k];
      if (s.busy) continue;
      {
        std::lock_guard<std::mutex> g(mu);
        if (pending.empty() || stop) break;
      }
      if (!FormRound(s)) {
        blocked = true;
        break;
      }
      const double t = Now();
      const int rc = Launch(s);
      progressed = true;
      {
        std::lock_guard<std::mutex> g(mu);
        if (first_admit_ms < 0) first_admit_ms = t;
        ++st.rounds;
        views_in_rounds += s.reqs.size();
      }
      if (rc != RT_OK) {
        fprintf(stderr, "rt_server: round launch failed: %s\n", rayt::LastError());
        (void)hipStreamSynchronize(s.stream);  // whatever was enqueued has finished before the arena is reused
        Retire(s, rc);
        continue;
      }
      s.busy = true;
      ++busy;
      std::lock_guard<std::mutex> g(mu);
      st.max_rounds_in_flight = std::max(st.max_rounds_in_flight, busy);
    }
    if (!progressed) {
      if (busy > 0) {
        // Rounds in flight: poll their events again without sleeping (a timed
        // wait rounds up to the kernel's timer slack, tens of microseconds).
        std::this_thread::yield();
        continue;
      }
      std::unique_lock<std::mutex> lk(mu);
      if (blocked)
        cv_work.wait_for(lk, std::chrono::milliseconds(2), [&] { return stop || !released.empty(); });
      else
        cv_work.wait_for(lk, std::chrono::milliseconds(2),
                         [&] { return stop || !pending.empty() || !released.empty(); });
    }
  }
}

extern "C" {

int rt_server_create(rt_ctx* ctx, const rt_server_config* cfg, rt_server** out) {
  RT_REQUIRE(ctx && cfg && out, RT_ERR_INVALID, "null argument");
  RT_REQUIRE(cfg->max_views_per_round >= 1 && cfg->max_views_per_round <= (1 << 16), RT_ERR_RANGE,
             "max_views_per_round must be in [1, 65536]");
  RT_REQUIRE(cfg->rounds_in_flight >= 1 && cfg->rounds_in_flight <= kMaxSlots, RT_ERR_RANGE,
             "rounds_in_flight must be in [1, 4]");
  RT_REQUIRE(cfg->framebuffer_bytes >= kBytesPerPixel && cfg->framebuffer_bytes <= (uint64_t(1) << 36), RT_ERR_RANGE,
             "framebuffer_bytes must be in [8, 64 GiB]");
  RT_REQUIRE(cfg->pool_bytes > 0, RT_ERR_RANGE, "pool_bytes must be positive");
  RT_REQUIRE(cfg->max_block_loads >= 0, RT_ERR_RANGE, "max_block_loads must be >= 0");
  RT_TRY_BEGIN
  rt_server* s = new rt_server();
  s->ctx = rayt::CtxRetain(ctx);
  s->cfg = *cfg;
  s->t0 = Clock::now();
  const uint64_t px = cfg->framebuffer_bytes / kBytesPerPixel;
  int rc = rt_buf_create(ctx, px * 4, &s->arena_image);
  if (rc == RT_OK) rc = rt_buf_create(ctx, px * 4, &s->arena_codes);
  if (rc == RT_OK && hipHostMalloc(reinterpret_cast<void**>(&s->pinned), px * 8, 0) != hipSuccess)
    rc = rayt::Fail(RT_ERR_NOMEM, "pinned framebuffer mirror");
  if (rc == RT_OK) {  // the kernels of small rounds write their images into the mirror
    s->pinned_image = new rt_buf();
    s->pinned_image->ctx = rayt::CtxRetain(ctx);
    s->pinned_image->ptr = s->pinned;
    s->pinned_image->size = px * 4;
    s->pinned_image->owned = false;
  }
  if (rc == RT_OK && hipStreamCreateWithFlags(&s->fx_stream, hipStreamNonBlocking) != hipSuccess)
    rc = rayt::Fail(RT_ERR_HIP, "server feedback stream");
  for (int k = 0; rc == RT_OK && k < cfg->rounds_in_flight; ++k) {
    Slot& sl = s->slots[k];
    if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess)
      rc = rayt::Fail(RT_ERR_HIP, "server stream");
    if (rc == RT_OK) rc = rayt::BatchCreate(ctx, cfg->max_views_per_round, &sl.batch);
    if (rc == RT_OK) rc = rt_batch_config(sl.batch, cfg->tile_shape, cfg->lds_levels, cfg->wgs_per_cu, 0);
  }
  if (rc != RT_OK) {
    rt_server_destroy(s);
    return rc;
  }
  s->arena = new ArenaAllocator(px);
  s->thread = std::thread([s] { s->Loop(); });
  *out = s;
  return RT_OK;
  RT_TRY_END
}

int rt_server_destroy(rt_server* s) {
  if (!s) return RT_OK;
  if (s->thread.joinable()) {
    {
      std::lock_guard<std::mutex> g(s->mu);
      s->stop = true;
      // pending requests complete as failed
      for (const Pending& p : s->pending) {
        s->completions.push_back(rt_completion{p.id, RT_ERR_STATE, 0, p.submit_ms, -1, s->Now()});
        ++s->retired;
        ++s->st.failed;
      }
      s->pending.clear();
    }
    s->cv_work.notify_all();
    s->thread.join();
  }
  for (Slot& sl : s->slots) {
    if (sl.stream) (void)hipStreamSynchronize(sl.stream);
    if (sl.batch) rt_batch_destroy(sl.batch);
    if (sl.done) (void)hipEventDestroy(sl.done);
    if (sl.stream) (void)hipStreamDestroy(sl.stream);
  }
  for (Scene& sc : s->scenes) {
    if (sc.cache) rt_cache_destroy(sc.cache);
    if (sc.loaded) (void)hipEventDestroy(sc.loaded);
  }
  s->fx.reset();
  if (s->fx_stream) (void)hipStreamDestroy(s->fx_stream);
  rt_buf_destroy(s->arena_image);
  rt_buf_destroy(s->pinned_image);
  rt_buf_destroy(s->arena_codes);
  if (s->pinned) (void)hipHostFree(s->pinned);
  delete s->arena;
  rt_ctx* ctx = s->ctx;
  delete s;
  rayt::CtxRelease(ctx);
  return RT_OK;
}

int rt_server_add_scene(rt_server* s, rt_tree* tree, int tree_depth) {
  return rt_server_add_scene_pool(s, tree, tree_depth, 0);
}

int rt_server_add_scene_pool(rt_server* s, rt_tree* tree, int tree_depth, int max_blocks) {
  RT_REQUIRE(s && tree, RT_ERR_INVALID, "null argument");
  RT_REQUIRE(tree_depth >= 0 && tree_depth <= 22, RT_ERR_RANGE, "tree_depth must be in [0, 22]");
  rt_tree_info info;
  int rc = rt_tree_info_get(tree, &info);
  if (rc) return rc;
  RT_REQUIRE(info.blocks_count > 0 && info.blocks_count <= uint32_t(rayt::kMaxBlocksCount), RT_ERR_RANGE,
             "scene block count out of range");
  Scene sc;
  sc.tree = tree;
  sc.depth = tree_depth ? tree_depth : 11;
  sc.max_blocks = max_blocks > 0 ? std::min(max_blocks, int(info.blocks_count)) : int(info.blocks_count);
  sc.partial = uint32_t(sc.max_blocks) < info.blocks_count;
  // channel data + 8 far pointers + one page-table word per block slot
  sc.pool_bytes = uint64_t(sc.max_blocks) * (info.bytes_in_block_content + 8 * 4 + 4);
  RT_REQUIRE(sc.pool_bytes <= s->cfg.pool_bytes, RT_ERR_RANGE, "scene block pool exceeds the server pool budget");
  std::lock_guard<std::mutex> g(s->mu);
  RT_REQUIRE(int(s->scenes.size()) < rayt::kMaxBatchScenes, RT_ERR_RANGE, "too many scenes");
  // The server thread reads `scenes` without the lock, so scenes are added only while it is idle.
  RT_REQUIRE(s->pending.empty() && s->submitted == s->retired, RT_ERR_STATE,
             "scenes are added before requests are submitted or after drain");
  s->scenes.push_back(sc);
  return int(s->scenes.size()) - 1;
}

int rt_server_submit_many(rt_server* s, const rt_request* reqs, int n, uint64_t* ids) {
  RT_REQUIRE(s && reqs && n >= 0, RT_ERR_INVALID, "bad argument");
  for (int i = 0; i < n; ++i) {
    const rt_request& r = reqs[i];
    RT_REQUIRE(r.width > 0 && r.height > 0 && int64_t(r.width) * r.height <= (int64_t(1) << 28), RT_ERR_RANGE,
               "bad request size");
    RT_REQUIRE(uint64_t(r.width) * r.height <= s->arena->capacity(), RT_ERR_RANGE,
               "request larger than the framebuffer arena");
    RT_REQUIRE(r.delivery == RT_DELIVER_COPY || r.delivery == RT_DELIVER_PINNED, RT_ERR_INVALID, "bad delivery");
    RT_REQUIRE((r.outputs & ~(RT_OUT_IMAGE | RT_OUT_CODES)) == 0, RT_ERR_INVALID, "bad outputs");
    RT_REQUIRE(r.delivery != RT_DELIVER_COPY || ((!(r.outputs & RT_OUT_IMAGE) || r.out_image) &&
                                                 (!(r.outputs & RT_OUT_CODES) || r.out_codes)),
               RT_ERR_INVALID, "copy delivery needs a destination for every requested output");
  }
  std::lock_guard<std::mutex> g(s->mu);
  RT_REQUIRE(!s->stop, RT_ERR_STATE, "server is stopping");
  for (int i = 0; i < n; ++i)
    RT_REQUIRE(reqs[i].scene >= 0 && reqs[i].scene < int(s->scenes.size()), RT_ERR_RANGE, "unknown scene");
  const double now = s->Now();
  for (int i = 0; i < n; ++i) {
    Pending p{s->next_id++, reqs[i], now};
    s->pending.push_back(p);
    if (ids) ids[i] = p.id;
  }
  s->submitted += uint64_t(n);
  s->st.submitted += uint64_t(n);
  s->cv_work.notify_one();
  return RT_OK;
}

int rt_server_submit(rt_server* s, const rt_request* req, uint64_t* id) { return rt_server_submit_many(s, req, 1, id); }

int rt_server_poll(rt_server* s, rt_completion* out, int cap, int timeout_ms) {
  RT_REQUIRE(s && out && cap > 0, RT_ERR_INVALID, "bad argument");
  std::unique_lock<std::mutex> lk(s->mu);
  if (s->completions.empty() && timeout_ms > 0)
    s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), [&] { return !s->completions.empty(); });
  int n = 0;
  while (n < cap && !s->completions.empty()) {
    out[n++] = s->completions.front();
    s->completions.pop_front();
  }
  return n;
}

int rt_server_release(rt_server* s, uint64_t id) {
  RT_REQUIRE(s, RT_ERR_INVALID, "null server");
  std::lock_guard<std::mutex> g(s->mu);
  auto it = s->held.find(id);
  RT_REQUIRE(it != s->held.end(), RT_ERR_INVALID, "no pinned result held for this id");
  s->released.push_back(it->second);
  s->held.erase(it);
  s->cv_work.notify_one();
  return RT_OK;
}

int rt_server_drain(rt_server* s, int timeout_ms) {
  RT_REQUIRE(s, RT_ERR_INVALID, "null server");
  std::unique_lock<std::mutex> lk(s->mu);
  const bool ok = s->cv_done.wait_for(lk, std::chrono::milliseconds(std::max(0, timeout_ms)),
                                      [&] { return s->retired == s->submitted; });
  RT_REQUIRE(ok, RT_ERR_STATE, "drain timed out");
  return RT_OK;
}

int rt_server_stats_get(rt_server* s, rt_server_stats* out) {
  RT_REQUIRE(s && out, RT_ERR_INVALID, "null argument");
  std::lock_guard<std::mutex> g(s->mu);
  rt_server_stats r = s->st;
  r.p50_ms = PercentileNearestRank(s->latencies, 0.50);
  r.p99_ms = PercentileNearestRank(s->latencies, 0.99);
  double sum = 0, mx = 0;
  for (double x : s->latencies) sum += x, mx = std::max(mx, x);
  r.mean_ms = s->latencies.empty() ? 0.0 : sum / double(s->latencies.size());
  r.max_ms = mx;
  r.span_ms = s->first_admit_ms >= 0 ? std::max(0.0, s->last_done_ms - s->first_admit_ms) : 0.0;
  r.mrays_per_s = r.span_ms > 0 ? double(r.primary_rays + r.shadow_rays) / (r.span_ms * 1e3) : 0.0;
  r.mean_views_per_round = r.rounds ? double(s->views_in_rounds) / double(r.rounds) : 0.0;
  *out = r;
  return RT_OK;
}

int rt_server_reset_stats(rt_server* s) {
  RT_REQUIRE(s, RT_ERR_INVALID, "null server");
  std::lock_guard<std::mutex> g(s->mu);
  const uint64_t submitted = s->st.submitted;
  s->st = rt_server_stats{};
  s->st.submitted = submitted;
  s->latencies.clear();
  s->first_admit_ms = -1;
  s->last_done_ms = 0;
  s->views_in_rounds = 0;
  return RT_OK;
}

}  // extern "C"
