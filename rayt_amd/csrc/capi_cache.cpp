// C ABI of the block cache (GPUOctreeCacheManager and the data it exposes).
#include <cstring>

#include "cache.h"
#include "tree_handle.h"

namespace {
// Finishes a tracer's pipelined frame before the cache's host state is read or changed.
int Settle(const rt_cache* c) { return (c && c->pending) ? rayt::FlushPending(c->pending) : RT_OK; }
}  // namespace

#define RT_SETTLE(c)                   \
  do {                                 \
    const int settle_rc_ = Settle(c);  \
    if (settle_rc_) return settle_rc_; \
  } while (0)

extern "C" {

int rt_cache_create(rt_ctx* ctx, rt_tree* tree, int max_blocks_count, rt_cache** out) {
  RT_REQUIRE(ctx && tree && out, RT_ERR_INVALID, "null argument");
  RT_REQUIRE(max_blocks_count > 0 && max_blocks_count <= rayt::kMaxBlocksCount, RT_ERR_RANGE,
             "max_blocks_count out of range");
  const auto& h = tree->src->header();
  RT_REQUIRE(uint64_t(max_blocks_count) * h.nodes_in_block < (1ull << 31), RT_ERR_RANGE,
             "pool would exceed 2^31 nodes (hit ids are int32)");
  RT_TRY_BEGIN
  rt_cache* c = new rt_cache();
  c->ctx = ctx;
  c->tree = tree;
  try {
    c->mgr.reset(new rayt::BlockCacheManager(ctx, tree->src.get(), max_blocks_count));
  } catch (...) {
    delete c;
    throw;
  }
  rayt::CtxRetain(ctx);
  *out = c;
  return RT_OK;
  RT_TRY_END
}

int rt_cache_destroy(rt_cache* cache) {
  if (!cache) return RT_OK;
  (void)Settle(cache);
  rt_ctx* ctx = cache->ctx;
  delete cache;
  rayt::CtxRelease(ctx);
  return RT_OK;
}

int rt_cache_initial_fill(rt_cache* c) {
  RT_REQUIRE(c, RT_ERR_INVALID, "null cache");
  RT_SETTLE(c);
  return c->mgr->InitialFillCache();
}

int rt_cache_root_node_index(const rt_cache* c) { return c ? c->mgr->root_node_index() : RT_ERR_INVALID; }
int rt_cache_nodes_in_block(const rt_cache* c) { return c ? c->mgr->nodes_in_block() : RT_ERR_INVALID; }
int rt_cache_max_blocks(const rt_cache* c) { return c ? c->mgr->cache().max_blocks_count() : RT_ERR_INVALID; }
int rt_cache_loaded_blocks(const rt_cache* c) {
  RT_REQUIRE(c, RT_ERR_INVALID, "null cache");
  RT_SETTLE(c);
  return c->mgr->cache().loaded_blocks_count();
}

rt_buf* rt_cache_channel(rt_cache* c, int index) {
  if (!c || index < 0 || index >= c->mgr->cache().pool().channels_count()) return nullptr;
  return c->mgr->cache().pool().channel(index);
}

rt_buf* rt_cache_channel_by_name(rt_cache* c, const char* name) {
  if (!c || !name) return nullptr;
  int idx = c->mgr->cache().pool().header().ChannelIndex(name);
  return idx < 0 ? nullptr : c->mgr->cache().pool().channel(idx);
}

rt_buf* rt_cache_far_pointers(rt_cache* c) { return c ? c->mgr->cache().pool().far_pointers() : nullptr; }

int rt_cache_block_slot(const rt_cache* c, int tree_block_index) {
  if (!c || tree_block_index < 0 || Settle(c)) return -1;
  return c->mgr->cache().BlockIndexInCache(uint32_t(tree_block_index));
}

int rt_cache_slot_block(const rt_cache* c, int slot) { return (c && !Settle(c)) ? c->mgr->cache().SlotBlock(slot) : -1; }

int rt_cache_start_session(rt_cache* c) {
  RT_REQUIRE(c, RT_ERR_INVALID, "null cache");
  RT_SETTLE(c);
  return c->mgr->StartRequestSession();
}
int rt_cache_start_transaction(rt_cache* c) {
  RT_REQUIRE(c, RT_ERR_INVALID, "null cache");
  RT_SETTLE(c);
  return c->mgr->StartRequestTransaction();
}
int rt_cache_mark_block_used(rt_cache* c, int slot) {
  RT_REQUIRE(c, RT_ERR_INVALID, "null cache");
  RT_SETTLE(c);
  return c->mgr->MarkBlockAsUsed(slot);
}
int rt_cache_mark_parent_used(rt_cache* c, int slot) {
  RT_REQUIRE(c, RT_ERR_INVALID, "null cache");
  RT_SETTLE(c);
  return c->mgr->MarkParentAsUsed(slot);
}
int rt_cache_request_block(rt_cache* c, int tree_block_index) {
  RT_REQUIRE(c, RT_ERR_INVALID, "null cache");
  RT_SETTLE(c);
  return c->mgr->RequestBlock(tree_block_index);
}
int rt_cache_transaction_filled(const rt_cache* c) {
  RT_REQUIRE(c, RT_ERR_INVALID, "null cache");
  RT_SETTLE(c);
  RT_REQUIRE(c->mgr->transaction_in_progress(), RT_ERR_STATE, "no transaction in progress");
  return c->mgr->TransactionFilledCache() ? 1 : 0;
}
int rt_cache_session_filled(const rt_cache* c) {
  RT_REQUIRE(c, RT_ERR_INVALID, "null cache");
  RT_SETTLE(c);
  RT_REQUIRE(c->mgr->session_in_progress(), RT_ERR_STATE, "no session in progress");
  return c->mgr->SessionFilledCache() ? 1 : 0;
}
int rt_cache_end_transaction(rt_cache* c) {
  RT_REQUIRE(c, RT_ERR_INVALID, "null cache");
  RT_SETTLE(c);
  return c->mgr->EndRequestTransaction();
}
int rt_cache_end_session(rt_cache* c) {
  RT_REQUIRE(c, RT_ERR_INVALID, "null cache");
  RT_SETTLE(c);
  return c->mgr->EndRequestSession();
}

int rt_cache_lru_order(const rt_cache* c, int* out, int cap) {
  RT_REQUIRE(c, RT_ERR_INVALID, "null cache");
  RT_SETTLE(c);
  std::vector<int> v = c->mgr->cache().LruOrder();
  for (int i = 0; i < cap && i < int(v.size()); ++i) out[i] = v[i];
  return int(v.size());
}

int rt_cache_counters(const rt_cache* c, uint64_t out[4]) {
  RT_REQUIRE(c && out, RT_ERR_INVALID, "null argument");
  RT_SETTLE(c);
  const auto& k = c->mgr->cache();
  out[0] = k.uploaded;
  out[1] = k.evicted;
  out[2] = k.patches;
  out[3] = k.far_writes;
  return RT_OK;
}

}  // extern "C"
