// HBM-resident block pool and its LRU block cache.
//
// BlockPool         ~ GPUOctreeData      (GPU data/gpu-octree-data.cpp:8-93)
// BlockCache        ~ GPUOctreeCache     (GPU data/gpu-octree-cache.cpp:11-313)
// BlockCacheManager ~ GPUOctreeCacheManager (GPU data/gpu-octree-cache-manager.cpp:11-212)
//
// The eviction policy, LRU markers, link conversion and parent-pointer
// rewrites are those of the reference. What differs is how bytes reach the
// device: blocks are staged raw (stored link format) in pinned host memory,
// copied with one transfer per batch, and a device kernel converts links to
// the cache format, scatters every channel into the slot and then applies
// the batch's parent-pointer and far-pointer writes in issue order.
#pragma once

#include <cstdint>
#include <functional>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "octree_format.h"
#include "rt_internal.h"

namespace rayt {

class BlockPool {
 public:
  BlockPool(rt_ctx* ctx, const StoredOctreeHeader& header, int max_blocks, int staging_blocks);
  ~BlockPool();
  BlockPool(const BlockPool&) = delete;
  BlockPool& operator=(const BlockPool&) = delete;

  int nodes_in_block() const { return nib_; }
  int max_blocks() const { return max_blocks_; }
  int channels_count() const { return int(channels_.size()); }
  rt_buf* channel(int c) { return channels_[c]; }
  rt_buf* far_pointers() { return far_; }
  rt_buf* slot_block_table() { return slot_block_; }
  const StoredOctreeHeader& header() const { return header_; }

  // Reserves staging for a block that goes to `slot`; returns the payload
  // area (BytesInBlockContent bytes) to fill with the stored block. Pending
  // link writes into `slot` issued earlier in this batch are dropped (the
  // whole slot is overwritten).
  uint8_t* StageBlock(int slot, uint32_t tree_block);
  // StageBlock + the copy of the payload. A payload that stays valid until
  // the next Flush (src_stable) is copied there, by several threads at once.
  void StageBlockCopy(int slot, uint32_t tree_block, const uint8_t* src, bool src_stable);
  // A 4-byte write into the node-links channel at node index `node`.
  void StageLinkWrite(uint64_t node, uint32_t value);
  void StageFarPointer(uint32_t index, uint32_t value);
  // Enqueues the pending batch (one H2D copy + scatter/patch kernels).
  int Flush();
  bool empty() const { return staged_slots_.empty() && link_writes_.empty() && far_writes_.empty(); }
  uint64_t content_bytes() const { return content_; }

 private:
  rt_ctx* ctx_;
  StoredOctreeHeader header_;
  int nib_, max_blocks_, staging_blocks_;
  uint64_t content_;
  std::vector<rt_buf*> channels_;
  rt_buf* far_ = nullptr;
  rt_buf* slot_block_ = nullptr;
  // pinned staging: [payloads | metadata]
  uint8_t* host_stage_ = nullptr;
  uint8_t* dev_stage_ = nullptr;
  uint64_t stage_bytes_ = 0;
  std::vector<int32_t> staged_slots_;
  std::vector<uint32_t> staged_tree_blocks_;
  std::vector<std::pair<uint64_t, uint32_t>> link_writes_;
  std::vector<std::pair<uint32_t, uint32_t>> far_writes_;
  std::vector<std::pair<uint8_t*, const uint8_t*>> pending_copies_;  // staging payload <- stable source
  // Slots staged in this batch: the index of their (latest) staged copy, and
  // the number of link writes issued before it (those into the slot are dead).
  std::unordered_map<int32_t, std::pair<size_t, size_t>> staged_at_;
  bool in_flight_ = false;
  // Persistent copy threads (started by the first large batch): once a batch
  // holds kStreamAfter payload copies they copy each one as it is staged, the
  // flushing thread joins them, and each chunk of kChunkBlocks payloads is
  // queued for transfer as soon as it is copied, so the copies overlap the
  // rest of the transaction and the transfer of earlier chunks.
  static constexpr int kChunkBlocks = 16;
  static constexpr size_t kStreamAfter = 8;
  struct CopyWorkers;
  std::unique_ptr<CopyWorkers> workers_;
  std::vector<int> chunk_expected_;  // per chunk of the batch: its copies made by the copy threads
  int chunks_issued_ = 0;            // chunks of the batch already transferred
  int CopyAndTransferPayloads(int nb);
  int IssueChunks(int upto);
  int WaitStaging();
};

class BlockCache {
 public:
  BlockCache(rt_ctx* ctx, const StoredOctreeHeader& header, int max_blocks);

  BlockPool& pool() { return pool_; }
  bool full() const { return first_free_ == max_blocks_; }
  int loaded_blocks_count() const { return first_free_; }
  int max_blocks_count() const { return max_blocks_; }
  bool IsBlockInCache(uint32_t block) const { return block < block_to_slot_.size() && block_to_slot_[block] >= 0; }
  bool IsParentInCache(uint32_t block) const {
    return block < block_to_parent_slot_.size() && block_to_parent_slot_[block] >= 0;
  }
  int BlockIndexInCache(uint32_t block) const { return IsBlockInCache(block) ? block_to_slot_[block] : -1; }
  int SlotBlock(int slot) const { return (slot >= 0 && slot < first_free_) ? info_[slot].block_index : -1; }

  void MarkBlockAsUsed(int slot);
  void MarkParentAsUsed(int index, bool global_index);

  // LRU markers (LRUMarker of gpu-octree-cache.h:14-30).
  int InsertLRUMarker();
  bool MarkerHit(int marker) const;
  void ReleaseMarker(int marker);

  // Uploads a block (stored link format in `payload`): picks a slot, evicting
  // the least recently used block when full, stages the bytes and the parent
  // pointer rewrites. Returns the slot. A payload that stays valid until the
  // pool's next Flush (payload_stable) is copied to staging at the flush.
  // `children`: the block's child blocks (the pointers of its fault links, in
  // link order) when the caller has them already; else the payload's links
  // are scanned for them here.
  int UploadBlock(const StoredOctreeBlockHeader& h, const uint8_t* payload, bool payload_stable = false,
                  const std::vector<uint32_t>* children = nullptr);
  std::vector<int> LruOrder() const;

  uint64_t uploaded = 0, evicted = 0, patches = 0, far_writes = 0;

 private:
  struct SlotInfo {
    int32_t block_index = -1;
    int32_t parent_slot = -1;
    int roots_count = 0;
    uint16_t ptr_index[8];
    uint32_t ptr_value[8];
    std::vector<int32_t> child_blocks;
  };
  int nib_, max_blocks_;
  BlockPool pool_;
  std::vector<int32_t> block_to_slot_, block_to_parent_slot_;
  std::vector<SlotInfo> info_;
  int first_free_ = 0;
  // Intrusive LRU list. Ids [0, max_blocks) are slots; larger ids are markers.
  // head_ is least recently used.
  std::vector<int32_t> prev_, next_;
  std::vector<uint8_t> marker_hit_;   // per marker id - max_blocks
  std::vector<int32_t> free_markers_;
  int32_t head_ = -1, tail_ = -1;
  bool IsMarker(int32_t id) const { return id >= max_blocks_; }
  void Unlink(int32_t id);
  void PushBack(int32_t id);
  void PushFront(int32_t id);
  void DequeueMarkers();
  int UnloadLRUBlock();
};

// A fixed set of worker threads that run the items of one job at a time
// together with the calling thread.
class ParallelJobs {
 public:
  explicit ParallelJobs(int workers);
  ~ParallelJobs();
  ParallelJobs(const ParallelJobs&) = delete;
  ParallelJobs& operator=(const ParallelJobs&) = delete;
  // Calls fn(i) for every i in [0, n) on the workers and the caller; returns when all have returned.
  void Run(size_t n, const std::function<void(size_t)>& fn);

 private:
  struct State;
  std::unique_ptr<State> st_;
};

class BlockCacheManager {
 public:
  BlockCacheManager(rt_ctx* ctx, BlockSource* source, int max_blocks);

  BlockCache& cache() { return cache_; }
  const BlockCache& cache() const { return cache_; }
  BlockSource* source() { return source_; }
  int root_node_index() const { return root_node_index_; }
  int nodes_in_block() const { return int(source_->header().nodes_in_block); }

  int InitialFillCache();
  int StartRequestSession();
  int StartRequestTransaction();
  int MarkBlockAsUsed(int slot);
  int MarkParentAsUsed(int slot);
  int RequestBlock(int block_index);
  bool TransactionFilledCache() const { return cache_.MarkerHit(transaction_marker_); }
  bool SessionFilledCache() const { return cache_.MarkerHit(session_marker_); }
  // wait_for_device = false leaves the transaction's uploads queued on the
  // context stream (work queued after them on that stream sees them; other
  // streams must synchronise); the reference's manager waits (WaitForAll).
  int EndRequestTransaction(bool wait_for_device = true);
  int EndRequestSession();
  bool session_in_progress() const { return session_; }
  bool transaction_in_progress() const { return transaction_; }

 private:
  rt_ctx* ctx_;
  BlockSource* source_;
  BlockCache cache_;
  int root_node_index_ = -1;
  bool session_ = false, transaction_ = false;
  int session_marker_ = -1, transaction_marker_ = -1;
  std::vector<int> transaction_blocks_;
  StoredOctreeBlock scratch_;
  bool fetched_in_place_ = false;  // the last FetchBlock points into the source's memory
  std::unique_ptr<ParallelJobs> scan_;  // child-block scans of a transaction's blocks
  std::vector<std::vector<uint32_t>> children_;
  const uint8_t* FetchBlock(uint32_t index);
  int LoadAndUpload(uint32_t index, const std::vector<uint32_t>* children = nullptr);
};

}  // namespace rayt

struct rt_cache {
  rt_ctx* ctx = nullptr;
  rt_tree* tree = nullptr;
  std::unique_ptr<rayt::BlockCacheManager> mgr;
  rt_tracer* pending = nullptr;  // a tracer with a pipelined frame to finish before any use
};
