// Block pool, LRU block cache and cache manager (see cache.h).
#include "cache.h"

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <queue>
#include <stdexcept>
#include <thread>
#include <unordered_set>

#include "pool_batch.h"

namespace rayt {

namespace {
rt_buf* MakeBuf(rt_ctx* ctx, uint64_t bytes) {
  rt_buf* b = nullptr;
  if (rt_buf_create(ctx, bytes, &b) != RT_OK) throw std::runtime_error(std::string("pool allocation: ") + rt_last_error());
  return b;
}
void Check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
}  // namespace

// --------------------------------------------------------------- BlockPool
BlockPool::BlockPool(rt_ctx* ctx, const StoredOctreeHeader& header, int max_blocks, int staging_blocks)
    : ctx_(ctx), header_(header), nib_(int(header.nodes_in_block)), max_blocks_(max_blocks),
      staging_blocks_(staging_blocks) {
  if (header.channels.empty() || header.channels[0].name != kNodeLinkChannelName ||
      header.channels[0].bytes_in_node != kNodeLinkSize)
    throw std::runtime_error("channel 0 must be 4-byte node links");
  if (header.channels.size() > kMaxPoolChannels) throw std::runtime_error("too many channels");
  content_ = uint64_t(header.SumBytesInNode()) * nib_;
  try {
    for (auto& c : header.channels) {
      rt_buf* b = MakeBuf(ctx, uint64_t(max_blocks) * nib_ * c.bytes_in_node);
      channels_.push_back(b);
      rt_buf_fill_u32(b, 0, b->size / 4, 0u);
    }
    far_ = MakeBuf(ctx, uint64_t(max_blocks) * 8 * 4);
    rt_buf_fill_u32(far_, 0, uint64_t(max_blocks) * 8, 0u);
    slot_block_ = MakeBuf(ctx, uint64_t(max_blocks) * 4);
    rt_buf_fill_u32(slot_block_, 0, uint64_t(max_blocks), 0xffffffffu);
    const uint64_t lw_cap = 16ull * staging_blocks + 16, fw_cap = 8ull * staging_blocks + 8;
    stage_bytes_ = uint64_t(staging_blocks) * content_ + uint64_t(staging_blocks) * 8 + lw_cap * 12 + fw_cap * 8 + 256;
    Check(hipHostMalloc(reinterpret_cast<void**>(&host_stage_), stage_bytes_, hipHostMallocDefault), "hipHostMalloc");
    Check(hipMalloc(reinterpret_cast<void**>(&dev_stage_), stage_bytes_), "hipMalloc staging");
    pending_copies_.reserve(size_t(staging_blocks));
    chunk_expected_.assign(size_t(staging_blocks + kChunkBlocks - 1) / kChunkBlocks, 0);
  } catch (...) {
    for (rt_buf* b : channels_) rt_buf_destroy(b);
    rt_buf_destroy(far_);
    rt_buf_destroy(slot_block_);
    if (host_stage_) hipHostFree(host_stage_);
    if (dev_stage_) hipFree(dev_stage_);
    throw;
  }
  Check(hipStreamSynchronize(ctx_->stream), "pool init");
}

BlockPool::~BlockPool() {
  workers_.reset();
  (void)hipStreamSynchronize(ctx_->stream);  // best effort, as a destructor
  for (rt_buf* b : channels_) rt_buf_destroy(b);
  rt_buf_destroy(far_);
  rt_buf_destroy(slot_block_);
  if (host_stage_) hipHostFree(host_stage_);
  if (dev_stage_) hipFree(dev_stage_);
}

int BlockPool::WaitStaging() {
  if (!in_flight_) return RT_OK;
  in_flight_ = false;
  RT_HIP_CHECK(hipStreamSynchronize(ctx_->stream));
  return RT_OK;
}

uint8_t* BlockPool::StageBlock(int slot, uint32_t tree_block) {
  if (slot < 0 || slot >= max_blocks_) throw std::runtime_error("slot out of range");
  if (int(staged_slots_.size()) == staging_blocks_ && Flush() != RT_OK) throw std::runtime_error(rt_last_error());
  if (WaitStaging() != RT_OK) throw std::runtime_error(rt_last_error());
  // An earlier copy of this slot in the batch is dead, and so are the link writes into the slot issued
  // before this copy (dropped at the flush).
  auto it = staged_at_.find(slot);
  if (it != staged_at_.end()) staged_slots_[it->second.first] = -1;
  staged_at_[slot] = {staged_slots_.size(), link_writes_.size()};
  uint8_t* p = host_stage_ + uint64_t(staged_slots_.size()) * content_;
  staged_slots_.push_back(slot);
  staged_tree_blocks_.push_back(tree_block);
  return p;
}

// Copy threads shared by the batches of one pool. A job is the list of a
// batch's payload copies, open while the batch is staged: StageBlockCopy
// publishes each copy as it is staged, so the threads copy while the
// transaction goes on, and the flush closes the list, joins in, and queues
// each chunk of kChunkBlocks payloads for transfer once its copies are done.
// Every participant takes the next copy with one atomic, so copies finish
// roughly in staging order; a thread whose index is past the published ones
// waits for more, and leaves once the job is closed below its index.
struct BlockPool::CopyWorkers {
  std::vector<std::thread> threads;
  std::mutex mu;
  std::condition_variable cv;
  uint64_t generation = 0;
  bool stop = false;
  // the current job
  const std::pair<uint8_t*, const uint8_t*>* items = nullptr;  // fixed for the job (reserved vector)
  uint64_t content = 0;
  const uint8_t* base = nullptr;  // host staging: an item's chunk is (dst - base) / (content * kChunkBlocks)
  std::vector<std::atomic<int>> chunk_done;
  std::atomic<size_t> next{0};
  std::atomic<size_t> avail{0};     // copies published
  std::atomic<bool> closed{false};  // avail is final
  std::atomic<int> active{0};       // workers inside the job
  bool open = false;                // a job is running (owner thread only)

  CopyWorkers(int count, size_t max_chunks) : chunk_done(max_chunks) {
    for (int i = 0; i < count; ++i) threads.emplace_back([this] { Run(); });
  }
  ~CopyWorkers() {
    Close(avail.load());  // drains what was published
    {
      std::lock_guard<std::mutex> lk(mu);
      stop = true;
    }
    cv.notify_all();
    for (std::thread& t : threads) t.join();
  }
  int ChunkOf(size_t i) const { return int(uint64_t(items[i].first - base) / (content * kChunkBlocks)); }
  void Begin(const std::pair<uint8_t*, const uint8_t*>* list, uint64_t bytes, const uint8_t* stage, size_t published) {
    {
      std::lock_guard<std::mutex> lk(mu);
      items = list;
      content = bytes;
      base = stage;
      for (auto& a : chunk_done) a.store(0, std::memory_order_relaxed);
      next.store(0);
      avail.store(published, std::memory_order_release);
      closed.store(false);
      active.store(int(threads.size()));
      ++generation;
    }
    open = true;
    cv.notify_all();
  }
  void Publish(size_t n) { avail.store(n, std::memory_order_release); }
  void Close(size_t n) {
    avail.store(n, std::memory_order_release);
    closed.store(true, std::memory_order_release);
  }
  // The index of the next copy to make, or false once the job is closed below it.
  bool Take(size_t& i) {
    i = next.fetch_add(1, std::memory_order_relaxed);
    while (i >= avail.load(std::memory_order_acquire)) {
      if (closed.load(std::memory_order_acquire) && i >= avail.load(std::memory_order_acquire)) return false;
      std::this_thread::yield();
    }
    return true;
  }
  void CopyOne(size_t i) {
    memcpy(items[i].first, items[i].second, content);
    chunk_done[size_t(ChunkOf(i))].fetch_add(1, std::memory_order_release);
  }
  void Run() {
    uint64_t seen = 0;
    for (;;) {
      {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return stop || generation != seen; });
        if (stop) return;
        seen = generation;
      }
      for (size_t i; Take(i);) CopyOne(i);
      active.fetch_sub(1, std::memory_order_acq_rel);
    }
  }
};

void BlockPool::StageBlockCopy(int slot, uint32_t tree_block, const uint8_t* src, bool src_stable) {
  uint8_t* dst = StageBlock(slot, tree_block);
  if (!src_stable) {
    memcpy(dst, src, content_);
    return;
  }
  pending_copies_.push_back({dst, src});  // capacity staging_blocks_: the job's list never moves
  ++chunk_expected_[size_t(staged_slots_.size() - 1) / kChunkBlocks];
  const size_t n = pending_copies_.size();
  if (workers_ && workers_->open) {
    workers_->Publish(n);
    // the chunks staged in full whose copies are done cross to the device while staging goes on
    if (IssueChunks(int(staged_slots_.size()) / kChunkBlocks) != RT_OK) throw std::runtime_error(rt_last_error());
  } else if (n == kStreamAfter) {  // a large batch: the copy threads start on it now
    if (!workers_) {
      const int hw = int(std::max(2u, std::thread::hardware_concurrency()));
      workers_.reset(new CopyWorkers(std::min(7, hw - 1), chunk_expected_.size()));
    }
    workers_->Begin(pending_copies_.data(), content_, host_stage_, n);
  }
}

// Transfers, in order, chunks [chunks_issued_, upto) whose copies are all done.
int BlockPool::IssueChunks(int upto) {
  const uint64_t end = uint64_t(staged_slots_.size()) * content_;
  while (chunks_issued_ < upto && workers_->chunk_done[size_t(chunks_issued_)].load(std::memory_order_acquire) ==
                                      chunk_expected_[size_t(chunks_issued_)]) {
    const uint64_t b = uint64_t(chunks_issued_) * kChunkBlocks * content_;
    const uint64_t e = std::min<uint64_t>(end, b + uint64_t(kChunkBlocks) * content_);
    RT_HIP_CHECK(hipMemcpyAsync(dev_stage_ + b, host_stage_ + b, e - b, hipMemcpyHostToDevice, ctx_->stream));
    ++chunks_issued_;
  }
  return RT_OK;
}

// The payload copies of a batch (a few MB per transaction out of core) and
// the transfers of the payload area [0, nb * content) to the device staging
// buffer, chunk by chunk as the chunks are filled.
int BlockPool::CopyAndTransferPayloads(int nb) {
  const size_t n = pending_copies_.size();
  const int nchunks = (nb + kChunkBlocks - 1) / kChunkBlocks;
  std::fill(chunk_expected_.begin(), chunk_expected_.begin() + nchunks, 0);
  if (!workers_ || !workers_->open) {  // a small batch: copy here, one transfer
    for (auto& cp : pending_copies_) memcpy(cp.first, cp.second, content_);
    pending_copies_.clear();
    if (nb == 0) return RT_OK;
    RT_HIP_CHECK(hipMemcpyAsync(dev_stage_, host_stage_, uint64_t(nb) * content_, hipMemcpyHostToDevice,
                                ctx_->stream));
    return RT_OK;
  }
  CopyWorkers& w = *workers_;
  w.Close(n);
  for (size_t i = 0; i < n; ++i) ++chunk_expected_[size_t(w.ChunkOf(i))];  // recounted: cleared above
  int rc = RT_OK;
  for (size_t i; w.Take(i);) {
    w.CopyOne(i);
    if (rc == RT_OK) rc = IssueChunks(nchunks);
  }
  while (w.active.load(std::memory_order_acquire) != 0) {  // the workers' last copies
    if (rc == RT_OK) rc = IssueChunks(nchunks);
    std::this_thread::yield();
  }
  if (rc == RT_OK) rc = IssueChunks(nchunks);
  const int issued = chunks_issued_;
  w.open = false;
  chunks_issued_ = 0;
  std::fill(chunk_expected_.begin(), chunk_expected_.begin() + nchunks, 0);
  pending_copies_.clear();
  return rc != RT_OK ? rc : (issued == nchunks ? RT_OK : Fail(RT_ERR_STATE, "payload chunks not transferred"));
}

void BlockPool::StageLinkWrite(uint64_t node, uint32_t value) {
  if (link_writes_.size() >= 16ull * staging_blocks_ + 16 && Flush() != RT_OK) throw std::runtime_error(rt_last_error());
  link_writes_.push_back({node, value});
}

void BlockPool::StageFarPointer(uint32_t index, uint32_t value) {
  if (far_writes_.size() >= 8ull * staging_blocks_ + 8 && Flush() != RT_OK) throw std::runtime_error(rt_last_error());
  far_writes_.push_back({index, value});
}

int BlockPool::Flush() {
  if (empty()) return RT_OK;
  int rc = WaitStaging();
  if (rc) return rc;
  // Last write wins per address; keep issue order for the survivors.
  auto dedupe = [](auto& v) {
    using T = typename std::decay<decltype(v)>::type::value_type;
    std::unordered_set<uint64_t> seen;
    std::vector<T> out;
    out.reserve(v.size());
    for (auto it = v.rbegin(); it != v.rend(); ++it)
      if (seen.insert(uint64_t(it->first)).second) out.push_back(*it);
    std::reverse(out.begin(), out.end());
    v.swap(out);
  };
  if (!staged_at_.empty()) {  // writes into a slot staged after them are overwritten by the slot's copy
    size_t k = 0;
    for (size_t i = 0; i < link_writes_.size(); ++i) {
      auto it = staged_at_.find(int32_t(link_writes_[i].first / uint64_t(nib_)));
      if (it == staged_at_.end() || i >= it->second.second) link_writes_[k++] = link_writes_[i];
    }
    link_writes_.resize(k);
  }
  dedupe(link_writes_);
  dedupe(far_writes_);
  const int nb = int(staged_slots_.size());
  const int nlw = int(link_writes_.size()), nfw = int(far_writes_.size());
  uint64_t off = uint64_t(nb) * content_;
  off = (off + 15) & ~uint64_t(15);
  const uint64_t off_slots = off;
  off += uint64_t(nb) * 4;
  const uint64_t off_tb = off;
  off += uint64_t(nb) * 4;
  off = (off + 7) & ~uint64_t(7);
  const uint64_t off_lwa = off;
  off += uint64_t(nlw) * 8;
  const uint64_t off_lwv = off;
  off += uint64_t(nlw) * 4;
  const uint64_t off_fwi = off;
  off += uint64_t(nfw) * 4;
  const uint64_t off_fwv = off;
  off += uint64_t(nfw) * 4;
  if (off > stage_bytes_) return Fail(RT_ERR_STATE, "staging overflow");
  rc = CopyAndTransferPayloads(nb);
  if (rc) return rc;
  memcpy(host_stage_ + off_slots, staged_slots_.data(), size_t(nb) * 4);
  memcpy(host_stage_ + off_tb, staged_tree_blocks_.data(), size_t(nb) * 4);
  for (int i = 0; i < nlw; ++i) {
    memcpy(host_stage_ + off_lwa + size_t(i) * 8, &link_writes_[i].first, 8);
    memcpy(host_stage_ + off_lwv + size_t(i) * 4, &link_writes_[i].second, 4);
  }
  for (int i = 0; i < nfw; ++i) {
    memcpy(host_stage_ + off_fwi + size_t(i) * 4, &far_writes_[i].first, 4);
    memcpy(host_stage_ + off_fwv + size_t(i) * 4, &far_writes_[i].second, 4);
  }
  RT_HIP_CHECK(hipMemcpyAsync(dev_stage_ + off_slots, host_stage_ + off_slots, off - off_slots,
                              hipMemcpyHostToDevice, ctx_->stream));
  in_flight_ = true;
  PoolBatchArgs a;
  a.stage = dev_stage_;
  a.nblocks = nb;
  a.content = content_;
  a.nib = nib_;
  a.nchan = int(channels_.size());
  for (int c = 0; c < a.nchan; ++c) {
    a.chan[c] = static_cast<uint8_t*>(channels_[c]->ptr);
    a.chan_bytes[c] = header_.channels[c].bytes_in_node;
    a.chan_off[c] = header_.ChannelOffsetInBlock(c);
  }
  a.slots = reinterpret_cast<const int32_t*>(dev_stage_ + off_slots);
  a.tree_blocks = reinterpret_cast<const uint32_t*>(dev_stage_ + off_tb);
  a.lw_addr = reinterpret_cast<const uint64_t*>(dev_stage_ + off_lwa);
  a.lw_val = reinterpret_cast<const uint32_t*>(dev_stage_ + off_lwv);
  a.nlw = nlw;
  a.fw_idx = reinterpret_cast<const uint32_t*>(dev_stage_ + off_fwi);
  a.fw_val = reinterpret_cast<const uint32_t*>(dev_stage_ + off_fwv);
  a.nfw = nfw;
  a.far_ptrs = static_cast<uint32_t*>(far_->ptr);
  a.slot_block = static_cast<int32_t*>(slot_block_->ptr);
  rc = LaunchPoolBatch(ctx_->stream, a);
  staged_slots_.clear();
  staged_tree_blocks_.clear();
  staged_at_.clear();
  link_writes_.clear();
  far_writes_.clear();
  return rc;
}

// -------------------------------------------------------------- BlockCache
static int StagingBlocksFor(const StoredOctreeHeader& h, int max_blocks) {
  const uint64_t content = uint64_t(h.SumBytesInNode()) * h.nodes_in_block;
  uint64_t n = (64ull << 20) / std::max<uint64_t>(content, 1);
  n = std::max<uint64_t>(n, 4);
  n = std::min<uint64_t>(n, uint64_t(max_blocks));
  return int(std::max<uint64_t>(n, 1));
}

BlockCache::BlockCache(rt_ctx* ctx, const StoredOctreeHeader& header, int max_blocks)
    : nib_(int(header.nodes_in_block)), max_blocks_(max_blocks),
      pool_(ctx, header, max_blocks, StagingBlocksFor(header, max_blocks)) {
  block_to_slot_.assign(header.blocks_count, -1);
  block_to_parent_slot_.assign(header.blocks_count, -1);
  info_.resize(max_blocks);
  prev_.assign(max_blocks, -1);
  next_.assign(max_blocks, -1);
}

void BlockCache::Unlink(int32_t id) {
  const int32_t p = prev_[id], n = next_[id];
  if (p >= 0) next_[p] = n; else head_ = n;
  if (n >= 0) prev_[n] = p; else tail_ = p;
  prev_[id] = next_[id] = -1;
}
void BlockCache::PushBack(int32_t id) {
  prev_[id] = tail_;
  next_[id] = -1;
  if (tail_ >= 0) next_[tail_] = id; else head_ = id;
  tail_ = id;
}
void BlockCache::PushFront(int32_t id) {
  next_[id] = head_;
  prev_[id] = -1;
  if (head_ >= 0) prev_[head_] = id; else tail_ = id;
  head_ = id;
}

void BlockCache::DequeueMarkers() {
  while (head_ >= 0 && IsMarker(head_)) {
    const int32_t m = head_;
    marker_hit_[m - max_blocks_] = 1;
    Unlink(m);
  }
}

void BlockCache::MarkBlockAsUsed(int slot) {
  // Touch the block and all its ancestors, child first, so a parent always
  // stays nearer the most-recently-used end than its children.
  for (int s = slot; uint32_t(s) != kRootBlockParent; s = info_[s].parent_slot) {
    if (s < 0 || s >= first_free_) throw std::runtime_error("MarkBlockAsUsed: slot not loaded");
    Unlink(s);
    PushBack(s);
  }
  DequeueMarkers();
}

void BlockCache::MarkParentAsUsed(int index, bool global_index) {
  if (global_index) {
    if (!IsParentInCache(uint32_t(index))) throw std::runtime_error("MarkParentAsUsed: parent not in cache");
    MarkBlockAsUsed(block_to_parent_slot_[index]);
  } else {
    if (index < 0 || index >= first_free_) throw std::runtime_error("MarkParentAsUsed: slot out of range");
    MarkBlockAsUsed(info_[index].parent_slot);
  }
}

int BlockCache::InsertLRUMarker() {
  int32_t id;
  if (!free_markers_.empty()) {
    id = free_markers_.back();
    free_markers_.pop_back();
  } else {
    id = int32_t(prev_.size());
    prev_.push_back(-1);
    next_.push_back(-1);
    marker_hit_.push_back(0);
  }
  marker_hit_[id - max_blocks_] = 0;
  PushBack(id);
  return id;
}

bool BlockCache::MarkerHit(int marker) const {
  if (marker < max_blocks_) return false;
  return marker_hit_[marker - max_blocks_] != 0;
}

void BlockCache::ReleaseMarker(int marker) {
  if (marker < max_blocks_) return;
  if (!marker_hit_[marker - max_blocks_]) Unlink(marker);
  free_markers_.push_back(marker);
}

int BlockCache::UnloadLRUBlock() {
  if (head_ < 0 || IsMarker(head_)) throw std::runtime_error("LRU head is not a block");
  const int slot = head_;
  SlotInfo& info = info_[slot];
  const int parent = info.parent_slot;
  if (uint32_t(parent) == kRootBlockParent) throw std::runtime_error("cannot evict the root block");
  for (int i = 0; i < info.roots_count; ++i) {
    uint32_t v = info.ptr_value[i];
    if (LinkMask(v) && !LinkFault(v))
      v = PackCacheNodeLink(LinkMask(v), false, LinkDup(v), false, NearPointerValue(info.ptr_index[i], LinkPtr(v)));
    pool_.StageLinkWrite(uint64_t(parent) * nib_ + info.ptr_index[i], v);
    ++patches;
  }
  for (int32_t c : info.child_blocks) block_to_parent_slot_[c] = -1;
  info.child_blocks.clear();
  Unlink(slot);
  block_to_slot_[info.block_index] = -1;
  DequeueMarkers();
  ++evicted;
  return slot;
}

int BlockCache::UploadBlock(const StoredOctreeBlockHeader& h, const uint8_t* payload, bool payload_stable,
                            const std::vector<uint32_t>* children) {
  if (h.block_index >= block_to_slot_.size()) throw std::runtime_error("block index out of range");
  if (IsBlockInCache(h.block_index)) throw std::runtime_error("block already in cache");
  int slot;
  if (first_free_ == max_blocks_) {
    if (h.parent_block_index != kRootBlockParent) MarkParentAsUsed(int(h.block_index), true);
    slot = UnloadLRUBlock();
  } else {
    slot = first_free_++;
  }
  SlotInfo& info = info_[slot];
  // Enumerate child blocks (fault links); link conversion happens on the device.
  auto add_child = [&](uint32_t child) {
    if (child >= block_to_parent_slot_.size()) throw std::runtime_error("fault link outside tree");
    if (block_to_parent_slot_[child] < 0) {
      block_to_parent_slot_[child] = slot;
      info.child_blocks.push_back(int32_t(child));
    }
  };
  if (children)
    for (uint32_t c : *children) add_child(c);
  else
    ForEachFaultLink(reinterpret_cast<const uint32_t*>(payload), nib_, [&](int, uint32_t l) { add_child(LinkPtr(l)); });
  pool_.StageBlockCopy(slot, h.block_index, payload, payload_stable);
  ++uploaded;
  if (h.parent_block_index != kRootBlockParent) {
    if (!IsParentInCache(h.block_index)) throw std::runtime_error("parent block not in cache");
    const int parent = block_to_parent_slot_[h.block_index];
    info.parent_slot = parent;
    info.roots_count = int(h.roots_count);
    for (uint32_t i = 0; i < h.roots_count && i < 8; ++i) {
      const StoredOctreeBlockRoot& r = h.roots[i];
      info.ptr_index[i] = r.parent_pointer_index;
      info.ptr_value[i] = r.parent_pointer_value;
      const int64_t absolute = int64_t(slot) * nib_ + r.pointed_child_index;
      const int64_t at = int64_t(parent) * nib_ + r.parent_pointer_index;
      bool far_flag;
      uint32_t ptr;
      if (NeedFarPointer(at, absolute)) {
        ptr = uint32_t(slot) * 8 + i;
        far_flag = true;
        pool_.StageFarPointer(ptr, uint32_t(absolute));
        ++far_writes;
      } else {
        far_flag = false;
        ptr = NearPointerValue(at, absolute);
      }
      // A root whose stored parent pointer is not a fault is a duplicate of a parent-block node.
      const uint32_t v = PackCacheNodeLink(LinkMask(r.parent_pointer_value), false, !LinkFault(r.parent_pointer_value),
                                           far_flag, ptr);
      pool_.StageLinkWrite(uint64_t(at), v);
      ++patches;
    }
  } else {
    info.parent_slot = int32_t(kRootBlockParent);
    info.roots_count = 0;
  }
  block_to_slot_[h.block_index] = slot;
  info.block_index = int32_t(h.block_index);
  PushFront(slot);
  MarkBlockAsUsed(slot);
  return slot;
}

std::vector<int> BlockCache::LruOrder() const {
  std::vector<int> out;
  for (int32_t id = head_; id >= 0; id = next_[id]) out.push_back(IsMarker(id) ? -1 : id);
  return out;
}

// ------------------------------------------------------------ ParallelJobs
struct ParallelJobs::State {
  std::vector<std::thread> threads;
  std::mutex mu;
  std::condition_variable cv, done_cv;
  uint64_t generation = 0;
  bool stop = false;
  const std::function<void(size_t)>* fn = nullptr;
  size_t n = 0;
  std::atomic<size_t> next{0};
  int active = 0;  // workers inside the current job

  void Work() {
    for (size_t i; (i = next.fetch_add(1)) < n;) (*fn)(i);
  }
  void Loop() {
    uint64_t seen = 0;
    for (;;) {
      {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return stop || generation != seen; });
        if (stop) return;
        seen = generation;
      }
      Work();
      {
        std::lock_guard<std::mutex> lk(mu);
        --active;
      }
      done_cv.notify_all();
    }
  }
};

ParallelJobs::ParallelJobs(int workers) : st_(new State()) {
  for (int i = 0; i < workers; ++i) st_->threads.emplace_back([s = st_.get()] { s->Loop(); });
}

ParallelJobs::~ParallelJobs() {
  {
    std::lock_guard<std::mutex> lk(st_->mu);
    st_->stop = true;
  }
  st_->cv.notify_all();
  for (std::thread& t : st_->threads) t.join();
}

void ParallelJobs::Run(size_t n, const std::function<void(size_t)>& fn) {
  State& s = *st_;
  {
    std::lock_guard<std::mutex> lk(s.mu);
    s.fn = &fn;
    s.n = n;
    s.next.store(0);
    s.active = int(s.threads.size());
    ++s.generation;
  }
  s.cv.notify_all();
  s.Work();
  std::unique_lock<std::mutex> lk(s.mu);
  s.done_cv.wait(lk, [&] { return s.active == 0; });
}

// ------------------------------------------------------- BlockCacheManager
BlockCacheManager::BlockCacheManager(rt_ctx* ctx, BlockSource* source, int max_blocks)
    : ctx_(ctx), source_(source), cache_(ctx, source->header(), max_blocks) {}

// The block's header and payload: read into scratch_, or, from a source that
// keeps records in memory, the header decoded and the payload read in place.
const uint8_t* BlockCacheManager::FetchBlock(uint32_t index) {
  if (const uint8_t* rec = source_->RecordPtr(index)) {
    DecodeBlockHeader(rec, &scratch_.header);
    fetched_in_place_ = true;
    return rec + kBlockHeaderSize;
  }
  fetched_in_place_ = false;
  return source_->LoadBlock(index, &scratch_) ? scratch_.data.data() : nullptr;
}

int BlockCacheManager::LoadAndUpload(uint32_t index, const std::vector<uint32_t>* children) {
  const uint8_t* payload = FetchBlock(index);
  if (!payload) return Fail(RT_ERR_IO, "failed to load block");
  const bool its_root = scratch_.header.parent_block_index == kRootBlockParent;
  const int root_in_block = scratch_.header.roots[0].pointed_child_index;
  const int slot = cache_.UploadBlock(scratch_.header, payload, fetched_in_place_, children);
  if (its_root) root_node_index_ = slot * nodes_in_block() + root_in_block;
  return RT_OK;
}

int BlockCacheManager::InitialFillCache() {
  RT_TRY_BEGIN
  std::queue<uint32_t> q;
  q.push(source_->header().root_block_index);
  const int nib = nodes_in_block();
  while (!cache_.full() && !q.empty()) {
    const uint32_t index = q.front();
    q.pop();
    if (cache_.IsBlockInCache(index)) continue;
    const uint8_t* payload = FetchBlock(index);
    if (!payload) return Fail(RT_ERR_IO, "failed to load block");
    const bool its_root = scratch_.header.parent_block_index == kRootBlockParent;
    const int root_in_block = scratch_.header.roots[0].pointed_child_index;
    ForEachFaultLink(reinterpret_cast<const uint32_t*>(payload), nib, [&](int, uint32_t l) { q.push(LinkPtr(l)); });
    const int slot = cache_.UploadBlock(scratch_.header, payload, fetched_in_place_);
    if (its_root) root_node_index_ = slot * nib + root_in_block;
  }
  int rc = cache_.pool().Flush();
  if (rc) return rc;
  RT_HIP_CHECK(hipStreamSynchronize(ctx_->stream));
  return RT_OK;
  RT_TRY_END
}

int BlockCacheManager::StartRequestSession() {
  RT_REQUIRE(!session_, RT_ERR_STATE, "session already in progress");
  session_ = true;
  session_marker_ = cache_.InsertLRUMarker();
  return RT_OK;
}

int BlockCacheManager::StartRequestTransaction() {
  RT_REQUIRE(session_ && !transaction_, RT_ERR_STATE, "transaction needs a session and no open transaction");
  transaction_ = true;
  transaction_marker_ = cache_.InsertLRUMarker();
  return RT_OK;
}

int BlockCacheManager::MarkBlockAsUsed(int slot) {
  RT_REQUIRE(transaction_, RT_ERR_STATE, "no transaction in progress");
  RT_TRY_BEGIN
  cache_.MarkBlockAsUsed(slot);
  return RT_OK;
  RT_TRY_END
}

int BlockCacheManager::MarkParentAsUsed(int slot) {
  RT_REQUIRE(transaction_, RT_ERR_STATE, "no transaction in progress");
  RT_TRY_BEGIN
  cache_.MarkParentAsUsed(slot, false);
  return RT_OK;
  RT_TRY_END
}

int BlockCacheManager::RequestBlock(int block_index) {
  RT_REQUIRE(transaction_, RT_ERR_STATE, "no transaction in progress");
  RT_REQUIRE(block_index >= 0 && uint32_t(block_index) < source_->header().blocks_count, RT_ERR_RANGE,
             "block index out of range");
  if (cache_.IsBlockInCache(uint32_t(block_index))) return RT_OK;
  RT_TRY_BEGIN
  transaction_blocks_.push_back(block_index);  // loaded at transaction end, after all parents are marked
  cache_.MarkParentAsUsed(block_index, true);
  return RT_OK;
  RT_TRY_END
}

int BlockCacheManager::EndRequestTransaction(bool wait_for_device) {
  RT_REQUIRE(transaction_, RT_ERR_STATE, "no transaction in progress");
  RT_TRY_BEGIN
  // A transaction that loads nothing (a resident frame) enqueues no device work: no flush, no wait.
  const bool device_work = !transaction_blocks_.empty() || !cache_.pool().empty();
  // The child blocks of blocks held in memory are found by scanning their links (32 KB of cold memory per
  // 8192-node block): scanned for the whole transaction at once on several threads, then registered in
  // load order by the loop below, as UploadBlock would have found them.
  const size_t nt = transaction_blocks_.size();
  const bool prescan = nt >= 8 && source_->RecordPtr(uint32_t(transaction_blocks_[0])) != nullptr;
  if (prescan) {
    if (!scan_) scan_.reset(new ParallelJobs(std::min(7, int(std::max(2u, std::thread::hardware_concurrency())) - 1)));
    if (children_.size() < nt) children_.resize(nt);
    const int nib = nodes_in_block();
    scan_->Run(nt, [&](size_t i) {
      std::vector<uint32_t>& kids = children_[i];
      kids.clear();
      const uint8_t* rec = source_->RecordPtr(uint32_t(transaction_blocks_[i]));
      if (!rec) return;
      ForEachFaultLink(reinterpret_cast<const uint32_t*>(rec + kBlockHeaderSize), nib,
                       [&](int, uint32_t l) { kids.push_back(LinkPtr(l)); });
    });
  }
  for (size_t i = 0; i < nt; ++i) {
    if (TransactionFilledCache()) break;
    int rc = LoadAndUpload(uint32_t(transaction_blocks_[i]), prescan ? &children_[i] : nullptr);
    if (rc) return rc;
  }
  transaction_blocks_.clear();
  if (device_work) {
    int rc = cache_.pool().Flush();
    if (rc) return rc;
    if (wait_for_device) RT_HIP_CHECK(hipStreamSynchronize(ctx_->stream));
  }
  transaction_ = false;
  cache_.ReleaseMarker(transaction_marker_);
  transaction_marker_ = -1;
  return RT_OK;
  RT_TRY_END
}

int BlockCacheManager::EndRequestSession() {
  RT_REQUIRE(session_ && !transaction_, RT_ERR_STATE, "no session, or a transaction is open");
  session_ = false;
  cache_.ReleaseMarker(session_marker_);
  session_marker_ = -1;
  return RT_OK;
}

}  // namespace rayt
