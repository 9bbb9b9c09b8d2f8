// Feedback extraction (RayTracingFeedbackExtractor,
// GPU_renderer/ray-tracing-feedback-extractor.cpp:9-205) and the GPU
// sort/scan primitives (GPU algorithms/gpu-sort.cpp, gpu-scan.cpp).
#pragma once

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "mark_codes.h"
#include "rt_internal.h"

namespace rayt {

int SortI32(hipStream_t s, int* data, int n, int* temp);
int ScanI32(hipStream_t s, int* data, int n);

class FeedbackExtractor {
 public:
  FeedbackExtractor(rt_ctx* ctx, int count);
  ~FeedbackExtractor();
  FeedbackExtractor(const FeedbackExtractor&) = delete;
  FeedbackExtractor& operator=(const FeedbackExtractor&) = delete;

  int count() const { return count_; }
  // The stream the extraction runs on (the context's unless set); the codes
  // it reads must be complete on the device when Process is called.
  void set_stream(hipStream_t s) { stream_ = s ? s : ctx_->stream; }
  int hits_to_extract() const { return hits_to_extract_; }
  void set_hits_to_extract(int v) { hits_to_extract_ = v; }
  void set_seed(uint64_t s) { rng_ = s; }
  uint64_t seed_state() const { return rng_; }

  // `codes` is the device faults_and_hits buffer (count_ ints). With
  // block_level the hit codes are already ((slot+1)<<1)|dup.
  int Process(int* codes, int nodes_in_block, bool block_level);
  // The same over the first n (<= count) codes. premarked: the block-level
  // codes were already marked into Targets() by the renderer that wrote them
  // (plain kernel), so only the bitmap walk runs; the codes stay in memory for
  // the sort path.
  int ProcessN(int* codes, int n, int nodes_in_block, bool block_level, bool premarked = false);
  // ProcessN in two halves, for a caller that overlaps the host half with
  // later device work: EnqueueN issues the marking, the walk and the copy of
  // its result; CompleteN waits for that copy and reads the lists. The codes
  // must not change in between (the sort path may read them).
  //
  // Premarked walks may alternate between two sets of bitmaps (`set` 0 or
  // 1), so that a renderer marks the next frame's codes into one set while
  // the walk of the previous frame reads and clears the other; such a walk
  // runs on `walk_stream` (the extractor's own stream when null) after the
  // event `after` (the end of the trace that marked its set).
  int EnqueueN(int* codes, int n, int nodes_in_block, bool block_level, bool premarked, int set = 0,
               hipStream_t walk_stream = nullptr, hipEvent_t after = nullptr);
  int CompleteN();
  // Where a renderer marks codes for the next premarked Process: the bitmaps
  // of `set`, with the bounds of the codes the pool can produce (tree blocks
  // for faults, 2 * slots + 2 for hit codes).
  MarkTargets Targets(uint32_t fault_limit, uint32_t hit_limit, int set = 0);

  const std::vector<int>& fault_blocks() const { return faults_; }
  const std::vector<int>& hit_blocks() const { return hits_; }
  const std::vector<int>& duplicate_hit_blocks() const { return dup_hits_; }
  int unique_count() const { return unique_count_; }
  int expected_faults() const { return expected_faults_; }

 private:
  rt_ctx* ctx_;
  hipStream_t stream_;
  int count_;
  int hits_to_extract_ = 4000;
  int expected_faults_ = 1 << 16;
  uint64_t rng_ = 0;
  int unique_count_ = 0, fault_count_ = 0;
  // the walk issued by EnqueueN that CompleteN finishes
  hipEvent_t copied_ = nullptr;
  int* q_codes_ = nullptr;
  int q_n_ = 0, q_nib_ = 0;
  bool q_block_level_ = false, q_empty_ = true;
  std::vector<int> faults_, hits_, dup_hits_;
  uint32_t* fault_bits_ = nullptr;
  uint32_t* hit_bits_ = nullptr;
  // the second set of bitmaps (allocated, cleared, on first use)
  uint32_t* fault_bits2_ = nullptr;
  uint32_t* hit_bits2_ = nullptr;
  int EnsureSecondSet();
  void* state_ = nullptr;
  void* host_state_ = nullptr;
  // Pinned head of the unique list, read with the state (kPrefixInts codes).
  static constexpr int kPrefixInts = 16384;
  void* host_prefix_ = nullptr;
  int prefix_len_ = 0;
  int* unique_ = nullptr;
  int* temp_ = nullptr;
  int* index_ = nullptr;
  void* partials_ = nullptr;  // int4 per marking workgroup: max fault, max hit, NO_HIT seen, overflow
  int max_mark_blocks_ = 1;
  uint32_t premark_fault_limit_ = 0, premark_hit_limit_ = 2;
  std::vector<int> host_sorted_, window_;

  void AddHit(int encoded) {
    if (encoded & 1)
      dup_hits_.push_back((encoded >> 1) - 1);
    else
      hits_.push_back((encoded >> 1) - 1);
  }
  uint32_t NextRandom();
  int UniqueBySort(int* codes, int n);
  int ReadLists();
};

}  // namespace rayt

struct rt_feedback {
  std::unique_ptr<rayt::FeedbackExtractor> fx;
  rt_tracer* pending = nullptr;  // the owning tracer while it has a pipelined frame to finish
};
