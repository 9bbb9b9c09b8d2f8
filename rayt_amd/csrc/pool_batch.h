// Arguments of the block-pool upload batch (host side in cache.cpp, kernels in pool_batch.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rayt {

constexpr size_t kMaxPoolChannels = 8;

struct PoolBatchArgs {
  const uint8_t* stage = nullptr;  // device copy of the staged block payloads (stored format)
  int nblocks = 0;
  uint64_t content = 0;  // payload bytes per block
  int nib = 0;
  int nchan = 0;
  uint8_t* chan[kMaxPoolChannels] = {};
  int chan_bytes[kMaxPoolChannels] = {};
  uint64_t chan_off[kMaxPoolChannels] = {};  // channel offset inside a payload
  const int32_t* slots = nullptr;            // destination slot per staged block (-1: superseded)
  const uint32_t* tree_blocks = nullptr;
  const uint64_t* lw_addr = nullptr;  // node-link writes (node index, value), one per address
  const uint32_t* lw_val = nullptr;
  int nlw = 0;
  const uint32_t* fw_idx = nullptr;  // far-pointer writes, one per index
  const uint32_t* fw_val = nullptr;
  int nfw = 0;
  uint32_t* far_ptrs = nullptr;
  int32_t* slot_block = nullptr;  // device page table: slot -> tree block
};

// Enqueues: (1) per staged block and node, stored->cache link conversion and
// scatter of every channel into the slot; (2) the link, far-pointer and page
// table writes. Returns RT_OK or an error code.
int LaunchPoolBatch(hipStream_t stream, const PoolBatchArgs& a);

}  // namespace rayt
