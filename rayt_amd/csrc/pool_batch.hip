// Device side of block uploads: link conversion + channel scatter, then the
// batch's pointer writes (gpu-octree-cache.cpp:121-141 for the conversion,
// :162-224 and :259-278 for the parent pointer rewrites computed on the host).
#include "pool_batch.h"
#include "rt_internal.h"

namespace rayt {

namespace {

__global__ __launch_bounds__(256) void ScatterBlocksKernel(PoolBatchArgs a) {
  const int b = blockIdx.y;
  const int slot = a.slots[b];
  if (slot < 0) return;
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= a.nib) return;
  const uint8_t* src = a.stage + uint64_t(b) * a.content;
  const uint64_t node = uint64_t(slot) * a.nib + n;
  uint32_t link = reinterpret_cast<const uint32_t*>(src)[n];
  const uint32_t mask = link & 255u;
  if (mask && !(link & (1u << 8))) {
    // stored near link (child index inside the block) -> cache near link
    const uint32_t ptr = link >> 11;
    const uint32_t near_value = uint32_t((1 << 20) + (int32_t(ptr) - n));
    link = (near_value << 11) | (link & (1u << 9)) | mask;
  }
  reinterpret_cast<uint32_t*>(a.chan[0])[node] = link;
  for (int c = 1; c < a.nchan; ++c) {
    const int bytes = a.chan_bytes[c];
    const uint8_t* s = src + a.chan_off[c] + uint64_t(n) * bytes;
    uint8_t* d = a.chan[c] + node * bytes;
    if ((bytes & 3) == 0) {
      for (int k = 0; k < bytes; k += 4)
        *reinterpret_cast<uint32_t*>(d + k) = *reinterpret_cast<const uint32_t*>(s + k);
    } else {
      for (int k = 0; k < bytes; ++k) d[k] = s[k];
    }
  }
}

__global__ __launch_bounds__(256) void ApplyWritesKernel(PoolBatchArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < a.nlw) reinterpret_cast<uint32_t*>(a.chan[0])[a.lw_addr[i]] = a.lw_val[i];
  if (i < a.nfw) a.far_ptrs[a.fw_idx[i]] = a.fw_val[i];
  if (i < a.nblocks && a.slots[i] >= 0) a.slot_block[a.slots[i]] = int32_t(a.tree_blocks[i]);
}

}  // namespace

int LaunchPoolBatch(hipStream_t stream, const PoolBatchArgs& a) {
  if (a.nblocks > 0) {
    const dim3 grid((a.nib + 255) / 256, a.nblocks);
    ScatterBlocksKernel<<<grid, dim3(256), 0, stream>>>(a);
    RT_HIP_CHECK(hipGetLastError());
  }
  const int n = a.nlw > a.nfw ? a.nlw : a.nfw;
  const int m = n > a.nblocks ? n : a.nblocks;
  if (m > 0) {
    ApplyWritesKernel<<<dim3((m + 255) / 256), dim3(256), 0, stream>>>(a);
    RT_HIP_CHECK(hipGetLastError());
  }
  return RT_OK;
}

}  // namespace rayt
