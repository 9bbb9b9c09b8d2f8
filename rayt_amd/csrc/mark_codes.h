// Marking of distinct feedback codes in the extractor's bitmaps, shared by
// the feedback extractor's marking kernel and the plain renderer, which marks
// the codes of each of its waves as it writes them (so a resident frame needs
// no separate pass over its codes before the bitmap walk).
//
// A code is a hit code ((slot + 1) << 1 | dup, block level), a fault
// (-block - 1) or NO_HIT (raytracing_pass.cl:38-55, common.cl:14). Hit codes
// below kHitDomain and fault blocks below kFaultDomain have a bit each; any
// other code sets `overflow`, and the extractor then takes the sort path over
// the codes in memory.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace rayt {

constexpr int kMarkNoHit = 0x7fffffff;
constexpr uint32_t kFaultDomain = 1u << 21;  // fault links carry 21-bit block indices
constexpr uint32_t kHitDomain = 1u << 23;    // hit codes handled by the bitmap path
constexpr uint32_t kMarkEmpty = 0xffffffffu;
constexpr uint32_t kMarkFaultKey = 0x40000000u;  // key = hit code, or fault block | this bit

// Where a renderer marks its codes. NO_HIT and "a code outside the bitmap
// domain" are marked as bits 0 and 1 of the hit bitmap (no hit code is 0 or
// 1: a hit code is (slot + 1) << 1 | dup). fault_limit / hit_limit bound the
// codes the pool can produce (tree blocks, 2 * slots + 2): the walk reads the
// bitmaps only below them, and a code at or above them is marked as outside
// the domain (the extractor then takes the sort path over the codes).
struct MarkTargets {
  uint32_t* fault_bits = nullptr;
  uint32_t* hit_bits = nullptr;
  uint32_t fault_limit = 0;
  uint32_t hit_limit = 0;
};
constexpr uint32_t kMarkNoHitKey = 0;
constexpr uint32_t kMarkOverflowKey = 1;

#ifdef __HIP__  // device code: the kernels' translation units only

__device__ __forceinline__ void MarkBit(uint32_t key, uint32_t* fault_bits, uint32_t* hit_bits) {
  if (key & kMarkFaultKey) {
    const uint32_t f = key & ~kMarkFaultKey;
    atomicOr(&fault_bits[f >> 5], 1u << (f & 31u));
  } else {
    atomicOr(&hit_bits[key >> 5], 1u << (key & 31u));
  }
}

// Marks the codes of one wave (valid: the lane has a code). The wave takes
// its distinct keys one at a time: the first pending lane's key is marked
// once by that lane and retires every lane holding it, so a wave issues one
// atomic per distinct key (a 64-pixel tile has a few) and no barrier.
__device__ __forceinline__ void MarkWaveCodes(int code, bool valid, const MarkTargets& t) {
  uint32_t key = kMarkEmpty;
  if (valid) {
    if (code == kMarkNoHit) {
      key = kMarkNoHitKey;
    } else if (code < 0) {
      const uint32_t f = uint32_t(-(code + 1));
      key = f < t.fault_limit ? (f | kMarkFaultKey) : kMarkOverflowKey;
    } else {
      key = uint32_t(code) < t.hit_limit ? uint32_t(code) : kMarkOverflowKey;
    }
  }
  bool pending = key != kMarkEmpty;
  uint64_t left = __ballot(pending);
  while (left) {
    const int leader = __ffsll(static_cast<unsigned long long>(left)) - 1;
    const uint32_t k = uint32_t(__shfl(int(key), leader, 64));
    if ((threadIdx.x & 63) == uint32_t(leader)) MarkBit(k, t.fault_bits, t.hit_bits);
    if (pending && key == k) pending = false;
    left = __ballot(pending);
  }
}

#endif  // __HIP__

}  // namespace rayt
