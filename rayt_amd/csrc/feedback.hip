// Feedback pipeline on gfx950: the reference's GPU algorithms (GPUSort,
// GPUScan, NodeToBlock / DifferenceFlag / Compress) and the extraction of the
// fault / hit block lists (RayTracingFeedbackExtractor).
//
// The extractor does not sort. Every feedback code lives in a small known
// domain (fault blocks < 2^21, hit codes < 2 * pool blocks + 2), so the set of
// distinct codes is marked in two bitmaps and a single workgroup walks the
// bitmaps in code order - fault bitmap from the highest block down, then the
// hit bitmap upwards, then NO_HIT - emitting exactly the array that
// sort + DifferenceFlag + Scan + Compress produce. Codes outside the domain
// fall back to the radix-sort path.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "block_scan.h"
#include "feedback.h"
#include "mark_codes.h"
#include "rt_internal.h"

namespace rayt {

namespace {

constexpr int kNoHit = kMarkNoHit;

// ---------------------------------------------------- feedback_compression.cl
__global__ void NodeToBlockKernel(int n, int block_size, int* __restrict__ data) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int d = data[i];
  if (d > 0 && d != kNoHit) data[i] = ((((d >> 1) - 1) / block_size + 1) << 1) | (d & 1);
}

__global__ void DifferenceFlagKernel(int n, const int* __restrict__ in, int* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = (i != 0 && in[i - 1] != in[i]) ? 1 : 0;
}

__global__ void CompressKernel(int n, const int* __restrict__ index, const int* __restrict__ in, int* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (i == 0)
    out[0] = in[0];
  else if (in[i - 1] != in[i])
    out[index[i]] = in[i];
}

// One pass of the merge sort of GPU algorithms/kernels/merge_sort.cl:47-73:
// the input holds sorted runs of `len` elements (len a power of two); each
// pair of runs [2m len, (2m+1) len) and [(2m+1) len, (2m+2) len) is merged
// into one sorted run of 2 len. An element's place is its index in its run
// plus the number of elements of the other run that precede it: strictly
// smaller ones when it comes from the first run, smaller or equal ones when
// it comes from the second (a binary search over the other run), so equal
// keys keep their order.
__global__ void MergePassKernel(int n, int len, const int* __restrict__ in, int* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int key = in[i];
  const int ii = i & (len - 1);
  const int run = i - ii;
  const int pair = run & ~(2 * len - 1);
  const bool first = run == pair;
  const int other = first ? run + len : pair;  // start of the other run
  const int other_end = min(n, other + len);
  int lo = other, hi = max(other, other_end);  // count of other-run elements that precede `key`
  while (lo < hi) {
    const int mid = lo + (hi - lo) / 2;
    const int v = in[mid];
    if (first ? v < key : v <= key)
      lo = mid + 1;
    else
      hi = mid;
  }
  out[pair + ii + (lo - other)] = key;
}

// One Hillis-Steele pass of GPU algorithms/kernels/scan.cl:2-12: out[i] =
// in[i] + in[i - len] where i >= len (with int32 wrap-around), else in[i].
__global__ void ScanPassKernel(int n, int len, const int* __restrict__ in, int* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t v = uint32_t(in[i]);
  if (i >= len) v += uint32_t(in[i - len]);
  out[i] = int(v);
}

// ------------------------------------------------------------- radix sort
// LSD radix sort of int32 keys (sign bit flipped), 8-bit digits, 4 passes.
// Per tile of 256 x 8 keys: a digit histogram (upsweep), one scan over all
// (digit, tile) counts, and a stable scatter that sorts the tile in LDS with
// eight 1-bit block-wide splits before writing runs out.
constexpr int kSortBlock = 256;
constexpr int kSortItems = 8;
constexpr int kSortTile = kSortBlock * kSortItems;

__device__ __forceinline__ uint32_t SortKey(int v) { return uint32_t(v) ^ 0x80000000u; }

__global__ __launch_bounds__(kSortBlock) void RadixUpsweepKernel(const int* __restrict__ in, int n, int shift,
                                                                 uint32_t* __restrict__ hist, int tiles) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const int base = blockIdx.x * kSortTile;
  for (int k = 0; k < kSortItems; ++k) {
    const int i = base + k * kSortBlock + threadIdx.x;
    if (i < n) atomicAdd(&h[(SortKey(in[i]) >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[threadIdx.x * tiles + blockIdx.x] = h[threadIdx.x];
}

// Exclusive scan of `n` uint32 in place with one workgroup of 1024 lanes.
__global__ __launch_bounds__(1024) void ScanSingleBlockKernel(uint32_t* __restrict__ data, int n) {
  __shared__ uint32_t waves[1024 / 64 + 1];
  const int per = (n + 1023) / 1024;
  const int b = threadIdx.x * per, e = min(n, b + per);
  uint32_t s = 0;
  for (int i = b; i < e; ++i) s += data[i];
  uint32_t total;
  uint32_t run = BlockExclusiveScan<1024>(s, waves, total);
  for (int i = b; i < e; ++i) {
    const uint32_t v = data[i];
    data[i] = run;
    run += v;
  }
}

__global__ __launch_bounds__(kSortBlock) void RadixScatterKernel(const int* __restrict__ in, int* __restrict__ out,
                                                                 int n, int shift, const uint32_t* __restrict__ hist,
                                                                 int tiles) {
  __shared__ uint32_t keys[kSortTile];
  __shared__ uint32_t tmp[kSortTile];
  __shared__ uint32_t waves[kSortBlock / 64 + 1];
  __shared__ uint32_t digit_start[256];
  __shared__ uint32_t digit_global[256];
  const int base = blockIdx.x * kSortTile;
  const int valid = min(kSortTile, n - base);
  // Blocked arrangement: thread t owns keys [t*8, t*8+8) of the tile, in tile order.
  uint32_t k[kSortItems];
  for (int j = 0; j < kSortItems; ++j) {
    const int i = threadIdx.x * kSortItems + j;
    k[j] = i < valid ? SortKey(in[base + i]) : 0xffffffffu;  // padding sorts last and is never written
  }
  // Stable sort of the tile by the current digit: 8 one-bit splits.
  for (int bit = 0; bit < 8; ++bit) {
    uint32_t zeros = 0;
    for (int j = 0; j < kSortItems; ++j) zeros += ((k[j] >> (shift + bit)) & 1u) ? 0u : 1u;
    uint32_t total_zeros;
    const uint32_t zero_before = BlockExclusiveScan<kSortBlock>(zeros, waves, total_zeros);
    uint32_t zpos = zero_before;
    uint32_t opos = total_zeros + threadIdx.x * kSortItems - zero_before;
    for (int j = 0; j < kSortItems; ++j) {
      if ((k[j] >> (shift + bit)) & 1u)
        tmp[opos++] = k[j];
      else
        tmp[zpos++] = k[j];
    }
    __syncthreads();
    for (int j = 0; j < kSortItems; ++j) k[j] = tmp[threadIdx.x * kSortItems + j];
    __syncthreads();
  }
  for (int j = 0; j < kSortItems; ++j) keys[threadIdx.x * kSortItems + j] = k[j];
  digit_global[threadIdx.x] = hist[threadIdx.x * tiles + blockIdx.x];
  digit_start[threadIdx.x] = 0xffffffffu;
  __syncthreads();
  // First position of each digit in the sorted tile.
  for (int j = 0; j < kSortItems; ++j) {
    const int i = threadIdx.x * kSortItems + j;
    if (i < valid) {
      const uint32_t d = (keys[i] >> shift) & 255u;
      if (i == 0 || ((keys[i - 1] >> shift) & 255u) != d) digit_start[d] = uint32_t(i);
    }
  }
  __syncthreads();
  for (int j = 0; j < kSortItems; ++j) {
    const int i = j * kSortBlock + threadIdx.x;  // striped for coalesced runs
    if (i < valid) {
      const uint32_t key = keys[i];
      const uint32_t d = (key >> shift) & 255u;
      out[digit_global[d] + (uint32_t(i) - digit_start[d])] = int(key ^ 0x80000000u);
    }
  }
}

// ------------------------------------------------------------------ scan
constexpr int kScanBlock = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanBlock * kScanItems;

__global__ __launch_bounds__(kScanBlock) void ScanReduceKernel(const int* __restrict__ in, int n,
                                                               uint32_t* __restrict__ sums) {
  __shared__ uint32_t waves[kScanBlock / 64 + 1];
  const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  uint32_t s = 0;
  for (int j = 0; j < kScanItems; ++j)
    if (base + j < n) s += uint32_t(in[base + j]);
  uint32_t total;
  BlockExclusiveScan<kScanBlock>(s, waves, total);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(kScanBlock) void ScanDownsweepKernel(const int* __restrict__ in, int* __restrict__ out,
                                                                  int n, const uint32_t* __restrict__ offsets) {
  __shared__ uint32_t waves[kScanBlock / 64 + 1];
  const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  uint32_t s = 0;
  for (int j = 0; j < kScanItems; ++j) {
    v[j] = base + j < n ? uint32_t(in[base + j]) : 0u;
    s += v[j];
  }
  uint32_t total;
  uint32_t run = BlockExclusiveScan<kScanBlock>(s, waves, total) + offsets[blockIdx.x];
  for (int j = 0; j < kScanItems; ++j) {
    run += v[j];
    if (base + j < n) out[base + j] = int(run);
  }
}

// ------------------------------------------------------ bitmap extraction
struct MarkState {
  int32_t max_fault;     // highest fault block seen (-1: none)
  int32_t max_hit;       // highest hit code seen (-1: none)
  int32_t no_hit;        // NO_HIT present
  int32_t overflow;      // a code outside the bitmap domain
  int32_t unique_count;  // outputs of the walk
  int32_t fault_count;
  int32_t pad[2];
};

__device__ __forceinline__ int WaveMax(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}

// Marks the distinct codes in the bitmaps. Each workgroup takes a contiguous
// chunk of kMarkItems codes per thread and inserts every code into an LDS
// hash set of the workgroup (one compare-and-swap per lane and code: lanes
// with equal codes collide on one LDS word, which the hardware serialises
// inside the instruction). After the chunk, every distinct value of the
// workgroup sets its bitmap bit with one atomic that returns nothing, so no
// wave waits on global memory and a hot bitmap word sees one atomic per
// workgroup. A value that finds no free slot within kMarkProbe probes sets its
// bit directly. The maxima / flags of each workgroup go to its own slot of
// `partials`, reduced by the walk.
constexpr int kMarkBlock = 512;
constexpr int kMarkItems = 16;
constexpr int kMarkBatch = 8;  // codes per thread loaded together
constexpr int kMarkSlots = 4096;
constexpr int kMarkProbe = 32;

__global__ __launch_bounds__(kMarkBlock) void MarkCodesKernel(const int* __restrict__ codes, int n, int block_size,
                                                              uint32_t* __restrict__ fault_bits,
                                                              uint32_t* __restrict__ hit_bits,
                                                              int4* __restrict__ partials) {
  __shared__ uint32_t table[kMarkSlots];
  __shared__ int red[4];  // max fault, max hit, NO_HIT seen, overflow
  for (int j = threadIdx.x; j < kMarkSlots; j += kMarkBlock) table[j] = kMarkEmpty;
  if (threadIdx.x < 4) red[threadIdx.x] = threadIdx.x < 2 ? -1 : 0;
  __syncthreads();
  const int lane = threadIdx.x & 63;
  int mf = -1, mh = -1;
  bool nohit = false, over = false;
  const int chunk = kMarkBlock * kMarkItems;
  for (int base = blockIdx.x * chunk; base < n; base += gridDim.x * chunk) {
    for (int it0 = 0; it0 < kMarkItems; it0 += kMarkBatch) {
      int loaded[kMarkBatch];  // the batch's loads are in flight before the first is used
#pragma unroll
      for (int k = 0; k < kMarkBatch; ++k) {
        const int i = base + (it0 + k) * kMarkBlock + threadIdx.x;  // coalesced: consecutive lanes, consecutive codes
        loaded[k] = i < n ? codes[i] : kNoHit;
      }
#pragma unroll
      for (int k = 0; k < kMarkBatch; ++k) {
        const bool valid = base + (it0 + k) * kMarkBlock + int(threadIdx.x) < n;
        int code = loaded[k];
        if (block_size > 0 && code > 0 && code != kNoHit)
          code = ((((code >> 1) - 1) / block_size + 1) << 1) | (code & 1);
        uint32_t key = kMarkEmpty;
        if (valid) {
          if (code == kNoHit) {
            nohit = true;
          } else if (code < 0) {
            const uint32_t f = uint32_t(-(code + 1));
            if (f >= kFaultDomain) {
              over = true;
            } else {
              key = f | kMarkFaultKey;
              mf = max(mf, int(f));
            }
          } else if (uint32_t(code) >= kHitDomain) {
            over = true;
          } else {
            key = uint32_t(code);
            mh = max(mh, code);
          }
        }
        if (key != kMarkEmpty) {
          uint32_t h = (key * 0x9E3779B1u) >> 20;  // 12-bit slot
          int p = 0;
          for (; p < kMarkProbe; ++p) {
            const uint32_t old = atomicCAS(&table[h], kMarkEmpty, key);
            if (old == kMarkEmpty || old == key) break;
            h = (h + 1) & (kMarkSlots - 1);
          }
          if (p == kMarkProbe) MarkBit(key, fault_bits, hit_bits);
        }
      }
    }
  }
  mf = WaveMax(mf);
  mh = WaveMax(mh);
  const bool any_nohit = __any(nohit), any_over = __any(over);
  if (lane == 0) {
    if (mf >= 0) atomicMax(&red[0], mf);
    if (mh >= 0) atomicMax(&red[1], mh);
    if (any_nohit) red[2] = 1;
    if (any_over) red[3] = 1;
  }
  __syncthreads();
  for (int j = threadIdx.x; j < kMarkSlots; j += kMarkBlock) {
    const uint32_t v = table[j];
    if (v != kMarkEmpty) MarkBit(v, fault_bits, hit_bits);
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = make_int4(red[0], red[1], red[2], red[3]);
}

// One workgroup reduces the marking partials, walks the bitmaps in ascending
// code order and writes the distinct codes, then clears what it read and
// leaves the totals in the state. premark_words > 0: the codes were marked by
// a renderer (MarkWaveCodes): the walk covers the first premark_words.x fault
// words and premark_words.y hit words, and reads NO_HIT and overflow from
// bits 0 and 1 of the hit bitmap instead of the partials.
//
// The state and the first prefix_len distinct codes also go straight to
// pinned host memory (host_state, host_prefix): the host reads them after the
// kernel, with no copy engine between this kernel and the next one.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void WalkBitmapsKernel(uint32_t* __restrict__ fault_bits,
                                                          uint32_t* __restrict__ hit_bits, MarkState* __restrict__ st,
                                                          int* __restrict__ unique_out,
                                                          const int4* __restrict__ partials, int nparts,
                                                          int2 premark_words, MarkState* __restrict__ host_state,
                                                          int* __restrict__ host_prefix, int prefix_len) {
  __shared__ uint32_t waves[BLOCK / 64 + 1];
  __shared__ int red[4];
  const bool premarked = premark_words.y > 0;
  if (threadIdx.x < 4) red[threadIdx.x] = threadIdx.x < 2 ? -1 : 0;
  __syncthreads();
  if (premarked) {
    if (threadIdx.x == 0) {
      const uint32_t w0 = hit_bits[0];
      red[2] = int(w0 >> kMarkNoHitKey) & 1;
      red[3] = int(w0 >> kMarkOverflowKey) & 1;
    }
  } else {
    int4 acc = make_int4(-1, -1, 0, 0);
    for (int j = threadIdx.x; j < nparts; j += BLOCK) {
      const int4 p = partials[j];
      acc.x = max(acc.x, p.x);
      acc.y = max(acc.y, p.y);
      acc.z |= p.z;
      acc.w |= p.w;
    }
    acc.x = WaveMax(acc.x);
    acc.y = WaveMax(acc.y);
    const bool z = __any(acc.z != 0), w = __any(acc.w != 0);
    if ((threadIdx.x & 63) == 0) {
      atomicMax(&red[0], acc.x);
      atomicMax(&red[1], acc.y);
      if (z) red[2] = 1;
      if (w) red[3] = 1;
    }
  }
  __syncthreads();
  const int max_fault = red[0], max_hit = red[1], no_hit = red[2], overflow = red[3];
  const int wf = premarked ? premark_words.x : max_fault >= 0 ? (max_fault >> 5) + 1 : 0;
  const int wh = premarked ? premark_words.y : max_hit >= 0 ? (max_hit >> 5) + 1 : 0;
  const uint32_t flag_bits = premarked ? (1u << kMarkNoHitKey) | (1u << kMarkOverflowKey) : 0u;
  const int L = wf + wh;
  const int per = (L + BLOCK - 1) / BLOCK;
  const int b = threadIdx.x * per, e = min(L, b + per);
  auto word = [&](int j) -> uint32_t {
    return j < wf ? fault_bits[wf - 1 - j] : j == wf ? hit_bits[0] & ~flag_bits : hit_bits[j - wf];
  };
  uint32_t cnt = 0;
  for (int j = b; j < e; ++j) cnt += __popc(word(j));
  uint32_t total;
  uint32_t pos = BlockExclusiveScan<BLOCK>(cnt, waves, total);
  uint32_t faults_total;
  {
    uint32_t fc = 0;
    for (int j = b; j < min(e, wf); ++j) fc += __popc(word(j));
    BlockExclusiveScan<BLOCK>(fc, waves, faults_total);
  }
  for (int j = b; j < e; ++j) {
    uint32_t w = word(j);
    if (j < wf) {
      const int fw = wf - 1 - j;
      while (w) {  // descending block index = ascending code
        const int bit = 31 - __clz(w);
        w &= ~(1u << bit);
        unique_out[pos++] = -(fw * 32 + bit) - 1;
      }
      fault_bits[fw] = 0u;
    } else {
      const int hw = j - wf;
      while (w) {
        const int bit = __ffs(w) - 1;
        w &= w - 1;
        unique_out[pos++] = hw * 32 + bit;
      }
      hit_bits[hw] = 0u;
    }
  }
  __syncthreads();
  const int u = int(total) + (no_hit ? 1 : 0);
  MarkState fin;
  fin.unique_count = u;
  fin.fault_count = int(faults_total);
  fin.max_fault = max_fault;
  fin.max_hit = max_hit;
  fin.no_hit = no_hit;
  fin.overflow = overflow;
  if (threadIdx.x == 0) {
    if (no_hit) unique_out[u - 1] = kNoHit;
    st->unique_count = fin.unique_count;
    st->fault_count = fin.fault_count;
    st->max_fault = fin.max_fault;
    st->max_hit = fin.max_hit;
    st->no_hit = fin.no_hit;
    st->overflow = fin.overflow;
  }
  if (host_state) {
    __syncthreads();  // the NO_HIT entry is in place
    const int m = min(u, prefix_len);
    for (int i = threadIdx.x; i < m; i += BLOCK) host_prefix[i] = unique_out[i];
    if (threadIdx.x == 0) {
      host_state->unique_count = fin.unique_count;
      host_state->fault_count = fin.fault_count;
      host_state->max_fault = fin.max_fault;
      host_state->max_hit = fin.max_hit;
      host_state->no_hit = fin.no_hit;
      host_state->overflow = fin.overflow;
    }
    __threadfence_system();
  }
}

int Grid(int64_t n, int block) { return int((n + block - 1) / block); }

}  // namespace

// ------------------------------------------------------------ host helpers
// The four 8-bit passes of the radix sort, ping-ponging between `data` and
// `temp` (the result ends in `data`); `hist` holds 256 counters per tile and
// is written in full by every upsweep, so it needs no clearing.
static void SortPasses(hipStream_t s, int* data, int n, int* temp, uint32_t* hist) {
  const int tiles = Grid(n, kSortTile);
  int* a = data;
  int* b = temp;
  for (int shift = 0; shift < 32; shift += 8) {
    RadixUpsweepKernel<<<dim3(tiles), dim3(kSortBlock), 0, s>>>(a, n, shift, hist, tiles);
    ScanSingleBlockKernel<<<dim3(1), dim3(1024), 0, s>>>(hist, tiles * 256);
    RadixScatterKernel<<<dim3(tiles), dim3(kSortBlock), 0, s>>>(a, b, n, shift, hist, tiles);
    std::swap(a, b);
  }
}

int SortI32(hipStream_t s, int* data, int n, int* temp) {
  if (n <= 1) return RT_OK;
  const int tiles = Grid(n, kSortTile);
  uint32_t* hist = nullptr;
  RT_HIP_CHECK(hipMallocAsync(reinterpret_cast<void**>(&hist), size_t(tiles) * 256 * 4, s));
  SortPasses(s, data, n, temp, hist);
  hipError_t e = hipGetLastError();
  (void)hipFreeAsync(hist, s);
  if (e != hipSuccess) return Fail(RT_ERR_HIP, hipGetErrorString(e));
  return RT_OK;
}

// Scratch of the sort that allocates nothing (rt_sort_i32_noalloc): the
// ping-pong buffer of kNoAllocSortMax ints, then the tile histograms. One
// copy per GPU, for the life of the process. In host code `scratch` names a
// host-side stand-in, never the device memory: the device address comes from
// hipGetSymbolAddress, once per device, when a context is created there.
constexpr int kNoAllocSortMax = RT_SORT_NOALLOC_MAX_N;
constexpr size_t kNoAllocHistWords = size_t((kNoAllocSortMax + kSortTile - 1) / kSortTile) * 256;
static __device__ __align__(256) uint32_t scratch[size_t(kNoAllocSortMax) + kNoAllocHistWords];

constexpr int kMaxScratchDevices = 64;
static std::mutex g_scratch_mutex;
static uint32_t* g_scratch_ptr[kMaxScratchDevices] = {};

int NoAllocScratchInit(int device) {
  RT_REQUIRE(device >= 0 && device < kMaxScratchDevices, RT_ERR_RANGE, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_ptr[device]) return RT_OK;
  int current = -1;
  RT_HIP_CHECK(hipGetDevice(&current));
  RT_REQUIRE(current == device, RT_ERR_STATE, "scratch set-up needs its device current");
  void* ptr = nullptr;
  RT_HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  RT_REQUIRE(ptr, RT_ERR_HIP, "hipGetSymbolAddress returned a null scratch address");
  g_scratch_ptr[device] = static_cast<uint32_t*>(ptr);
  return RT_OK;
}

int ScanI32(hipStream_t s, int* data, int n) {
  if (n <= 0) return RT_OK;
  const int tiles = Grid(n, kScanTile);
  uint32_t* sums = nullptr;
  RT_HIP_CHECK(hipMallocAsync(reinterpret_cast<void**>(&sums), size_t(tiles) * 4, s));
  ScanReduceKernel<<<dim3(tiles), dim3(kScanBlock), 0, s>>>(data, n, sums);
  ScanSingleBlockKernel<<<dim3(1), dim3(1024), 0, s>>>(sums, tiles);
  ScanDownsweepKernel<<<dim3(tiles), dim3(kScanBlock), 0, s>>>(data, data, n, sums);
  hipError_t e = hipGetLastError();
  (void)hipFreeAsync(sums, s);
  if (e != hipSuccess) return Fail(RT_ERR_HIP, hipGetErrorString(e));
  return RT_OK;
}

// --------------------------------------------------------- FeedbackExtractor
FeedbackExtractor::FeedbackExtractor(rt_ctx* ctx, int count)
    : ctx_(CtxRetain(ctx)), stream_(ctx->stream), count_(count) {
  auto alloc = [&](void** p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) throw std::runtime_error(std::string("hipMalloc: ") + hipGetErrorString(e));
  };
  alloc(reinterpret_cast<void**>(&fault_bits_), kFaultDomain / 8);
  alloc(reinterpret_cast<void**>(&hit_bits_), kHitDomain / 8);
  alloc(reinterpret_cast<void**>(&state_), sizeof(MarkState));
  alloc(reinterpret_cast<void**>(&unique_), size_t(std::max(count, 1) + 1) * 4);
  alloc(reinterpret_cast<void**>(&temp_), size_t(std::max(count, 1)) * 4);
  alloc(reinterpret_cast<void**>(&index_), size_t(std::max(count, 1)) * 4);
  max_mark_blocks_ = std::max(1, ctx->compute_units) * 8;
  alloc(&partials_, size_t(max_mark_blocks_) * sizeof(int4));
  (void)hipMemsetAsync(fault_bits_, 0, kFaultDomain / 8, ctx_->stream);
  (void)hipMemsetAsync(hit_bits_, 0, kHitDomain / 8, ctx_->stream);
  MarkState init{-1, -1, 0, 0, 0, 0, {0, 0}};
  (void)hipMemcpyAsync(state_, &init, sizeof(init), hipMemcpyHostToDevice, ctx_->stream);
  hipError_t e = hipHostMalloc(reinterpret_cast<void**>(&host_state_), sizeof(MarkState), hipHostMallocDefault);
  if (e != hipSuccess) throw std::runtime_error("hipHostMalloc");
  e = hipHostMalloc(&host_prefix_, size_t(kPrefixInts) * 4, hipHostMallocDefault);
  if (e != hipSuccess) throw std::runtime_error("hipHostMalloc");
  host_sorted_.resize(size_t(std::max(count, 1)) + 1);
  e = hipEventCreateWithFlags(&copied_, hipEventDisableTiming);
  if (e != hipSuccess) throw std::runtime_error("hipEventCreateWithFlags");
  (void)hipStreamSynchronize(ctx_->stream);
}

FeedbackExtractor::~FeedbackExtractor() {
  (void)hipStreamSynchronize(stream_);
  (void)hipStreamSynchronize(ctx_->stream);
  (void)hipFree(fault_bits_);
  (void)hipFree(hit_bits_);
  if (fault_bits2_) (void)hipFree(fault_bits2_);
  if (hit_bits2_) (void)hipFree(hit_bits2_);
  (void)hipFree(state_);
  (void)hipFree(unique_);
  (void)hipFree(temp_);
  (void)hipFree(index_);
  (void)hipFree(partials_);
  (void)hipHostFree(host_state_);
  (void)hipHostFree(host_prefix_);
  if (copied_) (void)hipEventDestroy(copied_);
  CtxRelease(ctx_);
}

uint32_t FeedbackExtractor::NextRandom() {
  uint64_t z = (rng_ += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return uint32_t(z >> 33);
}

// Sorted distinct codes via the reference's own sequence of steps (sort,
// difference flags, scan, compress), used when codes leave the bitmap domain.
int FeedbackExtractor::UniqueBySort(int* codes, int n) {
  hipStream_t s = stream_;
  int rc = SortI32(s, codes, n, temp_);
  if (rc) return rc;
  DifferenceFlagKernel<<<dim3(Grid(n, 256)), dim3(256), 0, s>>>(n, codes, index_);
  rc = ScanI32(s, index_, n);
  if (rc) return rc;
  CompressKernel<<<dim3(Grid(n, 256)), dim3(256), 0, s>>>(n, index_, codes, unique_);
  RT_HIP_CHECK(hipGetLastError());
  int last = 0;
  RT_HIP_CHECK(hipMemcpyAsync(&last, index_ + (n - 1), 4, hipMemcpyDeviceToHost, s));
  RT_HIP_CHECK(hipStreamSynchronize(s));
  unique_count_ = last + 1;
  // fault count = number of negative distinct codes
  std::vector<int> tmp(unique_count_);
  RT_HIP_CHECK(hipMemcpy(tmp.data(), unique_, size_t(unique_count_) * 4, hipMemcpyDeviceToHost));
  fault_count_ = int(std::lower_bound(tmp.begin(), tmp.end(), 0) - tmp.begin());
  return RT_OK;
}

int FeedbackExtractor::Process(int* codes, int nodes_in_block, bool block_level) {
  return ProcessN(codes, count_, nodes_in_block, block_level);
}

int FeedbackExtractor::EnsureSecondSet() {
  if (fault_bits2_) return RT_OK;
  RT_HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&fault_bits2_), kFaultDomain / 8));
  RT_HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&hit_bits2_), kHitDomain / 8));
  RT_HIP_CHECK(hipMemsetAsync(fault_bits2_, 0, kFaultDomain / 8, ctx_->stream));
  RT_HIP_CHECK(hipMemsetAsync(hit_bits2_, 0, kHitDomain / 8, ctx_->stream));
  RT_HIP_CHECK(hipStreamSynchronize(ctx_->stream));
  return RT_OK;
}

MarkTargets FeedbackExtractor::Targets(uint32_t fault_limit, uint32_t hit_limit, int set) {
  premark_fault_limit_ = std::min(fault_limit, kFaultDomain);
  premark_hit_limit_ = std::max(2u, std::min(hit_limit, kHitDomain));
  if (set && EnsureSecondSet() != RT_OK) throw std::runtime_error(rt_last_error());
  MarkTargets t;
  t.fault_bits = set ? fault_bits2_ : fault_bits_;
  t.hit_bits = set ? hit_bits2_ : hit_bits_;
  t.fault_limit = premark_fault_limit_;
  t.hit_limit = premark_hit_limit_;
  return t;
}

int FeedbackExtractor::ProcessN(int* codes, int n, int nodes_in_block, bool block_level, bool premarked) {
  const int rc = EnqueueN(codes, n, nodes_in_block, block_level, premarked);
  return rc ? rc : CompleteN();
}

int FeedbackExtractor::EnqueueN(int* codes, int n, int nodes_in_block, bool block_level, bool premarked, int set,
                                hipStream_t walk_stream, hipEvent_t after) {
  RT_REQUIRE(n >= 0 && n <= count_, RT_ERR_RANGE, "more codes than the extractor was created for");
  RT_REQUIRE(set == 0 || (set == 1 && premarked && fault_bits2_), RT_ERR_INVALID,
             "the second bitmap set holds premarked codes only (mark into Targets(..., 1) first)");
  hipStream_t s = walk_stream ? walk_stream : stream_;
  uint32_t* fault_bits = set ? fault_bits2_ : fault_bits_;
  uint32_t* hit_bits = set ? hit_bits2_ : hit_bits_;
  q_codes_ = codes;
  q_n_ = n;
  q_nib_ = nodes_in_block;
  q_block_level_ = block_level;
  q_empty_ = n <= 0;
  if (q_empty_) return RT_OK;
  if (after) RT_HIP_CHECK(hipStreamWaitEvent(s, after, 0));
  const int bs = block_level ? 0 : nodes_in_block;
  MarkState* st = static_cast<MarkState*>(state_);
  // Premarked: the renderer marked these block-level codes into Targets() as it wrote them.
  const int grid = premarked ? 1 : std::min(Grid(n, kMarkBlock * kMarkItems), max_mark_blocks_);
  if (!premarked)
    MarkCodesKernel<<<dim3(grid), dim3(kMarkBlock), 0, s>>>(codes, n, bs, fault_bits, hit_bits,
                                                     static_cast<int4*>(partials_));
  const int2 words = premarked ? make_int2(int((premark_fault_limit_ + 31) / 32), int((premark_hit_limit_ + 31) / 32))
                               : make_int2(0, 0);
  // The walk writes the state and, speculatively, the head of the unique
  // list (usually all of it) into pinned host memory itself.
  prefix_len_ = std::min(n + 1, kPrefixInts);
  // A premarked walk over a few thousand words takes one 256-thread workgroup;
  // larger walks and the reduction of many partials take 1024 threads.
  if (premarked && words.x + words.y <= 4096)
    WalkBitmapsKernel<256><<<dim3(1), dim3(256), 0, s>>>(fault_bits, hit_bits, st, unique_,
                                                         static_cast<const int4*>(partials_), 0, words,
                                                         static_cast<MarkState*>(host_state_),
                                                         static_cast<int*>(host_prefix_), prefix_len_);
  else
    WalkBitmapsKernel<1024><<<dim3(1), dim3(1024), 0, s>>>(fault_bits, hit_bits, st, unique_,
                                                           static_cast<const int4*>(partials_), premarked ? 0 : grid,
                                                           words, static_cast<MarkState*>(host_state_),
                                                           static_cast<int*>(host_prefix_), prefix_len_);
  RT_HIP_CHECK(hipGetLastError());
  RT_HIP_CHECK(hipEventRecord(copied_, s));
  return RT_OK;
}

int FeedbackExtractor::CompleteN() {
  if (q_empty_) {
    faults_.clear();
    hits_.clear();
    dup_hits_.clear();
    unique_count_ = 0;
    return RT_OK;
  }
  hipStream_t s = stream_;
  RT_HIP_CHECK(hipEventSynchronize(copied_));
  const MarkState& hs = *reinterpret_cast<const MarkState*>(host_state_);
  if (hs.overflow) {
    prefix_len_ = 0;  // the sort path rewrites the list
    if (!q_block_level_) NodeToBlockKernel<<<dim3(Grid(q_n_, 256)), dim3(256), 0, s>>>(q_n_, q_nib_, q_codes_);
    int rc = UniqueBySort(q_codes_, q_n_);
    if (rc) return rc;
  } else {
    unique_count_ = hs.unique_count;
    fault_count_ = hs.fault_count;
  }
  return ReadLists();
}

// ProcessUnique + AdjustExpectedFaults (ray-tracing-feedback-extractor.cpp:108-185):
// the reads of growing strides until the first non-fault are replayed from the
// known fault count, then one window of hits is read from a random start.
int FeedbackExtractor::ReadLists() {
  hipStream_t s = stream_;
  faults_.clear();
  hits_.clear();
  dup_hits_.clear();
  const int U = unique_count_, F = fault_count_;
  int stride = expected_faults_, size = 0;
  while (size < U && (size == 0 || size <= F)) {  // sorted[size-1] < 0  <=>  size-1 < F
    stride = std::min(stride, U - size);
    size += stride;
    stride *= 2;
  }
  int window_start = 0, window_len = 0, hits_length = 0;
  if (size < U) {
    hits_length = U - size;
    window_start = int(NextRandom() % uint32_t(hits_length));
    window_len = std::min(hits_to_extract_, hits_length);
  }
  std::vector<std::pair<int, int>> segs;
  if (window_len > 0) {
    if (window_start + window_len <= hits_length) {
      segs.push_back({window_start, window_len});
    } else {
      segs.push_back({0, window_start + window_len - hits_length});
      segs.push_back({window_start, hits_length - window_start});
    }
  }
  window_.resize(size_t(window_len));
  // Everything the reads below need lies in the prefix fetched with the state
  // when the list is short; otherwise read it as the reference does.
  if (size + hits_length <= prefix_len_) {
    const int* pre = static_cast<const int*>(host_prefix_);
    std::copy(pre, pre + size, host_sorted_.begin());
    int w = 0;
    for (auto& sg : segs) {
      std::copy(pre + size + sg.first, pre + size + sg.first + sg.second, window_.begin() + w);
      w += sg.second;
    }
  } else {
    if (size > 0)
      RT_HIP_CHECK(hipMemcpyAsync(host_sorted_.data(), unique_, size_t(size) * 4, hipMemcpyDeviceToHost, s));
    int w = 0;
    for (auto& sg : segs) {
      RT_HIP_CHECK(hipMemcpyAsync(window_.data() + w, unique_ + size + sg.first, size_t(sg.second) * 4,
                                  hipMemcpyDeviceToHost, s));
      w += sg.second;
    }
    RT_HIP_CHECK(hipStreamSynchronize(s));
  }
  for (int i = 0; i < size; ++i) {
    const int v = host_sorted_[i];
    if (v < 0)
      faults_.push_back(-v - 1);
    else if (v != kNoHit)
      AddHit(v);
  }
  for (int i = 0; i < window_len; ++i)
    if (window_[i] != kNoHit) AddHit(window_[i]);
  if (expected_faults_ > 1 && int(faults_.size()) < expected_faults_)
    expected_faults_ /= 2;
  else if (int(faults_.size()) > expected_faults_)
    expected_faults_ *= 2;
  return RT_OK;
}

}  // namespace rayt

// -------------------------------------------------------------------- C ABI
namespace {
// Finishes the owning tracer's pipelined frame before the extractor's state is read or changed.
int Settle(const rt_feedback* fb) { return (fb && fb->pending) ? rayt::FlushPending(fb->pending) : RT_OK; }
}  // namespace

extern "C" {

int rt_node_to_block(rt_ctx* ctx, int n, int block_size, rt_buf* data) {
  RT_REQUIRE(ctx && data, RT_ERR_INVALID, "null argument");
  RT_REQUIRE(n >= 0 && uint64_t(n) * 4 <= data->size, RT_ERR_RANGE, "n exceeds buffer");
  RT_REQUIRE(block_size > 0, RT_ERR_INVALID, "block_size must be positive");
  if (!n) return RT_OK;
  rayt::NodeToBlockKernel<<<dim3(rayt::Grid(n, 256)), dim3(256), 0, ctx->stream>>>(n, block_size,
                                                                                   static_cast<int*>(data->ptr));
  RT_HIP_CHECK(hipGetLastError());
  return RT_OK;
}

int rt_difference_flag(rt_ctx* ctx, int n, const rt_buf* in, rt_buf* out) {
  RT_REQUIRE(ctx && in && out, RT_ERR_INVALID, "null argument");
  RT_REQUIRE(n >= 0 && uint64_t(n) * 4 <= in->size && uint64_t(n) * 4 <= out->size, RT_ERR_RANGE, "n exceeds buffer");
  if (!n) return RT_OK;
  rayt::DifferenceFlagKernel<<<dim3(rayt::Grid(n, 256)), dim3(256), 0, ctx->stream>>>(
      n, static_cast<const int*>(in->ptr), static_cast<int*>(out->ptr));
  RT_HIP_CHECK(hipGetLastError());
  return RT_OK;
}

int rt_compress(rt_ctx* ctx, int n, const rt_buf* index, const rt_buf* in, rt_buf* out) {
  RT_REQUIRE(ctx && index && in && out, RT_ERR_INVALID, "null argument");
  RT_REQUIRE(n >= 0 && uint64_t(n) * 4 <= in->size && uint64_t(n) * 4 <= index->size && uint64_t(n) * 4 <= out->size,
             RT_ERR_RANGE, "n exceeds buffer");
  if (!n) return RT_OK;
  rayt::CompressKernel<<<dim3(rayt::Grid(n, 256)), dim3(256), 0, ctx->stream>>>(
      n, static_cast<const int*>(index->ptr), static_cast<const int*>(in->ptr), static_cast<int*>(out->ptr));
  RT_HIP_CHECK(hipGetLastError());
  return RT_OK;
}

int rt_merge_pass_i32(rt_ctx* ctx, int n, int length, const rt_buf* in, rt_buf* out) {
  RT_REQUIRE(ctx && in && out, RT_ERR_INVALID, "null argument");
  RT_REQUIRE(n >= 0 && uint64_t(n) * 4 <= in->size && uint64_t(n) * 4 <= out->size, RT_ERR_RANGE,
             "n exceeds buffer");
  RT_REQUIRE(length > 0 && (length & (length - 1)) == 0 && length <= (1 << 29), RT_ERR_RANGE,
             "run length must be a power of two");
  RT_REQUIRE(in->ptr != out->ptr, RT_ERR_INVALID, "a merge pass needs distinct buffers");
  if (!n) return RT_OK;
  rayt::MergePassKernel<<<dim3(rayt::Grid(n, 256)), dim3(256), 0, ctx->stream>>>(
      n, length, static_cast<const int*>(in->ptr), static_cast<int*>(out->ptr));
  RT_HIP_CHECK(hipGetLastError());
  return RT_OK;
}

int rt_scan_pass_i32(rt_ctx* ctx, int n, int length, const rt_buf* in, rt_buf* out) {
  RT_REQUIRE(ctx && in && out, RT_ERR_INVALID, "null argument");
  RT_REQUIRE(n >= 0 && uint64_t(n) * 4 <= in->size && uint64_t(n) * 4 <= out->size, RT_ERR_RANGE,
             "n exceeds buffer");
  RT_REQUIRE(length > 0, RT_ERR_RANGE, "pass length must be positive");
  RT_REQUIRE(in->ptr != out->ptr, RT_ERR_INVALID, "a scan pass needs distinct buffers");
  if (!n) return RT_OK;
  rayt::ScanPassKernel<<<dim3(rayt::Grid(n, 256)), dim3(256), 0, ctx->stream>>>(
      n, length, static_cast<const int*>(in->ptr), static_cast<int*>(out->ptr));
  RT_HIP_CHECK(hipGetLastError());
  return RT_OK;
}

int rt_sort_i32(rt_ctx* ctx, rt_buf* data, int n, rt_buf* temp) {
  RT_REQUIRE(ctx && data && temp, RT_ERR_INVALID, "null argument");
  RT_REQUIRE(n >= 0 && uint64_t(n) * 4 <= data->size && uint64_t(n) * 4 <= temp->size, RT_ERR_RANGE,
             "n exceeds buffer");
  return rayt::SortI32(ctx->stream, static_cast<int*>(data->ptr), n, static_cast<int*>(temp->ptr));
}

int rt_sort_i32_noalloc(rt_ctx* ctx, rt_buf* data, int n) {
  RT_REQUIRE(ctx && data, RT_ERR_INVALID, "null argument");
  RT_REQUIRE(n >= 0 && uint64_t(n) * 4 <= data->size, RT_ERR_RANGE, "n exceeds buffer");
  RT_REQUIRE(n <= rayt::kNoAllocSortMax, RT_ERR_RANGE,
             "n exceeds the no-allocation sort's maximum of RT_SORT_NOALLOC_MAX_N (8388608) elements");
  if (n <= 1) return RT_OK;
  RT_REQUIRE(ctx->device >= 0 && ctx->device < rayt::kMaxScratchDevices, RT_ERR_RANGE, "device index");
  uint32_t* base = nullptr;
  {
    std::lock_guard<std::mutex> lock(rayt::g_scratch_mutex);
    base = rayt::g_scratch_ptr[ctx->device];
  }
  RT_REQUIRE(base, RT_ERR_STATE, "no scratch address for this device (context creation sets it up)");
  rayt::SortPasses(ctx->stream, static_cast<int*>(data->ptr), n, reinterpret_cast<int*>(base),
                   base + rayt::kNoAllocSortMax);
  RT_HIP_CHECK(hipGetLastError());
  return RT_OK;
}

int rt_scan_i32(rt_ctx* ctx, rt_buf* data, int n, rt_buf* temp) {
  RT_REQUIRE(ctx && data, RT_ERR_INVALID, "null argument");
  RT_REQUIRE(n >= 0 && uint64_t(n) * 4 <= data->size, RT_ERR_RANGE, "n exceeds buffer");
  (void)temp;
  return rayt::ScanI32(ctx->stream, static_cast<int*>(data->ptr), n);
}

int rt_feedback_create(rt_ctx* ctx, int count, rt_feedback** out) {
  RT_REQUIRE(ctx && out, RT_ERR_INVALID, "null argument");
  RT_REQUIRE(count >= 0, RT_ERR_RANGE, "negative count");
  RT_TRY_BEGIN
  rt_feedback* f = new rt_feedback();
  try {
    f->fx.reset(new rayt::FeedbackExtractor(ctx, count));
  } catch (...) {
    delete f;
    throw;
  }
  *out = f;
  return RT_OK;
  RT_TRY_END
}

int rt_feedback_destroy(rt_feedback* fb) {
  (void)Settle(fb);
  delete fb;
  return RT_OK;
}

int rt_feedback_set_hits_to_extract(rt_feedback* fb, int value) {
  RT_REQUIRE(fb, RT_ERR_INVALID, "null feedback");
  if (const int rc = Settle(fb)) return rc;
  RT_REQUIRE(value >= 0, RT_ERR_RANGE, "negative hits_to_extract");
  fb->fx->set_hits_to_extract(value);
  return RT_OK;
}

int rt_feedback_set_seed(rt_feedback* fb, uint64_t seed) {
  RT_REQUIRE(fb, RT_ERR_INVALID, "null feedback");
  if (const int rc = Settle(fb)) return rc;
  fb->fx->set_seed(seed);
  return RT_OK;
}

int rt_feedback_process(rt_feedback* fb, rt_buf* codes, int nodes_in_block, int block_level) {
  RT_REQUIRE(fb && codes, RT_ERR_INVALID, "null argument");
  if (const int rc = Settle(fb)) return rc;
  RT_REQUIRE(uint64_t(fb->fx->count()) * 4 <= codes->size, RT_ERR_RANGE, "codes buffer smaller than count");
  RT_REQUIRE(block_level || nodes_in_block > 0, RT_ERR_INVALID, "nodes_in_block must be positive");
  RT_TRY_BEGIN
  return fb->fx->Process(static_cast<int*>(codes->ptr), nodes_in_block, block_level != 0);
  RT_TRY_END
}

int rt_feedback_list(const rt_feedback* fb, int which, int* out, int cap) {
  RT_REQUIRE(fb, RT_ERR_INVALID, "null feedback");
  if (const int rc = Settle(fb)) return rc;
  const std::vector<int>* v = which == 0 ? &fb->fx->fault_blocks() : which == 1 ? &fb->fx->hit_blocks()
                                                                               : &fb->fx->duplicate_hit_blocks();
  RT_REQUIRE(which >= 0 && which <= 2, RT_ERR_RANGE, "which must be 0, 1 or 2");
  for (int i = 0; i < cap && i < int(v->size()); ++i) out[i] = (*v)[i];
  return int(v->size());
}

int rt_feedback_stats(const rt_feedback* fb, int* unique_count, int* expected_faults) {
  RT_REQUIRE(fb, RT_ERR_INVALID, "null feedback");
  if (const int rc = Settle(fb)) return rc;
  if (unique_count) *unique_count = fb->fx->unique_count();
  if (expected_faults) *expected_faults = fb->fx->expected_faults();
  return RT_OK;
}

}  // extern "C"
