// Stored / cache octree formats, block records and block sources.
//
// Link layouts (CPU data/octree-constants.h:7-53, doc/storage data format.txt,
// doc/cache data format.txt):
//   stored: ptr<<11 | dup<<9 | fault<<8 | children_mask
//   cache : ptr<<11 | far<<10 | dup<<9 | fault<<8 | children_mask
// A near cache pointer is 2^20 + (pointed - pointer); a far one indexes the
// far-pointer table. Offsets into files and pools are 64-bit here (the
// reference keeps them in int, which caps pools at 2 GB).
#pragma once

#ifdef __SSE2__
#include <emmintrin.h>
#endif

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

namespace rayt {

constexpr const char* kNodeLinkChannelName = "node links";
constexpr int kNodeLinkSize = 4;          // octree-constants.h:8
constexpr int kBlockHeaderSize = 76;      // octree-constants.h:9
constexpr uint32_t kRootBlockParent = (1u << 24) - 1;  // octree-constants.h:10
constexpr int kMaxBlocksCount = (1 << 23) - 1;         // octree-constants.h:11
constexpr uint32_t kMaxLinkPtr = 1u << 21;

inline uint32_t PackStoredNodeLink(uint32_t children_mask, bool fault, bool dup, uint32_t ptr) {
  return (ptr << 11) | ((dup ? 1u : 0u) << 9) | ((fault ? 1u : 0u) << 8) | (children_mask & 255u);
}
inline uint32_t LinkMask(uint32_t link) { return link & 255u; }
inline bool LinkFault(uint32_t link) { return (link >> 8) & 1u; }

// Calls f(i, link) for every link among links[0, n) with children (mask != 0)
// that are not loaded (fault bit): the child blocks of a block. A block has
// few of them among thousands of links, so the scan tests four links per
// SSE2 compare, (link & 0x1ff) > 0x100, and visits only the hits.
template <class F>
inline void ForEachFaultLink(const uint32_t* links, int n, F f) {
  int i = 0;
#ifdef __SSE2__
  const __m128i low9 = _mm_set1_epi32(0x1ff), fault_only = _mm_set1_epi32(0x100);
  for (; i + 16 <= n; i += 16) {
    int bits = 0;
    for (int k = 0; k < 4; ++k) {
      const __m128i v = _mm_loadu_si128(reinterpret_cast<const __m128i*>(links + i + 4 * k));
      const __m128i hit = _mm_cmpgt_epi32(_mm_and_si128(v, low9), fault_only);
      bits |= _mm_movemask_ps(_mm_castsi128_ps(hit)) << (4 * k);
    }
    while (bits) {
      const int k = __builtin_ctz(bits);
      bits &= bits - 1;
      f(i + k, links[i + k]);
    }
  }
#endif
  for (; i < n; ++i)
    if (LinkMask(links[i]) && LinkFault(links[i])) f(i, links[i]);
}
inline bool LinkDup(uint32_t link) { return (link >> 9) & 1u; }
inline bool LinkFar(uint32_t link) { return (link >> 10) & 1u; }
inline uint32_t LinkPtr(uint32_t link) { return link >> 11; }

// octree-constants.h:33-39
inline bool NeedFarPointer(int64_t pointer_index, int64_t pointed_index) {
  int64_t d = pointer_index - pointed_index;
  return (d < 0 ? -d : d) >= (1 << 20);
}
inline uint32_t NearPointerValue(int64_t pointer_index, int64_t pointed_index) {
  return static_cast<uint32_t>((1 << 20) + (pointed_index - pointer_index));
}
inline uint32_t PackCacheNodeLink(uint32_t children_mask, bool fault, bool dup, bool far_flag, uint32_t ptr) {
  return (ptr << 11) | ((far_flag ? 1u : 0u) << 10) | ((dup ? 1u : 0u) << 9) |
         ((fault ? 1u : 0u) << 8) | (children_mask & 255u);
}

struct StoredOctreeChannel {
  int bytes_in_node = 0;
  std::string name;
  bool operator==(const StoredOctreeChannel& o) const {
    return bytes_in_node == o.bytes_in_node && name == o.name;
  }
};

struct StoredOctreeHeader {
  uint32_t nodes_in_block = 0;
  uint32_t blocks_count = 0;
  uint32_t root_block_index = 0;
  uint32_t block_stride = 0;  // bytes per block record (header + payload + padding)
  std::vector<StoredOctreeChannel> channels;  // channels[0] is "node links"
  int SumBytesInNode() const {
    int s = 0;
    for (auto& c : channels) s += c.bytes_in_node;
    return s;
  }
  int ChannelIndex(const std::string& name) const {
    for (size_t i = 0; i < channels.size(); ++i)
      if (channels[i].name == name) return static_cast<int>(i);
    return -1;
  }
  // Byte offset of channel c inside a block payload (payload is SoA per block:
  // all node links, then channel 1 for every node, ...).
  uint64_t ChannelOffsetInBlock(int c) const {
    uint64_t off = 0;
    for (int i = 0; i < c; ++i) off += uint64_t(channels[i].bytes_in_node) * nodes_in_block;
    return off;
  }
};

struct StoredOctreeBlockRoot {
  uint16_t parent_pointer_index = 0;  // within parent block
  uint16_t pointed_child_index = 0;   // within this block
  uint32_t parent_pointer_value = 0;  // stored link of the parent node when this block is absent
};

struct StoredOctreeBlockHeader {
  uint32_t block_index = 0;
  uint32_t parent_block_index = 0;
  uint32_t roots_count = 0;
  StoredOctreeBlockRoot roots[8];
};

struct StoredOctreeBlock {
  std::vector<uint8_t> data;  // channel payload, SoA inside the block
  StoredOctreeBlockHeader header;
  const uint32_t* links() const { return reinterpret_cast<const uint32_t*>(data.data()); }
  uint32_t* links() { return reinterpret_cast<uint32_t*>(data.data()); }
};

// Serialises a block header into its 76-byte record (little-endian).
void EncodeBlockHeader(const StoredOctreeBlockHeader& h, uint8_t* out76);
void DecodeBlockHeader(const uint8_t* in76, StoredOctreeBlockHeader* h);
// Serialises the file header (20 bytes + channel records).
std::vector<uint8_t> EncodeFileHeader(const StoredOctreeHeader& h);

// Anything blocks can be loaded from: a .tree file, or a tree held in host RAM.
class BlockSource {
 public:
  virtual ~BlockSource() = default;
  virtual const StoredOctreeHeader& header() const = 0;
  // Fills out_block (resized when needed); false on I/O failure or bad index.
  virtual bool LoadBlock(uint32_t index, StoredOctreeBlock* out_block) = 0;
  // Pointer to the raw record (76-byte header + payload) when the source keeps
  // records in memory, else nullptr.
  virtual const uint8_t* RecordPtr(uint32_t) const { return nullptr; }
  uint64_t BytesInBlockContent() const {
    return uint64_t(header().SumBytesInNode()) * header().nodes_in_block;
  }
};

// Reads a .tree file (storage data format.txt). Mirrors StoredOctreeLoader
// (CPU data/stored-octree-loader.cpp:9-98) with 64-bit file offsets.
class StoredOctreeLoader : public BlockSource {
 public:
  explicit StoredOctreeLoader(const std::string& file_name);
  ~StoredOctreeLoader() override;
  bool ok() const { return ok_; }
  const std::string& error() const { return error_; }
  const StoredOctreeHeader& header() const override { return header_; }
  bool LoadBlock(uint32_t index, StoredOctreeBlock* out_block) override;

 private:
  FILE* f_ = nullptr;
  bool ok_ = false;
  std::string error_;
  StoredOctreeHeader header_;
  uint64_t header_size_ = 0;
};

// Where a builder puts its blocks: one record (76-byte header + payload) per
// block index, in any order. Put returns false on failure (e.g. I/O).
class BlockSink {
 public:
  virtual ~BlockSink() = default;
  virtual bool Put(uint32_t index, std::unique_ptr<uint8_t[]> record) = 0;
};

// Writes block records straight into a .tree file at their final offsets
// (header + index * stride), so a build need not hold the tree in memory.
// Finish writes the file header once the block count is known.
class TreeFileWriter : public BlockSink {
 public:
  TreeFileWriter(const std::string& path, const StoredOctreeHeader& header);
  ~TreeFileWriter() override;
  bool ok() const { return f_ != nullptr; }
  bool Put(uint32_t index, std::unique_ptr<uint8_t[]> record) override;
  bool Finish(uint32_t blocks_count, std::string* err);

 private:
  FILE* f_ = nullptr;
  StoredOctreeHeader header_;
  uint64_t header_size_ = 0, record_size_ = 0;
  bool failed_ = false;
};

// A complete tree in host RAM, one record per block in file layout. This is
// the host-side block store for out-of-core streaming, and the builder's
// output. WriteFile produces a .tree file readable by StoredOctreeLoader.
class MemoryTree : public BlockSource, public BlockSink {
 public:
  MemoryTree() = default;
  const StoredOctreeHeader& header() const override { return header_; }
  StoredOctreeHeader& mutable_header() { return header_; }
  bool LoadBlock(uint32_t index, StoredOctreeBlock* out_block) override;
  const uint8_t* RecordPtr(uint32_t index) const override {
    return index < records_.size() ? records_[index].get() : nullptr;
  }
  // Record layout: 76-byte header + payload (BytesInBlockContent()).
  uint64_t record_size() const { return kBlockHeaderSize + BytesInBlockContent(); }
  void PutBlock(const StoredOctreeBlockHeader& h, const uint8_t* payload);
  bool Put(uint32_t index, std::unique_ptr<uint8_t[]> record) override;
  bool WriteFile(const std::string& path, std::string* err) const;
  uint64_t total_nodes = 0;  // nodes in the source tree (informational)

 private:
  StoredOctreeHeader header_;
  std::vector<std::unique_ptr<uint8_t[]>> records_;
};

}  // namespace rayt
