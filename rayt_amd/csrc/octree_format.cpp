// .tree file codec and block sources (CPU data/stored-octree-loader.cpp:9-98,
// CPU data/stored-octree-raw-writer.cpp:10-99, doc/storage data format.txt).
#include "octree_format.h"

#include <cerrno>

namespace rayt {

static inline void put_u32(uint8_t* p, uint32_t v) { memcpy(p, &v, 4); }
static inline void put_u16(uint8_t* p, uint16_t v) { memcpy(p, &v, 2); }
static inline uint32_t get_u32(const uint8_t* p) { uint32_t v; memcpy(&v, p, 4); return v; }
static inline uint16_t get_u16(const uint8_t* p) { uint16_t v; memcpy(&v, p, 2); return v; }

void EncodeBlockHeader(const StoredOctreeBlockHeader& h, uint8_t* out) {
  put_u32(out + 0, h.block_index);
  put_u32(out + 4, h.parent_block_index);
  put_u32(out + 8, h.roots_count);
  for (int i = 0; i < 8; ++i) {
    put_u16(out + 12 + i * 8 + 0, h.roots[i].parent_pointer_index);
    put_u16(out + 12 + i * 8 + 2, h.roots[i].pointed_child_index);
    put_u32(out + 12 + i * 8 + 4, h.roots[i].parent_pointer_value);
  }
}

void DecodeBlockHeader(const uint8_t* in, StoredOctreeBlockHeader* h) {
  h->block_index = get_u32(in + 0);
  h->parent_block_index = get_u32(in + 4);
  h->roots_count = get_u32(in + 8);
  for (int i = 0; i < 8; ++i) {
    h->roots[i].parent_pointer_index = get_u16(in + 12 + i * 8 + 0);
    h->roots[i].pointed_child_index = get_u16(in + 12 + i * 8 + 2);
    h->roots[i].parent_pointer_value = get_u32(in + 12 + i * 8 + 4);
  }
}

std::vector<uint8_t> EncodeFileHeader(const StoredOctreeHeader& h) {
  std::vector<uint8_t> out(20);
  put_u32(&out[0], h.nodes_in_block);
  put_u32(&out[4], h.blocks_count);
  put_u32(&out[8], h.root_block_index);
  put_u32(&out[12], h.block_stride);
  put_u32(&out[16], static_cast<uint32_t>(h.channels.size()));
  for (auto& c : h.channels) {
    size_t p = out.size();
    out.resize(p + 8 + c.name.size());
    put_u32(&out[p], static_cast<uint32_t>(c.bytes_in_node));
    put_u32(&out[p + 4], static_cast<uint32_t>(c.name.size()));
    memcpy(&out[p + 8], c.name.data(), c.name.size());
  }
  return out;
}

static bool pread_all(FILE* f, uint64_t pos, size_t len, void* dst) {
  if (fseeko(f, static_cast<off_t>(pos), SEEK_SET) != 0) return false;
  return fread(dst, 1, len, f) == len;
}

StoredOctreeLoader::StoredOctreeLoader(const std::string& file_name) {
  f_ = fopen(file_name.c_str(), "rb");
  if (!f_) {
    error_ = "cannot open " + file_name;
    return;
  }
  uint8_t fixed[20];
  if (!pread_all(f_, 0, 20, fixed)) {
    error_ = "failed to read header";
    return;
  }
  header_.nodes_in_block = get_u32(fixed + 0);
  header_.blocks_count = get_u32(fixed + 4);
  header_.root_block_index = get_u32(fixed + 8);
  header_.block_stride = get_u32(fixed + 12);
  uint32_t channels = get_u32(fixed + 16);
  uint64_t pos = 20;
  if (channels == 0 || channels > 64) {
    error_ = "bad channel count";
    return;
  }
  for (uint32_t i = 0; i < channels; ++i) {
    uint8_t rec[8];
    if (!pread_all(f_, pos, 8, rec)) {
      error_ = "failed to read channel header";
      return;
    }
    pos += 8;
    StoredOctreeChannel c;
    c.bytes_in_node = static_cast<int>(get_u32(rec));
    uint32_t l = get_u32(rec + 4);
    if (l > 4096) {
      error_ = "bad channel name length";
      return;
    }
    c.name.resize(l);
    if (l && !pread_all(f_, pos, l, &c.name[0])) {
      error_ = "failed to read channel name";
      return;
    }
    pos += l;
    header_.channels.push_back(c);
  }
  header_size_ = pos;
  if (header_.channels[0].name != kNodeLinkChannelName || header_.channels[0].bytes_in_node != kNodeLinkSize) {
    error_ = "first channel must be 4-byte node links";
    return;
  }
  if (header_.block_stride < kBlockHeaderSize + BytesInBlockContent()) {
    error_ = "block stride smaller than block record";
    return;
  }
  ok_ = true;
}

StoredOctreeLoader::~StoredOctreeLoader() {
  if (f_) fclose(f_);
}

bool StoredOctreeLoader::LoadBlock(uint32_t index, StoredOctreeBlock* out) {
  if (!ok_ || index >= header_.blocks_count || !out) return false;
  uint64_t pos = header_size_ + uint64_t(header_.block_stride) * index;
  uint8_t hdr[kBlockHeaderSize];
  if (!pread_all(f_, pos, kBlockHeaderSize, hdr)) return false;
  DecodeBlockHeader(hdr, &out->header);
  uint64_t content = BytesInBlockContent();
  if (out->data.size() != content) out->data.resize(content);
  return pread_all(f_, pos + kBlockHeaderSize, content, out->data.data());
}

bool MemoryTree::LoadBlock(uint32_t index, StoredOctreeBlock* out) {
  if (index >= records_.size() || !records_[index] || !out) return false;
  const uint8_t* rec = records_[index].get();
  DecodeBlockHeader(rec, &out->header);
  uint64_t content = BytesInBlockContent();
  if (out->data.size() != content) out->data.resize(content);
  memcpy(out->data.data(), rec + kBlockHeaderSize, content);
  return true;
}

void MemoryTree::PutBlock(const StoredOctreeBlockHeader& h, const uint8_t* payload) {
  if (records_.size() <= h.block_index) records_.resize(h.block_index + 1);
  uint64_t rs = record_size();
  std::unique_ptr<uint8_t[]> rec(new uint8_t[rs]);
  EncodeBlockHeader(h, rec.get());
  memcpy(rec.get() + kBlockHeaderSize, payload, BytesInBlockContent());
  records_[h.block_index] = std::move(rec);
}

bool MemoryTree::Put(uint32_t index, std::unique_ptr<uint8_t[]> record) {
  if (records_.size() <= index) records_.resize(size_t(index) + 1);
  records_[index] = std::move(record);
  return true;
}

TreeFileWriter::TreeFileWriter(const std::string& path, const StoredOctreeHeader& header) : header_(header) {
  f_ = fopen(path.c_str(), "w+b");
  record_size_ = kBlockHeaderSize + uint64_t(header.SumBytesInNode()) * header.nodes_in_block;
  header_.block_stride = static_cast<uint32_t>(record_size_);
  header_size_ = EncodeFileHeader(header_).size();
}

TreeFileWriter::~TreeFileWriter() {
  if (f_) fclose(f_);
}

bool TreeFileWriter::Put(uint32_t index, std::unique_ptr<uint8_t[]> record) {
  if (!f_ || failed_) return false;
  const uint64_t pos = header_size_ + uint64_t(index) * record_size_;
  if (fseeko(f_, off_t(pos), SEEK_SET) != 0 || fwrite(record.get(), 1, record_size_, f_) != record_size_)
    failed_ = true;
  return !failed_;
}

bool TreeFileWriter::Finish(uint32_t blocks_count, std::string* err) {
  if (!f_ || failed_) {
    if (err) *err = "writing the tree file failed";
    return false;
  }
  header_.blocks_count = blocks_count;
  const std::vector<uint8_t> hb = EncodeFileHeader(header_);
  bool ok = fseeko(f_, 0, SEEK_SET) == 0 && fwrite(hb.data(), 1, hb.size(), f_) == hb.size();
  ok = (fclose(f_) == 0) && ok;
  f_ = nullptr;
  if (!ok && err) *err = "writing the tree file header failed";
  return ok;
}

bool MemoryTree::WriteFile(const std::string& path, std::string* err) const {
  FILE* f = fopen(path.c_str(), "wb");
  if (!f) {
    if (err) *err = "cannot create " + path;
    return false;
  }
  StoredOctreeHeader h = header_;
  h.block_stride = static_cast<uint32_t>(record_size());
  std::vector<uint8_t> hb = EncodeFileHeader(h);
  bool ok = fwrite(hb.data(), 1, hb.size(), f) == hb.size();
  for (uint32_t i = 0; ok && i < h.blocks_count; ++i) {
    if (i >= records_.size() || !records_[i]) {
      ok = false;
      if (err) *err = "missing block record";
      break;
    }
    ok = fwrite(records_[i].get(), 1, record_size(), f) == record_size();
  }
  if (fclose(f) != 0) ok = false;
  if (!ok && err && err->empty()) *err = "write failed";
  return ok;
}

}  // namespace rayt
