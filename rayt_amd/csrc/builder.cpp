// BFS blocking builder (see builder.h). The sibling-group data structure, the
// order in which groups are expanded, the block-index allocation and the
// duplicate-subtree rule follow stored-octree-builder-impl.h:46-521 step by
// step; the node source is a precomputed NodeTree instead of a callback data
// source, so node data is copied rather than computed on demand.
#include "builder.h"

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <mutex>
#include <optional>
#include <queue>
#include <stdexcept>
#include <thread>

namespace rayt {

namespace {

struct Group;

struct BNode {
  uint32_t tree_node = 0;
  int children_mask = -1;  // -1: not expanded yet
  Group* child = nullptr;
  int fault_block = -1;
  bool duplicate = false;
  int frontier_id = -1;  // parallel build: this node's subtree is built by a worker (FrontierResult index)

  uint32_t NodeLink(int child_index_in_block) const {
    if (!children_mask) return PackStoredNodeLink(0, false, false, 0);
    uint32_t ptr = child ? uint32_t(child_index_in_block) : uint32_t(fault_block);
    return PackStoredNodeLink(uint32_t(children_mask), !child, duplicate, ptr);
  }
};

// The nodes of a sibling group (the children of one node: at most 8), held in
// the group itself so that a group is one allocation.
class GroupNodes {
 public:
  size_t size() const { return n_; }
  void push_back(const BNode& b) {
    if (n_ == 8) throw std::logic_error("sibling group of more than 8 nodes");
    v_[n_++] = b;
  }
  BNode& back() { return v_[n_ - 1]; }
  BNode& operator[](size_t i) { return v_[i]; }
  const BNode& operator[](size_t i) const { return v_[i]; }
  BNode* begin() { return v_; }
  BNode* end() { return v_ + n_; }
  const BNode* begin() const { return v_; }
  const BNode* end() const { return v_ + n_; }

 private:
  BNode v_[8];
  size_t n_ = 0;
};

struct Group {
  int index_in_block = -1;
  const NodeTree* src = nullptr;  // the tree the nodes' tree_node ids index
  GroupNodes nodes;
};

struct Subtree {
  Group* allocated_root = nullptr;  // what to free after commit
  Group* root = nullptr;            // root group inside the block
  Group* parent = nullptr;          // group holding the parent node (in the parent block)
  int child_index = 0;
  int node_count = 0;
};

struct BlockingState {
  std::optional<Subtree> root_subtree;
  std::vector<Subtree> remaining;
};

// What the parallel build's replay learns about a frontier node (a node of
// the root block whose subtree a worker built): no children, a subtree small
// enough to be placed by the parent side, or blocks already committed.
struct FrontierOutcome {
  bool has_children = false;
  bool full = false;  // committed as blocks; the node's fault_block is set
  std::optional<Subtree> small;
};

class Builder {
 public:
  // `channels`: the data channels (without node links) of every node source.
  Builder(const std::vector<StoredOctreeChannel>& channels, int nib, BlockSink* out)
      : channels_(channels), nib_(nib), out_(out) {
    int data_size = 0;
    for (auto& c : channels) data_size += c.bytes_in_node;
    payload_bytes_ = size_t(nib) * (kNodeLinkSize + data_size);
    ch_offsets_.push_back(size_t(nib) * kNodeLinkSize);
    for (size_t c = 1; c < channels.size(); ++c)
      ch_offsets_.push_back(ch_offsets_.back() + size_t(nib) * channels[c - 1].bytes_in_node);
  }

  void Build(const NodeTree& t) {
    Group* root = Alloc();
    root->src = &t;
    root->nodes.push_back(BNode());
    root->nodes.back().tree_node = 0;
    int block_index = blocks_count_++;
    auto s = BuildRecursive(root, nullptr, 0, block_index, kRootBlockParent);
    if (s) {
      --blocks_count_;
      CommitBlock({*s}, blocks_count_++, kRootBlockParent);
      ++leaf_blocks_;
    }
  }

  BuildReport report;
  int blocks_count_ = 0;
  int leaf_blocks_ = 0, non_leaf_blocks_ = 0;

  // Parallel build hooks. expand_hook: called before a node of `lazy_tree` is
  // expanded (the tree grows on demand). frontier_hook: builds the subtree
  // below a node with frontier_id >= 0 instead of this builder. node_done:
  // called once ProcessBlockNodes has finished a node (its subtree is built).
  const NodeTree* lazy_tree = nullptr;
  std::function<void(uint32_t)> expand_hook;
  std::function<FrontierOutcome(Group*, int)> frontier_hook;
  std::function<void(Group*, int)> node_done;

  Group* Alloc() {
    ++report.allocated_groups;
    return new Group();
  }
  void DeallocSubtree(Group* g) {
    for (auto& n : g->nodes)
      if (n.child) DeallocSubtree(n.child);
    Dealloc(g);
  }

  // Expands node n of group-source `src`: sets its children mask and, when
  // it has children, a new group with one entry per child in bit order.
  // Returns that group.
  Group* Expand(BNode& n, const NodeTree* src) {
    if (src == lazy_tree && expand_hook) expand_hook(n.tree_node);
    uint8_t m = src->mask[n.tree_node];
    n.children_mask = m;
    if (!m) return nullptr;
    Group* g = Alloc();
    g->src = src;
    uint32_t c = src->first_child[n.tree_node];
    for (int j = 0; j < 8; ++j)
      if (m & (1u << j)) {
        BNode b;
        b.tree_node = c++;
        g->nodes.push_back(b);
      }
    return g;
  }

  // The breadth-first fill of a block from `root` (BuildRecursive up to the
  // point where the block is full): *full says whether it stopped there.
  Subtree BfsFill(Group* root, Group* parent, int child_index, bool* full) {
    Subtree res;
    res.child_index = child_index;
    res.parent = parent;
    res.root = res.allocated_root = root;
    res.node_count = int(root->nodes.size());
    std::vector<Group*>& qu = bfs_queue_;  // FIFO: entries [head, size)
    qu.clear();
    qu.push_back(root);
    size_t head = 0;
    *full = false;
    while (head < qu.size() && !*full) {
      Group* cur = qu[head++];
      for (size_t i = 0; i < cur->nodes.size(); ++i) {
        BNode& n = cur->nodes[i];
        if (res.node_count + 8 <= nib_) {
          Group* g = Expand(n, cur->src);
          if (g) {
            n.child = g;
            qu.push_back(g);
            res.node_count += int(g->nodes.size());
          }
        } else {
          *full = true;
          break;
        }
      }
    }
    return res;
  }

  // The rest of BuildRecursive for a full block: its frontier subtrees, the
  // leaf block of its leftovers, then the block itself.
  void FinishFullBlock(Subtree& res, int block_index, uint32_t parent_block_index) {
    BlockingState s = ProcessBlockNodes(res.root, block_index, nullptr, 0);
    if (!s.remaining.empty()) {
      CommitBlock(s.remaining, blocks_count_++, uint32_t(block_index));
      ++leaf_blocks_;
    }
    CommitBlock({res}, block_index, parent_block_index);
    ++non_leaf_blocks_;
  }

  void FillIndexInBlockRecursive(Group* g, int& index) {
    if (g->index_in_block != -1) throw std::logic_error("group indexed twice");
    g->index_in_block = index;
    index += int(g->nodes.size());
    for (auto& n : g->nodes)
      if (n.child) FillIndexInBlockRecursive(n.child, index);
  }

  std::optional<Subtree> BuildRecursive(Group* root, Group* parent, int child_index, int block_index,
                                        uint32_t parent_block_index) {
    bool full = false;
    Subtree res = BfsFill(root, parent, child_index, &full);
    if (full) {
      int idx = 0;
      FillIndexInBlockRecursive(res.root, idx);
      FinishFullBlock(res, block_index, parent_block_index);
      return std::nullopt;
    }
    return res;
  }

  void CommitBlock(const std::vector<Subtree>& subtrees, int block_index, uint32_t parent_block_index) {
    if (log_ && ++committed_ % 16384 == 0) {  // progress of long builds (RAYT_BUILD_LOG)
      std::fprintf(stderr, "[rayt build] %d blocks committed\n", committed_);
      std::fflush(stderr);
    }
    if (subtrees.empty() || subtrees.size() > 8) throw std::logic_error("block with 0 or >8 roots");
    if (uint32_t(block_index) >= kMaxLinkPtr) throw std::runtime_error("more than 2^21 blocks");
    int node_count = 0;
    hdr_ = StoredOctreeBlockHeader();
    hdr_.block_index = uint32_t(block_index);
    hdr_.parent_block_index = parent_block_index;
    hdr_.roots_count = uint32_t(subtrees.size());
    const uint64_t rs = kBlockHeaderSize + payload_bytes_;
    std::unique_ptr<uint8_t[]> rec(new uint8_t[rs]);
    uint8_t* payload = rec.get() + kBlockHeaderSize;
    memset(payload, 0, payload_bytes_);
    for (size_t si = 0; si < subtrees.size(); ++si) {
      const Subtree& st = subtrees[si];
      if (st.root->index_in_block != -1) {
        if (subtrees.size() != 1) throw std::logic_error("pre-indexed subtree shares a block");
      } else {
        int idx = node_count;
        FillIndexInBlockRecursive(st.root, idx);
      }
      if (st.parent) SetFaultBlock(st.parent->nodes[st.child_index], block_index);
      node_count += st.node_count;
      hdr_.roots[si].pointed_child_index = uint16_t(st.root->index_in_block);
      if (st.parent) {
        const BNode& pn = st.parent->nodes[st.child_index];
        hdr_.roots[si].parent_pointer_index = uint16_t(st.parent->index_in_block + st.child_index);
        hdr_.roots[si].parent_pointer_value = pn.NodeLink(pn.child ? pn.child->index_in_block : 0);
      }
      PackBlockRecursive(st.root, payload);
      DeallocSubtree(st.allocated_root);
    }
    if (node_count > nib_) throw std::logic_error("block overflow");
    EncodeBlockHeader(hdr_, rec.get());
    if (!out_->Put(uint32_t(block_index), std::move(rec))) throw std::runtime_error("block output failed");
  }

 private:
  std::vector<StoredOctreeChannel> channels_;
  int nib_;
  BlockSink* out_;
  bool log_ = std::getenv("RAYT_BUILD_LOG") != nullptr;
  int committed_ = 0;
  size_t payload_bytes_ = 0;
  std::vector<size_t> ch_offsets_;
  std::vector<Group*> bfs_queue_;
  StoredOctreeBlockHeader hdr_;

  void Dealloc(Group* g) {
    ++report.deallocated_groups;
    delete g;
  }

  void SetFaultBlock(BNode& n, int block) {
    if (n.children_mask && !n.child) n.fault_block = block;
    if (n.child)
      for (auto& c : n.child->nodes) SetFaultBlock(c, block);
  }

  // Copies one channel value (sizes 1, 2, 4 and 8 as fixed-size moves).
  static void CopyValue(uint8_t* dst, const uint8_t* src, int s) {
    switch (s) {
      case 1: *dst = *src; break;
      case 2: memcpy(dst, src, 2); break;
      case 4: memcpy(dst, src, 4); break;
      case 8: memcpy(dst, src, 8); break;
      default: memcpy(dst, src, size_t(s));
    }
  }

  void PackBlockRecursive(Group* g, uint8_t* payload) {
    for (size_t ni = 0; ni < g->nodes.size(); ++ni) {
      BNode& n = g->nodes[ni];
      int idx = g->index_in_block + int(ni);
      uint32_t link = n.NodeLink(n.child ? n.child->index_in_block : 0);
      memcpy(&payload[size_t(idx) * kNodeLinkSize], &link, 4);
      const uint8_t* src = &g->src->data[size_t(n.tree_node) * g->src->data_size];
      size_t off = 0;
      for (size_t c = 0; c < channels_.size(); ++c) {
        const int s = channels_[c].bytes_in_node;
        CopyValue(&payload[ch_offsets_[c] + size_t(s) * idx], src + off, s);
        off += s;
      }
      if (n.child) PackBlockRecursive(n.child, payload);
    }
  }

  static int SumSize(const std::vector<Subtree>& a) {
    int r = 0;
    for (auto& s : a) r += s.node_count;
    return r;
  }

  // Greedy packing of subtree groups into leaf blocks; leaves fewer than 8
  // subtrees and fewer than nodes_in_block nodes in out_remainder.
  void GroupIntoBlocks(std::vector<std::vector<Subtree>> groups, uint32_t parent_block_index,
                       std::vector<Subtree>& out_remainder) {
    int total_nodes = 0, total_subtrees = 0;
    for (auto& g : groups) total_nodes += SumSize(g), total_subtrees += int(g.size());
    if (total_nodes < nib_ && total_subtrees < 8) {
      for (auto& g : groups) out_remainder.insert(out_remainder.end(), g.begin(), g.end());
      return;
    }
    // Ascending by node count. Ties keep their original order (a stable sort,
    // where the reference's std::sort leaves tie order to the library): the
    // (count, original position) keys are distinct, so a plain sort of them
    // gives the stable order.
    std::vector<std::pair<int, int>> order(groups.size());  // (node count, original index)
    for (size_t i = 0; i < groups.size(); ++i) order[i] = {SumSize(groups[i]), int(i)};
    std::sort(order.begin(), order.end());
    std::vector<bool> taken(order.size(), false);
    while (total_nodes >= nib_ || total_subtrees >= 8) {
      int gn = 0, gs = 0;
      std::vector<Subtree> group;
      for (int i = int(order.size()) - 1; i >= 0; --i) {
        if (taken[i]) continue;
        const std::vector<Subtree>& g = groups[order[i].second];
        const int n = order[i].first, s = int(g.size());
        if (gn + n <= nib_ && gs + s <= 8) {
          gn += n;
          gs += s;
          group.insert(group.end(), g.begin(), g.end());
          taken[i] = true;
        }
      }
      if (group.empty()) throw std::logic_error("subtree group does not fit a block");
      CommitBlock(group, blocks_count_++, parent_block_index);
      ++leaf_blocks_;
      total_nodes -= gn;
      total_subtrees -= gs;
    }
    for (size_t i = 0; i < order.size(); ++i)
      if (!taken[i]) {
        const std::vector<Subtree>& g = groups[order[i].second];
        out_remainder.insert(out_remainder.end(), g.begin(), g.end());
      }
  }

 public:
  BlockingState ProcessBlockNodes(Group* root, int block_index, Group* parent, int child_index) {
    bool all_roots = true;
    std::vector<Subtree> root_subtrees(root->nodes.size());
    std::vector<Group*> useless_roots;
    std::vector<std::vector<Subtree>> subtree_groups;
    int rooted_nodes_count = int(root->nodes.size());

    for (size_t ni = 0; ni < root->nodes.size(); ++ni) {
      root_subtrees[ni].root = nullptr;
      // (a worker may be writing a frontier node's children mask: only its frontier id is read before the
      // hook has waited for the worker)
      if (root->nodes[ni].frontier_id >= 0 && frontier_hook) {
        FrontierOutcome o = frontier_hook(root, int(ni));
        if (o.full) {
          all_roots = false;
        } else if (o.has_children) {
          root_subtrees[ni] = *o.small;
          subtree_groups.push_back({*o.small});
          rooted_nodes_count += o.small->node_count;
        }
      } else if (root->nodes[ni].children_mask == -1) {
        Group* g = Expand(root->nodes[ni], root->src);
        if (g) {
          int new_block_index = blocks_count_++;
          auto r = BuildRecursive(g, root, int(ni), new_block_index, uint32_t(block_index));
          if (!r) {
            root->nodes[ni].fault_block = new_block_index;
            all_roots = false;
          } else {
            --blocks_count_;
            root_subtrees[ni] = *r;
            subtree_groups.push_back({*r});
            rooted_nodes_count += r->node_count;
          }
        }
      }
      BNode& n = root->nodes[ni];
      if (n.children_mask == 0) {
        root_subtrees[ni].root = nullptr;
        root_subtrees[ni].node_count = 0;
      } else if (n.child) {
        BlockingState s = ProcessBlockNodes(n.child, block_index, root, int(ni));
        if (!s.root_subtree) {
          all_roots = false;
        } else {
          root_subtrees[ni] = *s.root_subtree;
          rooted_nodes_count += s.root_subtree->node_count;
          if (s.remaining.empty()) useless_roots.push_back(s.root_subtree->allocated_root);
        }
        if (!s.remaining.empty()) subtree_groups.push_back(s.remaining);
      }
      if (node_done) node_done(root, int(ni));
    }

    BlockingState res;
    if (all_roots && rooted_nodes_count < nib_ && parent) {
      // The whole group and everything pending below it fits one block: offer a
      // duplicate of it to the parent side.
      Subtree ns;
      ns.child_index = child_index;
      ns.parent = parent;
      ns.node_count = rooted_nodes_count;
      ns.root = ns.allocated_root = Alloc();
      ns.root->nodes = root->nodes;
      ns.root->src = root->src;
      ns.root->index_in_block = -1;
      for (size_t i = 0; i < ns.root->nodes.size(); ++i) {
        BNode& c = ns.root->nodes[i];
        c.duplicate = true;
        if (c.children_mask) c.child = root_subtrees[i].root;
      }
      res.root_subtree = ns;
      if (subtree_groups.size() == 1) {
        res.remaining = subtree_groups[0];
        res.remaining[0].allocated_root = ns.root;
      } else if (subtree_groups.size() > 1) {
        res.remaining.push_back(ns);
      }
    } else {
      for (Group* g : useless_roots) DeallocSubtree(g);
      GroupIntoBlocks(subtree_groups, uint32_t(block_index), res.remaining);
    }
    return res;
  }
};

}  // namespace

std::unique_ptr<MemoryTree> BuildOctree(const NodeTree& tree, int nodes_in_block, BuildReport* report,
                                        std::string* err) {
  if (nodes_in_block < 9 || nodes_in_block > 65536) {
    if (err) *err = "nodes_in_block must be in [9, 65536]";
    return nullptr;
  }
  if (tree.size() == 0) {
    if (err) *err = "empty tree";
    return nullptr;
  }
  std::unique_ptr<MemoryTree> out(new MemoryTree());
  StoredOctreeHeader& h = out->mutable_header();
  h.nodes_in_block = uint32_t(nodes_in_block);
  h.root_block_index = 0;
  h.channels.push_back({kNodeLinkSize, kNodeLinkChannelName});
  for (auto& c : tree.channels) h.channels.push_back(c);
  h.block_stride = uint32_t(kBlockHeaderSize + uint64_t(h.SumBytesInNode()) * nodes_in_block);
  Builder b(tree.channels, nodes_in_block, out.get());
  try {
    b.Build(tree);
  } catch (const std::exception& e) {
    if (err) *err = e.what();
    return nullptr;
  }
  h.blocks_count = uint32_t(b.blocks_count_);
  out->total_nodes = tree.size();
  if (report) {
    *report = b.report;
    report->nodes = tree.size();
    report->blocks = uint32_t(b.blocks_count_);
    report->leaf_blocks = uint32_t(b.leaf_blocks_);
    report->non_leaf_blocks = uint32_t(b.non_leaf_blocks_);
  }
  return out;
}

// ---------------------------------------------------------------- parallel
namespace {

// The records of one worker's blocks, numbered from 0 (its frontier child's
// block) in the order the sequential builder would allocate them.
class LocalBlocks : public BlockSink {
 public:
  bool Put(uint32_t index, std::unique_ptr<uint8_t[]> record) override {
    blocks.emplace_back(index, std::move(record));
    return true;
  }
  std::vector<std::pair<uint32_t, std::unique_ptr<uint8_t[]>>> blocks;
};

// What a worker built below one frontier node of the root block.
struct FrontierResult {
  bool ready = false;
  std::string error;
  uint8_t mask = 0;
  float values[6] = {0, 0, 0, 0, 0, 0};  // the frontier node's unquantised normal and colour
  std::unique_ptr<NodeTree> tree;         // the subtree (node 0 = the frontier node)
  std::optional<Subtree> small;           // a subtree left for the parent side to place
  LocalBlocks local;                      // or its committed blocks
  uint32_t count = 0;
  BuildReport report;
};

// The top of the tree, grown as the root block's breadth-first fill expands
// it; values (6 floats per node) are filled bottom-up by the replay.
struct TopTree {
  NodeTree t;
  std::vector<TerrainCell> cells;
  std::vector<float> values;
  void Add(const TerrainCell& c) {
    t.mask.push_back(0);
    t.first_child.push_back(0);
    t.data.resize(t.data.size() + 8, 0);
    cells.push_back(c);
  }
};

// Renumbers a worker's block record into the global range starting at
// `base`: its index, its parent (except for the worker's first block, whose
// parent is the root block), and every fault pointer, which names a block of
// the same worker.
bool Relocate(uint8_t* rec, uint32_t base, int nib, std::string* err) {
  StoredOctreeBlockHeader h;
  DecodeBlockHeader(rec, &h);
  const bool first = h.block_index == 0;
  h.block_index += base;
  if (h.block_index >= kMaxLinkPtr) {
    *err = "more than 2^21 blocks";
    return false;
  }
  if (!first) h.parent_block_index += base;
  for (uint32_t i = 0; i < h.roots_count; ++i)
    if (LinkFault(h.roots[i].parent_pointer_value)) h.roots[i].parent_pointer_value += base << 11;
  EncodeBlockHeader(h, rec);
  uint32_t* links = reinterpret_cast<uint32_t*>(rec + kBlockHeaderSize);
  ForEachFaultLink(links, nib, [&](int i, uint32_t) { links[i] += base << 11; });
  return true;
}

void AddReport(BuildReport* to, const BuildReport& r) {
  to->leaf_blocks += r.leaf_blocks;
  to->non_leaf_blocks += r.non_leaf_blocks;
  to->allocated_groups += r.allocated_groups;
  to->deallocated_groups += r.deallocated_groups;
}

}  // namespace

bool BuildTerrainBlocks(const TerrainParams& p, int nodes_in_block, BlockSink* sink, StoredOctreeHeader* header,
                        BuildReport* report, std::string* err) {
  std::string e;
  if (!err) err = &e;
  if (nodes_in_block < 9 || nodes_in_block > 65536) {
    *err = "nodes_in_block must be in [9, 65536]";
    return false;
  }
  const int nib = nodes_in_block;
  TerrainSource src(p);
  StoredOctreeHeader& h = *header;
  h = StoredOctreeHeader();
  h.nodes_in_block = uint32_t(nib);
  h.root_block_index = 0;
  h.channels.push_back({kNodeLinkSize, kNodeLinkChannelName});
  for (auto& c : src.channels()) h.channels.push_back(c);
  h.block_stride = uint32_t(kBlockHeaderSize + uint64_t(h.SumBytesInNode()) * nib);
  BuildReport rep;

  // Phase A: the root block's breadth-first fill over the top of the tree, grown on demand.
  TopTree top;
  top.t.max_depth = p.depth;
  top.t.channels = src.channels();
  top.t.data_size = 8;
  top.Add(src.Root());
  Builder mb(src.channels(), nib, sink);
  mb.lazy_tree = &top.t;
  mb.expand_hook = [&](uint32_t x) {
    TerrainCell kids[8];
    uint8_t m = 0;
    const TerrainCell c = top.cells[x];
    const int n = src.Children(c, kids, &m);
    top.t.mask[x] = m;
    top.t.first_child[x] = n ? uint32_t(top.cells.size()) : 0;
    for (int k = 0; k < n; ++k) top.Add(kids[k]);
  };
  Group* root = mb.Alloc();
  root->src = &top.t;
  root->nodes.push_back(BNode());
  mb.blocks_count_ = 1;  // block 0 is the root block
  bool full = false;
  Subtree res = mb.BfsFill(root, nullptr, 0, &full);
  if (!full) {  // the whole tree fits the root block: block it as a whole
    mb.DeallocSubtree(root);
    float v[6];
    const NodeTree whole = src.Subtree(src.Root(), src.threads(), v);
    Builder b(src.channels(), nib, sink);
    try {
      b.Build(whole);
    } catch (const std::exception& ex) {
      *err = ex.what();
      return false;
    }
    h.blocks_count = uint32_t(b.blocks_count_);
    rep = b.report;
    rep.nodes = whole.size();
    rep.blocks = h.blocks_count;
    rep.leaf_blocks = uint32_t(b.leaf_blocks_);
    rep.non_leaf_blocks = uint32_t(b.non_leaf_blocks_);
    AddReport(&rep, mb.report);
    if (report) *report = rep;
    return true;
  }
  int idx = 0;
  mb.FillIndexInBlockRecursive(res.root, idx);
  // The root block's unexpanded nodes, in the order ProcessBlockNodes reaches them.
  std::vector<std::pair<Group*, int>> frontier;
  std::function<void(Group*)> collect = [&](Group* g) {
    for (size_t ni = 0; ni < g->nodes.size(); ++ni) {
      BNode& n = g->nodes[ni];
      if (n.children_mask == -1) {
        n.frontier_id = int(frontier.size());
        frontier.emplace_back(g, int(ni));
      } else if (n.child) {
        collect(n.child);
      }
    }
  };
  collect(res.root);
  top.values.assign(top.cells.size() * 6, 0.f);

  // Phase B: workers build the frontier subtrees, in order, at most `window` ahead of the replay.
  std::vector<FrontierResult> results(frontier.size());
  std::mutex mu;
  std::condition_variable cv;
  std::atomic<size_t> next{0};
  size_t consumed = 0;
  bool abort = false;
  const int threads = std::max(1, src.threads());
  const size_t window = size_t(8) * threads;
  auto work = [&]() {
    for (;;) {
      const size_t f = next.fetch_add(1);
      if (f >= frontier.size()) return;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return abort || f < consumed + window; });
        if (abort) return;
      }
      FrontierResult& r = results[f];
      try {
        Group* owner = frontier[f].first;
        const int ni = frontier[f].second;
        BNode& fn = owner->nodes[ni];
        r.tree.reset(new NodeTree(src.Subtree(top.cells[fn.tree_node], 1, r.values)));
        r.mask = r.tree->mask[0];
        fn.children_mask = r.mask;
        if (r.mask) {
          Builder wb(src.channels(), nib, &r.local);
          wb.blocks_count_ = 1;  // local block 0: the frontier child's block
          Group* g = wb.Alloc();
          g->src = r.tree.get();
          uint32_t c = r.tree->first_child[0];
          for (int j = 0; j < 8; ++j)
            if (r.mask & (1u << j)) {
              BNode b;
              b.tree_node = c++;
              g->nodes.push_back(b);
            }
          auto sub = wb.BuildRecursive(g, owner, ni, 0, 0);
          if (sub) r.small = sub;
          else r.count = uint32_t(wb.blocks_count_);
          r.report = wb.report;
          r.report.leaf_blocks = uint32_t(wb.leaf_blocks_);
          r.report.non_leaf_blocks = uint32_t(wb.non_leaf_blocks_);
        }
      } catch (const std::exception& ex) {
        r.error = ex.what();
      }
      {
        std::lock_guard<std::mutex> lk(mu);
        r.ready = true;
      }
      cv.notify_all();
    }
  };

  // Phase C: the sequential builder's processing of the root block, replayed
  // with the workers' results in its own order: each frontier's blocks get
  // the indices the sequential build would have given them.
  uint64_t worker_nodes = 0;
  const bool log = std::getenv("RAYT_BUILD_LOG") != nullptr;
  mb.frontier_hook = [&](Group* owner, int ni) -> FrontierOutcome {
    BNode& fn = owner->nodes[ni];
    const size_t f = size_t(fn.frontier_id);
    FrontierResult& r = results[f];
    {
      std::unique_lock<std::mutex> lk(mu);
      cv.wait(lk, [&] { return r.ready; });
    }
    if (!r.error.empty()) throw std::runtime_error(r.error);
    worker_nodes += r.tree->size();
    AddReport(&rep, r.report);
    FrontierOutcome o;
    o.has_children = r.mask != 0;
    if (r.mask && r.small) {
      o.small = r.small;
    } else if (r.mask) {
      o.full = true;
      const uint32_t base = uint32_t(mb.blocks_count_);
      mb.blocks_count_ += int(r.count);
      fn.fault_block = int(base);
      std::string rerr;
      for (auto& b : r.local.blocks) {
        if (!Relocate(b.second.get(), base, nib, &rerr)) throw std::runtime_error(rerr);
        if (!sink->Put(base + b.first, std::move(b.second))) throw std::runtime_error("block output failed");
      }
      r.local.blocks.clear();
      r.local.blocks.shrink_to_fit();
      r.tree.reset();  // its blocks are out; nothing refers to its nodes any more
    }
    {
      std::lock_guard<std::mutex> lk(mu);
      consumed = f + 1;
    }
    cv.notify_all();
    if (log && (f + 1) % std::max<size_t>(1, frontier.size() / 20) == 0) {  // progress of long builds
      std::fprintf(stderr, "[rayt build] %zu of %zu root-block subtrees, %d blocks\n", f + 1, frontier.size(),
                   mb.blocks_count_);
      std::fflush(stderr);
    }
    return o;
  };
  // Node data of the root block, bottom-up as ProcessBlockNodes finishes each node.
  mb.node_done = [&](Group* g, int ni) {
    if (g->src != &top.t) return;
    const BNode& n = g->nodes[ni];
    float* v = &top.values[size_t(n.tree_node) * 6];
    if (n.frontier_id >= 0) {
      memcpy(v, results[size_t(n.frontier_id)].values, sizeof(float) * 6);
    } else if (n.children_mask == 0) {
      src.LeafValues(top.cells[n.tree_node], v);
    } else {
      float kids[48];
      int cnt = 0;
      for (const BNode& c : n.child->nodes) memcpy(&kids[6 * cnt++], &top.values[size_t(c.tree_node) * 6], 24);
      TerrainSource::InnerValues(kids, cnt, v);
    }
    TerrainSource::PackValues(v, &top.t.data[size_t(n.tree_node) * 8]);
  };

  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t) pool.emplace_back(work);
  bool ok = true;
  try {
    mb.FinishFullBlock(res, 0, kRootBlockParent);
  } catch (const std::exception& ex) {
    *err = ex.what();
    ok = false;
  }
  {
    std::lock_guard<std::mutex> lk(mu);
    abort = true;
  }
  cv.notify_all();
  for (auto& t : pool) t.join();
  if (!ok) return false;
  h.blocks_count = uint32_t(mb.blocks_count_);
  AddReport(&rep, mb.report);
  rep.leaf_blocks += uint32_t(mb.leaf_blocks_);
  rep.non_leaf_blocks += uint32_t(mb.non_leaf_blocks_);
  rep.nodes = uint64_t(top.cells.size()) - frontier.size() + worker_nodes;
  rep.blocks = h.blocks_count;
  if (report) *report = rep;
  return true;
}

}  // namespace rayt
