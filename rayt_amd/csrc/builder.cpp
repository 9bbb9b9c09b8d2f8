// BFS blocking builder (see builder.h). The sibling-group data structure, the
// order in which groups are expanded, the block-index allocation and the
// duplicate-subtree rule follow stored-octree-builder-impl.h:46-521 step by
// step; the node source is a precomputed NodeTree instead of a callback data
// source, so node data is copied rather than computed on demand.
#include "builder.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <optional>
#include <queue>
#include <stdexcept>

namespace rayt {

namespace {

struct Group;

struct BNode {
  uint32_t tree_node = 0;
  int children_mask = -1;  // -1: not expanded yet
  Group* child = nullptr;
  int fault_block = -1;
  bool duplicate = false;

  uint32_t NodeLink(int child_index_in_block) const {
    if (!children_mask) return PackStoredNodeLink(0, false, false, 0);
    uint32_t ptr = child ? uint32_t(child_index_in_block) : uint32_t(fault_block);
    return PackStoredNodeLink(uint32_t(children_mask), !child, duplicate, ptr);
  }
};

struct Group {
  int index_in_block = -1;
  std::vector<BNode> nodes;
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

class Builder {
 public:
  Builder(const NodeTree& t, int nib, MemoryTree* out) : t_(t), nib_(nib), out_(out) {
    payload_.resize(size_t(nib) * (kNodeLinkSize + t.data_size));
    ch_offsets_.push_back(size_t(nib) * kNodeLinkSize);
    for (size_t c = 1; c < t.channels.size(); ++c)
      ch_offsets_.push_back(ch_offsets_.back() + size_t(nib) * t.channels[c - 1].bytes_in_node);
  }

  void Build() {
    Group* root = Alloc();
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

 private:
  const NodeTree& t_;
  int nib_;
  MemoryTree* out_;
  bool log_ = std::getenv("RAYT_BUILD_LOG") != nullptr;
  int committed_ = 0;
  std::vector<uint8_t> payload_;
  std::vector<size_t> ch_offsets_;
  StoredOctreeBlockHeader hdr_;

  Group* Alloc() {
    ++report.allocated_groups;
    return new Group();
  }
  void Dealloc(Group* g) {
    ++report.deallocated_groups;
    delete g;
  }
  void DeallocSubtree(Group* g) {
    for (auto& n : g->nodes)
      if (n.child) DeallocSubtree(n.child);
    Dealloc(g);
  }

  // Expands node n: sets its children mask and, when it has children, a new
  // group with one entry per child in bit order. Returns that group.
  Group* Expand(BNode& n) {
    uint8_t m = t_.mask[n.tree_node];
    n.children_mask = m;
    if (!m) return nullptr;
    Group* g = Alloc();
    uint32_t c = t_.first_child[n.tree_node];
    for (int j = 0; j < 8; ++j)
      if (m & (1u << j)) {
        BNode b;
        b.tree_node = c++;
        g->nodes.push_back(b);
      }
    return g;
  }

  void SetFaultBlock(BNode& n, int block) {
    if (n.children_mask && !n.child) n.fault_block = block;
    if (n.child)
      for (auto& c : n.child->nodes) SetFaultBlock(c, block);
  }

  void FillIndexInBlockRecursive(Group* g, int& index) {
    if (g->index_in_block != -1) throw std::logic_error("group indexed twice");
    g->index_in_block = index;
    index += int(g->nodes.size());
    for (auto& n : g->nodes)
      if (n.child) FillIndexInBlockRecursive(n.child, index);
  }

  void PackBlockRecursive(Group* g) {
    for (size_t ni = 0; ni < g->nodes.size(); ++ni) {
      BNode& n = g->nodes[ni];
      int idx = g->index_in_block + int(ni);
      uint32_t link = n.NodeLink(n.child ? n.child->index_in_block : 0);
      memcpy(&payload_[size_t(idx) * kNodeLinkSize], &link, 4);
      const uint8_t* src = &t_.data[size_t(n.tree_node) * t_.data_size];
      size_t off = 0;
      for (size_t c = 0; c < t_.channels.size(); ++c) {
        int s = t_.channels[c].bytes_in_node;
        memcpy(&payload_[ch_offsets_[c] + size_t(s) * idx], src + off, s);
        off += s;
      }
      if (n.child) PackBlockRecursive(n.child);
    }
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
    std::fill(payload_.begin(), payload_.end(), 0);
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
      PackBlockRecursive(st.root);
      DeallocSubtree(st.allocated_root);
    }
    if (node_count > nib_) throw std::logic_error("block overflow");
    out_->PutBlock(hdr_, payload_.data());
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
    // where the reference's std::sort leaves tie order to the library).
    std::stable_sort(groups.begin(), groups.end(),
                     [](const std::vector<Subtree>& a, const std::vector<Subtree>& b) { return SumSize(a) < SumSize(b); });
    while (total_nodes >= nib_ || total_subtrees >= 8) {
      int gn = 0, gs = 0;
      std::vector<Subtree> group;
      for (int i = int(groups.size()) - 1; i >= 0; --i) {
        int n = SumSize(groups[i]), s = int(groups[i].size());
        if (gn + n <= nib_ && gs + s <= 8) {
          gn += n;
          gs += s;
          group.insert(group.end(), groups[i].begin(), groups[i].end());
          groups.erase(groups.begin() + i);
        }
      }
      if (group.empty()) throw std::logic_error("subtree group does not fit a block");
      CommitBlock(group, blocks_count_++, parent_block_index);
      ++leaf_blocks_;
      total_nodes -= gn;
      total_subtrees -= gs;
    }
    for (auto& g : groups) out_remainder.insert(out_remainder.end(), g.begin(), g.end());
  }

  BlockingState ProcessBlockNodes(Group* root, int block_index, Group* parent, int child_index) {
    bool all_roots = true;
    std::vector<Subtree> root_subtrees(root->nodes.size());
    std::vector<Group*> useless_roots;
    std::vector<std::vector<Subtree>> subtree_groups;
    int rooted_nodes_count = int(root->nodes.size());

    for (size_t ni = 0; ni < root->nodes.size(); ++ni) {
      root_subtrees[ni].root = nullptr;
      if (root->nodes[ni].children_mask == -1) {
        Group* g = Expand(root->nodes[ni]);
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

  std::optional<Subtree> BuildRecursive(Group* root, Group* parent, int child_index, int block_index,
                                        uint32_t parent_block_index) {
    Subtree res;
    res.child_index = child_index;
    res.parent = parent;
    res.root = res.allocated_root = root;
    res.node_count = int(root->nodes.size());

    std::queue<Group*> qu;
    qu.push(root);
    bool full = false;
    while (!qu.empty() && !full) {
      Group* cur = qu.front();
      qu.pop();
      for (size_t i = 0; i < cur->nodes.size(); ++i) {
        BNode& n = cur->nodes[i];
        if (res.node_count + 8 <= nib_) {
          Group* g = Expand(n);
          if (g) {
            n.child = g;
            qu.push(g);
            res.node_count += int(g->nodes.size());
          }
        } else {
          full = true;
          break;
        }
      }
    }

    if (full) {
      int idx = 0;
      FillIndexInBlockRecursive(res.root, idx);
      BlockingState s = ProcessBlockNodes(res.root, block_index, nullptr, 0);
      if (!s.remaining.empty()) {
        CommitBlock(s.remaining, blocks_count_++, uint32_t(block_index));
        ++leaf_blocks_;
      }
      CommitBlock({res}, block_index, parent_block_index);
      ++non_leaf_blocks_;
      return std::nullopt;
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
  Builder b(tree, nodes_in_block, out.get());
  try {
    b.Build();
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

}  // namespace rayt
