// Blocking of an in-memory octree into the stored block format.
//
// Restates StoredOctreeBuilder (CPU data/stored-octree-builder-impl.h:46-521):
// BFS-fill a block until one more sibling group could overflow it, recurse
// into the frontier, pack small leftover subtrees (<= 8 roots per block) into
// leaf blocks, and duplicate small parent-side subtrees into child blocks.
// Block numbering follows the same allocation order.
#pragma once

#include <memory>

#include "octree_format.h"
#include "scene.h"

namespace rayt {

struct BuildReport {
  uint64_t nodes = 0;
  uint32_t blocks = 0;
  uint32_t leaf_blocks = 0;
  uint32_t non_leaf_blocks = 0;
  uint64_t allocated_groups = 0;
  uint64_t deallocated_groups = 0;
};

// Builds a blocked tree in host RAM. Returns nullptr and sets *err on failure
// (e.g. more than 2^21 blocks, which fault links cannot address).
std::unique_ptr<MemoryTree> BuildOctree(const NodeTree& tree, int nodes_in_block, BuildReport* report,
                                        std::string* err);

// Builds the terrain scene of `p` block by block, in parallel: the root
// block is filled from the top of the tree; the subtree below each of its
// unexpanded nodes is generated and blocked by one of p.threads workers,
// numbering its blocks from 0; the sequential builder's processing of the
// root block is then replayed in its own order with those results, giving
// each worker's blocks the indices the sequential build gives them. The
// blocks go to `sink` as soon as they are numbered (workers run at most a
// few subtrees ahead), so the whole node tree is never held in memory, and
// they equal BuildOctree(GenerateTerrainTree(p)) byte for byte. *header gets
// the file header (block count included).
bool BuildTerrainBlocks(const TerrainParams& p, int nodes_in_block, BlockSink* sink, StoredOctreeHeader* header,
                        BuildReport* report, std::string* err);

}  // namespace rayt
