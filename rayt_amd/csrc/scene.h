// Procedural scene sources: an in-memory octree (children contiguous, BFS
// order) generated from a seeded description. The builder blocks it into the
// stored format.
//
// Sources:
//  * analytic sphere with "node index" / "node index % 211" channels, the
//    test source of scene sources/stored-octree-test.cpp:237-313;
//  * seeded fBm height-field terrain plus spheres and axis-aligned boxes with
//    "normals" (char4) and "colors" (uchar4) channels laid out like the OBJ
//    importer's channels (scene sources/import-obj-scene.cpp:516-526,611-629).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "octree_format.h"

namespace rayt {

struct NodeTree {
  int max_depth = 0;
  std::vector<StoredOctreeChannel> channels;  // data channels, without node links
  int data_size = 0;                          // bytes of channel data per node
  std::vector<uint8_t> mask;                  // children mask per node
  std::vector<uint32_t> first_child;          // index of the first child (children contiguous, in bit order)
  std::vector<uint8_t> data;                  // node-major channel data (data_size bytes per node)
  uint64_t size() const { return mask.size(); }
};

// Analytic sphere, as the builder data source of stored-octree-test.cpp:237-313
// (float arithmetic, straddle test on the 8 cell corners, node ids 8*parent+child).
NodeTree GenerateSphereTree(float cx, float cy, float cz, float radius, int level);

struct TerrainParams {
  int depth = 8;
  uint64_t seed = 1;
  int octaves = 4;
  float height_lo = 0.0f;
  float height_hi = 0.6f;
  float base_frequency = 3.0f;
  int spheres = 0;
  float sphere_rmin = 0.05f, sphere_rmax = 0.2f;
  int boxes = 0;
  float box_emin = 0.01f, box_emax = 0.1f;
  int threads = 0;  // 0 = hardware concurrency
};

NodeTree GenerateTerrainTree(const TerrainParams& p);

// A cell of the terrain scene: integer coordinates at `level` and the
// primitives whose surface it straddles (bit 0 the height field, then the
// spheres, then the boxes).
struct TerrainCell {
  uint32_t i = 0, j = 0, k = 0;
  int level = 0;
  uint64_t prims[2] = {0, 0};
};

// The terrain scene as a node source: the children of any cell, and any
// cell's subtree as a NodeTree. GenerateTerrainTree is Subtree(Root()).
// Subtrees of different cells can be generated concurrently.
class TerrainSource {
 public:
  explicit TerrainSource(const TerrainParams& p);
  ~TerrainSource();
  TerrainSource(const TerrainSource&) = delete;
  TerrainSource& operator=(const TerrainSource&) = delete;
  int depth() const;
  int threads() const;
  const std::vector<StoredOctreeChannel>& channels() const;
  // The data channels of every terrain scene: "normals" (char4), "colors" (uchar4).
  static std::vector<StoredOctreeChannel> Channels();
  TerrainCell Root() const;
  // Children in child-index order (cell (2i+(c&1), 2j+((c>>1)&1), 2k+(c>>2))
  // for child c); returns their count and sets *mask. A cell at the full
  // depth has none.
  int Children(const TerrainCell& c, TerrainCell out[8], uint8_t* mask) const;
  // The subtree of `c` (node 0 = c) with the node data bytes of the full tree;
  // *root_values gets node 0's unquantised normal and colour (6 floats).
  NodeTree Subtree(const TerrainCell& c, int threads, float root_values[6]) const;
  // Unquantised normal and colour of a leaf cell, and of an inner node from
  // its children's (in child order); the byte packing of the node data.
  void LeafValues(const TerrainCell& c, float out[6]) const;
  static void InnerValues(const float* children, int count, float out[6]);
  static void PackValues(const float v[6], uint8_t out[8]);

 private:
  struct Impl;
  Impl* impl_;
};

}  // namespace rayt
