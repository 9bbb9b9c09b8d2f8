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

}  // namespace rayt
