// Wavefront OBJ scenes as octree sources.
//
// Restates ImportObjScene's data source (scene sources/import-obj-scene.cpp:
// 17-638): the OBJ reader (ReadObjFile :334-479), the MTL reader
// (MaterialLibrary :153-280), textures sampled as CImg 1.4.9's
// linear_atXY (libs/CImg.h:11725-11742), the triangle/box overlap test
// (util/intersections.cpp:7-61) and the per-node channel data of
// BuilderDataSource (:529-638). The reference drives that source depth first
// through the builder; here the tree is expanded one level at a time over
// worker threads into a NodeTree (BFS order, children contiguous), which the
// blocking builder then turns into stored blocks. Node data depends only on a
// node's box, its triangles and its children's data, so the order of
// expansion does not change a byte.
#pragma once

#include <cstdint>
#include <string>

#include "scene.h"

namespace rayt {

struct ObjImportReport {
  uint64_t faces = 0;          // triangles after fan splitting
  uint64_t materials = 0;      // including the default material
  uint64_t unsupported = 0;    // lines and options skipped as unsupported
  uint64_t triangle_refs = 0;  // (node, triangle) pairs over every level
};

// Reads `obj_path` (and its MTL libraries and PNM textures) and voxelises it
// to depth `max_level`. threads 0 = hardware concurrency. Returns false with
// *err set where the reference crashes (malformed numbers, bad faces, unknown
// material names, unreadable textures) and for texture formats other than
// PNM. On success `out` has the channels "normals" (char4) and "colors"
// (uchar4).
bool ImportObjTree(const std::string& obj_path, int max_level, int threads, NodeTree* out, ObjImportReport* report,
                   std::string* err);

}  // namespace rayt
