// C ABI over block sources: .tree files and trees built in host RAM.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>

#include "builder.h"
#include "obj_import.h"
#include "rt_internal.h"
#include "tree_handle.h"

using rayt::BlockSource;

namespace {

using BuildClock = std::chrono::steady_clock;

// With RAYT_BUILD_LOG set, large builds report their phases on stderr.
void BuildLog(const char* what, uint64_t count, BuildClock::time_point t0) {
  if (!std::getenv("RAYT_BUILD_LOG")) return;
  const double s = std::chrono::duration<double>(BuildClock::now() - t0).count();
  std::fprintf(stderr, "[rayt build] %s: %llu in %.1f s\n", what, static_cast<unsigned long long>(count), s);
  std::fflush(stderr);
}

int FinishBuild(const rayt::NodeTree& t, int nodes_in_block, rt_tree** out) {
  std::string err;
  rayt::BuildReport rep;
  const auto t0 = BuildClock::now();
  std::unique_ptr<rayt::MemoryTree> mt = rayt::BuildOctree(t, nodes_in_block, &rep, &err);
  if (!mt) return rayt::Fail(RT_ERR_INVALID, "build failed: " + err);
  BuildLog("blocks", rep.blocks, t0);
  if (rep.allocated_groups != rep.deallocated_groups)
    return rayt::Fail(RT_ERR_STATE, "builder leaked sibling groups");
  rt_tree* h = new rt_tree();
  h->mem = mt.get();
  h->src = std::move(mt);
  *out = h;
  return RT_OK;
}

int CheckTerrain(const rt_terrain_params* p, int max_depth, rayt::TerrainParams* tp) {
  RT_REQUIRE(p, RT_ERR_INVALID, "null argument");
  RT_REQUIRE(p->depth >= 1 && p->depth <= max_depth, RT_ERR_RANGE,
             "terrain depth must be in [1, " + std::to_string(max_depth) + "]");
  RT_REQUIRE(p->octaves >= 1 && p->octaves <= 16, RT_ERR_RANGE, "octaves must be in [1, 16]");
  RT_REQUIRE(p->spheres >= 0 && p->boxes >= 0 && p->spheres + p->boxes <= 127, RT_ERR_RANGE,
             "at most 127 spheres + boxes");
  tp->depth = p->depth;
  tp->seed = p->seed;
  tp->octaves = p->octaves;
  tp->height_lo = p->height_lo;
  tp->height_hi = p->height_hi;
  tp->spheres = p->spheres;
  tp->sphere_rmin = p->sphere_rmin;
  tp->sphere_rmax = p->sphere_rmax;
  tp->boxes = p->boxes;
  tp->box_emin = p->box_emin;
  tp->box_emax = p->box_emax;
  tp->threads = p->threads;
  return RT_OK;
}

int FinishReport(const rayt::BuildReport& rep, BuildClock::time_point t0) {
  BuildLog("terrain blocks (parallel builder)", rep.blocks, t0);
  BuildLog("terrain nodes", rep.nodes, t0);
  RT_REQUIRE(rep.allocated_groups == rep.deallocated_groups, RT_ERR_STATE, "builder leaked sibling groups");
  return RT_OK;
}

}  // namespace

extern "C" {

int rt_tree_open(const char* path, rt_tree** out) {
  RT_REQUIRE(path && out, RT_ERR_INVALID, "null argument");
  RT_TRY_BEGIN
  std::unique_ptr<rayt::StoredOctreeLoader> l(new rayt::StoredOctreeLoader(path));
  if (!l->ok()) return rayt::Fail(RT_ERR_IO, l->error());
  rt_tree* h = new rt_tree();
  h->src = std::move(l);
  *out = h;
  return RT_OK;
  RT_TRY_END
}

int rt_tree_build_sphere(float cx, float cy, float cz, float radius, int level, int nodes_in_block, rt_tree** out) {
  RT_REQUIRE(out, RT_ERR_INVALID, "null out");
  RT_REQUIRE(level >= 0 && level <= 12, RT_ERR_RANGE, "sphere level must be in [0, 12]");
  RT_TRY_BEGIN
  rayt::NodeTree t = rayt::GenerateSphereTree(cx, cy, cz, radius, level);
  return FinishBuild(t, nodes_in_block, out);
  RT_TRY_END
}

int rt_tree_build_terrain(const rt_terrain_params* p, int nodes_in_block, rt_tree** out) {
  RT_REQUIRE(out, RT_ERR_INVALID, "null out");
  rayt::TerrainParams tp;
  int rc = CheckTerrain(p, 15, &tp);
  if (rc) return rc;
  RT_TRY_BEGIN
  const auto t0 = BuildClock::now();
  std::unique_ptr<rayt::MemoryTree> mt(new rayt::MemoryTree());
  rayt::StoredOctreeHeader h;
  rayt::BuildReport rep;
  std::string err;
  if (!rayt::BuildTerrainBlocks(tp, nodes_in_block, mt.get(), &h, &rep, &err))
    return rayt::Fail(RT_ERR_INVALID, "build failed: " + err);
  if ((rc = FinishReport(rep, t0))) return rc;
  mt->mutable_header() = h;
  mt->total_nodes = rep.nodes;
  rt_tree* t = new rt_tree();
  t->mem = mt.get();
  t->src = std::move(mt);
  *out = t;
  return RT_OK;
  RT_TRY_END
}

int rt_tree_build_terrain_levelwise(const rt_terrain_params* p, int nodes_in_block, rt_tree** out) {
  RT_REQUIRE(out, RT_ERR_INVALID, "null out");
  rayt::TerrainParams tp;
  int rc = CheckTerrain(p, 14, &tp);
  if (rc) return rc;
  RT_TRY_BEGIN
  const auto t0 = BuildClock::now();
  rayt::NodeTree t = rayt::GenerateTerrainTree(tp);
  BuildLog("terrain nodes", t.size(), t0);
  return FinishBuild(t, nodes_in_block, out);
  RT_TRY_END
}

int rt_tree_build_terrain_file(const rt_terrain_params* p, int nodes_in_block, const char* path, rt_tree** out) {
  RT_REQUIRE(out && path, RT_ERR_INVALID, "null argument");
  rayt::TerrainParams tp;
  int rc = CheckTerrain(p, 15, &tp);
  if (rc) return rc;
  RT_TRY_BEGIN
  const auto t0 = BuildClock::now();
  rayt::StoredOctreeHeader h;
  h.nodes_in_block = uint32_t(nodes_in_block);
  h.channels.push_back({rayt::kNodeLinkSize, rayt::kNodeLinkChannelName});
  for (auto& c : rayt::TerrainSource::Channels()) h.channels.push_back(c);
  rayt::TreeFileWriter w(path, h);
  RT_REQUIRE(w.ok(), RT_ERR_IO, std::string("cannot create ") + path);
  rayt::BuildReport rep;
  std::string err;
  rayt::StoredOctreeHeader built;
  if (!rayt::BuildTerrainBlocks(tp, nodes_in_block, &w, &built, &rep, &err))
    return rayt::Fail(RT_ERR_IO, "build failed: " + err);
  if (!w.Finish(built.blocks_count, &err)) return rayt::Fail(RT_ERR_IO, err);
  if ((rc = FinishReport(rep, t0))) return rc;
  return rt_tree_open(path, out);
  RT_TRY_END
}

int rt_tree_import_obj(const char* obj_path, int max_level, int nodes_in_block, int threads, rt_tree** out) {
  RT_REQUIRE(obj_path && out, RT_ERR_INVALID, "null argument");
  RT_REQUIRE(max_level >= 0 && max_level <= 20, RT_ERR_RANGE, "max_level must be in [0, 20]");
  RT_TRY_BEGIN
  const auto t0 = BuildClock::now();
  rayt::NodeTree t;
  rayt::ObjImportReport rep;
  std::string err;
  if (!rayt::ImportObjTree(obj_path, max_level, threads, &t, &rep, &err)) return rayt::Fail(RT_ERR_IO, err);
  BuildLog("obj nodes", t.size(), t0);
  return FinishBuild(t, nodes_in_block, out);
  RT_TRY_END
}

int rt_tree_write(const rt_tree* tree, const char* path) {
  RT_REQUIRE(tree && path, RT_ERR_INVALID, "null argument");
  RT_REQUIRE(tree->mem, RT_ERR_STATE, "only trees built in memory can be written");
  std::string err;
  if (!tree->mem->WriteFile(path, &err)) return rayt::Fail(RT_ERR_IO, err);
  return RT_OK;
}

int rt_tree_info_get(const rt_tree* tree, rt_tree_info* out) {
  RT_REQUIRE(tree && out, RT_ERR_INVALID, "null argument");
  const auto& h = tree->src->header();
  out->nodes_in_block = h.nodes_in_block;
  out->blocks_count = h.blocks_count;
  out->root_block_index = h.root_block_index;
  out->channels_count = uint32_t(h.channels.size());
  out->bytes_in_block_content = tree->src->BytesInBlockContent();
  out->source_nodes = tree->mem ? tree->mem->total_nodes : 0;
  return RT_OK;
}

int rt_tree_channel(const rt_tree* tree, int index, int* bytes_in_node, char* name, int name_cap) {
  RT_REQUIRE(tree, RT_ERR_INVALID, "null tree");
  const auto& h = tree->src->header();
  RT_REQUIRE(index >= 0 && size_t(index) < h.channels.size(), RT_ERR_RANGE, "channel index out of range");
  if (bytes_in_node) *bytes_in_node = h.channels[index].bytes_in_node;
  if (name && name_cap > 0) {
    size_t n = std::min(size_t(name_cap - 1), h.channels[index].name.size());
    memcpy(name, h.channels[index].name.data(), n);
    name[n] = 0;
  }
  return RT_OK;
}

int rt_tree_load_block(rt_tree* tree, uint32_t index, uint8_t* header76, uint8_t* payload, uint64_t payload_cap) {
  RT_REQUIRE(tree, RT_ERR_INVALID, "null tree");
  RT_REQUIRE(index < tree->src->header().blocks_count, RT_ERR_RANGE, "block index out of range");
  RT_TRY_BEGIN
  rayt::StoredOctreeBlock b;
  if (!tree->src->LoadBlock(index, &b)) return rayt::Fail(RT_ERR_IO, "failed to load block");
  if (header76) rayt::EncodeBlockHeader(b.header, header76);
  if (payload) {
    RT_REQUIRE(payload_cap >= b.data.size(), RT_ERR_RANGE, "payload buffer too small");
    memcpy(payload, b.data.data(), b.data.size());
  }
  return RT_OK;
  RT_TRY_END
}

int rt_tree_destroy(rt_tree* tree) {
  delete tree;
  return RT_OK;
}

}  // extern "C"
