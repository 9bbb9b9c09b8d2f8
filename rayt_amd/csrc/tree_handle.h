// The rt_tree handle: a block source, plus the in-memory tree when it was built here.
#pragma once

#include <memory>

#include "octree_format.h"

struct rt_tree {
  std::unique_ptr<rayt::BlockSource> src;
  rayt::MemoryTree* mem = nullptr;  // non-owning view of src when it is a MemoryTree
};
