"""The parallel terrain builder against the level-by-level one.

rt_tree_build_terrain blocks the subtrees below the root block on worker
threads, numbers their blocks from 0 and renumbers them in the sequential
builder's order (stored-octree-builder-impl.h:348-459); rt_tree_build_terrain_file
streams the same blocks into a .tree file. Both must equal, byte for byte, the
sequential builder over the whole node tree generated in memory first
(rt_tree_build_terrain_levelwise), whatever the worker count.
"""
import hashlib
import os

import numpy as np
import pytest

import oracle as orc
from rayt_amd import api


def _blocks(tree):
    info = tree.info()
    h = hashlib.sha256()
    for b in range(info.blocks_count):
        hdr, payload = tree.load_block(b)
        h.update(hdr.tobytes())
        h.update(payload.tobytes())
    return info.blocks_count, h.hexdigest()


# (depth, nodes per block, spheres, boxes, octaves, threads): tiny blocks (many small subtrees packed into
# leaf blocks, duplicates), the reference's 8192-node blocks, primitives of every kind, one and many workers
CONFIGS = [(5, 16, 0, 0, 4, 3), (6, 64, 2, 3, 4, 1), (7, 128, 2, 4, 4, 4), (8, 512, 8, 0, 4, 8),
           (9, 1024, 0, 16, 6, 8), (9, 64, 3, 3, 6, 5), (10, 8192, 0, 64, 4, 8)]


@pytest.mark.parametrize("depth,nib,spheres,boxes,octaves,threads", CONFIGS)
def test_parallel_builder_equals_levelwise(depth, nib, spheres, boxes, octaves, threads):
    spec = api.TerrainSpec(depth=depth, seed=depth + 1, octaves=octaves, spheres=spheres, boxes=boxes,
                           threads=threads)
    par = api.StoredOctree.terrain(spec, nib)
    seq = api.StoredOctree.terrain_levelwise(spec, nib)
    assert par.info().source_nodes == seq.info().source_nodes
    assert par.info().blocks_count > 1
    assert _blocks(par) == _blocks(seq)


def test_worker_count_does_not_change_the_blocks():
    base = api.TerrainSpec(depth=9, seed=4, octaves=6, boxes=8, threads=1)
    one = _blocks(api.StoredOctree.terrain(base, 256))
    for threads in (2, 7, 16):
        spec = api.TerrainSpec(depth=9, seed=4, octaves=6, boxes=8, threads=threads)
        assert _blocks(api.StoredOctree.terrain(spec, 256)) == one


def test_file_build_equals_memory_build(tmp_path):
    spec = api.TerrainSpec(depth=9, seed=7, octaves=6, spheres=2, boxes=2, threads=6)
    mem = api.StoredOctree.terrain(spec, 1024)
    path = str(tmp_path / "d9.tree")
    ftree = api.StoredOctree.terrain_file(spec, path, 1024)
    info = mem.info()
    assert ftree.info().blocks_count == info.blocks_count
    assert _blocks(ftree) == _blocks(mem)
    # the file is the header plus one record per block, as StoredOctreeRawWriter lays it out
    stride = 76 + info.bytes_in_block_content
    hdr_len = 20 + sum(8 + len(name) for _, name in mem.channels())
    assert os.path.getsize(path) == hdr_len + info.blocks_count * stride
    # and the oracle's own loader reads it
    t = orc.Tree(path)
    assert (t.nodes_in_block, t.blocks, t.root_block) == (1024, info.blocks_count, info.root_block_index)
    oh, od = t.load_block(info.blocks_count - 1)
    h, p = mem.load_block(info.blocks_count - 1)
    assert np.array_equal(od[:len(p)], p) and int(oh[0]) == info.blocks_count - 1


def test_one_block_tree_and_depth_limits():
    tiny = api.TerrainSpec(depth=2, seed=1, octaves=2, threads=4)
    assert _blocks(api.StoredOctree.terrain(tiny, 8192)) == _blocks(api.StoredOctree.terrain_levelwise(tiny, 8192))
    with pytest.raises(api.RaytError):
        api.StoredOctree.terrain(api.TerrainSpec(depth=16), 8192)
    with pytest.raises(api.RaytError):
        api.StoredOctree.terrain_levelwise(api.TerrainSpec(depth=15), 8192)
