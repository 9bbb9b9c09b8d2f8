"""The reference's structural checkers, restated in the oracle, over trees
written by the product's builders:

  * StoredOctreeTraverser + CheckTree (CPU data/stored-octree-traverser.cpp:33-120):
    every block's roots hang under the parent node whose stored link they
    record, and no fault remains once every block is loaded;
  * CheckTestOctreeSphereWithLoader (scene sources/stored-octree-test.cpp:191-210):
    the sphere test scene read back through the loader equals the analytic
    sphere, node by node, with its node-index channels;
  * DebugCheckConstantDepthTree (experimental/debug-check-tree.cpp:9-47):
    every leaf of a surface voxelisation is at the full depth.

Each checker is itself pinned by mistakes planted in a copy of the file.
"""
import os
import shutil
import struct

import numpy as np
import pytest

import oracle as orc
from rayt_amd import api
from tests.scenes import sphere_tree, terrain_tree

CENTER, RADIUS = (0.3, 0.4, 0.65), 0.23


def _layout(path):
    """(header size, record stride, nodes per block, channel byte sizes) of a .tree file."""
    with open(path, "rb") as f:
        nib, blocks, root, stride, nch = struct.unpack("<5I", f.read(20))
        pos, sizes = 20, []
        for _ in range(nch):
            b, ln = struct.unpack("<2I", f.read(8))
            f.read(ln)
            sizes.append(b)
            pos += 8 + ln
    return pos, stride, nib, sizes


def _patch(src, dst, block, offset_in_record, fmt, fn):
    """Copies src to dst and rewrites one value of block `block`'s record (offset from the record start)."""
    shutil.copyfile(src, dst)
    hdr, stride, _, _ = _layout(src)
    pos = hdr + block * stride + offset_in_record
    with open(dst, "r+b") as f:
        f.seek(pos)
        v = struct.unpack(fmt, f.read(struct.calcsize(fmt)))[0]
        f.seek(pos)
        f.write(struct.pack(fmt, fn(v)))
    return orc.Tree(dst)


def _links(path, block):
    hdr, stride, nib, _ = _layout(path)
    with open(path, "rb") as f:
        f.seek(hdr + block * stride + 76)
        return np.frombuffer(f.read(4 * nib), np.uint32)


@pytest.mark.parametrize("level,nib", [(3, 131), (5, 16), (7, 111), (8, 737)])
def test_sphere_tree_through_the_loader(level, nib):
    _, path = sphere_tree(CENTER, RADIUS, level, nib)
    t = orc.Tree(path)
    assert orc.check_sphere_tree(t, CENTER, RADIUS, level) == 0
    err, nodes = orc.traverse_check(t)
    assert err == 0 and nodes > 4 ** (level - 1)
    assert orc.check_constant_depth(t, level)[0] == 0
    # a different sphere, level or channel set fails
    assert orc.check_sphere_tree(t, (0.31, 0.4, 0.65), RADIUS, level) != 0
    assert orc.check_sphere_tree(t, CENTER, RADIUS, level + 1) != 0


def test_sphere_checker_catches_planted_mistakes(tmp_path):
    _, path = sphere_tree(CENTER, RADIUS, 6, 64)
    hdr, stride, nib, sizes = _layout(path)
    t = orc.Tree(path)
    last = t.blocks - 1
    # a node index byte of the last block changed: channel data differs (15)
    ch1 = 76 + 4 * nib
    bad = _patch(path, str(tmp_path / "a.tree"), last, ch1, "<I", lambda v: v ^ 1)
    assert orc.check_sphere_tree(bad, CENTER, RADIUS, 6) == 15
    # the % 211 byte (16)
    bad = _patch(path, str(tmp_path / "b.tree"), last, ch1 + 4 * nib, "<B", lambda v: v ^ 1)
    assert orc.check_sphere_tree(bad, CENTER, RADIUS, 6) == 16
    # a root's recorded parent link no longer matches its parent node (traverser error 3)
    bad = _patch(path, str(tmp_path / "c.tree"), last, 12 + 4, "<I", lambda v: v ^ (1 << 9))
    assert orc.check_sphere_tree(bad, CENTER, RADIUS, 6) == 23
    assert orc.traverse_check(bad)[0] == 3


def test_traverser_catches_a_dangling_fault(tmp_path):
    """A child-bearing link turned into a fault to a block that does not hang there: the subtree below it is
    lost, so a fault remains (5) or a later block's parent node is missing (2)."""
    _, path = terrain_tree(7, 128, 0, 4)
    links = _links(path, 0)
    k = next(i for i, l in enumerate(links) if (l & 255) and not (l >> 8) & 1)
    bad = _patch(path, str(tmp_path / "d.tree"), 0, 76 + 4 * k, "<I", lambda v: (v & 0xFF) | (1 << 8) | (7 << 11))
    assert orc.traverse_check(bad)[0] in (2, 5)


@pytest.mark.parametrize("depth,nib,spheres,boxes", [(6, 64, 2, 3), (8, 512, 8, 0), (9, 8192, 0, 16)])
def test_builder_output_traverses_and_has_constant_depth(depth, nib, spheres, boxes):
    """Trees of the parallel builder: the traverser reaches exactly the builder's node count with no fault left,
    and every leaf is at the full depth (a cell straddling a surface always has a straddling child)."""
    tree, path = terrain_tree(depth, nib, spheres, boxes)
    t = orc.Tree(path)
    err, nodes = orc.traverse_check(t)
    assert err == 0
    assert nodes == tree.info().source_nodes
    errors, checked = orc.check_constant_depth(t, depth)
    assert errors == 0 and checked == nodes
    assert orc.check_constant_depth(t, depth - 1)[0] > 0  # "too deep"
    assert orc.check_constant_depth(t, depth + 1)[0] > 0  # "not deep enough"


def test_file_built_tree_traverses(tmp_path):
    spec = api.TerrainSpec(depth=8, seed=3, octaves=6, boxes=4, threads=4)
    path = str(tmp_path / "d8.tree")
    api.StoredOctree.terrain_file(spec, path, 256)
    mem = api.StoredOctree.terrain(spec, 256)
    err, nodes = orc.traverse_check(orc.Tree(path))
    assert err == 0 and nodes == mem.info().source_nodes
    assert os.path.getsize(path) > 0
