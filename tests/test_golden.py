"""The committed golden vectors of tests/golden/ (see make_golden.py): the
builder reproduces the tree byte for byte, the oracle reproduces every vector,
and the GPU path reproduces them too (hit ids and t through the pixels and
codes they determine)."""
import hashlib
import os
import sys

import numpy as np
import pytest

import oracle as orc
from rayt_amd import api

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
import make_golden as g  # noqa: E402

G = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sphere5_n16.npz"))


@pytest.fixture(scope="module")
def tree_and_path():
    tree, path = g.tree_file()
    yield tree, path
    os.remove(path)


def test_builder_reproduces_golden_tree(tree_and_path):
    _, path = tree_and_path
    assert hashlib.sha256(open(path, "rb").read()).digest() == G["tree_sha256"].tobytes()


def test_oracle_reproduces_golden_vectors(tree_and_path):
    _, path = tree_and_path
    t = orc.Tree(path)
    c = orc.Cache(t, t.blocks)
    c.initial_fill()
    assert c.root() == int(G["root"])
    normals = c.channel_bytes(1, 4).view(np.uint32)
    colors = c.channel_bytes(2, 4).view(np.uint32)
    for k, (yaw, pitch, pos) in enumerate(g.CAMERAS):
        vpi = np.asarray(g.camera(yaw, pitch, pos), np.float32).reshape(16)
        np.testing.assert_array_equal(vpi.view(np.uint32), G["view_proj_inv_bits"][k])
        r = orc.render_plain(c.links(), c.far(), c.root(), normals, colors, g.NIB, g.W, g.H, vpi, flags=1,
                             want_hits=True)
        np.testing.assert_array_equal(r["hit_node"], G["hit_node"][k])
        np.testing.assert_array_equal(r["hit_t"].view(np.uint32), G["hit_t_bits"][k])
        np.testing.assert_array_equal(r["image"], G["image"][k])
        np.testing.assert_array_equal(r["codes"], G["codes"][k])
    small = orc.Cache(t, g.SMALL_POOL)
    small.initial_fill()
    fr = small.render_frame(g.W, g.H, G["view_proj_inv_bits"][0].view(np.float32), flags=1, seed=7)
    assert [len(p) for p in fr["pass_faults"]] == G["frame_pass_counts"].tolist()
    assert [x for p in fr["pass_faults"] for x in p] == G["frame_faults"].tolist()
    np.testing.assert_array_equal(fr["image"], G["frame_image"])


@pytest.mark.gpu
def test_gpu_reproduces_golden_vectors(gpu_ctx, tree_and_path):
    tree, _ = tree_and_path
    m = api.GPUOctreeCacheManager(tree.info().blocks_count, tree, gpu_ctx)
    m.InitialFillCache()
    assert m.root_node_index() == int(G["root"])
    for k in range(len(g.CAMERAS)):
        vpi = G["view_proj_inv_bits"][k].view(np.float32)
        out, codes, _ = api.render_plain(gpu_ctx, m, g.W, g.H, vpi, flags=api.L.RT_FLAG_SHADOWS)
        np.testing.assert_array_equal(out.read_all(np.uint32).reshape(g.H, g.W), G["image"][k])
        np.testing.assert_array_equal(codes.read_all(np.int32).reshape(g.H, g.W), G["codes"][k])
    small = api.GPUOctreeCacheManager(g.SMALL_POOL, tree, gpu_ctx)
    small.InitialFillCache()
    tr = api.GPURayTracer(small, g.W, g.H)
    tr.set_frame_time_limit(1e9)
    tr.set_lod_voxel_size(0.0)
    tr.set_flags(api.L.RT_FLAG_SHADOWS)
    tr.set_seed(7)
    tr.RenderFrame(G["view_proj_inv_bits"][0].view(np.float32))
    pf = tr.pass_faults()
    assert [len(p) for p in pf] == G["frame_pass_counts"].tolist()
    assert [x for p in pf for x in p] == G["frame_faults"].tolist()
    np.testing.assert_array_equal(tr.out_image(), G["frame_image"])
