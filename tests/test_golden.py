"""The committed golden vectors of tests/golden/ (see make_golden.py): the
builder reproduces each tree byte for byte, the oracle reproduces every
vector, and the GPU path reproduces them too (hit ids and t through the
pixels and codes they determine)."""
import hashlib
import os
import sys

import numpy as np
import pytest

import oracle as orc
from rayt_amd import api

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, HERE)
import make_golden as g  # noqa: E402

NAMES = sorted(g.SCENES)


@pytest.fixture(scope="module", params=NAMES)
def scene(request):
    name = request.param
    tree, path = g.tree_file(name)
    yield name, tree, path, np.load(os.path.join(HERE, f"{name}.npz"))
    os.remove(path)


def test_builder_reproduces_golden_tree(scene):
    _, _, path, G = scene
    assert hashlib.sha256(open(path, "rb").read()).digest() == G["tree_sha256"].tobytes()


def test_oracle_reproduces_golden_vectors(scene):
    name, _, path, G = scene
    t = orc.Tree(path)
    assert t.blocks == int(G["blocks"])
    c = orc.Cache(t, t.blocks)
    c.initial_fill()
    assert c.root() == int(G["root"])
    normals, colors = g.channels(c, g.SCENES[name][1])
    flags = int(G["flags"])
    for k, (yaw, pitch, pos) in enumerate(g.CAMERAS):
        vpi = np.asarray(g.camera(yaw, pitch, pos), np.float32).reshape(16)
        np.testing.assert_array_equal(vpi.view(np.uint32), G["view_proj_inv_bits"][k])
        r = orc.render_plain(c.links(), c.far(), c.root(), normals, colors, g.NIB, g.W, g.H, vpi, flags=flags,
                             want_hits=True)
        np.testing.assert_array_equal(r["hit_node"], G["hit_node"][k])
        np.testing.assert_array_equal(r["hit_t"].view(np.uint32), G["hit_t_bits"][k])
        np.testing.assert_array_equal(r["image"], G["image"][k])
        np.testing.assert_array_equal(r["codes"], G["codes"][k])
    small = orc.Cache(t, int(G["frame_pool"]))
    small.initial_fill()
    fr = small.render_frame(g.W, g.H, G["view_proj_inv_bits"][0].view(np.float32), flags=flags, seed=7)
    assert [len(p) for p in fr["pass_faults"]] == G["frame_pass_counts"].tolist()
    assert [x for p in fr["pass_faults"] for x in p] == G["frame_faults"].tolist()
    np.testing.assert_array_equal(fr["image"], G["frame_image"])


@pytest.mark.gpu
def test_gpu_reproduces_golden_vectors(gpu_ctx, scene):
    name, tree, _, G = scene
    m = api.GPUOctreeCacheManager(tree.info().blocks_count, tree, gpu_ctx)
    m.InitialFillCache()
    assert m.root_node_index() == int(G["root"])
    flags = int(G["flags"])
    for k in range(len(g.CAMERAS)):
        vpi = G["view_proj_inv_bits"][k].view(np.float32)
        out, codes, _ = api.render_plain(gpu_ctx, m, g.W, g.H, vpi, flags=flags)
        np.testing.assert_array_equal(out.read_all(np.uint32).reshape(g.H, g.W), G["image"][k])
        np.testing.assert_array_equal(codes.read_all(np.int32).reshape(g.H, g.W), G["codes"][k])
    small = api.GPUOctreeCacheManager(int(G["frame_pool"]), tree, gpu_ctx)
    small.InitialFillCache()
    tr = api.GPURayTracer(small, g.W, g.H)
    tr.set_frame_time_limit(1e9)
    tr.set_lod_voxel_size(0.0)
    tr.set_flags(flags)
    tr.set_seed(7)
    tr.RenderFrame(G["view_proj_inv_bits"][0].view(np.float32))
    pf = tr.pass_faults()
    assert [len(p) for p in pf] == G["frame_pass_counts"].tolist()
    assert [x for p in pf for x in p] == G["frame_faults"].tolist()
    np.testing.assert_array_equal(tr.out_image(), G["frame_image"])
