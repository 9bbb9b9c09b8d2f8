"""GPU parity of the full RenderFrame loop (passes + feedback + cache
transactions + finish) against the oracle's restatement of
gpu-ray-tracer.cpp:66-227 on an undersized block pool, frame after frame:
per-pass fault lists, the image, every pool word and the LRU order must be
identical."""
import numpy as np
import pytest

import oracle as orc
from rayt_amd import api
from tests.scenes import POSES, camera_vpi, terrain_tree

pytestmark = pytest.mark.gpu


def _run(gpu_ctx, depth, nib, boxes, pool_frac, W, H, flags, max_passes, lod, poses, hits_to_extract=4000):
    tree, path = terrain_tree(depth, nib, 0, boxes)
    info = tree.info()
    max_blocks = max(4, int(info.blocks_count * pool_frac))
    m = api.GPUOctreeCacheManager(max_blocks, tree, gpu_ctx)
    m.InitialFillCache()
    tr = api.GPURayTracer(m, W, H)
    tr.set_frame_time_limit(1e9)
    tr.set_max_passes(max_passes)
    tr.set_lod_voxel_size(lod)
    tr.set_flags(flags)
    tr.set_seed(99)
    t = orc.Tree(path)
    c = orc.Cache(t, max_blocks)
    c.initial_fill()
    assert m.root_node_index() == c.root()
    rng, ef = 99, 1 << 16
    total_uploads = 0
    most_in_frame = 0
    for k, (yaw, pitch, pos) in enumerate(poses):
        vpi = camera_vpi(yaw, pitch, pos, aspect=W / H)
        tr.RenderFrame(vpi)
        r = c.render_frame(W, H, vpi, lod=lod, flags=flags, max_passes=max_passes, hits_to_extract=hits_to_extract,
                           seed=rng, expected_faults=ef)
        rng, ef = r["rng"], r["expected_faults"]
        assert tr.pass_faults() == r["pass_faults"], f"fault lists differ in frame {k}"
        st = tr.frame_stats()
        assert st[0] == r["passes"]
        np.testing.assert_array_equal(tr.out_image(), r["image"], err_msg=f"image differs in frame {k}")
        pool = m.pool_arrays()
        np.testing.assert_array_equal(pool["links"], c.links(), err_msg=f"pool links differ after frame {k}")
        np.testing.assert_array_equal(pool["far"], c.far())
        np.testing.assert_array_equal(pool["normals"], c.channel_bytes(1, 4))
        assert m.lru_order() == c.lru_order(), f"LRU order differs after frame {k}"
        most_in_frame = max([most_in_frame] + [len(p) for p in tr.pass_faults()])
        total_uploads = m.counters()[0]
    return total_uploads, m.counters(), most_in_frame


def test_frame_out_of_core_small_pool(gpu_ctx):
    uploads, counters, most = _run(gpu_ctx, depth=9, nib=512, boxes=16, pool_frac=0.35, W=320, H=240, flags=0,
                                   max_passes=6, lod=0.0, poses=POSES[:4])
    assert counters[1] > 0, "the pool was expected to evict blocks"
    # a transaction that requests blocks over several 16-block staging chunks (copy threads, chunked transfers)
    assert most >= 32, most


def test_frame_out_of_core_shadows_lod(gpu_ctx):
    _run(gpu_ctx, depth=9, nib=1024, boxes=16, pool_frac=0.5, W=256, H=192, flags=api.L.RT_FLAG_SHADOWS,
         max_passes=4, lod=1.0, poses=POSES[2:6])


def test_frame_tiny_pool_session_fills(gpu_ctx):
    """A pool far smaller than the working set: transactions stop when the LRU marker is hit."""
    _run(gpu_ctx, depth=8, nib=256, boxes=8, pool_frac=0.08, W=160, H=120, flags=0, max_passes=0, lod=0.0,
         poses=POSES[:3], hits_to_extract=50)


def test_frame_resident_single_pass(gpu_ctx):
    """Whole tree resident: one pass (fused kernel), no faults, LRU touched by hits."""
    _run(gpu_ctx, depth=8, nib=8192, boxes=0, pool_frac=1.0, W=320, H=240, flags=api.L.RT_FLAG_SHADOWS,
         max_passes=0, lod=0.0, poses=POSES[:3])


def _resident_run(gpu_ctx, tree, pipelined, W, H, poses, query_at=None):
    info = tree.info()
    m = api.GPUOctreeCacheManager(info.blocks_count, tree, gpu_ctx)
    m.InitialFillCache()
    tr = api.GPURayTracer(m, W, H)
    tr.set_pipelining(pipelined)
    tr.set_frame_time_limit(1e9)
    tr.set_lod_voxel_size(0.0)
    tr.set_flags(api.L.RT_FLAG_SHADOWS)
    tr.set_seed(5)
    mid = None
    for k, (yaw, pitch, pos) in enumerate(poses):
        tr.RenderFrame(camera_vpi(yaw, pitch, pos, aspect=W / H))
        if k == query_at:  # through the cache and the extractor: a pending frame is finished first
            mid = (m.lru_order(), tr.feedback().hit_blocks(), m.counters())
    fb = tr.feedback()
    return dict(mid=mid, lru=m.lru_order(), stats=tr.frame_stats(), faults=tr.pass_faults(), image=tr.out_image(),
                codes=tr.codes(), frames=tr.trace_totals()[1], hits=fb.hit_blocks(), dups=fb.duplicate_hit_blocks(),
                seed_state=fb.stats())


def test_pipelined_resident_frames_equal_unpipelined(gpu_ctx):
    """Frames queued back to back (the host half of each finished behind the next trace) leave exactly the
    state of frames rendered to the end one at a time: LRU order, lists, statistics, image and codes."""
    tree, _ = terrain_tree(8, 4096, 0, 8)
    W, H = 320, 240
    poses = POSES + POSES[:3]
    a = _resident_run(gpu_ctx, tree, True, W, H, poses, query_at=3)
    b = _resident_run(gpu_ctx, tree, False, W, H, poses, query_at=3)
    assert a["frames"] == b["frames"] == len(poses)
    for key in ("lru", "stats", "faults", "hits", "dups", "seed_state"):
        assert a[key] == b[key], key
    assert a["mid"][0] == b["mid"][0] and a["mid"][1] == b["mid"][1] and a["mid"][2] == b["mid"][2]
    np.testing.assert_array_equal(a["image"], b["image"])
    np.testing.assert_array_equal(a["codes"], b["codes"])
    assert len(a["hits"]) > 0


def test_two_tracers_on_one_cache_take_turns(gpu_ctx):
    """A second tracer's frame on the same cache finishes the first one's pending frame before its own."""
    tree, _ = terrain_tree(8, 4096, 0, 8)
    info = tree.info()
    orders = []
    for pipelined in (True, False):
        m = api.GPUOctreeCacheManager(info.blocks_count, tree, gpu_ctx)
        m.InitialFillCache()
        t1, t2 = api.GPURayTracer(m, 200, 150), api.GPURayTracer(m, 120, 90)
        for tr in (t1, t2):
            tr.set_pipelining(pipelined)
            tr.set_lod_voxel_size(0.0)
            tr.set_seed(3)
        for k, (yaw, pitch, pos) in enumerate(POSES):
            (t1 if k % 2 == 0 else t2).RenderFrame(camera_vpi(yaw, pitch, pos))
        orders.append((m.lru_order(), t1.frame_stats(), t2.frame_stats()))
    assert orders[0] == orders[1]


def test_scheduled_tiles_render_what_raster_order_renders(gpu_ctx):
    """Resident frames take their tiles longest first from the previous frame's wave costs; every frame,
    across launch-shape changes (which restart from raster order), equals the stateless plain render."""
    tree, _ = terrain_tree(8, 4096, 0, 8)
    info = tree.info()
    m = api.GPUOctreeCacheManager(info.blocks_count, tree, gpu_ctx)
    m.InitialFillCache()
    W, H = 328, 200  # partial tiles at the right and bottom edges
    tr = api.GPURayTracer(m, W, H)
    tr.set_lod_voxel_size(0.0)
    tr.set_flags(api.L.RT_FLAG_SHADOWS)
    try:
        for k, (waves, levels) in enumerate([(1, 10), (1, 10), (1, 10), (4, 10), (4, 10), (1, 12), (1, 12)]):
            api.plain_config(gpu_ctx, levels, waves, 0)
            yaw, pitch, pos = POSES[k % len(POSES)]
            vpi = camera_vpi(yaw, pitch, pos, aspect=W / H)
            tr.RenderFrame(vpi)
            img, codes = tr.out_image(), tr.codes()
            out, cod, _ = api.render_plain(gpu_ctx, m, W, H, vpi, lod=0.0, flags=api.L.RT_FLAG_SHADOWS)
            np.testing.assert_array_equal(img, out.read_all(np.uint32).reshape(H, W), err_msg=f"frame {k}")
            np.testing.assert_array_equal(codes, cod.read_all(np.int32).reshape(H, W), err_msg=f"codes, frame {k}")
    finally:
        api.plain_config(gpu_ctx, 16, 4, 0)
